"""fastmode (eig_fpca) against the oracle's restatement of kjg_fpca and against
the exact eigenpairs it converges to.

Bars: with the same start panel, eigenvalues within 1e-10 relative and
sign-normalised eigenvectors within 1e-8 of oracle.fpca (numpy SVDs in place
of the reference's GSL/LAPACK ones); with the built-in generator, top-K
eigenvalues within 1e-9 relative of the exact eigenvalues of X^T X / ncols and
eigenvectors within 1e-6.  Parity with the reference's own output is unpinned:
its start panel comes from GSL's generator, which is not in this image.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    import eig_amd
    e = eig_amd.Engine(0)
    yield e
    e.close()


def _panel(seed, M, npop, per, fst=0.1, miss=0.02):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.05, 0.5, size=M)
    cols = []
    for k in range(npop):
        f = fst * (1 + 0.3 * k)
        pk = rng.beta(p * (1 - f) / f, (1 - p) * (1 - f) / f)
        cols.append(rng.binomial(2, pk[:, None], size=(M, per)))
    g = np.concatenate(cols, axis=1).astype(np.int32)
    g[rng.random(g.shape) < miss] = -1
    return g


def _oracle_X(rows, xindex, nind, live=None):
    st = O.snpstats(rows, xindex)
    lf = O.fastmode_levels(st["mean"], st["fancy"], st["status"])
    X = O.fastmode_matrix(rows, xindex, lf, nind)
    return X if live is None else X[live.astype(bool)]


def test_fpca_matches_oracle_with_same_start_panel(eng):
    g = _panel(7, 4001, 4, 45)
    g[10, :] = -1
    g[11, :] = 0
    N = g.shape[1]
    rows = O.pack_genotypes(g)
    xindex = np.array([i for i in range(N) if i % 7 != 3], dtype=np.int32)
    live = np.ones(g.shape[0], np.uint8)
    live[100:160] = 0
    # four populations: three converged axes (an unconverged Ritz value in the noise
    # bulk depends on rounding, in the reference's SVD as much as here)
    K, L, I = 3, 8, 5
    omega = np.random.default_rng(9).standard_normal((len(xindex), L))
    _say(f"fpca vs oracle N={N} nrows={len(xindex)} M={g.shape[0]} K={K} L={L} I={I}")
    eng.load_packed(rows, N)
    eng.select(xindex)
    eng.snpstats(live=live)
    ev, vec = eng.fpca(K, L, I, omega=omega)
    X = _oracle_X(rows, xindex, N, live)
    rev, rvec = O.fpca(X, K, L, I, omega)
    assert np.allclose(ev, rev, rtol=1e-10), (ev, rev)
    assert np.max(np.abs(vec - rvec)) < 1e-8
    assert np.allclose(np.sum(vec * vec, axis=1), 1.0, atol=1e-13)


@pytest.mark.parametrize("N,M,K,L,I", [(300, 20000, 10, 20, 10), (129, 3000, 2, 6, 8)])
def test_fpca_builtin_generator_reaches_exact_eigenpairs(eng, N, M, K, L, I):
    g = _panel(N, M, 12, N // 12 + 1)[:, :N]
    rows = O.pack_genotypes(g)
    xindex = np.arange(N, dtype=np.int32)
    _say(f"fpca exact N={N} M={M} K={K} L={L} I={I}")
    eng.load_packed(rows, N)
    eng.select(xindex)
    eng.snpstats()
    ev, vec = eng.fpca(K, L, I, seed=1234)
    X = _oracle_X(rows, xindex, N)
    w, v = np.linalg.eigh(X.T @ X / X.shape[0])
    w, v = w[::-1][:K], O.sign_rule(v[:, ::-1][:, :K].T)
    assert np.allclose(ev, w, rtol=1e-9), (ev, w)
    assert np.max(np.abs(vec - v)) < 1e-6
    # a second call replays the captured iteration and gives the same bits
    ev2, vec2 = eng.fpca(K, L, I, seed=1234)
    assert np.array_equal(ev, ev2) and np.array_equal(vec, vec2)


# L=40: 64-column panel products in the power iteration; L=10, I=2: a B product of
# at most 32 columns (its 32-column X^T B group)
@pytest.mark.parametrize("N,M,K,L,I", [(300, 20000, 10, 20, 10), (200, 4000, 3, 8, 5),
                                       (200, 4000, 4, 40, 2), (150, 3000, 3, 10, 2)])
def test_fpca_split_bf16_panels_match_fp64(eng, N, M, K, L, I):
    """The split-BF16 panel products (4 products, FP32 within 512-sample/SNP blocks)
    against the FP64 form with the same start panel.  Bars: eigenvalues 1e-6
    relative, eigenvectors 1e-5."""
    import eig_amd
    g = _panel(N + 1, M, 12, N // 12 + 1)[:, :N]
    rows = O.pack_genotypes(g)
    xindex = np.arange(N, dtype=np.int32)
    _say(f"fpca bf16 vs fp64 N={N} M={M} K={K} L={L} I={I}")
    eng.load_packed(rows, N)
    eng.select(xindex)
    eng.snpstats()
    ev64, v64 = eng.fpca(K, L, I, seed=99, form=eig_amd.FORM_FP64)
    ev16, v16 = eng.fpca(K, L, I, seed=99, form=eig_amd.FORM_BF16X4)
    assert np.allclose(ev16, ev64, rtol=1e-6), (ev16, ev64)
    assert np.max(np.abs(v16 - v64)) < 1e-5
    ev16b, v16b = eng.fpca(K, L, I, seed=99, form=eig_amd.FORM_BF16X4)
    assert np.array_equal(ev16, ev16b) and np.array_equal(v16, v16b)


def test_fastmode_selection_above_the_grm_limit(eng):
    """fastmode takes more samples than the N x N path allows (the reference checks
    nrows < 46341 only outside fastmode, smartpca.c:955-988)."""
    import eig_amd
    N, M = 46400, 256
    rng = np.random.default_rng(5)
    p = rng.uniform(0.1, 0.5, size=M)
    g = rng.binomial(2, p[:, None], size=(M, N)).astype(np.int32)
    rows = O.pack_genotypes(g)
    _say(f"fastmode above the GRM limit N={N} M={M}")
    eng.load_packed(rows, N)
    eng.select(np.arange(N, dtype=np.int32))
    eng.snpstats()
    with pytest.raises(eig_amd.EigError, match="GRM"):
        eng.xtx(0, 0, want_matrix=False)
    ev, vec = eng.fpca(2, 4, 2, seed=3)
    assert np.all(np.isfinite(ev)) and ev[0] > 0 and vec.shape == (2, N)
    assert np.allclose(np.sum(vec * vec, axis=1), 1.0, atol=1e-12)


def test_fpca_rejects_bad_shapes(eng):
    import eig_amd
    g = _panel(1, 500, 2, 20)
    rows = O.pack_genotypes(g)
    eng.load_packed(rows, 40)
    eng.select(np.arange(40, dtype=np.int32))
    eng.snpstats()
    with pytest.raises(eig_amd.EigError):
        eng.fpca(4, 4, 2)          # L must exceed K
    with pytest.raises(eig_amd.EigError):
        eng.fpca(4, 30, 9)         # (I+1)*L > 256


def test_fpca_split_bf16_calls_with_changing_k_and_mask(eng):
    """Consecutive split-BF16 fastmode calls that change K (the size of the eigenvector
    region of the workspace) and the SNP mask (the 1/ncols scale) at the same L and I
    re-capture the power-iteration graph: each result matches the FP64 form."""
    import eig_amd
    g = _panel(31, 6000, 4, 60)
    N = g.shape[1]
    rows = O.pack_genotypes(g)
    eng.load_packed(rows, N)
    eng.select(np.arange(N, dtype=np.int32))
    live = np.ones(6000, np.uint8)
    for K, cut in ((10, 0), (3, 0), (3, 1500), (6, 0)):
        live[:] = 1
        live[:cut] = 0
        eng.snpstats(live=live)
        _say(f"fpca bf16x4 K={K} masked={cut}")
        e64, v64 = eng.fpca(K, 12, 3, seed=5, form=eig_amd.FORM_FP64)
        eb, vb = eng.fpca(K, 12, 3, seed=5, form=eig_amd.FORM_BF16X4)
        # the three population axes (the bulk Ritz pairs of a 3-iteration run are not
        # converged, and their form-to-form differences say nothing about the graph)
        assert np.max(np.abs(eb[:3] - e64[:3]) / np.abs(e64[:3])) <= 1e-6, (K, cut)
        assert np.max(np.abs(vb[:3] - v64[:3])) <= 1e-5, (K, cut)
