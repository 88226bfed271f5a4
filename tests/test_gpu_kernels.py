"""GPU parity of each kernel family against the CPU oracle (marked gpu).

Bars: bit-exact for decode, counts, missingness, keep flags and the FP64 column
statistics (built without FMA contraction); FP64 XTX within 1e-12 of the
largest entry; split-BF16 XTX within 2e-6 relative (its stated error budget);
eigenvalues within 1e-10 * lambda_max of LAPACK dspev; sign-normalised
eigenvectors within 1e-8; loadings / lsqproject within 1e-9 relative.
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


def _rand_panel(seed, M, N, miss=0.03):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.02, 0.5, size=M)
    g = rng.binomial(2, p[:, None], size=(M, N)).astype(np.int32)
    g[rng.random((M, N)) < miss] = -1
    g[3 % M, :] = 1 if M > 3 else g[3 % M, :]
    return g


def test_synth_matches_cpu_restatement(eng):
    from eig_amd import synth
    P = synth.config_panel(3, nsnp=700, scale=0.01)  # admixed, 250 samples
    _say(f"synth M={P.nsnp} N={P.nind}")
    P.to_device(eng)
    got = eng.fetch_packed()
    want = P.packed()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("N,sub", [(5, False), (131, False), (300, True), (1031, True)])
def test_snpstats_bit_exact(eng, N, sub):
    g = _rand_panel(N, 517, N)
    g[7, :] = -1
    rows = O.pack_genotypes(g)
    xindex = np.arange(N, dtype=np.int32)
    if sub:
        xindex = xindex[(xindex % 5) != 2]
    _say(f"snpstats N={N} nrows={len(xindex)}")
    eng.load_packed(rows, N)
    eng.select(xindex)
    for altnorm in (1, 0):
        st = eng.snpstats(altnorm=altnorm)
        ref = O.snpstats(rows, xindex, altnorm=altnorm)
        for k in ("c0", "c1", "nmiss", "status"):
            assert np.array_equal(st[k], ref[k]), k
        ok = ref["status"] != 2
        assert np.array_equal(st["mean"][ok], ref["mean"][ok])
        assert np.array_equal(st["fancy"][ok], ref["fancy"][ok])


def test_snpstats_live_mask(eng):
    g = _rand_panel(11, 300, 77)
    rows = O.pack_genotypes(g)
    eng.load_packed(rows, 77)
    eng.select(np.arange(77, dtype=np.int32))
    live = np.ones(300, np.uint8)
    live[::3] = 0
    st = eng.snpstats(live=live)
    assert np.all(st["status"][::3] == 3)
    ref = O.snpstats(rows, np.arange(77, dtype=np.int32))
    assert np.array_equal(st["c0"][1::3], ref["c0"][1::3])


def test_snpstats_without_fetch_feeds_xtx_the_same(eng):
    """Statistics left on the device (no host outputs) give the same XTX bits as a
    call that also copies them out."""
    g = _rand_panel(12, 2000, 90)
    rows = O.pack_genotypes(g)
    eng.load_packed(rows, 90)
    eng.select(np.arange(90, dtype=np.int32))
    st = eng.snpstats()
    _, m1 = eng.xtx(0)
    assert eng.snpstats(fetch=False) is None
    _, m2 = eng.xtx(0)
    assert np.array_equal(m1, m2)
    assert np.array_equal(st["c0"], O.snpstats(rows, np.arange(90, dtype=np.int32))["c0"])


@pytest.mark.parametrize("N,M", [(100, 3000), (300, 4500), (129, 257)])
def test_xtx_fp64_matches_oracle(eng, N, M):
    g = _rand_panel(N + M, M, N)
    rows = O.pack_genotypes(g)
    xindex = np.arange(N, dtype=np.int32)[np.arange(N) % 11 != 4]
    _say(f"xtx fp64 N={N} M={M}")
    eng.load_packed(rows, N)
    eng.select(xindex)
    eng.snpstats()
    y, mat = eng.xtx(0)
    ref = O.snpstats(rows, xindex)
    lower = O.xtx_lookup(rows, xindex, ref["levels"], ref["status"])
    full, yr = O.finalize_xtx(lower, len(xindex))
    assert abs(y - yr) <= 1e-12 * abs(yr)
    assert np.max(np.abs(mat - full)) <= 1e-12 * np.max(np.abs(full))
    assert np.array_equal(mat, mat.T)


# 3 products: the dropped lo*lo term biases every genotype-sharing pair (~2e-6).
# 4 products with 256-SNP FP32 blocks (the chosen form): left with the hi+lo
# residual (~2^-17 of each level) that walks randomly over SNPs, about
# 2^-16/sqrt(M) = 1.6e-7 per entry at M=9000.  Long FP32 blocks add a rounding
# drift that grows with the block length (see the error sweep in DESIGN).
@pytest.mark.parametrize("form,block,tol", [(1, 4096, 2e-6), (1, 64, 2e-6), (2, 256, 5e-7),
                                            (2, 64, 5e-7), (2, 48, 5e-7)])
def test_xtx_split_bf16_error_budget(eng, form, block, tol):
    N, M = 256, 9000
    g = _rand_panel(5, M, N)
    rows = O.pack_genotypes(g)
    xindex = np.arange(N, dtype=np.int32)
    _say(f"xtx split-bf16 form={form} N={N} M={M} block={block}")
    eng.load_packed(rows, N)
    eng.select(xindex)
    eng.snpstats()
    _, m64 = eng.xtx(0)
    _, m16 = eng.xtx(form, block)
    rel = np.max(np.abs(m16 - m64)) / np.max(np.abs(m64))
    assert rel < tol, rel
    assert np.array_equal(m16, m16.T)


@pytest.mark.parametrize("n", [1, 2, 3, 33, 77, 300])
def test_eigh_matches_lapack(eng, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    a = a @ a.T / n + np.diag(rng.uniform(0, 1, n))
    nvec = min(n, 10)
    _say(f"eigh n={n}")
    w, v = eng.eigh(a, nvec)
    wr, vr = O.dspev_lower(a)
    assert np.max(np.abs(w - wr)) <= 1e-10 * max(1.0, abs(wr[0]))
    vr = O.sign_rule(vr[:nvec])
    gaps = np.abs(np.diff(wr[:nvec + 1])) if n > nvec else np.abs(np.diff(wr))
    for j in range(nvec):
        # compare only well separated eigenvectors
        sep = min([abs(wr[j] - wr[k]) for k in range(n) if k != j] or [1.0])
        if sep > 1e-6 * abs(wr[0]):
            assert np.max(np.abs(v[j] - vr[j])) < 1e-8 / min(1.0, sep), j
    assert np.allclose(v @ v.T, np.eye(nvec), atol=1e-10)
    assert np.allclose(a @ v.T, v.T * w[:nvec], atol=1e-9 * max(1.0, abs(wr[0])))


def test_eigh_clustered_and_structured(eng):
    # a population-structure-like matrix: few large eigenvalues, clustered bulk
    rng = np.random.default_rng(9)
    n = 257
    u = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = np.concatenate([[30, 12, 12 * (1 + 1e-9), 5], np.ones(n - 4)])
    a = (u * lam) @ u.T
    a = (a + a.T) / 2
    w, v = eng.eigh(a, 4)
    assert np.allclose(np.sort(w)[::-1], np.sort(lam)[::-1], atol=1e-10 * 30)
    assert np.allclose(v @ v.T, np.eye(4), atol=1e-9)
    assert np.allclose(a @ v.T, v.T * w[:4], atol=1e-8)


@pytest.mark.parametrize("n", [1001, 2305])
def test_eigh_both_column_launch_shapes(eng, n):
    """Columns with trailing order >= 1,000 run as two kernels (symv, then w-column),
    shorter ones as one launch with the in-launch hand-off (n = 1,001: only the first
    column is long; n = 2,305: 1,305 long columns, odd and even first rows).  The graph
    is captured on the first call and replayed on the second, which must give the
    same bits.  Bars: eigenvalues within 1e-10 * lambda_max of LAPACK (numpy's
    dsyevd), residuals 1e-9 * lambda_max, orthonormal vectors."""
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    a = a @ a.T / n + np.diag(rng.uniform(0, 1, n))
    _say(f"eigh n={n} (both launch shapes)")
    w, v = eng.eigh(a, 10)
    wr = np.linalg.eigvalsh(a)[::-1]
    assert np.max(np.abs(w - wr)) <= 1e-10 * wr[0]
    assert np.allclose(v @ v.T, np.eye(10), atol=1e-10)
    assert np.max(np.abs(a @ v.T - v.T * w[:10])) <= 1e-9 * wr[0]
    w2, v2 = eng.eigh(a, 10)
    assert np.array_equal(w, w2) and np.array_equal(v, v2)


def test_eigh_xtx_pipeline(eng):
    N, M = 200, 4000
    g = _rand_panel(21, M, N)
    rows = O.pack_genotypes(g)
    xindex = np.arange(N, dtype=np.int32)
    eng.load_packed(rows, N)
    eng.select(xindex)
    eng.snpstats()
    _, mat = eng.xtx(0)
    w, v = eng.eigh_xtx(10)
    wr, vr = O.dspev_lower(mat)
    assert np.max(np.abs(w - wr)) < 1e-10 * wr[0]
    assert abs(w.sum() - (N - 1)) < 1e-8


def test_loadings_and_lsqproject_match_oracle(eng):
    N, M, K = 150, 3000, 6
    g = _rand_panel(33, M, N, miss=0.1)
    g[:, 5] = -1                      # individual with no data at all
    g[10:, 6] = -1                    # only 10 observed -> fine (K=6)
    g[K:, 7] = -1                     # K observed -> ignored
    rows = O.pack_genotypes(g)
    xindex = np.array([i for i in range(N) if i % 3 != 1 and i != 5], dtype=np.int32)
    _say(f"loadings/lsqproject N={N} M={M} K={K}")
    eng.load_packed(rows, N)
    eng.select(xindex)
    st = eng.snpstats()
    _, mat = eng.xtx(0)
    w, v = eng.eigh_xtx(K)
    ref = O.snpstats(rows, xindex)
    ign = (ref["status"] != 0).astype(np.int32)
    ff = np.zeros((K, M))
    O.lib().orc_loadings(rows.reshape(-1), rows.shape[1], M, xindex, len(xindex), ign, 1, 1,
                         np.ascontiguousarray(v), K, ff.reshape(-1))
    gff, gfx = eng.loadings(v)
    assert np.max(np.abs(gff - ff)) <= 1e-11 * np.max(np.abs(ff))
    xm = ref["mean"] * ref["fancy"]
    fxv = np.zeros((K, len(xindex))); fx = np.zeros(K)
    O.lib().orc_fxscal(rows.reshape(-1), rows.shape[1], M, xindex, len(xindex), ign, xm,
                       np.ascontiguousarray(ref["fancy"]), ff.reshape(-1), K, fxv.reshape(-1), fx)
    assert np.allclose(gfx, fx, rtol=1e-11)
    use = np.ones(N, np.int32)
    use[5] = 0
    a, b, s = eng.lsqproject(use, K)
    ra = np.zeros((K, N)); rb = np.zeros((K, N)); rs = np.zeros(N, np.int32)
    O.lib().orc_lsqproj(rows.reshape(-1), rows.shape[1], M, N, use, ign, xm,
                        np.ascontiguousarray(ref["fancy"]), ff.reshape(-1), fx, K,
                        ra.reshape(-1), rb.reshape(-1), rs)
    assert np.array_equal(s, rs)
    assert s[7] == 1 and s[5] == 3
    ok = rs == 0
    assert np.allclose(a[:, ok], ra[:, ok], rtol=1e-9, atol=1e-12)
    assert np.allclose(b[:, ok], rb[:, ok], rtol=1e-9, atol=1e-12)


def test_eigh_rejects_non_finite_input_before_any_launch(eng):
    """eig_eigh checks the caller's matrix on the host: a NaN or Inf entry is an
    argument error (EIG_E_ARG) and no kernel runs on it."""
    import eig_amd
    for bad in (np.nan, np.inf, -np.inf):
        a = np.eye(6)
        a[2, 4] = a[4, 2] = bad
        _say(f"eigh non-finite {bad}")
        with pytest.raises(eig_amd.EigError, match=r"rc=-1: .*not finite"):
            eng.eigh(a, 2)
    w, v = eng.eigh(np.diag([3.0, 2.0, 1.0]), 1)  # the context still works
    assert np.allclose(w, [3.0, 2.0, 1.0]) and np.allclose(np.abs(v[0]), [1.0, 0.0, 0.0])


@pytest.mark.parametrize("n,kind", [(300, "random"), (257, "clustered"), (700, "structure")])
def test_eigh_all_vectors_in_batches(eng, n, kind):
    """eigvecs' contract (eigsubs.c:39-55): every eigenvector.  Past 64 vectors the
    eigensolver runs batches of 64 and orthogonalises each against the earlier batches'
    vectors of its cluster; checked against LAPACK dspev: eigenvalues, orthonormality of
    all n vectors, residuals, and the vectors of well separated eigenvalues."""
    rng = np.random.default_rng(n)
    if kind == "random":
        a = rng.standard_normal((n, n))
        a = a @ a.T / n
    else:
        u = np.linalg.qr(rng.standard_normal((n, n)))[0]
        if kind == "clustered":  # a 150-fold eigenvalue spanning three batches
            lam = np.concatenate([np.linspace(40, 5, 50), np.ones(150), np.linspace(0.5, 0.01, n - 200)])
        else:  # few large eigenvalues over a Marchenko-Pastur-like bulk, as an XTX has
            lam = np.concatenate([[60, 30, 12], rng.uniform(0.8, 1.2, n - 3)])
        a = (u * lam) @ u.T
        a = (a + a.T) / 2
    _say(f"eigh all vectors n={n} {kind}")
    w, v = eng.eigh(a, n)
    wr, vr = O.dspev_lower(a)
    lam_max = abs(wr[0])
    assert np.max(np.abs(w - wr)) <= 1e-10 * lam_max
    orth = np.max(np.abs(v @ v.T - np.eye(n)))
    res = np.max(np.abs(a @ v.T - v.T * w))
    _say(f"all vectors n={n} {kind}: |V^T V - I| {orth:.2e}, residual {res:.2e}")
    assert orth <= 1e-10
    assert res <= 1e-9 * lam_max
    vr = O.sign_rule(vr)
    for j in range(n):
        sep = min(abs(wr[j] - wr[k]) for k in range(n) if k != j)
        if sep > 1e-3 * lam_max:
            assert np.max(np.abs(v[j] - vr[j])) < 1e-8 / min(1.0, sep), j
