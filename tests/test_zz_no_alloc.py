"""eig_lsqproject_noalloc (the allocation-free variant of eig_lsqproject) against
eig_lsqproject through the C-ABI: the same status codes and the same bits, on two
calls in a row, the second with a smaller panel (fewer individuals and SNPs), and an
input beyond its static workspace rejected with EIG_E_ARG."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _setup(eng, seed, M, N, K, random_axes=False):
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.05, 0.5, M)
    g = np.concatenate([rng.binomial(2, rng.beta(p * 9, (1 - p) * 9)[:, None], size=(M, N // 3))
                        for _ in range(3)], axis=1).astype(np.int32)
    N = g.shape[1]
    g[rng.random(g.shape) < 0.05] = -1
    g[:, 7] = -1  # an individual with no data: status 1
    rows = O.pack_genotypes(g)
    eng.load_packed(rows, N)
    eng.select(np.arange(N, dtype=np.int32))
    eng.snpstats(fetch=False)
    if random_axes:  # loadings of arbitrary unit axes: no N x N matrix needed
        vec = rng.standard_normal((K, N))
        vec /= np.linalg.norm(vec, axis=1, keepdims=True)
    else:
        eng.xtx(0, 0, want_matrix=False)
        ev, vec = eng.eigh_xtx(K)
    eng.loadings(vec, want_ffvecs=False)
    use = np.ones(N, np.int32)
    use[::13] = 0
    return N, use


def test_noalloc_lsqproject_matches_lsqproject_twice():
    import eig_amd
    eng = eig_amd.Engine(0)
    L = eig_amd.lib()
    try:
        for seed, M, N, K in ((1, 30000, 2400, 10), (2, 9000, 600, 6)):
            N, use = _setup(eng, seed, M, N, K)
            _say(f"noalloc lsqproject N={N} M={M} K={K}")
            a, b, s = eng.lsqproject(use, K)
            a2 = np.full((K, N), -7.0); b2 = np.full((K, N), -7.0); s2 = np.zeros(N, np.int32)
            a[:, (s == 1) | (s == 3)] = -7.0  # untouched by both entry points
            b[:, (s == 1) | (s == 3)] = -7.0
            rc = L.eig_lsqproject_noalloc(eng.h, use, a2.reshape(-1), b2.reshape(-1), s2)
            assert rc == 0, L.eig_last_error(eng.h)
            assert np.array_equal(s2, s)
            assert (s == 1).any() and (s == 3).any()
            assert np.array_equal(a2, a) and np.array_equal(b2, b)
    finally:
        eng.close()


def test_noalloc_lsqproject_rejects_inputs_beyond_its_workspace():
    import eig_amd
    eng = eig_amd.Engine(0)
    L = eig_amd.lib()
    try:
        N, use = _setup(eng, 3, 300, 45000, 16, random_axes=True)
        _say(f"noalloc lsqproject oversize N={N}")
        a = np.zeros((16, N)); b = np.zeros((16, N)); s = np.zeros(N, np.int32)
        rc = L.eig_lsqproject_noalloc(eng.h, use, a.reshape(-1), b.reshape(-1), s)
        assert rc == -1
        assert b"static workspace" in L.eig_last_error(eng.h)
    finally:
        eng.close()
