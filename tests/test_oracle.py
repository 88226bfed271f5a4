"""Pins the CPU oracle against the reference's own fixtures and against mathematics.

Known-answer vectors come from the reference tree (copied under tests/golden/):
EIGENSTRAT/example.* (.eval, .pca.evec), CONVERTF/example.* (one 7x5 genotype
table in packed GENO, EIGENSTRAT text and PLINK .bed form), POPGEN/grmjunk.
Closed forms and LAPACK (numpy's OpenBLAS) pin what the fixtures leave open.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O


def _ids(path):
    return [l.split()[0] for l in open(path) if l.split()]


def test_geno_header_hash_kat(golden):
    d = os.path.join(golden, "convertf")
    ind = _ids(os.path.join(d, "example.ind"))
    snp = _ids(os.path.join(d, "example.snp"))
    assert O.hasharr(ind) & 0xFFFFFFFF == 0xD28EE92A
    assert O.hasharr(snp) & 0xFFFFFFFF == 0x7A6C3691
    hdr = open(os.path.join(d, "example.packedancestrymapgeno"), "rb").read(48)
    assert hdr == O.geno_header(len(ind), len(snp), O.hasharr(ind), O.hasharr(snp))


def test_decode_kat_three_formats(golden):
    d = os.path.join(golden, "convertf")
    ind = _ids(os.path.join(d, "example.ind"))
    snp = _ids(os.path.join(d, "example.snp"))
    packed = O.read_geno_packed(os.path.join(d, "example.packedancestrymapgeno"), snp, ind)
    text = [l.strip() for l in open(os.path.join(d, "example.eigenstratgeno")) if l.strip()]
    want = np.array([[int(c) for c in t] for t in text], dtype=np.int32)
    assert np.array_equal(O.decode_rows(packed, 5), want)
    bed = O.read_bed(os.path.join(d, "example.bed"), 5, 7)
    assert np.array_equal(O.decode_rows(bed, 5), want)
    # first two records, MSB-first: 0x54 0x00 -> 11100 ; 0x19 0x80 -> 01212
    assert list(packed[0, :2]) == [0x54, 0x00] and list(packed[1, :2]) == [0x19, 0x80]


def test_decode_brute_force():
    rng = np.random.default_rng(1)
    g = rng.integers(-1, 3, size=(17, 203)).astype(np.int32)
    rows = O.pack_genotypes(g)
    assert rows.shape[1] == 51
    assert np.array_equal(O.decode_rows(rows, 203), g)


def test_hash_rejects_changed_ids(golden, tmp_path):
    d = os.path.join(golden, "convertf")
    ind = _ids(os.path.join(d, "example.ind"))
    snp = _ids(os.path.join(d, "example.snp"))
    with pytest.raises(ValueError, match="indiv file has changed"):
        O.read_geno_packed(os.path.join(d, "example.packedancestrymapgeno"), snp,
                           ind[:4] + ["OTHER"])
    with pytest.raises(ValueError, match="snp file has changed"):
        O.read_geno_packed(os.path.join(d, "example.packedancestrymapgeno"), snp[::-1], ind)
    raw = open(os.path.join(d, "example.packedancestrymapgeno"), "rb").read()
    p = tmp_path / "t.geno"
    p.write_bytes(raw[:-10])
    with pytest.raises(ValueError, match="truncated"):
        O.read_geno_packed(str(p), snp, ind)


def _rand_geno(rng, M, N, miss=0.05):
    p = rng.uniform(0.05, 0.5, size=M)
    g = rng.binomial(2, p[:, None], size=(M, N)).astype(np.int32)
    g[rng.random((M, N)) < miss] = -1
    return g


@pytest.mark.parametrize("altnorm", [1, 0])
def test_snpstats_brute_force(altnorm):
    rng = np.random.default_rng(2)
    g = _rand_geno(rng, 300, 97)
    g[5, :] = -1           # all missing
    g[6, :] = 0            # monomorphic
    g[7, :] = 2
    rows = O.pack_genotypes(g)
    xindex = np.arange(0, 97, 2, dtype=np.int32)
    st = O.snpstats(rows, xindex, altnorm=altnorm)
    sub = g[:, xindex]
    obs = sub >= 0
    c0 = np.where(obs, sub, 0).sum(1)
    c1 = np.where(obs, 2 - sub, 0).sum(1)
    nm = (~obs).sum(1)
    assert np.array_equal(st["c0"], c0) and np.array_equal(st["c1"], c1)
    assert np.array_equal(st["nmiss"], nm)
    ok = nm < len(xindex)
    mean = c0[ok] / (len(xindex) - nm[ok])
    p = mean / 2 if altnorm else (c0[ok] + 1.0) / (2.0 * (len(xindex) - nm[ok]) + 2.0)
    v = p * (1 - p)
    fancy = np.where(v > 0, 1 / np.sqrt(np.where(v > 0, v, 1)), 1.0)
    assert np.array_equal(st["mean"][ok], mean)
    assert np.array_equal(st["fancy"][ok], fancy)
    lv = (np.arange(3)[None, :] - mean[:, None]) * fancy[:, None]
    assert np.array_equal(st["levels"][ok], lv)
    assert st["status"][5] == 2 and st["status"][6] == 1 and st["status"][7] == 1
    t = np.minimum(c0, c1)
    assert np.array_equal(st["status"] != 0, (t < 1) | (nm == len(xindex)))


def test_xtx_lookup_matches_dense_product():
    rng = np.random.default_rng(3)
    g = _rand_geno(rng, 211, 83)
    g[3, :] = 1
    rows = O.pack_genotypes(g)
    xindex = np.array([i for i in range(83) if i % 7 != 3], dtype=np.int32)
    st = O.snpstats(rows, xindex)
    lower = O.xtx_lookup(rows, xindex, st["levels"], st["status"], nthreads=3)
    sub = g[:, xindex]
    z = np.where(sub >= 0, np.take_along_axis(st["levels"], np.clip(sub, 0, 2), 1), 0.0)
    z[st["status"] != 0] = 0.0
    want = z.T @ z
    n = len(xindex)
    full, y = O.finalize_xtx(lower, n)
    assert np.allclose(full * y, want, rtol=1e-12, atol=1e-9)
    assert np.isclose(np.trace(full), n - 1, rtol=1e-13)
    # thread split does not change bits
    lower1 = O.xtx_lookup(rows, xindex, st["levels"], st["status"], nthreads=1)
    assert np.array_equal(lower, lower1)


def test_grmjunk_kat(golden):
    """POPGEN/par.example (altnormstyle NO): grm = XTX_scaled * n / trace."""
    d = os.path.join(golden, "eigenstrat")
    par = dict(genotypename="example.geno", snpname="example.snp", indivname="example.ind",
               altnormstyle="NO", numoutevec="2")
    r = O.smartpca(par, d)
    xtx = r["xtx"]
    n = xtx.shape[0]
    grm = xtx * (n / np.trace(xtx))
    want = {}
    for line in open(os.path.join(golden, "popgen", "grmjunk")):
        a, b, nsnp, v = line.split()
        want[(int(a) - 1, int(b) - 1)] = float(v)
        assert int(nsnp) == 7
    for (a, b), v in want.items():
        assert abs(grm[a, b] - v) <= 5.1e-7


@pytest.mark.parametrize("parname", ["eigenstrat", "popgen"])
def test_example_pipeline_kat(golden, parname):
    d = os.path.join(golden, "eigenstrat")
    par = O.read_par(os.path.join(d, "example.pca.par"))
    if parname == "popgen":  # POPGEN/par.example: same data, default outlier settings
        par = dict(genotypename="example.geno", snpname="example.snp",
                   indivname="example.ind", altnormstyle="NO", numoutevec="2")
        want_evec = open(os.path.join(golden, "popgen", "example.evec")).read().splitlines()
        want_eval = open(os.path.join(golden, "popgen", "example.eval")).read().splitlines()
    else:
        want_evec = open(os.path.join(d, "example.pca.evec")).read().splitlines()
        want_eval = open(os.path.join(d, "example.eval")).read().splitlines()
    r = O.smartpca(par, d)
    assert r["eval_lines"] == want_eval
    assert r["evec_lines"] == want_evec


def test_dspev_against_numpy_eigh():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((60, 60))
    a = a @ a.T
    w, v = O.dspev_lower(a)
    w2, v2 = np.linalg.eigh(a)
    assert np.allclose(w, w2[::-1], rtol=1e-12, atol=1e-10)
    assert np.all(np.diff(w) <= 0)
    v2 = v2[:, ::-1].T
    assert np.allclose(O.sign_rule(v), O.sign_rule(v2), atol=1e-9)
    assert np.allclose(a @ v.T, v.T * w, atol=1e-9)


def test_ridoutlier_planted():
    rng = np.random.default_rng(5)
    n = 200
    ev = rng.standard_normal((3, n))
    ev[1, 17] = 40.0
    ev[2, 90] = -35.0
    ev[0, 90] = 30.0
    bad, vno, sc = O.ridoutlier(ev, 3, 6.0, 0)
    z = (ev - ev.mean(1, keepdims=True)) / ev.std(1, keepdims=True)
    flagged = np.nonzero((np.abs(z) > 6.0).any(0))[0]
    assert list(bad) == list(flagged) == [17, 90]
    assert list(vno) == [1, 0]
    assert np.isclose(sc[0], z[1, 17]) and np.isclose(sc[1], z[0, 90])
    assert len(O.ridoutlier(ev[:, :2], 3, 6.0, 0)[0]) == 0  # n < 3
    assert len(O.ridoutlier(ev, 3, 6.0, 2)[0]) == 0        # mode 2 off


def test_ridoutlier_leave_one_out_hand_worked():
    """outliermode 1 (smartsubs.c:59-82): sample j's z-score uses the mean of the other
    n-1 values and their squared deviations summed over those n-1 but divided by n.
    Row 0: 200 values +-1 (mean 0) and 6.0 at index 50, so z_50 = 6 sqrt(201/200) =
    6.01497 > 6; a divisor of n-1 would give exactly 6.0 (not flagged), and the
    ordinary z of outliermode 0 gives 5.51 (not flagged either)."""
    n = 201
    ev = np.zeros((2, n))
    ev[0] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    ev[0, 50:] = np.roll(ev[0, 50:], 1)  # keep 100 of each sign around index 50
    ev[0, 50] = 6.0
    assert np.sum(np.delete(ev[0], 50)) == 0.0
    # row 1: a smaller outlier at 120 on top of a spread-out background
    ev[1] = np.linspace(-1.0, 1.0, n)
    ev[1, 120] = -9.0
    bad, vno, sc = O.ridoutlier(ev, 2, 6.0, 1)
    assert list(bad) == [50, 120] and list(vno) == [0, 1]
    assert sc[0] == pytest.approx(6.0 * np.sqrt(201 / 200), rel=1e-14)
    rest = np.delete(ev[1], 120)
    mu = rest.mean()
    z = (-9.0 - mu) / np.sqrt(np.sum((rest - mu) ** 2) / n)
    assert sc[1] == pytest.approx(z, rel=1e-13)
    # every other sample stays well inside the threshold; mode 0 flags only row 1's
    assert len(O.ridoutlier(ev, 1, 6.0, 0)[0]) == 0
    assert list(O.ridoutlier(ev, 2, 6.0, 0)[0]) == [120]


def _lsq_setup(rng, M=400, N=60, K=4, miss=0.0):
    g = _rand_geno(rng, M, N, miss)
    rows = O.pack_genotypes(g)
    xindex = np.arange(N, dtype=np.int32)
    st = O.snpstats(rows, xindex)
    ign = (st["status"] != 0).astype(np.int32)
    ffvecs = rng.standard_normal((K, M))
    ffvecs[:, ign == 1] = 0.0
    fxscal = rng.uniform(0.5, 2.0, K)
    xmean = st["mean"] * st["fancy"]
    return g, rows, st, ign, np.ascontiguousarray(ffvecs), fxscal, xmean


def test_lsqproj_matches_least_squares():
    rng = np.random.default_rng(6)
    K, M, N = 4, 400, 60
    g, rows, st, ign, ff, fx, xmean = _lsq_setup(rng, M, N, K, miss=0.2)
    a = np.zeros((K, N)); b = np.zeros((K, N)); status = np.zeros(N, np.int32)
    use = np.ones(N, np.int32)
    O.lib().orc_lsqproj(rows.reshape(-1), rows.shape[1], M, N, use, ign, xmean,
                        np.ascontiguousarray(st["fancy"]), ff.reshape(-1), fx, K,
                        a.reshape(-1), b.reshape(-1), status)
    E = (fx[:, None] * ff).T
    for q in range(N):
        obs = (g[:, q] >= 0) & (ign == 0)
        x = np.where(g[:, q] >= 0, g[:, q] * st["fancy"] - xmean, 0.0)
        sol = np.linalg.lstsq(E[obs], x[obs], rcond=None)[0]
        assert np.allclose(a[:, q], sol, rtol=1e-9, atol=1e-12)
        assert np.allclose(b[:, q], fx * (ff @ x), rtol=1e-12)
    assert np.all(status == 0)


def test_lsqproj_sparse_and_singular_rules():
    rng = np.random.default_rng(7)
    K, M, N = 3, 50, 6
    g, rows, st, ign, ff, fx, xmean = _lsq_setup(rng, M, N, K)
    g = g.copy()
    g[K + 1:, 1] = -1                       # sample 1: K+1 observed -> solvable
    g[K:, 2] = -1                           # sample 2: K observed -> ignored (kk <= K)
    rows = O.pack_genotypes(g)
    st2 = O.snpstats(rows, np.arange(N, dtype=np.int32))
    ff[:, :] = np.where(np.arange(M)[None, :] < 25, ff, 0.0)
    ff[2, :] = 0.0                          # column 2 of E is zero: Gram singular
    a = np.zeros((K, N)); b = np.zeros((K, N)); status = np.zeros(N, np.int32)
    O.lib().orc_lsqproj(rows.reshape(-1), rows.shape[1], M, N, np.ones(N, np.int32), ign,
                        st2["mean"] * st2["fancy"], np.ascontiguousarray(st2["fancy"]),
                        ff.reshape(-1), fx, K, a.reshape(-1), b.reshape(-1), status)
    assert status[2] == 1
    assert np.all(a[:, 2] == 0.0)
    # every solvable sample hits the singular Gram: regans stays at its zero start
    assert status[0] == 2 and np.all(a[:, 0] == 0.0)


def _planted_panel(seed, M, npop, per, fst=0.1, miss=0.02):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.05, 0.5, size=M)
    cols = []
    for k in range(npop):
        a = p * (1 - fst) / fst
        b = (1 - p) * (1 - fst) / fst
        pk = rng.beta(a, b)
        cols.append(rng.binomial(2, pk[:, None], size=(M, per)))
    g = np.concatenate(cols, axis=1).astype(np.int32)
    g[rng.random(g.shape) < miss] = -1
    return g


def test_fastmode_levels_match_the_normal_mode_levels():
    """setgval's table (through xmean/xfancy and /sqrt(2)) against the
    column statistics' own levels (g - mean) * fancy: the same numbers up to the
    rounding of the round trip (gval.c:73-80 vs smartpca.c:2282-2313)."""
    g = _planted_panel(3, 400, 3, 30)
    g[5, :] = -1  # no observed genotype
    g[6, :] = 2   # monomorphic
    rows = O.pack_genotypes(g)
    xindex = np.arange(g.shape[1], dtype=np.int32)
    for altnorm in (1, 0):
        st = O.snpstats(rows, xindex, altnorm=altnorm)
        lf = O.fastmode_levels(st["mean"], st["fancy"], st["status"])
        assert np.all(lf[5] == 0) and np.all(lf[:, 3] == 0)
        mean, fancy = st["mean"], st["fancy"]
        for k in range(3):
            ref = (k - mean) * fancy / np.sqrt(2.0)
            ok = st["status"] != 2
            assert np.allclose(lf[ok, k], ref[ok], rtol=1e-14, atol=1e-15)
        assert lf[6, 2] == 0.0  # an all-2 SNP contributes nothing where observed


def test_fpca_restatement_converges_to_the_exact_eigenpairs():
    """kjg_fpca's randomized block Krylov method against the exact eigen-
    decomposition of X^T X / m (its fixed point): top-K eigenvalues and
    eigenvectors of a planted 4-population panel."""
    g = _planted_panel(11, 3000, 4, 40)
    rows = O.pack_genotypes(g)
    n = g.shape[1]
    xindex = np.arange(n, dtype=np.int32)
    st = O.snpstats(rows, xindex)
    lf = O.fastmode_levels(st["mean"], st["fancy"], st["status"])
    X = O.fastmode_matrix(rows, xindex, lf, n)
    K, L, I = 3, 6, 6
    omega = np.random.default_rng(5).standard_normal((n, L))
    ev, vec = O.fpca(X, K, L, I, omega)
    w, v = np.linalg.eigh(X.T @ X / X.shape[0])
    w, v = w[::-1][:K], O.sign_rule(v[:, ::-1][:, :K].T)
    assert np.allclose(ev, w, rtol=1e-10)
    assert np.max(np.abs(vec - v)) < 1e-7
    # fewer iterations leave a visible gap: the check has teeth
    ev1, _ = O.fpca(X, K, K + 1, 1, omega[:, :K + 1])
    assert np.max(np.abs(ev1 - w) / w) > 1e-9
