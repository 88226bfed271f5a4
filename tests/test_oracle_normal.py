"""Pins of the oracle's non-lookup column path (usepopsformissing, SNP weights,
ldregress) and of its lsqproject fill, on the CPU.

Each check compares the oracle with something it does not restate: the lookup
path's levels (already pinned by the reference's own example outputs), hand-worked
columns, numpy least squares, or an exact-duplicate panel whose answer is known.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _panel(seed, M, sizes, miss=0.03, fst=0.08):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.1, 0.5, M)
    cols = []
    for s in sizes:
        pk = rng.beta(p * (1 - fst) / fst, (1 - p) * (1 - fst) / fst)
        cols.append(rng.binomial(2, pk[:, None], size=(M, s)))
    g = np.concatenate(cols, axis=1).astype(np.int32)
    g[rng.random(g.shape) < miss] = -1
    types = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    return g, types


def test_columns_without_fill_or_weights_are_the_lookup_levels():
    g, types = _panel(1, 400, [30, 40, 25])
    rows = O.pack_genotypes(g)
    xi = np.arange(g.shape[1], dtype=np.int32)
    st = O.snpstats(rows, xi)
    nc = O.normal_columns(rows, xi)
    assert np.array_equal(nc["status"], st["status"])
    assert np.array_equal(nc["c0"], st["c0"]) and np.array_equal(nc["nmiss"], st["nmiss"])
    kept = st["status"] == 0
    lev = st["levels"]
    want = np.where(g >= 0, np.take_along_axis(lev[:, :3], np.maximum(g, 0), axis=1), 0.0)
    assert np.array_equal(nc["cols"][kept], want[kept])
    assert np.array_equal(nc["levels"][kept, :3], lev[kept, :3])
    # and the dense XTX equals the partial-sum lookup XTX
    full, y = O.dense_xtx(nc["cols"][kept])
    ref, yr = O.finalize_xtx(O.xtx_lookup(rows, xi, lev, st["status"]), len(xi))
    assert abs(y - yr) <= 1e-12 * yr
    assert np.max(np.abs(full - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_usepopsformissing_fill_hand_worked():
    # one SNP, pops A = [0, 1, 2, miss], B = [2, 2, miss, miss], C = [miss, miss]
    g = np.array([[0, 1, 2, -1, 2, 2, -1, -1, -1, -1]], np.int32)
    types = np.array([0, 0, 0, 0, 1, 1, 1, 1, 2, 2], np.int32)
    rows = O.pack_genotypes(g)
    xi = np.arange(10, dtype=np.int32)
    nc = O.normal_columns(rows, xi, xtypes=types)
    y = np.array([0, 1, 2, 1, 2, 2, 2, 2], float)  # filled; C stays missing
    mean = y.sum() / 8
    p = 0.5 * mean
    fancy = 1.0 / math.sqrt(p * (1 - p))
    want = np.concatenate([(y - mean) * fancy, [0.0, 0.0]])
    assert nc["status"][0] == 0 and nc["nmiss"][0] == 2 and nc["c0"][0] == 7
    assert np.allclose(nc["cols"][0], want, rtol=0, atol=1e-15)
    assert nc["mean"][0] == mean and nc["fancy"][0] == fancy
    # getcolxf's plain levels for the loadings: mean over the 5 observed genotypes
    pm = 7 / 5
    pf = 1.0 / math.sqrt(0.5 * pm * (1 - 0.5 * pm))
    assert np.allclose(nc["levels"][0, :3], (np.arange(3) - pm) * pf, rtol=0, atol=1e-15)
    # a maxmissing of 1 keeps the SNP without the fill's help only when filled
    assert O.normal_columns(rows, xi, xtypes=types, maxmissing=1)["status"][0] == 1
    assert O.normal_columns(rows, xi, xtypes=types, maxmissing=2)["status"][0] == 0


def test_weights_scale_columns():
    g, _ = _panel(2, 300, [50, 50])
    rows = O.pack_genotypes(g)
    xi = np.arange(100, dtype=np.int32)
    w = np.random.default_rng(3).uniform(0.2, 3.0, 300)
    a = O.normal_columns(rows, xi)
    b = O.normal_columns(rows, xi, weight=w)
    assert np.array_equal(b["status"], a["status"])
    assert np.array_equal(b["cols"], a["cols"] * w[:, None])


def test_ldregress_drops_exact_duplicates():
    """Every SNP twice in a row, no regressions (ldr2lo above any correlation): a second
    copy is dropped exactly when its squared correlation with the first, (z.z / 2(n-1))^2,
    exceeds ldr2hi; the XTX is then the lookup XTX of the panel without those copies."""
    g, _ = _panel(4, 300, [60, 60, 60], miss=0.02)
    gd = np.repeat(g, 2, axis=0)
    rows = O.pack_genotypes(gd)
    xi = np.arange(g.shape[1], dtype=np.int32)
    n = len(xi)
    nc = O.normal_columns(rows, xi)
    kept = np.nonzero(nc["status"] == 0)[0]
    z2 = np.einsum("ij,ij->i", nc["cols"], nc["cols"])
    drop = np.array([m % 2 == 1 and (z2[m] / (2 * (n - 1))) ** 2 > 0.95 for m in kept])
    assert drop.sum() >= len(kept) // 4
    M2 = gd.shape[0]
    chrom = np.ones(M2, np.int32); ppos = np.arange(M2) * 100.0; gpos = ppos * 1e-8
    res, ynum = O.ldregress_columns(nc["cols"], kept, chrom, ppos, gpos, 5, r2lo=10.0, r2hi=0.95)
    assert ynum == float((~drop).sum())
    assert np.array_equal(np.any(res != 0, axis=1), ~drop)
    full, _ = O.dense_xtx(res)
    rows1 = O.pack_genotypes(gd[kept[~drop]])
    st1 = O.snpstats(rows1, xi)
    ref, _ = O.finalize_xtx(O.xtx_lookup(rows1, xi, st1["levels"], st1["status"]), n)
    assert np.max(np.abs(full - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_ldregress_residuals_are_least_squares_on_the_reference_window():
    rng = np.random.default_rng(5)
    n, M, W = 80, 40, 4
    base = rng.standard_normal((M, n))
    cols = base.copy()
    for m in range(1, M):  # correlated neighbours
        cols[m] = 0.8 * cols[m - 1] + 0.6 * base[m]
    cols -= cols.mean(axis=1, keepdims=True)
    chrom = np.array([1] * 25 + [2] * 15, np.int32)
    ppos = np.arange(M) * 100.0
    gpos = ppos * 1e-8
    kept = np.arange(M)
    res, ynum = O.ldregress_columns(cols, kept, chrom, ppos, gpos, W, ldposlimit=250,
                                    r2lo=-1.0, r2hi=1e9)
    want_y = 0.0
    for k in range(M):
        # the buffer after SNP k-1: at most W SNPs of SNP k's chromosome, each within
        # 250 bases of SNP k-1
        win = [j for j in range(max(0, k - W), k) if chrom[j] == chrom[k]
               and ppos[k - 1] - ppos[j] <= 250]
        if win:
            S = cols[win].T
            coef, *_ = np.linalg.lstsq(S, cols[k], rcond=None)
            r = cols[k] - S @ coef
        else:
            r = cols[k]
        assert np.allclose(res[k], r, rtol=0, atol=1e-10), k
        want_y += r @ r / (cols[k] @ cols[k])
    assert abs(ynum - want_y) <= 1e-10 * want_y


def test_lsqproj_fill_reduces_and_matches_lstsq():
    g, types = _panel(6, 500, [30, 30, 30], miss=0.1)
    N = g.shape[1]
    rows = O.pack_genotypes(g)
    xi = np.arange(N, dtype=np.int32)
    st = O.snpstats(rows, xi)
    ign = (st["status"] != 0).astype(np.int32)
    K = 3
    rng = np.random.default_rng(7)
    ff = rng.standard_normal((K, 500)) * (ign == 0)
    fx = rng.uniform(0.5, 2.0, K)
    xm = st["mean"] * st["fancy"]
    fancy = np.ascontiguousarray(st["fancy"])
    use = np.ones(N, np.int32)
    L = O.lib()

    def run(fill, itype):
        a = np.zeros((K, N)); b = np.zeros((K, N)); s = np.zeros(N, np.int32)
        if fill:
            L.orc_lsqproj_fill(rows.reshape(-1), rows.shape[1], 500, N, use, itype, 3, ign, xm,
                               fancy, ff.reshape(-1), fx, K, a.reshape(-1), b.reshape(-1), s)
        else:
            L.orc_lsqproj(rows.reshape(-1), rows.shape[1], 500, N, use, ign, xm, fancy,
                          ff.reshape(-1), fx, K, a.reshape(-1), b.reshape(-1), s)
        return a, b, s
    # without types the fill changes nothing
    a0, b0, s0 = run(False, None)
    a1, b1, s1 = run(True, np.full(N, -1, np.int32))
    assert np.array_equal(s0, s1) and np.array_equal(a0, a1) and np.array_equal(b0, b1)
    # with types: individual q's filled row solved by numpy
    a2, b2, s2 = run(True, types)
    kept = ign == 0
    E = (fx[:, None] * ff).T  # M x K
    for q in (0, 31, 77):
        raw = g[:, q].astype(float)
        for m in np.nonzero((raw < 0) & kept)[0]:
            obs = (g[m] >= 0) & (types == types[q])
            if obs.sum() > 0:
                raw[m] = g[m][obs].sum() / obs.sum()
        sel = (raw >= 0) & kept
        x = raw[sel] * st["fancy"][sel] - xm[sel]
        coef, *_ = np.linalg.lstsq(E[sel], x, rcond=None)
        assert s2[q] == 0
        assert np.allclose(a2[:, q], coef, rtol=1e-9, atol=1e-12)
