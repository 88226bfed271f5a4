"""The non-lookup column path on the GPU (eig_xtx_normal, eig_lsqproject_fill) against
the oracle's restatement of getcolxz / ldregx / bumpldvv / domult_increment_normal and
of lsqproj's fill (pinned in test_oracle_normal.py).

Panels carry linkage: exact duplicate SNPs (squared correlation ~1, dropped by ldr2hi)
and near duplicates (a tenth of the genotypes redrawn, regressed out), on two
chromosomes.  Bars: counts, missing counts and keep status bit-exact; XTX within
1e-10 of its largest entry; trace scale and the GRM SNP count within 1e-10 relative;
the CLI's .eval/.evec to print precision.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def eng():
    import eig_amd
    e = eig_amd.Engine(0)
    yield e
    e.close()


def ld_panel(seed, M, sizes, miss=0.04, fst=0.06):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.08, 0.5, M)
    cols = []
    for s in sizes:
        pk = rng.beta(p * (1 - fst) / fst, (1 - p) * (1 - fst) / fst)
        cols.append(rng.binomial(2, pk[:, None], size=(M, s)))
    g = np.concatenate(cols, axis=1).astype(np.int32)
    for m in range(1, M):
        if m % 7 == 3:
            g[m] = g[m - 1]  # exact duplicate
        elif m % 7 == 5:
            g[m] = g[m - 1]
            redo = rng.random(g.shape[1]) < 0.1
            g[m, redo] = rng.binomial(2, 0.3, size=redo.sum())
    g[rng.random(g.shape) < miss] = -1
    types = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    chrom = np.where(np.arange(M) < M // 2, 1, 2).astype(np.int32)
    ppos = 1000.0 + 150.0 * np.arange(M)
    gpos = ppos * 1e-8
    return g, types, chrom, ppos, gpos


CASES = {
    "usepops": dict(usepops=True),
    "weights": dict(weights=True),
    "ldregress": dict(ld=12),
    "ldregress_window": dict(ld=30, ldposlimit=900),
    "all": dict(usepops=True, weights=True, ld=8),
}


@pytest.mark.parametrize("case", list(CASES))
def test_xtx_normal_matches_oracle(eng, case):
    opt = CASES[case]
    g, types, chrom, ppos, gpos = ld_panel(11, 2400, [60, 45, 70, 40])
    N, M = g.shape[1], g.shape[0]
    rows = O.pack_genotypes(g)
    xi = np.array([i for i in range(N) if i % 11 != 4], dtype=np.int32)
    xt = types[xi]
    w = np.random.default_rng(2).uniform(0.3, 2.0, M) if opt.get("weights") else None
    ld = opt.get("ld", 0)
    lposl = opt.get("ldposlimit", 1000000000)
    _say(f"xtx_normal {case} N={len(xi)} M={M}")
    eng.load_packed(rows, N)
    eng.select(xi)
    eng.snpstats(fetch=False)
    y, yn, st, X = eng.xtx_normal(xtypes=xt if opt.get("usepops") else None, npop=4, weight=w,
                                  ldregress=ld, chrom=chrom, ppos=ppos, gpos=gpos,
                                  ldposlimit=lposl)
    nc = O.normal_columns(rows, xi, xt if opt.get("usepops") else None, w)
    for k in ("status", "c0", "c1", "nmiss"):
        assert np.array_equal(st[k], nc[k]), k
    kept = np.nonzero(nc["status"] == 0)[0]
    if ld:
        Z, ynum = O.ldregress_columns(nc["cols"], kept, chrom, ppos, gpos, ld, ldposlimit=lposl)
        ndrop = int(np.sum(~np.any(Z != 0, axis=1)))
        _say(f"{case}: {ndrop} of {len(kept)} SNPs dropped, ynumsnps {ynum:.6f} (gpu {yn:.6f})")
        assert ndrop >= len(kept) // 20  # most of the exact duplicates
    else:
        Z, ynum = nc["cols"][kept], float(len(kept))
    ref, yr = O.dense_xtx(Z)
    assert abs(y - yr) <= 1e-10 * yr
    assert abs(yn - ynum) <= 1e-10 * ynum
    assert np.array_equal(X, X.T)
    err = np.max(np.abs(X - ref)) / np.max(np.abs(ref))
    _say(f"{case}: XTX max rel. error {err:.2e}")
    assert err <= 1e-10


def test_normal_path_feeds_loadings_and_lsqproject_fill(eng):
    g, types, chrom, ppos, gpos = ld_panel(12, 1800, [50, 50, 50])
    N, M = g.shape[1], g.shape[0]
    rows = O.pack_genotypes(g)
    xi = np.arange(N, dtype=np.int32)
    _say(f"normal path + lsqproject fill N={N} M={M}")
    eng.load_packed(rows, N)
    eng.select(xi)
    eng.snpstats(fetch=False)
    eng.xtx_normal(xtypes=types, npop=3, want_matrix=False)
    ev, vec = eng.eigh_xtx(3)
    nc = O.normal_columns(rows, xi, types)
    ref, _ = O.dense_xtx(nc["cols"][nc["status"] == 0])
    evr, vr = O.dspev_lower(ref)
    assert np.max(np.abs(ev - evr)) <= 1e-10 * evr[0]
    K = 3
    ign = (nc["status"] != 0).astype(np.int32)
    ff, fx = eng.loadings(vec, ignore=ign.astype(np.uint8))
    off = np.zeros((K, M))
    L = O.lib()
    L.orc_loadings(rows.reshape(-1), rows.shape[1], M, xi, N, ign, 1, 1, np.ascontiguousarray(vec),
                   K, off.reshape(-1))
    assert np.max(np.abs(ff - off)) <= 1e-11 * np.max(np.abs(off))
    xm = nc["mean"] * nc["fancy"]
    fancy = np.ascontiguousarray(nc["fancy"])
    fxv = np.zeros((K, N)); fxr = np.zeros(K)
    L.orc_fxscal(rows.reshape(-1), rows.shape[1], M, xi, N, ign, xm, fancy, off.reshape(-1), K,
                 fxv.reshape(-1), fxr)
    assert np.allclose(fx, fxr, rtol=1e-11)
    use = np.ones(N, np.int32)
    use[::17] = 0
    itype = types.copy()
    itype[5] = -2  # an individual without a type: no fill
    a, b, s = eng.lsqproject_fill(use, itype, 3, K)
    ra = np.zeros((K, N)); rb = np.zeros((K, N)); rs = np.zeros(N, np.int32)
    L.orc_lsqproj_fill(rows.reshape(-1), rows.shape[1], M, N, use, itype, 3, ign, xm, fancy,
                       off.reshape(-1), fxr, K, ra.reshape(-1), rb.reshape(-1), rs)
    assert np.array_equal(s, rs)
    ok = rs == 0
    assert np.allclose(a[:, ok], ra[:, ok], rtol=1e-9, atol=1e-12)
    assert np.allclose(b[:, ok], rb[:, ok], rtol=1e-9, atol=1e-12)


def _write_panel(prefix, g, types, chrom, ppos):
    from eig_amd import synth
    M, N = g.shape
    inds = ["I%05d" % i for i in range(N)]
    pops = ["POP%d" % t for t in types]
    snps = ["rs%06d" % m for m in range(M)]
    with open(prefix + ".ind", "w") as f:
        for i, p in zip(inds, pops):
            f.write("%20s U %10s\n" % (i, p))
    with open(prefix + ".snp", "w") as f:
        for m in range(M):
            f.write("%20s %3d %12.6f %12d A G\n" % (snps[m], chrom[m], 0.0, int(ppos[m])))
    rows = O.pack_genotypes(g)
    rlen = rows.shape[1]
    hdr = ("GENO %7d %7d %x %x" % (N, M, synth.hasharr(inds), synth.hasharr(snps))).encode()
    with open(prefix + ".geno", "wb") as f:
        f.write(hdr.ljust(rlen, b"\0"))
        f.write(rows.tobytes())
    return snps


@pytest.mark.parametrize("case", ["usepops", "ldregress", "weights_usepops_ld"])
def test_cli_normal_path_matches_oracle(tmp_path, case):
    from test_cli import _parse_evec, run_cli, write_par
    g, types, chrom, ppos, gpos = ld_panel(13, 2100, [40, 55, 45])
    prefix = str(tmp_path / "ld")
    snps = _write_panel(prefix, g, types, chrom, ppos)
    par = dict(genotypename=prefix + ".geno", snpname=prefix + ".snp", indivname=prefix + ".ind",
               evecoutname="out.evec", evaloutname="out.eval", numoutevec=4, numoutlieriter=2,
               numoutlierevec=3, grmoutname="grm")
    if "usepops" in case:
        par["usepopsformissing"] = "YES"
    if "ld" in case:
        par.update(ldregress=10, ldposlimit=2000)
    if "weights" in case:
        w = np.random.default_rng(4).uniform(0.5, 1.5, len(snps))
        with open(tmp_path / "w.txt", "w") as f:
            f.write("# weights\n")
            for s_, x in zip(snps, w):
                f.write(f"{s_} {x:.6f}\n")
        par["weightname"] = str(tmp_path / "w.txt")
    write_par(tmp_path / "par", **par)
    _say(f"cli normal path {case}")
    run_cli(str(tmp_path / "par"), tmp_path)
    ref = O.smartpca({k: str(v) for k, v in par.items()}, str(tmp_path))
    got_eval = np.array([float(x) for x in (tmp_path / "out.eval").read_text().split()])
    want_eval = np.array([float(x) for x in ref["eval_lines"]])
    assert len(got_eval) == len(want_eval)
    assert np.max(np.abs(got_eval - want_eval)) <= 1.01e-6
    gh, grows = _parse_evec((tmp_path / "out.evec").read_text())
    wh, wrows = _parse_evec("\n".join(ref["evec_lines"]))
    assert np.allclose(gh, wh, atol=1.01e-3)
    assert [r[0] for r in grows] == [r[0] for r in wrows]
    gg = np.array([r[1] for r in grows])[:, :2]
    ww = np.array([r[1] for r in wrows])[:, :2]
    assert np.max(np.abs(gg - ww)) <= 2.01e-4
    grm = [l.split() for l in (tmp_path / "grm").read_text().splitlines()]
    assert {int(r[2]) for r in grm} == {int(ref["ynumsnps"])}
