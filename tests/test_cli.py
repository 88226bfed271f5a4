"""The smartpca drop-in end to end: same parfile, same inputs, same .evec/.eval.

GPU tests run the CLI (eig_amd/bin/smartpca, C host + libeig_amd.so) and compare
its files with the reference's own example outputs (byte for byte) and with the
CPU oracle's pipeline on seeded synthetic panels (numeric tolerance stated per
test).  CPU tests cover the fault paths that end before any device call and the
exported C-ABI.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "eig_amd", "bin", "smartpca")
HDR = os.path.join(ROOT, "include", "eig_abi.h")


def run_cli(par_path, cwd, expect_ok=True):
    r = subprocess.run([CLI, "-p", par_path], cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    if expect_ok:
        assert r.returncode == 0, r.stderr
    return r


def write_par(path, **kv):
    with open(path, "w") as f:
        for k, v in kv.items():
            f.write(f"{k}: {v}\n")


def _copy_example(golden, tmp_path):
    for f in ("example.geno", "example.snp", "example.ind"):
        shutil.copy(os.path.join(golden, "eigenstrat", f), tmp_path / f)
    d = os.path.join(golden, "convertf")
    shutil.copy(os.path.join(d, "example.packedancestrymapgeno"), tmp_path / "packed.geno")
    shutil.copy(os.path.join(d, "example.snp"), tmp_path / "packed.snp")
    shutil.copy(os.path.join(d, "example.ind"), tmp_path / "packed.ind")
    shutil.copy(os.path.join(d, "example.bed"), tmp_path / "example.bed")
    shutil.copy(os.path.join(d, "example.pedsnp"), tmp_path / "example.bim")
    shutil.copy(os.path.join(d, "example.pedind"), tmp_path / "example.fam")


# ------------------------------------------------------------------ CPU tests
def test_cli_and_library_are_built():
    assert os.path.exists(CLI), "run python -m eig_amd.build"
    r = subprocess.run([CLI], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 0 and "no parameters" in r.stdout + r.stderr


def test_library_exports_every_declared_symbol():
    import ctypes
    import eig_amd
    lib = ctypes.CDLL(eig_amd.LIB_PATH)
    decl = re.findall(r"^\w[\w\s\*]*?\b(eig_\w+)\s*\(", open(HDR).read(), re.M)
    assert len(decl) >= 14
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) <= set(eig_amd.SIGNATURES), set(decl) - set(eig_amd.SIGNATURES)


def test_duplicate_parameter_is_fatal(tmp_path):
    p = tmp_path / "par"
    p.write_text("genotypename: a\ngenotypename: b\n")
    r = run_cli(str(p), tmp_path, expect_ok=False)
    assert r.returncode != 0 and "duplicate parameter" in r.stderr


@pytest.mark.parametrize("key,val", [("killr2", "YES"), ("xregionname", "x.txt"),
                                     ("shrinkmode", "YES"), ("hwfilter", "0.001"),
                                     ("popsizelimit", "20")])
def test_unimplemented_reference_parameter_is_fatal(tmp_path, key, val):
    """A reference key that would change the PCA and that this build does not
    implement stops the run before any input is read, instead of being ignored."""
    p = tmp_path / "par"
    p.write_text(f"genotypename: a\nsnpname: b\nindivname: c\n{key}: {val}\n")
    r = run_cli(str(p), tmp_path, expect_ok=False)
    assert r.returncode != 0 and f"{key}: is not supported" in r.stderr


def test_reference_default_of_an_unimplemented_parameter_is_accepted(golden, tmp_path):
    """Setting such a key to the reference's default is not an error: the run goes on
    to read the inputs (here it stops at the deliberately truncated genotype file)."""
    _copy_example(golden, tmp_path)
    raw = (tmp_path / "packed.geno").read_bytes()
    (tmp_path / "packed.geno").write_bytes(raw[:-20])
    write_par(tmp_path / "par", genotypename="packed.geno", snpname="packed.snp",
              indivname="packed.ind", killr2="NO", shrinkmode="NO", popsizelimit=-1)
    r = run_cli(str(tmp_path / "par"), tmp_path, expect_ok=False)
    assert r.returncode != 0 and "length mismatch" in r.stderr


def test_parfile_macros_are_substituted_longest_first(tmp_path):
    """getpars.c:378-449: upper-case keys are macros replaced in the other values,
    longest name first (so D2 is not read as D followed by "2")."""
    (tmp_path / "par").write_text(
        "D: /nonexistent-d\n"
        f"D2: {tmp_path}/sub\n"
        "genotypename: D2/x.geno\n"
        "snpname: D2/x.snp\n"
        "indivname: D/x.ind\n")
    r = run_cli(str(tmp_path / "par"), tmp_path, expect_ok=False)
    out = r.stdout + r.stderr
    assert r.returncode != 0
    assert f"genotypename: {tmp_path}/sub/x.geno" in out
    assert "indivname: /nonexistent-d/x.ind" in out
    assert f"can't open {tmp_path}/sub/x.snp" in out


def test_geno_hash_mismatch_is_fatal(golden, tmp_path):
    _copy_example(golden, tmp_path)
    lines = (tmp_path / "packed.ind").read_text().splitlines()
    lines[0] = lines[0].replace("SAMPLE0", "SAMPLEX")
    (tmp_path / "packed.ind").write_text("\n".join(lines) + "\n")
    write_par(tmp_path / "par", genotypename="packed.geno", snpname="packed.snp",
              indivname="packed.ind")
    r = run_cli(str(tmp_path / "par"), tmp_path, expect_ok=False)
    assert r.returncode != 0 and "indiv file has changed" in r.stderr


def test_truncated_geno_is_fatal(golden, tmp_path):
    _copy_example(golden, tmp_path)
    raw = (tmp_path / "packed.geno").read_bytes()
    (tmp_path / "packed.geno").write_bytes(raw[:-20])
    write_par(tmp_path / "par", genotypename="packed.geno", snpname="packed.snp",
              indivname="packed.ind")
    r = run_cli(str(tmp_path / "par"), tmp_path, expect_ok=False)
    assert r.returncode != 0 and "length mismatch" in r.stderr


def test_unordered_snps_with_packed_input_is_fatal(golden, tmp_path):
    _copy_example(golden, tmp_path)
    lines = (tmp_path / "packed.snp").read_text().splitlines()
    lines[1], lines[2] = lines[2], lines[1]
    (tmp_path / "packed.snp").write_text("\n".join(lines) + "\n")
    write_par(tmp_path / "par", genotypename="packed.geno", snpname="packed.snp",
              indivname="packed.ind", hashcheck="NO")
    r = run_cli(str(tmp_path / "par"), tmp_path, expect_ok=False)
    assert r.returncode != 0 and "out of order" in r.stderr


# ------------------------------------------------------------------ GPU tests
@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["eigenstrat", "packed", "bed"])
def test_cli_reproduces_reference_example(golden, tmp_path, fmt):
    _copy_example(golden, tmp_path)
    inputs = {"eigenstrat": ("example.geno", "example.snp", "example.ind"),
              "packed": ("packed.geno", "packed.snp", "packed.ind"),
              "bed": ("example.bed", "example.bim", "example.fam")}[fmt]
    write_par(tmp_path / "par", genotypename=inputs[0], snpname=inputs[1], indivname=inputs[2],
              evecoutname="out.evec", evaloutname="out.eval", altnormstyle="NO",
              numoutevec=2, numoutlieriter=5, numoutlierevec=2, outliersigmathresh=6.0,
              familynames="NO", grmoutname="grm")
    print(f"[gpu-case] cli example {fmt}", flush=True)
    run_cli(str(tmp_path / "par"), tmp_path)
    d = os.path.join(golden, "eigenstrat")
    # the null eigenvalue of the centred matrix is a rounding-level zero whose sign
    # is noise (LAPACK dspev landed on -1e-16): compare it up to that sign
    def nozero(t):
        return t.replace("-0.000000", " 0.000000")
    assert nozero((tmp_path / "out.eval").read_text()) == \
        nozero(open(os.path.join(d, "example.eval")).read())
    assert (tmp_path / "out.evec").read_text() == open(os.path.join(d, "example.pca.evec")).read()
    got = [l.split() for l in (tmp_path / "grm").read_text().splitlines()]
    want = [l.split() for l in open(os.path.join(golden, "popgen", "grmjunk"))]
    assert [g[:3] for g in got] == [w[:3] for w in want]
    assert np.allclose([float(g[3]) for g in got], [float(w[3]) for w in want], atol=1.1e-6)
    assert (tmp_path / "grm.id").read_text() == open(os.path.join(golden, "popgen", "grmjunk.id")).read()


def _synthetic(tmp_path, nsnp, scale, cfg=1, plant=None, miss_rows=None):
    from eig_amd import synth
    P = synth.config_panel(cfg, nsnp=nsnp, scale=scale)
    prefix = str(tmp_path / "syn")
    synth.write_files(P, prefix)
    if plant is not None:
        # replace individual `plant` by an extreme genotype vector (all hom-alt)
        raw = bytearray(open(prefix + ".geno", "rb").read())
        rlen = max(48, (2 * P.nind + 7) // 8)
        for m in range(P.nsnp):
            off = rlen * (m + 1) + plant // 4
            sh = 2 * (3 - plant % 4)
            raw[off] = (raw[off] & ~(3 << sh)) | (2 << sh)
        open(prefix + ".geno", "wb").write(bytes(raw))
    return P, prefix


def _parse_evec(text):
    lines = text.splitlines()
    head = [float(x) for x in lines[0].split()[1:]]
    rows = [(l.split()[0], [float(x) for x in l.split()[1:-1]], l.split()[-1]) for l in lines[1:]]
    return head, rows


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "bf16x4"])
def test_cli_matches_oracle_on_synthetic_panel(tmp_path, form):
    P, prefix = _synthetic(tmp_path, 6000, 0.3)
    par = dict(genotypename=prefix + ".geno", snpname=prefix + ".snp",
               indivname=prefix + ".ind", evecoutname="out.evec", evaloutname="out.eval",
               numoutevec=10, numoutlieriter=0, xtxform=form)
    write_par(tmp_path / "par", **par)
    print(f"[gpu-case] cli synthetic N={P.nind} M={P.nsnp} form={form}", flush=True)
    run_cli(str(tmp_path / "par"), tmp_path)
    ref = O.smartpca({k: str(v) for k, v in par.items()}, str(tmp_path))
    got_eval = np.array([float(x) for x in (tmp_path / "out.eval").read_text().split()])
    want_eval = np.array([float(x) for x in ref["eval_lines"]])
    assert np.max(np.abs(got_eval - want_eval)) <= 2e-6
    gh, grows = _parse_evec((tmp_path / "out.evec").read_text())
    wh, wrows = _parse_evec("\n".join(ref["evec_lines"]))
    assert np.allclose(gh, wh, atol=1.01e-3)
    assert [r[0] for r in grows] == [r[0] for r in wrows]
    assert [r[2] for r in grows] == [r[2] for r in wrows]
    diff = np.max(np.abs(np.array([r[1] for r in grows]) - np.array([r[1] for r in wrows])))
    assert diff <= 2.01e-4, diff  # %10.4f printing: at most a last-digit flip


@pytest.mark.gpu
def test_cli_poplist_axis_and_projection_match_oracle(tmp_path):
    """poplistname (egsubs.c:86-109, smartpca.c:774-892): the PCA axes come from the
    listed populations only; every other sample is projected by lsqproject.  .eval and
    .evec (axis and projected rows) against the oracle, to print precision."""
    P, prefix = _synthetic(tmp_path, 4000, 0.2)
    (tmp_path / "pops.txt").write_text("P01\nP03\n")
    par = dict(genotypename=prefix + ".geno", snpname=prefix + ".snp",
               indivname=prefix + ".ind", evecoutname="out.evec", evaloutname="out.eval",
               numoutevec=4, numoutlieriter=0, poplistname=str(tmp_path / "pops.txt"))
    write_par(tmp_path / "par", **par)
    print(f"[gpu-case] cli poplist N={P.nind} M={P.nsnp}", flush=True)
    run_cli(str(tmp_path / "par"), tmp_path)
    ref = O.smartpca({k: str(v) for k, v in par.items()}, str(tmp_path))
    assert len(ref["xindex"]) < P.nind  # the axes use a subset
    got_eval = np.array([float(x) for x in (tmp_path / "out.eval").read_text().split()])
    want_eval = np.array([float(x) for x in ref["eval_lines"]])
    assert len(got_eval) == len(want_eval)
    assert np.max(np.abs(got_eval - want_eval)) <= 1.01e-6
    gh, grows = _parse_evec((tmp_path / "out.evec").read_text())
    wh, wrows = _parse_evec("\n".join(ref["evec_lines"]))
    assert np.allclose(gh, wh, atol=1.01e-3)
    assert [r[0] for r in grows] == [r[0] for r in wrows]
    assert len(grows) == P.nind  # projected samples are written too
    diff = np.max(np.abs(np.array([r[1] for r in grows]) - np.array([r[1] for r in wrows])))
    assert diff <= 2.01e-4, diff


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_cli_outlier_removal_matches_oracle(tmp_path, mode):
    """Outlier iterations with the ordinary z-score (outliermode 0) and the
    leave-one-out z-score (outliermode 1, smartsubs.c:59-82)."""
    P, prefix = _synthetic(tmp_path, 3000, 0.15, plant=7)
    par = dict(genotypename=prefix + ".geno", snpname=prefix + ".snp",
               indivname=prefix + ".ind", evecoutname="out.evec", evaloutname="out.eval",
               outlieroutname="out.outl", numoutevec=4, numoutlieriter=3, numoutlierevec=4,
               outliersigmathresh=6.0, outliermode=mode)
    write_par(tmp_path / "par", **par)
    print(f"[gpu-case] cli outliers N={P.nind} M={P.nsnp} outliermode={mode}", flush=True)
    run_cli(str(tmp_path / "par"), tmp_path)
    ref = O.smartpca({k: str(v) for k, v in par.items()}, str(tmp_path))
    got = (tmp_path / "out.outl").read_text().splitlines()
    want = ["REMOVED outlier %s iter %d evec %d sigmage %.3f pop: %s" % r for r in ref["removed"]]
    assert len(want) >= 1
    assert [g.split()[:7] for g in got] == [w.split()[:7] for w in want]
    # the score's sign follows the eigenvector's sign (LAPACK's in the reference, the
    # sign rule here): magnitudes are compared
    assert np.allclose([abs(float(g.split()[8])) for g in got],
                       [abs(float(w.split()[8])) for w in want], atol=2e-3)
    ids = [l.split()[0] for l in (tmp_path / "out.evec").read_text().splitlines()[1:]]
    assert P.ids()[0][7] not in ids


@pytest.mark.gpu
def test_cli_grm_binary_snp_weights_and_best_snps_match_oracle(tmp_path):
    """dumpgrmbin (smartpca.c:3705-3767): float32 XTX[a][b] / (trace/nrows) over the
    lower triangle and the SNP count as int32 in .N.bin; writesnpeigs
    (smartpca.c:3038-3104): each axis's loadings scaled to mean square 1 over the
    pass's SNPs, the 10 largest |weights| per axis as 'eigbestsnp' lines, then one
    line per kept SNP.  Expected values from the oracle's XTX and loadings."""
    P, prefix = _synthetic(tmp_path, 3000, 0.15)
    par = dict(genotypename=prefix + ".geno", snpname=prefix + ".snp",
               indivname=prefix + ".ind", evecoutname="out.evec", evaloutname="out.eval",
               numoutevec=3, numoutlieriter=0, grmoutname="grm", grmbinary="YES",
               snpweightoutname="out.snpw")
    write_par(tmp_path / "par", **par)
    print(f"[gpu-case] cli grm binary + snp weights N={P.nind} M={P.nsnp}", flush=True)
    out = run_cli(str(tmp_path / "par"), tmp_path).stdout
    ref = O.smartpca({k: str(v) for k, v in par.items()}, str(tmp_path))
    n = len(ref["xindex"])
    X = ref["xtx"]
    want = (X[np.tril_indices(n)] / (np.trace(X) / n)).astype(np.float32)
    got = np.fromfile(tmp_path / "grm.bin", dtype=np.float32)
    assert got.shape == want.shape
    # one float32 rounding apart at most
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-3)) <= 2.4e-7
    cnt = np.fromfile(tmp_path / "grm.N.bin", dtype=np.int32)
    assert cnt.shape == want.shape and np.all(cnt == int(ref["ynumsnps"]))
    # SNP weights
    snp = O.read_snp(prefix + ".snp")
    W = ref["ffvecs"] * np.sqrt(ref["ffvecs"].shape[1] / np.sum(ref["ffvecs"] ** 2, axis=1))[:, None]
    keep = ref["snp_ignore_final"] == 0
    cols = ref["cols"]
    lines = (tmp_path / "out.snpw").read_text().splitlines()
    assert [l.split()[0] for l in lines] == [snp["ids"][m] for m in cols[keep]]
    assert [int(l.split()[1]) for l in lines] == [int(snp["chrom"][m]) for m in cols[keep]]
    gw = np.array([[float(x) for x in l.split()[3:]] for l in lines])
    assert np.max(np.abs(gw - W[:, keep].T)) <= 1.01e-3
    best = re.findall(r"^eigbestsnp\s+(\d+)\s+(\S+)\s+\d+\s+\d+\s+(\S+)$", out, re.M)
    assert len(best) == 30
    for j in range(3):
        a = np.abs(W[j]) * keep - (~keep)
        top = np.argsort(-a, kind="stable")[:10]
        assert [b[1] for b in best if int(b[0]) == j + 1] == [snp["ids"][cols[t]] for t in top]
        assert np.allclose([float(b[2]) for b in best if int(b[0]) == j + 1], a[top], atol=1.01e-3)


@pytest.mark.gpu
def test_cli_fastmode_writes_the_randomized_pca(tmp_path):
    """fastmode: YES -> .evec only (header eigenvalues S^2/ncols, unit-norm rows);
    checked against the exact eigenpairs of X^T X / ncols that the randomized
    method converges to (parity with the reference's GSL start panel is unpinned)."""
    P, prefix = _synthetic(tmp_path, 4000, 0.1)
    par = dict(genotypename=prefix + ".geno", snpname=prefix + ".snp",
               indivname=prefix + ".ind", evecoutname="out.evec", evaloutname="out.eval",
               numoutevec=2, fastmode="YES", fastiter=8, seed=77)
    write_par(tmp_path / "par", **par)
    print(f"[gpu-case] cli fastmode N={P.nind} M={P.nsnp}", flush=True)
    r = run_cli(str(tmp_path / "par"), tmp_path)
    assert "end of smartpca(fastmode)" in r.stdout
    assert not (tmp_path / "out.eval").exists()
    head, rows = _parse_evec((tmp_path / "out.evec").read_text())
    rows_g = P.packed()
    xindex = np.arange(P.nind, dtype=np.int32)
    st = O.snpstats(rows_g, xindex)
    lf = O.fastmode_levels(st["mean"], st["fancy"], st["status"])
    X = O.fastmode_matrix(rows_g, xindex, lf, P.nind)
    w, v = np.linalg.eigh(X.T @ X / X.shape[0])
    w, v = w[::-1][:2], O.sign_rule(v[:, ::-1][:, :2].T)
    assert np.allclose(head, w, atol=1.01e-3)
    got = np.array([r_[1] for r_ in rows]).T
    assert np.max(np.abs(got - v)) <= 1.01e-4 + 1e-6
    assert [r_[0] for r_ in rows] == P.ids()[0]


@pytest.mark.gpu
def test_cli_population_statistics_match_the_reference_log(golden, tmp_path):
    """POPGEN's example (par.example over the CONVERTF genotypes, here read from their
    .bed conversion): the anova p-value of each eigenvector, the population-difference
    chi-square and its p-value, the sample-coordinate kurtosis and the population sizes
    printed by the reference (POPGEN/example.log, kept as example_log.txt); and the
    Tracy-Widom block, whose statistics the tests of stats.c pin against the
    reference's twstats.  The SNP-loading kurtosis column of that log (smartpca
    version 13050) differs from what the current source computes
    (kurtosis(ffvecs), smartpca.c:1486-1501, version 18140): it is checked against
    scipy's kurtosis of the oracle's loadings instead."""
    _copy_example(golden, tmp_path)
    write_par(tmp_path / "par", genotypename="example.bed", snpname="example.bim",
              indivname="example.fam", evecoutname="out.evec", evaloutname="out.eval",
              altnormstyle="NO", numoutevec=2, familynames="NO", grmoutname="grm")
    print("[gpu-case] cli population statistics (POPGEN example)", flush=True)
    out = run_cli(str(tmp_path / "par"), tmp_path).stdout
    log = open(os.path.join(golden, "popgen", "example_log.txt")).read()

    def grab(text, pat):
        return re.findall(pat, text, re.M)
    for pat in (r"^\s+eigenvector_\d_Case_Control_\s+(\S+)",
                r"^popdifference:\s+Case\s+Control\s+(\S+)\s+(\S+)\s+(\d+)\s+(\d+)",
                r"^ eigenvector\s+(\d)\s+\S+\s+(\S+)$",
                r"^population:\s+(\d+)\s+(\S+)\s+(\d+)$"):
        want = grab(log, pat)
        assert want, pat
        assert grab(out, pat) == want, pat
    from scipy import stats as S
    ref = O.smartpca(dict(genotypename="packed.geno", snpname="packed.snp",
                          indivname="packed.ind", altnormstyle="NO", numoutevec="2",
                          familynames="NO"), str(tmp_path))
    got = [float(x) for x in grab(out, r"^ eigenvector\s+\d\s+(\S+)\s+\S+$")]
    want = [S.kurtosis(ref["ffvecs"][j], fisher=False) for j in range(2)]
    assert np.allclose(got, want, atol=1.01e-3), (got, want)
    assert "## Tracy-Widom statistics: rows: 5  cols: 7" in out
    assert "## end of Fst statistics between populations" in out


@pytest.mark.gpu
def test_cli_deleted_snp_log_follows_the_reference(tmp_path):
    """deletsnpoutname (logdeletedsnp, mcio.c:4893-4900; smartpca.c:801-852, 1131-1139):
    chromosome filters first, then the position filters each on its own line (also for
    SNPs already dropped), then the no-data and the keep-rule drops, in the reference's
    '%-40s  %-40s' format."""
    from eig_amd import synth
    P = synth.config_panel(1, nsnp=1200, scale=0.1)
    prefix = str(tmp_path / "del")
    synth.write_files(P, prefix)
    snp_lines = open(prefix + ".snp").read().splitlines()
    # chromosome 23 (X), 0 and 30 on a few SNPs; positions as written by write_files
    for m, ch in ((3, 23), (4, 0), (5, 30)):
        f = snp_lines[m].split()
        f[1] = str(ch)
        snp_lines[m] = " ".join(f)
    open(prefix + ".snp", "w").write("\n".join(snp_lines) + "\n")
    # a monomorphic SNP on chromosome 2 (dropped by the keep rule); order within the
    # chromosome is unchanged, so the packed order check still holds
    raw = bytearray(open(prefix + ".geno", "rb").read())
    rlen = max(48, (2 * P.nind + 7) // 8)
    mono = [m for m in range(P.nsnp) if snp_lines[m].split()[1] == "2"][7]
    raw[rlen * (mono + 1): rlen * (mono + 2)] = bytes(rlen)  # all code 0
    open(prefix + ".geno", "wb").write(bytes(raw))
    write_par(tmp_path / "par", genotypename=prefix + ".geno", snpname=prefix + ".snp",
              indivname=prefix + ".ind", evecoutname="out.evec", numoutevec=2,
              numoutlieriter=0, chrom=2, lopos=1000 + 100 * (mono - 20),
              hipos=1000 + 100 * (mono + 20), deletsnpoutname="deleted.txt", hashcheck="NO",
              pordercheck="NO")
    print("[gpu-case] cli deleted-SNP log", flush=True)
    run_cli(str(tmp_path / "par"), tmp_path)
    ids = [l.split()[0] for l in snp_lines]
    chrom = [int(l.split()[1]) for l in snp_lines]
    pos = [int(l.split()[3]) for l in snp_lines]
    want = []
    for m in range(P.nsnp):
        why = ("chrom-X" if chrom[m] == 23 else "chrom-0" if chrom[m] == 0
               else "chrom-big" if chrom[m] > 23 else None)
        if why:
            want.append((ids[m], why))
    for m in range(P.nsnp):
        if chrom[m] != 2:
            want.append((ids[m], "not-chrom"))
        if pos[m] < 1000 + 100 * (mono - 20):
            want.append((ids[m], "lopos"))
        if pos[m] > 1000 + 100 * (mono + 20):
            want.append((ids[m], "hipos"))
    want.append((ids[mono], "minallelecnt"))
    got = (tmp_path / "deleted.txt").read_text().splitlines()
    assert got == ["%-40s  %-40s" % w for w in want]
