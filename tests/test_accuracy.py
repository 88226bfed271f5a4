"""Accuracy of the split-BF16 form against the FP64 oracle, at the workload sizes.

* N=1,000 x M=50,000 (three Balding-Nichols populations, 1% missing): the engine's
  BF16X4 and FP64 eigenvalues against LAPACK dspev of the oracle's FP64 XTX (the
  reference's partial-sum lookup order), top-10 relative error <= 1e-7 for BF16X4 and
  <= 1e-11 for FP64; the separated top eigenvectors' subspace angle; and the drop-in
  CLI's .evec/.eval against the oracle's pipeline for both forms.
* N=5,000 x M=600,000 (ten populations, 2% missing), generated in HBM: properties that
  need no CPU pass over the genotypes -- BF16X4 vs FP64 top-10 <= 1e-7, sum of the
  eigenvalues = N-1, eigen-residuals, orthonormality, and the FP64 eigenvalues against
  numpy's eigvalsh of the same matrix.
* The admixed 25-population panel (5% admixed, 3% missing) at N=1,000 through the CLI
  with 5 outlier iterations, 10 outlier eigenvectors and 6 sigma, two planted
  outliers: the same removals, eigenvalues and coordinates as the oracle.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOP = 10


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def eng():
    import eig_amd
    e = eig_amd.Engine(0)
    yield e
    e.close()


def _max_sin_angle(a, b):
    """Largest principal-angle sine between the row spaces of a and b (k x n):
    || (I - Qa Qa^T) Qb ||_2, accurate for small angles (unlike sqrt(1 - cos^2))."""
    qa, _ = np.linalg.qr(a.T)
    qb, _ = np.linalg.qr(b.T)
    return float(np.linalg.norm(qb - qa @ (qa.T @ qb), 2))


@pytest.fixture(scope="module")
def cfg1():
    from eig_amd import synth
    P = synth.config_panel(1)
    rows = P.packed()
    return P, rows


def test_cfg1_full_size_engine_vs_oracle(eng, cfg1):
    import eig_amd
    P, rows = cfg1
    N, M = P.nind, P.nsnp
    _say(f"cfg1 engine N={N} M={M}")
    xindex = np.arange(N, dtype=np.int32)
    st = O.snpstats(rows, xindex)
    lower = O.xtx_lookup(rows, xindex, st["levels"], st["status"], nthreads=16)
    full, y = O.finalize_xtx(lower, N)
    ev_o, vec_o = O.dspev_lower(full)
    eng.load_packed(rows, N)
    eng.select(xindex)
    eng.snpstats(fetch=False)
    res = {}
    for form, name in ((eig_amd.FORM_FP64, "fp64"), (eig_amd.FORM_BF16X4, "bf16x4")):
        eng.xtx(form, 256, want_matrix=False)
        ev, vec = eng.eigh_xtx(TOP)
        rel = np.abs(ev[:TOP] - ev_o[:TOP]) / np.abs(ev_o[:TOP])
        res[name] = (ev, vec, float(rel.max()))
        _say(f"cfg1 {name}: top-{TOP} eigenvalue rel. error vs oracle {rel.max():.3e}, "
             f"all-eigenvalue max abs error {np.max(np.abs(ev - ev_o)):.3e}")
    assert res["fp64"][2] <= 1e-11
    assert res["bf16x4"][2] <= 1e-7  # the headline accuracy bar
    assert np.max(np.abs(res["bf16x4"][0] - ev_o)) <= 1e-5  # every eigenvalue, 6 d.p. output
    # the two population axes are well separated from the bulk: their subspace is
    # determined to high accuracy; report the top-10 subspace angle as well
    vo = vec_o[:TOP]
    for name in ("fp64", "bf16x4"):
        s2 = _max_sin_angle(res[name][1][:2], vo[:2])
        s10 = _max_sin_angle(res[name][1], vo)
        _say(f"cfg1 {name}: sin(angle) top-2 {s2:.3e} top-10 {s10:.3e}")
        assert s2 <= (1e-9 if name == "fp64" else 1e-6)


@pytest.mark.parametrize("form", ["fp64", "bf16x4"])
def test_cfg1_full_size_cli_vs_oracle(tmp_path, form):
    from eig_amd import synth
    from test_cli import _parse_evec, run_cli, write_par
    P = synth.config_panel(1)
    prefix = str(tmp_path / "cfg1")
    synth.write_files(P, prefix)
    par = dict(genotypename=prefix + ".geno", snpname=prefix + ".snp", indivname=prefix + ".ind",
               evecoutname="out.evec", evaloutname="out.eval", numoutevec=10, numoutlieriter=0,
               xtxform=form)
    write_par(tmp_path / "par", **par)
    _say(f"cfg1 cli N={P.nind} M={P.nsnp} form={form}")
    run_cli(str(tmp_path / "par"), tmp_path)
    ref = O.smartpca({k: str(v) for k, v in par.items()}, str(tmp_path), nthreads=16)
    got_eval = np.array([float(x) for x in (tmp_path / "out.eval").read_text().split()])
    want_eval = np.array([float(x) for x in ref["eval_lines"]])
    assert np.max(np.abs(got_eval - want_eval)) <= 1.01e-6
    gh, grows = _parse_evec((tmp_path / "out.evec").read_text())
    wh, wrows = _parse_evec("\n".join(ref["evec_lines"]))
    assert np.allclose(gh, wh, atol=1.01e-3)
    assert [r[0] for r in grows] == [r[0] for r in wrows]
    g = np.array([r[1] for r in grows])
    w = np.array([r[1] for r in wrows])
    # the two population axes to print precision; the bulk axes' coordinates are
    # ill-conditioned (eigenvalue gaps ~1e-3), their eigenvalues are checked above
    assert np.max(np.abs(g[:, :2] - w[:, :2])) <= 2.01e-4


def test_cfg2_full_size_properties(eng):
    import eig_amd
    from eig_amd import synth
    P = synth.config_panel(2)
    N, M = P.nind, P.nsnp
    _say(f"cfg2 properties N={N} M={M}")
    P.to_device(eng)
    eng.select(np.arange(N, dtype=np.int32))
    eng.snpstats(fetch=False)
    y64, A = eng.xtx(eig_amd.FORM_FP64, 256, want_matrix=True)
    ev64, v64 = eng.eigh_xtx(TOP)
    yb, Ab = eng.xtx(eig_amd.FORM_BF16X4, 256, want_matrix=True)
    evb, vb = eng.eigh_xtx(TOP)
    assert np.array_equal(A, A.T) and np.array_equal(Ab, Ab.T)
    # trace normalisation: sum of eigenvalues = trace = N - 1
    assert abs(np.trace(A) - (N - 1)) <= 1e-9 * N
    assert abs(ev64.sum() - (N - 1)) <= 1e-8 * N
    assert abs(evb.sum() - (N - 1)) <= 1e-8 * N
    lam = ev64[0]
    for ev, v, mat, name in ((ev64, v64, A, "fp64"), (evb, vb, Ab, "bf16x4")):
        r = mat @ v.T - v.T * ev[:TOP]
        res = float(np.max(np.linalg.norm(r, axis=0)))
        orth = float(np.max(np.abs(v @ v.T - np.eye(TOP))))
        _say(f"cfg2 {name}: max ||Av - lv|| {res:.3e}, max |V^T V - I| {orth:.3e}")
        assert res <= 1e-10 * lam
        assert orth <= 1e-10
    rel = np.abs(evb[:TOP] - ev64[:TOP]) / np.abs(ev64[:TOP])
    xerr = float(np.max(np.abs(Ab - A)) / np.max(np.abs(A)))
    _say(f"cfg2 bf16x4 vs fp64: top-{TOP} eigenvalue rel. error {rel.max():.3e}, "
         f"XTX max rel. error {xerr:.3e}, sin(angle) top-{TOP} {_max_sin_angle(vb, v64):.3e}")
    assert rel.max() <= 1e-7
    w = np.linalg.eigvalsh(A)[::-1]
    err = float(np.max(np.abs(w - ev64)))
    _say(f"cfg2 fp64 eigenvalues vs numpy eigvalsh: max abs error {err:.3e} (lambda_max {lam:.3f})")
    assert err <= 1e-11 * lam


@pytest.mark.parametrize("form", ["fp64", "bf16x4"])
def test_cfg3_admixed_outlier_iterations_cli_vs_oracle(tmp_path, form):
    from eig_amd import synth
    from test_cli import _parse_evec, run_cli, write_par
    P = synth.config_panel(3, nsnp=20000, scale=0.05)
    assert P.nind == 1000 and P.admixed.sum() == 50
    prefix = str(tmp_path / "cfg3")
    synth.write_files(P, prefix)
    # two planted outliers: one sample all hom-alt, one all hom-ref
    raw = bytearray(open(prefix + ".geno", "rb").read())
    rlen = max(48, (2 * P.nind + 7) // 8)
    for who, code in ((11, 2), (503, 0)):
        for m in range(P.nsnp):
            off = rlen * (m + 1) + who // 4
            sh = 2 * (3 - who % 4)
            raw[off] = (raw[off] & ~(3 << sh)) | (code << sh)
    open(prefix + ".geno", "wb").write(bytes(raw))
    par = dict(genotypename=prefix + ".geno", snpname=prefix + ".snp", indivname=prefix + ".ind",
               evecoutname="out.evec", evaloutname="out.eval", outlieroutname="out.outl",
               numoutevec=10, numoutlieriter=5, numoutlierevec=10, outliersigmathresh=6.0,
               xtxform=form)
    write_par(tmp_path / "par", **par)
    _say(f"cfg3 admixed cli N={P.nind} M={P.nsnp} form={form}")
    run_cli(str(tmp_path / "par"), tmp_path)
    ref = O.smartpca({k: str(v) for k, v in par.items()}, str(tmp_path), nthreads=16)
    got = (tmp_path / "out.outl").read_text().splitlines()
    want = ["REMOVED outlier %s iter %d evec %d sigmage %.3f pop: %s" % r for r in ref["removed"]]
    _say(f"cfg3 {form}: {len(want)} removals by the oracle")
    assert len(want) >= 1
    assert [g.split()[:7] for g in got] == [w.split()[:7] for w in want]
    # the score's sign follows the eigenvector's sign, which the reference leaves to
    # LAPACK and this build fixes by its sign rule: magnitudes are compared
    assert np.allclose([abs(float(g.split()[8])) for g in got],
                       [abs(float(w.split()[8])) for w in want], atol=2e-3)
    got_eval = np.array([float(x) for x in (tmp_path / "out.eval").read_text().split()])
    want_eval = np.array([float(x) for x in ref["eval_lines"]])
    assert len(got_eval) == len(want_eval)
    assert np.max(np.abs(got_eval - want_eval)) <= (1.01e-6 if form == "fp64" else 1e-5)
    gh, grows = _parse_evec((tmp_path / "out.evec").read_text())
    wh, wrows = _parse_evec("\n".join(ref["evec_lines"]))
    assert np.allclose(gh, wh, atol=1.01e-3)
    assert [r[0] for r in grows] == [r[0] for r in wrows]
    # coordinates on the axes separated from the bulk by a clear eigenvalue gap
    sep = [j for j in range(len(wh) - 1) if wh[j] - wh[j + 1] > 0.05 * wh[j + 1]]
    k = (max(sep) + 1) if sep else 1
    g = np.array([r[1] for r in grows])[:, :k]
    w = np.array([r[1] for r in wrows])[:, :k]
    _say(f"cfg3 {form}: {k} separated axes, max coordinate difference {np.max(np.abs(g - w)):.2e}")
    assert np.max(np.abs(g - w)) <= 2.01e-4
