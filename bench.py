#!/usr/bin/env python3
"""Benchmark: smartpca XTX + eigensolve at N=5,000 samples x M=600,000 SNPs on one MI355X.

One step = one pass of the hot path over the resident panel: sample selection and
tile re-pack, column statistics (decode, counts, normalisation), XTX accumulation
and the FP64 eigensolve (all eigenvalues + top-10 eigenvectors).  The chosen form
(split-BF16 SYRK) is the headline value; the FP64 form runs the same steps
beside it.  Genotypes are generated directly in HBM (seeded Balding-Nichols,
10 populations x 500, Fst ~ U(0.01,0.1), p ~ U(0.01,0.5), 2% missing) before the
timed region.  The lsqproject serving leg reports p50/p99 latency at 1,000
samples/s Poisson arrivals against the same panel (and at 50x that rate, and a
saturating burst).  The same line carries the reference CPU pass (a sampled estimate
at this shape and a full run at N=1,000 x M=50,000), one fastmode call per form at
N=100,000 x M=500,000 and one pass per form at N=20,000 x M=1,000,000.

Other modes: --fastmode and --cfg3 run those shapes in full; --cpu-full times the
reference CPU pass in full at the headline shape.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "XTX+eig s: split-BF16 vs FP64 on GPU at N=5k,M=600k; lsqproject p99 ms"
PEAK_BF16 = 2500.0  # TFLOP/s dense, MI355X
PEAK_FP64 = 78.6
PEAK_HBM = 8000.0   # GB/s


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


BF16_BLOCK = 256  # SNPs per FP32 accumulation block of the split-BF16 forms


def one_pass(eng, xindex, form, nvec):
    eng.select(xindex)
    eng.snpstats(fetch=False)  # the statistics stay in HBM, where XTX reads them
    eng.xtx(form, BF16_BLOCK, want_matrix=False)
    ev, vec = eng.eigh_xtx(nvec)
    return ev, vec


def host_threads():
    """The reference's thread rule (all cores up to 8, else one fewer, smartpca.c:1042-1069),
    capped at the 16 host cores a one-GPU box gives a job."""
    n = os.cpu_count() or 1
    n = n if n <= 8 else min(n - 1, 127)
    return min(n, 16)


def cpu_baseline(eng, panel, nthreads):
    """Reference algorithm on the host (the oracle's C restatement): partial-sum
    lookup XTX on a SNP sample + LAPACK dspev at a reduced order, scaled to the
    full workload (XTX linear in M, dspev cubic in N)."""
    from oracle import oracle as O
    N, M = panel.nind, panel.nsnp
    ms = 8000
    rows = eng.fetch_packed(0, ms)
    xindex = np.arange(N, dtype=np.int32)
    t0 = time.perf_counter()
    st = O.snpstats(rows, xindex)
    O.xtx_lookup(rows, xindex, st["levels"], st["status"], nthreads=nthreads)
    t_x = time.perf_counter() - t0
    ne = 2500
    rng = np.random.default_rng(1)
    a = rng.standard_normal((ne, ne))
    a = a @ a.T
    t0 = time.perf_counter()
    O.dspev_lower(a)
    t_e = time.perf_counter() - t0
    est = t_x * (M / ms) + t_e * (N / ne) ** 3
    return dict(value=round(est, 2), unit="s", cores=nthreads, kind="port",
                sample=f"lookup XTX on {ms} of {M} SNPs x {N} samples ({t_x:.2f}s, x{M / ms:g}) + "
                       f"LAPACK dspev at n={ne} ({t_e:.2f}s, x(N/n)^3); per pass")


def cpu_cfg1_full(nthreads, runs=3):
    """The reference CPU pass in full at N=1,000 x M=50,000 (3 populations, 1% missing):
    column statistics, partial-sum lookup XTX, trace scaling and LAPACK dspev; median
    of `runs`, in seconds."""
    from eig_amd import synth
    from oracle import oracle as O
    P = synth.config_panel(1)
    rows = P.packed()
    xindex = np.arange(P.nind, dtype=np.int32)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        st = O.snpstats(rows, xindex)
        full, _ = O.finalize_xtx(O.xtx_lookup(rows, xindex, st["levels"], st["status"],
                                              nthreads=nthreads), P.nind)
        O.dspev_lower(full)
        times.append(time.perf_counter() - t0)
    return dict(s=round(float(np.median(times)), 3), runs=runs, cores=nthreads,
                N=P.nind, M=P.nsnp, lapack="numpy's bundled scipy_openblas64 dspev")


def cpu_full(eng, panel, nthreads):
    """The reference CPU pass in full at the headline workload: column statistics and
    partial-sum lookup XTX over every SNP, trace scaling and LAPACK dspev of the N x N
    matrix, once (minutes of host time)."""
    from oracle import oracle as O
    N, M = panel.nind, panel.nsnp
    xindex = np.arange(N, dtype=np.int32)
    rows = eng.fetch_packed(0, M)
    t0 = time.perf_counter()
    st = O.snpstats(rows, xindex)
    lower = O.xtx_lookup(rows, xindex, st["levels"], st["status"], nthreads=nthreads)
    t_x = time.perf_counter() - t0
    log(f"cpu full pass: statistics + lookup XTX {t_x:.1f} s")
    t1 = time.perf_counter()
    full, _ = O.finalize_xtx(lower, N)
    del lower
    O.dspev_lower(full)
    t_e = time.perf_counter() - t1
    log(f"cpu full pass: trace scaling + dspev {t_e:.1f} s")
    return dict(s=round(t_x + t_e, 2), xtx_s=round(t_x, 2), eig_s=round(t_e, 2), cores=nthreads,
                N=N, M=M, lapack="numpy's bundled scipy_openblas64 dspev")


def serve_leg(eng, P, vec, seconds, rate=1000.0, pool_n=4096):
    """lsqproject serving against the resident panel: Poisson arrivals at `rate`
    samples/s, request sizes log-uniform on [1, 512], query samples from the panel's
    populations with per-sample missingness ~ U(0, 0.9), rows resident in HBM.
    Returns latency percentiles (submit -> result ready), samples/s, and the same
    for a host-memory (PCIe-inclusive) run, a run at 50x the rate (batches shared by
    requests that arrive while others are on the device) and a saturating burst."""
    import threading

    import eig_amd
    from eig_amd import synth
    K = vec.shape[0]
    N = P.nind
    eng.loadings(vec, want_ffvecs=False)
    a, b, st = eng.lsqproject(np.ones(N, np.int32), K)
    eigscale = np.array([np.dot(a[j], b[j]) / np.dot(a[j], a[j]) for j in range(K)])
    stride = ((P.nsnp + 63) // 64) * 16
    pop, mt = synth.serving_queries(P, pool_n, seed=0xE16 + 5)
    pool = eng.dev_alloc(pool_n * stride)
    eig_amd.synth_queries(eng, P.thr, pop, mt, 0xE16 + 5, pool, stride)
    host_ptr, host = eng.host_alloc(pool_n * stride)  # pinned
    eng.memcpy(host_ptr, pool, pool_n * stride)
    host_np = host.reshape(pool_n, stride)
    srv = eig_amd.Server(eng, eigscale=eigscale, nlanes=4, max_batch=1024)
    info0 = srv.info()
    log(f"server up: chunk {info0['chunk']} grid {info0['grid']}")
    rng = np.random.default_rng(0xE16 + 5)

    def schedule(dur, r=rate):
        t, out = 0.0, []
        while True:
            n = int(min(512, max(1, np.floor(np.exp(rng.uniform(0, np.log(513)))))))
            t += rng.exponential(1.0 / (r / 81.9))
            if t > dur:
                return out
            out.append((t, n, int(rng.integers(0, pool_n - n))))

    def run(reqs, on_device):
        lat, done = [], []
        q = []
        cv = threading.Condition()

        def poller():
            k = 0
            while k < len(reqs):
                with cv:
                    while len(q) <= k:
                        cv.wait()
                    t = q[k]
                co, stt, ms = srv.poll(t)
                lat.append(ms)
                done.append(time.perf_counter())
                k += 1
        th = threading.Thread(target=poller)
        th.start()
        t0 = time.perf_counter()
        for (ta, n, off) in reqs:
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            if on_device:
                t = srv.submit(device_ptr=pool + off * stride, nq=n, stride=stride)
            else:
                t = srv.submit(host_np[off:off + n])
            with cv:
                q.append(t)
                cv.notify()
        th.join()
        wall = max(done) - t0
        ns = sum(n for _, n, _ in reqs)
        lat = np.array(lat)
        return dict(p50_ms=round(float(np.percentile(lat, 50)), 3),
                    p99_ms=round(float(np.percentile(lat, 99)), 3),
                    max_ms=round(float(lat.max()), 3), requests=len(reqs), samples=ns,
                    samples_per_s=round(ns / wall, 1))
    run(schedule(1.0), True)  # warm-up
    b0 = srv.info()["batches"]
    res = run(schedule(seconds), True)
    res["batches"] = srv.info()["batches"] - b0
    log(f"serving (HBM-resident rows): {res}")
    pcie = run(schedule(seconds / 2), False)
    log(f"serving (host rows, PCIe-inclusive): {pcie}")
    # under load: 50x the headline rate (about 30 % of saturation), where requests
    # arrive while others are on the device and share batches
    b0 = srv.info()["batches"]
    loaded = run(schedule(seconds / 4, 50.0 * rate), True)
    loaded["batches"] = srv.info()["batches"] - b0
    loaded["requests_per_batch"] = round(loaded["requests"] / max(loaded["batches"], 1), 2)
    log(f"serving under load ({50.0 * rate:g} samples/s offered): {loaded}")
    burst = [(0.0, n, off) for (_, n, off) in schedule(seconds / 2)]
    sat = run(burst, True)
    log(f"serving saturation burst: {sat}")
    srv.stop()
    eng.host_free(host_ptr)
    eng.dev_free(pool)
    return dict(resident=res, host_rows=pcie, loaded=loaded, burst=dict(samples=sat["samples"],
                samples_per_s=sat["samples_per_s"], p99_ms=sat["p99_ms"]),
                chunk=info0["chunk"], grid=info0["grid"])


def fastmode_leg(args, steps=None, warmup=None):
    """fastmode at the 50-population panel's shape (N=100,000 x M=500,000, 1% missing;
    --scale / --nsnp shrink it): `warmup` untimed and `steps` timed calls per form
    (the run's --steps / --warmup unless given)."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    import eig_amd
    from eig_amd import synth
    P = synth.config_panel(4, nsnp=None if args.nsnp == 600000 else args.nsnp, scale=args.scale)
    N, M = P.nind, P.nsnp
    log(f"fastmode: generating N={N} M={M} in HBM")
    eng = eig_amd.Engine(0)
    P.to_device(eng)
    eng.select(np.arange(N, dtype=np.int32))
    eng.snpstats()
    out = {"metric": "fastmode s per call (K=10, L=20, I=10)", "unit": "s", "n_gpus": 1,
           "config": {"workload": "fastmode randomized PCA, 50 Balding-Nichols pops, 1% missing",
                      "N": N, "M": M, "K": 10, "L": 20, "I": 10}, "forms": {}}
    ref = None
    for form, name in ((eig_amd.FORM_FP64, "fp64"), (eig_amd.FORM_BF16X4, "bf16x4")):
        for _ in range(warmup):
            eng.fpca(10, 20, 10, seed=1234, form=form)
        times = []
        for s in range(steps):
            t0 = time.perf_counter()
            ev, vec = eng.fpca(10, 20, 10, seed=1234, form=form)
            times.append(time.perf_counter() - t0)
            log(f"fastmode {name} call {s}: {times[-1]:.3f} s (power {eng.last_ms('fpca_power'):.1f} ms, "
                f"finish {eng.last_ms('fpca_finish'):.1f} ms)")
        rec = {"s": round(float(np.mean(times)), 4), "power_ms": eng.last_ms("fpca_power"),
               "finish_ms": eng.last_ms("fpca_finish"), "evals": [round(float(x), 6) for x in ev]}
        if ref is None:
            ref = (ev, vec)
        else:
            rec["eval_relerr_vs_fp64"] = float(np.max(np.abs(ev - ref[0]) / np.abs(ref[0])))
            rec["evec_maxabs_vs_fp64"] = float(np.max(np.abs(vec - ref[1])))
        out["forms"][name] = rec
    eng.close()
    out["speedup_vs_fp64"] = round(out["forms"]["fp64"]["s"] / out["forms"]["bf16x4"]["s"], 3)
    return out


def ridoutlier(evecs, nle, thresh):
    """The z-score outlier rule (smartsubs.c:18-93, outliermode 0): sample i is removed
    when |v_i - mean(v)| / sqrt(mean((v - mean)^2)) exceeds thresh on any of the top
    nle eigenvectors."""
    bad = np.zeros(evecs.shape[1], dtype=bool)
    for e in range(nle):
        v = evecs[e]
        w = v - v.mean()
        bad |= np.abs(w / np.sqrt(np.dot(w, w) / len(w))) > thresh
    return np.nonzero(bad)[0]


def cfg3_leg(args, max_passes=6):
    """The 25-population admixed panel at N=20,000 x M=1,000,000 (5% admixed, 3% missing)
    through smartpca's outlier loop (5 iterations, 6 sigma, 10 outlier eigenvectors),
    generated in HBM: each form's passes (statistics, XTX, eigensolve) timed, the
    removals, and the final top-10 eigenvalues of the two forms compared."""
    import eig_amd
    from eig_amd import synth
    P = synth.config_panel(3, nsnp=None if args.nsnp == 600000 else args.nsnp, scale=args.scale)
    N, M = P.nind, P.nsnp
    log(f"cfg3: generating N={N} M={M} in HBM")
    eng = eig_amd.Engine(0)
    P.to_device(eng)
    out = {"metric": "smartpca outlier loop s (5 iterations, 6 sigma, 10 eigenvectors)",
           "unit": "s", "n_gpus": 1,
           "config": {"workload": "25 Balding-Nichols pops x 800, 5% admixed (Dirichlet over 3), "
                                  "3% missing", "N": N, "M": M}, "forms": {}}
    for form, name in ((eig_amd.FORM_BF16X4, "bf16x4"), (eig_amd.FORM_FP64, "fp64")):
        keep = np.ones(N, dtype=bool)
        dead = np.zeros(M, dtype=bool)
        passes, removed = [], []
        t_all = time.perf_counter()
        for it in range(1, max_passes + 1):
            xindex = np.nonzero(keep)[0].astype(np.int32)
            t0 = time.perf_counter()
            eng.select(xindex)
            st = eng.snpstats(live=(~dead).astype(np.uint8))
            eng.xtx(form, BF16_BLOCK, want_matrix=False)
            ev, vec = eng.eigh_xtx(10)
            dt = time.perf_counter() - t0
            dead |= st["status"] != 0
            passes.append(dict(nrows=len(xindex), s=round(dt, 3), syrk_ms=eng.last_ms("syrk"),
                               tridiag_ms=eng.last_ms("tridiag"), trace_sum=float(ev.sum())))
            log(f"cfg3 {name} pass {it}: nrows {len(xindex)} {dt:.3f} s")
            if it == max_passes:
                break
            bad = ridoutlier(vec, min(10, len(xindex) - 1), 6.0)
            if len(bad) == 0:
                break
            removed.append([int(x) for x in xindex[bad]])
            keep[xindex[bad]] = False
        out["forms"][name] = dict(s=round(time.perf_counter() - t_all, 3), passes=passes,
                                  removed=removed, top10=[float(x) for x in ev[:10]])
    b, f = out["forms"]["bf16x4"], out["forms"]["fp64"]
    out["same_removals"] = b["removed"] == f["removed"]
    if out["same_removals"]:
        out["top10_eval_relerr_bf16x4_vs_fp64"] = float(np.max(
            np.abs(np.array(b["top10"]) - np.array(f["top10"])) / np.abs(np.array(f["top10"]))))
    out["speedup_vs_fp64"] = round(f["s"] / b["s"], 3)
    eng.close()
    return out


def main():
    global BF16_BLOCK  # set from --bf16-block; one_pass and the legs read it
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsnp", type=int, default=600000)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--forms", default="bf16x4,bf16x3,fp64",
                    help="forms to run; anything but all three prints stage times only "
                         "(for profiler runs)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="instead of the GPU passes, run the reference CPU pass once in full at "
                         "the headline workload and print its time")
    ap.add_argument("--no-extra", action="store_true",
                    help="leave out the one-call fastmode and one-pass N=20,000 summaries")
    ap.add_argument("--cfg3", action="store_true",
                    help="instead of the XTX pass, run the 25-population admixed panel at "
                         "20,000 x 1,000,000 through the outlier loop in both forms")
    ap.add_argument("--fastmode", action="store_true",
                    help="instead of the XTX pass, time fastmode (K=10, L=20, I=10) on the "
                         "100,000 x 500,000 panel in both panel-product forms")
    ap.add_argument("--bf16-block", type=int, default=BF16_BLOCK,
                    help="SNPs per FP32 accumulation block of the split-BF16 forms (a multiple "
                         "of 16; the headline line is quoted at 256)")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="first copy this many MiB device-to-device (the known-size copy "
                         "that scales the HBM byte counters of a --pmc run)")
    args = ap.parse_args()
    BF16_BLOCK = args.bf16_block
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import eig_amd
    from eig_amd import synth
    if args.fastmode:
        print(json.dumps(fastmode_leg(args)))
        return
    if args.cfg3:
        print(json.dumps(cfg3_leg(args)))
        return
    nvec = 10
    P = synth.config_panel(2, nsnp=args.nsnp, scale=args.scale)
    N, M = P.nind, P.nsnp
    log(f"generating N={N} M={M} in HBM")
    eng = eig_amd.Engine(0)
    P.to_device(eng)
    xindex = np.arange(N, dtype=np.int32)
    if args.cpu_full:
        out = {"metric": "reference CPU pass s (statistics, lookup XTX, dspev)", "unit": "s",
               "config": {"workload": "Balding-Nichols 10 pops x 500, 2% missing", "N": N, "M": M},
               "kind": "port", **cpu_full(eng, P, host_threads())}
        eng.close()
        print(json.dumps(out))
        return
    if args.copy_calib_mb:
        nb = args.copy_calib_mb << 20
        a, b = eng.dev_alloc(nb), eng.dev_alloc(nb)
        eng.memcpy(b, a, nb)
        eng.dev_free(a)
        eng.dev_free(b)
    res = {}
    forms = [(f, n) for f, n in ((eig_amd.FORM_BF16X4, "bf16x4"), (eig_amd.FORM_BF16X3, "bf16x3"),
                                 (eig_amd.FORM_FP64, "fp64")) if n in args.forms.split(",")]
    for form, name in forms:
        for _ in range(args.warmup):
            one_pass(eng, xindex, form, nvec)
        times, syrk_ms, eig_ms = [], [], []
        for s in range(args.steps):
            t0 = time.perf_counter()
            ev, vec = one_pass(eng, xindex, form, nvec)
            times.append(time.perf_counter() - t0)
            syrk_ms.append(eng.last_ms("syrk"))
            eig_ms.append(sum(eng.last_ms(k) for k in ("tridiag", "bisect", "invit", "eigvec")))
            log(f"{name} step {s}: {times[-1]*1e3:.1f} ms (syrk {syrk_ms[-1]:.2f} ms, "
                f"eig {eig_ms[-1]:.2f} ms)")
        res[name] = dict(t=float(np.mean(times)), syrk_ms=float(np.mean(syrk_ms)),
                         eig_ms=float(np.mean(eig_ms)), ev=ev, vec=vec,
                         stats_ms=eng.last_ms("snpstats"), select_ms=eng.last_ms("select"),
                         parts={k: eng.last_ms(k) for k in ("tridiag", "bisect", "invit", "eigvec",
                                                          "xtx_finalize")})
    if len(res) != 3:
        part = {k: {"s": v["t"], "syrk_ms": v["syrk_ms"], "eig_ms": v["eig_ms"], **v["parts"]}
                for k, v in res.items()}
        if "fp64" in res:  # top-10 eigenvalue relative error of each split form vs FP64
            f = res["fp64"]
            for k, v in res.items():
                if k != "fp64":
                    part[k]["top10_eval_relerr_vs_fp64"] = float(np.max(
                        np.abs(v["ev"][:nvec] - f["ev"][:nvec]) / np.abs(f["ev"][:nvec])))
        print(json.dumps({"forms": part, "bf16_block": BF16_BLOCK}))
        eng.close()
        return
    b, b3, f = res["bf16x4"], res["bf16x3"], res["fp64"]
    top = slice(0, nvec)

    def errs(x):
        rel = float(np.max(np.abs(x["ev"][top] - f["ev"][top]) / np.abs(f["ev"][top])))
        cos = float(np.abs(np.sum(x["vec"] * f["vec"], axis=1)).min())
        allabs = float(np.max(np.abs(x["ev"] - f["ev"])))
        return rel, cos, allabs
    relerr, cosmin, allabs = errs(b)
    relerr3, cosmin3, allabs3 = errs(b3)
    mfma_flops = 4.0 * N * (N + 1) * M   # the four BF16 products on MFMA
    alg_flops = 3.0 * N * (N + 1) * M    # the compensated SYRK's three products
    useful_flops = 1.0 * N * (N + 1) * M  # the XTX itself (lower triangle, 2 flop/MAC)
    syrk_s = b["syrk_ms"] * 1e-3
    tri_bytes = 8 * (N - 1) * N * (2 * N - 1) // 6  # sum over L = 1 .. N-1 of L^2 doubles
    tri_ach = tri_bytes / (b["parts"]["tridiag"] * 1e-3) / 1e9
    # decode + normalise: GENO rows read (rlen each) and 28 B of per-SNP statistics written
    rlen = max(48, (2 * N + 7) // 8)
    dec_bytes = M * rlen + 28 * M
    dec_ms = b["select_ms"] + b["stats_ms"]
    # HBM bytes per launch of the dominant stage's kernels, from rocprofv3 --pmc runs
    # (FETCH_SIZE, WRITE_SIZE) of this workload scaled by a 1 GiB copy: a stored figure
    traffic, tsrc = None, None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath) and (N, M) == (5000, 600000):
        t = json.load(open(tpath)).get("tridiag_pass")
        if t:
            traffic = int(t["read_bytes"] + t["write_bytes"])
            tsrc = t["source"] + " (stored, not this run)"
    out = {
        "metric": METRIC, "value": round(b["t"], 5), "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(b["t"] * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (hi+lo, 4 MFMA products, fp32 per %d-SNP block, fp64 across; fp64 "
                 "eigensolve)" % BF16_BLOCK,
        "data": "synthetic",
        "config": {"workload": "XTX+eigensolve, Balding-Nichols 10 pops x 500, 2% missing",
                   "N": N, "M": M, "numoutevec": nvec, "bf16_block": BF16_BLOCK,
                   "form": "split-BF16 hi+lo, 4 products"},
        "fp64_s": round(f["t"], 5),
        "speedup_vs_fp64": round(f["t"] / b["t"], 3),
        "top10_eval_relerr_vs_fp64": relerr,
        "top10_evec_min_abs_cos_vs_fp64": cosmin,
        "all_eval_max_abs_err_vs_fp64": allabs,
        "bf16x3_s": round(b3["t"], 5),
        "bf16x3_top10_eval_relerr_vs_fp64": relerr3,
        "bf16x3_all_eval_max_abs_err_vs_fp64": allabs3,
        "stages_ms": {"bf16x4": {"select": b["select_ms"], "snpstats": b["stats_ms"],
                                 "syrk": b["syrk_ms"], **b["parts"]},
                      "bf16x3": {"syrk": b3["syrk_ms"], **b3["parts"]},
                      "fp64": {"syrk": f["syrk_ms"], **f["parts"]}},
        # the dominant stage of the headline pass: the tridiagonalisation, one graph
        # launch per pass (k_symv / k_wcol / k_column / k_syr2k); its symv streams each
        # column's L x L trailing square once (L = N-1 .. 1); HIP-event time of the
        # launch on the context stream, per-column latency included
        "roofline": {"bound": "hbm", "achieved": round(tri_ach, 1), "peak": PEAK_HBM,
                     "unit": "GB/s", "frac": round(tri_ach / PEAK_HBM, 4), "traffic": traffic,
                     "traffic_source": tsrc,
                     "kernel": "tridiagonalisation graph (k_symv<4>, k_wcol, k_column, k_syr2k)",
                     "algorithmic_bytes": tri_bytes, "ms": round(b["parts"]["tridiag"], 3)},
        "syrk_roofline": {"kernel": "k_syrk_bf16<true, true>", "bound": "mfma", "ms": round(b["syrk_ms"], 3),
                          "peak": PEAK_BF16, "unit": "TFLOP/s",
                          "mfma_achieved": round(mfma_flops / syrk_s / 1e12, 1),
                          "mfma_frac": round(mfma_flops / syrk_s / 1e12 / PEAK_BF16, 4),
                          "alg_3product_achieved": round(alg_flops / syrk_s / 1e12, 1),
                          "alg_3product_frac": round(alg_flops / syrk_s / 1e12 / PEAK_BF16, 4),
                          "useful_rate": round(useful_flops / syrk_s / 1e12, 1),
                          "mfma_flops": mfma_flops, "alg_flops": alg_flops},
        "fp64_syrk_roofline": {"kernel": "k_syrk_f64",
                               "achieved": round(useful_flops / (f["syrk_ms"] * 1e-3) / 1e12, 2),
                               "peak": PEAK_FP64, "unit": "TFLOP/s",
                               "frac": round(useful_flops / (f["syrk_ms"] * 1e-3) / 1e12 / PEAK_FP64, 4)},
        "decode_normalise": {"kernels": "k_select + k_snpstats", "bytes": dec_bytes,
                             "ms": round(dec_ms, 4),
                             "achieved_GBps": round(dec_bytes / (dec_ms * 1e-3) / 1e9, 1),
                             "frac": round(dec_bytes / (dec_ms * 1e-3) / 1e9 / PEAK_HBM, 4)},
    }
    if not args.no_serve:
        log("serving leg")
        sv = serve_leg(eng, P, f["vec"], args.serve_seconds)
        out["lsqproject_p99_ms"] = sv["resident"]["p99_ms"]
        out["lsqproject_p50_ms"] = sv["resident"]["p50_ms"]
        out["lsqproject_samples_per_s"] = sv["resident"]["samples_per_s"]
        out["serving"] = sv
    if not args.no_cpu:
        nthreads = host_threads()
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(eng, P, nthreads)
        log("cpu pass in full at N=1,000 x M=50,000")
        out["cpu_cfg1_full"] = cpu_cfg1_full(nthreads)
    eng.close()
    if not args.no_extra:
        # the other headline shapes, one call / one pass each per form (their own
        # flags run them in full)
        log("fastmode at N=100,000 x M=500,000, one call per form")
        fm = fastmode_leg(args, steps=1, warmup=1)
        out["fastmode"] = {"config": fm["config"], "speedup_vs_fp64": fm["speedup_vs_fp64"],
                           **{k: {kk: v[kk] for kk in ("s", "power_ms", "finish_ms",
                                                       "eval_relerr_vs_fp64") if kk in v}
                              for k, v in fm["forms"].items()}}
        log("N=20,000 x M=1,000,000, one pass per form")
        c3 = cfg3_leg(args, max_passes=1)
        out["cfg3_one_pass"] = {"config": c3["config"], "speedup_vs_fp64": c3["speedup_vs_fp64"],
                                "top10_eval_relerr_bf16x4_vs_fp64":
                                    c3.get("top10_eval_relerr_bf16x4_vs_fp64"),
                                **{k: v["passes"][0] for k, v in c3["forms"].items()}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
