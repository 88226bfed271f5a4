"""lsqproject serving (eig_server_*) against the oracle, and batch invariance.

Bars: synthetic query rows bit-exact with their CPU restatement; per-sample
coordinates within 1e-9 relative (atol 1e-12) of the oracle's lsqproj restatement
(orc_lsqproj, smartpca.c:4606-4757) fed the same loadings, status codes equal;
a request's result bit-identical alone, batched with other requests, and
submitted from host or device memory.
"""
import sys
import threading

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _panel_model(eng, N=240, M=3001, K=6, seed=5):
    """Loadings of a small synthetic panel on the engine; returns the oracle's inputs."""
    from eig_amd import synth
    P = synth.config_panel(2, nsnp=M, scale=N / 5000)
    P.to_device(eng)
    rows = P.packed()
    xindex = np.arange(P.nind, dtype=np.int32)
    eng.select(xindex)
    eng.snpstats()
    eng.xtx(0, 0, want_matrix=False)
    _, vec = eng.eigh_xtx(K)
    ff, fx = eng.loadings(vec)
    ref = O.snpstats(rows, xindex)
    return dict(P=P, rows=rows, ff=ff, fx=fx, K=K, ign=(ref["status"] != 0).astype(np.int32),
                xm=ref["mean"] * ref["fancy"], fancy=np.ascontiguousarray(ref["fancy"]))


def _oracle_project(model, g):
    """orc_lsqproj over query codes g [nq, M] (3 = missing)."""
    K = model["K"]
    nq, M = g.shape
    rows = O.pack_genotypes(np.ascontiguousarray(g.T))
    use = np.ones(nq, np.int32)
    ra = np.zeros((K, nq)); rb = np.zeros((K, nq)); rs = np.zeros(nq, np.int32)
    O.lib().orc_lsqproj(rows.reshape(-1), rows.shape[1], M, nq, use, model["ign"], model["xm"],
                        model["fancy"], np.ascontiguousarray(model["ff"]).reshape(-1),
                        np.ascontiguousarray(model["fx"]), K, ra.reshape(-1), rb.reshape(-1), rs)
    return ra.T, rs


@pytest.fixture(scope="module")
def eng():
    import eig_amd
    e = eig_amd.Engine(0)
    yield e
    e.close()


def test_synth_queries_match_cpu_restatement(eng):
    import eig_amd
    from eig_amd import synth
    P = synth.config_panel(2, nsnp=777, scale=0.01)
    pop, mt = synth.serving_queries(P, 37, seed=11)
    stride = ((P.nsnp + 63) // 64) * 16
    _say(f"synth queries M={P.nsnp} nq=37 stride={stride}")
    buf = eng.dev_alloc(37 * stride)
    eig_amd.synth_queries(eng, P.thr, pop, mt, 99, buf, stride)
    got = np.zeros((37, stride), np.uint8)
    eng.memcpy(got, buf, got.nbytes)
    eng.dev_free(buf)
    want = synth.pack_samples(synth.query_genotypes(P.thr, pop, mt, 99), stride)
    assert np.array_equal(got, want)


def test_server_matches_oracle(eng):
    import eig_amd
    from eig_amd import synth
    model = _panel_model(eng)
    P, K = model["P"], model["K"]
    M = P.nsnp
    # queries: the panel's own individuals, then seeded samples with heavy and
    # extreme missingness (all missing; exactly K observed; K + 1 observed)
    own = P.genotypes().T.copy()
    pop, mt = synth.serving_queries(P, 60, seed=3, miss_lo=0.0, miss_hi=0.95)
    qg = synth.query_genotypes(P.thr, pop, mt, 1234)
    qg[0, :] = 3
    kept = np.flatnonzero(model["ign"] == 0)
    qg[1, :] = 3
    qg[1, kept[:K]] = 1
    qg[2, :] = 3
    qg[2, kept[10:10 + 3 * K]] = 2
    g = np.concatenate([own, qg])
    _say(f"server vs oracle M={M} nq={len(g)} K={K}")
    want, wst = _oracle_project(model, g)
    with eig_amd.Server(eng, nlanes=2, max_batch=512) as srv:
        assert srv.info()["chunk"] % 64 == 0
        rows = synth.pack_samples(g)
        t1 = srv.submit(rows[:P.nind])
        t2 = srv.submit(rows[P.nind:])
        c1, s1, _ = srv.poll(t1)
        c2, s2, _ = srv.poll(t2)
    co = np.concatenate([c1, c2]); st = np.concatenate([s1, s2])
    assert np.array_equal(st, wst), (st, wst)
    assert st[P.nind] == 1 and st[P.nind + 1] == 1 and st[P.nind + 2] == 0
    ok = st == 0
    assert np.allclose(co[ok], want[ok], rtol=1e-9, atol=1e-12)
    assert np.all(co[st == 1] == 0.0)


def test_server_result_same_alone_and_in_the_mix(eng):
    import eig_amd
    from eig_amd import synth
    model = _panel_model(eng, N=200, M=20000, K=10)
    P = model["P"]
    npool = 1200
    stride = ((P.nsnp + 63) // 64) * 16
    pop, mt = synth.serving_queries(P, npool, seed=21)
    base = eng.dev_alloc(npool * stride)
    eig_amd.synth_queries(eng, P.thr, pop, mt, 77, base, stride)
    host = np.zeros((npool, stride), np.uint8)
    eng.memcpy(host, base, host.nbytes)
    r0, nr = 500, 37  # the request under test
    _say(f"server invariance M={P.nsnp} pool={npool}")
    with eig_amd.Server(eng, nlanes=4, max_batch=1024) as srv:
        alone = srv.poll(srv.submit(device_ptr=base + r0 * stride, nq=nr, stride=stride))
        from_host = srv.poll(srv.submit(host[r0:r0 + nr]))
        b0 = srv.info()["batches"]
        rng = np.random.default_rng(4)
        sizes = [max(1, int(x)) for x in np.exp(rng.uniform(0, np.log(512), size=24))]
        offs = [int(rng.integers(0, npool - n)) for n in sizes]
        sizes[12], offs[12] = nr, r0
        tickets, mine = [], None
        lock = threading.Lock()

        def client(lo, hi):
            for i in range(lo, hi):
                t = srv.submit(device_ptr=base + offs[i] * stride, nq=sizes[i], stride=stride)
                with lock:
                    tickets.append((i, t))
        th = [threading.Thread(target=client, args=(k * 8, k * 8 + 8)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for i, t in tickets:
            res = srv.poll(t)
            if i == 12:
                mine = res
        nb = srv.info()["batches"] - b0
    eng.dev_free(base)
    assert mine is not None
    for other in (from_host, mine):
        assert np.array_equal(alone[0], other[0])
        assert np.array_equal(alone[1], other[1])
    assert np.all(alone[1] == 0)
    assert nb < len(tickets), "requests were not batched"


def test_server_rejects_bad_requests(eng):
    import eig_amd
    _panel_model(eng, N=120, M=640, K=3)
    with eig_amd.Server(eng, nlanes=1, max_batch=64, chunk=128) as srv:
        info = srv.info()
        assert info["chunk"] == 128
        rows = np.zeros((65, srv.rowbytes), np.uint8)
        with pytest.raises(eig_amd.EigError):
            srv.submit(rows)                          # more than max_batch
        with pytest.raises(eig_amd.EigError):
            srv.submit(rows[:, :-1][:2])              # row shorter than ceil(nsnp/4)
        with pytest.raises(eig_amd.EigError):
            srv.submit(device_ptr=8, nq=1, stride=srv.rowpad)  # misaligned device rows
        t = srv.submit(np.full((1, srv.rowbytes), 0xFF, np.uint8))  # all missing
        co, st, _ = srv.poll(t)
        assert st[0] == 1 and np.all(co == 0)
