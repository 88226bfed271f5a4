"""MI355X smartpca engine: Python binding of libeig_amd.so (the C-ABI in
include/eig_abi.h).  The compute path is the HIP library; this module only moves
arrays across the boundary.  It raises if the library is missing -- there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libeig_amd.so")
CLI_PATH = os.path.join(PKG, "bin", "smartpca")

FORM_FP64 = 0
FORM_BF16X3 = 1
FORM_BF16X4 = 2
SNP_KEPT, SNP_DROPPED, SNP_ALLMISS, SNP_DEAD = 0, 1, 2, 3

_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p

# name -> (restype, argtypes); mirrors include/eig_abi.h
SIGNATURES = {
    "eig_open": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "eig_close": (None, [_vp]),
    "eig_last_error": (C.c_char_p, [_vp]),
    "eig_last_ms": (C.c_double, [_vp, C.c_char_p]),
    "eig_load_packed": (C.c_int, [_vp, _u8p, C.c_int64, C.c_int64, C.c_int32]),
    "eig_synth_packed": (C.c_int, [_vp, C.c_int64, C.c_int32, C.c_int32, _u32p, _i32p, _u32p,
                                   C.c_uint32, C.c_uint64]),
    "eig_fetch_packed": (C.c_int, [_vp, C.c_int64, C.c_int64, _u8p]),
    "eig_select": (C.c_int, [_vp, _i32p, C.c_int32]),
    "eig_snpstats": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _vp,
                               _vp, _vp, _vp, _vp, _vp]),  # outputs may be NULL
    "eig_xtx": (C.c_int, [_vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), _vp]),
    "eig_xtx_normal": (C.c_int, [_vp, _vp, C.c_int32, _vp, C.c_int32, _vp, _vp, _vp, C.c_int32,
                                 C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, _vp,
                                 _vp, _vp, _vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 _vp]),
    "eig_popcounts": (C.c_int, [_vp, _i32p, C.c_int32, _vp, _vp, _i32p]),
    "eig_eigh_xtx": (C.c_int, [_vp, C.c_int32, _f64p, _vp]),
    "eig_eigh": (C.c_int, [_vp, _f64p, C.c_int32, C.c_int32, _f64p, _vp]),
    "eig_loadings": (C.c_int, [_vp, _f64p, C.c_int32, _vp, _vp, _vp]),
    "eig_lsqproject": (C.c_int, [_vp, _i32p, _f64p, _f64p, _i32p]),
    "eig_lsqproject_fill": (C.c_int, [_vp, _i32p, _i32p, C.c_int32, _f64p, _f64p, _i32p]),
    "eig_lsqproject_noalloc": (C.c_int, [_vp, _i32p, _f64p, _f64p, _i32p]),
    "eig_fpca": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _vp, _f64p, _vp]),
    "eig_fpca_form": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _vp, C.c_int32,
                                _f64p, _vp]),
    "eig_dev_alloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "eig_dev_free": (C.c_int, [_vp, _vp]),
    "eig_host_alloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "eig_host_free": (C.c_int, [_vp, _vp]),
    "eig_memcpy": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "eig_server_start": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "eig_server_submit": (C.c_int, [_vp, _vp, C.c_int64, C.c_int32, C.c_int32,
                                    C.POINTER(C.c_int64)]),
    "eig_server_poll": (C.c_int, [_vp, C.c_int64, C.c_int32, _vp, _vp, C.POINTER(C.c_double)]),
    "eig_server_info": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64)]),
    "eig_server_error": (C.c_char_p, [_vp]),
    "eig_server_stop": (None, [_vp]),
    "eig_synth_queries": (C.c_int, [_vp, C.c_int64, C.c_int32, C.c_int32, _u32p, _i32p, _u32p,
                                    C.c_uint64, _vp, C.c_int64]),
}

EIG_PENDING = 1


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run python -m eig_amd.build (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class EigError(RuntimeError):
    pass


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Engine:
    """One device context (eig_ctx)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = _vp()
        rc = self._L.eig_open(device, C.byref(h))
        if rc != 0:
            raise EigError(f"eig_open({device}) failed with {rc}")
        self.h = h
        self.nsnp = 0
        self.nind = 0
        self.nrows = 0

    def close(self):
        if self.h:
            self._L.eig_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            msg = self._L.eig_last_error(self.h)
            raise EigError(f"{what}: rc={rc}: {msg.decode() if msg else ''}")

    def last_ms(self, what: str) -> float:
        return self._L.eig_last_ms(self.h, what.encode())

    # ---- device / pinned memory (this package loads no other HIP runtime)
    def dev_alloc(self, nbytes: int) -> int:
        p = _vp()
        self._chk(self._L.eig_dev_alloc(self.h, nbytes, C.byref(p)), "eig_dev_alloc")
        return p.value

    def dev_free(self, ptr: int):
        self._chk(self._L.eig_dev_free(self.h, C.c_void_p(ptr)), "eig_dev_free")

    def host_alloc(self, nbytes: int):
        """Pinned host buffer as (address, uint8 numpy view)."""
        p = _vp()
        self._chk(self._L.eig_host_alloc(self.h, nbytes, C.byref(p)), "eig_host_alloc")
        view = np.ctypeslib.as_array((C.c_uint8 * nbytes).from_address(p.value))
        return p.value, view

    def host_free(self, ptr: int):
        self._chk(self._L.eig_host_free(self.h, C.c_void_p(ptr)), "eig_host_free")

    def memcpy(self, dst, src, nbytes: int):
        """dst / src: integer addresses (device or host) or numpy arrays."""
        d = dst.ctypes.data if isinstance(dst, np.ndarray) else dst
        s_ = src.ctypes.data if isinstance(src, np.ndarray) else src
        self._chk(self._L.eig_memcpy(self.h, C.c_void_p(d), C.c_void_p(s_), nbytes), "eig_memcpy")

    def load_packed(self, rows: np.ndarray, nind: int):
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        self._chk(self._L.eig_load_packed(self.h, rows.reshape(-1), rows.shape[1], rows.shape[0],
                                          nind), "eig_load_packed")
        self.nsnp, self.nind = rows.shape[0], nind

    def synth(self, nsnp, nind, thr, anc, cum, miss_thr, seed):
        thr = np.ascontiguousarray(thr, dtype=np.uint32)
        npop = thr.shape[0]
        self._chk(self._L.eig_synth_packed(self.h, nsnp, nind, npop, thr.reshape(-1),
                                           np.ascontiguousarray(anc, dtype=np.int32).reshape(-1),
                                           np.ascontiguousarray(cum, dtype=np.uint32).reshape(-1),
                                           int(miss_thr), int(seed)), "eig_synth_packed")
        self.nsnp, self.nind = nsnp, nind

    def fetch_packed(self, m0=0, count=None):
        if count is None:
            count = self.nsnp - m0
        rlen = max(48, (2 * self.nind + 7) // 8)
        out = np.zeros((count, rlen), dtype=np.uint8)
        self._chk(self._L.eig_fetch_packed(self.h, m0, count, out.reshape(-1)), "eig_fetch_packed")
        return out

    def select(self, xindex):
        xi = np.ascontiguousarray(xindex, dtype=np.int32)
        self._chk(self._L.eig_select(self.h, xi, len(xi)), "eig_select")
        self.nrows = len(xi)

    def snpstats(self, live=None, altnorm=1, fancynorm=1, minallelecnt=1, maxmissing=9999999,
                 fetch=True):
        """Column statistics on the device.  fetch=False leaves them there (the XTX,
        loadings and lsqproject kernels read them in place) and returns None."""
        M = self.nsnp
        keys = ("c0", "c1", "nmiss", "mean", "fancy", "status")
        types = (np.int32, np.int32, np.int32, np.float64, np.float64, np.int32)
        out = {k: np.zeros(M, t) for k, t in zip(keys, types)} if fetch else None
        lv = None if live is None else np.ascontiguousarray(live, dtype=np.uint8)
        args = [_ptr(out[k]) if fetch else None for k in keys]
        self._chk(self._L.eig_snpstats(self.h, _ptr(lv), altnorm, fancynorm, minallelecnt,
                                       maxmissing, *args), "eig_snpstats")
        return out

    def xtx(self, form=FORM_FP64, block=0, want_matrix=True):
        y = C.c_double(0)
        mat = np.zeros((self.nrows, self.nrows)) if want_matrix else None
        self._chk(self._L.eig_xtx(self.h, form, block, C.byref(y), _ptr(mat)), "eig_xtx")
        return y.value, mat

    def xtx_normal(self, xtypes=None, npop=0, weight=None, ldregress=0, chrom=None, ppos=None,
                   gpos=None, ldposlimit=1000000000, ldlimit=9999.0, ldr2lo=0.01, ldr2hi=0.95,
                   minallelecnt=1, maxmissing=9999999, want_matrix=True):
        """The non-lookup column path (usepopsformissing / weights / ldregress): returns
        (trace_scale, ynumsnps, stats dict, matrix or None)."""
        M = self.nsnp
        st = {k: np.zeros(M, np.int32) for k in ("c0", "c1", "nmiss", "status")}
        arr = lambda a, t: None if a is None else np.ascontiguousarray(a, dtype=t)
        xt, w = arr(xtypes, np.int32), arr(weight, np.float64)
        ch, pp, gp = arr(chrom, np.int32), arr(ppos, np.float64), arr(gpos, np.float64)
        y, yn = C.c_double(0), C.c_double(0)
        mat = np.zeros((self.nrows, self.nrows)) if want_matrix else None
        self._chk(self._L.eig_xtx_normal(self.h, _ptr(xt), npop, _ptr(w), ldregress, _ptr(ch),
                                         _ptr(pp), _ptr(gp), ldposlimit, ldlimit, ldr2lo, ldr2hi,
                                         minallelecnt, maxmissing, _ptr(st["c0"]), _ptr(st["c1"]),
                                         _ptr(st["nmiss"]), _ptr(st["status"]), C.byref(y),
                                         C.byref(yn), _ptr(mat)), "eig_xtx_normal")
        return y.value, yn.value, st, mat

    def popcounts(self, xtypes, npop, haploid=None, xsnp=None):
        xt = np.ascontiguousarray(xtypes, dtype=np.int32)
        h = None if haploid is None else np.ascontiguousarray(haploid, dtype=np.uint8)
        x = None if xsnp is None else np.ascontiguousarray(xsnp, dtype=np.uint8)
        out = np.zeros((self.nsnp, npop, 2), np.int32)
        self._chk(self._L.eig_popcounts(self.h, xt, npop, _ptr(h), _ptr(x), out.reshape(-1)),
                  "eig_popcounts")
        return out

    def eigh_xtx(self, nvec):
        ev = np.zeros(self.nrows)
        vec = np.zeros((nvec, self.nrows)) if nvec > 0 else None
        self._chk(self._L.eig_eigh_xtx(self.h, nvec, ev, _ptr(vec)), "eig_eigh_xtx")
        return ev, vec

    def eigh(self, a, nvec):
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        ev = np.zeros(n)
        vec = np.zeros((nvec, n)) if nvec > 0 else None
        self._chk(self._L.eig_eigh(self.h, a.reshape(-1), n, nvec, ev, _ptr(vec)), "eig_eigh")
        return ev, vec

    def loadings(self, evecs, ignore=None, want_ffvecs=True):
        ev = np.ascontiguousarray(evecs, dtype=np.float64)
        K = ev.shape[0]
        ign = None if ignore is None else np.ascontiguousarray(ignore, dtype=np.uint8)
        ff = np.zeros((K, self.nsnp)) if want_ffvecs else None
        fx = np.zeros(K)
        self._chk(self._L.eig_loadings(self.h, ev.reshape(-1), K, _ptr(ign), _ptr(ff), _ptr(fx)),
                  "eig_loadings")
        return ff, fx

    def fpca(self, K, L, I, seed=0, omega=None, form=0):
        """fastmode randomized PCA over the selected samples (after snpstats);
        form FORM_FP64 or FORM_BF16X4 (split-BF16 panel products)."""
        om = None if omega is None else np.ascontiguousarray(omega, dtype=np.float64)
        ev = np.zeros(K)
        vec = np.zeros((K, self.nrows))
        self._chk(self._L.eig_fpca_form(self.h, K, L, I, seed, _ptr(om), form, ev, _ptr(vec)),
                  "eig_fpca_form")
        return ev, vec

    def lsqproject(self, use, K):
        use = np.ascontiguousarray(use, dtype=np.int32)
        n = len(use)
        a = np.zeros((K, n)); b = np.zeros((K, n)); st = np.zeros(n, np.int32)
        self._chk(self._L.eig_lsqproject(self.h, use, a.reshape(-1), b.reshape(-1), st),
                  "eig_lsqproject")
        return a, b, st


    def lsqproject_fill(self, use, itype, npop, K):
        use = np.ascontiguousarray(use, dtype=np.int32)
        it = np.ascontiguousarray(itype, dtype=np.int32)
        n = len(use)
        a = np.zeros((K, n)); b = np.zeros((K, n)); st = np.zeros(n, np.int32)
        self._chk(self._L.eig_lsqproject_fill(self.h, use, it, npop, a.reshape(-1), b.reshape(-1),
                                              st), "eig_lsqproject_fill")
        return a, b, st


class Server:
    """lsqproject server (eig_server_*) started from an Engine whose loadings are
    current.  Query rows are sample-major packed genotypes (see eig_abi.h)."""

    def __init__(self, eng: Engine, eigscale=None, nlanes=0, max_batch=0, chunk=0):
        self._L = eng._L
        es = None if eigscale is None else np.ascontiguousarray(eigscale, dtype=np.float64)
        h = _vp()
        rc = self._L.eig_server_start(eng.h, _ptr(es), nlanes, max_batch, chunk, C.byref(h))
        if rc != 0:
            eng._chk(rc, "eig_server_start")
        self.h = h
        self._keep = {}
        info = self.info()
        self.K, self.nsnp, self.rowpad = info["K"], info["nsnp"], info["rowpad"]
        self.rowbytes = (self.nsnp + 3) // 4

    def _chk(self, rc, what):
        if rc != 0:
            msg = self._L.eig_server_error(self.h)
            raise EigError(f"{what}: rc={rc}: {msg.decode() if msg else ''}")

    def info(self):
        K = C.c_int32(); M = C.c_int64(); rp = C.c_int64(); ch = C.c_int32(); g = C.c_int32()
        nb = C.c_int64(); ns = C.c_int64()
        self._chk(self._L.eig_server_info(self.h, C.byref(K), C.byref(M), C.byref(rp),
                                          C.byref(ch), C.byref(g), C.byref(nb), C.byref(ns)),
                  "eig_server_info")
        return dict(K=K.value, nsnp=M.value, rowpad=rp.value, chunk=ch.value, grid=g.value,
                    batches=nb.value, samples=ns.value)

    def submit(self, rows=None, nq=None, stride=None, device_ptr=None):
        """Host rows: uint8 array [nq, >= ceil(nsnp/4)] (kept alive until polled).
        Device rows: device_ptr (int address), stride and nq."""
        t = C.c_int64()
        if device_ptr is not None:
            rc = self._L.eig_server_submit(self.h, C.c_void_p(device_ptr), stride, nq, 1,
                                           C.byref(t))
            self._chk(rc, "eig_server_submit")
        else:
            rows = np.ascontiguousarray(rows, dtype=np.uint8)
            rc = self._L.eig_server_submit(self.h, _ptr(rows), rows.shape[1], rows.shape[0], 0,
                                           C.byref(t))
            self._chk(rc, "eig_server_submit")
            self._keep[t.value] = rows
            nq = rows.shape[0]
        self._keep.setdefault(t.value, None)
        self._keep[("n", t.value)] = nq
        return t.value

    def poll(self, ticket, wait=True):
        """(coords [nq, K], status [nq], latency_ms) or None while pending."""
        nq = self._keep[("n", ticket)]
        co = np.zeros((nq, self.K)); st = np.zeros(nq, np.int32); lat = C.c_double()
        rc = self._L.eig_server_poll(self.h, ticket, 1 if wait else 0, _ptr(co), _ptr(st),
                                     C.byref(lat))
        if rc == EIG_PENDING:
            return None
        self._keep.pop(ticket, None)
        self._keep.pop(("n", ticket), None)
        self._chk(rc, "eig_server_poll")
        return co, st, lat.value

    def stop(self):
        if self.h:
            self._L.eig_server_stop(self.h)
            self.h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()


def synth_queries(eng: Engine, thr, pop, miss_thr, seed, rows_dev_ptr, stride):
    """eig_synth_queries into caller device memory (integer address)."""
    thr = np.ascontiguousarray(thr, dtype=np.uint32)
    pop = np.ascontiguousarray(pop, dtype=np.int32)
    miss_thr = np.ascontiguousarray(miss_thr, dtype=np.uint32)
    eng._chk(eng._L.eig_synth_queries(eng.h, thr.shape[1], len(pop), thr.shape[0], thr.reshape(-1),
                                      pop, miss_thr, seed, C.c_void_p(rows_dev_ptr), stride),
             "eig_synth_queries")
