"""CPU oracle for the smartpca hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  The product (eig_amd/, the smartpca CLI) never
imports, links or runs anything under oracle/.

It restates the reference EIGENSOFT 8.0 smartpca FP64 path:
  * file readers: .snp (mcio.c:415-540), .ind (mcio.c:982-1035), packed GENO with
    the header hash check (mcio.c:2769-2879), EIGENSTRAT text (mcio.c:2528-2656),
    PLINK .bed (mcio.c:3893-4009);
  * the sample/SNP filters of main (smartpca.c:774-946);
  * the outlier loop (smartpca.c:1077-1265): column stats, XTX by the partial-sum
    lookup, symmetrisation and trace scaling, eigvecs = LAPACK dspev on -XTX
    (eigsubs.c:39-55, eigx.c:97-117), ridoutlier;
  * loadings, fxscal, lsqproject and eigscale (smartpca.c:1438-1566);
  * the .eval / .evec writers (smartpca.c:1424-1437, 1568-1591).
The numeric kernels live in oracle.c (liboracle.so); dspev is the real LAPACK
routine from numpy's bundled OpenBLAS (scipy_openblas64, ILP64).
"""
from __future__ import annotations

import ctypes as C
import glob
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LAPACK = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_hashit.restype = C.c_int32
        L.orc_hashit.argtypes = [C.c_char_p]
        L.orc_hasharr.restype = C.c_int32
        L.orc_hasharr.argtypes = [C.POINTER(C.c_char_p), C.c_int32]
        L.orc_decode_row.argtypes = [u8p, C.c_int32, i32p]
        L.orc_snpstats.argtypes = [u8p, C.c_int64, C.c_int64, i32p, C.c_int32, C.c_int32,
                                   C.c_int32, C.c_int32, C.c_int32, i32p, i32p, i32p,
                                   f64p, f64p, f64p, i32p]
        L.orc_xtx_lookup.restype = C.c_int
        L.orc_xtx_lookup.argtypes = [u8p, C.c_int64, C.c_int64, i32p, C.c_int32, f64p,
                                     i32p, C.c_int32, f64p]
        L.orc_finalize_xtx.restype = C.c_double
        L.orc_finalize_xtx.argtypes = [f64p, C.c_int32, f64p]
        L.orc_ridoutlier.restype = C.c_int32
        L.orc_ridoutlier.argtypes = [f64p, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                     i32p, i32p, f64p]
        L.orc_loadings.argtypes = [u8p, C.c_int64, C.c_int64, i32p, C.c_int32, i32p,
                                   C.c_int32, C.c_int32, f64p, C.c_int32, f64p]
        L.orc_fxscal.argtypes = [u8p, C.c_int64, C.c_int64, i32p, C.c_int32, i32p, f64p,
                                 f64p, f64p, C.c_int32, f64p, f64p]
        L.orc_lsqproj.argtypes = [u8p, C.c_int64, C.c_int64, C.c_int32, i32p, i32p, f64p,
                                  f64p, f64p, f64p, C.c_int32, f64p, f64p, i32p]
        L.orc_eigscale.argtypes = [f64p, f64p, C.c_int32, i32p, C.c_int32, C.c_int32, f64p]
        L.orc_lsqproj_fill.argtypes = [u8p, C.c_int64, C.c_int64, C.c_int32, i32p, i32p,
                                       C.c_int32, i32p, f64p, f64p, f64p, f64p, C.c_int32, f64p,
                                       f64p, i32p]
        L.orc_solvit.restype = C.c_int
        L.orc_solvit.argtypes = [f64p, f64p, C.c_int32, f64p]
        _LIB = L
    return _LIB


# ---------------------------------------------------------------- LAPACK dspev
def _lapack():
    global _LAPACK
    if _LAPACK is None:
        d = os.path.join(os.path.dirname(np.__file__), "..", "numpy.libs")
        cands = glob.glob(os.path.join(d, "libscipy_openblas64_*.so"))
        if not cands:
            raise RuntimeError("numpy's bundled OpenBLAS (scipy_openblas64) not found")
        _LAPACK = C.CDLL(cands[0])
    return _LAPACK


def dspev_lower(a_full: np.ndarray):
    """eigvecs (eigsubs.c:39-55): dspev('V','L') of -A in LAPACK packed lower form,
    negated back -> eigenvalues descending, evecs[j] = eigenvector j (unit norm)."""
    n = a_full.shape[0]
    neg = -np.asarray(a_full, dtype=np.float64)
    # packsym (eigsubs.c:145): row-major upper rows == column-major lower columns
    iu = np.triu_indices(n)
    ap = np.ascontiguousarray(neg[iu])
    w = np.zeros(n)
    z = np.zeros((n, n))  # column-major Z(ldz,n) == row-major z[j][i] = comp i of vec j
    work = np.zeros(3 * n)
    L = _lapack()
    fn = L.scipy_dspev_64_
    i64 = C.c_int64
    jobz, uplo = C.c_char(b"V"), C.c_char(b"L")
    nn, ldz, info = i64(n), i64(n), i64(0)
    fn.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(i64), f64p, f64p, f64p,
                   C.POINTER(i64), f64p, C.POINTER(i64), C.c_size_t, C.c_size_t]
    fn(C.byref(jobz), C.byref(uplo), C.byref(nn), ap, w, z.reshape(-1), C.byref(ldz),
       work, C.byref(info), 1, 1)
    if info.value != 0:
        raise RuntimeError(f"dspev info={info.value}")
    return -w, z


def sign_rule(evecs: np.ndarray) -> np.ndarray:
    """Each eigenvector's largest-|.| component (first on ties) made positive.
    LAPACK's sign choice is implementation-defined; both sides apply this rule."""
    out = np.array(evecs, dtype=np.float64, copy=True)
    for j in range(out.shape[0]):
        k = int(np.argmax(np.abs(out[j])))
        if out[j, k] < 0:
            out[j] = -out[j]
    return out


# ---------------------------------------------------------------- small kernels
def hashit(s: str) -> int:
    return lib().orc_hashit(s.encode())


def hasharr(ids) -> int:
    arr = (C.c_char_p * len(ids))(*[i.encode() for i in ids])
    return lib().orc_hasharr(arr, len(ids))


def decode_rows(rows: np.ndarray, nind: int) -> np.ndarray:
    """Returns int32 [M, nind] with -1 for missing (getgtypes)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    out = np.empty((rows.shape[0], nind), dtype=np.int32)
    for m in range(rows.shape[0]):
        lib().orc_decode_row(np.ascontiguousarray(rows[m]), nind, out[m])
    return out


def snpstats(rows, xindex, altnorm=1, fancynorm=1, minallelecnt=1, maxmissing=9999999):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    M, rlen = rows.shape
    xindex = np.ascontiguousarray(xindex, dtype=np.int32)
    n = len(xindex)
    c0 = np.zeros(M, np.int32); c1 = np.zeros(M, np.int32); nm = np.zeros(M, np.int32)
    mean = np.zeros(M); fancy = np.zeros(M); lev = np.zeros((M, 3)); st = np.zeros(M, np.int32)
    lib().orc_snpstats(rows.reshape(-1), rlen, M, xindex, n, altnorm, fancynorm, minallelecnt,
                       maxmissing, c0, c1, nm, mean, fancy, lev.reshape(-1), st)
    return dict(c0=c0, c1=c1, nmiss=nm, mean=mean, fancy=fancy, levels=lev, status=st)


def xtx_lookup(rows, xindex, levels, status, nthreads=8):
    """Packed lower triangle of sum_m z_m z_m^T, reference summation order."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    M, rlen = rows.shape
    xindex = np.ascontiguousarray(xindex, dtype=np.int32)
    n = len(xindex)
    lower = np.zeros(n * (n + 1) // 2)
    rc = lib().orc_xtx_lookup(rows.reshape(-1), rlen, M, xindex, n,
                              np.ascontiguousarray(levels, dtype=np.float64).reshape(-1),
                              np.ascontiguousarray(status, dtype=np.int32), nthreads, lower)
    if rc != 0:
        raise MemoryError("orc_xtx_lookup")
    return lower


def fastmode_levels(mean, fancy, status):
    """setgval's normalisation table (gval.c:66-80): for g = 0,1,2
    ((g - xmean/xfancy) * xfancy) / sqrt(2) with xmean = mean*fancy
    (getcolxz, smartpca.c:3205-3206); column 3 (missing) is 0.  SNPs with no
    observed genotype (status 2) are all zero.  Returns float64 [M, 4]."""
    mean = np.asarray(mean, dtype=np.float64)
    fancy = np.asarray(fancy, dtype=np.float64)
    lf = np.zeros((len(mean), 4))
    ok = np.asarray(status) != 2
    xm = mean[ok] * fancy[ok]
    mu = xm / fancy[ok]
    r2 = math.sqrt(2.0)
    for k in range(3):
        y = float(k) - mu
        y = y * fancy[ok]
        lf[ok, k] = y / r2
    return lf


def fastmode_matrix(rows, xindex, lf, nind):
    """X [ncols, nrows]: X[m][i] = lf[m][g] (kjg_geno_get_normalized_row, gval.c:172-199)."""
    g = decode_rows(rows, nind)[:, np.asarray(xindex)]
    g = np.where(g < 0, 3, g)
    return np.take_along_axis(lf, g, axis=1)


def fpca(X, K, L, I, omega):
    """kjg_fpca (ksrc/kjg_fpca.c:24-101) restated with numpy: X is ncols x nrows,
    omega the nrows x L start panel (the reference draws it from GSL's default
    generator, kjg_gsl.c:96-186).  Q_i = X G (XTXA, 104-128), G <- X^T Q_i / m
    (gsl_matrix_scale at 52), Q_I = X G (XA, 130-153); Q <- left singular vectors
    of Q (dgesvd jobu 'O', kjg_gsl.c:198-209); B = X^T Q (XTB, 155-178); the
    eigenvectors are the first K left singular vectors of B and the eigenvalues
    S^2 / m (79-95).  Returns (evals [K], evecs [K, nrows]) with the largest-|.|
    component of each eigenvector made positive."""
    m, n = X.shape
    Q = np.zeros((m, (I + 1) * L))
    G = np.array(omega, dtype=np.float64)
    for i in range(I):
        Qi = X @ G
        Q[:, i * L:(i + 1) * L] = Qi
        G = (X.T @ Qi) * (1.0 / m)
    Q[:, I * L:] = X @ G
    U = np.linalg.svd(Q, full_matrices=False)[0]
    B = X.T @ U
    UB, S = np.linalg.svd(B, full_matrices=False)[:2]
    return S[:K] ** 2 / m, sign_rule(UB[:, :K].T)


# ---------------------------------------------------------------- non-lookup columns
def _seqsum(x):
    """Sum in index order (the reference's plain loop), not numpy's pairwise sum."""
    x = np.asarray(x, dtype=np.float64)
    return float(np.add.accumulate(x)[-1]) if x.size else 0.0


def normal_columns(rows, xindex, xtypes=None, weight=None, altnorm=1, fancynorm=1,
                   minallelecnt=1, maxmissing=9999999, live=None):
    """getcolxz + fvadjust (smartpca.c:3129-3216, 2236-2279), the weight scaling
    (1150-1152) and the keep rule (1131-1144) for every SNP of rows over the samples
    xindex.  xtypes (len(xindex)): population index per sample for usepopsformissing.
    Returns dict(cols [M, n] -- zero for dropped SNPs, status, c0, c1, nmiss (after the
    fill), mean (filled column), fancy, levels (plain getcolxf levels of kept SNPs))."""
    g_all = decode_rows(rows, int(np.max(xindex)) + 1)[:, xindex]
    M, n = g_all.shape
    out = dict(cols=np.zeros((M, n)), status=np.zeros(M, np.int32), c0=np.zeros(M, np.int32),
               c1=np.zeros(M, np.int32), nmiss=np.zeros(M, np.int32), mean=np.zeros(M),
               fancy=np.zeros(M), levels=np.zeros((M, 4)))
    for m in range(M):
        if live is not None and not live[m]:
            out["status"][m] = 3
            continue
        g = g_all[m]
        obs = g >= 0
        c0 = int(g[obs].sum()); c1 = int((2 - g[obs]).sum())
        x = g.astype(np.float64)
        tt = int((~obs).sum())
        if xtypes is not None and tt > 0:
            k = np.asarray(xtypes)
            npop = int(k.max()) + 1
            psum = np.bincount(k[obs], weights=g[obs], minlength=npop)
            pnum = np.bincount(k[obs], minlength=npop).astype(np.float64)
            tt = 0
            for j in np.nonzero(~obs)[0]:
                if pnum[k[j]] > 0.5:
                    x[j] = psum[k[j]] / pnum[k[j]]
                else:
                    tt += 1
        have = x >= 0
        ynum = float(have.sum())
        out["c0"][m], out["c1"][m], out["nmiss"][m] = c0, c1, tt
        if ynum == 0:
            out["status"][m] = 2
            continue
        ysum = _seqsum(x[have])
        mean = ysum / ynum
        fancy = 1.0
        if fancynorm:
            p = 0.5 * mean if altnorm else (ysum + 1.0) / (2.0 * ynum + 2.0)
            y = p * (1.0 - p)
            if y > 0.0:
                fancy = 1.0 / math.sqrt(y)
        out["mean"][m], out["fancy"][m] = mean, fancy
        t = min(c0, c1)
        if t < minallelecnt or tt > maxmissing or t == 0:
            out["status"][m] = 1
            continue
        col = np.where(have, x - mean, 0.0) * fancy
        if weight is not None:
            col = col * weight[m]
        out["cols"][m] = col
        # getcolxf's plain levels (no fill)
        pm = c0 / float(n - int((~obs).sum()))
        pf = 1.0
        if fancynorm:
            p = 0.5 * pm if altnorm else (c0 + 1.0) / (2.0 * (n - int((~obs).sum())) + 2.0)
            y = p * (1.0 - p)
            if y > 0.0:
                pf = 1.0 / math.sqrt(y)
        out["levels"][m, :3] = (np.arange(3.0) - pm) * pf
    return out


def regressit(eq, rhs):
    """regressit (regsubs.c:8-74) through solvit/choldc/cholsl (nicksrc/linsubs.c:296-393):
    eq m x n, normal equations summed in row order; None when not positive definite."""
    m, n = eq.shape
    co = np.zeros((n, n)); rr = np.zeros(n)
    for i in range(m):
        rr += eq[i] * rhs[i]
        co += np.outer(eq[i], eq[i])
    ans = np.zeros(n)
    ret = lib().orc_solvit(np.ascontiguousarray(co).reshape(-1), np.ascontiguousarray(rr), n, ans)
    return ans if ret >= 0 else None


def ldregress_columns(cols, kept, chrom, ppos, gpos, maxld, ldposlimit=1000000000,
                      ldlimit=9999.0, r2lo=0.01, r2hi=0.95):
    """The ldregress loop of smartpca.c:1146-1170 over the kept SNPs in order: ldregx
    (2751-2826) on the reference's buffer, then bumpldvv (2691-2748).  cols [M, n]
    normalised columns, kept: SNP indices in order.  Returns (residual columns [len(kept), n],
    ynumsnps)."""
    n = cols.shape[1]
    buf, ids = [], []
    lastc = None
    res = np.zeros((len(kept), n))
    ynum = 0.0
    for k, m in enumerate(kept):
        cc = cols[m]
        if lastc != chrom[m]:
            buf, ids = [], []
        # ldregx
        rc, passed = 0, []
        for src in buf:
            ssum = 0.0
            for ii in range(n):  # same order as the reference's loop
                ssum += cc[ii] * src[ii]
            r2 = (ssum / (2 * (n - 1))) ** 2
            if r2 > r2hi:
                rc = 2
                break
            if r2 > r2lo:
                passed.append(src)
        if rc == 2:
            ww = np.zeros(n)
        elif passed:
            S = np.array(passed)  # rsize_pass x n
            ans = regressit(S.T, cc)
            regans = ans if ans is not None else np.zeros(len(passed))
            www = np.zeros(n)
            for a in range(len(passed)):
                www += regans[a] * S[a]
            ww = cc - www
            rc = 1
        else:
            ww = cc.copy()
        if rc < 2:
            ynum += float(np.dot(ww, ww)) / float(np.dot(cc, cc))
        newsrc = cc if rc < 2 else ww
        # bumpldvv: drop the oldest while out of range, append (dropping one when full)
        while buf:
            f = ids[0]
            pdiff = int(np.rint(ppos[m] - ppos[f]))
            gdiff = gpos[m] - gpos[f]
            if pdiff <= ldposlimit and gdiff <= ldlimit:
                break
            buf.pop(0); ids.pop(0)
        if len(buf) == maxld:
            buf.pop(0); ids.pop(0)
        buf.append(newsrc.copy()); ids.append(m)
        lastc = chrom[m]
        res[k] = ww
    return res, ynum


def dense_xtx(cols):
    """domult_increment_normal (smartpca.c:3498-3561) + symit2 + trace scaling
    (480-508, 1222-1236) of the columns [ncols, n]: returns (full XTX / y, y)."""
    X = cols.T @ cols
    n = X.shape[0]
    y = float(np.trace(X)) / (n - 1)
    return X / y, y


def finalize_xtx(lower, n):
    full = np.zeros((n, n))
    y = lib().orc_finalize_xtx(np.ascontiguousarray(lower), n, full.reshape(-1))
    return full, y


def ridoutlier(evecs, neigs, thresh, mode=0):
    evecs = np.ascontiguousarray(evecs, dtype=np.float64)
    n = evecs.shape[1]
    bad = np.zeros(n, np.int32); vno = np.zeros(n, np.int32); sc = np.zeros(n)
    nb = lib().orc_ridoutlier(evecs.reshape(-1), n, neigs, thresh, mode, bad, vno, sc)
    return bad[:nb].copy(), vno[:nb].copy(), sc[:nb].copy()


# ---------------------------------------------------------------- file formats
def read_snp(path):
    """.snp: ID chrom gpos ppos [ref var] (mcio.c:415-540)."""
    ids, chrom, gpos, ppos = [], [], [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            ids.append(t[0]); chrom.append(str2chrom(t[1]))
            gpos.append(float(t[2])); ppos.append(float(t[3]))
    gpos = np.array(gpos); ppos = np.array(ppos)
    if len(gpos) and gpos.max() <= 0.00001:
        gpos = 1.0e-8 * ppos
    elif len(gpos) and gpos.max() > 100:
        gpos = gpos / 100.0
    return dict(ids=ids, chrom=np.array(chrom, np.int32), gpos=gpos, ppos=ppos)


def str2chrom(s):
    u = s.upper()
    if u.startswith("CHR"):
        u = u[3:]
    if u == "X":
        return 23
    if u == "Y":
        return 24
    if u in ("MT", "M"):
        return 90
    if u == "XY":
        return 91
    try:
        return int(u)
    except ValueError:
        return 0


def read_ind(path):
    ids, sex, pop = [], [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            ids.append(t[0]); sex.append(t[1][0]); pop.append(t[2])
    return dict(ids=ids, sex=sex, pop=pop)


def rlen_for(nind):
    return max(48, (2 * nind + 7) // 8)


def read_geno_packed(path, snp_ids, ind_ids, hashcheck=True):
    """Packed GENO (mcio.c:2769-2879).  Returns uint8 [M, rlen]."""
    nind, nsnp = len(ind_ids), len(snp_ids)
    rlen = rlen_for(nind)
    with open(path, "rb") as f:
        hdr = f.read(rlen)
        if hdr[:4] != b"GENO":
            raise ValueError("not a packed GENO file")
        if hashcheck:
            t = hdr.split(b"\0")[0].split()
            xn, xs, xih, xsh = int(t[1]), int(t[2]), int(t[3], 16), int(t[4], 16)
            if xn != nind:
                raise ValueError("number of individuals mismatch")
            if xs != nsnp:
                raise ValueError("number of SNPs mismatch")
            if xih != (hasharr(ind_ids) & 0xFFFFFFFF):
                raise ValueError("indiv file has changed since genotype file was created")
            if xsh != (hasharr(snp_ids) & 0xFFFFFFFF):
                raise ValueError("snp file has changed since genotype file was created")
        data = f.read(rlen * nsnp)
    if len(data) != rlen * nsnp:
        raise ValueError("truncated packed genotype file")
    return np.frombuffer(data, dtype=np.uint8).reshape(nsnp, rlen).copy()


def pack_genotypes(g: np.ndarray) -> np.ndarray:
    """int [M, N] with -1/3 missing -> GENO rows (MSB-first, 0x00 padding)."""
    M, N = g.shape
    rlen = rlen_for(N)
    codes = np.where(g < 0, 3, g).astype(np.uint8)
    pad = np.zeros((M, rlen * 4), np.uint8)
    pad[:, :N] = codes
    pad = pad.reshape(M, rlen, 4)
    return (pad[:, :, 0] << 6 | pad[:, :, 1] << 4 | pad[:, :, 2] << 2 | pad[:, :, 3]).astype(np.uint8)


def read_eigenstrat(path, nind):
    rows = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s:
                continue
            rows.append([(-1 if c == "9" else int(c)) for c in s[:nind]])
    return pack_genotypes(np.array(rows, dtype=np.int32))


_BED_MAP = np.array([2, -1, 1, 0], dtype=np.int32)  # ancval: bed 00,01,10,11


def read_bed(path, nind, nsnp):
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:3] != bytes([0x6C, 0x1B, 0x01]):
        raise ValueError("bad .bed magic")
    blen = (nind + 3) // 4
    body = np.frombuffer(raw[3:3 + blen * nsnp], dtype=np.uint8).reshape(nsnp, blen)
    sh = np.array([0, 2, 4, 6], dtype=np.uint8)
    codes = ((body[:, :, None] >> sh) & 3).reshape(nsnp, blen * 4)[:, :nind]
    return pack_genotypes(_BED_MAP[codes])


def geno_header(nind, nsnp, ihash, shash):
    s = "GENO %7d %7d %x %x" % (nind, nsnp, ihash & 0xFFFFFFFF, shash & 0xFFFFFFFF)
    return s.encode().ljust(rlen_for(nind), b"\0")


# ---------------------------------------------------------------- parfile
def read_par(path):
    """key: value lines, '#' comments (nicksrc/getpars.c)."""
    out = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            t = line.split(None, 1)
            k = t[0].rstrip(":")
            out[k] = t[1].strip() if len(t) > 1 else ""
    return out


def _yes(v):
    v = str(v).strip().upper()
    if v in ("YES", "Y", "TRUE"):
        return 1
    if v in ("NO", "N", "FALSE"):
        return 0
    return int(float(v))


# ---------------------------------------------------------------- pipeline
def smartpca(par: dict, base_dir: str = ".", geno_rows=None, nthreads=8, verbose=False):
    """Runs the reference FP64 smartpca path.  par holds parfile keys (no colon).
    Returns dict(evals, evec_lines, eval_lines, removed, ...)."""
    def P(k, d=None):
        return par.get(k, d)

    def path(k):
        v = P(k)
        return v if v is None or os.path.isabs(v) else os.path.join(base_dir, v)

    numeigs = int(P("numoutevec", P("numeigs", 10)))
    numoutliter = int(P("numoutlieriter", P("numoutliter", 5)))
    numoutleigs = int(P("numoutlierevec", 10))
    thresh = float(P("outliersigmathresh", P("outlthresh", 6.0)))
    outliermode = int(P("outliermode", 0))
    altnorm = _yes(P("altnormstyle", "YES"))
    fancynorm = _yes(P("usenorm", P("fancynorm", "YES")))
    minallelecnt = int(P("minallelecnt", 1))
    maxmissing = int(P("maxmissing", 9999999))
    noxdata = _yes(P("noxdata", "YES"))
    numchrom = int(P("numchrom", 22))
    hiprec = _yes(P("hiprecision", P("hiprec", "NO")))
    regmode = _yes(P("lsqproject", P("regmode", "YES")))
    usepops = _yes(P("usepopsformissing", "NO"))
    ldregress = int(P("ldregress", P("nsnpldregress", 0)))
    ldlimit = float(P("maxdistldregress", P("ldlimit", 9999.0)))
    ldposlimit = int(P("ldposlimit", 1000000000))
    ldr2lo, ldr2hi = float(P("ldr2lo", 0.01)), float(P("ldr2hi", 0.95))
    if ldlimit <= 0 or ldposlimit <= 0:
        ldregress = 0
    weightname = P("weightname")
    normal = usepops or ldregress > 0 or weightname is not None

    snp = read_snp(path("snpname"))
    ind = read_ind(path("indivname"))
    N, M = len(ind["ids"]), len(snp["ids"])
    if geno_rows is None:
        gp = path("genotypename")
        with open(gp, "rb") as f:
            head = f.read(4)
        if head == b"GENO":
            geno_rows = read_geno_packed(gp, snp["ids"], ind["ids"],
                                         hashcheck=_yes(P("hashcheck", "YES")))
        elif head[:3] == bytes([0x6C, 0x1B, 0x01]):
            geno_rows = read_bed(gp, N, M)
        else:
            geno_rows = read_eigenstrat(gp, N)
    rlen = geno_rows.shape[1]

    ind_ignore = np.array([p == "Ignore" for p in ind["pop"]], dtype=bool)
    # poplistname -> axis pops (egsubs.c:86-109); default: all non-ignored
    poplist = P("poplistname")
    if poplist:
        with open(path("poplistname")) as f:
            axis_pops = {l.split()[0] for l in f if l.split() and not l.startswith("#")}
        aff = np.array([p in axis_pops for p in ind["pop"]]) & ~ind_ignore
    else:
        aff = ~ind_ignore
    g_all = decode_rows(geno_rows, N)  # [M, N] raw codes, -1 missing
    snp_ignore = np.zeros(M, dtype=bool)
    ch = snp["chrom"]
    snp_ignore |= (noxdata == 1) & (ch == numchrom + 1)
    snp_ignore |= ch == 0
    snp_ignore |= ch > numchrom + 1
    # nodata: <= 1 valid genotype among non-ignored axis samples (smartpca.c:865-871)
    valid_aff = ((g_all >= 0) & (aff & ~ind_ignore)[None, :]).sum(axis=1)
    snp_ignore |= (~snp_ignore) & (valid_aff <= 1)
    # individuals with no valid genotype over kept SNPs are ignored (smartpca.c:885-897)
    vv = ((g_all >= 0) & (~snp_ignore)[:, None]).sum(axis=0)
    ind_ignore |= vv == 0
    keep_snps = np.nonzero(~snp_ignore)[0]
    rows = geno_rows[keep_snps]
    # axis populations in first-appearance order (makeeglist) and each sample's index
    if poplist:
        eglist = [l.split()[0] for l in open(path("poplistname")) if l.split() and not l.startswith("#")]
    else:
        eglist = []
        for i in range(N):
            if not ind_ignore[i] and ind["pop"][i] not in eglist:
                eglist.append(ind["pop"][i])
    popidx = np.array([eglist.index(p) if p in eglist else -1 for p in ind["pop"]], np.int32)
    xt_quirk = np.full(N, -2, np.int32)
    weight = None
    if weightname:
        ids_k = [snp["ids"][m] for m in keep_snps]
        weight = np.ones(len(keep_snps))
        wp = weightname if os.path.isabs(weightname) else os.path.join(base_dir, weightname)
        for line in open(wp):
            f = line.split()
            if not f or f[0].startswith("#") or f[0] in ("SNP_ID", "Indiv_ID", "Chr"):
                continue
            if len(f) > 1 and f[0] in ids_k:
                weight[ids_k.index(f[0])] = float(f[1])
    ynumsnps = 0.0

    g_k = g_all[keep_snps]
    Mk = len(keep_snps)
    snp_dead = np.zeros(Mk, dtype=bool)  # cupt->ignore set during passes

    removed = []
    numoutiter = numoutliter + 1 if numoutliter >= 1 else 1
    omode = outliermode if numoutliter >= 1 else 2
    for outliter in range(1, numoutiter + 1):
        xindex = np.nonzero(~ind_ignore & aff)[0].astype(np.int32)
        nrows = len(xindex)
        # loadsnpx: live SNPs with >= 1 valid genotype over non-ignored indivs
        live = ~snp_dead & (((g_k >= 0) & (~ind_ignore)[None, :]).sum(axis=1) > 0)
        cols = np.nonzero(live)[0]
        xt_quirk[:nrows] = popidx[xindex]
        if normal:
            st = normal_columns(rows[cols], xindex, popidx[xindex] if usepops else None,
                                None if weight is None else weight[cols], altnorm, fancynorm,
                                minallelecnt, maxmissing)
            killed = st["status"] != 0
            snp_dead[cols[killed]] = True
            kc = np.nonzero(~killed)[0]
            if ldregress > 0:
                ks = keep_snps[cols]
                Z, ynumsnps = ldregress_columns(
                    st["cols"], kc, snp["chrom"][ks], snp["ppos"][ks], snp["gpos"][ks],
                    ldregress, ldposlimit, ldlimit, ldr2lo, ldr2hi)
            else:
                Z, ynumsnps = st["cols"][kc], float(len(kc))
            full, y = dense_xtx(Z)
        else:
            st = snpstats(rows[cols], xindex, altnorm, fancynorm, minallelecnt, maxmissing)
            killed = st["status"] != 0
            snp_dead[cols[killed]] = True
            lower = xtx_lookup(rows[cols], xindex, st["levels"], st["status"], nthreads)
            full, y = finalize_xtx(lower, nrows)
            ynumsnps = float(int((~killed).sum()))
        if y == -1.0:
            raise RuntimeError("bad XTX matrix")
        if y == -2.0:
            raise RuntimeError("XTX has zero trace (perhaps no data)")
        evals, evecs = dspev_lower(full)
        if outliter > numoutliter:
            break
        nle = min(numoutleigs, nrows - 1)
        bad, vno, sc = ridoutlier(evecs, nle, thresh, omode)
        if len(bad) == 0:
            break
        for b, v, s in zip(bad, vno, sc):
            i = int(xindex[b])
            removed.append((ind["ids"][i], outliter, int(v), float(s), ind["pop"][i]))
            ind_ignore[i] = True

    ncols = len(cols)
    K = min(numeigs, nrows, ncols)
    ev_sign = sign_rule(evecs[:K])
    res = dict(evals=evals, evecs=ev_sign, xindex=xindex, removed=removed, K=K,
               xtx=full, trace_scale=y, stats=st, cols=keep_snps[cols], ynumsnps=ynumsnps)
    eval_lines = ["%12.6f" % v for v in evals]
    head = "%20s " % "#eigvals:" + "".join("%9.3f " % evals[j] for j in range(K))
    if K == 0:
        res.update(eval_lines=eval_lines, evec_lines=[head])
        return res
    crow = np.ascontiguousarray(rows[cols])
    ignore_c = np.ascontiguousarray(killed.astype(np.int32))
    ffvecs = np.zeros((K, ncols))
    L = lib()
    L.orc_loadings(crow.reshape(-1), rlen, ncols, xindex, nrows, ignore_c, altnorm, fancynorm,
                   np.ascontiguousarray(ev_sign), K, ffvecs.reshape(-1))
    xmean = st["mean"] * st["fancy"]
    xfancy = np.ascontiguousarray(st["fancy"])
    fxvecs = np.zeros((K, nrows)); fxscal = np.zeros(K)
    L.orc_fxscal(crow.reshape(-1), rlen, ncols, xindex, nrows, ignore_c, xmean, xfancy,
                 ffvecs.reshape(-1), K, fxvecs.reshape(-1), fxscal)
    acoeffs = np.zeros((K, N)); bcoeffs = np.zeros((K, N)); status = np.zeros(N, np.int32)
    use = (~ind_ignore).astype(np.int32)
    if regmode:
        if usepops:
            L.orc_lsqproj_fill(crow.reshape(-1), rlen, ncols, N, use, np.ascontiguousarray(xt_quirk),
                               len(eglist), ignore_c, xmean, xfancy, ffvecs.reshape(-1), fxscal, K,
                               acoeffs.reshape(-1), bcoeffs.reshape(-1), status)
        else:
            L.orc_lsqproj(crow.reshape(-1), rlen, ncols, N, use, ignore_c, xmean, xfancy,
                          ffvecs.reshape(-1), fxscal, K, acoeffs.reshape(-1),
                          bcoeffs.reshape(-1), status)
        eigscale = np.zeros(K)
        L.orc_eigscale(acoeffs.reshape(-1), bcoeffs.reshape(-1), N, xindex, nrows, K, eigscale)
        acoeffs = acoeffs * eigscale[:, None]
    ind_ignore |= status == 1
    lines = [head]
    fmt = "%12.6f  " if hiprec else "%10.4f  "
    for i in range(N):
        if ind_ignore[i]:
            continue
        lines.append("%20s " % ind["ids"][i] + "".join(fmt % acoeffs[j, i] for j in range(K))
                     + "%15s" % ind["pop"][i])
    res.update(eval_lines=eval_lines, evec_lines=lines, ffvecs=ffvecs, fxscal=fxscal,
               acoeffs=acoeffs, bcoeffs=bcoeffs, lsq_status=status,
               eigscale=eigscale if regmode else None, xmean=xmean, xfancy=xfancy,
               snp_ignore_final=ignore_c)
    return res
