/*
 * oracle.c -- CPU restatement of the reference smartpca FP64 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU engine is compared
 * against; it is imported only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product (eig_amd/, the smartpca
 * CLI) links or calls it.
 *
 * Every function restates one piece of the EIGENSOFT 8.0 algorithm and cites the
 * reference file:line it follows.  The summation orders of the reference are kept
 * where they decide bits (the partial-sum lookup XTX, the Cholesky solve), so the
 * oracle reproduces the reference's numbers, not only its mathematics.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_hashit / orc_hasharr  : GENO header KAT "d28ee92a 7a6c3691" (CONVERTF example)
 *   orc_decode_row            : CONVERTF packed/EIGENSTRAT/bed KATs, brute force
 *   orc_snpstats              : integer brute force; closed-form levels
 *   orc_xtx_lookup            : numpy Z^T Z (exact-sum comparison), POPGEN grmjunk
 *   orc_ridoutlier            : planted-outlier closed form
 *   orc_loadings / orc_lsqproj: EIGENSTRAT example .evec KAT, LAPACK dposv, the
 *                               no-missing orthogonal-projection identity
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- hashing of IDs for the GENO header: admutils.c:651-682 ---------------- */
int32_t orc_hashit(const char *s)
{
    uint32_t h = 0; /* int arithmetic with wraparound, done unsigned to stay defined */
    for (; *s; ++s) h = h * 23u + (uint32_t)(int32_t)(signed char)*s;
    return (int32_t)h;
}

int32_t orc_hasharr(const char *const *ids, int32_t n)
{
    uint32_t h = 0;
    for (int32_t i = 0; i < n; ++i) {
        h *= 17u;
        h ^= (uint32_t)orc_hashit(ids[i]);
    }
    return (int32_t)h;
}

/* ---- 2-bit decode: admutils.c:718-735 (rbuff) + 575-592 (3 -> -1) -------------- */
void orc_decode_row(const uint8_t *row, int32_t nind, int32_t *g)
{
    for (int32_t k = 0; k < nind; ++k) {
        int v = (row[k >> 2] >> ((3 - (k & 3)) << 1)) & 3;
        g[k] = (v == 3) ? -1 : v;
    }
}

/* ---- per-SNP column statistics: smartpca.c:3219-3301 + 2282-2313 ----------------
 * Over the axis samples xindex[0..nrows): c0 = sum g, c1 = sum (2-g), nmiss.
 * mean = c0/(nrows-nmiss); p = mean/2 (altnorm) or (c0+1)/(2 ynum + 2);
 * fancy = 1/sqrt(p(1-p)) when positive else 1 (fancynorm) ; levels[g] = (g-mean)*fancy.
 * status: 0 kept, 1 dropped by the keep rule of smartpca.c:1131-1144, 2 all missing. */
void orc_snpstats(const uint8_t *rows, int64_t rlen, int64_t nsnp, const int32_t *xindex,
                  int32_t nrows, int32_t altnorm, int32_t fancynorm, int32_t minallelecnt,
                  int32_t maxmissing, int32_t *c0o, int32_t *c1o, int32_t *nmisso,
                  double *meano, double *fancyo, double *levels3, int32_t *status)
{
    for (int64_t m = 0; m < nsnp; ++m) {
        const uint8_t *row = rows + m * rlen;
        int32_t c0 = 0, c1 = 0, nmiss = 0;
        for (int32_t j = 0; j < nrows; ++j) {
            int32_t k = xindex[j];
            int v = (row[k >> 2] >> ((3 - (k & 3)) << 1)) & 3;
            if (v == 3) { ++nmiss; continue; }
            c0 += v;
            c1 += 2 - v;
        }
        double *lv = levels3 + 3 * m;
        c0o[m] = c0; c1o[m] = c1; nmisso[m] = nmiss;
        if (nmiss == nrows) {
            /* fvadjust_binary returns -999 -> getcolxz_binary1 zeroes everything */
            meano[m] = 0.0; fancyo[m] = 0.0;
            lv[0] = lv[1] = lv[2] = 0.0;
            status[m] = 2;
            continue;
        }
        double ynum = (double)(nrows - nmiss);
        double ysum = (double)c0;
        double ymean = ysum / ynum;
        double yfancy = 1.0;
        if (fancynorm) {
            double p = 0.5 * ymean;
            if (!altnorm) p = (ysum + 1.0) / (2.0 * ynum + 2.0);
            double y = p * (1.0 - p);
            if (y > 0.0) yfancy = 1.0 / sqrt(y);
        }
        lv[0] = (-ymean) * yfancy;
        lv[1] = (1.0 - ymean) * yfancy;
        lv[2] = (2.0 - ymean) * yfancy;
        meano[m] = ymean;
        fancyo[m] = yfancy;
        int32_t t = c0 < c1 ? c0 : c1;
        status[m] = ((t < minallelecnt) || (nmiss > maxmissing) || (t == 0)) ? 1 : 0;
    }
}

/* ---- XTX by the partial-sum lookup: smartpca.c:3304-3328, 3361-3495 -------------
 * Kept SNPs are taken in order, 20 per block.  Each sample gets a 64-bit word of
 * 3-bit codes (hom-ref 0, het 2, hom-alt 3) plus a missing mask (7); SNP b of a
 * block sits at bit 3b + b/5, so every 16-bit chunk carries 5 SNPs.  For each
 * chunk a 32768-entry table holds the sum over its 5 SNPs of the level products
 * t_a * t_b selected by (code_i + code_j) | mask_i | mask_j.  The lower triangle
 * element (i, j) adds the four chunk lookups in the order w0+w1+w2+w3. */
struct xtx_job {
    double *lower;
    const uint64_t *cols, *mmask;
    const double *lut;
    int32_t r0, r1;
};

static void *xtx_rows(void *arg)
{
    struct xtx_job *jb = (struct xtx_job *)arg;
    const double *w0 = jb->lut, *w1 = jb->lut + 32768, *w2 = jb->lut + 65536,
                 *w3 = jb->lut + 98304;
    for (int64_t i = jb->r0; i < jb->r1; ++i) {
        double *out = jb->lower + (i * (i + 1)) / 2;
        uint64_t gi = jb->cols[i], mi = jb->mmask[i];
        for (int64_t j = 0; j <= i; ++j) {
            uint64_t f = (jb->cols[j] + gi) | jb->mmask[j] | mi;
            out[j] += w0[(uint16_t)f] + w1[(uint16_t)(f >> 16)] + w2[(uint16_t)(f >> 32)] +
                      w3[f >> 48];
        }
    }
    return NULL;
}

static void build_lut(const double *tb /* 3 per slot, 20 slots */, double *lut)
{
    for (int q = 0; q < 4; ++q) {
        double inc[5][8];
        for (int s = 0; s < 5; ++s) {
            const double *t = tb + 3 * (5 * q + s);
            inc[s][0] = t[0] * t[0];
            inc[s][1] = 0.0;
            inc[s][2] = t[0] * t[1];
            inc[s][3] = t[0] * t[2];
            inc[s][4] = t[1] * t[1];
            inc[s][5] = t[1] * t[2];
            inc[s][6] = t[2] * t[2];
            inc[s][7] = 0.0;
        }
        double *d = lut + 32768 * q;
        /* index bits: [14:12]=slot4 [11:9]=slot3 [8:6]=slot2 [5:3]=slot1 [2:0]=slot0,
         * summed high slot first as in the reference's nested loops */
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b)
                for (int c = 0; c < 8; ++c)
                    for (int e = 0; e < 8; ++e)
                        for (int f = 0; f < 8; ++f)
                            *d++ = (((inc[4][a] + inc[3][b]) + inc[2][c]) + inc[1][e]) + inc[0][f];
    }
}

/* Bounds splitting rows [0,n) into pieces of about equal triangle area. */
static void tri_split(int32_t n, int32_t pieces, int32_t *b)
{
    b[0] = 0;
    double tot = (double)n * (n + 1) / 2.0;
    for (int32_t p = 1; p < pieces; ++p) {
        double target = tot * p / pieces;
        int32_t r = (int32_t)floor(sqrt(2.0 * target));
        if (r > n) r = n;
        if (r < b[p - 1]) r = b[p - 1];
        b[p] = r;
    }
    b[pieces] = n;
}

static void flush_block(double *lower, uint64_t *cols, uint64_t *mmask, double *tb, double *lut,
                        int32_t nrows, int32_t nthreads)
{
    build_lut(tb, lut);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads > nrows) nthreads = nrows > 0 ? nrows : 1;
    int32_t b[257];
    tri_split(nrows, nthreads, b);
    pthread_t th[256];
    struct xtx_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].lower = lower; jobs[t].cols = cols; jobs[t].mmask = mmask; jobs[t].lut = lut;
        jobs[t].r0 = b[t]; jobs[t].r1 = b[t + 1];
        if (t > 0) pthread_create(&th[t], NULL, xtx_rows, &jobs[t]);
    }
    xtx_rows(&jobs[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    memset(cols, 0, sizeof(uint64_t) * (size_t)nrows);
    memset(mmask, 0, sizeof(uint64_t) * (size_t)nrows);
    memset(tb, 0, sizeof(double) * 60);
}

/* lower: packed lower triangle, nrows(nrows+1)/2 doubles, zeroed by the caller.
 * Only SNPs with status==0 are accumulated (dropped SNPs take no slot). */
int orc_xtx_lookup(const uint8_t *rows, int64_t rlen, int64_t nsnp, const int32_t *xindex,
                   int32_t nrows, const double *levels3, const int32_t *status,
                   int32_t nthreads, double *lower)
{
    uint64_t *cols = calloc((size_t)nrows, sizeof(uint64_t));
    uint64_t *mmask = calloc((size_t)nrows, sizeof(uint64_t));
    double *lut = malloc(sizeof(double) * 131072);
    double tb[60];
    if (!cols || !mmask || !lut) { free(cols); free(mmask); free(lut); return -1; }
    memset(tb, 0, sizeof tb);
    int slot = 0;
    for (int64_t m = 0; m < nsnp; ++m) {
        if (status[m] != 0) continue;
        const uint8_t *row = rows + m * rlen;
        unsigned sh = 3u * slot + slot / 5;
        for (int32_t j = 0; j < nrows; ++j) {
            int32_t k = xindex[j];
            int v = (row[k >> 2] >> ((3 - (k & 3)) << 1)) & 3;
            if (v == 1) cols[j] |= (uint64_t)2 << sh;
            else if (v == 2) cols[j] |= (uint64_t)3 << sh;
            else if (v == 3) mmask[j] |= (uint64_t)7 << sh;
        }
        memcpy(tb + 3 * slot, levels3 + 3 * m, 3 * sizeof(double));
        if (++slot == 20) {
            flush_block(lower, cols, mmask, tb, lut, nrows, nthreads);
            slot = 0;
        }
    }
    if (slot > 0) flush_block(lower, cols, mmask, tb, lut, nrows, nthreads);
    free(cols); free(mmask); free(lut);
    return 0;
}

/* ---- symit2 + trace normalisation: smartpca.c:480-508, 1222-1236 ---------------
 * Unpacks the lower triangle to a full row-major n x n matrix and divides by
 * trace/(n-1) (as a multiply by 1/y).  Returns the scale y, or a negative value
 * for the reference's two fatal cases (NaN trace: -1, trace <= 0: -2). */
double orc_finalize_xtx(const double *lower, int32_t n, double *full)
{
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j <= i; ++j) {
            double v = lower[(i * (i + 1)) / 2 + j];
            full[i * n + j] = v;
            full[j * n + i] = v;
        }
    double tr = 0.0;
    for (int64_t i = 0; i < n; ++i) tr += full[i * n + i];
    double y = tr / (double)(n - 1);
    if (isnan(y)) return -1.0;
    if (y <= 0.0) return -2.0;
    double s = 1.0 / y;
    for (int64_t i = 0; i < (int64_t)n * n; ++i) full[i] *= s;
    return y;
}

/* ---- outlier test: smartsubs.c:18-93 ---------------------------------------------
 * evecs: neigs rows of n (eigenvector j at evecs + j*n).  mode 0: z-score with
 * population s.d.; mode 1: leave-one-out mean/s.d.; mode > 1: off.  Writes the
 * flagged rows in increasing order to bad[], and for each the first eigenvector
 * index and score that flagged it.  Returns the count. */
int32_t orc_ridoutlier(const double *evecs, int32_t n, int32_t neigs, double thresh,
                       int32_t mode, int32_t *bad, int32_t *vecno, double *score)
{
    if (mode > 1 || n < 3) return 0;
    int32_t *flag = calloc((size_t)n, sizeof(int32_t));
    int32_t *vno = malloc(sizeof(int32_t) * (size_t)n);
    double *sc = calloc((size_t)n, sizeof(double));
    double *w = malloc(sizeof(double) * (size_t)n);
    double *w2 = malloc(sizeof(double) * (size_t)n);
    for (int32_t j = 0; j < n; ++j) vno[j] = -1;
    for (int32_t e = 0; e < neigs; ++e) {
        memcpy(w, evecs + (int64_t)e * n, sizeof(double) * (size_t)n);
        if (mode == 0) {
            double s = 0.0;
            for (int32_t j = 0; j < n; ++j) s += w[j];
            double mu = s / (double)n;
            for (int32_t j = 0; j < n; ++j) w[j] += -mu;
            double s2 = 0.0;
            for (int32_t j = 0; j < n; ++j) s2 += w[j] * w[j];
            double sd = sqrt(s2 / (double)n);
            double r = 1.0 / sd;
            for (int32_t j = 0; j < n; ++j) w[j] *= r;
            for (int32_t j = 0; j < n; ++j)
                if (fabs(w[j]) > thresh) {
                    flag[j] = 1;
                    if (vno[j] < 0) { vno[j] = e; sc[j] = w[j]; }
                }
        } else {
            for (int32_t j = 0; j < n; ++j) {
                double yy = w[j];
                w[j] = 0.0;
                double s = 0.0;
                for (int32_t i = 0; i < n; ++i) s += w[i];
                double mu = s / (double)(n - 1);
                for (int32_t i = 0; i < n; ++i) w2[i] = w[i] + (-mu);
                w2[j] = 0.0;
                double s2 = 0.0;
                for (int32_t i = 0; i < n; ++i) s2 += w2[i] * w2[i];
                double sd = sqrt(s2 / (double)n);
                double z = (yy - mu) / sd;
                if (fabs(z) > thresh) {
                    flag[j] = 1;
                    if (vno[j] < 0) { vno[j] = e; sc[j] = z; }
                }
                w[j] = yy;
            }
        }
    }
    int32_t nb = 0;
    for (int32_t j = 0; j < n; ++j)
        if (flag[j]) { bad[nb] = j; vecno[nb] = vno[j]; score[nb] = sc[j]; ++nb; }
    free(flag); free(vno); free(sc); free(w); free(w2);
    return nb;
}

/* ---- SNP loadings: smartpca.c:1438-1494, getcolxf 3564-3595, setfvecs qpsubs.c:489
 * fvecs = 10 * evecs (K rows of nrows).  For each SNP column (over xindex) that is
 * not ignored: z = (g - mean) * fancy with missing 0 (fvadjust), and
 * ffvecs[j*ncols + m] = sum_k fvecs[j][k] * z[k] in sample order. */
void orc_loadings(const uint8_t *rows, int64_t rlen, int64_t ncols, const int32_t *xindex,
                  int32_t nrows, const int32_t *ignore, int32_t altnorm, int32_t fancynorm,
                  const double *evecs, int32_t K, double *ffvecs)
{
    double *cc = malloc(sizeof(double) * (size_t)nrows);
    double *fv = malloc(sizeof(double) * (size_t)nrows * K);
    for (int64_t i = 0; i < (int64_t)nrows * K; ++i) fv[i] = evecs[i] * 10.0;
    for (int64_t m = 0; m < ncols; ++m) {
        for (int j = 0; j < K; ++j) ffvecs[(int64_t)j * ncols + m] = 0.0;
        if (ignore[m]) continue;
        const uint8_t *row = rows + m * rlen;
        double ynum = 0.0, ysum = 0.0;
        for (int32_t k = 0; k < nrows; ++k) {
            int32_t s = xindex[k];
            int v = (row[s >> 2] >> ((3 - (s & 3)) << 1)) & 3;
            cc[k] = (v == 3) ? -1.0 : (double)v;
            if (v != 3) { ynum += 1.0; ysum += (double)v; }
        }
        if (ynum == 0.0) {
            /* Unreachable on the reference path: a SNP with no observed axis genotype
             * is ignored in the final pass (smartpca.c:1131-1144) before loadings, and
             * fvadjust would leave getcolxf's scale uninitialised.  Loading 0. */
            continue;
        }
        double ymean = ysum / ynum, yfancy = 1.0;
        for (int32_t k = 0; k < nrows; ++k) cc[k] = (cc[k] < 0.0) ? 0.0 : cc[k] - ymean;
        if (fancynorm) {
            double p = 0.5 * ymean;
            if (!altnorm) p = (ysum + 1.0) / (2.0 * ynum + 2.0);
            double y = p * (1.0 - p);
            if (y > 0.0) yfancy = 1.0 / sqrt(y);
        }
        for (int32_t k = 0; k < nrows; ++k) cc[k] *= yfancy;
        for (int j = 0; j < K; ++j) {
            double acc = 0.0;
            for (int32_t k = 0; k < nrows; ++k) acc += fv[(int64_t)j * nrows + k] * cc[k];
            ffvecs[(int64_t)j * ncols + m] = acc;
        }
    }
    free(cc); free(fv);
}

/* Row of one individual as fixxrow sees it: qpsubs.c:322-349.  Ignored SNPs and
 * missing genotypes read -1; xrow = g*fancy, then (< -0.001 -> 0) else -= xmean
 * where xmean = mean*fancy. */
static void fix_row(const uint8_t *rows, int64_t rlen, int64_t ncols, int32_t ind,
                    const int32_t *ignore, const double *xmean, const double *xfancy,
                    double *raw, double *xrow)
{
    for (int64_t m = 0; m < ncols; ++m) {
        double g = -1.0;
        if (!ignore[m]) {
            const uint8_t *row = rows + m * rlen;
            int v = (row[ind >> 2] >> ((3 - (ind & 3)) << 1)) & 3;
            g = (v == 3) ? -1.0 : (double)v;
        }
        raw[m] = g;
        double y = g * xfancy[m];
        xrow[m] = (y < -0.001) ? 0.0 : y - xmean[m];
    }
}

/* ---- fxvecs / fxscal: smartpca.c:1504-1525 ---------------------------------------
 * fxvecs[j*nrows + i] = <fixed row of xindex[i], ffvecs_j>; fxscal_j = 1/sqrt(sum_i fxvecs^2).
 * xmean here is mean*fancy per SNP (as stored by getcolxz_binary1). */
void orc_fxscal(const uint8_t *rows, int64_t rlen, int64_t ncols, const int32_t *xindex,
                int32_t nrows, const int32_t *ignore, const double *xmean,
                const double *xfancy, const double *ffvecs, int32_t K, double *fxvecs,
                double *fxscal)
{
    double *raw = malloc(sizeof(double) * (size_t)ncols);
    double *xr = malloc(sizeof(double) * (size_t)ncols);
    for (int j = 0; j < K; ++j) fxscal[j] = 0.0;
    for (int32_t i = 0; i < nrows; ++i) {
        fix_row(rows, rlen, ncols, xindex[i], ignore, xmean, xfancy, raw, xr);
        for (int j = 0; j < K; ++j) {
            double acc = 0.0;
            const double *fp = ffvecs + (int64_t)j * ncols;
            for (int64_t m = 0; m < ncols; ++m) acc += xr[m] * fp[m];
            fxvecs[(int64_t)j * nrows + i] = acc;
            fxscal[j] += acc * acc;
        }
    }
    for (int j = 0; j < K; ++j) fxscal[j] = 1.0 / sqrt(fxscal[j]);
    free(raw); free(xr);
}

/* ---- Cholesky solve: nicksrc/linsubs.c:296-393 (solvit, choldc, cholsl) ---------- */
static int chol_solve(const double *prod, const double *rhs, int n, double *ans)
{
    double a[64 * 64], p[64], b[64];
    for (int i = 0; i < n * n; ++i) if (isnan(prod[i])) return -4;
    for (int i = 0; i < n; ++i) if (isnan(rhs[i])) return -3;
    memcpy(a, prod, sizeof(double) * (size_t)(n * n));
    memcpy(b, rhs, sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) p[i] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            double s = a[i * n + j];
            for (int k = i - 1; k >= 0; --k) s -= a[i * n + k] * a[j * n + k];
            if (i == j) {
                if (s <= 0.0) return -1;
                p[i] = sqrt(s);
            } else {
                a[j * n + i] = s / p[i];
            }
        }
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = i - 1; k >= 0; --k) s -= a[i * n + k] * ans[k];
        ans[i] = s / p[i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = ans[i];
        for (int k = i + 1; k < n; ++k) s -= a[k * n + i] * ans[k];
        ans[i] = s / p[i];
    }
    return 1;
}

/* ---- lsqproject: smartpca.c:4606-4757 + regressit regsubs.c:8-74 ----------------
 * For every individual ind (in .ind order) with use[ind] != 0: over SNPs that are
 * not ignored and not missing for ind, E row = fxscal .* ffvecs[:,m], rhs = fixed
 * x; solve (E^T E) c = E^T x.  Fewer than K+1 observed SNPs -> status 1 (ignored,
 * coordinates untouched).  A failed Cholesky leaves the previous answer in place
 * (regans persists across individuals), status 2.  bcoeffs = fxscal_j <x, ffvecs_j>.
 * acoeffs/bcoeffs are K x numindivs, row j at j*numindivs. */
void orc_lsqproj(const uint8_t *rows, int64_t rlen, int64_t ncols, int32_t numindivs,
                 const int32_t *use, const int32_t *ignore, const double *xmean,
                 const double *xfancy, const double *ffvecs, const double *fxscal, int32_t K,
                 double *acoeffs, double *bcoeffs, int32_t *status)
{
    double *raw = malloc(sizeof(double) * (size_t)ncols);
    double *xr = malloc(sizeof(double) * (size_t)ncols);
    double regans[64];
    for (int j = 0; j < K; ++j) regans[j] = 0.0;
    for (int32_t q = 0; q < numindivs; ++q) {
        status[q] = 0;
        if (!use[q]) { status[q] = 3; continue; }
        fix_row(rows, rlen, ncols, q, ignore, xmean, xfancy, raw, xr);
        double co[64 * 64], rr[64];
        for (int i = 0; i < K * K; ++i) co[i] = 0.0;
        for (int i = 0; i < K; ++i) rr[i] = 0.0;
        int64_t kk = 0;
        double e[64];
        for (int64_t m = 0; m < ncols; ++m) {
            if (raw[m] < 0.0) continue;
            for (int j = 0; j < K; ++j) e[j] = fxscal[j] * ffvecs[(int64_t)j * ncols + m];
            double x = xr[m];
            /* regressit: rr[j] += e_j x ; co[j][k] = co[k][j] += e_j e_k for k >= j */
            for (int j = 0; j < K; ++j) {
                rr[j] += e[j] * x;
                for (int k = j; k < K; ++k) co[j * K + k] = co[k * K + j] += e[j] * e[k];
            }
            ++kk;
        }
        if (kk <= K) { status[q] = 1; continue; }
        double ans[64];
        int ret = chol_solve(co, rr, K, ans);
        if (ret >= 0) memcpy(regans, ans, sizeof(double) * (size_t)K);
        else status[q] = 2;
        for (int j = 0; j < K; ++j) acoeffs[(int64_t)j * numindivs + q] = regans[j];
    }
    /* bcoeffs over the individuals that are still in use */
    for (int32_t q = 0; q < numindivs; ++q) {
        if (!use[q] || status[q] == 1) continue;
        fix_row(rows, rlen, ncols, q, ignore, xmean, xfancy, raw, xr);
        for (int j = 0; j < K; ++j) {
            double acc = 0.0;
            const double *fp = ffvecs + (int64_t)j * ncols;
            for (int64_t m = 0; m < ncols; ++m) acc += xr[m] * fp[m];
            bcoeffs[(int64_t)j * numindivs + q] = fxscal[j] * acc;
        }
    }
    free(raw); free(xr);
}

/* ---- seteigscale: smartpca.c:4758-4782 --------------------------------------------
 * eigscale_j = <a_j, b_j> / <a_j, a_j> over the axis samples xindex. */
void orc_eigscale(const double *acoeffs, const double *bcoeffs, int32_t numindivs,
                  const int32_t *xindex, int32_t nrows, int32_t K, double *eigscale)
{
    for (int j = 0; j < K; ++j) {
        double ab = 0.0, aa = 0.0;
        for (int32_t i = 0; i < nrows; ++i) {
            double a = acoeffs[(int64_t)j * numindivs + xindex[i]];
            double b = bcoeffs[(int64_t)j * numindivs + xindex[i]];
            ab += a * b;
            aa += a * a;
        }
        eigscale[j] = ab / aa;
    }
}

/* ---- lsqproject with usepopsformissing's fill: smartpca.c:4606-4757 ----------------
 * itype[q]: individual q's population type (< 0: none).  First loop (4637-4651): an
 * individual in use with a type and <= K observed genotypes over the SNPs not ignored
 * is ignored (status 1).  Fill table (4655-4678): per type and SNP, the sum and count
 * of observed genotypes over the individuals still in use with a type.  Second loop:
 * a missing genotype of an individual with a type takes sum/count when count > 0.1,
 * then fixxrow; the regression runs over the SNPs whose (filled) genotype is >= 0. */
void orc_lsqproj_fill(const uint8_t *rows, int64_t rlen, int64_t ncols, int32_t numindivs,
                      const int32_t *use, const int32_t *itype, int32_t npop,
                      const int32_t *ignore, const double *xmean, const double *xfancy,
                      const double *ffvecs, const double *fxscal, int32_t K, double *acoeffs,
                      double *bcoeffs, int32_t *status)
{
    double *raw = malloc(sizeof(double) * (size_t)ncols);
    double *xr = malloc(sizeof(double) * (size_t)ncols);
    double *xnum = calloc((size_t)npop * ncols, sizeof(double));
    double *xden = calloc((size_t)npop * ncols, sizeof(double));
    double regans[64];
    for (int j = 0; j < K; ++j) regans[j] = 0.0;
    for (int32_t q = 0; q < numindivs; ++q) {
        status[q] = use[q] ? 0 : 3;
        if (!use[q] || itype[q] < 0) continue;
        fix_row(rows, rlen, ncols, q, ignore, xmean, xfancy, raw, xr);
        int64_t x = 0;
        for (int64_t m = 0; m < ncols; ++m) x += raw[m] >= 0.0;
        if (x <= K) status[q] = 1;
    }
    for (int32_t q = 0; q < numindivs; ++q) {
        if (status[q] != 0 || itype[q] < 0) continue;
        fix_row(rows, rlen, ncols, q, ignore, xmean, xfancy, raw, xr);
        const int32_t t = itype[q];
        for (int64_t m = 0; m < ncols; ++m)
            if (raw[m] >= 0.0) { xnum[t * ncols + m] += raw[m]; xden[t * ncols + m] += 1.0; }
    }
    for (int pass = 0; pass < 2; ++pass)
        for (int32_t q = 0; q < numindivs; ++q) {
            if (status[q] == 1 || status[q] == 3) continue;
            fix_row(rows, rlen, ncols, q, ignore, xmean, xfancy, raw, xr);
            const int32_t t = itype[q];
            if (t >= 0)
                for (int64_t m = 0; m < ncols; ++m) {
                    const int64_t u = t * ncols + m;
                    if (raw[m] < 0.0 && xden[u] > 0.1) raw[m] = xnum[u] / xden[u];
                }
            for (int64_t m = 0; m < ncols; ++m) {
                double y = raw[m] * xfancy[m];
                xr[m] = (y < -0.001) ? 0.0 : y - xmean[m];
            }
            if (pass == 1) { /* bcoeffs */
                for (int j = 0; j < K; ++j) {
                    double acc = 0.0;
                    const double *fp = ffvecs + (int64_t)j * ncols;
                    for (int64_t m = 0; m < ncols; ++m) acc += xr[m] * fp[m];
                    bcoeffs[(int64_t)j * numindivs + q] = fxscal[j] * acc;
                }
                continue;
            }
            double co[64 * 64], rr[64], e[64];
            for (int i = 0; i < K * K; ++i) co[i] = 0.0;
            for (int i = 0; i < K; ++i) rr[i] = 0.0;
            int64_t kk = 0;
            for (int64_t m = 0; m < ncols; ++m) {
                if (raw[m] < 0.0) continue;
                for (int j = 0; j < K; ++j) e[j] = fxscal[j] * ffvecs[(int64_t)j * ncols + m];
                for (int j = 0; j < K; ++j) {
                    rr[j] += e[j] * xr[m];
                    for (int k = j; k < K; ++k) co[j * K + k] = co[k * K + j] += e[j] * e[k];
                }
                ++kk;
            }
            if (kk <= K) { status[q] = 1; continue; }
            double ans[64];
            if (chol_solve(co, rr, K, ans) >= 0) memcpy(regans, ans, sizeof(double) * (size_t)K);
            else status[q] = 2;
            for (int j = 0; j < K; ++j) acoeffs[(int64_t)j * numindivs + q] = regans[j];
        }
    free(raw); free(xr); free(xnum); free(xden);
}

/* ---- solvit for the ldregress oracle (normal equations of up to 64 unknowns) ------ */
int orc_solvit(const double *prod, const double *rhs, int32_t n, double *ans)
{
    return chol_solve(prod, rhs, n, ans);
}
