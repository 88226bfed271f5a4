// The non-lookup column path of the XTX pass: explicit FP64 columns instead of the
// 2-bit level lookup, for the three reference options that need them --
//   usepopsformissing: a missing genotype takes the mean genotype of the sample's
//     population at that SNP (getcolxz, smartpca.c:3129-3216);
//   weightname: each column is scaled by its SNP's weight (smartpca.c:1150-1152);
//   ldregress: each column is replaced by its residual after regression on the
//     preceding kept SNPs of the same chromosome within the distance limits whose
//     squared correlation exceeds ldr2lo; a column correlated above ldr2hi with any
//     of them is dropped (ldregx, bumpldvv, smartpca.c:2691-2826, 1155-1170);
// then XTX_lower += sum over SNPs of r r^T (domult_increment_normal,
// smartpca.c:3498-3561) and the same trace scaling as the lookup path.
//
// Device work per chunk of kept SNPs: the columns are materialised from the tile
// layout (k_normcols), their Gram matrix and the residual columns are FP64 MFMA
// products (eig_gemm_f64), the drop decisions run in SNP order in one block
// (k_ld_scan), each SNP's normal equations are solved by one block (k_ld_solve,
// choldc/cholsl order), and the XTX update is a lower-tile FP64 MFMA product.
// Built without FMA contraction: the column statistics round as the reference's C.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "eig_internal.h"

namespace {
constexpr int kMaxPop = 1024;  // populations of usepopsformissing (LDS tables)

__device__ inline uint32_t sel_code(const uint32_t *__restrict__ sel, int64_t M, int32_t i,
                                    int64_t m) {
  const uint32_t w = sel[((int64_t)(i >> 7) * M + m) * kTileWords + ((i & 127) >> 4)];
  return (w >> (2 * (i & 15))) & 3u;
}

template <typename T>
__device__ inline T block_sum(T v, T *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  T r = red[0];
  __syncthreads();
  return r;
}

// Per SNP (one block each): observed counts, population sums, the filled column's
// mean and normalisation (fvadjust, smartpca.c:2236-2279), the keep rule
// (smartpca.c:1131-1144) with the missing count after the fill, and the plain
// levels of getcolxf (smartpca.c:3564-3595) for the loadings.
__global__ __launch_bounds__(256) void k_normstats(
    const uint32_t *__restrict__ sel, int64_t M, int32_t nrows, const uint8_t *__restrict__ live,
    const int32_t *__restrict__ xtypes, int32_t npop, int32_t altnorm, int32_t fancynorm,
    int32_t minallelecnt, int32_t maxmissing, int32_t *__restrict__ c0o, int32_t *__restrict__ c1o,
    int32_t *__restrict__ nmisso, int32_t *__restrict__ stato, double *__restrict__ meano,
    double *__restrict__ fancyo, double *__restrict__ lev) {
  __shared__ int32_t psum[kMaxPop], pnum[kMaxPop];
  __shared__ double dred[256];
  __shared__ int32_t ired[256];
  const int64_t m = blockIdx.x;
  const int t = threadIdx.x;
  if (live && !live[m]) {
    if (t == 0) {
      c0o[m] = c1o[m] = nmisso[m] = 0;
      stato[m] = EIG_SNP_DEAD;
      meano[m] = fancyo[m] = 0.0;
      lev[4 * m] = lev[4 * m + 1] = lev[4 * m + 2] = lev[4 * m + 3] = 0.0;
    }
    return;
  }
  const bool pops = xtypes != nullptr;
  if (pops)
    for (int k = t; k < npop; k += 256) psum[k] = pnum[k] = 0;
  __syncthreads();
  int32_t c0 = 0, c1 = 0, nm = 0;
  for (int32_t i = t; i < nrows; i += 256) {
    const uint32_t g = sel_code(sel, M, i, m);
    if (g == 3u) {
      ++nm;
      continue;
    }
    c0 += (int32_t)g;
    c1 += 2 - (int32_t)g;
    if (pops) {
      atomicAdd(&psum[xtypes[i]], (int32_t)g);
      atomicAdd(&pnum[xtypes[i]], 1);
    }
  }
  c0 = block_sum(c0, ired);
  c1 = block_sum(c1, ired);
  nm = block_sum(nm, ired);
  // the filled column: observed g, the population mean for a missing genotype of a
  // population with data, else missing
  double ysum = 0.0;
  int32_t ynum = 0, tt = 0;
  for (int32_t i = t; i < nrows; i += 256) {
    const uint32_t g = sel_code(sel, M, i, m);
    if (g != 3u) {
      ysum += (double)g;
      ++ynum;
    } else if (pops && pnum[xtypes[i]] > 0) {
      ysum += (double)psum[xtypes[i]] / (double)pnum[xtypes[i]];
      ++ynum;
    } else {
      ++tt;
    }
  }
  ysum = block_sum(ysum, dred);
  ynum = block_sum(ynum, ired);
  tt = block_sum(tt, ired);
  if (t != 0) return;
  auto norm = [&](double s, double num, double &mean, double &fancy) {
    mean = s / num;
    fancy = 1.0;
    if (fancynorm) {
      double p = 0.5 * mean;
      if (!altnorm) p = (s + 1.0) / (2.0 * num + 2.0);
      const double y = p * (1.0 - p);
      if (y > 0.0) fancy = 1.0 / sqrt(y);
    }
  };
  int32_t status;
  double mean = 0.0, fancy = 0.0;
  if (ynum == 0) {
    status = EIG_SNP_ALLMISS;  // getcolxz returns -1: xmean = xfancy = 0
  } else {
    norm(ysum, (double)ynum, mean, fancy);
    const int32_t mn = c0 < c1 ? c0 : c1;
    status = (mn < minallelecnt || tt > maxmissing || mn == 0) ? EIG_SNP_DROPPED : EIG_SNP_KEPT;
  }
  c0o[m] = c0;
  c1o[m] = c1;
  nmisso[m] = tt;
  stato[m] = status;
  meano[m] = mean;
  fancyo[m] = fancy;
  double l0 = 0.0, l1 = 0.0, l2 = 0.0;
  if (status == EIG_SNP_KEPT) {  // getcolxf: the plain column, no fill
    double pm, pf;
    norm((double)c0, (double)(nrows - nm), pm, pf);
    l0 = (0.0 - pm) * pf;
    l1 = (1.0 - pm) * pf;
    l2 = (2.0 - pm) * pf;
  }
  lev[4 * m] = l0;
  lev[4 * m + 1] = l1;
  lev[4 * m + 2] = l2;
  lev[4 * m + 3] = 0.0;
}

// Columns of the SNPs idx[b] (kept, in order): Z[b * ldz + i] = ((y_i - mean) * fancy) * w
// with the filled values y of k_normstats, missing (unfilled) 0; rows nrows..ldz-1 are 0.
__global__ __launch_bounds__(256) void k_normcols(
    const uint32_t *__restrict__ sel, int64_t M, int32_t nrows, const int32_t *__restrict__ xtypes,
    int32_t npop, const double *__restrict__ mean, const double *__restrict__ fancy,
    const double *__restrict__ weight, const int32_t *__restrict__ idx, int64_t ldz,
    double *__restrict__ Z) {
  __shared__ int32_t psum[kMaxPop], pnum[kMaxPop];
  const int64_t m = idx[blockIdx.x];
  const int t = threadIdx.x;
  const bool pops = xtypes != nullptr;
  if (pops) {
    for (int k = t; k < npop; k += 256) psum[k] = pnum[k] = 0;
    __syncthreads();
    for (int32_t i = t; i < nrows; i += 256) {
      const uint32_t g = sel_code(sel, M, i, m);
      if (g != 3u) {
        atomicAdd(&psum[xtypes[i]], (int32_t)g);
        atomicAdd(&pnum[xtypes[i]], 1);
      }
    }
    __syncthreads();
  }
  const double mu = mean[m], f = fancy[m], w = weight ? weight[m] : 1.0;
  double *z = Z + (int64_t)blockIdx.x * ldz;
  for (int64_t i = t; i < ldz; i += 256) {
    double v = 0.0;
    if (i < nrows) {
      const uint32_t g = sel_code(sel, M, (int32_t)i, m);
      bool have = g != 3u;
      double y = (double)g;
      if (!have && pops && pnum[xtypes[i]] > 0) {
        y = (double)psum[xtypes[i]] / (double)pnum[xtypes[i]];
        have = true;
      }
      if (have) {
        v = (y - mu) * f;
        if (weight) v = v * w;
      }
    }
    z[i] = v;
  }
}

// ldregress drop decisions, in kept-SNP order, one block.  SNP k (position k of the
// kept list) regresses on positions ws[k] .. k-1; G is the Gram matrix of the chunk's
// columns (positions h0 .. h0+nc-1, ld nc).  A source dropped earlier holds a zero
// column in the reference's buffer, so it neither drops nor enters a regression.
// Output per SNP of the chunk: dropped[k] and its sources' positions (oldest first),
// nsrc = -1 for a dropped SNP.  The chunk's drop flags live in LDS while it runs.
constexpr int kMaxLdNc = 8192;  // chunk + window positions held in LDS
__global__ __launch_bounds__(256) void k_ld_scan(const double *__restrict__ G, int64_t nc,
                                                 int64_t h0, int64_t k0, int64_t k1,
                                                 const int64_t *__restrict__ ws, double scale,
                                                 double r2lo, double r2hi,
                                                 uint8_t *__restrict__ dropped,
                                                 int32_t *__restrict__ nsrc,
                                                 int32_t *__restrict__ src, int32_t W) {
  __shared__ uint8_t dl[kMaxLdNc];
  __shared__ int32_t cnt[256];
  __shared__ int32_t total;
  const int t = threadIdx.x;
  for (int64_t p = t; p < nc; p += 256) dl[p] = (h0 + p < k0) ? dropped[h0 + p] : 0;
  __syncthreads();
  for (int64_t k = k0; k < k1; ++k) {
    const int64_t a = ws[k], len = k - a;
    const double *gk = G + (k - h0) * nc;
    int kill = 0;
    if (t == 0) total = 0;
    __syncthreads();
    for (int64_t base = 0; base < len; base += 256) {
      const int64_t p = a + base + t;
      int sel = 0;
      if (base + t < len && !dl[p - h0]) {
        const double c = gk[p - h0] * scale;
        const double r2 = c * c;
        if (r2 > r2hi) kill = 1;
        else if (r2 > r2lo) sel = 1;
      }
      // compact the selected sources in position order (inclusive prefix sum)
      cnt[t] = sel;
      __syncthreads();
      for (int o = 1; o < 256; o <<= 1) {
        const int v = t >= o ? cnt[t - o] : 0;
        __syncthreads();
        cnt[t] += v;
        __syncthreads();
      }
      const int32_t before = total;
      if (sel) src[(k - k0) * (int64_t)W + before + cnt[t] - 1] = (int32_t)p;
      __syncthreads();
      if (t == 0) total = before + cnt[255];
      __syncthreads();
    }
    kill = __syncthreads_or(kill);
    if (t == 0) {
      dl[k - h0] = (uint8_t)kill;
      dropped[k] = (uint8_t)kill;
      nsrc[k - k0] = kill ? -1 : total;
    }
    __syncthreads();
  }
}

// One block per SNP of the chunk: the residual's coefficient column Bc[:, c] (ld nc,
// c = k - k0): e_k minus the least-squares coefficients on the sources, from the
// normal equations in G solved as choldc/cholsl do (nicksrc/linsubs.c:330-393); a
// dropped SNP gets a zero column (zero residual), a non positive definite system
// leaves the coefficients 0 (ldregx's fresh regans).  work: nc_max^2 + 2 nc_max doubles
// per block.
__global__ __launch_bounds__(256) void k_ld_solve(const double *__restrict__ G, int64_t nc,
                                                  int64_t h0, int64_t k0,
                                                  const int32_t *__restrict__ nsrc,
                                                  const int32_t *__restrict__ src, int32_t W,
                                                  double *__restrict__ work,
                                                  double *__restrict__ Bc) {
  const int64_t c = blockIdx.x, k = k0 + c;
  const int t = threadIdx.x;
  double *col = Bc + c * nc;
  for (int64_t r = t; r < nc; r += 256) col[r] = 0.0;
  __syncthreads();
  const int32_t s = nsrc[c];
  if (s < 0) return;  // dropped: zero residual
  if (t == 0) col[k - h0] = 1.0;
  if (s == 0) return;
  const int32_t *sp = src + c * (int64_t)W;
  double *a = work + c * ((int64_t)W * W + 2 * W), *p = a + (int64_t)W * W, *x = p + W;
  for (int32_t e = t; e < s * s; e += 256) {
    const int32_t i = e / s, j = e % s;
    a[e] = G[(sp[i] - h0) + (sp[j] - h0) * nc];
  }
  __shared__ int bad;
  if (t == 0) bad = 0;
  __syncthreads();
  for (int32_t i = 0; i < s; ++i) {
    // row i of the factor: the diagonal first, then the column below it in parallel
    if (t == 0) {
      double sum = a[i * s + i];
      for (int32_t q = i - 1; q >= 0; --q) sum -= a[i * s + q] * a[i * s + q];
      if (sum <= 0.0) bad = 1;
      else p[i] = sqrt(sum);
    }
    __syncthreads();
    if (bad) return;  // coefficients stay 0: residual = the column itself
    for (int32_t j = i + 1 + t; j < s; j += 256) {
      double sum = a[i * s + j];
      for (int32_t q = i - 1; q >= 0; --q) sum -= a[i * s + q] * a[j * s + q];
      a[j * s + i] = sum / p[i];
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int32_t i = 0; i < s; ++i) {
      double sum = G[(sp[i] - h0) + (k - h0) * nc];
      for (int32_t q = i - 1; q >= 0; --q) sum -= a[i * s + q] * x[q];
      x[i] = sum / p[i];
    }
    for (int32_t i = s - 1; i >= 0; --i) {
      double sum = x[i];
      for (int32_t q = i + 1; q < s; ++q) sum -= a[q * s + i] * x[q];
      x[i] = sum / p[i];
    }
    for (int32_t i = 0; i < s; ++i) col[sp[i] - h0] = -x[i];
  }
}

// squared norms of the nb columns of R (ld ldz, n rows), in row order per column
__global__ __launch_bounds__(256) void k_colnorm2(const double *__restrict__ R, int64_t ldz,
                                                  int32_t n, double *__restrict__ out) {
  __shared__ double red[256];
  const double *r = R + (int64_t)blockIdx.x * ldz;
  double s = 0.0;
  for (int32_t i = threadIdx.x; i < n; i += 256) s += r[i] * r[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// trace of the n x n matrix X (column-major, lower tiles formed), in diagonal order
__global__ void k_dense_trace(const double *__restrict__ X, int32_t n, double *__restrict__ tr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int32_t i = 0; i < n; ++i) s += X[(int64_t)i * n + i];
  *tr = s;
}

// X[i][j] = X[j][i] = X_lower[max][min] / y for the whole matrix (symmetric result)
__global__ __launch_bounds__(256) void k_dense_scale_mirror(double *__restrict__ X, int32_t n,
                                                            const double *__restrict__ tr) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int64_t i = e / n, j = e % n;
  if (i < j) return;  // each lower element writes itself and its mirror
  const double y = *tr / (double)(n - 1);
  const double v = X[i + j * (int64_t)n] / y;
  X[i + j * (int64_t)n] = v;
  X[j + i * (int64_t)n] = v;
}

// lsqproj's fill table (smartpca.c:4656-4678): over the individuals q with
// itype[q] >= 0, fill[m * npop + t] = mean observed genotype of type t at SNP m, or NaN
// when the type has none there (xden <= 0.1).  One block per SNP.
__global__ __launch_bounds__(256) void k_lsq_filltab(const uint8_t *__restrict__ geno,
                                                     int64_t rlen, int32_t nind,
                                                     const int32_t *__restrict__ itype,
                                                     int32_t npop, double *__restrict__ fill) {
  __shared__ int32_t psum[kMaxPop], pnum[kMaxPop];
  const int64_t m = blockIdx.x;
  const int t = threadIdx.x;
  for (int k = t; k < npop; k += 256) psum[k] = pnum[k] = 0;
  __syncthreads();
  const uint8_t *row = geno + m * rlen;
  for (int32_t q = t; q < nind; q += 256) {
    const int32_t ty = itype[q];
    if (ty < 0) continue;
    const uint32_t g = (row[q >> 2] >> ((3 - (q & 3)) << 1)) & 3u;
    if (g == 3u) continue;
    atomicAdd(&psum[ty], (int32_t)g);
    atomicAdd(&pnum[ty], 1);
  }
  __syncthreads();
  for (int k = t; k < npop; k += 256)
    fill[m * npop + k] = pnum[k] > 0 ? (double)psum[k] / (double)pnum[k] : __builtin_nan("");
}

// Allele counts per SNP and population over the selected samples (one block per SNP):
// out[(m * npop + t) * 2 + {0, 1}] = sum of g and of 2 - g over the observed samples of
// population t; on an X-chromosome SNP a haploid sample counts g/2 and 1 - g/2 and a
// heterozygous call of one is skipped (getrawcolx, qpsubs.c:251-283; fstcolyy 1205-1262)
__global__ __launch_bounds__(256) void k_popcounts(const uint32_t *__restrict__ sel, int64_t M,
                                                   int32_t nrows,
                                                   const int32_t *__restrict__ xtypes,
                                                   int32_t npop,
                                                   const uint8_t *__restrict__ haploid,
                                                   const uint8_t *__restrict__ xsnp,
                                                   int32_t *__restrict__ out) {
  __shared__ int32_t ca[kMaxPop], cb[kMaxPop];
  const int64_t m = blockIdx.x;
  const int t = threadIdx.x;
  for (int k = t; k < npop; k += 256) ca[k] = cb[k] = 0;
  __syncthreads();
  const bool onx = xsnp && xsnp[m];
  for (int32_t i = t; i < nrows; i += 256) {
    const int32_t ty = xtypes[i];
    if (ty < 0 || ty >= npop) continue;
    int32_t g = (int32_t)sel_code(sel, M, i, m);
    if (g == 3) continue;
    int32_t a = g, b = 2 - g;
    if (onx && haploid && haploid[i]) {
      if (g == 1) continue;
      a = g / 2;
      b = 1 - a;
    }
    atomicAdd(&ca[ty], a);
    atomicAdd(&cb[ty], b);
  }
  __syncthreads();
  for (int k = t; k < npop; k += 256) {
    out[(m * npop + k) * 2] = ca[k];
    out[(m * npop + k) * 2 + 1] = cb[k];
  }
}

// observed genotypes of each individual over the SNPs not ignored (numvalids,
// admutils.c:312-340), one thread per individual
__global__ __launch_bounds__(256) void k_lsq_count(const uint8_t *__restrict__ geno, int64_t rlen,
                                                   int64_t M, int32_t nind,
                                                   const uint8_t *__restrict__ ignore,
                                                   int32_t *__restrict__ nobs) {
  const int32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nind) return;
  int32_t c = 0;
  for (int64_t m = 0; m < M; ++m) {
    if (ignore[m]) continue;
    const uint32_t g = (geno[m * rlen + (q >> 2)] >> ((3 - (q & 3)) << 1)) & 3u;
    c += g != 3u;
  }
  nobs[q] = c;
}
}  // namespace

int eig_lsq_filltab(eig_ctx *ctx, const int32_t *use, const int32_t *itype, int32_t npop,
                    int32_t K, double **fill, const int32_t **d_itype) {
  if (npop <= 0 || npop > kMaxPop)
    return eig_set_err(ctx, EIG_E_ARG, "lsqproject fill: 1 <= npop <= %d", kMaxPop);
  const int64_t M = ctx->nsnp;
  const int32_t nind = ctx->nind;
  for (int32_t q = 0; q < nind; ++q)
    if (itype[q] >= npop) return eig_set_err(ctx, EIG_E_ARG, "lsqproject fill: itype[%d] >= npop", q);
  hipStream_t st = ctx->stream;
  EIG_TRY(eig_ensure(ctx, ctx->fill, 8 * (size_t)M * npop + 12 * (size_t)nind + 256));
  double *d_fill = (double *)ctx->fill.p;
  int32_t *d_it = (int32_t *)(d_fill + (size_t)M * npop), *d_eff = d_it + nind, *d_nobs = d_eff + nind;
  EIG_HIP(ctx, hipMemcpyAsync(d_it, itype, 4 * (size_t)nind, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_lsq_count, dim3((unsigned)((nind + 255) / 256)), dim3(256), 0, st,
                     (const uint8_t *)ctx->geno.p, ctx->rlen, M, nind,
                     (const uint8_t *)ctx->ignore_dev.p, d_nobs);
  EIG_HIP(ctx, hipGetLastError());
  std::vector<int32_t> nobs(nind), eff(nind);
  EIG_HIP(ctx, hipMemcpyAsync(nobs.data(), d_nobs, 4 * (size_t)nind, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  // the table's contributors: individuals in use, of a type >= 0, and not dropped by
  // lsqproj's first loop (<= K observed genotypes, smartpca.c:4637-4651)
  for (int32_t q = 0; q < nind; ++q)
    eff[q] = (use[q] && itype[q] >= 0 && nobs[q] > K) ? itype[q] : -1;
  EIG_HIP(ctx, hipMemcpyAsync(d_eff, eff.data(), 4 * (size_t)nind, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_lsq_filltab, dim3((unsigned)M), dim3(256), 0, st,
                     (const uint8_t *)ctx->geno.p, ctx->rlen, nind, (const int32_t *)d_eff, npop,
                     d_fill);
  EIG_HIP(ctx, hipGetLastError());
  *fill = d_fill;
  *d_itype = d_it;
  return EIG_OK;
}

extern "C" int eig_xtx_normal(eig_ctx *ctx, const int32_t *xtypes, int32_t npop,
                              const double *weight, int32_t ldregress, const int32_t *chrom,
                              const double *ppos, const double *gpos, int32_t ldposlimit,
                              double ldlimit, double ldr2lo, double ldr2hi, int32_t minallelecnt,
                              int32_t maxmissing, int32_t *c0, int32_t *c1, int32_t *nmiss,
                              int32_t *status, double *trace_scale, double *ynumsnps,
                              double *xtx_out) {
  if (!ctx) return EIG_E_ARG;
  if (!ctx->have_stats)
    return eig_set_err(ctx, EIG_E_STATE, "eig_xtx_normal before eig_snpstats");
  if (xtypes && (npop <= 0 || npop > kMaxPop))
    return eig_set_err(ctx, EIG_E_ARG, "eig_xtx_normal: 1 <= npop <= %d", kMaxPop);
  if (ldregress < 0 || (ldregress > 0 && (!chrom || !ppos || !gpos)))
    return eig_set_err(ctx, EIG_E_ARG, "eig_xtx_normal: ldregress needs chrom, ppos and gpos");
  const int32_t n = ctx->nrows;
  if (n >= 46341)
    return eig_set_err(ctx, EIG_E_ARG, "size of GRM matrix too large: nrows: %d", n);
  if (n < 2) return eig_set_err(ctx, EIG_E_ARG, "eig_xtx_normal: nrows < 2");
  const int64_t M = ctx->nsnp;
  hipStream_t st = ctx->stream;
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  ctx->have_xtx = ctx->have_load = false;
  if (xtypes)
    for (int32_t i = 0; i < n; ++i)
      if (xtypes[i] < 0 || xtypes[i] >= npop)
        return eig_set_err(ctx, EIG_E_ARG, "eig_xtx_normal: xtypes[%d] out of range", i);

  // ---- per-SNP statistics with the fill, keep rule, plain levels
  const int32_t W = ldregress;
  const size_t o_xt = 0, o_w = (4 * (size_t)n + 255) & ~(size_t)255;
  EIG_TRY(eig_ensure(ctx, ctx->scratch, o_w + 8 * (size_t)M + 256));
  int32_t *d_xt = xtypes ? (int32_t *)((char *)ctx->scratch.p + o_xt) : nullptr;
  double *d_w = weight ? (double *)((char *)ctx->scratch.p + o_w) : nullptr;
  if (d_xt) EIG_HIP(ctx, hipMemcpyAsync(d_xt, xtypes, 4 * (size_t)n, hipMemcpyHostToDevice, st));
  if (d_w) EIG_HIP(ctx, hipMemcpyAsync(d_w, weight, 8 * (size_t)M, hipMemcpyHostToDevice, st));
  eig_time_begin(ctx);
  hipLaunchKernelGGL(k_normstats, dim3((unsigned)M), dim3(256), 0, st, (const uint32_t *)ctx->sel.p,
                     M, n, ctx->have_live ? (const uint8_t *)ctx->live.p : nullptr,
                     (const int32_t *)d_xt, npop, ctx->altnorm, ctx->fancynorm, minallelecnt,
                     maxmissing, (int32_t *)ctx->c0.p, (int32_t *)ctx->c1.p,
                     (int32_t *)ctx->nmiss.p, (int32_t *)ctx->status.p, (double *)ctx->mean.p,
                     (double *)ctx->fancy.p, (double *)ctx->lev.p);
  EIG_HIP(ctx, hipGetLastError());
  std::vector<int32_t> hst(M);
  EIG_HIP(ctx, hipMemcpyAsync(hst.data(), ctx->status.p, 4 * (size_t)M, hipMemcpyDeviceToHost, st));
  struct {
    int32_t *h;
    DevBuf *d;
  } outs[] = {{c0, &ctx->c0}, {c1, &ctx->c1}, {nmiss, &ctx->nmiss}};
  for (auto &o : outs)
    if (o.h) EIG_HIP(ctx, hipMemcpyAsync(o.h, o.d->p, 4 * (size_t)M, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  if (status) std::memcpy(status, hst.data(), 4 * (size_t)M);
  std::vector<int32_t> kl;  // kept SNPs in order
  for (int64_t m = 0; m < M; ++m)
    if (hst[m] == EIG_SNP_KEPT) kl.push_back((int32_t)m);
  const int64_t nk = (int64_t)kl.size();

  // ---- ldregress windows: the reference's buffer before SNP k holds kept positions
  // ws[k] .. k-1 (bumpldvv: cleared when the chromosome changes, oldest dropped while
  // too far from the newest SNP or when full)
  std::vector<int64_t> ws(nk + 1, 0);
  if (W > 0) {
    int64_t numld = 0;
    int32_t lastc = -1;
    bool first = true;
    for (int64_t k = 0; k < nk; ++k) {
      const int32_t m = kl[k];
      if (first || lastc != chrom[m]) numld = 0;
      ws[k] = k - numld;
      while (numld > 0) {
        const int32_t f = kl[k - numld];
        const long pd = lrint(ppos[m] - ppos[f]);
        const double gd = gpos[m] - gpos[f];
        if (pd <= ldposlimit && gd <= ldlimit) break;
        --numld;
      }
      if (numld == W) --numld;
      ++numld;
      lastc = chrom[m];
      first = false;
    }
  } else {
    for (int64_t k = 0; k < nk; ++k) ws[k] = k;
  }

  // ---- chunks of kept SNPs
  const int64_t B = W > 0 ? 256 : 1024;
  if (W > 0 && B + W > kMaxLdNc)
    return eig_set_err(ctx, EIG_E_ARG, "eig_xtx_normal: ldregress <= %d", kMaxLdNc - 256);
  const int64_t ldz = n;
  const int64_t ncmax = B + W;
  size_t need = 0;
  auto add = [&](size_t bytes) {
    size_t o = need;
    need += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_idx = add(4 * (size_t)ncmax), o_Z = add(8 * (size_t)ldz * ncmax),
               o_R = add(W > 0 ? 8 * (size_t)ldz * B : 0),
               o_G = add(W > 0 ? 8 * (size_t)ncmax * ncmax : 0),
               o_Bc = add(W > 0 ? 8 * (size_t)ncmax * B : 0),
               o_work = add(W > 0 ? 8 * (size_t)B * ((size_t)W * W + 2 * W) : 0),
               o_ws = add(W > 0 ? 8 * (size_t)(nk + 1) : 0), o_drop = add(W > 0 ? (size_t)nk + 1 : 0),
               o_ns = add(4 * (size_t)B), o_src = add(W > 0 ? 4 * (size_t)B * W : 0),
               o_nrm = add(8 * (size_t)B);
  EIG_TRY(eig_ensure(ctx, ctx->scratch2, need));
  EIG_TRY(eig_ensure(ctx, ctx->xtx, 8 * (size_t)n * n));
  char *base = (char *)ctx->scratch2.p;
  int32_t *d_idx = (int32_t *)(base + o_idx);
  double *d_Z = (double *)(base + o_Z), *d_R = (double *)(base + o_R), *d_G = (double *)(base + o_G);
  double *d_Bc = (double *)(base + o_Bc), *d_work = (double *)(base + o_work);
  int64_t *d_ws = (int64_t *)(base + o_ws);
  uint8_t *d_drop = (uint8_t *)(base + o_drop);
  int32_t *d_ns = (int32_t *)(base + o_ns), *d_src = (int32_t *)(base + o_src);
  double *d_nrm = (double *)(base + o_nrm);
  double *X = (double *)ctx->xtx.p;
  if (W > 0) {
    EIG_HIP(ctx, hipMemcpyAsync(d_ws, ws.data(), 8 * (size_t)nk, hipMemcpyHostToDevice, st));
    EIG_HIP(ctx, hipMemsetAsync(d_drop, 0, (size_t)nk + 1, st));
  }
  double ysnps = 0.0;
  bool first = true;
  std::vector<double> rn(B), zn(B);
  std::vector<int32_t> hns(B);
  const double scale = 1.0 / (2.0 * (double)(n - 1));  // ldregx: cor = dot / (2 (n-1))
  for (int64_t k0 = 0; k0 < nk; k0 += B) {
    const int64_t k1 = std::min(nk, k0 + B), nb = k1 - k0;
    int64_t h0 = k0;
    for (int64_t k = k0; k < k1; ++k) h0 = std::min(h0, ws[k]);
    const int64_t nc = k1 - h0;
    EIG_HIP(ctx, hipMemcpyAsync(d_idx, kl.data() + h0, 4 * (size_t)nc, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_normcols, dim3((unsigned)nc), dim3(256), 0, st,
                       (const uint32_t *)ctx->sel.p, M, n, (const int32_t *)d_xt, npop,
                       (const double *)ctx->mean.p, (const double *)ctx->fancy.p,
                       (const double *)d_w, (const int32_t *)d_idx, ldz, d_Z);
    EIG_HIP(ctx, hipGetLastError());
    const double *R = d_Z + (k0 - h0) * ldz;  // no regression: the columns themselves
    if (W > 0) {
      EIG_TRY(eig_gemm_f64(ctx, st, true, false, nc, nc, n, 1.0, d_Z, ldz, 0, d_Z, ldz, 0, 0.0, d_G,
                           nc, 0, 1));
      hipLaunchKernelGGL(k_ld_scan, dim3(1), dim3(256), 0, st, (const double *)d_G, nc, h0, k0, k1,
                         (const int64_t *)d_ws, scale, ldr2lo, ldr2hi, d_drop, d_ns, d_src, W);
      hipLaunchKernelGGL(k_ld_solve, dim3((unsigned)nb), dim3(256), 0, st, (const double *)d_G, nc,
                         h0, k0, (const int32_t *)d_ns, (const int32_t *)d_src, W, d_work, d_Bc);
      EIG_HIP(ctx, hipGetLastError());
      EIG_TRY(eig_gemm_f64(ctx, st, false, false, n, nb, nc, 1.0, d_Z, ldz, 0, d_Bc, nc, 0, 0.0,
                           d_R, ldz, 0, 1));
      hipLaunchKernelGGL(k_colnorm2, dim3((unsigned)nb), dim3(256), 0, st, (const double *)d_R, ldz,
                         n, d_nrm);
      EIG_HIP(ctx, hipGetLastError());
      R = d_R;
      // ynumsnps += |r|^2 / |z|^2 for each SNP not dropped (smartpca.c:1157-1162)
      EIG_HIP(ctx, hipMemcpyAsync(rn.data(), d_nrm, 8 * (size_t)nb, hipMemcpyDeviceToHost, st));
      EIG_HIP(ctx, hipMemcpyAsync(hns.data(), d_ns, 4 * (size_t)nb, hipMemcpyDeviceToHost, st));
      EIG_HIP(ctx, hipMemcpy2DAsync(zn.data(), 8, d_G + (k0 - h0) * (nc + 1), 8 * (nc + 1), 8, nb,
                                    hipMemcpyDeviceToHost, st));
    }
    EIG_TRY(eig_gemm_f64(ctx, st, false, true, n, n, nb, 1.0, R, ldz, 0, R, ldz, 0,
                         first ? 0.0 : 1.0, X, n, 0, 1, true));
    first = false;
    if (W > 0) {
      EIG_HIP(ctx, hipStreamSynchronize(st));
      for (int64_t c = 0; c < nb; ++c)
        if (hns[c] >= 0) ysnps += rn[c] / zn[c];
    } else {
      ysnps += (double)nb;
    }
  }
  if (first) EIG_HIP(ctx, hipMemsetAsync(X, 0, 8 * (size_t)n * n, st));
  eig_time_end(ctx, "xtx_normal");
  double *d_tr = d_nrm;
  hipLaunchKernelGGL(k_dense_trace, dim3(1), dim3(64), 0, st, (const double *)X, n, d_tr);
  double tr = 0.0;
  EIG_HIP(ctx, hipMemcpyAsync(&tr, d_tr, 8, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  const double y = tr / (double)(n - 1);
  if (trace_scale) *trace_scale = y;
  if (ynumsnps) *ynumsnps = ysnps;
  if (std::isnan(y)) return eig_set_err(ctx, EIG_E_NUMERIC, "bad XTX matrix");
  if (y <= 0.0) return eig_set_err(ctx, EIG_E_NUMERIC, "XTX has zero trace (perhaps no data)");
  const int64_t tot = (int64_t)n * n;
  hipLaunchKernelGGL(k_dense_scale_mirror, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, X,
                     n, (const double *)d_tr);
  EIG_HIP(ctx, hipGetLastError());
  if (xtx_out) {
    EIG_HIP(ctx, hipMemcpyAsync(xtx_out, X, 8 * (size_t)tot, hipMemcpyDeviceToHost, st));
  }
  EIG_HIP(ctx, hipStreamSynchronize(st));
  ctx->have_xtx = true;
  return EIG_OK;
}

extern "C" int eig_popcounts(eig_ctx *ctx, const int32_t *xtypes, int32_t npop,
                             const uint8_t *haploid, const uint8_t *xsnp, int32_t *counts) {
  if (!ctx || !xtypes || !counts) return eig_set_err(ctx, EIG_E_ARG, "eig_popcounts: args");
  if (!ctx->have_sel) return eig_set_err(ctx, EIG_E_STATE, "eig_popcounts before eig_select");
  if (npop <= 0 || npop > kMaxPop)
    return eig_set_err(ctx, EIG_E_ARG, "eig_popcounts: 1 <= npop <= %d", kMaxPop);
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t M = ctx->nsnp;
  const int32_t n = ctx->nrows;
  hipStream_t st = ctx->stream;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_xt = 0, o_h = up(4 * (size_t)n), o_x = o_h + up((size_t)n),
               o_out = o_x + up((size_t)M);
  const size_t out_b = 8 * (size_t)M * npop;
  EIG_TRY(eig_ensure(ctx, ctx->scratch, o_out + out_b));
  char *base = (char *)ctx->scratch.p;
  EIG_HIP(ctx, hipMemcpyAsync(base + o_xt, xtypes, 4 * (size_t)n, hipMemcpyHostToDevice, st));
  if (haploid) EIG_HIP(ctx, hipMemcpyAsync(base + o_h, haploid, (size_t)n, hipMemcpyHostToDevice, st));
  if (xsnp) EIG_HIP(ctx, hipMemcpyAsync(base + o_x, xsnp, (size_t)M, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_popcounts, dim3((unsigned)M), dim3(256), 0, st, (const uint32_t *)ctx->sel.p,
                     M, n, (const int32_t *)(base + o_xt), npop,
                     haploid ? (const uint8_t *)(base + o_h) : nullptr,
                     xsnp ? (const uint8_t *)(base + o_x) : nullptr, (int32_t *)(base + o_out));
  EIG_HIP(ctx, hipGetLastError());
  EIG_HIP(ctx, hipMemcpyAsync(counts, base + o_out, out_b, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  return EIG_OK;
}
