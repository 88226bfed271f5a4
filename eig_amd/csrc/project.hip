// SNP loadings, projection scales and lsqproject on the device.
//
//  loadings : ffvecs[j][m] = sum_i 10 v_j[i] z_m[i].  z takes three values per
//             SNP, so each (sample tile, SNP) first bins the eigenvector entries
//             by genotype code: S_g[j] = sum_{i: g_i = g} v_j[i]; then
//             ffvecs[j][m] = 10 * sum_g level_g S_g[j].
//  fxscal   : fxvecs[j][i] = sum_m x_i[m] ffvecs[j][m] with the reference's row
//             fix-up x = g*fancy - mean*fancy (missing or ignored -> 0).
//  lsqproject: per individual, over observed SNPs, the K x K Gram of
//             e_m = fxscal .* ffvecs[:,m] and the right-hand side sum_m e_m x_m,
//             reduced over fixed SNP chunks in chunk order, then the Cholesky
//             solve of nicksrc/linsubs.c (choldc/cholsl) per individual.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "eig_internal.h"

constexpr int kLsqChunk = 4096;  // SNPs per lsqproject partial

// run the statement with the compile-time constant KC equal to the runtime K (1..16)
#define EIG_KCASE(n, ...) \
  case n: {               \
    constexpr int KC = n; \
    __VA_ARGS__;          \
  } break;
#define EIG_KSWITCH(k, ...)                                                                  \
  switch (k) {                                                                               \
    EIG_KCASE(1, __VA_ARGS__) EIG_KCASE(2, __VA_ARGS__) EIG_KCASE(3, __VA_ARGS__)             \
    EIG_KCASE(4, __VA_ARGS__) EIG_KCASE(5, __VA_ARGS__) EIG_KCASE(6, __VA_ARGS__)             \
    EIG_KCASE(7, __VA_ARGS__) EIG_KCASE(8, __VA_ARGS__) EIG_KCASE(9, __VA_ARGS__)             \
    EIG_KCASE(10, __VA_ARGS__) EIG_KCASE(11, __VA_ARGS__) EIG_KCASE(12, __VA_ARGS__)          \
    EIG_KCASE(13, __VA_ARGS__) EIG_KCASE(14, __VA_ARGS__) EIG_KCASE(15, __VA_ARGS__)          \
    EIG_KCASE(16, __VA_ARGS__)                                                                \
    default:                                                                                 \
      return eig_set_err(ctx, EIG_E_ARG, "K out of range");                                  \
  }

// ---------------------------------------------------------------- loadings
// block = 256 threads handles 64 SNPs; thread (m = t>>2, part = t&3) covers
// samples 32*part .. 32*part+31 of each sample tile.
template <int K>
__global__ __launch_bounds__(256) void k_loadings(const uint32_t *__restrict__ sel, int64_t M,
                                                  int32_t nst, int32_t nrows,
                                                  const double *__restrict__ ev,
                                                  const double *__restrict__ lev,
                                                  const uint8_t *__restrict__ ignore,
                                                  double *__restrict__ ff) {
  __shared__ double fv[kMaxK][kTileS];
  __shared__ double red[4][64][kMaxK];
  const int t = threadIdx.x, ml = t >> 2, part = t & 3;
  const int64_t m = (int64_t)blockIdx.x * 64 + ml;
  double S[3][K];
#pragma unroll
  for (int g = 0; g < 3; ++g)
#pragma unroll
    for (int j = 0; j < K; ++j) S[g][j] = 0.0;
  for (int32_t st = 0; st < nst; ++st) {
    __syncthreads();
    for (int idx = t; idx < K * kTileS; idx += 256) {
      int j = idx / kTileS, i = idx % kTileS;
      int32_t gi = st * kTileS + i;
      fv[j][i] = gi < nrows ? ev[(int64_t)j * nrows + gi] : 0.0;
    }
    __syncthreads();
    if (m < M) {
      const uint32_t *wp = sel + ((int64_t)st * M + m) * kTileWords + 2 * part;
      uint32_t w0 = wp[0], w1 = wp[1];
      for (int s = 0; s < 32; ++s) {
        uint32_t code = ((s < 16 ? w0 : w1) >> (2 * (s & 15))) & 3u;
        int i = 32 * part + s;
        if (code < 3) {
#pragma unroll
          for (int j = 0; j < K; ++j) {
            double v = fv[j][i];
            S[0][j] += code == 0 ? v : 0.0;
            S[1][j] += code == 1 ? v : 0.0;
            S[2][j] += code == 2 ? v : 0.0;
          }
        }
      }
    }
  }
  // combine the 4 parts of each SNP in fixed order
  double out[K];
  const double *L = lev + 4 * (m < M ? m : 0);
#pragma unroll
  for (int j = 0; j < K; ++j) out[j] = L[0] * S[0][j] + L[1] * S[1][j] + L[2] * S[2][j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < K; ++j) red[part][ml][j] = out[j];
  __syncthreads();
  if (part == 0 && m < M) {
    bool ign = ignore[m] != 0;
    for (int j = 0; j < K; ++j) {
      double s = ((red[0][ml][j] + red[1][ml][j]) + red[2][ml][j]) + red[3][ml][j];
      ff[(int64_t)j * M + m] = ign ? 0.0 : 10.0 * s;
    }
  }
}

// ---------------------------------------------------------------- fxvecs
// block (sample tile st, SNP chunk): thread i < 128 -> sample st*128+i.
constexpr int kFxChunk = 2048;
template <int K>
__global__ __launch_bounds__(128) void k_fxvecs(const uint32_t *__restrict__ sel, int64_t M,
                                                int32_t nst, int32_t nrows,
                                                const double *__restrict__ ff,
                                                const double *__restrict__ fancy,
                                                const double *__restrict__ xmean,
                                                const uint8_t *__restrict__ ignore,
                                                double *__restrict__ partial) {
  __shared__ double sf[64][kMaxK + 2];
  const int i = threadIdx.x;
  const int32_t st = blockIdx.x % nst;
  const int64_t chunk = blockIdx.x / nst;
  const int64_t m0 = chunk * kFxChunk, m1 = min(M, m0 + kFxChunk);
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  for (int64_t mb = m0; mb < m1; mb += 64) {
    __syncthreads();
    for (int idx = i; idx < 64 * (K + 2); idx += 128) {
      int r = idx / (K + 2), c = idx % (K + 2);
      int64_t m = mb + r;
      double v = 0.0;
      if (m < m1) {
        if (c < K) v = ff[(int64_t)c * M + m];
        else if (c == K) v = ignore[m] ? 0.0 : fancy[m];
        else v = ignore[m] ? 0.0 : xmean[m];
      }
      sf[r][c] = v;
    }
    __syncthreads();
    const int64_t lim = min((int64_t)64, m1 - mb);
    for (int r = 0; r < lim; ++r) {
      uint32_t wv = sel[((int64_t)st * M + mb + r) * kTileWords + (i >> 4)];
      uint32_t code = (wv >> (2 * (i & 15))) & 3u;
      double g = (code == 3u) ? -1.0 : (double)code;
      double y = g * sf[r][K];
      double x = (y < -0.001) ? 0.0 : y - sf[r][K + 1];
#pragma unroll
      for (int j = 0; j < K; ++j) acc[j] += x * sf[r][j];
    }
  }
  const int32_t gi = st * kTileS + i;
  if (gi < nrows)
#pragma unroll
    for (int j = 0; j < K; ++j) partial[(chunk * K + j) * (int64_t)nrows + gi] = acc[j];
}

__global__ void k_fxscal(const double *__restrict__ partial, int64_t nchunk, int32_t nrows,
                         int32_t K, double *__restrict__ fxvecs, double *__restrict__ fxscal) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double s2 = 0.0;
  for (int32_t i = threadIdx.x; i < nrows; i += 256) {
    double s = 0.0;
    for (int64_t c = 0; c < nchunk; ++c) s += partial[(c * K + j) * (int64_t)nrows + i];
    fxvecs[(int64_t)j * nrows + i] = s;
    s2 += s * s;
  }
  red[threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) fxscal[j] = 1.0 / sqrt(red[0]);
}

// ---------------------------------------------------------------- lsqproject
// Individuals read from GENO rows (all loaded samples).  Per (individual block of
// 64, SNP chunk): Gram upper triangle (K(K+1)/2), rhs (K), observed count.
template <int K>
__global__ __launch_bounds__(64) void k_lsq_gram(const uint8_t *__restrict__ geno, int64_t rlen,
                                                 int64_t M, int32_t nind,
                                                 const double *__restrict__ ff,
                                                 const double *__restrict__ fxscal,
                                                 const double *__restrict__ fancy,
                                                 const double *__restrict__ xmean,
                                                 const uint8_t *__restrict__ ignore,
                                                 const double *__restrict__ fill,
                                                 const int32_t *__restrict__ itype, int32_t npop,
                                                 double *__restrict__ partial, int32_t nblk,
                                                 int64_t chunk0) {
  __shared__ double se[64][kMaxK + 2];
  const int li = threadIdx.x;
  const int32_t q = (blockIdx.x % nblk) * 64 + li;
  const int64_t chunk = chunk0 + blockIdx.x / nblk;
  const int64_t m0 = chunk * kLsqChunk, m1 = min(M, m0 + kLsqChunk);
  constexpr int nP = K * (K + 1) / 2;
  double G[nP];
  double r[K];
#pragma unroll
  for (int p = 0; p < nP; ++p) G[p] = 0.0;
#pragma unroll
  for (int j = 0; j < K; ++j) r[j] = 0.0;
  double kk = 0.0, nobs = 0.0;
  const int32_t ty = (fill && q < nind) ? itype[q] : -1;
  for (int64_t mb = m0; mb < m1; mb += 64) {
    __syncthreads();
    {
      int64_t m = mb + li;
      bool ok = m < m1;
#pragma unroll
      for (int j = 0; j < K; ++j) se[li][j] = ok ? fxscal[j] * ff[(int64_t)j * M + m] : 0.0;
      se[li][K] = ok ? (ignore[m] ? -1.0 : fancy[m]) : -1.0;
      se[li][K + 1] = ok ? xmean[m] : 0.0;
    }
    __syncthreads();
    const int64_t lim = min((int64_t)64, m1 - mb);
    if (q < nind) {
      for (int rr = 0; rr < lim; ++rr) {
        double fz = se[rr][K];
        if (fz < 0.0) continue;  // ignored SNP: reads as missing
        uint32_t b = geno[(mb + rr) * rlen + (q >> 2)];
        uint32_t code = (b >> ((3 - (q & 3)) << 1)) & 3u;
        double gv = (double)code;
        if (code == 3u) {
          // usepopsformissing: the mean genotype of the individual's population type
          // (lsqproj, smartpca.c:4680-4690); without one the SNP is unobserved
          if (ty < 0) continue;
          gv = fill[(mb + rr) * npop + ty];
          if (isnan(gv)) continue;
        } else {
          nobs += 1.0;
        }
        double x = gv * fz;
        x = (x < -0.001) ? 0.0 : x - se[rr][K + 1];
        kk += 1.0;
        double e[K];
#pragma unroll
        for (int a = 0; a < K; ++a) e[a] = se[rr][a];
#pragma unroll
        for (int a = 0; a < K; ++a) r[a] += e[a] * x;
#pragma unroll
        for (int a = 0; a < K; ++a)
#pragma unroll
          for (int c = a; c < K; ++c) G[a * K - a * (a - 1) / 2 + (c - a)] += e[a] * e[c];
      }
    }
  }
  if (q < nind) {
    double *out = partial + ((chunk - chunk0) * (int64_t)nind + q) * (nP + K + 2);
#pragma unroll
    for (int p = 0; p < nP; ++p) out[p] = G[p];
#pragma unroll
    for (int j = 0; j < K; ++j) out[nP + j] = r[j];
    out[nP + K] = kk;
    out[nP + K + 1] = nobs;
  }
}

// reduce chunks in order, Cholesky solve (choldc/cholsl order of operations)
// itype (may be NULL): with the fill, an individual of type >= 0 is also ignored when
// its observed genotypes alone number <= K (lsqproj's first loop, smartpca.c:4637-4651).
__global__ void k_lsq_solve(const double *__restrict__ partial, int64_t nchunk, int32_t nind,
                            int32_t K, const int32_t *__restrict__ use,
                            const int32_t *__restrict__ itype, double *__restrict__ ans,
                            double *__restrict__ rhs_out, int32_t *__restrict__ st) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nind) return;
  const int nP = K * (K + 1) / 2, W = nP + K + 2;
  double acc[kMaxK * (kMaxK + 1) / 2 + kMaxK + 2];
  for (int p = 0; p < W; ++p) acc[p] = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) {
    const double *pp = partial + (c * (int64_t)nind + q) * W;
    for (int p = 0; p < W; ++p) acc[p] += pp[p];
  }
  for (int j = 0; j < K; ++j) rhs_out[(int64_t)j * nind + q] = acc[nP + j];
  if (!use[q]) {
    st[q] = 3;
    return;
  }
  if (acc[nP + K] <= (double)K || (itype && itype[q] >= 0 && acc[nP + K + 1] <= (double)K)) {
    st[q] = 1;
    return;
  }
  double a[kMaxK][kMaxK], p[kMaxK], x[kMaxK];
  {
    int idx = 0;
    for (int i = 0; i < K; ++i)
      for (int j = i; j < K; ++j) {
        a[i][j] = acc[idx];
        a[j][i] = acc[idx];
        ++idx;
      }
  }
  bool bad = false;
  for (int i = 0; i < K * K; ++i)
    if (isnan(a[i / K][i % K])) bad = true;
  for (int i = 0; i < K; ++i)
    if (isnan(acc[nP + i])) bad = true;
  for (int i = 0; i < K && !bad; ++i) {
    for (int j = i; j < K; ++j) {
      double s = a[i][j];
      for (int k = i - 1; k >= 0; --k) s -= a[i][k] * a[j][k];
      if (i == j) {
        if (s <= 0.0) {
          bad = true;
          break;
        }
        p[i] = sqrt(s);
      } else {
        a[j][i] = s / p[i];
      }
    }
  }
  if (bad) {
    st[q] = 2;
    return;
  }
  for (int i = 0; i < K; ++i) {
    double s = acc[nP + i];
    for (int k = i - 1; k >= 0; --k) s -= a[i][k] * x[k];
    x[i] = s / p[i];
  }
  for (int i = K - 1; i >= 0; --i) {
    double s = x[i];
    for (int k = i + 1; k < K; ++k) s -= a[k][i] * x[k];
    x[i] = s / p[i];
  }
  for (int j = 0; j < K; ++j) ans[(int64_t)j * nind + q] = x[j];
  st[q] = 0;
}

extern "C" int eig_loadings(eig_ctx *ctx, const double *evecs, int32_t K, const uint8_t *ignore,
                            double *ffvecs, double *fxscal) {
  if (!ctx || !evecs || K <= 0 || K > kMaxK)
    return eig_set_err(ctx, EIG_E_ARG, "eig_loadings: K must be in [1, %d]", kMaxK);
  if (!ctx->have_stats) return eig_set_err(ctx, EIG_E_STATE, "eig_loadings before eig_snpstats");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t M = ctx->nsnp;
  const int32_t n = ctx->nrows;
  hipStream_t st = ctx->stream;
  // ignore mask (host), xmean = mean*fancy (device, computed here)
  std::vector<uint8_t> ign(M);
  if (ignore) {
    std::memcpy(ign.data(), ignore, M);
  } else {
    std::vector<int32_t> s(M);
    EIG_HIP(ctx, hipMemcpy(s.data(), ctx->status.p, 4 * M, hipMemcpyDeviceToHost));
    for (int64_t m = 0; m < M; ++m) ign[m] = s[m] != EIG_SNP_KEPT;
  }
  std::vector<double> mean(M), fancy(M), xm(M);
  EIG_HIP(ctx, hipMemcpy(mean.data(), ctx->mean.p, 8 * M, hipMemcpyDeviceToHost));
  EIG_HIP(ctx, hipMemcpy(fancy.data(), ctx->fancy.p, 8 * M, hipMemcpyDeviceToHost));
  for (int64_t m = 0; m < M; ++m) xm[m] = mean[m] * fancy[m];
  EIG_TRY(eig_ensure(ctx, ctx->ignore_dev, M));
  EIG_TRY(eig_ensure(ctx, ctx->ffvecs, 8 * (size_t)K * M));
  EIG_TRY(eig_ensure(ctx, ctx->fxscal, 8 * (size_t)(kMaxK + M)));  // fxscal | xmean
  EIG_TRY(eig_ensure(ctx, ctx->projw, 8 * (size_t)K * n));
  double *d_fx = (double *)ctx->fxscal.p, *d_xm = d_fx + kMaxK;
  EIG_HIP(ctx, hipMemcpyAsync(ctx->ignore_dev.p, ign.data(), M, hipMemcpyHostToDevice, st));
  EIG_HIP(ctx, hipMemcpyAsync(d_xm, xm.data(), 8 * M, hipMemcpyHostToDevice, st));
  EIG_HIP(ctx, hipMemcpyAsync(ctx->projw.p, evecs, 8 * (size_t)K * n, hipMemcpyHostToDevice, st));
  eig_time_begin(ctx);
  EIG_KSWITCH(K, hipLaunchKernelGGL((k_loadings<KC>), dim3((unsigned)((M + 63) / 64)), dim3(256),
                                    0, st, (const uint32_t *)ctx->sel.p, M, ctx->nst, n,
                                    (const double *)ctx->projw.p, (const double *)ctx->lev.p,
                                    (const uint8_t *)ctx->ignore_dev.p, (double *)ctx->ffvecs.p));
  EIG_HIP(ctx, hipGetLastError());
  const int64_t nchunk = (M + kFxChunk - 1) / kFxChunk;
  EIG_TRY(eig_ensure(ctx, ctx->scratch2, 8 * (size_t)(nchunk * K * n + (size_t)K * n)));
  double *d_part = (double *)ctx->scratch2.p, *d_fxv = d_part + nchunk * K * n;
  EIG_KSWITCH(K, hipLaunchKernelGGL((k_fxvecs<KC>), dim3((unsigned)(nchunk * ctx->nst)), dim3(128),
                                    0, st, (const uint32_t *)ctx->sel.p, M, ctx->nst, n,
                                    (const double *)ctx->ffvecs.p, (const double *)ctx->fancy.p,
                                    (const double *)d_xm, (const uint8_t *)ctx->ignore_dev.p,
                                    d_part));
  hipLaunchKernelGGL(k_fxscal, dim3((unsigned)K), dim3(256), 0, st, (const double *)d_part,
                     nchunk, n, K, d_fxv, d_fx);
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "loadings");
  if (ffvecs)
    EIG_HIP(ctx, hipMemcpyAsync(ffvecs, ctx->ffvecs.p, 8 * (size_t)K * M, hipMemcpyDeviceToHost, st));
  if (fxscal) EIG_HIP(ctx, hipMemcpyAsync(fxscal, d_fx, 8 * (size_t)K, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  ctx->last_K = K;
  ctx->have_load = true;
  return EIG_OK;
}

static int lsqproject(eig_ctx *ctx, const int32_t *use, const int32_t *itype, int32_t npop,
                      double *acoeffs, double *bcoeffs, int32_t *status) {
  if (!ctx || !use) return eig_set_err(ctx, EIG_E_ARG, "eig_lsqproject: args");
  if (!ctx->have_load) return eig_set_err(ctx, EIG_E_STATE, "eig_lsqproject before eig_loadings");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t M = ctx->nsnp;
  const int32_t nind = ctx->nind, K = ctx->last_K;
  hipStream_t st = ctx->stream;
  const int nP = K * (K + 1) / 2, W = nP + K + 2;
  double *d_fill = nullptr;
  const int32_t *d_itype = nullptr;
  if (itype) EIG_TRY(eig_lsq_filltab(ctx, use, itype, npop, K, &d_fill, &d_itype));
  const int64_t nchunk = (M + kLsqChunk - 1) / kLsqChunk;
  const int32_t nblk = (nind + 63) / 64;
  size_t part_b = 8 * (size_t)nchunk * nind * W;
  size_t out_b = 8 * (size_t)2 * K * nind + 4 * (size_t)nind * 2 + 256;
  EIG_TRY(eig_ensure(ctx, ctx->scratch2, part_b));
  EIG_TRY(eig_ensure(ctx, ctx->proj_out, out_b));
  double *d_ans = (double *)ctx->proj_out.p, *d_rhs = d_ans + (size_t)K * nind;
  int32_t *d_st = (int32_t *)(d_rhs + (size_t)K * nind), *d_use = d_st + nind;
  EIG_HIP(ctx, hipMemcpyAsync(d_use, use, 4 * (size_t)nind, hipMemcpyHostToDevice, st));
  EIG_HIP(ctx, hipMemsetAsync(d_ans, 0, 8 * (size_t)K * nind, st));
  const double *d_fx = (const double *)ctx->fxscal.p, *d_xm = d_fx + kMaxK;
  eig_time_begin(ctx);
  EIG_KSWITCH(K, hipLaunchKernelGGL((k_lsq_gram<KC>), dim3((unsigned)(nchunk * nblk)), dim3(64),
                                    0, st, (const uint8_t *)ctx->geno.p, ctx->rlen, M, nind,
                                    (const double *)ctx->ffvecs.p, d_fx,
                                    (const double *)ctx->fancy.p, d_xm,
                                    (const uint8_t *)ctx->ignore_dev.p, (const double *)d_fill,
                                    d_itype, npop, (double *)ctx->scratch2.p, nblk, (int64_t)0));
  hipLaunchKernelGGL(k_lsq_solve, dim3((unsigned)((nind + 127) / 128)), dim3(128), 0, st,
                     (const double *)ctx->scratch2.p, nchunk, nind, K, (const int32_t *)d_use,
                     d_itype, d_ans, d_rhs, d_st);
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "lsqproject");
  std::vector<double> ans((size_t)K * nind), rhs((size_t)K * nind);
  std::vector<int32_t> sv(nind);
  EIG_HIP(ctx, hipMemcpyAsync(ans.data(), d_ans, 8 * (size_t)K * nind, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipMemcpyAsync(rhs.data(), d_rhs, 8 * (size_t)K * nind, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipMemcpyAsync(sv.data(), d_st, 4 * (size_t)nind, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  // regans persists across individuals in the reference (smartpca.c:4656,
  // regsubs.c:51-53): a failed solve keeps the previous individual's answer.
  double last[kMaxK];
  for (int j = 0; j < K; ++j) last[j] = 0.0;
  for (int32_t q = 0; q < nind; ++q) {
    if (status) status[q] = sv[q];
    if (sv[q] == 0)
      for (int j = 0; j < K; ++j) last[j] = ans[(size_t)j * nind + q];
    if (sv[q] == 0 || sv[q] == 2)
      for (int j = 0; j < K; ++j)
        if (acoeffs) acoeffs[(size_t)j * nind + q] = last[j];
    if (bcoeffs && (sv[q] == 0 || sv[q] == 2))
      for (int j = 0; j < K; ++j) bcoeffs[(size_t)j * nind + q] = rhs[(size_t)j * nind + q];
  }
  return EIG_OK;
}

extern "C" int eig_lsqproject(eig_ctx *ctx, const int32_t *use, double *acoeffs,
                              double *bcoeffs, int32_t *status) {
  return lsqproject(ctx, use, nullptr, 0, acoeffs, bcoeffs, status);
}

extern "C" int eig_lsqproject_fill(eig_ctx *ctx, const int32_t *use, const int32_t *itype,
                                   int32_t npop, double *acoeffs, double *bcoeffs,
                                   int32_t *status) {
  if (!itype) return eig_set_err(ctx, EIG_E_ARG, "eig_lsqproject_fill: itype is required");
  return lsqproject(ctx, use, itype, npop, acoeffs, bcoeffs, status);
}

// ---------------------------------------------------------------- allocation-free lsqproject
namespace {
__global__ void k_zero(double *__restrict__ p, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) p[e] = 0.0;
}

// acc[q][p] += part[c][q][p] for the group's chunks c in order: the same sequence of
// additions as k_lsq_solve's loop over all chunks
__global__ void k_lsq_accum(double *__restrict__ acc, const double *__restrict__ part,
                            int64_t nchunk, int64_t per) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= per) return;
  double v = acc[e];
  for (int64_t c = 0; c < nchunk; ++c) v += part[c * per + e];
  acc[e] = v;
}

// the stale-answer rule of the reference, in individual order, one thread per axis j:
// outputs of individuals with status 1 or 3 keep the caller's values
__global__ void k_lsq_stale(const double *__restrict__ ans, const double *__restrict__ rhs,
                            const int32_t *__restrict__ st, int32_t nind, int32_t K,
                            double *__restrict__ A, double *__restrict__ B) {
  const int j = threadIdx.x;
  if (j >= K) return;
  double last = 0.0;
  for (int32_t q = 0; q < nind; ++q) {
    const int32_t s = st[q];
    if (s == 0) last = ans[(int64_t)j * nind + q];
    if (s == 0 || s == 2) {
      if (A) A[(int64_t)j * nind + q] = last;
      if (B) B[(int64_t)j * nind + q] = rhs[(int64_t)j * nind + q];
    }
  }
}
}  // namespace

extern "C" int eig_lsqproject_noalloc(eig_ctx *ctx, const int32_t *use, double *acoeffs,
                                      double *bcoeffs, int32_t *status) {
  if (!ctx || !use) return eig_set_err(ctx, EIG_E_ARG, "eig_lsqproject_noalloc: args");
  if (!ctx->have_load)
    return eig_set_err(ctx, EIG_E_STATE, "eig_lsqproject_noalloc before eig_loadings");
  if (!ctx->noalloc) return eig_set_err(ctx, EIG_E_STATE, "eig_lsqproject_noalloc: no workspace");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t M = ctx->nsnp;
  const int32_t nind = ctx->nind, K = ctx->last_K;
  hipStream_t st = ctx->stream;
  const int nP = K * (K + 1) / 2, W = nP + K + 2;
  const int64_t nchunk = (M + kLsqChunk - 1) / kLsqChunk;
  const int32_t nblk = (nind + 63) / 64;
  // workspace: acc | ans | rhs | A | B | status | use | partials of G chunks
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t per = (size_t)nind * W;
  const size_t o_acc = 0, o_ans = up(8 * per), o_rhs = o_ans + up(8 * (size_t)K * nind),
               o_A = o_rhs + up(8 * (size_t)K * nind), o_B = o_A + up(8 * (size_t)K * nind),
               o_st = o_B + up(8 * (size_t)K * nind), o_use = o_st + up(4 * (size_t)nind),
               o_part = o_use + up(4 * (size_t)nind);
  if (o_part + 8 * per > ctx->noalloc_bytes)
    return eig_set_err(ctx, EIG_E_ARG,
                       "eig_lsqproject_noalloc: %d individuals x K=%d need more than the %zu-byte "
                       "static workspace", nind, K, ctx->noalloc_bytes);
  const int64_t G = std::min<int64_t>(nchunk, (int64_t)((ctx->noalloc_bytes - o_part) / (8 * per)));
  char *base = (char *)ctx->noalloc;
  double *d_acc = (double *)(base + o_acc), *d_ans = (double *)(base + o_ans),
         *d_rhs = (double *)(base + o_rhs), *d_A = (double *)(base + o_A),
         *d_B = (double *)(base + o_B), *d_part = (double *)(base + o_part);
  int32_t *d_st = (int32_t *)(base + o_st), *d_use = (int32_t *)(base + o_use);
  const size_t kn = (size_t)K * nind;
  EIG_HIP(ctx, hipMemcpyAsync(d_use, use, 4 * (size_t)nind, hipMemcpyHostToDevice, st));
  // the caller's current outputs go in, so entries the rule leaves alone come back as they were
  if (acoeffs) EIG_HIP(ctx, hipMemcpyAsync(d_A, acoeffs, 8 * kn, hipMemcpyHostToDevice, st));
  if (bcoeffs) EIG_HIP(ctx, hipMemcpyAsync(d_B, bcoeffs, 8 * kn, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_zero, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, st, d_acc,
                     (int64_t)per);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)((kn + 255) / 256)), dim3(256), 0, st, d_ans,
                     (int64_t)kn);
  const double *d_fx = (const double *)ctx->fxscal.p, *d_xm = d_fx + kMaxK;
  for (int64_t c0 = 0; c0 < nchunk; c0 += G) {
    const int64_t gc = std::min<int64_t>(G, nchunk - c0);
    EIG_KSWITCH(K, hipLaunchKernelGGL((k_lsq_gram<KC>), dim3((unsigned)(gc * nblk)), dim3(64), 0,
                                      st, (const uint8_t *)ctx->geno.p, ctx->rlen, M, nind,
                                      (const double *)ctx->ffvecs.p, d_fx,
                                      (const double *)ctx->fancy.p, d_xm,
                                      (const uint8_t *)ctx->ignore_dev.p, (const double *)nullptr,
                                      (const int32_t *)nullptr, 0, d_part, nblk, c0));
    hipLaunchKernelGGL(k_lsq_accum, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, st, d_acc,
                       (const double *)d_part, gc, (int64_t)per);
  }
  hipLaunchKernelGGL(k_lsq_solve, dim3((unsigned)((nind + 127) / 128)), dim3(128), 0, st,
                     (const double *)d_acc, (int64_t)1, nind, K, (const int32_t *)d_use,
                     (const int32_t *)nullptr, d_ans, d_rhs, d_st);
  hipLaunchKernelGGL(k_lsq_stale, dim3(1), dim3(64), 0, st, (const double *)d_ans,
                     (const double *)d_rhs, (const int32_t *)d_st, nind, K, acoeffs ? d_A : nullptr,
                     bcoeffs ? d_B : nullptr);
  EIG_HIP(ctx, hipGetLastError());
  if (status) EIG_HIP(ctx, hipMemcpyAsync(status, d_st, 4 * (size_t)nind, hipMemcpyDeviceToHost, st));
  if (acoeffs) EIG_HIP(ctx, hipMemcpyAsync(acoeffs, d_A, 8 * kn, hipMemcpyDeviceToHost, st));
  if (bcoeffs) EIG_HIP(ctx, hipMemcpyAsync(bcoeffs, d_B, 8 * kn, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  return EIG_OK;
}
