// fastmode: randomized block-Krylov PCA of the normalised genotype matrix
// (kjg_fpca, ksrc/kjg_fpca.c:24-101, with setgval's normalisation, gval.c:31-87).
//
// X is the ncols x nrows matrix of SNP rows, X[m][i] = ((g - mean) * fancy) / sqrt(2)
// with mean = xmean / xfancy exactly as setgval recomputes it, and 0 for a missing
// genotype.  Omega (nrows x L) is a Gaussian panel; for i < I
//     Q_i = X Omega_i,   Omega_{i+1} = X^T Q_i / ncols,
// Q_I = X Omega_I; the columns of Q = [Q_0 .. Q_I] are orthonormalised (the
// reference takes the SVD of Q; here block Gram-Schmidt with reorthogonalisation,
// each block orthonormalised by Gram + Jacobi eigendecomposition, gives an
// orthonormal basis U of the same range); then B = X^T U, and the top-K left
// singular vectors of B (from the eigen-decomposition of B^T B) are the
// eigenvectors, S^2 / ncols the eigenvalues (kjg_fpca.c:79-95).
//
// The two panel products read X straight from the 2-bit sample-tile layout and
// multiply on FP64 MFMA (16x16x4), or, in the split-BF16 form, on BF16 MFMA
// (32x32x16) over hi/lo splits; both are persistent kernels whose grid and
// split-K factor are picked by a timing tuner on the first call for a shape, and
// the 2I+1 products of the power iteration are replayed as one captured graph.
// The instances of up to 32 columns load each step's operands while the previous
// step's MFMAs run (two LDS buffers, one barrier per step).  The tall-skinny Gram
// and basis products run on the library's FP64 MFMA GEMM (gemm.hip).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "eig_internal.h"

typedef __attribute__((__vector_size__(4 * sizeof(double)))) double f64x4;

namespace {
constexpr int kKC = 32;       // samples (X A) or SNPs (X^T B) per LDS stage
constexpr int kMaxCT = 16;    // column tiles of 16 -> at most 256 columns
constexpr int kJacMax = 256;  // order limit of the one-block Jacobi eigensolver
}  // namespace

// ---------------------------------------------------------------- levels
// lf[m][g] for g = 0,1,2 and lf[m][3] = 0 (missing); SNPs not live or with no
// observed genotype get zero rows.  Built without contraction (gval.c:73-80).
__global__ void k_fpca_levels(const double *__restrict__ mean, const double *__restrict__ fancy,
                              const int32_t *__restrict__ status, int64_t M,
                              double *__restrict__ lf) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  double *o = lf + 4 * m;
  const int32_t st = status[m];
  if (st == EIG_SNP_DEAD || st == EIG_SNP_ALLMISS) {
    o[0] = o[1] = o[2] = o[3] = 0.0;
    return;
  }
  const double f = fancy[m];
  const double xm = mean[m] * f;  // xmean[col] = pmean * yfancy (smartpca.c:3205)
  const double mu = xm / f;       // mean = xmean[col] / xfancy[col] (gval.c:74)
  const double r2 = sqrt(2.0);
  for (int k = 0; k < 3; ++k) {
    double y = (double)k - mu;
    y *= f;
    o[k] = y / r2;
  }
  o[3] = 0.0;
}

// ---------------------------------------------------------------- Gaussian panel
__device__ __host__ inline uint64_t fp_mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

// Omega[i][c] ~ N(0,1) for i < n, c < L (Box-Muller on a counter hash); zero padding.
__global__ void k_fpca_gauss(double *__restrict__ A, int32_t npad, int32_t lda, int32_t n,
                             int32_t L, uint64_t seed) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)npad * lda) return;
  const int32_t i = (int32_t)(idx / lda), c = (int32_t)(idx % lda);
  double z = 0.0;
  if (i < n && c < L) {
    const uint64_t key = (uint64_t)i * (uint64_t)L + (uint64_t)c;
    const uint64_t h1 = fp_mix64(seed + (key + 1) * 0x9E3779B97F4A7C15ull);
    const uint64_t h2 = fp_mix64(h1 ^ 0xD1B54A32D192ED03ull);
    const double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;          // [0, 1)
    z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  A[idx] = z;
}

// ---------------------------------------------------------------- Y = X A
// Persistent over SNP tiles of 128 (4 waves x 32 SNPs); each wave holds a
// 32 x 16CT FP64 accumulator.  A: npad x lda (rows = samples, zero padded);
// Y rows m, columns [0, C) at Y[m * ldy + c].
template <int CT>
__global__ __launch_bounds__(256) void k_fpca_xa(const uint32_t *__restrict__ sel, int64_t M,
                                                 int32_t nst, const double *__restrict__ lf,
                                                 const double *__restrict__ A, int32_t lda,
                                                 double *__restrict__ Y, int64_t ldy, int32_t C) {
  __shared__ double sA[kKC][16 * CT];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ntiles = (M + 127) / 128;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t mb = tile * 128 + 32 * w;
    int64_t r[2];
    double lv[2][3];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      r[t] = mb + 16 * t + (lane & 15);
      const bool ok = r[t] < M;
#pragma unroll
      for (int g = 0; g < 3; ++g) lv[t][g] = ok ? lf[4 * r[t] + g] : 0.0;
    }
    f64x4 acc[2][CT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[t][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int32_t st = 0; st < nst; ++st) {
      for (int kc = 0; kc < kTileS; kc += kKC) {
        __syncthreads();
        const double *src = A + (int64_t)(st * kTileS + kc) * lda;
        for (int idx = threadIdx.x; idx < kKC * 16 * CT; idx += 256) {
          const int k = idx / (16 * CT), c = idx % (16 * CT);
          sA[k][c] = src[(int64_t)k * lda + c];
        }
        __syncthreads();
        uint32_t wd[2][2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          if (r[t] < M) {
            const uint2 v = *(const uint2 *)(sel + ((int64_t)st * M + r[t]) * kTileWords + kc / 16);
            wd[t][0] = v.x;
            wd[t][1] = v.y;
          } else {
            wd[t][0] = wd[t][1] = ~0u;
          }
        }
#pragma unroll
        for (int ks = 0; ks < kKC / 4; ++ks) {
          const int sh = 2 * ((ks & 3) * 4 + (lane >> 4));
          double a[2];
#pragma unroll
          for (int t = 0; t < 2; ++t) {
            const uint32_t code = (wd[t][ks >> 2] >> sh) & 3u;
            a[t] = code == 0 ? lv[t][0] : (code == 1 ? lv[t][1] : (code == 2 ? lv[t][2] : 0.0));
          }
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            const double b = sA[4 * ks + (lane >> 4)][16 * c + (lane & 15)];
            acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b, acc[0][c], 0, 0, 0);
            acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b, acc[1][c], 0, 0, 0);
          }
        }
      }
    }
    // accumulator layout: lane l, element e -> row l/16 + 4e, column l%16
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t row = mb + 16 * t + (lane >> 4) + 4 * e;
          const int col = 16 * c + (lane & 15);
          if (row < M && col < C) Y[row * ldy + col] = acc[t][c][e];
        }
  }
}

// The same Y = X A with the next sample stage's loads (the panel rows and each
// lane's genotype words) issued before the current stage's MFMAs: the panel stage
// goes global -> registers -> one of two LDS buffers, one barrier per stage.  For
// CT <= 2, as k_fpca_xtb_pf.
template <int CT>
__global__ __launch_bounds__(256, 4) void k_fpca_xa_pf(const uint32_t *__restrict__ sel,
                                                       int64_t M, int32_t nst,
                                                       const double *__restrict__ lf,
                                                       const double *__restrict__ A, int32_t lda,
                                                       double *__restrict__ Y, int64_t ldy,
                                                       int32_t C) {
  constexpr int kAE = kKC * 16 * CT / 256;  // panel elements per thread per stage
  constexpr int kStT = kTileS / kKC;        // stages per sample tile
  __shared__ double sA[2][kKC][16 * CT];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int64_t ntiles = (M + 127) / 128;
  const int32_t nstage = nst * kStT;
  double ra[kAE];
  uint2 rw[2];
  int buf = 0;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t mb = tile * 128 + 32 * w;
    int64_t r[2];
    double lv[2][3];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      r[t] = mb + 16 * t + (lane & 15);
      const bool ok = r[t] < M;
#pragma unroll
      for (int g = 0; g < 3; ++g) lv[t][g] = ok ? lf[4 * r[t] + g] : 0.0;
    }
    auto fetch = [&](int32_t q) {
      const int32_t st = q / kStT, kc = (q % kStT) * kKC;
      const double *src = A + (int64_t)(st * kTileS + kc) * lda;
#pragma unroll
      for (int j = 0; j < kAE; ++j) {
        const int idx = tid + 256 * j;
        ra[j] = src[(int64_t)(idx / (16 * CT)) * lda + idx % (16 * CT)];
      }
#pragma unroll
      for (int t = 0; t < 2; ++t)
        rw[t] = r[t] < M ? *(const uint2 *)(sel + ((int64_t)st * M + r[t]) * kTileWords + kc / 16)
                         : make_uint2(~0u, ~0u);
    };
    f64x4 acc[2][CT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[t][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    for (int32_t q = 0; q < nstage; ++q) {
#pragma unroll
      for (int j = 0; j < kAE; ++j) {
        const int idx = tid + 256 * j;
        sA[buf][idx / (16 * CT)][idx % (16 * CT)] = ra[j];
      }
      const uint32_t wd[2][2] = {{rw[0].x, rw[0].y}, {rw[1].x, rw[1].y}};
      // the other buffer was last read before this barrier's previous instance
      __syncthreads();
      if (q + 1 < nstage) fetch(q + 1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int ks = 0; ks < kKC / 4; ++ks) {
        const int sh = 2 * ((ks & 3) * 4 + (lane >> 4));
        double a[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const uint32_t code = (wd[t][ks >> 2] >> sh) & 3u;
          a[t] = code == 0 ? lv[t][0] : (code == 1 ? lv[t][1] : (code == 2 ? lv[t][2] : 0.0));
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          const double b = sA[buf][4 * ks + (lane >> 4)][16 * c + (lane & 15)];
          acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b, acc[0][c], 0, 0, 0);
          acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b, acc[1][c], 0, 0, 0);
        }
      }
      buf ^= 1;
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t row = mb + 16 * t + (lane >> 4) + 4 * e;
          const int col = 16 * c + (lane & 15);
          if (row < M && col < C) Y[row * ldy + col] = acc[t][c][e];
        }
  }
}

// ---------------------------------------------------------------- partial X^T B
// Persistent over (sample tile, SNP split) items; a block covers one sample tile
// (4 waves x 32 samples) and SNPs [s*mchunk, min(M, (s+1)*mchunk)).  B: rows m,
// columns [0, C) at B[m * ldb + c] (columns >= C read as 0).  part[s][i][c],
// i < nst*128, c < 16CT (row stride ldp).
template <int CT>
__global__ __launch_bounds__(256) void k_fpca_xtb(const uint32_t *__restrict__ sel, int64_t M,
                                                  int32_t nst, const double *__restrict__ lf,
                                                  const double *__restrict__ B, int64_t ldb,
                                                  int32_t C, int32_t S, int64_t mchunk,
                                                  double *__restrict__ part, int32_t ldp) {
  __shared__ double sB[kKC][16 * CT];
  __shared__ double sL[kKC][4];
  __shared__ uint32_t sW[kKC][kTileWords];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t items = (int64_t)nst * S;
  const int64_t npad = (int64_t)nst * kTileS;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int32_t st = (int32_t)(it % nst), s = (int32_t)(it / nst);
    const int64_t m0 = (int64_t)s * mchunk, m1 = min(M, m0 + mchunk);
    f64x4 acc[2][CT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[t][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int64_t mb = m0; mb < m1; mb += kKC) {
      __syncthreads();
      for (int idx = threadIdx.x; idx < kKC * 16 * CT; idx += 256) {
        const int k = idx / (16 * CT), c = idx % (16 * CT);
        const int64_t m = mb + k;
        sB[k][c] = (m < m1 && c < C) ? B[m * ldb + c] : 0.0;
      }
      if (threadIdx.x < kKC * 4) {
        const int k = threadIdx.x >> 2, g = threadIdx.x & 3;
        const int64_t m = mb + k;
        sL[k][g] = m < m1 ? lf[4 * m + g] : 0.0;
      } else if (threadIdx.x < kKC * 4 + kKC * kTileWords / 2) {
        const int q = threadIdx.x - kKC * 4;  // one uint2 of one SNP's tile record
        const int k = q >> 2, wv = (q & 3) * 2;
        const int64_t m = mb + k;
        uint2 v = make_uint2(~0u, ~0u);
        if (m < m1) v = *(const uint2 *)(sel + ((int64_t)st * M + m) * kTileWords + wv);
        sW[k][wv] = v.x;
        sW[k][wv + 1] = v.y;
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < kKC / 4; ++ks) {
        const int k = 4 * ks + (lane >> 4);
        double a[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int sl = 32 * w + 16 * t + (lane & 15);  // sample within the tile
          const uint32_t code = (sW[k][sl >> 4] >> (2 * (sl & 15))) & 3u;
          a[t] = sL[k][code];
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          const double b = sB[k][16 * c + (lane & 15)];
          acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b, acc[0][c], 0, 0, 0);
          acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b, acc[1][c], 0, 0, 0);
        }
      }
    }
    double *out = part + ((int64_t)s * npad + (int64_t)st * kTileS) * ldp;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 32 * w + 16 * t + (lane >> 4) + 4 * e;
          out[(int64_t)row * ldp + 16 * c + (lane & 15)] = acc[t][c][e];
        }
  }
}

// The same partial X^T B with the next SNP stage's loads issued before the current
// stage's MFMAs: the stage goes global -> registers -> one of two LDS buffers, so
// one barrier per stage suffices and the load latency hides behind 8 MFMA steps.
// For CT <= 2, whose prefetch registers fit beside the accumulators at four waves
// per SIMD (wider instances spill there).
template <int CT>
__global__ __launch_bounds__(256, 4) void k_fpca_xtb_pf(const uint32_t *__restrict__ sel, int64_t M,
                                                     int32_t nst, const double *__restrict__ lf,
                                                     const double *__restrict__ B, int64_t ldb,
                                                     int32_t C, int32_t S, int64_t mchunk,
                                                     double *__restrict__ part, int32_t ldp) {
  constexpr int kBE = kKC * 16 * CT / 256;  // B elements per thread per stage
  static_assert(kKC * 4 == 128 && kKC * kTileWords / 2 == 128, "stage loads take 256 threads");
  __shared__ double sB[2][kKC][16 * CT];
  __shared__ double sL[2][kKC][4];
  __shared__ uint32_t sW[2][kKC][kTileWords];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int64_t items = (int64_t)nst * S;
  const int64_t npad = (int64_t)nst * kTileS;
  double rb[kBE];
  double rl = 0.0;
  uint2 rw = make_uint2(~0u, ~0u);
  int buf = 0;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int32_t st = (int32_t)(it % nst), s = (int32_t)(it / nst);
    const int64_t m0 = (int64_t)s * mchunk, m1 = min(M, m0 + mchunk);
    // threads 0..127: level (k, g) = (tid/4, tid%4); 128..255: uint2 q of SNP k's record
    auto fetch = [&](int64_t mb) {
#pragma unroll
      for (int j = 0; j < kBE; ++j) {
        const int idx = tid + 256 * j;
        const int k = idx / (16 * CT), c = idx % (16 * CT);
        const int64_t m = mb + k;
        rb[j] = (m < m1 && c < C) ? B[m * ldb + c] : 0.0;
      }
      if (tid < 128) {
        const int64_t m = mb + (tid >> 2);
        rl = m < m1 ? lf[4 * m + (tid & 3)] : 0.0;
      } else {
        const int q = tid - 128;
        const int64_t m = mb + (q >> 2);
        rw = make_uint2(~0u, ~0u);
        if (m < m1) rw = *(const uint2 *)(sel + ((int64_t)st * M + m) * kTileWords + (q & 3) * 2);
      }
    };
    f64x4 acc[2][CT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[t][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    if (m0 < m1) fetch(m0);
    for (int64_t mb = m0; mb < m1; mb += kKC) {
#pragma unroll
      for (int j = 0; j < kBE; ++j) {
        const int idx = tid + 256 * j;
        sB[buf][idx / (16 * CT)][idx % (16 * CT)] = rb[j];
      }
      if (tid < 128) {
        sL[buf][tid >> 2][tid & 3] = rl;
      } else {
        const int q = tid - 128;
        sW[buf][q >> 2][(q & 3) * 2] = rw.x;
        sW[buf][q >> 2][(q & 3) * 2 + 1] = rw.y;
      }
      // the other buffer was last read before this barrier's previous instance
      __syncthreads();
      if (mb + kKC < m1) fetch(mb + kKC);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int ks = 0; ks < kKC / 4; ++ks) {
        const int k = 4 * ks + (lane >> 4);
        double a[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int sl = 32 * w + 16 * t + (lane & 15);  // sample within the tile
          const uint32_t code = (sW[buf][k][sl >> 4] >> (2 * (sl & 15))) & 3u;
          a[t] = sL[buf][k][code];
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          const double b = sB[buf][k][16 * c + (lane & 15)];
          acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b, acc[0][c], 0, 0, 0);
          acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b, acc[1][c], 0, 0, 0);
        }
      }
      buf ^= 1;
    }
    double *out = part + ((int64_t)s * npad + (int64_t)st * kTileS) * ldp;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 32 * w + 16 * t + (lane >> 4) + 4 * e;
          out[(int64_t)row * ldp + 16 * c + (lane & 15)] = acc[t][c][e];
        }
  }
}

// out[i][c] = scale * sum_s part[s][i][c] (splits in order) for c < min(ldp, ldo);
// padding columns of out beyond ldp are left alone.
__global__ void k_fpca_reduce(const double *__restrict__ part, int32_t S, int64_t npad,
                              int32_t ldp, double scale, double *__restrict__ out, int32_t ldo) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npad * ldp) return;
  const int64_t i = idx / ldp;
  const int c = (int)(idx % ldp);
  if (c >= ldo) return;
  double s = 0.0;
  for (int32_t k = 0; k < S; ++k) s += part[(int64_t)k * npad * ldp + idx];
  out[i * ldo + c] = scale * s;
}

// ---------------------------------------------------------------- split-BF16 panels
// The chosen form of the panel products: every operand is split into bf16 hi + lo
// and the four products hi*hi, hi*lo, lo*hi, lo*lo run on v_mfma_f32_32x32x16_bf16,
// in FP32 within a block of the reduction and promoted to FP64 across blocks (as
// the split-BF16 XTX).  The genotype operand is decoded straight into MFMA operand
// registers from LDS tables of level pairs; the dense panel is split once per
// product into transposed bf16 hi / lo arrays.
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float fp_f32x16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 fp_bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t fp_u32x4;

// fastmode levels lf as bf16 hi (low half) | lo (high half) per (SNP, code)
__global__ void k_fpca_levbf(const double *__restrict__ lf, int64_t M, uint32_t *__restrict__ lb) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 4 * M) return;
  const double v = lf[e];
  const uint16_t hi = eig_bf16_rne((float)v);
  const uint16_t lo = eig_bf16_rne((float)(v - (double)eig_bf16_to_f32(hi)));
  lb[e] = (uint32_t)hi | ((uint32_t)lo << 16);
}

// th/tl[c][r] = bf16 hi / lo of src[r * lds + c] for r < rows, c < C; zero up to
// rpad rows and cpad columns (row stride of th/tl: rpad)
__global__ void k_fpca_split_t(const double *__restrict__ src, int64_t rows, int64_t lds,
                               int32_t C, int64_t rpad, int32_t cpad, uint16_t *__restrict__ th,
                               uint16_t *__restrict__ tl) {
  __shared__ double tile[32][33];
  const int64_t r0 = (int64_t)blockIdx.x * 32;
  const int32_t c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 8 rows at a time
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = r0 + k;
    const int32_t c = c0 + tx;
    tile[k][tx] = (r < rows && c < C) ? src[r * lds + c] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int32_t c = c0 + k;
    const int64_t r = r0 + tx;
    if (c >= cpad || r >= rpad) continue;
    const double v = tile[tx][k];
    const uint16_t hi = eig_bf16_rne((float)v);
    const uint16_t lo = eig_bf16_rne((float)(v - (double)eig_bf16_to_f32(hi)));
    th[(int64_t)c * rpad + r] = hi;
    tl[(int64_t)c * rpad + r] = lo;
  }
}

// Y = X A.  Persistent over SNP tiles of 128 (4 waves x 32 SNPs); a wave's 32 SNPs x
// 32CT columns accumulate in FP32 over blkt sample tiles, then in FP64.  The table of
// SNP s (row s, stride 17 against bank conflicts) maps the codes c = code(2p) |
// code(2p+1) << 2 of two consecutive samples to their two bf16 hi (table 0) or lo
// (table 1) parts -- one dword of the operand.  Ah/Al: [column][sample], row stride ldt.
template <int CT>
__global__ __launch_bounds__(256, 2) void k_fpca_xa_bf16(
    const uint32_t *__restrict__ sel, int64_t M, int32_t nst, const uint32_t *__restrict__ lbf,
    const uint16_t *__restrict__ Ah, const uint16_t *__restrict__ Al, int64_t ldt,
    double *__restrict__ Y, int64_t ldy, int32_t C, int32_t blkt) {
  // two tables 8708 bytes apart: no paired LDS load spans that, so every lookup
  // lands in its operand register
  __shared__ uint32_t tt[2][128 * 17 + 1];
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][32 * CT][128 + 8];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, h = lane >> 5, r = lane & 31;
  const int64_t ntiles = (M + 127) / 128;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t mb = tile * 128;
    __syncthreads();  // the previous tile's table reads are done
    for (int e = t; e < 128 * 16; e += 256) {
      const int sl = e >> 4, c = e & 15;
      const int64_t m = mb + sl;
      uint32_t e0 = 0u, e1 = 0u;
      if (m < M) {
        e0 = lbf[4 * m + (c & 3)];
        e1 = lbf[4 * m + (c >> 2)];
      }
      tt[0][sl * 17 + c] = (e0 & 0xFFFFu) | (e1 << 16);
      tt[1][sl * 17 + c] = (e0 >> 16) | (e1 & 0xFFFF0000u);
    }
    const int64_t m = mb + 32 * w + r;  // this lane's SNP
    const uint32_t rowoff = (uint32_t)((32 * w + r) * 17);
    fp_f32x16 acc[CT];
    double dacc[CT][16];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        acc[ct][q] = 0.f;
        dacc[ct][q] = 0.0;
      }
    auto promote = [&]() {
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          dacc[ct][q] += (double)acc[ct][q];
          acc[ct][q] = 0.f;
        }
    };
    int32_t left = blkt;
    for (int32_t st = 0; st < nst; ++st) {
      __syncthreads();  // the previous sample tile's panel reads are done
      for (int q = t; q < 2 * 32 * CT * 16; q += 256) {
        const int hl = q / (32 * CT * 16), rem = q % (32 * CT * 16), c = rem >> 4, ch = rem & 15;
        const uint16_t *src = (hl ? Al : Ah) + (int64_t)c * ldt + (int64_t)st * kTileS + ch * 8;
        *(uint4 *)&sA[hl][c][ch * 8] = *(const uint4 *)src;
      }
      uint32_t wd[8];
      if (m < M) {
        const uint4 a = *(const uint4 *)(sel + ((int64_t)st * M + m) * kTileWords);
        const uint4 b = *(const uint4 *)(sel + ((int64_t)st * M + m) * kTileWords + 4);
        wd[0] = a.x; wd[1] = a.y; wd[2] = a.z; wd[3] = a.w;
        wd[4] = b.x; wd[5] = b.y; wd[6] = b.z; wd[7] = b.w;
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) wd[j] = 0xFFFFFFFFu;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 8; ++j) {  // 16-sample steps of the tile
        const uint32_t bits = wd[j] >> (16 * h);
        fp_u32x4 xh, xl;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t c = (bits >> (4 * q)) & 15u;
          xh[q] = tt[0][rowoff + c];
          xl[q] = tt[1][rowoff + c];
        }
        const fp_bf16x8 ah = __builtin_bit_cast(fp_bf16x8, xh), al = __builtin_bit_cast(fp_bf16x8, xl);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const fp_bf16x8 bh = *(const fp_bf16x8 *)&sA[0][32 * ct + r][16 * j + 8 * h];
          const fp_bf16x8 bl = *(const fp_bf16x8 *)&sA[1][32 * ct + r][16 * j + 8 * h];
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bl, acc[ct], 0, 0, 0);
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[ct], 0, 0, 0);
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[ct], 0, 0, 0);
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[ct], 0, 0, 0);
        }
      }
      if (--left == 0) {
        promote();
        left = blkt;
      }
    }
    promote();
    // accumulator: lane l, element q -> row 8(q>>2) + 4(l>>5) + (q&3), column l&31
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int64_t row = mb + 32 * w + 8 * (q >> 2) + 4 * h + (q & 3);
        const int col = 32 * ct + r;
        if (row < M && col < C) Y[row * ldy + col] = dacc[ct][q];
      }
  }
}

// The same Y = X A with the next sample tile's panel and genotype words loaded
// before the current tile's MFMAs: the panel goes global -> registers -> one of two
// LDS buffers, one barrier per tile.  For CT == 1, whose two panel buffers fit
// beside the tables at two blocks per CU.
template <int CT>
__global__ __launch_bounds__(256, 2) void k_fpca_xa_bf16_pf(
    const uint32_t *__restrict__ sel, int64_t M, int32_t nst, const uint32_t *__restrict__ lbf,
    const uint16_t *__restrict__ Ah, const uint16_t *__restrict__ Al, int64_t ldt,
    double *__restrict__ Y, int64_t ldy, int32_t C, int32_t blkt) {
  constexpr int kPE = 2 * 32 * CT * 16 / 256;  // panel uint4 per thread per tile
  // tables entry-major (entry c of SNP s at c * 128 + s): the 32 lanes of a half-wave
  // are 32 SNPs, so their lookups fall in 32 different banks whatever their codes; the
  // two tables 8196 bytes apart, which no paired LDS load spans
  __shared__ uint32_t tt[2][16 * 128 + 1];
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][2][32 * CT][128 + 8];  // [buf][hi/lo]
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, h = lane >> 5, r = lane & 31;
  const int64_t ntiles = (M + 127) / 128;
  fp_u32x4 pa[kPE], wa, wb;
  int buf = 0;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t mb = tile * 128;
    __syncthreads();  // the previous tile's table reads are done
    for (int e = t; e < 128 * 16; e += 256) {
      const int sl = e & 127, c = e >> 7;
      const int64_t m = mb + sl;
      uint32_t e0 = 0u, e1 = 0u;
      if (m < M) {
        e0 = lbf[4 * m + (c & 3)];
        e1 = lbf[4 * m + (c >> 2)];
      }
      tt[0][e] = (e0 & 0xFFFFu) | (e1 << 16);
      tt[1][e] = (e0 >> 16) | (e1 & 0xFFFF0000u);
    }
    const int64_t m = mb + 32 * w + r;  // this lane's SNP
    const uint32_t rowoff = (uint32_t)(32 * w + r);
    auto fetch = [&](int32_t st) {
#pragma unroll
      for (int j = 0; j < kPE; ++j) {
        const int q = t + 256 * j;
        const int hl = q / (32 * CT * 16), rem = q % (32 * CT * 16), c = rem >> 4, ch = rem & 15;
        pa[j] = *(const fp_u32x4 *)((hl ? Al : Ah) + (int64_t)c * ldt + (int64_t)st * kTileS + ch * 8);
      }
      // a SNP past M reads SNP M-1's words (in bounds) and takes all-missing codes
      const uint32_t *ws = sel + ((int64_t)st * M + (m < M ? m : M - 1)) * kTileWords;
      const fp_u32x4 x0 = *(const fp_u32x4 *)ws, x1 = *(const fp_u32x4 *)(ws + 4);
      const fp_u32x4 none = {~0u, ~0u, ~0u, ~0u};
      wa = m < M ? x0 : none;
      wb = m < M ? x1 : none;
    };
    fp_f32x16 acc[CT];
    double dacc[CT][16];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        acc[ct][q] = 0.f;
        dacc[ct][q] = 0.0;
      }
    auto promote = [&]() {
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          dacc[ct][q] += (double)acc[ct][q];
          acc[ct][q] = 0.f;
        }
    };
    fetch(0);
    int32_t left = blkt;
    for (int32_t st = 0; st < nst; ++st) {
#pragma unroll
      for (int j = 0; j < kPE; ++j) {
        const int q = t + 256 * j;
        const int hl = q / (32 * CT * 16), rem = q % (32 * CT * 16), c = rem >> 4, ch = rem & 15;
        *(fp_u32x4 *)&sA[buf][hl][c][ch * 8] = pa[j];
      }
      const fp_u32x4 ca = wa, cb = wb;  // this tile's words; wa, wb take the next tile's
      // the other buffer was last read before this barrier's previous instance
      __syncthreads();
      if (st + 1 < nst) fetch(st + 1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 8; ++j) {  // 16-sample steps of the tile
        const uint32_t bits = (j < 4 ? ca[j & 3] : cb[j & 3]) >> (16 * h);
        fp_u32x4 xh, xl;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t c = (bits >> (4 * q)) & 15u;
          xh[q] = tt[0][c * 128 + rowoff];
          xl[q] = tt[1][c * 128 + rowoff];
        }
        const fp_bf16x8 ah = __builtin_bit_cast(fp_bf16x8, xh), al = __builtin_bit_cast(fp_bf16x8, xl);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const fp_bf16x8 bh = *(const fp_bf16x8 *)&sA[buf][0][32 * ct + r][16 * j + 8 * h];
          const fp_bf16x8 bl = *(const fp_bf16x8 *)&sA[buf][1][32 * ct + r][16 * j + 8 * h];
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bl, acc[ct], 0, 0, 0);
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[ct], 0, 0, 0);
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[ct], 0, 0, 0);
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[ct], 0, 0, 0);
        }
      }
      if (--left == 0) {
        promote();
        left = blkt;
      }
      buf ^= 1;
    }
    promote();
    // accumulator: lane l, element q -> row 8(q>>2) + 4(l>>5) + (q&3), column l&31
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int64_t row = mb + 32 * w + 8 * (q >> 2) + 4 * h + (q & 3);
        const int col = 32 * ct + r;
        if (row < M && col < C) Y[row * ldy + col] = dacc[ct][q];
      }
  }
}

// Partial X^T B.  Items (sample tile, SNP split, column group of 32 NCT); a block covers
// one sample tile (4 waves x 32 samples) x 32 NCT columns over SNPs [s*mchunk, ...)
// (mchunk a multiple of 64).  The genotype operand comes from the grouped layout through
// per-group tables of level pairs (the split-BF16 XTX's scheme); the panel operand
// from Bh/Bl [column][SNP] (row stride ldt >= 64 * ceil(M/64)), staged per group.
// part[s][i][c] with row stride ldp.
constexpr int kFpTab = 4 * 4 * 2 * 16;
template <int NCT>
__global__ __launch_bounds__(256, 2) void k_fpca_xtb_bf16(
    const uint2 *__restrict__ selg, int64_t G64, int64_t M, int32_t nst,
    const uint32_t *__restrict__ lbf, const uint16_t *__restrict__ Bh,
    const uint16_t *__restrict__ Bl, int64_t ldt, int32_t ncg, int32_t S, int64_t mchunk,
    int32_t blkg, double *__restrict__ part, int32_t ldp) {
  __shared__ uint32_t tab[2][2][kFpTab + 1];  // [buf][hi/lo][entry], as the XTX kernel
  constexpr int GW = 32 * NCT;  // columns per group
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][2][GW][64 + 8];  // [buf][hi/lo][col][snp]
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, h = lane >> 5, r = lane & 31;
  const int64_t items = (int64_t)nst * S * ncg;
  const int64_t npad = (int64_t)nst * kTileS;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int32_t st = (int32_t)(it % nst);
    const int64_t rest = it / nst;
    const int32_t s = (int32_t)(rest % S), cg = (int32_t)(rest / S);
    const int64_t m_begin = (int64_t)s * mchunk;
    const int64_t m_end = min(M, m_begin + mchunk);
    if (m_begin >= m_end) {  // an empty split (M not filling S chunks): zero partial
      double *out = part + ((int64_t)s * npad + (int64_t)st * kTileS) * ldp;
      for (int e = t; e < kTileS * GW; e += 256) out[(int64_t)(e / GW) * ldp + GW * cg + e % GW] = 0.0;
      continue;
    }
    const int32_t ng = (int32_t)((m_end - m_begin + kGrp - 1) / kGrp);
    const int64_t g0 = m_begin / kGrp;
    const uint2 *src = selg + (((int64_t)st * G64 + g0) * kTileS + 32 * w + r) * 2 + h;
    auto lev_fetch = [&](int32_t gi, uint32_t (&lv)[4]) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = t + 256 * u;
        const int s4 = e >> 7, q = (e >> 5) & 3, hh = (e >> 4) & 1, c = e & 15;
        const int64_t m0 = m_begin + (int64_t)gi * kGrp + 16 * s4 + 2 * (4 * hh + q);
        const uint32_t e0 = lbf[(m0 < m_end ? m0 : m_begin) * 4 + (c & 3)];
        const uint32_t e1 = lbf[(m0 + 1 < m_end ? m0 + 1 : m_begin) * 4 + (c >> 2)];
        lv[2 * u] = m0 < m_end ? e0 : 0u;
        lv[2 * u + 1] = m0 + 1 < m_end ? e1 : 0u;
      }
    };
    auto tab_store = [&](int buf, const uint32_t (&lv)[4]) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = t + 256 * u;
        const uint32_t e0 = lv[2 * u], e1 = lv[2 * u + 1];
        tab[buf][0][e] = (e0 & 0xFFFFu) | (e1 << 16);
        tab[buf][1][e] = (e0 >> 16) | (e1 & 0xFFFF0000u);
      }
    };
    // panel chunk of group gi: 2 x GW columns x 64 SNPs, 2 NCT uint4 per thread (hi and
    // lo rows 32u + (t >> 3), u < NCT)
    const int pc = t >> 3, pch = t & 7;
    const uint16_t *ph0 = Bh + (int64_t)(GW * cg + pc) * ldt + m_begin + pch * 8;
    const uint16_t *pl0 = Bl + (int64_t)(GW * cg + pc) * ldt + m_begin + pch * 8;
    const int64_t p32 = 32 * ldt;
    fp_u32x4 pvh[NCT], pvl[NCT];
    auto pan_fetch = [&](int32_t gi) {
      const int64_t o = (int64_t)gi * kGrp;
#pragma unroll
      for (int u = 0; u < NCT; ++u) {
        pvh[u] = *(const fp_u32x4 *)(ph0 + u * p32 + o);
        pvl[u] = *(const fp_u32x4 *)(pl0 + u * p32 + o);
      }
    };
    auto pan_store = [&](int buf) {
#pragma unroll
      for (int u = 0; u < NCT; ++u) {
        *(fp_u32x4 *)&sB[buf][0][32 * u + pc][pch * 8] = pvh[u];
        *(fp_u32x4 *)&sB[buf][1][32 * u + pc][pch * 8] = pvl[u];
      }
    };
    fp_f32x16 acc[NCT];
    double dacc[NCT][16];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        acc[ct][q] = 0.f;
        dacc[ct][q] = 0.0;
      }
    auto promote = [&]() {
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          dacc[ct][q] += (double)acc[ct][q];
          acc[ct][q] = 0.f;
        }
    };
    uint32_t lv[4];
    uint2 cur = src[0], nxt = cur;
    __syncthreads();  // the previous item's reads of tab and sB are done
    lev_fetch(0, lv);
    pan_fetch(0);
    tab_store(0, lv);
    pan_store(0);
    __syncthreads();
    int32_t left = blkg;
    for (int32_t gi = 0; gi < ng; ++gi) {
      const int buf = gi & 1;
      const int32_t gn = gi + 1 < ng ? gi + 1 : gi;  // the last group loads itself again
      const uint32_t *th = tab[buf][0], *tl = tab[buf][1];
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        if (s4 == 0) {
          lev_fetch(gn, lv);
          pan_fetch(gn);
          nxt = src[(int64_t)gn * (2 * kTileS)];
        }
        const uint32_t word = (s4 < 2) ? cur.x : cur.y;
        fp_u32x4 xh, xl;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t c = (word >> (16 * (s4 & 1) + 4 * q)) & 15u;
          const uint32_t idx = (uint32_t)(((s4 * 4 + q) * 2) * 16) + 16u * (uint32_t)h + c;
          xh[q] = th[idx];
          xl[q] = tl[idx];
        }
        const fp_bf16x8 ah = __builtin_bit_cast(fp_bf16x8, xh), al = __builtin_bit_cast(fp_bf16x8, xl);
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
          const fp_bf16x8 bh = *(const fp_bf16x8 *)&sB[buf][0][32 * ct + r][16 * s4 + 8 * h];
          const fp_bf16x8 bl = *(const fp_bf16x8 *)&sB[buf][1][32 * ct + r][16 * s4 + 8 * h];
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bl, acc[ct], 0, 0, 0);
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[ct], 0, 0, 0);
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[ct], 0, 0, 0);
          acc[ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[ct], 0, 0, 0);
        }
      }
      if (--left == 0) {
        promote();
        left = blkg;
      }
      tab_store(buf ^ 1, lv);
      pan_store(buf ^ 1);
      cur = nxt;
      __syncthreads();
    }
    promote();
    double *out = part + ((int64_t)s * npad + (int64_t)st * kTileS) * ldp;
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int row = 32 * w + 8 * (q >> 2) + 4 * h + (q & 3);
        out[(int64_t)row * ldp + GW * cg + 32 * ct + r] = dacc[ct][q];
      }
  }
}

// ---------------------------------------------------------------- Jacobi
// All eigenpairs of a symmetric n x n matrix (n <= 256, column-major, lda) by
// cyclic two-sided Jacobi with round-robin pairing, in one block.  On exit the
// diagonal of A holds the eigenvalues and V (n x n, ldv) their eigenvectors in
// its columns (unsorted).  Stops after a sweep in which no pair had
// |a_pq| > tol * sqrt(|a_pp a_qq|), or after max_sweeps.
__global__ __launch_bounds__(1024) void k_jacobi(double *__restrict__ A, int32_t lda,
                                                 double *__restrict__ V, int32_t ldv, int32_t n,
                                                 int32_t max_sweeps, double tol,
                                                 int32_t *__restrict__ sweeps_out) {
  __shared__ int32_t idx[kJacMax];
  __shared__ int32_t pp[kJacMax / 2], pq[kJacMax / 2];
  __shared__ double cs[kJacMax / 2], sn[kJacMax / 2];
  __shared__ int rotated;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int np = n + (n & 1);  // an odd order gets a dummy index n
  const int half = np / 2;
  for (int e = tid; e < n * n; e += nt) {
    const int i = e % n, j = e / n;
    V[(int64_t)j * ldv + i] = (i == j) ? 1.0 : 0.0;
  }
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    if (tid == 0) rotated = 0;
    for (int e = tid; e < np; e += nt) idx[e] = e;
    __syncthreads();
    for (int round = 0; round < np - 1; ++round) {
      for (int k = tid; k < half; k += nt) {
        int p = idx[k], q = idx[np - 1 - k];
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = A[(int64_t)q * lda + p];
          const double app = A[(int64_t)p * lda + p], aqq = A[(int64_t)q * lda + q];
          if (fabs(apq) > tol * sqrt(fabs(app * aqq)) && apq != 0.0) {
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            rotated = 1;
          }
        }
        pp[k] = p;
        pq[k] = q < n ? q : p;
        cs[k] = c;
        sn[k] = q < n ? s : 0.0;
      }
      __syncthreads();
      // columns: A <- A J, V <- V J
      for (int e = tid; e < n * half; e += nt) {
        const int i = e % n, k = e / n;
        const double c = cs[k], s = sn[k];
        if (s == 0.0) continue;
        const int p = pp[k], q = pq[k];
        double *ap = A + (int64_t)p * lda + i, *aq = A + (int64_t)q * lda + i;
        const double x = *ap, y = *aq;
        *ap = c * x - s * y;
        *aq = s * x + c * y;
        double *vp = V + (int64_t)p * ldv + i, *vq = V + (int64_t)q * ldv + i;
        const double u = *vp, v = *vq;
        *vp = c * u - s * v;
        *vq = s * u + c * v;
      }
      __syncthreads();
      // rows: A <- J^T A
      for (int e = tid; e < n * half; e += nt) {
        const int j = e % n, k = e / n;
        const double c = cs[k], s = sn[k];
        if (s == 0.0) continue;
        const int p = pp[k], q = pq[k];
        double *ap = A + (int64_t)j * lda + p, *aq = A + (int64_t)j * lda + q;
        const double x = *ap, y = *aq;
        *ap = c * x - s * y;
        *aq = s * x + c * y;
      }
      __syncthreads();
      // exact zeros for the annihilated pairs
      for (int k = tid; k < half; k += nt)
        if (sn[k] != 0.0) {
          A[(int64_t)pq[k] * lda + pp[k]] = 0.0;
          A[(int64_t)pp[k] * lda + pq[k]] = 0.0;
        }
      // next pairing: keep idx[0], rotate idx[1..np-1] by one (np <= 256 < threads)
      {
        const int e = tid + 1;
        int v = 0;
        if (e < np) v = (e == 1) ? idx[np - 1] : idx[e - 1];
        __syncthreads();
        if (e < np) idx[e] = v;
        __syncthreads();
      }
    }
    if (!rotated) break;
    __syncthreads();
  }
  if (tid == 0) *sweeps_out = sweep;
}

// ---------------------------------------------------------------- host side
struct FpcaTune {
  int64_t M = -1;
  int32_t nst = -1, CT = -1;
  int32_t grid = 0, S = 0;
};

struct FpcaState {
  hipGraphExec_t gexec = nullptr;
  // everything the captured power-iteration graph bakes in: buffer addresses (the
  // workspace base and every offset in it), shape, form, launch shapes and the 1/ncols
  // scale; a call whose signature differs re-captures
  std::vector<int64_t> gsig;
  FpcaTune xa[kMaxCT + 1], xtb[kMaxCT + 1];
  // held for the whole call: the graph and the tuners are shared by
  // every context on the device
  std::mutex mu;
};
static FpcaState g_fpca[16];

#define FP_KCASE(n, ...) \
  case n: {              \
    constexpr int CT = n; \
    __VA_ARGS__;         \
  } break;
#define FP_CTSWITCH(ctx, ct, ...)                                                            \
  switch (ct) {                                                                              \
    FP_KCASE(1, __VA_ARGS__) FP_KCASE(2, __VA_ARGS__) FP_KCASE(3, __VA_ARGS__)               \
    FP_KCASE(4, __VA_ARGS__) FP_KCASE(5, __VA_ARGS__) FP_KCASE(6, __VA_ARGS__)               \
    FP_KCASE(7, __VA_ARGS__) FP_KCASE(8, __VA_ARGS__) FP_KCASE(9, __VA_ARGS__)               \
    FP_KCASE(10, __VA_ARGS__) FP_KCASE(11, __VA_ARGS__) FP_KCASE(12, __VA_ARGS__)            \
    FP_KCASE(13, __VA_ARGS__) FP_KCASE(14, __VA_ARGS__) FP_KCASE(15, __VA_ARGS__)            \
    FP_KCASE(16, __VA_ARGS__)                                                                \
    default:                                                                                 \
      return eig_set_err(ctx, EIG_E_ARG, "fpca: column count out of range");                 \
  }


namespace {
struct Plan {
  int64_t M;
  int32_t nst, npad;
  const uint32_t *sel;
  const double *lf;
};
}  // namespace

static int launch_xa(eig_ctx *ctx, const Plan &P, int CT, int grid, const double *A, int32_t lda,
                     double *Y, int64_t ldy, int32_t C, hipStream_t st) {
  FP_CTSWITCH(ctx, CT,
              if constexpr (CT <= 2) hipLaunchKernelGGL((k_fpca_xa_pf<CT>), dim3(grid), dim3(256),
                                                        0, st, P.sel, P.M, P.nst, P.lf, A, lda, Y,
                                                        ldy, C);
              else hipLaunchKernelGGL((k_fpca_xa<CT>), dim3(grid), dim3(256), 0, st, P.sel, P.M,
                                      P.nst, P.lf, A, lda, Y, ldy, C));
  return EIG_OK;
}

static int launch_xtb(eig_ctx *ctx, const Plan &P, int CT, int grid, int S, const double *B,
                      int64_t ldb, int32_t C, double *part, double *out, int32_t ldo, double scale,
                      hipStream_t st) {
  const int64_t mchunk = ((P.M + S - 1) / S + kKC - 1) / kKC * kKC;
  const int32_t ldp = 16 * CT;
  FP_CTSWITCH(ctx, CT,
              if constexpr (CT <= 2) hipLaunchKernelGGL((k_fpca_xtb_pf<CT>), dim3(grid), dim3(256),
                                                        0, st, P.sel, P.M, P.nst, P.lf, B, ldb, C,
                                                        S, mchunk, part, ldp);
              else hipLaunchKernelGGL((k_fpca_xtb<CT>), dim3(grid), dim3(256), 0, st, P.sel, P.M,
                                      P.nst, P.lf, B, ldb, C, S, mchunk, part, ldp));
  const int64_t tot = (int64_t)P.npad * ldp;
  hipLaunchKernelGGL(k_fpca_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     (const double *)part, S, (int64_t)P.npad, ldp, scale, out, ldo);
  return EIG_OK;
}

template <class Fn>
static int time_launch(eig_ctx *ctx, Fn &&fn, float &ms) {
  hipStream_t st = ctx->stream;
  EIG_HIP(ctx, hipEventRecord(ctx->ev0, st));
  int rc = fn();
  if (rc != EIG_OK) return rc;
  EIG_HIP(ctx, hipEventRecord(ctx->ev1, st));
  EIG_HIP(ctx, hipEventSynchronize(ctx->ev1));
  EIG_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  return EIG_OK;
}

// Timing tuner for X A: persistent grid for this shape (A: npad x 16CT, Y scratch).
static int tune_xa(eig_ctx *ctx, FpcaState &F, const Plan &P, int CT, const double *A,
                   int32_t lda, double *Y, int64_t ldy) {
  FpcaTune &ta = F.xa[CT];
  hipStream_t st = ctx->stream;
  const int cu = ctx->num_cu;
  if (ta.M != P.M || ta.nst != P.nst) {
    float best = 1e30f;
    for (int gm : {1, 2, 4}) {
      const int grid = (int)std::min<int64_t>((int64_t)gm * cu, (P.M + 127) / 128);
      float ms = 0.f;
      EIG_TRY(time_launch(
          ctx, [&] { return launch_xa(ctx, P, CT, grid, A, lda, Y, ldy, 16 * CT, st); }, ms));
      if (ms < best) {
        best = ms;
        ta.grid = grid;
      }
    }
    ta.M = P.M;
    ta.nst = P.nst;
  }
  return EIG_OK;
}

// Timing tuner for X^T B: persistent grid and split count (B: M x >= 16CT, out scratch).
static int tune_xtb(eig_ctx *ctx, FpcaState &F, const Plan &P, int CT, const double *B,
                    int64_t ldb, double *part, int32_t Smax, double *out, int32_t ldo) {
  FpcaTune &tb = F.xtb[CT];
  hipStream_t st = ctx->stream;
  const int cu = ctx->num_cu;
  if (tb.M != P.M || tb.nst != P.nst) {
    float best = 1e30f;
    for (int S : {1, 2, 4, 8}) {
      if (S > Smax) continue;
      for (int gm : {1, 2, 4}) {
        const int grid = (int)std::min<int64_t>((int64_t)gm * cu, (int64_t)P.nst * S);
        float ms = 0.f;
        EIG_TRY(time_launch(
            ctx,
            [&] { return launch_xtb(ctx, P, CT, grid, S, B, ldb, 16 * CT, part, out, ldo, 1.0, st); },
            ms));
        if (ms < best) {
          best = ms;
          tb.grid = grid;
          tb.S = S;
        }
      }
    }
    tb.M = P.M;
    tb.nst = P.nst;
  }
  return EIG_OK;
}

// Eigen-decomposition of the n x n symmetric device matrix G (column-major, ld n):
// eigenvalues descending in w (host), eigenvectors in the columns of V (device,
// ld n) in the same order.
// The same cyclic Jacobi for orders up to kJacLds, with A and V held in LDS (the
// global-memory version pays an L2 round trip per rotation round).
constexpr int kJacLds = 48;
__global__ __launch_bounds__(256) void k_jacobi_lds(double *__restrict__ A, int32_t lda,
                                                    double *__restrict__ V, int32_t ldv, int32_t n,
                                                    int32_t max_sweeps, double tol,
                                                    int32_t *__restrict__ sweeps_out) {
  __shared__ double sA[kJacLds][kJacLds + 1], sV[kJacLds][kJacLds + 1];
  __shared__ int32_t idx[kJacLds];
  __shared__ int32_t pp[kJacLds / 2], pq[kJacLds / 2];
  __shared__ double cs[kJacLds / 2], sn[kJacLds / 2];
  __shared__ int rotated;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int np = n + (n & 1);
  const int half = np / 2;
  // sA[j][i] = A[j * lda + i] (column j), sV likewise
  for (int e = tid; e < n * n; e += nt) {
    const int i = e % n, j = e / n;
    sA[j][i] = A[(int64_t)j * lda + i];
    sV[j][i] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    if (tid == 0) rotated = 0;
    for (int e = tid; e < np; e += nt) idx[e] = e;
    __syncthreads();
    for (int round = 0; round < np - 1; ++round) {
      for (int k = tid; k < half; k += nt) {
        int p = idx[k], q = idx[np - 1 - k];
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = sA[q][p], app = sA[p][p], aqq = sA[q][q];
          if (fabs(apq) > tol * sqrt(fabs(app * aqq)) && apq != 0.0) {
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            rotated = 1;
          }
        }
        pp[k] = p;
        pq[k] = q < n ? q : p;
        cs[k] = c;
        sn[k] = q < n ? s : 0.0;
      }
      __syncthreads();
      for (int e = tid; e < n * half; e += nt) {  // columns: A <- A J, V <- V J
        const int i = e % n, k = e / n;
        const double c = cs[k], s = sn[k];
        if (s == 0.0) continue;
        const int p = pp[k], q = pq[k];
        const double x = sA[p][i], y = sA[q][i];
        sA[p][i] = c * x - s * y;
        sA[q][i] = s * x + c * y;
        const double u = sV[p][i], v = sV[q][i];
        sV[p][i] = c * u - s * v;
        sV[q][i] = s * u + c * v;
      }
      __syncthreads();
      for (int e = tid; e < n * half; e += nt) {  // rows: A <- J^T A
        const int j = e % n, k = e / n;
        const double c = cs[k], s = sn[k];
        if (s == 0.0) continue;
        const int p = pp[k], q = pq[k];
        const double x = sA[j][p], y = sA[j][q];
        sA[j][p] = c * x - s * y;
        sA[j][q] = s * x + c * y;
      }
      __syncthreads();
      for (int k = tid; k < half; k += nt)
        if (sn[k] != 0.0) {
          sA[pq[k]][pp[k]] = 0.0;
          sA[pp[k]][pq[k]] = 0.0;
        }
      {
        const int e = tid + 1;
        int v = 0;
        if (e < np) v = (e == 1) ? idx[np - 1] : idx[e - 1];
        __syncthreads();
        if (e < np) idx[e] = v;
        __syncthreads();
      }
    }
    if (!rotated) break;
  }
  for (int e = tid; e < n * n; e += nt) {
    const int i = e % n, j = e / n;
    A[(int64_t)j * lda + i] = sA[j][i];
    V[(int64_t)j * ldv + i] = sV[j][i];
  }
  if (tid == 0) sweeps_out[0] = sweep;
}

// V column k = work column ord[k] (n x n, column-major)
__global__ void k_gather_cols(double *__restrict__ V, const double *__restrict__ work,
                              const int32_t *__restrict__ ord, int32_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int32_t k = (int32_t)(e / n), i = (int32_t)(e % n);
  V[e] = work[(int64_t)ord[k] * n + i];
}

// C (m x n, column-major, ld ldc) = sum over s of part[s] (m x n, ld m), in order
__global__ void k_sum_parts(const double *__restrict__ part, int32_t S, int32_t m, int32_t n,
                            double *__restrict__ C, int32_t ldc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t mn = (int64_t)m * n;
  if (e >= mn) return;
  double v = 0.0;
  for (int32_t sp = 0; sp < S; ++sp) v += part[(int64_t)sp * mn + e];
  C[(e / m) * ldc + (e % m)] = v;
}

// C (m x n, column-major, ld ldc) = A B^T, A m x k (ld lda) and B n x k (ld ldb) column-major,
// for a long k: k is cut into S equal chunks (one batched product) and a remainder,
// whose partial products are summed in chunk order.  A single product of these shapes
// has only a few output tiles to spread a long k over.
static int gemm_nt_long(eig_ctx *ctx, int32_t m, int32_t n, int64_t k, const double *A,
                        int64_t lda, const double *B, int64_t ldb, double *C, int32_t ldc,
                        double *part, size_t part_doubles, hipStream_t st) {
  const int64_t mn = (int64_t)m * n;
  int64_t S = std::min<int64_t>(64, k / 4096);
  S = std::min<int64_t>(S, (int64_t)(part_doubles / std::max<int64_t>(mn, 1)) - 1);
  if (S < 2)
    return eig_gemm_f64(ctx, st, false, true, m, n, k, 1.0, A, lda, 0, B, ldb, 0, 0.0, C, ldc, 0, 1);
  const int64_t kc = k / S, rem = k - S * kc;
  EIG_TRY(eig_gemm_f64(ctx, st, false, true, m, n, kc, 1.0, A, lda, kc * lda, B, ldb, kc * ldb, 0.0,
                       part, m, mn, (int32_t)S));
  int32_t nparts = (int32_t)S;
  if (rem > 0) {
    EIG_TRY(eig_gemm_f64(ctx, st, false, true, m, n, rem, 1.0, A + S * kc * lda, lda, 0,
                         B + S * kc * ldb, ldb, 0, 0.0, part + S * mn, m, 0, 1));
    ++nparts;
  }
  hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st,
                     (const double *)part, nparts, m, n, C, ldc);
  EIG_HIP(ctx, hipGetLastError());
  return EIG_OK;
}

static int small_eigh(eig_ctx *ctx, double *G, int32_t n, double *V, std::vector<double> &w,
                      double *work, int32_t *d_sweeps) {
  if (n > kJacMax) return eig_set_err(ctx, EIG_E_ARG, "fpca: basis larger than %d", kJacMax);
  hipStream_t st = ctx->stream;
  if (n <= kJacLds)
    hipLaunchKernelGGL(k_jacobi_lds, dim3(1), dim3(256), 0, st, G, n, work, n, n, 30, 1e-15,
                       d_sweeps);
  else
    hipLaunchKernelGGL(k_jacobi, dim3(1), dim3(1024), 0, st, G, n, work, n, n, 30, 1e-15,
                       d_sweeps);
  EIG_HIP(ctx, hipGetLastError());
  std::vector<double> a((size_t)n * n);
  int32_t sweeps = 0;
  EIG_HIP(ctx, hipMemcpyAsync(a.data(), G, sizeof(double) * n * n, hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipMemcpyAsync(&sweeps, d_sweeps, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  if (sweeps >= 30) return eig_set_err(ctx, EIG_E_CONVERGE, "fpca: Jacobi did not converge");
  std::vector<int32_t> ord(n);
  for (int32_t i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
    return a[(size_t)x * n + x] > a[(size_t)y * n + y];
  });
  w.resize(n);
  for (int32_t k = 0; k < n; ++k) w[k] = a[(size_t)ord[k] * n + ord[k]];
  int32_t *d_ord = d_sweeps + 1;  // kJacMax entries after the sweep count
  EIG_HIP(ctx, hipMemcpyAsync(d_ord, ord.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_gather_cols, dim3((unsigned)(((int64_t)n * n + 255) / 256)), dim3(256), 0,
                     st, V, (const double *)work, (const int32_t *)d_ord, n);
  EIG_HIP(ctx, hipGetLastError());
  EIG_HIP(ctx, hipStreamSynchronize(st));  // ord lives on this host frame
  return EIG_OK;
}

// Columns j < r of V scaled by 1 / sqrt(w_j) (device, column-major ld n).
__global__ void k_scale_cols(double *__restrict__ V, int32_t n, int32_t r,
                             const double *__restrict__ s) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * r) return;
  V[idx] *= s[idx / n];
}

// X A in the split-BF16 form: A (npad x lda, rows = samples) split into th/tl
// [cxa][npad], then the persistent product into Y (rows = SNPs, ld ldy)
static int launch_xa_bf16(eig_ctx *ctx, const Plan &P, const uint32_t *lbf, uint16_t *th,
                          uint16_t *tl, int32_t cxa, const double *A, int32_t lda, double *Y,
                          int64_t ldy, int32_t C, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_fpca_split_t, dim3((unsigned)(P.npad / 32), (unsigned)(cxa / 32)),
                     dim3(256), 0, st, A, (int64_t)P.npad, (int64_t)lda, C, (int64_t)P.npad, cxa,
                     th, tl);
  if (cxa == 32)
    hipLaunchKernelGGL(k_fpca_xa_bf16_pf<1>, dim3(grid), dim3(256), 0, st, P.sel, P.M, P.nst, lbf,
                       (const uint16_t *)th, (const uint16_t *)tl, (int64_t)P.npad, Y, ldy, C, 4);
  else
    hipLaunchKernelGGL(k_fpca_xa_bf16<2>, dim3(grid), dim3(256), 0, st, P.sel, P.M, P.nst, lbf,
                       (const uint16_t *)th, (const uint16_t *)tl, (int64_t)P.npad, Y, ldy, C, 4);
  EIG_HIP(ctx, hipGetLastError());
  return EIG_OK;
}

// scale * X^T B in the split-BF16 form: B (M x ldb, rows = SNPs) split into th/tl
// [64 ncg][Mpad]; the SNP range is cut into S chunks (multiples of 64) chosen like the
// XTX's split count; the partials are summed in split order into out (ld ldo)
static int launch_xtb_bf16(eig_ctx *ctx, const Plan &P, const uint2 *selg, const uint32_t *lbf,
                           uint16_t *th, uint16_t *tl, int64_t Mpad, const double *B, int64_t ldb,
                           int32_t C, double *part, size_t part_doubles, double *out, int32_t ldo,
                           double scale, const int64_t slots[2], hipStream_t st) {
  // 32-column groups for a panel of up to 32 columns (the power iteration's at L <= 32),
  // else 64-column groups, which decode X^T once per 64 columns
  const int nct = C <= 32 ? 1 : 2, gw = 32 * nct;
  const int32_t ncg = (C + gw - 1) / gw, cp = gw * ncg;
  hipLaunchKernelGGL(k_fpca_split_t, dim3((unsigned)(Mpad / 32), (unsigned)(cp / 32)), dim3(256),
                     0, st, B, P.M, ldb, C, Mpad, cp, th, tl);
  const int64_t G64 = Mpad / kGrp;
  const int64_t smax = std::max<int64_t>(
      1, std::min<int64_t>({64, G64, (int64_t)(part_doubles / ((size_t)P.npad * cp))}));
  int64_t chunk = Mpad, nsplit = 1;
  double best = -1.0;
  for (int64_t sp = 1; sp <= smax; ++sp) {
    int64_t c = (P.M + sp - 1) / sp;
    c = ((c + kGrp - 1) / kGrp) * kGrp;
    const int64_t ns = (P.M + c - 1) / c;
    if (ns != sp) continue;
    const double rounds = (double)((P.nst * ns * ncg + slots[nct - 1] - 1) / slots[nct - 1]);
    const double cost = rounds * (double)c / (double)P.M + 0.001 * (double)ns;
    if (best < 0.0 || cost < best - 1e-12) {
      best = cost;
      chunk = c;
      nsplit = ns;
    }
  }
  const int64_t items = (int64_t)P.nst * nsplit * ncg;
  const unsigned grid = (unsigned)std::min<int64_t>(items, slots[nct - 1]);
  if (nct == 1)
    hipLaunchKernelGGL(k_fpca_xtb_bf16<1>, dim3(grid), dim3(256), 0, st, selg, G64, P.M, P.nst,
                       lbf, (const uint16_t *)th, (const uint16_t *)tl, Mpad, ncg, (int32_t)nsplit,
                       chunk, 8, part, cp);
  else
    hipLaunchKernelGGL(k_fpca_xtb_bf16<2>, dim3(grid), dim3(256), 0, st, selg, G64, P.M, P.nst,
                       lbf, (const uint16_t *)th, (const uint16_t *)tl, Mpad, ncg, (int32_t)nsplit,
                       chunk, 8, part, cp);
  const int64_t tot = (int64_t)P.npad * cp;
  hipLaunchKernelGGL(k_fpca_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     (const double *)part, (int32_t)nsplit, (int64_t)P.npad, cp, scale, out, ldo);
  EIG_HIP(ctx, hipGetLastError());
  return EIG_OK;
}

extern "C" int eig_fpca(eig_ctx *ctx, int32_t K, int32_t L, int32_t I, uint64_t seed,
                        const double *omega, double *evals, double *evecs) {
  return eig_fpca_form(ctx, K, L, I, seed, omega, EIG_FORM_FP64, evals, evecs);
}

extern "C" int eig_fpca_form(eig_ctx *ctx, int32_t K, int32_t L, int32_t I, uint64_t seed,
                             const double *omega, int32_t form, double *evals, double *evecs) {
  if (!ctx || K <= 0 || L <= K || I <= 0 || !evals)
    return eig_set_err(ctx, EIG_E_ARG, "eig_fpca: need 0 < K < L and I > 0");
  if (form != EIG_FORM_FP64 && form != EIG_FORM_BF16X4)
    return eig_set_err(ctx, EIG_E_ARG, "eig_fpca: form must be FP64 or BF16X4");
  const bool bf = form == EIG_FORM_BF16X4;
  if (bf && L > 64) return eig_set_err(ctx, EIG_E_ARG, "eig_fpca: the BF16X4 form needs L <= 64");
  const int32_t Ctot = (I + 1) * L;
  if (L > 16 * kMaxCT || Ctot > kJacMax)
    return eig_set_err(ctx, EIG_E_ARG, "eig_fpca: L <= %d and (I+1)*L <= %d", 16 * kMaxCT,
                       kJacMax);
  if (!ctx->have_stats) return eig_set_err(ctx, EIG_E_STATE, "eig_fpca before eig_snpstats");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  FpcaState &F = g_fpca[ctx->device & 15];
  std::lock_guard<std::mutex> lock(F.mu);
  hipStream_t st = ctx->stream;
  const int64_t M = ctx->nsnp;
  const int32_t n = ctx->nrows, nst = ctx->nst, npad = nst * kTileS;
  // ncols: SNPs on the path (not removed by the caller's filters)
  std::vector<int32_t> status(M);
  EIG_HIP(ctx, hipMemcpy(status.data(), ctx->status.p, 4 * M, hipMemcpyDeviceToHost));
  int64_t ncols = 0;
  for (int64_t m = 0; m < M; ++m) ncols += status[m] != EIG_SNP_DEAD;
  if (ncols == 0) return eig_set_err(ctx, EIG_E_ARG, "eig_fpca: no SNPs");

  const int CTL = (L + 15) / 16, lda = 16 * CTL;
  const int64_t ldq = ((Ctot + 15) / 16) * 16;
  const int Smax = 8;
  // workspace: lf | Omega x2 | Q | part | G | V | Vwork | U | B | small
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t off = 0;
  auto add = [&](size_t b) {
    size_t o = off;
    off += up(b);
    return o;
  };
  const size_t o_lf = add(sizeof(double) * 4 * M);
  const size_t o_a0 = add(sizeof(double) * npad * lda), o_a1 = add(sizeof(double) * npad * lda);
  const size_t o_q = add(sizeof(double) * M * ldq);
  const size_t o_u = add(sizeof(double) * M * ldq);
  const size_t o_part = add(sizeof(double) * (size_t)Smax * npad * 16 * kMaxCT);
  const size_t o_b = add(sizeof(double) * npad * ldq);
  const size_t o_g = add(sizeof(double) * kJacMax * kJacMax);
  const size_t o_v = add(sizeof(double) * kJacMax * kJacMax);
  const size_t o_vw = add(sizeof(double) * kJacMax * kJacMax);
  const size_t o_sm = add(sizeof(double) * (kJacMax + 64) + sizeof(int32_t) * (kJacMax + 1));
  const size_t o_cb = add(sizeof(double) * kJacMax * kJacMax);
  const size_t o_ev = add(sizeof(double) * (size_t)K * npad);
  // split-BF16 form: levels, the X A panel (32 * ceil(L/32) columns) and the X^T B
  // panel (up to 64 * ceil(C/64) columns, C <= (I+1)L) as transposed bf16 hi / lo
  const int64_t Mpad = ((M + kGrp - 1) / kGrp) * kGrp;
  const int32_t cxa = 32 * ((L + 31) / 32), cxb = 64 * ((Ctot + 63) / 64);
  const size_t o_lbf = bf ? add(sizeof(uint32_t) * 4 * M) : 0;
  const size_t o_ta = bf ? add(sizeof(uint16_t) * 2 * (size_t)cxa * npad) : 0;
  const size_t o_tb = bf ? add(sizeof(uint16_t) * 2 * (size_t)cxb * Mpad) : 0;
  void *old = ctx->scratch2.p;
  EIG_TRY(eig_ensure(ctx, ctx->scratch2, off));
  char *base = (char *)ctx->scratch2.p;
  double *lf = (double *)(base + o_lf), *A0 = (double *)(base + o_a0), *A1 = (double *)(base + o_a1);
  double *Q = (double *)(base + o_q), *U = (double *)(base + o_u), *part = (double *)(base + o_part);
  double *Bm = (double *)(base + o_b), *G = (double *)(base + o_g), *V = (double *)(base + o_v);
  double *Vw = (double *)(base + o_vw), *sm = (double *)(base + o_sm), *Ev = (double *)(base + o_ev);
  int32_t *d_sweeps = (int32_t *)(sm + kJacMax);
  double *Cb = (double *)(base + o_cb);
  const size_t part_doubles = (size_t)Smax * npad * 16 * kMaxCT;
  uint32_t *lbf = bf ? (uint32_t *)(base + o_lbf) : nullptr;
  uint16_t *tah = bf ? (uint16_t *)(base + o_ta) : nullptr, *tal = tah ? tah + (size_t)cxa * npad : nullptr;
  uint16_t *tbh = bf ? (uint16_t *)(base + o_tb) : nullptr, *tbl = tbh ? tbh + (size_t)cxb * Mpad : nullptr;


  eig_time_begin(ctx);
  hipLaunchKernelGGL(k_fpca_levels, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st,
                     (const double *)ctx->mean.p, (const double *)ctx->fancy.p,
                     (const int32_t *)ctx->status.p, M, lf);
  if (omega) {
    std::vector<double> h((size_t)npad * lda, 0.0);
    for (int32_t i = 0; i < n; ++i)
      for (int32_t c = 0; c < L; ++c) h[(size_t)i * lda + c] = omega[(size_t)i * L + c];
    EIG_HIP(ctx, hipMemcpyAsync(A0, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, st));
    EIG_HIP(ctx, hipStreamSynchronize(st));
  } else {
    const int64_t tot = (int64_t)npad * lda;
    hipLaunchKernelGGL(k_fpca_gauss, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, A0,
                       npad, lda, n, L, seed);
  }
  EIG_HIP(ctx, hipGetLastError());
  Plan P{M, nst, npad, (const uint32_t *)ctx->sel.p, lf};
  int64_t bslots[2] = {0, 0};  // resident X^T B blocks, 32- and 64-column groups
  int bxa_grid = 0;
  if (bf) {
    hipLaunchKernelGGL(k_fpca_levbf, dim3((unsigned)((4 * M + 255) / 256)), dim3(256), 0, st,
                       (const double *)lf, M, lbf);
    EIG_HIP(ctx, hipGetLastError());
    EIG_TRY(eig_ensure_selg(ctx));
    int occ = 0;
    EIG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_fpca_xtb_bf16<1>, 256, 0));
    bslots[0] = (int64_t)std::max(occ, 1) * std::max(ctx->num_cu, 1);
    occ = 0;
    EIG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_fpca_xtb_bf16<2>, 256, 0));
    bslots[1] = (int64_t)std::max(occ, 1) * std::max(ctx->num_cu, 1);
    occ = 0;
    EIG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_fpca_xa_bf16<2>, 256, 0));
    bxa_grid = (int)std::min<int64_t>((M + 127) / 128, (int64_t)std::max(occ, 1) * ctx->num_cu);
  } else {
    // the tuners write scratch outputs (U and Bm), never Q or the panels in use
    EIG_TRY(tune_xa(ctx, F, P, CTL, A0, lda, U, ldq));
    EIG_TRY(tune_xtb(ctx, F, P, CTL, U, ldq, part, Smax, Bm, lda));
  }
  auto xa = [&](const double *A_, double *Y_) -> int {
    if (bf)
      return launch_xa_bf16(ctx, P, lbf, tah, tal, cxa, A_, lda, Y_, ldq, L, bxa_grid, st);
    return launch_xa(ctx, P, CTL, F.xa[CTL].grid, A_, lda, Y_, ldq, L, st);
  };
  auto xtb = [&](const double *B_, double *out_) -> int {
    if (bf)
      return launch_xtb_bf16(ctx, P, (const uint2 *)ctx->selg.p, lbf, tbh, tbl, Mpad, B_, ldq, L,
                             part, part_doubles, out_, lda, 1.0 / (double)ncols, bslots, st);
    return launch_xtb(ctx, P, CTL, F.xtb[CTL].grid, F.xtb[CTL].S, B_, ldq, L, part, out_, lda,
                      1.0 / (double)ncols, st);
  };
  // ---- power iterations (captured once per shape and buffer set, then replayed)
  double inv_ncols = 1.0 / (double)ncols;
  int64_t inv_bits;
  std::memcpy(&inv_bits, &inv_ncols, sizeof inv_bits);
  const std::vector<int64_t> sig = {
      (int64_t)(uintptr_t)ctx->sel.p, (int64_t)(uintptr_t)ctx->selg.p, (int64_t)(uintptr_t)base,
      (int64_t)off, (int64_t)o_lf, (int64_t)o_a0, (int64_t)o_a1, (int64_t)o_q, (int64_t)o_part,
      (int64_t)o_lbf, (int64_t)o_ta, (int64_t)o_tb, M, nst, K, L, I, (int64_t)bf, inv_bits,
      bf ? ((int64_t)bslots[0] * 100000 + bslots[1]) * 100000 + bxa_grid
         : (int64_t)F.xa[CTL].grid * 1000 + F.xtb[CTL].S};
  (void)old;
  const bool same = F.gexec && F.gsig == sig;
  auto power = [&]() -> int {
    double *cur = A0, *nxt = A1;
    for (int32_t i = 0; i < I; ++i) {
      EIG_TRY(xa(cur, Q + (size_t)i * L));
      EIG_TRY(xtb(Q + (size_t)i * L, nxt));
      std::swap(cur, nxt);
    }
    EIG_TRY(xa(cur, Q + (size_t)I * L));
    return EIG_OK;
  };
  if (getenv("EIG_NO_GRAPH")) {
    EIG_TRY(power());
  } else {
    if (!same) {
      if (F.gexec) hipGraphExecDestroy(F.gexec);
      F.gexec = nullptr;
      hipGraph_t graph;
      EIG_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = power();
      hipError_t e = hipStreamEndCapture(st, &graph);
      if (rc != EIG_OK) return rc;
      if (e != hipSuccess) return eig_set_err(ctx, EIG_E_HIP, "fpca capture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&F.gexec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e != hipSuccess) return eig_set_err(ctx, EIG_E_HIP, "fpca graph: %s", hipGetErrorString(e));
      F.gsig = sig;
    }
    EIG_HIP(ctx, hipGraphLaunch(F.gexec, st));
  }
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "fpca_power");

  // ---- orthonormal basis U of range(Q), block by block (kjg_fpca takes the SVD of
  // Q): each Krylov block is projected against the basis so far twice, orthonormalised
  // from its column-scaled Gram matrix, projected once more and orthonormalised again.
  // The blocks differ in scale by powers of the top eigenvalue, so one Gram matrix of
  // all of Q would lose the small directions that the SVD keeps.
  eig_time_begin(ctx);
  std::vector<double> w;
  int32_t r = 0;
  // W (nc columns, column-major nc x M at Wp, ld ldq) -= U_r (U_r^T W)
  auto project = [&](double *Wp, int32_t nc) -> int {
    if (r == 0) return EIG_OK;
    EIG_TRY(gemm_nt_long(ctx, r, nc, M, U, ldq, Wp, ldq, Cb, r, part, part_doubles, st));
    EIG_TRY(eig_gemm_f64(ctx, st, true, false, nc, M, r, -1.0, Cb, r, 0, U, ldq, 0, 1.0, Wp, ldq,
                         0, 1));
    return EIG_OK;
  };
  // Out (keep columns) = orthonormal basis of the nc columns at Wp
  auto orth = [&](const double *Wp, int32_t nc, double *Out, int32_t &keep) -> int {
    keep = 0;
    EIG_TRY(gemm_nt_long(ctx, nc, nc, M, Wp, ldq, Wp, ldq, G, nc, part, part_doubles, st));
    std::vector<double> g((size_t)nc * nc), dsc(nc);
    EIG_HIP(ctx, hipMemcpyAsync(g.data(), G, sizeof(double) * g.size(), hipMemcpyDeviceToHost, st));
    EIG_HIP(ctx, hipStreamSynchronize(st));
    for (int32_t i = 0; i < nc; ++i) {
      const double gi = g[(size_t)i * nc + i];
      dsc[i] = gi > 0.0 ? sqrt(gi) : 1.0;
    }
    for (int32_t jc = 0; jc < nc; ++jc)
      for (int32_t i = 0; i < nc; ++i) g[(size_t)jc * nc + i] /= dsc[i] * dsc[jc];
    EIG_HIP(ctx, hipMemcpyAsync(G, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, st));
    EIG_TRY(small_eigh(ctx, G, nc, V, w, Vw, d_sweeps));
    if (!(w[0] > 0.0)) return EIG_OK;
    while (keep < nc && w[keep] > 1e-13 * w[0]) ++keep;
    std::vector<double> v((size_t)nc * keep);
    EIG_HIP(ctx, hipMemcpyAsync(v.data(), V, sizeof(double) * v.size(), hipMemcpyDeviceToHost, st));
    EIG_HIP(ctx, hipStreamSynchronize(st));
    for (int32_t k = 0; k < keep; ++k)
      for (int32_t i = 0; i < nc; ++i) v[(size_t)k * nc + i] /= sqrt(w[k]) * dsc[i];
    EIG_HIP(ctx, hipMemcpyAsync(V, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, st));
    EIG_TRY(eig_gemm_f64(ctx, st, true, false, keep, M, nc, 1.0, V, nc, 0, Wp, ldq, 0, 0.0, Out,
                         ldq, 0, 1));
    return EIG_OK;
  };
  for (int32_t blk = 0; blk <= I; ++blk) {
    double *Ub = U + r, *Qb = Q + (size_t)blk * L;
    EIG_HIP(ctx, hipMemcpy2DAsync(Ub, sizeof(double) * ldq, Qb, sizeof(double) * ldq,
                                  sizeof(double) * L, (size_t)M, hipMemcpyDeviceToDevice, st));
    EIG_TRY(project(Ub, L));
    EIG_TRY(project(Ub, L));
    int32_t k1 = 0, k2 = 0;
    EIG_TRY(orth(Ub, L, Qb, k1));  // into the block's own (now free) columns of Q
    if (k1 == 0) continue;
    EIG_TRY(project(Qb, k1));
    EIG_TRY(orth(Qb, k1, Ub, k2));
    r += k2;
  }
  EIG_HIP(ctx, hipStreamSynchronize(st));
  const double *src = U;
  if (r < K) return eig_set_err(ctx, EIG_E_NUMERIC, "fpca: Krylov basis rank %d < K", r);
  // ---- B = X^T U  (npad x r, ld ldq), then the top-K left singular vectors of B
  if (bf) {
    // 64-column groups, and a remainder of at most 32 columns in a 32-column group, which
    // pads less (a 64-column group took 20 ms, a 32-column one 12 ms at N=100,000 x
    // M=500,000) when its 32 output columns stay within the row stride ldq
    const int32_t rem = r % 64;
    const int32_t r64 = (rem != 0 && rem <= 32 && r - rem + 32 <= ldq) ? r - rem : r;
    if (r64 > 0)
      EIG_TRY(launch_xtb_bf16(ctx, P, (const uint2 *)ctx->selg.p, lbf, tbh, tbl, Mpad, src, ldq,
                              r64, part, part_doubles, Bm, (int32_t)ldq, 1.0, bslots, st));
    if (r64 < r)
      EIG_TRY(launch_xtb_bf16(ctx, P, (const uint2 *)ctx->selg.p, lbf, tbh, tbl, Mpad, src + r64,
                              ldq, r - r64, part, part_doubles, Bm + r64, (int32_t)ldq, 1.0, bslots,
                              st));
  } else {
    // in 32-column pieces: the wide instances hold all r columns' accumulators and run
    // one wave per SIMD, at about half the FP64 rate of k_fpca_xtb<2> (four waves per
    // SIMD), which the power iteration has tuned already when L <= 32.  A last piece
    // whose 32 columns would pass ldq (a multiple of 16 >= r) takes 16.
    for (int32_t c0 = 0; c0 < r; c0 += 32) {
      const int32_t cw = std::min<int32_t>(32, r - c0);
      const int ct = c0 + 32 <= ldq ? 2 : 1;
      EIG_TRY(tune_xtb(ctx, F, P, ct, src + c0, ldq, part, Smax, Bm + c0, (int32_t)ldq));
      EIG_TRY(launch_xtb(ctx, P, ct, F.xtb[ct].grid, F.xtb[ct].S, src + c0, ldq, cw, part,
                         Bm + c0, (int32_t)ldq, 1.0, st));
    }
  }
  EIG_TRY(gemm_nt_long(ctx, r, r, npad, Bm, ldq, Bm, ldq, G, r, part, part_doubles, st));
  // top-K eigenpairs of B^T B (order r up to 256) by the FP64 eigensolver: all
  // eigenvalues, K eigenvectors into V (column-major r x K)
  {
    std::vector<double> evh((size_t)K * r);
    w.assign(r, 0.0);
    EIG_TRY(eig_run_eigh(ctx, G, r, K, w.data(), evh.data(), true));
    EIG_HIP(ctx, hipMemcpyAsync(V, evh.data(), sizeof(double) * evh.size(), hipMemcpyHostToDevice,
                                st));
    EIG_HIP(ctx, hipStreamSynchronize(st));
  }
  std::vector<double> s(K);
  for (int32_t j = 0; j < K; ++j) {
    if (!(w[j] > 0.0)) return eig_set_err(ctx, EIG_E_NUMERIC, "fpca: singular value is zero");
    s[j] = 1.0 / sqrt(w[j]);
  }
  EIG_HIP(ctx, hipMemcpyAsync(sm, s.data(), sizeof(double) * K, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_scale_cols, dim3((unsigned)(((int64_t)r * K + 255) / 256)), dim3(256), 0,
                     st, V, r, K, (const double *)sm);
  // Ev (column-major K x npad == row-major npad x K) = V_K^T B^T
  EIG_TRY(eig_gemm_f64(ctx, st, true, false, K, npad, r, 1.0, V, r, 0, Bm, ldq, 0, 0.0, Ev, K, 0,
                       1));
  std::vector<double> ev((size_t)npad * K);
  EIG_HIP(ctx, hipMemcpyAsync(ev.data(), Ev, sizeof(double) * ev.size(), hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  eig_time_end(ctx, "fpca_finish");
  for (int32_t j = 0; j < K; ++j) {
    evals[j] = w[j] / (double)ncols;  // S^2 / m (kjg_fpca.c:91-94)
    if (!evecs) continue;
    // unit norm, largest-|.| component positive
    double nrm = 0.0, big = 0.0;
    for (int32_t i = 0; i < n; ++i) {
      const double x = ev[(size_t)i * K + j];
      nrm += x * x;
      if (fabs(x) > fabs(big)) big = x;
    }
    const double f = (big < 0.0 ? -1.0 : 1.0) / sqrt(nrm);
    for (int32_t i = 0; i < n; ++i) evecs[(size_t)j * n + i] = f * ev[(size_t)i * K + j];
  }
  return EIG_OK;
}
