// Column-major FP64 GEMM on the FP64 matrix cores (v_mfma_f64_16x16x4f64), with the
// dgemm argument conventions: for each batch b,
//     C_b = alpha * op(A_b) * op(B_b) + beta * C_b,   op(X) = X or X^T,
// A_b = A + b * sA (and so on).  The engine's small and skinny products use it: the
// fastmode basis products (Gram matrices of long panels, panel updates, the final
// eigenvector product), and the dense-column XTX of the non-lookup column path.
//
// One block of 256 threads forms a 64 x 64 tile of C; each wave a 32 x 32 quarter as
// 2 x 2 MFMA tiles.  op(A) and op(B) tiles of 16 k-steps are staged in LDS (two
// buffers, one barrier per step) with each thread loading along the operand's
// contiguous dimension.  The MFMA takes the B tile as its first operand, so a lane's
// results are 16 consecutive rows of one column of C: the stores are coalesced in
// column-major order.
#include "eig_internal.h"

typedef __attribute__((__vector_size__(4 * sizeof(double)))) double f64x4;

namespace {
constexpr int kGM = 64, kGN = 64, kGK = 16;

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_f64(int64_t m, int64_t n, int64_t k, double alpha,
                                                  const double *__restrict__ A, int64_t lda,
                                                  int64_t sA, const double *__restrict__ B,
                                                  int64_t ldb, int64_t sB, double beta,
                                                  double *__restrict__ C, int64_t ldc, int64_t sC,
                                                  int32_t gm, int32_t lower) {
  const int64_t bi = blockIdx.x % gm, bj = blockIdx.x / gm;
  if (lower && bi < bj) return;  // whole block: strictly upper tile
  A += (int64_t)blockIdx.z * sA;
  B += (int64_t)blockIdx.z * sB;
  C += (int64_t)blockIdx.z * sC;
  __shared__ double As[2][kGK][kGM + 1];
  __shared__ double Bs[2][kGK][kGN + 1];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
  const int64_t i0 = bi * kGM, j0 = bj * kGN;
  double ra[4], rb[4];
  // element q of this thread's share of the op(A) tile (row ia, step ka) and of the
  // op(B) tile (step kb, column jb); the contiguous dimension runs along the threads
  auto a_idx = [&](int q, int &ii, int &kk) {
    if (!TA) { ii = t & 63; kk = (t >> 6) + 4 * q; }
    else { kk = t & 15; ii = (t >> 4) + 16 * q; }
  };
  auto b_idx = [&](int q, int &kk, int &jj) {
    if (!TB) { kk = t & 15; jj = (t >> 4) + 16 * q; }
    else { jj = t & 63; kk = (t >> 6) + 4 * q; }
  };
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int ii, kk, jj, kb;
      a_idx(q, ii, kk);
      const int64_t gi = i0 + ii, gk = k0 + kk;
      ra[q] = (gi < m && gk < k) ? (TA ? A[gk + gi * lda] : A[gi + gk * lda]) : 0.0;
      b_idx(q, kb, jj);
      const int64_t gj = j0 + jj, gkb = k0 + kb;
      rb[q] = (gj < n && gkb < k) ? (TB ? B[gj + gkb * ldb] : B[gkb + gj * ldb]) : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int ii, kk, jj, kb;
      a_idx(q, ii, kk);
      As[buf][kk][ii] = ra[q];
      b_idx(q, kb, jj);
      Bs[buf][kb][jj] = rb[q];
    }
  };
  f64x4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = (f64x4){0.0, 0.0, 0.0, 0.0};
  const int64_t nk = (k + kGK - 1) / kGK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int buf = (int)(kt & 1);
    if (kt + 1 < nk) load((kt + 1) * kGK);
#pragma unroll
    for (int s = 0; s < kGK / 4; ++s) {
      const int kr = 4 * s + (lane >> 4);
      double av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) av[x] = As[buf][kr][32 * wr + 16 * x + (lane & 15)];
#pragma unroll
      for (int y = 0; y < 2; ++y) bv[y] = Bs[buf][kr][32 * wc + 16 * y + (lane & 15)];
      // D = op(B)^T op(A)^T tile: lane l, result r holds C[l % 16][l / 16 + 4 r] of the
      // 16 x 16 sub-tile (the layout of v_mfma_f64_16x16x4f64 with the operands swapped)
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[y], av[x], acc[x][y], 0, 0, 0);
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = i0 + 32 * wr + 16 * x + (lane & 15);
        const int64_t col = j0 + 32 * wc + 16 * y + (lane >> 4) + 4 * r;
        if (row < m && col < n) {
          double *c = C + row + col * ldc;
          const double v = alpha * acc[x][y][r];
          *c = beta == 0.0 ? v : v + beta * *c;
        }
      }
}
}  // namespace

int eig_gemm_f64(eig_ctx *ctx, hipStream_t st, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                 double alpha, const double *A, int64_t lda, int64_t sA, const double *B,
                 int64_t ldb, int64_t sB, double beta, double *C, int64_t ldc, int64_t sC,
                 int32_t batch, bool lower) {
  if (m <= 0 || n <= 0 || batch <= 0) return EIG_OK;
  if (lower && m != n) return eig_set_err(ctx, EIG_E_ARG, "gemm: lower tiles need m == n");
  if (ldc < m || lda < (ta ? k : m) || ldb < (tb ? n : k))
    return eig_set_err(ctx, EIG_E_ARG, "gemm: leading dimension too small");
  const int64_t gm = (m + kGM - 1) / kGM, gn = (n + kGN - 1) / kGN;
  if (gm * gn > 0x7fffffff || batch > 65535)
    return eig_set_err(ctx, EIG_E_ARG, "gemm: grid too large");
  dim3 grid((unsigned)(gm * gn), 1, (unsigned)batch);
  const int32_t g32 = (int32_t)gm;
  const int32_t lw = lower ? 1 : 0;
  if (!ta && !tb)
    hipLaunchKernelGGL((k_gemm_f64<false, false>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda,
                       sA, B, ldb, sB, beta, C, ldc, sC, g32, lw);
  else if (!ta && tb)
    hipLaunchKernelGGL((k_gemm_f64<false, true>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda,
                       sA, B, ldb, sB, beta, C, ldc, sC, g32, lw);
  else if (ta && !tb)
    hipLaunchKernelGGL((k_gemm_f64<true, false>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda,
                       sA, B, ldb, sB, beta, C, ldc, sC, g32, lw);
  else
    hipLaunchKernelGGL((k_gemm_f64<true, true>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda,
                       sA, B, ldb, sB, beta, C, ldc, sC, g32, lw);
  EIG_HIP(ctx, hipGetLastError());
  return EIG_OK;
}
