// XTX = Z^T Z over the kept SNPs of the selected samples, for the lower-triangle
// 128x128 sample tiles, with a fixed split of the SNP range (split-K) whose FP64
// partials are summed in split order -- results do not depend on scheduling.
//
// Z[m][i] = level of sample i's genotype at SNP m (missing -> 0), decoded on the
// fly from the 2-bit tile layout; nothing but packed codes and a 4-entry level
// table per SNP is read from HBM.
//
//  form FP64   : levels in FP64, v_mfma_f64_16x16x4 accumulation in FP64.
//  form BF16X3 : each level split into bf16 hi + lo; hi*hi + hi*lo + lo*hi on
//                v_mfma_f32_32x32x16_bf16 into FP32 accumulators, which are
//                added into FP64 accumulators every `blk` SNPs.
//  form BF16X4 : as BF16X3 plus lo*lo.  The dropped lo*lo term is not noise: every
//                pair of samples sharing a genotype at a SNP loses lo^2 > 0, a
//                systematic bias (~3e-6 of the diagonal) that dominates the error
//                of the top eigenvalues; the fourth product removes it.
#include <cmath>
#include <vector>

#include "eig_internal.h"

typedef __attribute__((__vector_size__(4 * sizeof(double)))) double f64x4;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;

__device__ inline void pair_to_tiles(int32_t p, int32_t &ti, int32_t &tj) {
  int32_t t = (int32_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while ((int64_t)t * (t + 1) / 2 > p) --t;
  while ((int64_t)(t + 1) * (t + 2) / 2 <= p) ++t;
  ti = t;
  tj = p - t * (t + 1) / 2;
}

// ---------------------------------------------------------------- FP64 form
constexpr int kKT64 = 16;  // SNPs per LDS stage

__global__ __launch_bounds__(256, 2) void k_syrk_f64(const uint32_t *__restrict__ sel,
                                                     int64_t M, const double *__restrict__ lev,
                                                     int32_t P, int64_t chunk,
                                                     double *__restrict__ part) {
  __shared__ __attribute__((aligned(16))) double zA[2][kKT64][kTileS];
  __shared__ __attribute__((aligned(16))) double zB[2][kKT64][kTileS];
  const int32_t pair = blockIdx.x % P;
  const int32_t split = blockIdx.x / P;
  int32_t ti, tj;
  pair_to_tiles(pair, ti, tj);
  const bool diag = (ti == tj);
  const int64_t m_begin = (int64_t)split * chunk;
  const int64_t m_end = min(M, m_begin + chunk);
  const int t = threadIdx.x;
  const int dk = t >> 4, dc = t & 15;  // decode: SNP row dk, samples 8dc..8dc+7
  const int lane = t & 63, w = t >> 6;
  const int wr = w >> 1, wc = w & 1;

  const uint32_t *selA = sel + (int64_t)ti * M * kTileWords;
  const uint32_t *selB = sel + (int64_t)tj * M * kTileWords;

  f64x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (f64x4){0.0, 0.0, 0.0, 0.0};

  uint32_t wa = 0xFFFFFFFFu, wb = 0xFFFFFFFFu;
  double l0 = 0, l1 = 0, l2 = 0;
  auto fetch = [&](int64_t m0) {
    int64_t m = m0 + dk;
    if (m < m_end) {
      wa = selA[m * kTileWords + (dc >> 1)];
      wb = diag ? wa : selB[m * kTileWords + (dc >> 1)];
      const double *L = lev + 4 * m;
      l0 = L[0];
      l1 = L[1];
      l2 = L[2];
    } else {
      wa = wb = 0xFFFFFFFFu;
      l0 = l1 = l2 = 0.0;
    }
  };
  auto decode = [&](int buf) {
    const int sh = 16 * (dc & 1);
    double va[8], vb[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      uint32_t ca = (wa >> (sh + 2 * s)) & 3u, cb = (wb >> (sh + 2 * s)) & 3u;
      va[s] = ca == 0 ? l0 : (ca == 1 ? l1 : (ca == 2 ? l2 : 0.0));
      vb[s] = cb == 0 ? l0 : (cb == 1 ? l1 : (cb == 2 ? l2 : 0.0));
    }
    double2 *pa = (double2 *)&zA[buf][dk][8 * dc];
#pragma unroll
    for (int s = 0; s < 4; ++s) pa[s] = make_double2(va[2 * s], va[2 * s + 1]);
    if (!diag) {
      double2 *pb = (double2 *)&zB[buf][dk][8 * dc];
#pragma unroll
      for (int s = 0; s < 4; ++s) pb[s] = make_double2(vb[2 * s], vb[2 * s + 1]);
    }
  };

  const int64_t nk = (m_end - m_begin + kKT64 - 1) / kKT64;
  if (nk > 0) {
    fetch(m_begin);
    decode(0);
  }
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int buf = (int)(kt & 1);
    if (kt + 1 < nk) fetch(m_begin + (kt + 1) * kKT64);
    double(*zb)[kTileS] = diag ? zA[buf] : zB[buf];
#pragma unroll
    for (int kk = 0; kk < kKT64 / 4; ++kk) {
      const int kr = 4 * kk + (lane >> 4);
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = zA[buf][kr][64 * wr + 16 * a + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = zb[kr][64 * wc + 16 * b + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    if (kt + 1 < nk) decode(buf ^ 1);
    __syncthreads();
  }
  double *out = part + ((int64_t)split * P + pair) * (kTileS * kTileS);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // v_mfma_f64_16x16x4: lane l, result r holds D[l/16 + 4r][l%16]
        int row = 64 * wr + 16 * a + (lane >> 4) + 4 * r;
        int col = 64 * wc + 16 * b + (lane & 15);
        out[row * kTileS + col] = acc[a][b][r];
      }
}

// ---------------------------------------------------------------- BF16x3 form
constexpr int kKT16 = 16;      // SNPs per LDS stage (one MFMA k-step)
constexpr int kRowB = 24;      // bf16 per LDS row: 16 + 8 pad -> 48 B stride

template <bool kLoLo>
__global__ __launch_bounds__(256, 2) void k_syrk_bf16x3(
    const uint32_t *__restrict__ sel, int64_t M, const uint32_t *__restrict__ levbf, int32_t P,
    int64_t chunk, int32_t blk, double *__restrict__ part) {
  // operand tiles [buf][operand A/B][hi/lo][sample][k]
  __shared__ __attribute__((aligned(16))) uint16_t zop[2][2][2][kTileS][kRowB];
  // raw codes [buf][operand][k][word] and levels [buf][k][4]
  __shared__ uint32_t raw[2][2][kKT16][kTileWords];
  __shared__ uint32_t lvl[2][kKT16][4];
  const int32_t pair = blockIdx.x % P;
  const int32_t split = blockIdx.x / P;
  int32_t ti, tj;
  pair_to_tiles(pair, ti, tj);
  const bool diag = (ti == tj);
  const int64_t m_begin = (int64_t)split * chunk;
  const int64_t m_end = min(M, m_begin + chunk);
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int wr = w >> 1, wc = w & 1;
  const uint32_t *selA = sel + (int64_t)ti * M * kTileWords;
  const uint32_t *selB = sel + (int64_t)tj * M * kTileWords;

  // fetch: thread t < 128 -> raw word (k = t>>3, word = t&7) of A; 128..255 -> of B;
  // threads 0..63 also fetch the level entry (k = t>>2, code = t&3).
  uint32_t rw = 0xFFFFFFFFu, rl = 0;
  auto fetch = [&](int64_t m0) {
    const int op = t >> 7, k = (t & 127) >> 3, wd = t & 7;
    int64_t m = m0 + k;
    rw = 0xFFFFFFFFu;
    if (m < m_end) rw = (op ? selB : selA)[m * kTileWords + wd];
    if (t < 64) {
      int64_t ml = m0 + (t >> 2);
      rl = (ml < m_end) ? levbf[ml * 4 + (t & 3)] : 0u;
    }
  };
  auto stash = [&](int buf) {
    const int op = t >> 7, k = (t & 127) >> 3, wd = t & 7;
    raw[buf][op][k][wd] = rw;
    if (t < 64) lvl[buf][t >> 2][t & 3] = rl;
  };
  // decode: thread t -> operand op = t>>7 (skipped for B on diagonal tiles),
  // sample i = t & 127, SNPs 0..15 of the stage -> 16 bf16 hi + 16 lo.
  auto decode = [&](int rbuf, int obuf) {
    const int op = t >> 7, i = t & 127;
    if (op == 1 && diag) return;
    const int wd = i >> 4, sh = 2 * (i & 15);
    uint32_t hv[8], lv[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) {
      uint32_t c0 = (raw[rbuf][op][2 * k2][wd] >> sh) & 3u;
      uint32_t c1 = (raw[rbuf][op][2 * k2 + 1][wd] >> sh) & 3u;
      uint32_t e0 = lvl[rbuf][2 * k2][c0], e1 = lvl[rbuf][2 * k2 + 1][c1];
      hv[k2] = (e0 & 0xFFFFu) | (e1 << 16);
      lv[k2] = (e0 >> 16) | (e1 & 0xFFFF0000u);
    }
    uint4 *ph = (uint4 *)&zop[obuf][op][0][i][0];
    uint4 *pl = (uint4 *)&zop[obuf][op][1][i][0];
    ph[0] = make_uint4(hv[0], hv[1], hv[2], hv[3]);
    ph[1] = make_uint4(hv[4], hv[5], hv[6], hv[7]);
    pl[0] = make_uint4(lv[0], lv[1], lv[2], lv[3]);
    pl[1] = make_uint4(lv[4], lv[5], lv[6], lv[7]);
  };

  f32x16 acc[2][2];
  double dacc[2][2][16];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc[a][b][r] = 0.f;
        dacc[a][b][r] = 0.0;
      }
    }

  const int64_t nk = (m_end - m_begin + kKT16 - 1) / kKT16;
  const int64_t kper = blk / kKT16;  // stages per FP32 block
  // prologue: raw(0) -> LDS, decode(0); raw(1) -> LDS
  if (nk > 0) {
    fetch(m_begin);
    stash(0);
  }
  __syncthreads();
  if (nk > 0) decode(0, 0);
  if (nk > 1) {
    fetch(m_begin + kKT16);
    stash(1);
  }
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int ob = (int)(kt & 1);
    if (kt + 2 < nk) fetch(m_begin + (kt + 2) * kKT16);
    // compute stage kt
    {
      const int opB = diag ? 0 : 1;
      const int kofs = 8 * (lane >> 5);
      bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int i = 64 * wr + 32 * a + (lane & 31);
        ah[a] = *(const bf16x8 *)&zop[ob][0][0][i][kofs];
        al[a] = *(const bf16x8 *)&zop[ob][0][1][i][kofs];
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int j = 64 * wc + 32 * b + (lane & 31);
        bh[b] = *(const bf16x8 *)&zop[ob][opB][0][j][kofs];
        bl[b] = *(const bf16x8 *)&zop[ob][opB][1][j][kofs];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          if (kLoLo)
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bl[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
        }
    }
    // promote the FP32 block sums into FP64 every blk SNPs and at the end
    if (((kt + 1) % kper) == 0 || kt + 1 == nk) {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            dacc[a][b][r] += (double)acc[a][b][r];
            acc[a][b][r] = 0.f;
          }
    }
    if (kt + 1 < nk) decode(ob ^ 1, ob ^ 1);
    if (kt + 2 < nk) stash(ob);
    __syncthreads();
  }
  double *out = part + ((int64_t)split * P + pair) * (kTileS * kTileS);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        int row = 64 * wr + 32 * a + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        int col = 64 * wc + 32 * b + (lane & 31);
        out[row * kTileS + col] = dacc[a][b][r];
      }
}

// ---------------------------------------------------------------- reduction
// trace of the unscaled matrix from the diagonal tiles (one block, fixed order)
__global__ void k_trace(const double *__restrict__ part, int32_t P, int32_t nsplit,
                        int32_t nrows, double *__restrict__ tr) {
  __shared__ double red[256];
  double s = 0.0;
  for (int32_t i = threadIdx.x; i < nrows; i += blockDim.x) {
    int32_t ti = i / kTileS, r = i % kTileS;
    int32_t p = ti * (ti + 1) / 2 + ti;
    double v = 0.0;
    for (int32_t sp = 0; sp < nsplit; ++sp)
      v += part[((int64_t)sp * P + p) * (kTileS * kTileS) + r * kTileS + r];
    s += v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) tr[0] = red[0];
}

// full symmetric matrix, scaled by 1/y with y = trace/(nrows-1)
__global__ void k_unpack_scale(const double *__restrict__ part, int32_t P, int32_t nsplit,
                               int32_t nrows, const double *__restrict__ tr,
                               double *__restrict__ xtx) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = nrows;
  if (idx >= n * n) return;
  int32_t i = (int32_t)(idx / n), j = (int32_t)(idx % n);
  int32_t a = i > j ? i : j, b = i > j ? j : i;  // lower-triangle element (a >= b)
  int32_t ta = a / kTileS, tb = b / kTileS;
  int32_t p = ta * (ta + 1) / 2 + tb;
  int64_t off = (int64_t)(a % kTileS) * kTileS + (b % kTileS);
  double v = 0.0;
  for (int32_t sp = 0; sp < nsplit; ++sp) v += part[((int64_t)sp * P + p) * (kTileS * kTileS) + off];
  double y = tr[0] / (double)(nrows - 1);
  xtx[idx] = v * (1.0 / y);
}

int eig_run_syrk(eig_ctx *ctx, int form, int bf16_block, double *trace_scale) {
  const int64_t M = ctx->nsnp;
  const int32_t T = ctx->nst;
  const int32_t P = T * (T + 1) / 2;
  const int32_t n = ctx->nrows;
  const bool split = (form == EIG_FORM_BF16X3 || form == EIG_FORM_BF16X4);
  if (split) {
    if (bf16_block <= 0) bf16_block = 4096;
    if (bf16_block % kKT16 != 0) return eig_set_err(ctx, EIG_E_ARG, "bf16_block % 16 != 0");
  }
  // Split the SNP range (split-K).  Every block of one launch runs equally long, so
  // the launch takes ceil(P*s / slots) rounds of M/s SNPs each, slots being the
  // blocks the chip holds at once: choose s to minimise that, plus a small charge
  // per split for the partial tiles it writes and the finaliser reads.  Chunks are
  // multiples of the FP32 block (bf16) or of 256 SNPs and hold at least 2048 SNPs.
  const int64_t align = split ? bf16_block : 256;
  int occ = 0;
  if (form == EIG_FORM_FP64)
    EIG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_syrk_f64, 256, 0));
  else if (form == EIG_FORM_BF16X3)
    EIG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_syrk_bf16x3<false>, 256, 0));
  else
    EIG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_syrk_bf16x3<true>, 256, 0));
  const int64_t slots = (int64_t)std::max(occ, 1) * std::max(ctx->num_cu, 1);
  const int64_t smax = std::max<int64_t>(1, std::min<int64_t>(32, M / std::max<int64_t>(align, 2048)));
  int64_t chunk = M, best = -1;
  double best_cost = 0.0;
  for (int64_t s = 1; s <= smax; ++s) {
    int64_t c = (M + s - 1) / s;
    c = ((c + align - 1) / align) * align;
    const int64_t ns = (M + c - 1) / c;
    if (ns != s) continue;  // alignment merged two splits: s is not reachable
    const double rounds = (double)((P * ns + slots - 1) / slots);
    const double cost = rounds * (double)c / (double)M + 0.001 * (double)ns;
    if (best < 0 || cost < best_cost - 1e-12) {
      best = s;
      best_cost = cost;
      chunk = c;
    }
  }
  int32_t nsplit = (int32_t)((M + chunk - 1) / chunk);
  size_t tile = (size_t)kTileS * kTileS;
  EIG_TRY(eig_ensure(ctx, ctx->part, sizeof(double) * (tile * P * nsplit + 64)));
  EIG_TRY(eig_ensure(ctx, ctx->xtx, sizeof(double) * (size_t)n * n));
  double *d_tr = (double *)ctx->part.p + tile * P * nsplit;
  unsigned grid = (unsigned)(P * nsplit);
  eig_time_begin(ctx);
  if (form == EIG_FORM_FP64) {
    hipLaunchKernelGGL(k_syrk_f64, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint32_t *)ctx->sel.p, M, (const double *)ctx->lev.p, P, chunk,
                       (double *)ctx->part.p);
  } else if (form == EIG_FORM_BF16X3) {
    hipLaunchKernelGGL(k_syrk_bf16x3<false>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint32_t *)ctx->sel.p, M, (const uint32_t *)ctx->levbf.p, P,
                       chunk, bf16_block, (double *)ctx->part.p);
  } else {
    hipLaunchKernelGGL(k_syrk_bf16x3<true>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint32_t *)ctx->sel.p, M, (const uint32_t *)ctx->levbf.p, P,
                       chunk, bf16_block, (double *)ctx->part.p);
  }
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "syrk");
  eig_time_begin(ctx);
  hipLaunchKernelGGL(k_trace, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->part.p,
                     P, nsplit, n, d_tr);
  int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(k_unpack_scale, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0,
                     ctx->stream, (const double *)ctx->part.p, P, nsplit, n,
                     (const double *)d_tr, (double *)ctx->xtx.p);
  EIG_HIP(ctx, hipGetLastError());
  double tr = 0.0;
  EIG_HIP(ctx, hipMemcpyAsync(&tr, d_tr, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  eig_time_end(ctx, "xtx_finalize");
  double y = tr / (double)(n - 1);
  if (trace_scale) *trace_scale = y;
  if (std::isnan(y)) return eig_set_err(ctx, EIG_E_NUMERIC, "bad XTX matrix");
  if (y <= 0.0) return eig_set_err(ctx, EIG_E_NUMERIC, "XTX has zero trace (perhaps no data)");
  ctx->ms["syrk_splits"] = nsplit;
  return EIG_OK;
}

extern "C" int eig_xtx(eig_ctx *ctx, int32_t form, int32_t bf16_block, double *trace_scale,
                       double *xtx_out) {
  if (!ctx) return EIG_E_ARG;
  if (form != EIG_FORM_FP64 && form != EIG_FORM_BF16X3 && form != EIG_FORM_BF16X4)
    return eig_set_err(ctx, EIG_E_ARG, "unknown XTX form %d", form);
  if (!ctx->have_stats) return eig_set_err(ctx, EIG_E_STATE, "eig_xtx before eig_snpstats");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  ctx->have_xtx = false;
  EIG_TRY(eig_run_syrk(ctx, form, bf16_block, trace_scale));
  if (xtx_out) {
    EIG_HIP(ctx, hipMemcpyAsync(xtx_out, ctx->xtx.p, sizeof(double) * (size_t)ctx->nrows * ctx->nrows,
                                hipMemcpyDeviceToHost, ctx->stream));
    EIG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  ctx->have_xtx = true;
  return EIG_OK;
}
