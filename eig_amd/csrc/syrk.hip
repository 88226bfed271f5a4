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
#include <numeric>
#include <type_traits>
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

// ---------------------------------------------------------------- split-BF16 forms
// Grouped code layout: per sample tile st, 64-SNP group g and sample i, 16 bytes
// selg[(st * G64 + g) * 128 + i] whose word 2h + (s >> 1), bits 16 (s & 1) + 2 j,
// hold the code of SNP 64 g + 16 s + 8 h + j (s = 16-SNP stage 0..3, h = half of
// the stage 0..1, j = 0..7); SNPs past M carry code 3.  A lane of fastmode's split-
// BF16 X^T B product that feeds MFMA k-slots 8h..8h+7 of a sample reads one 8-byte
// word pair per group.

__global__ __launch_bounds__(256) void k_regroup(const uint32_t *__restrict__ sel, int64_t M,
                                                 int64_t G64, int64_t ngroups,
                                                 uint4 *__restrict__ selg) {
  __shared__ uint32_t w[2][kGrp * kTileWords];
  const int half = threadIdx.x >> 7, i = threadIdx.x & 127;
  const int64_t gid = (int64_t)blockIdx.x * 2 + half;  // over nst * G64 groups
  const bool live = gid < ngroups;
  const int64_t st = live ? gid / G64 : 0, g = live ? gid % G64 : 0;
  const uint32_t *src = sel + (st * M + g * kGrp) * kTileWords;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = i + 128 * q;  // SNP e >> 3 of the group, word e & 7
    w[half][e] = (live && g * kGrp + (e >> 3) < M) ? src[e] : 0xFFFFFFFFu;
  }
  __syncthreads();
  if (!live) return;
  const int wd = i >> 4, sh = 2 * (i & 15);
  uint32_t o[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t c = (w[half][(16 * s + 8 * h + j) * kTileWords + wd] >> sh) & 3u;
        o[2 * h + (s >> 1)] |= c << (16 * (s & 1) + 2 * j);
      }
  selg[gid * kTileS + i] = make_uint4(o[0], o[1], o[2], o[3]);
}

// Split-BF16 SYRK.  Block = 128x128 tile pair, 4 waves of 64x64; every wave
// decodes its own MFMA operands straight into registers from a per-group table of
// level pairs in LDS: for stage s, SNP pair p = 4h + q (SNPs 2p, 2p+1 of the
// stage) and combo c = code(2p) | code(2p+1) << 2, the table holds the bf16 hi
// parts of both levels in one word and the lo parts in another -- exactly one
// dword of the bf16x8 operand.  One table per 64-SNP group, built by the block from
// the bf16 level entries and double-buffered: one barrier per group.
//   table word index ((s * 4 + q) * 2 + h) * 16 + c: the two halves h of a wave
//   read disjoint bank ranges.
constexpr int kTabWords = 4 * 4 * 2 * 16;

// Byte-code layout for the split-BF16 SYRK: per sample tile st, 64-SNP group g,
// sample i and k-slot half h, 16 bytes at selb[((st * G64 + g) * 128 + i) * 2 + h];
// byte 4 s + q is the LDS byte offset, within the (stage s, pair q) row of the level
// table, of the entry for this sample's code pair: 64 h + 4 c, c = code(2p) |
// code(2p+1) << 2 for SNPs 2p, 2p+1 of stage s (p = 4h + q).  A lookup is then one
// byte extract; the stage/pair row and the buffer are immediate offsets.
__global__ __launch_bounds__(256) void k_regroup8(const uint32_t *__restrict__ sel, int64_t M,
                                                  int64_t G64, int64_t ngroups,
                                                  uint4 *__restrict__ selb) {
  __shared__ uint32_t w[2][kGrp * kTileWords];
  const int half = threadIdx.x >> 7, i = threadIdx.x & 127;
  const int64_t gid = (int64_t)blockIdx.x * 2 + half;  // over nst * G64 groups
  const bool live = gid < ngroups;
  const int64_t st = live ? gid / G64 : 0, g = live ? gid % G64 : 0;
  const uint32_t *src = sel + (st * M + g * kGrp) * kTileWords;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = i + 128 * q;  // SNP e >> 3 of the group, word e & 7
    w[half][e] = (live && g * kGrp + (e >> 3) < M) ? src[e] : 0xFFFFFFFFu;
  }
  __syncthreads();
  if (!live) return;
  const int wd = i >> 4, sh = 2 * (i & 15);
  auto code = [&](int m) { return (w[half][m * kTileWords + wd] >> sh) & 3u; };
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    uint32_t o[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      uint32_t word = 0u;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = 16 * s + 2 * (4 * h + q);
        const uint32_t c = code(m) | (code(m + 1) << 2);
        word |= (64u * (uint32_t)h + 4u * c) << (8 * q);
      }
      o[s] = word;
    }
    selb[(gid * kTileS + i) * 2 + h] = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// The SYRK over the byte-code layout: per lookup one byte extract and two LDS loads
// whose table (hi, lo), stage row and buffer are immediate offsets (the group loop is
// unrolled by two so the buffer is a constant); 2 % faster than 4-bit codes with the
// lane's table offset added per lookup (measured at N=5,000 x M=600,000).
// kGrpBlk: the FP32 block is a whole number of groups, so promotions happen only at
// group ends and the four stages of a group form one branch-free block of code.
template <bool kLoLo, bool kGrpBlk>
__global__ __launch_bounds__(256, 2) void k_syrk_bf16(
    const uint4 *__restrict__ selb, int64_t G64, int64_t M, const uint32_t *__restrict__ levbf,
    int32_t P, int64_t chunk, int32_t blk, double *__restrict__ part) {
  // [buf][hi/lo][entry]; the lo table sits 2052 bytes after the hi table, an offset
  // no paired LDS load can span, so each lookup lands in its operand register
  __shared__ uint32_t tab[2][2][kTabWords + 1];
  const int32_t pair = blockIdx.x % P;
  const int32_t split = blockIdx.x / P;
  int32_t ti, tj;
  pair_to_tiles(pair, ti, tj);
  const int64_t m_begin = (int64_t)split * chunk;  // a multiple of kGrp
  const int64_t m_end = min(M, m_begin + chunk);
  const int32_t ng = (int32_t)((m_end - m_begin + kGrp - 1) / kGrp);
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int wr = w >> 1, wc = w & 1;
  const int h = lane >> 5;
  const int64_t g0 = m_begin / kGrp;
  const uint4 *src = selb + (((int64_t)ti * G64 + g0) * kTileS + 64 * wr + (lane & 31)) * 2 + h;
  const int64_t offB = ((int64_t)(tj - ti) * G64 * kTileS + 64 * (wc - wr)) * 2;
  const int64_t roff[4] = {0, 64, offB, offB + 64};
  auto lev_fetch = [&](int32_t gi, uint32_t (&lv)[4]) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = t + 256 * u;
      const int s = e >> 7, q = (e >> 5) & 3, hh = (e >> 4) & 1, c = e & 15;
      const int64_t m0 = m_begin + (int64_t)gi * kGrp + 16 * s + 2 * (4 * hh + q);
      const uint32_t e0 = levbf[(m0 < m_end ? m0 : m_begin) * 4 + (c & 3)];
      const uint32_t e1 = levbf[(m0 + 1 < m_end ? m0 + 1 : m_begin) * 4 + (c >> 2)];
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
  auto promote = [&]() {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          dacc[a][b][r] += (double)acc[a][b][r];
          acc[a][b][r] = 0.f;
        }
  };
  const int32_t kper = kGrpBlk ? blk / kGrp : blk / 16;
  int32_t left = kper;
  uint32_t lv[4];
  uint4 cur[4], nxt[4];
  if (ng > 0) {
    lev_fetch(0, lv);
#pragma unroll
    for (int r = 0; r < 4; ++r) cur[r] = src[roff[r]];
    tab_store(0, lv);
  }
  __syncthreads();
  const char *tbase = (const char *)&tab[0][0][0];
  constexpr int kTabBytes = (kTabWords + 1) * 4;  // one table; hi then lo, per buffer
  auto group = [&](auto bufc, int32_t gi) {
    constexpr int buf = decltype(bufc)::value;
    const int32_t gn = gi + 1 < ng ? gi + 1 : gi;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s == 2) {
        lev_fetch(gn, lv);
#pragma unroll
        for (int r = 0; r < 4; ++r) nxt[r] = src[roff[r] + (int64_t)gn * (2 * kTileS)];
      }
      auto row = [&](int r, bf16x8 &hi, bf16x8 &lo) {
        const uint32_t word = s == 0 ? cur[r].x : s == 1 ? cur[r].y : s == 2 ? cur[r].z : cur[r].w;
        uint32_t hv[4], lw[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t off = (word >> (8 * q)) & 0xFFu;
          const char *pe = tbase + off + (2 * buf * kTabBytes + (s * 4 + q) * 128);
          hv[q] = *(const uint32_t *)pe;
          lw[q] = *(const uint32_t *)(pe + kTabBytes);
        }
        typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;
        u32x4 hh = {hv[0], hv[1], hv[2], hv[3]}, ll = {lw[0], lw[1], lw[2], lw[3]};
        hi = __builtin_bit_cast(bf16x8, hh);
        lo = __builtin_bit_cast(bf16x8, ll);
      };
      bf16x8 bh[2], bl[2];
      row(2, bh[0], bl[0]);
      row(3, bh[1], bl[1]);
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        bf16x8 ah, al;
        row(a, ah, al);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          if (kLoLo)
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bl[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[b], acc[a][b], 0, 0, 0);
        }
      }
      if (!kGrpBlk && --left == 0) {
        promote();
        left = kper;
      }
    }
    if (kGrpBlk && --left == 0) {
      promote();
      left = kper;
    }
    tab_store(buf ^ 1, lv);
#pragma unroll
    for (int r = 0; r < 4; ++r) cur[r] = nxt[r];
    __syncthreads();
  };
  for (int32_t gi = 0; gi < ng; gi += 2) {
    group(std::integral_constant<int, 0>{}, gi);
    if (gi + 1 < ng) group(std::integral_constant<int, 1>{}, gi + 1);
  }
  promote();
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
// trace of the summed tiles, per diagonal tile: block ti sums rows ti*128 .. ti*128+127
// (each row's splits in split order, four splits' loads in flight at once) into trp[ti]
__global__ __launch_bounds__(kTileS) void k_trace(const double *__restrict__ part, int32_t P,
                                                  int32_t nsplit, int32_t nrows,
                                                  double *__restrict__ trp) {
  __shared__ double red[kTileS];
  const int32_t ti = blockIdx.x, r = threadIdx.x, i = ti * kTileS + r;
  const int64_t stride = (int64_t)P * (kTileS * kTileS);
  const int32_t p = ti * (ti + 1) / 2 + ti;
  const double *d = part + (int64_t)p * (kTileS * kTileS) + r * kTileS + r;
  double v = 0.0;
  if (i < nrows) {
    int32_t sp = 0;
    for (; sp + 3 < nsplit; sp += 4) {
      const double u0 = d[sp * stride], u1 = d[(sp + 1) * stride], u2 = d[(sp + 2) * stride],
                   u3 = d[(sp + 3) * stride];
      v += u0;
      v += u1;
      v += u2;
      v += u3;
    }
    for (; sp < nsplit; ++sp) v += d[sp * stride];
  }
  red[r] = v;
  __syncthreads();
  for (int o = kTileS / 2; o > 0; o >>= 1) {
    if (r < o) red[r] += red[r + o];
    __syncthreads();
  }
  if (r == 0) trp[ti] = red[0];
}

// the trace from the per-tile sums trp[0 .. nt-1], in a fixed order (the same bits in
// every block); blocks of 256 threads
__device__ inline double trace_total(const double *__restrict__ trp, int32_t nt, double *red) {
  double v = 0.0;
  for (int32_t t = threadIdx.x; t < nt; t += 256) v += trp[t];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  return red[0];
}

// full symmetric matrix, scaled by 1/y with y = trace/(nrows-1): one block per tile
// pair sums its splits in split order (coalesced tile reads) and writes the tile
// and, through LDS in 32-row slabs, its transpose.  Every block forms the trace from
// the per-tile sums; block 0 also stores it in tr[0] for the host.
__global__ __launch_bounds__(256) void k_unpack_scale(const double *__restrict__ part, int32_t P,
                                                      int32_t nsplit, int32_t nrows,
                                                      double *__restrict__ tr,
                                                      double *__restrict__ xtx) {
  __shared__ double sl[32][kTileS + 1];
  __shared__ double red[256];
  const int32_t p = blockIdx.x;
  int32_t ti, tj;
  pair_to_tiles(p, ti, tj);
  const double trace = trace_total(tr + 1, (nrows + kTileS - 1) / kTileS, red);
  if (p == 0 && threadIdx.x == 0) tr[0] = trace;
  const double inv = 1.0 / (trace / (double)(nrows - 1));
  const int64_t n = nrows;
  const int t = threadIdx.x;
  for (int r0 = 0; r0 < kTileS; r0 += 32) {
    // slab rows r0..r0+31 of the tile: thread t covers columns t & 127, rows t >> 7 + 2k
    __syncthreads();
    // the splits are summed in split order, two splits' 32 loads in flight at once
    const double *base = part + (int64_t)p * (kTileS * kTileS) + (r0 + (t >> 7)) * kTileS + (t & 127);
    const int64_t stride = (int64_t)P * (kTileS * kTileS);
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = 0.0;
    int32_t sp = 0;
    for (; sp + 1 < nsplit; sp += 2) {
      const double *b0 = base + sp * stride, *b1 = b0 + stride;
      double u0[16], u1[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        u0[k] = b0[2 * k * kTileS];
        u1[k] = b1[2 * k * kTileS];
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        v[k] += u0[k];
        v[k] += u1[k];
      }
    }
    if (sp < nsplit) {
      const double *b0 = base + sp * stride;
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] += b0[2 * k * kTileS];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int rr = (t >> 7) + 2 * k, c = t & 127;
      const double val = v[k] * inv;
      sl[rr][c] = val;
      const int64_t gi = (int64_t)ti * kTileS + r0 + rr, gj = (int64_t)tj * kTileS + c;
      // lower triangle only: the upper is the transpose of it, so the matrix is
      // exactly symmetric (a diagonal tile's upper half is not used)
      if (gi < n && gj <= gi) xtx[gi * n + gj] = val;
    }
    __syncthreads();
    // transpose: row (tile column c) of the slab's 32 values -> xtx[tj*128 + c][ti*128 + r0 ..]
    for (int k = 0; k < 16; ++k) {
      const int idx = t + 256 * k, c = idx >> 5, rr = idx & 31;
      const int64_t gi = (int64_t)tj * kTileS + c, gj = (int64_t)ti * kTileS + r0 + rr;
      if (gj < n && gi < gj) xtx[gi * n + gj] = sl[rr][c];
    }
  }
}

// the per-64-SNP code layout of the current selection (built once per selection)
int eig_ensure_selg(eig_ctx *ctx) {
  if (ctx->have_selg) return EIG_OK;
  const int64_t M = ctx->nsnp;
  const int64_t G64 = (M + kGrp - 1) / kGrp;
  const int64_t ngroups = (int64_t)ctx->nst * G64;
  EIG_TRY(eig_ensure(ctx, ctx->selg, (size_t)ngroups * kTileS * 16));
  eig_time_begin(ctx);
  hipLaunchKernelGGL(k_regroup, dim3((unsigned)((ngroups + 1) / 2)), dim3(256), 0, ctx->stream,
                     (const uint32_t *)ctx->sel.p, M, G64, ngroups, (uint4 *)ctx->selg.p);
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "regroup");
  ctx->have_selg = true;
  return EIG_OK;
}

static int eig_ensure_selb(eig_ctx *ctx) {
  if (ctx->have_selb) return EIG_OK;
  const int64_t M = ctx->nsnp;
  const int64_t G64 = (M + kGrp - 1) / kGrp;
  const int64_t ngroups = (int64_t)ctx->nst * G64;
  EIG_TRY(eig_ensure(ctx, ctx->selb, (size_t)ngroups * kTileS * 32));
  eig_time_begin(ctx);
  hipLaunchKernelGGL(k_regroup8, dim3((unsigned)((ngroups + 1) / 2)), dim3(256), 0, ctx->stream,
                     (const uint32_t *)ctx->sel.p, M, G64, ngroups, (uint4 *)ctx->selb.p);
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "regroup");
  ctx->have_selb = true;
  return EIG_OK;
}

int eig_run_syrk(eig_ctx *ctx, int form, int bf16_block, double *trace_scale) {
  const int64_t M = ctx->nsnp;
  const int32_t T = ctx->nst;
  const int32_t P = T * (T + 1) / 2;
  const int32_t n = ctx->nrows;
  const bool split = (form == EIG_FORM_BF16X3 || form == EIG_FORM_BF16X4);
  if (split) {
    if (bf16_block <= 0) bf16_block = 4096;
    if (bf16_block % 16 != 0) return eig_set_err(ctx, EIG_E_ARG, "bf16_block % 16 != 0");
  }
  // Split the SNP range (split-K).  Every block of one launch runs equally long, so
  // the launch takes ceil(P*s / slots) rounds of M/s SNPs each, slots being the
  // blocks the chip holds at once: choose s to minimise that, plus a small charge
  // per split for the partial tiles it writes and the finaliser reads.  Chunks are
  // multiples of the FP32 block (bf16) or of 256 SNPs and hold at least 2048 SNPs.
  const int64_t align = split ? std::lcm<int64_t>(bf16_block, kGrp) : 256;
  const bool grpblk = split && bf16_block % kGrp == 0;
  typedef void (*bf16_kernel)(const uint4 *, int64_t, int64_t, const uint32_t *, int32_t, int64_t,
                              int32_t, double *);
  bf16_kernel kb = nullptr;
  if (form == EIG_FORM_BF16X3)
    kb = grpblk ? k_syrk_bf16<false, true> : k_syrk_bf16<false, false>;
  else if (form == EIG_FORM_BF16X4)
    kb = grpblk ? k_syrk_bf16<true, true> : k_syrk_bf16<true, false>;
  int occ = 0;
  if (form == EIG_FORM_FP64)
    EIG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_syrk_f64, 256, 0));
  else
    EIG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kb, 256, 0));
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
  // after the partial tiles: the trace, then the T per-tile diagonal sums
  EIG_TRY(eig_ensure(ctx, ctx->part, sizeof(double) * (tile * P * nsplit + 64 + T)));
  EIG_TRY(eig_ensure(ctx, ctx->xtx, sizeof(double) * (size_t)n * n));
  double *d_tr = (double *)ctx->part.p + tile * P * nsplit;
  unsigned grid = (unsigned)(P * nsplit);
  const int64_t G64 = (M + kGrp - 1) / kGrp;
  if (split) EIG_TRY(eig_ensure_selb(ctx));
  eig_time_begin(ctx);
  if (form == EIG_FORM_FP64) {
    hipLaunchKernelGGL(k_syrk_f64, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint32_t *)ctx->sel.p, M, (const double *)ctx->lev.p, P, chunk,
                       (double *)ctx->part.p);
  } else {
    hipLaunchKernelGGL(kb, dim3(grid), dim3(256), 0, ctx->stream, (const uint4 *)ctx->selb.p,
                       G64, M, (const uint32_t *)ctx->levbf.p, P, chunk, bf16_block,
                       (double *)ctx->part.p);
  }
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "syrk");
  eig_time_begin(ctx);
  hipLaunchKernelGGL(k_trace, dim3((unsigned)T), dim3(kTileS), 0, ctx->stream,
                     (const double *)ctx->part.p, P, nsplit, n, d_tr + 1);
  hipLaunchKernelGGL(k_unpack_scale, dim3((unsigned)P), dim3(256), 0, ctx->stream,
                     (const double *)ctx->part.p, P, nsplit, n, d_tr, (double *)ctx->xtx.p);
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
  // the N x N matrix is limited as the reference's (smartpca.c:984-988); fastmode is
  // not, so the selection itself may be larger
  if (ctx->nrows >= 46341)
    return eig_set_err(ctx, EIG_E_ARG, "size of GRM matrix too large: nrows: %d", ctx->nrows);
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
