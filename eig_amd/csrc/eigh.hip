// FP64 symmetric eigensolver on the device.
//
//  1. Householder tridiagonalisation, blocked by panels of kNB columns (the
//     dsytrd/dlatrd scheme).  The working matrix is row-major and symmetric with
//     a padded leading dimension, so "column c below the diagonal" is row c right
//     of the diagonal and every access is a contiguous 16-byte pair.  Each column c
//     (panel position j) runs in two parts:
//       symv : the reflector scalars of column c (dlarfg), then y = A_trail v
//              with v rebuilt on the fly from the unscaled column x; per-block
//              partial y.v;
//       wcol : w = tau (y - V q - W p) + alpha v, where alpha = -tau/2 w.v is
//              known up front from w.v = tau (y.v - 2 p.q); in the same pass the
//              next row x = A[c+1] - (V W^T + W V^T)[c+1] and per-block partials
//              of its norm and of the next column's panel dots V^T x, W^T x.
//     Trailing orders of 1,000 and more run the parts as two kernels (k_symv,
//     k_wcol); shorter ones run them in one launch, k_column, whose w-column
//     blocks load what they can up front and wait on a ticket counter for the
//     symv blocks.
//     Cross-block sums are per-block partials reduced in a fixed order (same
//     bits in every block, independent of scheduling).  After each panel
//     k_syr2k applies A -= V W^T + W V^T to the trailing matrix.  The reflector of
//     row c is left in A[c][c+1:], with tau[c].  The launch sequence for one n is
//     captured once into a hipGraph and replayed.
//  2. All eigenvalues of the tridiagonal T: Sturm counts on a shared grid of the
//     Gershgorin interval, then 9-way multisection (8 lanes per eigenvalue, one
//     count each).
//  3. The top nvec eigenvectors of T by inverse iteration (LU with partial
//     pivoting), Gram-Schmidt inside clusters of close eigenvalues.
//  4. Back-transform X = Q Y in compact-WY form Q_p = I - V_p T_p V_p^T, last
//     panel first; all T_p are formed up front (the Gram matrices V_p^T V_p over
//     row ranges, then T_p per panel), then per panel one kernel applies the
//     previous panel and writes per-block dots V_p^T X, and a kernel with one block
//     per vector sums them in block order into Z_p = T_p V_p^T X.
//  Wave-level sums exchange lanes by DPP and permlane swaps (xlane), not through LDS.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "eig_internal.h"

constexpr int kNB = 32;      // panel width
constexpr int kMaxVec = 64;  // max eigenvectors computed

enum {
  S_GL = 0, S_GU, S_PIVMIN, S_TNORM,
  S_COUNT
};

struct EigWs {
  double *a;  // n x lda working matrix (destroyed); columns n..lda are zero
  int32_t lda;
  double *d, *e, *e2, *tau;
  double *V, *W;  // n x kNB panel vectors (row-major)
  double *xtmp;   // current reflector column, unscaled (length lda, zero padded)
  double *y;
  double *sc;
  double *part1;          // per-block partial dots of the back-transform
  double *pw;             // [2][maxw][kPW] column hand-off partials (by column parity)
  double *ps;             // [maxblk] per-block y.v of the symv part
  int32_t *sync;          // [0] hand-off ticket counter, [1] hand-off timeout flag
  int32_t maxw;
  double *eval;           // ascending
  double *X;              // nvec x n eigenvectors
  double *lu;             // n x 5 x kMaxVec LU factors of inverse iteration
  double *Z;              // kNB x kMaxVec
  double *Tall;           // panels x kNB x kNB
};

// Exchange of a double with lane (lane ^ o) of the wave, for o a power of two below 64
// that is a constant once the caller's loop is unrolled, without the LDS crossbar of
// ds_bpermute: o = 1, 2 and 8 by one DPP move (quad permutes; a rotation by 8 within a
// 16-lane row), o = 4 by two DPP moves (row shifts by 4 either way) and a select, o = 16
// and 32 by v_permlane16_swap / v_permlane32_swap of the value with itself and a
// select.  The values moved are those of __shfl_xor(v, o), so every sum keeps its bits.
__device__ __forceinline__ int xlane32(int v, int o) {
  const int lane = threadIdx.x & 63;
  switch (o) {
    case 1: return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
    case 2: return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
    case 4: {
      const int dn = __builtin_amdgcn_mov_dpp(v, 0x114, 0xF, 0xF, false);  // row_shr:4
      const int up = __builtin_amdgcn_mov_dpp(v, 0x104, 0xF, 0xF, false);  // row_shl:4
      return (lane & 4) ? dn : up;
    }
    case 8: return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, false);  // row_ror:8
    case 16: {  // rows 1, 3 of the first operand swap with rows 0, 2 of the second
      const auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
      return (int)((lane & 16) ? r[0] : r[1]);
    }
    default: {  // lanes 32-63 of the first operand swap with lanes 0-31 of the second
      const auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
      return (int)((lane & 32) ? r[0] : r[1]);
    }
  }
}
__device__ __forceinline__ double xlane(double v, int o) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)xlane32((int)(uint32_t)u, o);
  const uint32_t hi = (uint32_t)xlane32((int)(uint32_t)(u >> 32), o);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// lane 0's value in every lane (all lanes of the wave active)
__device__ __forceinline__ double lane0(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// sum over a 256-thread block; result valid in every thread.  The wave sums are trees
// over lane ^ o (lane 0 receives from lane o at each level, as with a shift down)
__device__ inline double block_sum(double v, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += xlane(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// ---- partials handed from part to part -------------------------------------
// Every cross-block sum is written as per-block partials and reduced, in block
// order, by the blocks of the next part:
//   wcol(c-1) -> pw[c & 1][b][kPW]: sum x^2 over rows >= c+2, x_{c+1} (alpha,
//                from block 0 only), and for q below column c's panel position
//                sum_{r >= c+2} V_q[r] x_r and sum W_q[r] x_r;
//   symv(c)   -> ps[b]: sum over its rows of y_r v_r.
constexpr int kPW = 2 + 2 * kNB;

// sum_{b = start, start+step, ... < end} p[b * stride], in that order, with the
// loads issued eight at a time so that their latencies overlap.  The first eight
// (load8) are straight-line and unconditional (indices past the end read entry
// end-1 and are discarded), so a caller can issue them together with its other
// loads and add them up later (sum_rest).  Needs start >= 0 and end >= 1.
__device__ inline void load8(const double *__restrict__ p, int64_t stride, int32_t start,
                             int32_t step, int32_t end, double (&v)[8]) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int32_t bb = start + u * step;
    const double x = p[(int64_t)min(bb, end - 1) * stride];
    v[u] = bb < end ? x : 0.0;
  }
}
__device__ __forceinline__ double sum8(const double (&v)[8]) {  // pairwise
  return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}
__device__ inline double sum_rest(const double *__restrict__ p, int64_t stride, int32_t start,
                                  int32_t step, int32_t end, const double (&v0)[8]) {
  double s = sum8(v0);
  for (int32_t b = start + 8 * step; b < end; b += 8 * step) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t bb = b + u * step;
      v[u] = bb < end ? p[(int64_t)bb * stride] : 0.0;
    }
    s += sum8(v);
  }
  return s;
}
__device__ inline double sum_strided(const double *__restrict__ p, int64_t stride, int32_t start,
                                     int32_t step, int32_t end) {
  if (end <= 0) return 0.0;
  double v[8];
  load8(p, stride, start, step, end, v);
  return sum_rest(p, stride, start, step, end, v);
}

constexpr int kWRows = 64;  // rows per w-column block (four threads per row)
__device__ inline int32_t wcol_blocks(int32_t n, int32_t c) {  // producer blocks of column c
  return c == 0 ? 1 : (n - c + kWRows - 1) / kWRows;
}

// dlarfg scalars of column c from the producers' partials, reduced by one wave in
// a fixed order (the same bits in every wave that calls it).  col_scalars_sums
// takes each lane's strided sums of the partials (sum over b = lane, lane+64, ...
// of pw[b][0] and pw[b][1]).
__device__ inline void col_scalars_sums(double s0, double s1, double &tau, double &scale,
                                        double &beta) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s0 += xlane(s0, o);
    s1 += xlane(s1, o);
  }
  s0 = lane0(s0);
  s1 = lane0(s1);
  const double alpha = s1, xn2 = s0;
  tau = 0.0;
  scale = 0.0;
  beta = alpha;
  if (xn2 > 0.0) {
    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    tau = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
}
__device__ inline void col_scalars(const EigWs &ws, int32_t n, int32_t c, double &tau,
                                   double &scale, double &beta) {
  const double *pw = ws.pw + (int64_t)(c & 1) * ws.maxw * kPW;
  const int32_t nw = wcol_blocks(n, c);
  const int lane = threadIdx.x & 63;
  col_scalars_sums(sum_strided(pw, kPW, lane, 64, nw), sum_strided(pw + 1, kPW, lane, 64, nw),
                   tau, scale, beta);
}

// column 0: x = row 0, d[0], and the partials of column 0 (one block)
__global__ __launch_bounds__(256) void k_col0(EigWs ws, int32_t n) {
  __shared__ double red[4];
  double s = 0.0;
  for (int32_t r = threadIdx.x; r < n; r += 256) {
    const double x = ws.a[r];
    ws.xtmp[r] = x;
    if (r >= 2) s += x * x;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) {
    ws.d[0] = ws.a[0];
    ws.pw[0] = s;
    ws.pw[1] = n > 1 ? ws.a[1] : 0.0;
    ws.sync[0] = 0;
    ws.sync[1] = 0;
  }
}

// a store that is coherent device-wide once the wave's vmcnt has drained (it
// writes through the XCD's L2), for the values the wcol part of the same launch
// reads (y and the y.v partials)
__device__ inline void store_dev(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- symv part
constexpr int kSymvRows = 8;      // symv rows per block in k_column
constexpr int kSymvRowsLong = 4;  // in k_symv (long columns): 106 VGPRs, 4 waves/SIMD;
                                  // 2 ms less than 8 rows at n = 5,000 (measured)
template <int R>
__device__ inline void symv_part(const EigWs &ws, int32_t n, int32_t c, int32_t j, int32_t blk) {
  __shared__ double red[4][R];
  __shared__ double sscale;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int32_t lda = ws.lda;
  const int32_t r0 = c + 1 + blk * R;
  // the block's own entries of x, for its rows of v at the end (loaded now, so no
  // memory round trip follows the stream)
  const double xown = (threadIdx.x < R && r0 + (int32_t)threadIdx.x < n) ? ws.xtmp[r0 + threadIdx.x] : 0.0;
  const int32_t a0 = (c + 1) & ~1;
  const int32_t npair = (((n + 1) & ~1) - a0) >> 1;
  const double2 *xp = (const double2 *)(ws.xtmp + a0);
  // rows past n read row r0 instead (in bounds, result discarded), so every load
  // in the loop is unconditional and all of them are in flight together
  const double2 *rp[R];
#pragma unroll
  for (int rr = 0; rr < R; ++rr)
    rp[rr] = (const double2 *)(ws.a + (int64_t)((r0 + rr) < n ? r0 + rr : r0) * lda + a0);
  double acc[R];
#pragma unroll
  for (int rr = 0; rr < R; ++rr) acc[rr] = 0.0;
  // the first pairs' loads go out before the reflector scalars are reduced
  int32_t p = threadIdx.x;
  double2 x0 = make_double2(0.0, 0.0), x1 = x0, a0v[R], a1v[R];
  const bool two = p + 256 < npair;
  if (two) {
    x0 = xp[p];
    x1 = xp[p + 256];
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      a0v[rr] = rp[rr][p];
      a1v[rr] = rp[rr][p + 256];
    }
  }
  if (wv == 0) {
    double tau, scale, beta;
    col_scalars(ws, n, c, tau, scale, beta);
    if (lane == 0) {
      sscale = scale;
      if (blk == 0) {
        ws.e[c] = beta;
        ws.e2[c] = beta * beta;
        ws.tau[c] = tau;
      }
    }
  }
  __syncthreads();
  const double scale = sscale;
  // v: 0 up to column c, 1 at c+1, x*scale beyond (x is zero past n); only the
  // first pair (p = 0) can hold columns <= c+1
  auto vpair = [&](const double2 xx, int32_t pp, double &v0, double &v1) {
    v0 = xx.x * scale;
    v1 = xx.y * scale;
    if (pp == 0) {
      const int32_t s0 = a0;
      v0 = (s0 <= c) ? 0.0 : ((s0 == c + 1) ? 1.0 : v0);
      v1 = (s0 + 1 == c + 1) ? 1.0 : v1;
    }
  };
  for (bool have = two; have;) {
    double v0, v1, w0, w1;
    vpair(x0, p, v0, v1);
    vpair(x1, p + 256, w0, w1);
#pragma unroll
    for (int rr = 0; rr < R; ++rr)
      acc[rr] += (a0v[rr].x * v0 + a0v[rr].y * v1) + (a1v[rr].x * w0 + a1v[rr].y * w1);
    p += 512;
    have = p + 256 < npair;
    if (have) {
      x0 = xp[p];
      x1 = xp[p + 256];
#pragma unroll
      for (int rr = 0; rr < R; ++rr) {
        a0v[rr] = rp[rr][p];
        a1v[rr] = rp[rr][p + 256];
      }
    }
  }
  if (p < npair) {
    const double2 x0 = xp[p];
    double2 a0v[R];
#pragma unroll
    for (int rr = 0; rr < R; ++rr) a0v[rr] = rp[rr][p];
    double v0, v1;
    vpair(x0, p, v0, v1);
#pragma unroll
    for (int rr = 0; rr < R; ++rr) acc[rr] += a0v[rr].x * v0 + a0v[rr].y * v1;
  }
  // butterfly transpose-reduce of the R row sums: after the exchanges at offsets
  // R/2, ..., 1 lane l holds row (l & (R-1)) summed over its group of R lanes; the
  // exchanges at R, 2R, ..., 32 sum the groups
  static_assert(R == 4 || R == 8, "R rows per block: 4 or 8");
#pragma unroll
  for (int o = R / 2; o > 0; o >>= 1) {
    const bool b = (lane & o) != 0;
#pragma unroll
    for (int k = 0; k < o; ++k) acc[k] = (b ? acc[k + o] : acc[k]) + xlane(b ? acc[k] : acc[k + o], o);
  }
  {
    double a = acc[0];
#pragma unroll
    for (int o = R; o < 64; o <<= 1) a += xlane(a, o);
    if (lane < R) red[wv][lane] = a;
  }
  __syncthreads();
  // per row (lanes 0 .. R-1 of wave 0): y, v into V[j] and A[c], and the row's y.v;
  // the block's y.v partial is their sum over the wave (a tree of R/2 .. 1 exchanges)
  if (wv == 0) {
    double val = 0.0;
    if (lane < R) {
      const int rr = lane;
      const int32_t r = r0 + rr;
      if (r < n) {
        const double vr = (r == c + 1) ? 1.0 : xown * scale;
        const double yr = (red[0][rr] + red[1][rr]) + (red[2][rr] + red[3][rr]);
        val = yr * vr;
        store_dev(ws.y + r, yr);
        ws.V[(int64_t)r * kNB + j] = vr;
        ws.a[(int64_t)c * lda + r] = vr;
      }
    }
#pragma unroll
    for (int o = R / 2; o > 0; o >>= 1) val += xlane(val, o);
    if (lane == 0) store_dev(ws.ps + blk, val);
  }
}

// ---------------------------------------------------------------- wcol part
// Panel vectors are stored row-major, V[r * kNB + q], so a thread loads its row's
// coefficients in one batch; entries of columns not yet in this panel are stale
// but always meet a zero coefficient.
// kEarly (a launch after the symv part's): y and the y.v partials are final before
// the launch starts, so their loads go out with the others and the part runs on one
// round of memory latency instead of two.
template <bool kEarly, class Wait>
__device__ inline void wcol_part(const EigWs &ws, int32_t n, int32_t c, int32_t j, int32_t blk,
                                 int32_t nps, Wait &&wait) {
  __shared__ double red[4][4];
  __shared__ double pqr[4][2 * kNB];  // raw panel dots per wave
  __shared__ double pq[2 * kNB];      // p_q = V_q . v, q_q = W_q . v (zero for q >= j)
  __shared__ double cc[2 * kNB];      // V_q[c+1], W_q[c+1] (zero for q >= j)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int sub = threadIdx.x & 3;  // this thread's quarter of the panel: q = 8 sub .. 8 sub + 7
  const int32_t c1 = c + 1;
  const int32_t lda = ws.lda;
  const double *pwc = ws.pw + (int64_t)(c & 1) * ws.maxw * kPW;
  const int32_t nw = wcol_blocks(n, c);
  // ---- loads of data from earlier launches go out before the hand-off wait: this
  // row's quarter of the panel vectors (column j, written by this launch's symv part,
  // is rebuilt from x below) and its entry of the trailing row c+1, its x, row c+1's
  // panel coefficients, and the panel-dot and norm partials
  const int32_t r = c1 + blk * kWRows + (int32_t)(threadIdx.x >> 2);
  const bool inr = r < n;
  const int32_t rs = inr ? r : c1;  // rows past n read row c+1 (in bounds, discarded)
  // first batches of the three partial sums: panel dots (component lane over blocks
  // wv, wv+4, ..., three batches of eight), and the norm and alpha of column c
  // (blocks lane, lane+64, ...)
  double lp[3][8], ln0[8], ln1[8];
#pragma unroll
  for (int t = 0; t < 3; ++t) load8(pwc + 2 + lane, kPW, wv + 32 * t, 4, nw, lp[t]);
  load8(pwc, kPW, lane, 64, nw, ln0);
  load8(pwc + 1, kPW, lane, 64, nw, ln1);
  double vrow[8], wrow[8];
  {
    const double2 *vp = (const double2 *)(ws.V + (int64_t)rs * kNB + 8 * sub);
    const double2 *wp = (const double2 *)(ws.W + (int64_t)rs * kNB + 8 * sub);
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const double2 a2 = vp[h], b2 = wp[h];
      vrow[2 * h] = a2.x;
      vrow[2 * h + 1] = a2.y;
      wrow[2 * h] = b2.x;
      wrow[2 * h + 1] = b2.y;
    }
  }
  double base = ws.a[(int64_t)c1 * lda + rs];
  const double xr = ws.xtmp[rs];
  double atc = 0.0;
  if (threadIdx.x < 2 * kNB) {
    const int q = threadIdx.x & (kNB - 1);
    atc = (threadIdx.x < kNB) ? ws.V[(int64_t)c1 * kNB + q] : ws.W[(int64_t)c1 * kNB + q];
  }
  double ye = 0.0, yc1e = 0.0, ype[8];
  if constexpr (kEarly) {
    ye = ws.y[rs];
    yc1e = ws.y[c1];
    load8(ws.ps, 1, threadIdx.x, 256, nps, ype);
  }
  double praw = 0.0;
  {
    const int comp = lane;  // 0..31 V dots, 32..63 W dots
    const int q = comp & (kNB - 1);
    // pairwise, so the chain is five additions long instead of 24
    double st[3];
#pragma unroll
    for (int t = 0; t < 3; ++t)
      st[t] = ((lp[t][0] + lp[t][1]) + (lp[t][2] + lp[t][3])) +
              ((lp[t][4] + lp[t][5]) + (lp[t][6] + lp[t][7]));
    praw = (st[0] + st[1]) + st[2];
    for (int32_t bb = wv + 96; bb < nw; bb += 4) praw += pwc[2 + comp + (int64_t)bb * kPW];
    praw = q < j ? praw : 0.0;
  }
  // every wave forms the scalars itself (same bits as the symv part's)
  double tau, scale;
  {
    double beta;
    col_scalars_sums(sum_rest(pwc, kPW, lane, 64, nw, ln0), sum_rest(pwc + 1, kPW, lane, 64, nw, ln1),
                     tau, scale, beta);
  }
  if (!wait()) return;
  // ---- the symv part's outputs: y and the per-block partials of y.v
  double yy = kEarly ? ye : ws.y[rs];
  const double yc1 = kEarly ? yc1e : ws.y[c1];
  double yv = kEarly ? sum_rest(ws.ps, 1, threadIdx.x, 256, nps, ype)
                     : sum_strided(ws.ps, 1, threadIdx.x, 256, nps);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) yv += xlane(yv, o);
  if (lane == 0) red[wv][0] = yv;
  pqr[wv][lane] = praw;
  __syncthreads();
  const double ydv = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
  if (threadIdx.x < 2 * kNB) {
    const int comp = threadIdx.x, q = comp & (kNB - 1);
    const double raw = (pqr[0][comp] + pqr[1][comp]) + (pqr[2][comp] + pqr[3][comp]);
    cc[comp] = q < j ? atc : 0.0;
    pq[comp] = q < j ? atc + scale * raw : 0.0;  // v = e_{c+1} + scale x beyond c+1
  }
  __syncthreads();
  // alpha = -tau/2 w.v with w.v = tau (y.v - 2 sum_q p_q q_q), and w_pre at row c+1
  // (same in every block).  pq and cc are zero from q = j on, so the sums run over
  // the whole panel.  Lane q (mod 32) forms term q and five exchanges (lane ^ 16 ..
  // lane ^ 1) add the terms as a tree: the same bits in every lane, wave and block.
  double spq, y1;
  {
    const int q = lane & (kNB - 1);
    double tq = pq[q] * pq[kNB + q];
    double ty = cc[q] * pq[kNB + q] + cc[kNB + q] * pq[q];
#pragma unroll
    for (int o = kNB / 2; o > 0; o >>= 1) {
      tq += xlane(tq, o);
      ty += xlane(ty, o);
    }
    spq = tq;
    y1 = yc1 - ty;
  }
  const double alpha = -0.5 * tau * (tau * (ydv - 2.0 * spq));
  const double wc1 = tau * y1;
  // the row's panel corrections: each thread sums its quarter, two exchanges add the
  // quarters (the same bits in all four threads of the row)
  {
    double py[8], pb[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int q = 8 * sub + k;
      py[k] = vrow[k] * pq[kNB + q] + wrow[k] * pq[q];
      pb[k] = cc[q] * wrow[k] + cc[kNB + q] * vrow[k];
    }
    // pairwise: three dependent additions instead of eight
    double sy = ((py[0] + py[1]) + (py[2] + py[3])) + ((py[4] + py[5]) + (py[6] + py[7]));
    double sbse = ((pb[0] + pb[1]) + (pb[2] + pb[3])) + ((pb[4] + pb[5]) + (pb[6] + pb[7]));
    sy += xlane(sy, 1);
    sbse += xlane(sbse, 1);
    sy += xlane(sy, 2);
    sbse += xlane(sbse, 2);
    yy -= sy;
    base -= sbse;
  }
  // this column's v_r, as the symv part formed it, joins vrow at j; the row's
  // W[r][j] joins wrow below
  const double vr = (r == c1) ? 1.0 : xr * scale;
#pragma unroll
  for (int k = 0; k < 8; ++k) vrow[k] = (8 * sub + k == j) ? vr : vrow[k];
  const double w = tau * yy;
  const double wj = w + alpha * vr;
  // row c+1 minus panel column j:  V[j][c1] (=1) W[j][r] + W[j][c1] V[j][r]
  const double x = base - w - wc1 * vr - 2.0 * alpha * vr;
  const bool live = inr && r >= c1 + 2;
  if (inr && sub == 0) {
    ws.W[(int64_t)r * kNB + j] = wj;
    ws.xtmp[r] = x;
    if (r == c1) ws.d[c1] = x;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) wrow[k] = (8 * sub + k == j) ? wj : wrow[k];
  // ---- this block's partials for column c+1, summed in a fixed order:
  // norm, the next alpha (row c+2, block 0), and if column c+1 continues the
  // panel, V_q . x and W_q . x for q <= j
  const int jn = ((c1 % kNB) != 0 && c1 + 1 < n) ? j + 1 : 0;
  const double xl = live ? x : 0.0;
  double *pwn = ws.pw + (int64_t)(c1 & 1) * ws.maxw * kPW + (int64_t)blk * kPW;
  {
    double v0 = sub == 0 ? xl * xl : 0.0, v1 = (sub == 0 && inr && r == c1 + 1) ? x : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      v0 += xlane(v0, o);
      v1 += xlane(v1, o);
    }
    if (lane == 0) {
      red[wv][1] = v0;
      red[wv][2] = v1;
    }
  }
  // panel dots V_q . x, W_q . x: a thread holds 16 products (its quarter of V, then
  // of W).  Butterfly transpose-reduce over the wave's 16 rows (lane bits 5..2): the
  // exchange at offset o keeps the half selected by bit o, so lane (i, sub) ends
  // with product i (i = (lane >> 2) & 15) summed over the rows; then one pass over
  // the four waves through LDS.
  {
    double t[16];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      t[k] = vrow[k] * xl;
      t[8 + k] = wrow[k] * xl;
    }
#pragma unroll
    for (int off = 32; off >= 4; off >>= 1) {
      const int half = off >> 2;
      const bool bsel = (lane & off) != 0;
#pragma unroll
      for (int k = 0; k < half; ++k)
        t[k] = (bsel ? t[k + half] : t[k]) + xlane(bsel ? t[k] : t[k + half], off);
    }
    const int i = (lane >> 2) & 15;
    const int comp = (i < 8) ? 8 * sub + i : kNB + 8 * sub + (i - 8);
    pqr[wv][comp] = t[0];
  }
  __syncthreads();
  if (threadIdx.x < 2 * kNB) {
    const int comp = threadIdx.x;
    const double sum = (pqr[0][comp] + pqr[1][comp]) + (pqr[2][comp] + pqr[3][comp]);
    pwn[2 + comp] = (comp & (kNB - 1)) < jn ? sum : 0.0;
  }
  if (threadIdx.x == 0) {
    pwn[0] = (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]);
    pwn[1] = (red[0][2] + red[1][2]) + (red[2][2] + red[3][2]);
  }
}

// The two parts as kernels of their own, for long columns: there the symv part
// streams a large trailing matrix and a launch boundary costs less than holding
// w-column blocks and draining device-coherent stores inside one launch.
template <int R>
__global__ __launch_bounds__(256) void k_symv(EigWs ws, int32_t n, int32_t c, int32_t j) {
  symv_part<R>(ws, n, c, j, (int32_t)blockIdx.x);
}
__global__ __launch_bounds__(256) void k_wcol(EigWs ws, int32_t n, int32_t c, int32_t j, int32_t nps) {
  wcol_part<true>(ws, n, c, j, (int32_t)blockIdx.x, nps, []() { return true; });
}

// ---------------------------------------------------------------- k_column
// One launch per column with gw + gsym blocks (gw = ceil((n-c-1)/64) w-column
// chunks, gsym = ceil((n-c-1)/8) symv row groups).  Blocks gw .. gw+gsym-1 run the
// symv part for their 8 rows and then add one to a ticket counter once their y and
// ps stores (device-coherent stores) have completed.  Blocks 0 .. gw-1 run the
// w-column part for one 64-row chunk each: they issue their loads of data from
// earlier launches at once, then wait until all gsym tickets are in, then read y
// and the y.v partials.  The waiters never hold up a symv block (k_column runs for
// trailing orders below kMergeBelow, so there are at most 16 of them, far fewer than
// the blocks the device holds at once), so the launch cannot deadlock; the wait is
// bounded anyway, and a timeout sets sync[1] for the host to report.  The waiters
// first sleep for quiet_ticks of the device's wall clock (most of the symv part's
// expected duration), then poll.  The counter is cleared by k_col0 and grows
// monotonically over the columns: column c's tickets are base .. base + gsym - 1.
// Ordering rests on the ISA, not on a release: the y and partial stores write through
// the XCD's L2 (sc1) and are drained (vmcnt 0) before the ticket is added, and the
// waiter polls with sc1 loads and runs an agent-scope acquire (L2 invalidate) before
// reading them.  A release would write back the whole L2 once per block.
constexpr int64_t kHandoffSpin = 1 << 22;
__global__ __launch_bounds__(256) void k_column(EigWs ws, int32_t n, int32_t c, int32_t j,
                                                int32_t base, int64_t quiet_ticks) {
  __shared__ int32_t ok_sh;
  const int32_t gw = (n - c - 1 + kWRows - 1) / kWRows;
  const int32_t gsym = (int32_t)gridDim.x - gw;
  if ((int32_t)blockIdx.x >= gw) {
    symv_part<kSymvRows>(ws, n, c, j, (int32_t)blockIdx.x - gw);
    // each wave drains its own device-coherent stores, then one thread adds the ticket
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_fetch_add(ws.sync, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const uint64_t t_start = wall_clock64();
  auto wait = [&]() -> bool {
    if (threadIdx.x == 0) {
      // no polling for most of the symv part's expected time: polls are device-
      // coherent loads and slow down the symv blocks' stream of the matrix
      while ((int64_t)(wall_clock64() - t_start) < quiet_ticks) __builtin_amdgcn_s_sleep(32);
      int32_t ok = 0;
      for (int64_t s = 0; s < kHandoffSpin; ++s) {
        if (__hip_atomic_load(ws.sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= base + gsym) {
          ok = 1;
          break;
        }
        if (__hip_atomic_load(ws.sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        __builtin_amdgcn_s_sleep(1);
      }
      if (!ok) __hip_atomic_store(ws.sync + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ok_sh = ok;
    }
    __syncthreads();
    if (!ok_sh) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
  };
  wcol_part<false>(ws, n, c, j, (int32_t)blockIdx.x, gsym, wait);
}

// A[r][s] -= sum_q V[r][q] W[s][q] + W[r][q] V[s][q] for r, s >= k0: one FP64 MFMA
// GEMM with K = 2 kNB per 64x64 tile, A operand [V_r | W_r], B operand [W_s | V_s]
// (4 waves x 32x32, MFMA f64 16x16x4).
typedef __attribute__((__vector_size__(4 * sizeof(double)))) double eig_f64x4;
constexpr int kS2K = 2 * kNB;        // GEMM depth
constexpr int kS2Ld = kS2K + 1;      // padded LDS row (conflict-free 16-row reads)

// 64 consecutive rows of two row-major n x kNB panels (P0 and P1), 16 bytes per load:
// element i of a thread covers row (t + 256 i) / 16, columns 2 ((t + 256 i) % 16) + {0, 1}.
// Rows at or past n load row n - 1 and are zeroed when stored.
struct S2Rows {
  double2 p[2][4];
};
__device__ __forceinline__ void s2_load(const double *P0, const double *P1, int32_t r0,
                                        int32_t n, S2Rows &s) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int e = threadIdx.x + 256 * i;
    const int64_t off = (int64_t)min(r0 + (e >> 4), n - 1) * kNB + 2 * (e & 15);
    s.p[0][i] = *(const double2 *)(P0 + off);
    s.p[1][i] = *(const double2 *)(P1 + off);
  }
}
// row o of `sm`: P0's kNB columns, then P1's; columns at or past nb are zero
__device__ __forceinline__ void s2_store(double (*sm)[kS2Ld], int32_t r0, int32_t n,
                                         int32_t nb, const S2Rows &s) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int e = threadIdx.x + 256 * i, o = e >> 4, q = 2 * (e & 15);
    const bool ok = r0 + o < n;
    const bool ok0 = ok && q < nb, ok1 = ok && q + 1 < nb;
    sm[o][q] = ok0 ? s.p[0][i].x : 0.0;
    sm[o][q + 1] = ok1 ? s.p[0][i].y : 0.0;
    sm[o][kNB + q] = ok0 ? s.p[1][i].x : 0.0;
    sm[o][kNB + q + 1] = ok1 ? s.p[1][i].y : 0.0;
  }
}

// One block per 64x64 tile.  The B operand rows [W_c | V_c] are staged in LDS; each
// lane's A operand stays in registers: the MFMA sums over k, so lane group g = lane / 16
// takes k = 16 g + ks at step ks (in A and B alike), and its A entries are 16
// consecutive doubles of a row of V (g < 2) or W.  33 KB of LDS and ~150 registers
// give three blocks per CU.
__global__ __launch_bounds__(256) void k_syr2k(EigWs ws, int32_t n, int32_t k0, int32_t nb) {
  __shared__ double sb[64][kS2Ld];
  const int32_t nt = (n - k0 + 63) / 64;
  const int32_t tr = blockIdx.x / nt, tc = blockIdx.x % nt;
  const int32_t rb = k0 + tr * 64, cb = k0 + tc * 64;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1, g = lane >> 4;
  const int32_t lda = ws.lda;
  // the B rows and the A entries are loaded together (both from L2-resident panels)
  S2Rows rbv;
  s2_load(ws.W, ws.V, cb, n, rbv);
  double2 av2[2][8];
  {
    const double *asrc = (g < 2 ? ws.V : ws.W) + 16 * (g & 1);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      // rows past n read row n - 1 (their results are not stored)
      const int32_t r = min(rb + 32 * wr + 16 * a + (lane & 15), n - 1);
      const double2 *src = (const double2 *)(asrc + (int64_t)r * kNB);
#pragma unroll
      for (int h = 0; h < 8; ++h) av2[a][h] = src[h];
    }
  }
  __builtin_amdgcn_sched_barrier(0);  // every load goes out before the LDS stores wait
  s2_store(sb, cb, n, nb, rbv);
  // panel columns at or past nb take no part (zero in A as in B)
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      const int q = 16 * (g & 1) + 2 * h;
      av2[a][h].x = q < nb ? av2[a][h].x : 0.0;
      av2[a][h].y = q + 1 < nb ? av2[a][h].y : 0.0;
    }
  __syncthreads();
  // the tile's old values (row, column of lane's accumulator elements) load while the
  // products are formed; indices past n are clamped, and those elements are never stored
  double old[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int32_t r = min(rb + 32 * wr + 16 * a + (lane >> 4) + 4 * e, n - 1);
        const int32_t c = min(cb + 32 * wc + 16 * b + (lane & 15), n - 1);
        old[a][b][e] = ws.a[(int64_t)r * lda + c];
      }
  __builtin_amdgcn_sched_barrier(0);  // the old values are in flight before the first product
  eig_f64x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = eig_f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < 16; ++ks) {
    double bv[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) bv[b] = sb[32 * wc + 16 * b + (lane & 15)][16 * g + ks];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const double av = (ks & 1) ? av2[a][ks >> 1].y : av2[a][ks >> 1].x;
#pragma unroll
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[b], acc[a][b], 0, 0, 0);
    }
  }
  // accumulator layout: lane l, element e -> row l/16 + 4e, column l%16
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int32_t r = rb + 32 * wr + 16 * a + (lane >> 4) + 4 * e;
        const int32_t c = cb + 32 * wc + 16 * b + (lane & 15);
        if (r < n && c < n) ws.a[(int64_t)r * lda + c] = old[a][b][e] - acc[a][b][e];
      }
}

// ---------------------------------------------------------------- eigenvalues
__global__ void k_gersh(EigWs ws, int32_t n) {
  __shared__ double lo[256], hi[256], mx[256], tnr[256];
  double gl = 1e300, gu = -1e300, emax = 0.0, tn = 0.0;
  for (int32_t i = threadIdx.x; i < n; i += blockDim.x) {
    double r = 0.0;
    if (i > 0) r += fabs(ws.e[i - 1]);
    if (i < n - 1) r += fabs(ws.e[i]);
    gl = fmin(gl, ws.d[i] - r);
    gu = fmax(gu, ws.d[i] + r);
    if (i < n - 1) emax = fmax(emax, ws.e2[i]);
    tn = fmax(tn, fabs(ws.d[i]) + r);
  }
  lo[threadIdx.x] = gl;
  hi[threadIdx.x] = gu;
  mx[threadIdx.x] = emax;
  tnr[threadIdx.x] = tn;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + o]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + o]);
      mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + o]);
      tnr[threadIdx.x] = fmax(tnr[threadIdx.x], tnr[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double eps = 2.220446049250313e-16, safmin = 2.2250738585072014e-308;
    double tnorm = fmax(fabs(lo[0]), fabs(hi[0]));
    double pivmin = safmin * fmax(1.0, mx[0]);
    ws.sc[S_GL] = lo[0] - 2.0 * eps * tnorm * n - 2.0 * pivmin;
    ws.sc[S_GU] = hi[0] + 2.0 * eps * tnorm * n + 2.0 * pivmin;
    ws.sc[S_PIVMIN] = pivmin;
    ws.sc[S_TNORM] = fmax(tnr[0], 1e-300);
  }
}

// Two Sturm counts (eigenvalues of T below x and below x2), interleaved.  The
// reciprocal is v_rcp_f64 refined by two Newton steps (relative error ~1e-16),
// which perturbs each pivot far below the bisection tolerance.  d and e2 are read
// a batch of four entries ahead of use.
// number of eigenvalues of T below x (Sturm count of T - x I), with q clamped away
// from zero by pivmin as in LAPACK's dstebz
__device__ inline int32_t sturm1(const double *__restrict__ d, const double *__restrict__ e2,
                                 int32_t n, double x, double pivmin) {
  int32_t cnt = 0;
  double q = d[0] - x;
  if (q <= pivmin) { ++cnt; q = fmin(q, -pivmin); }
  auto step = [&](double di, double ei) {
    double r = __builtin_amdgcn_rcp(q);
    r = fma(r, fma(-q, r, 1.0), r);
    q = (di - x) - ei * r;
    if (q <= pivmin) { ++cnt; q = fmin(q, -pivmin); }
  };
  // batches of four steps in two alternating register sets: the next batch's
  // (wave-uniform, scalar) loads are issued before the current batch's steps, so their
  // latency hides behind the recurrence
  int32_t i = 1;
  double ad[4], af[4], bd[4], bf[4];
  auto ld = [&](int32_t j0, double (&xd)[4], double (&xf)[4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xd[u] = d[j0 + u];
      xf[u] = e2[j0 - 1 + u];
    }
  };
  auto run = [&](const double (&xd)[4], const double (&xf)[4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) step(xd[u], xf[u]);
  };
  if (i + 3 < n) {
    ld(i, ad, af);
    for (;;) {
      // set a holds steps i .. i+3 (all < n)
      if (i + 7 >= n) {
        run(ad, af);
        i += 4;
        break;
      }
      ld(i + 4, bd, bf);
      run(ad, af);
      i += 4;
      // set b holds steps i .. i+3
      if (i + 7 >= n) {
        run(bd, bf);
        i += 4;
        break;
      }
      ld(i + 4, ad, af);
      run(bd, bf);
      i += 4;
    }
  }
  for (; i < n; ++i) step(d[i], e2[i - 1]);
  return cnt;
}

// kMS lanes per eigenvalue (k-th smallest): each round the lanes evaluate kMS interior
// points of [lo, hi] and keep the sub-interval holding eigenvalue k.  One count per
// lane; at N=5,000 (after the grid) 8 lanes took 2.52 ms, 4 lanes 3.16, 16 lanes 2.94
// and 32 lanes 3.70: more lanes cut the rounds but add each round's work.
constexpr int kMS = 8;
// Shared first level: Sturm counts at kMsGrid evenly spaced interior points of the
// Gershgorin interval, one per thread (each a count over the whole tridiagonal, like
// one round of the multisection for kMsGrid / 16 eigenvalues).  Every eigenvalue then
// starts from the grid interval that brackets it, kMsGrid + 1 times narrower than the
// Gershgorin interval, which saves about four rounds.
constexpr int kMsGrid = 65536;
__device__ __forceinline__ double msgrid_x(double glo, double ghi, int32_t j) {
  return glo + (ghi - glo) / (double)(kMsGrid + 1) * (double)(j + 1);
}
__global__ __launch_bounds__(256) void k_msgrid(EigWs ws, int32_t n, int32_t *__restrict__ gcnt) {
  const int32_t j = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);  // kMsGrid threads
  gcnt[j] = sturm1(ws.d, ws.e2, n, msgrid_x(ws.sc[S_GL], ws.sc[S_GU], j), ws.sc[S_PIVMIN]);
}
// gcnt: k_msgrid's counts, or null to start from the Gershgorin interval
__global__ __launch_bounds__(256) void k_multisect(EigWs ws, int32_t n,
                                                   const int32_t *__restrict__ gcnt) {
  const int32_t k = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kMS);
  const int sub = threadIdx.x % kMS;
  const bool active = k < n;
  double lo = ws.sc[S_GL], hi = ws.sc[S_GU];
  if (gcnt != nullptr && active) {
    // the grid interval bracketing eigenvalue k: a ends as a point (or the lower bound,
    // a = -1) whose count is <= k, b as one (or the upper bound, b = kMsGrid) whose
    // count is > k, adjacent -- a valid start whatever the counts' rounding
    int32_t a = -1, b = kMsGrid;
    while (b - a > 1) {
      const int32_t m = (a + b) >> 1;
      if (gcnt[m] <= k) a = m; else b = m;
    }
    const double glo = lo, ghi = hi;
    if (a >= 0) lo = msgrid_x(glo, ghi, a);
    if (b < kMsGrid) hi = msgrid_x(glo, ghi, b);
  }
  const double pivmin = ws.sc[S_PIVMIN];
  const double eps = 2.220446049250313e-16;
  // absolute floor eps * ||T||_1, the default of LAPACK's dstebz (ABSTOL <= 0): an
  // eigenvalue at zero (the centred matrix has one) converges in ~13 rounds instead
  // of running to the round cap with a relative tolerance alone
  const double atol = eps * ws.sc[S_TNORM];
  for (int it = 0; it < 96; ++it) {
    const double tol = 2.0 * eps * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin + atol;
    const bool done = (hi - lo <= tol);
    // all kMS lanes of a group share lo/hi, so `done` is uniform within a group
    if (__all(done || !active)) break;
    const double w = (hi - lo) / (double)(kMS + 1);
    const double x = lo + w * (double)(sub + 1);
    int32_t cnt = 0;
    if (active && !done) cnt = sturm1(ws.d, ws.e2, n, x, pivmin);
    // new lo = largest point with count <= k ; new hi = smallest point with count > k
    double cand_lo = (cnt <= k) ? x : -1e300;
    double cand_hi = (cnt > k) ? x : 1e300;
    for (int o = kMS / 2; o > 0; o >>= 1) {
      cand_lo = fmax(cand_lo, __shfl_xor(cand_lo, o, kMS));
      cand_hi = fmin(cand_hi, __shfl_xor(cand_hi, o, kMS));
    }
    if (!done) {
      double nlo = fmax(lo, cand_lo), nhi = fmin(hi, cand_hi);
      if (nhi < nlo) nhi = nlo;
      lo = nlo;
      hi = nhi;
    }
  }
  if (active && sub == 0) ws.eval[k] = 0.5 * (lo + hi);
}

// ---------------------------------------------------------------- inverse iteration
// 1/a by the hardware reciprocal and two Newton steps (error a few ulp): the LU of
// T - lambda I keeps pivot reciprocals, so no division sits on a serial chain
__device__ inline double rcp_nr(double a) {
  double r = __builtin_amdgcn_rcp(a);
  r = fma(r, fma(-a, r, 1.0), r);
  return fma(r, fma(-a, r, 1.0), r);
}

constexpr int kInvB = 16;  // recurrence steps per batch of loads
// Inverse iteration for the eigenvalues j0 .. j0+nvec-1 (descending order) of the
// tridiagonal, one lane per vector.  flags select the stages of this launch: kInvFactor
// (the pivoted LU of T - lambda I), kInvStart (start vectors), kInvSolve and kInvGs (per
// iteration: the two triangular solves, then Gram-Schmidt against the batch's earlier
// vectors of the same cluster and normalisation), for `iters` iterations.  A single
// batch runs all stages in one launch; longer runs of vectors go in batches with the
// Gram-Schmidt against earlier batches between a launch's solves and its GS.
enum { kInvFactor = 1, kInvStart = 2, kInvSolve = 4, kInvGs = 8, kInvAll = 15 };
// The LU factors are interleaved over the vectors, so that one recurrence step's
// accesses of all lanes share a cache line: field f of step i of vector kv sits at
// lu[(i * kLuF + f) * kMaxVec + kv] (f: 0 pivot reciprocal, 1 and 2 the U
// superdiagonals, 3 multiplier, 4 pivot flag).
constexpr int kLuF = 5;
struct LuCol {
  double *p;  // lu + f * kMaxVec + kv
  __device__ double &operator[](int32_t i) const { return p[(int64_t)i * (kLuF * kMaxVec)]; }
};
__global__ __launch_bounds__(256) void k_invit(EigWs ws, int32_t n, int32_t nvec, int32_t j0,
                                               int32_t flags, int32_t iters) {
  __shared__ double red[4];
  __shared__ double dots[2];
  const double tnorm = ws.sc[S_TNORM];
  const double eps = 2.220446049250313e-16;
  const double tiny = eps * tnorm;
  const int tid = threadIdx.x;
  if ((flags & kInvFactor) && tid < nvec) {
    const int kv = tid;
    const double lam = ws.eval[n - 1 - (j0 + kv)];
    const LuCol u0{ws.lu + kv}, u1{ws.lu + kMaxVec + kv}, u2{ws.lu + 2 * kMaxVec + kv},
        ml{ws.lu + 3 * kMaxVec + kv}, pv{ws.lu + 4 * kMaxVec + kv};  // u0: pivot reciprocals
    double acur = ws.d[0] - lam, bcur = (n > 1) ? ws.e[0] : 0.0;
    // the recurrence is serial; its operands are loaded kInvB steps at a time, one
    // batch ahead of the batch being computed, so memory latency stays hidden
    auto ldf = [&](int32_t i0, double (&eb)[kInvB + 1], double (&db)[kInvB]) {
#pragma unroll
      for (int k = 0; k <= kInvB; ++k) eb[k] = (i0 + k < n - 1) ? ws.e[i0 + k] : 0.0;
#pragma unroll
      for (int k = 0; k < kInvB; ++k) db[k] = (i0 + k + 1 < n) ? ws.d[i0 + k + 1] : 0.0;
    };
    // full: the whole batch is inside the recurrence (no per-step bound check).  The
    // pivot choice is a select, not a branch: lanes (vectors) pivot differently, and
    // both reciprocals are formed, that of l off the serial chain.
    auto runf = [&](bool full, int32_t i0, const double (&eb)[kInvB + 1],
                    const double (&db)[kInvB]) {
#pragma unroll
      for (int k = 0; k < kInvB; ++k) {
        const int32_t i = i0 + k;
        if (full || i < n - 1) {
          const double l = eb[k];
          const double an = db[k] - lam, bn = eb[k + 1];
          const bool sw = fabs(acur) < fabs(l);  // row interchange
          double a0 = acur;
          if (fabs(a0) < tiny) a0 = copysign(tiny, a0 == 0.0 ? 1.0 : a0);
          const double rl = rcp_nr(l), ra = rcp_nr(a0);
          const double r = sw ? rl : ra;
          const double m = (sw ? acur : l) * r;
          u0[i] = r; u1[i] = sw ? an : bcur; u2[i] = sw ? bn : 0.0; ml[i] = m; pv[i] = sw ? 1.0 : 0.0;
          const double na = sw ? bcur - m * an : an - m * bcur;
          bcur = sw ? -m * bn : bn;
          acur = na;
        }
      }
    };
    {
      double eA[kInvB + 1], dA[kInvB], eB[kInvB + 1], dB[kInvB];
      ldf(0, eA, dA);
      for (int32_t i0 = 0; i0 < n - 1; i0 += 2 * kInvB) {
        ldf(i0 + kInvB, eB, dB);
        if (i0 + kInvB <= n - 1) runf(true, i0, eA, dA); else runf(false, i0, eA, dA);
        ldf(i0 + 2 * kInvB, eA, dA);
        if (i0 + 2 * kInvB <= n - 1) runf(true, i0 + kInvB, eB, dB); else runf(false, i0 + kInvB, eB, dB);
      }
    }
    if (fabs(acur) < tiny) acur = copysign(tiny, acur == 0.0 ? 1.0 : acur);
    u0[n - 1] = rcp_nr(acur);
    u1[n - 1] = 0.0; u2[n - 1] = 0.0; ml[n - 1] = 0.0; pv[n - 1] = 0.0;
  }
  // start vectors (a fixed hash of index and vector), filled by the whole block
  if (flags & kInvStart)
    for (int kv = 0; kv < nvec; ++kv) {
      double *x = ws.X + (int64_t)kv * n;
      const int kg = j0 + kv;
      for (int32_t i = tid; i < n; i += 256) {
        uint32_t h = (uint32_t)(i + 1) * 2654435761u ^ (uint32_t)(kg + 7) * 40503u;
        h ^= h >> 15; h *= 2246822519u; h ^= h >> 13;
        x[i] = 0.5 + (double)(h & 0xFFFFFF) / 16777216.0;
      }
    }
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
    if ((flags & kInvSolve) && tid < nvec) {
      const int kv = tid;
      const LuCol u0{ws.lu + kv}, u1{ws.lu + kMaxVec + kv}, u2{ws.lu + 2 * kMaxVec + kv},
          ml{ws.lu + 3 * kMaxVec + kv}, pv{ws.lu + 4 * kMaxVec + kv};
      double *x = ws.X + (int64_t)kv * n;
      double carry = x[0];
      // forward: P L; batch i0 reads x[i0+1 .. i0+kInvB] and writes x[i0 .. i0+kInvB-1],
      // so the next batch's loads (issued before this batch runs) read values no
      // earlier batch writes
      auto ldx = [&](int32_t i0, double (&xb)[kInvB], double (&pb)[kInvB], double (&mb)[kInvB]) {
#pragma unroll
        for (int k = 0; k < kInvB; ++k) {
          const bool in = i0 + k < n - 1;
          xb[k] = in ? x[i0 + k + 1] : 0.0;
          pb[k] = in ? pv[i0 + k] : 0.0;
          mb[k] = in ? ml[i0 + k] : 0.0;
        }
      };
      auto runx = [&](bool full, int32_t i0, const double (&xb)[kInvB], const double (&pb)[kInvB],
                      const double (&mb)[kInvB]) {
#pragma unroll
        for (int k = 0; k < kInvB; ++k) {
          const int32_t i = i0 + k;
          if (full || i < n - 1) {
            const bool sw = pb[k] != 0.0;
            const double lo = sw ? xb[k] : carry, nx = sw ? carry : xb[k];
            x[i] = lo;
            carry = nx - mb[k] * lo;
          }
        }
      };
      {
        double xA[kInvB], pA[kInvB], mA[kInvB], xB[kInvB], pB[kInvB], mB[kInvB];
        ldx(0, xA, pA, mA);
        for (int32_t i0 = 0; i0 < n - 1; i0 += 2 * kInvB) {
          ldx(i0 + kInvB, xB, pB, mB);
          if (i0 + kInvB <= n - 1) runx(true, i0, xA, pA, mA); else runx(false, i0, xA, pA, mA);
          ldx(i0 + 2 * kInvB, xA, pA, mA);
          if (i0 + 2 * kInvB <= n - 1) runx(true, i0 + kInvB, xB, pB, mB);
          else runx(false, i0 + kInvB, xB, pB, mB);
        }
      }
      x[n - 1] = carry;
      // backward: U; batch i0 reads and writes x[i0 - kInvB + 1 .. i0] only
      double x2 = 0.0, x1 = x[n - 1] * u0[n - 1];
      x[n - 1] = x1;
      auto ldb = [&](int32_t i0, double (&xb)[kInvB], double (&ab)[kInvB], double (&bb)[kInvB],
                     double (&cb)[kInvB]) {
#pragma unroll
        for (int k = 0; k < kInvB; ++k) {
          const int32_t i = i0 - k;
          const bool in = i >= 0;
          xb[k] = in ? x[i] : 0.0;
          ab[k] = in ? u0[i] : 1.0;
          bb[k] = in ? u1[i] : 0.0;
          cb[k] = in ? u2[i] : 0.0;
        }
      };
      auto runb = [&](bool full, int32_t i0, const double (&xb)[kInvB], const double (&ab)[kInvB],
                      const double (&bb)[kInvB], const double (&cb)[kInvB]) {
#pragma unroll
        for (int k = 0; k < kInvB; ++k) {
          const int32_t i = i0 - k;
          if (full || i >= 0) {
            const double xi = (xb[k] - bb[k] * x1 - cb[k] * x2) * ab[k];
            x[i] = xi;
            x2 = x1;
            x1 = xi;
          }
        }
      };
      {
        double xA[kInvB], aA[kInvB], bA[kInvB], cA[kInvB];
        double xB[kInvB], aB[kInvB], bB[kInvB], cB[kInvB];
        ldb(n - 2, xA, aA, bA, cA);
        for (int32_t i0 = n - 2; i0 >= 0; i0 -= 2 * kInvB) {
          ldb(i0 - kInvB, xB, aB, bB, cB);
          if (i0 - kInvB + 1 >= 0) runb(true, i0, xA, aA, bA, cA); else runb(false, i0, xA, aA, bA, cA);
          ldb(i0 - 2 * kInvB, xA, aA, bA, cA);
          if (i0 - 2 * kInvB + 1 >= 0) runb(true, i0 - kInvB, xB, aB, bB, cB);
          else runb(false, i0 - kInvB, xB, aB, bB, cB);
        }
      }
    }
    __syncthreads();
    if (!(flags & kInvGs)) continue;
    for (int kv = 0; kv < nvec; ++kv) {
      double *x = ws.X + (int64_t)kv * n;
      const double lam = ws.eval[n - 1 - (j0 + kv)];
      for (int pv2 = 0; pv2 < kv; ++pv2) {
        const double lp = ws.eval[n - 1 - (j0 + pv2)];
        if (fabs(lam - lp) > 1e-3 * tnorm) continue;
        const double *p = ws.X + (int64_t)pv2 * n;
        double s = 0.0;
        for (int32_t i = tid; i < n; i += 256) s += x[i] * p[i];
        s = block_sum(s, red);
        for (int32_t i = tid; i < n; i += 256) x[i] -= s * p[i];
        __syncthreads();
      }
      double s = 0.0;
      for (int32_t i = tid; i < n; i += 256) s += x[i] * x[i];
      s = block_sum(s, red);
      if (tid == 0) dots[0] = 1.0 / sqrt(s);
      __syncthreads();
      const double inv = dots[0];
      for (int32_t i = tid; i < n; i += 256) x[i] *= inv;
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------- back-transform
// Panel p holds reflector rows c0 = p*kNB .. c0+nb-1 (v_q in A[c0+q][c0+q+1:],
// v_q[c0+q+1] = 1).  k_bt_gram forms G = V^T V over the panel's rows in kBtSplit
// row ranges (one block each, partial G into part1), k_bt_prep sums a panel's partials
// in range order and forms T (dlarft forward, columnwise:
// T[0:j, j] = -tau_j T[0:j, 0:j] G[0:j, j]) into Tall[p].  part1 holds
// (p_last + 1) * kBtSplit * kNB^2 doubles: 16 KNB^2 per 32 matrix rows, within its
// 512 doubles per row.
constexpr int kBtRows = 32;
constexpr int kBtSplit = 16;   // row ranges per panel in k_bt_gram
constexpr int kBtGRows = 128;  // rows staged per round in k_bt_gram
__device__ __forceinline__ int32_t bt_gram_rows(int32_t n, int32_t c0) {
  const int32_t rows = n - c0 - 1;  // reflector rows c0 + 1 .. n - 1
  return (rows + kBtSplit * kBtGRows - 1) / (kBtSplit * kBtGRows) * kBtGRows;
}
__global__ __launch_bounds__(256) void k_bt_gram(EigWs ws, int32_t n, int32_t last) {
  __shared__ double sv[kNB][kBtGRows + 1];
  const int32_t p = blockIdx.x, s = blockIdx.y;
  const int32_t c0 = p * kNB;
  const int32_t nb = min(kNB, last - c0 + 1);
  const int32_t per = bt_gram_rows(n, c0);
  const int32_t r_beg = c0 + 1 + s * per, r_end = min(n, r_beg + per);
  if (r_beg >= r_end) return;
  const int32_t lda = ws.lda;
  // element k of a thread: reflector q = (t + 256 k) / 128, row offset (t + 256 k) % 128;
  // indices are clamped so every load is unconditional, and masked when staged
  auto load = [&](int32_t rb, double (&v)[16]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int idx = threadIdx.x + 256 * k, q = idx / kBtGRows, i = idx % kBtGRows;
      const int32_t cq = c0 + min(q, nb - 1), r = min(rb + i, n - 1);
      v[k] = ws.a[(int64_t)cq * lda + r];
    }
  };
  double g[4] = {0.0, 0.0, 0.0, 0.0}, cur[16], nxt[16];
  load(r_beg, cur);
  for (int32_t rb = r_beg; rb < r_end; rb += kBtGRows) {
    __syncthreads();  // the previous round's reads of sv are done
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int idx = threadIdx.x + 256 * k, q = idx / kBtGRows, i = idx % kBtGRows;
      const int32_t r = rb + i;
      sv[q][i] = (q < nb && r < r_end && r > c0 + q) ? cur[k] : 0.0;
    }
    __syncthreads();
    if (rb + kBtGRows < r_end) load(rb + kBtGRows, nxt);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int idx = threadIdx.x + 256 * k;
      const int q1 = idx / kNB, q2 = idx % kNB;
      double a = 0.0;
#pragma unroll 8
      for (int i = 0; i < kBtGRows; ++i) a += sv[q1][i] * sv[q2][i];
      g[k] += a;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) cur[k] = nxt[k];
  }
  double *out = ws.part1 + ((int64_t)p * kBtSplit + s) * (kNB * kNB);
#pragma unroll
  for (int k = 0; k < 4; ++k) out[threadIdx.x + 256 * k] = g[k];
}
__global__ __launch_bounds__(256) void k_bt_prep(EigWs ws, int32_t n, int32_t last) {
  __shared__ double G[kNB][kNB + 1], T[kNB][kNB + 1];
  const int32_t p = blockIdx.x;
  const int32_t c0 = p * kNB;
  const int32_t nb = min(kNB, last - c0 + 1);
  const int32_t per = bt_gram_rows(n, c0);
  const int32_t nsp = (n - c0 - 1 + per - 1) / per;  // row ranges holding rows
  const double *src = ws.part1 + (int64_t)p * kBtSplit * (kNB * kNB);
  double g[4] = {0.0, 0.0, 0.0, 0.0};
  for (int32_t s = 0; s < nsp; ++s)
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] += src[(int64_t)s * (kNB * kNB) + threadIdx.x + 256 * k];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int idx = threadIdx.x + 256 * k;
    G[idx / kNB][idx % kNB] = g[k];
  }
  for (int idx = threadIdx.x; idx < kNB * kNB; idx += 256) T[idx / kNB][idx % kNB] = 0.0;
  __syncthreads();
  for (int jj = 0; jj < nb; ++jj) {
    const double tj = ws.tau[c0 + jj];
    // T[i][jj] for i < jj in parallel; T is upper triangular
    double s = 0.0;
    const int i = threadIdx.x;
    if (i < jj)
      for (int k = i; k < jj; ++k) s += T[i][k] * G[k][jj];
    __syncthreads();
    if (i < jj) T[i][jj] = -tj * s;
    if (i == jj) T[jj][jj] = tj;
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < kNB * kNB; idx += 256)
    ws.Tall[(int64_t)p * kNB * kNB + idx] = T[idx / kNB][idx % kNB];
}

// Fused step for panel p: X[r] -= sum_q v_{p+1,q}[r] Z[q] (previous panel, if any),
// then this block's partial dots V_p^T X; k_bt_reduce forms Z <- T_p (V_p^T X).
__global__ __launch_bounds__(256) void k_bt_step(EigWs ws, int32_t n, int32_t p, int32_t last,
                                                 int32_t nvec, int32_t apply_prev) {
  __shared__ double szbuf[kNB * kMaxVec];
  __shared__ double sv[kNB][kBtRows + 1];   // this panel's reflectors over the block's rows
  __shared__ double svn[kNB][kBtRows + 1];  // the previous (higher) panel's
  __shared__ double sx[kMaxVec][kBtRows + 1];
  double(*sz)[kMaxVec] = (double(*)[kMaxVec])szbuf;
  const int32_t lda = ws.lda;
  const int32_t c0 = p * kNB;
  const int32_t nb = min(kNB, last - c0 + 1);
  const int32_t rbase = c0 + 1 + blockIdx.x * kBtRows;
  const int32_t rows = min(kBtRows, n - rbase);
  const int32_t c0n = c0 + kNB;  // previous (higher) panel
  const int32_t nbn = apply_prev ? min(kNB, last - c0n + 1) : 0;
  // every operand goes to LDS in one round of loads: Z, the block's rows of X, and
  // both panels' reflector entries over those rows.  All loads are issued before the
  // first LDS store (indices clamped into range; entries outside are stored as zero).
  constexpr int kZU = kNB * kMaxVec / 256, kXU = kMaxVec * kBtRows / 256;
  constexpr int kVU = kNB * kBtRows / 256;
  double zv[kZU], xv[kXU], vv[kVU], vnv[kVU] = {};
  const int t = threadIdx.x;
  if (apply_prev) {
#pragma unroll
    for (int u = 0; u < kZU; ++u) zv[u] = ws.Z[t + 256 * u];  // the whole kNB x kMaxVec block
  }
#pragma unroll
  for (int u = 0; u < kXU; ++u) {
    const int idx = t + 256 * u, o = min(idx / kBtRows, nvec - 1), i = min(idx % kBtRows, rows - 1);
    if (256 * u < nvec * kBtRows) xv[u] = ws.X[(int64_t)o * n + rbase + i];
  }
#pragma unroll
  for (int u = 0; u < kVU; ++u) {
    const int idx = t + 256 * u, q = idx / kBtRows, i = min(idx % kBtRows, rows - 1);
    vv[u] = ws.a[(int64_t)(c0 + min(q, nb - 1)) * lda + rbase + i];
    if (nbn > 0) vnv[u] = ws.a[(int64_t)(c0n + min(q, nbn - 1)) * lda + rbase + i];
  }
  __builtin_amdgcn_sched_barrier(0);
  if (apply_prev) {
#pragma unroll
    for (int u = 0; u < kZU; ++u) szbuf[t + 256 * u] = zv[u];
  }
#pragma unroll
  for (int u = 0; u < kXU; ++u) {
    const int idx = t + 256 * u, o = idx / kBtRows, i = idx % kBtRows;
    if (256 * u < nvec * kBtRows && o < nvec) sx[o][i] = i < rows ? xv[u] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < kVU; ++u) {
    const int idx = t + 256 * u, q = idx / kBtRows, i = idx % kBtRows;
    const int32_t r = rbase + i;
    sv[q][i] = (q < nb && i < rows && r > c0 + q) ? vv[u] : 0.0;
    svn[q][i] = (q < nbn && i < rows && r > c0n + q) ? vnv[u] : 0.0;
  }
  __syncthreads();
  if (apply_prev) {
    for (int idx = threadIdx.x; idx < nvec * kBtRows; idx += 256) {
      const int o = idx / kBtRows, i = idx % kBtRows;
      if (i >= rows) continue;
      double s = 0.0;
      for (int q = 0; q < nbn; ++q) s += svn[q][i] * sz[q][o];
      sx[o][i] -= s;
      ws.X[(int64_t)o * n + rbase + i] = sx[o][i];
    }
  }
  __syncthreads();  // sx updated
  // this block's partial dots V_p^T X over its rows; k_bt_reduce sums them
  const int nd = kNB * nvec;
  for (int idx = threadIdx.x; idx < nd; idx += 256) {
    const int q = idx / nvec, o = idx % nvec;
    double s = 0.0;
    for (int i = 0; i < rows; ++i) s += sv[q][i] * sx[o][i];
    ws.part1[(int64_t)blockIdx.x * nd + idx] = s;
  }
}

// Z_p = T_p (V_p^T X) from the partial dots of k_bt_step's nblk blocks: one block per
// vector o.  Thread (g, k) sums the partials
// b = g, g + 32, ... of entry (k, o), all its loads in one or two rounds; the 32 group
// sums are then added in group order and multiplied by T_p.
__global__ __launch_bounds__(1024) void k_bt_reduce(EigWs ws, int32_t p, int32_t last,
                                                    int32_t nvec, int32_t nblk) {
  __shared__ double acc[32][kNB + 1];
  __shared__ double col[kNB];
  __shared__ double sT[kNB][kNB + 1];
  const int o = blockIdx.x;
  const int nd = kNB * nvec;
  const int32_t nb = min(kNB, last - p * kNB + 1);
  const int k = threadIdx.x % kNB, g = threadIdx.x / kNB;
  const double tv = ws.Tall[(int64_t)p * kNB * kNB + threadIdx.x];  // T_p, one entry each
  acc[g][k] = sum_strided(ws.part1 + k * nvec + o, nd, g, 32, nblk);
  sT[g][k] = tv;
  __syncthreads();
  if (threadIdx.x < kNB) {
    double s = acc[0][k];
    for (int gg = 1; gg < 32; ++gg) s += acc[gg][k];
    col[k] = s;
  }
  __syncthreads();
  if (threadIdx.x < kNB) {
    const int q = threadIdx.x;
    double s = 0.0;
    for (int kk = q; kk < nb; ++kk) s += sT[q][kk] * col[kk];
    ws.Z[q * kMaxVec + o] = s;
  }
}

// final application of panel 0's Z
__global__ __launch_bounds__(256) void k_bt_apply0(EigWs ws, int32_t n, int32_t last,
                                                   int32_t nvec) {
  const int32_t r = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t nb = min(kNB, last + 1);
  double v[kNB];
#pragma unroll
  for (int q = 0; q < kNB; ++q) v[q] = (q < nb && r > q) ? ws.a[(int64_t)q * ws.lda + r] : 0.0;
  for (int o = 0; o < nvec; ++o) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kNB; ++q) s += v[q] * ws.Z[q * kMaxVec + o];
    ws.X[(int64_t)o * n + r] -= s;
  }
}

// normalise + sign rule (largest |.| component positive, first on ties)
__global__ __launch_bounds__(256) void k_finish_vec(EigWs ws, int32_t n) {
  __shared__ double red[4];
  __shared__ double bv[256];
  __shared__ int32_t bi[256];
  __shared__ double fac;
  double *x = ws.X + (int64_t)blockIdx.x * n;
  double s = 0.0, best = -1.0;
  // start from an element this thread owns (or 0): the index stays in bounds even
  // when no comparison succeeds (a vector of NaNs)
  int32_t besti = (int32_t)threadIdx.x < n ? (int32_t)threadIdx.x : 0;
  for (int32_t i = threadIdx.x; i < n; i += 256) {
    s += x[i] * x[i];
    double a = fabs(x[i]);
    if (a > best) {
      best = a;
      besti = i;
    }
  }
  s = block_sum(s, red);
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = besti;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      double b2 = bv[threadIdx.x + o];
      int32_t i2 = bi[threadIdx.x + o];
      if (b2 > bv[threadIdx.x] || (b2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
        bv[threadIdx.x] = b2;
        bi[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double f = 1.0 / sqrt(s);
    const int32_t ib = bi[0];
    if (ib >= 0 && ib < n && x[ib] < 0.0) f = -f;
    fac = f;
  }
  __syncthreads();
  for (int32_t i = threadIdx.x; i < n; i += 256) x[i] *= fac;
}

// ---------------------------------------------------------------- driver
// wall-clock ticks the w-column blocks of k_column sleep before polling: 80% of the
// symv part's streaming time at 6.5 TB/s (its measured rate on large trailing
// matrices), nothing below 2 us
static int64_t quiet_ticks(int32_t L, int32_t clock_khz) {
  const double us = 0.8 * (double)L * (double)L * 8.0 / 6.5e12 * 1e6;
  return us < 2.0 ? 0 : (int64_t)(us * clock_khz / 1000.0);
}

// Trailing order below which k_column runs (one launch per column); longer columns
// run k_symv + k_wcol.  Measured on MI355X: merging at every length costs ~10 us per
// column on 5,000-long columns (the w-column blocks held for the whole stream and the
// store drain) and saves a few us per column on short ones.  With the 64-row
// w-column chunks, cuts of 500 to 1,000 were best for n = 3,000 and 5,000 (1 to
// 3,000 tried; never merging costs 1.1 ms at n = 5,000).
// EIG_NO_MERGE=1 runs every column as k_symv + k_wcol: counter-collection runs use it,
// since no launch then waits on other blocks of its own grid.
constexpr int32_t kMergeBelow = 1000;
static void launch_tridiag(EigWs ws, int32_t n, int32_t clock_khz, hipStream_t st) {
  const int32_t merge_below = getenv("EIG_NO_MERGE") != nullptr ? 1 : kMergeBelow;
  hipLaunchKernelGGL(k_col0, dim3(1), dim3(256), 0, st, ws, n);
  int32_t base = 0;  // hand-off tickets taken by the earlier columns
  for (int32_t c = 0; c + 1 < n; ++c) {
    const int32_t j = c % kNB;
    const int32_t L = n - c - 1;
    const int32_t gsym = (L + kSymvRows - 1) / kSymvRows, gw = (L + kWRows - 1) / kWRows;
    if (L >= merge_below) {
      const int32_t g4 = (L + kSymvRowsLong - 1) / kSymvRowsLong;
      hipLaunchKernelGGL(k_symv<kSymvRowsLong>, dim3((unsigned)g4), dim3(256), 0, st, ws, n, c, j);
      hipLaunchKernelGGL(k_wcol, dim3((unsigned)gw), dim3(256), 0, st, ws, n, c, j, g4);
    } else {
      hipLaunchKernelGGL(k_column, dim3((unsigned)(gw + gsym)), dim3(256), 0, st, ws, n, c, j, base,
                         quiet_ticks(L, clock_khz));
      base += gsym;
    }
    // panel end: update the trailing matrix before the next panel's first symv
    if (j == kNB - 1 && c + 2 < n) {
      const int32_t k1 = c + 1;
      const int32_t nt = (n - k1 + 63) / 64;
      hipLaunchKernelGGL(k_syr2k, dim3((unsigned)(nt * nt)), dim3(256), 0, st, ws, n, k1, kNB);
    }
  }
}

struct TridiagGraph {
  int32_t n = -1;
  void *a = nullptr;
  hipGraphExec_t exec = nullptr;
  std::mutex mu;  // every context on the device shares the cached graph
};
static TridiagGraph g_graph[16];  // per device

int eig_run_eigh(eig_ctx *ctx, const double *d_a, int32_t n, int32_t nvec, double *evals,
                 double *evecs, bool internal) {
  if (n <= 0) return eig_set_err(ctx, EIG_E_ARG, "eigh: n <= 0");
  if (nvec < 0 || nvec > n) return eig_set_err(ctx, EIG_E_ARG, "eigh: nvec must be in [0, n]");
  if (internal && nvec > kMaxVec)
    return eig_set_err(ctx, EIG_E_ARG, "eigh: internal calls take at most %d vectors", kMaxVec);
  // an internal call leaves the caller's stage timing (same events) running
  auto tbeg = [&]() { if (!internal) eig_time_begin(ctx); };
  auto tend = [&](const char *what) { if (!internal) eig_time_end(ctx, what); };
  const int64_t N = n;
  const int32_t lda = (int32_t)(((N + 7) / 8) * 8);
  const int64_t maxblk = (N + kSymvRowsLong - 1) / kSymvRowsLong + 8;  // symv partials
  const int64_t maxnv = (int64_t)kNB * kMaxVec;
  const int64_t maxw = (N + kWRows - 1) / kWRows + 2;
  const int64_t npanel = (N + kNB - 1) / kNB + 1;
  size_t need = 0;
  auto add = [&](size_t cnt) {
    size_t o = need;
    need += ((cnt * 8 + 255) / 256) * 256;
    return o;
  };
  size_t o_a = add((size_t)N * lda), o_d = add(N), o_e = add(N), o_e2 = add(N),
         o_tau = add(N), o_V = add(kNB * N), o_W = add(kNB * N), o_x = add(lda), o_y = add(N),
         o_sc = add(S_COUNT), o_p1 = add((size_t)maxblk * maxnv),
         o_ev = add(N),
         o_X = add((size_t)kMaxVec * N), o_lu = add((size_t)kMaxVec * kLuF * N),
         o_Z = add(kNB * kMaxVec), o_T = add((size_t)npanel * kNB * kNB),
         o_pw = add((size_t)2 * maxw * kPW), o_ps = add((size_t)maxblk), o_sync = add(8);
  void *old = ctx->ework.p;
  EIG_TRY(eig_ensure(ctx, ctx->ework, need));
  char *base = (char *)ctx->ework.p;
  EigWs ws;
  ws.a = (double *)(base + o_a);
  ws.lda = lda;
  ws.d = (double *)(base + o_d);
  ws.e = (double *)(base + o_e);
  ws.e2 = (double *)(base + o_e2);
  ws.tau = (double *)(base + o_tau);
  ws.V = (double *)(base + o_V);
  ws.W = (double *)(base + o_W);
  ws.xtmp = (double *)(base + o_x);
  ws.y = (double *)(base + o_y);
  ws.sc = (double *)(base + o_sc);
  ws.part1 = (double *)(base + o_p1);
  ws.eval = (double *)(base + o_ev);
  ws.X = (double *)(base + o_X);
  ws.lu = (double *)(base + o_lu);
  ws.Z = (double *)(base + o_Z);
  ws.Tall = (double *)(base + o_T);
  ws.pw = (double *)(base + o_pw);
  ws.ps = (double *)(base + o_ps);
  ws.sync = (int32_t *)(base + o_sync);
  ws.maxw = (int32_t)maxw;
  hipStream_t st = ctx->stream;
  if (lda != n) EIG_HIP(ctx, hipMemsetAsync(ws.a, 0, sizeof(double) * N * lda, st));
  EIG_HIP(ctx, hipMemcpy2DAsync(ws.a, sizeof(double) * lda, d_a, sizeof(double) * N,
                                sizeof(double) * N, N, hipMemcpyDeviceToDevice, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.sc, 0, sizeof(double) * S_COUNT, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.e, 0, sizeof(double) * N, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.e2, 0, sizeof(double) * N, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.tau, 0, sizeof(double) * N, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.V, 0, sizeof(double) * kNB * N, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.W, 0, sizeof(double) * kNB * N, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.xtmp, 0, sizeof(double) * lda, st));

  // ---- tridiagonalisation (graph replay for n above a few hundred)
  int32_t clock_khz = 0;
  EIG_HIP(ctx, hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
  tbeg();
  const bool use_graph = !internal && n >= 256 && getenv("EIG_NO_GRAPH") == nullptr;
  if (use_graph) {
    TridiagGraph &G = g_graph[ctx->device & 15];
    std::lock_guard<std::mutex> lock(G.mu);
    if (G.exec == nullptr || G.n != n || G.a != (void *)ws.a || old != ctx->ework.p) {
      if (G.exec) hipGraphExecDestroy(G.exec);
      G.exec = nullptr;
      hipGraph_t graph;
      EIG_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      launch_tridiag(ws, n, clock_khz, st);
      EIG_HIP(ctx, hipStreamEndCapture(st, &graph));
      EIG_HIP(ctx, hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0));
      hipGraphDestroy(graph);
      G.n = n;
      G.a = ws.a;
    }
    EIG_HIP(ctx, hipGraphLaunch(G.exec, st));
  } else {
    launch_tridiag(ws, n, clock_khz, st);
  }
  EIG_HIP(ctx, hipGetLastError());
  tend("tridiag");
  // ---- eigenvalues
  tbeg();
  hipLaunchKernelGGL(k_gersh, dim3(1), dim3(256), 0, st, ws, n);
  {
    int64_t threads = (int64_t)n * kMS;
    // the grid's counts go to the inverse iteration's LU area, unused until then
    // (kMaxVec * 5 * N doubles hold the kMsGrid counts from N = 103 on)
    int32_t *gcnt = nullptr;
    if ((int64_t)kMaxVec * kLuF * N * 2 >= kMsGrid && n >= 128) {
      gcnt = (int32_t *)ws.lu;
      hipLaunchKernelGGL(k_msgrid, dim3(kMsGrid / 256), dim3(256), 0, st, ws, n, gcnt);
    }
    hipLaunchKernelGGL(k_multisect, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, ws, n,
                       (const int32_t *)gcnt);
  }
  EIG_HIP(ctx, hipGetLastError());
  tend("bisect");
  // ---- eigenvectors: one batch of at most kMaxVec (inverse iteration, back-transform,
  // sign rule), or, for more vectors, batches of kMaxVec in eigenvalue order with
  // each iteration's vectors orthogonalised against the earlier batches' vectors of
  // the same cluster (classical Gram-Schmidt, twice, as FP64 MFMA products) between
  // the solves and the batch's own Gram-Schmidt
  const int32_t last = n - 2;  // last reflector row
  const int32_t p_last = last >= 0 ? last / kNB : -1;
  auto back_transform = [&](int32_t nb) -> int {
    if (last >= 0) {
      for (int32_t p = p_last; p >= 0; --p) {
        const int32_t c0 = p * kNB;
        const unsigned g = (unsigned)((n - c0 - 1 + kBtRows - 1) / kBtRows);
        hipLaunchKernelGGL(k_bt_step, dim3(g), dim3(256), 0, st, ws, n, p, last, nb,
                           (int32_t)(p < p_last));
        hipLaunchKernelGGL(k_bt_reduce, dim3((unsigned)nb), dim3(1024), 0, st, ws, p, last, nb,
                           (int32_t)g);
      }
      hipLaunchKernelGGL(k_bt_apply0, dim3((unsigned)((n - 1 + 255) / 256)), dim3(256), 0, st, ws,
                         n, last, nb);
    }
    hipLaunchKernelGGL(k_finish_vec, dim3((unsigned)nb), dim3(256), 0, st, ws, n);
    EIG_HIP(ctx, hipGetLastError());
    return EIG_OK;
  };
  tbeg();
  if (nvec > 0 && nvec <= kMaxVec) {
    // two iterations: the eigenvalues are accurate to a few ulps, so one solve already
    // leaves contamination of order eps ||T|| / gap from outside a cluster, and the
    // second squares it (the parity tests pass with one; two keep a margin)
    hipLaunchKernelGGL(k_invit, dim3(1), dim3(256), 0, st, ws, n, nvec, 0, (int32_t)kInvAll, 2);
    tend("invit");
    tbeg();
    if (last >= 0) {
      hipLaunchKernelGGL(k_bt_gram, dim3((unsigned)(p_last + 1), kBtSplit), dim3(256), 0, st, ws, n,
                         last);
      hipLaunchKernelGGL(k_bt_prep, dim3((unsigned)(p_last + 1)), dim3(256), 0, st, ws, n, last);
    }
    EIG_TRY(back_transform(nvec));
  } else if (nvec > kMaxVec) {
    // the T-basis vectors of all batches (vector j at j*N) and the CGS coefficients
    EIG_TRY(eig_ensure(ctx, ctx->evall, sizeof(double) * ((size_t)N * nvec + (size_t)kMaxVec * nvec)));
    double *XT = (double *)ctx->evall.p, *Cc = XT + (size_t)N * nvec;
    std::vector<double> ev(N);
    double sc[S_COUNT];
    EIG_HIP(ctx, hipMemcpyAsync(ev.data(), ws.eval, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    EIG_HIP(ctx, hipMemcpyAsync(sc, ws.sc, sizeof(sc), hipMemcpyDeviceToHost, st));
    EIG_HIP(ctx, hipStreamSynchronize(st));
    const double tol = 1e-3 * sc[S_TNORM];  // the cluster criterion of k_invit
    if (last >= 0) {
      hipLaunchKernelGGL(k_bt_gram, dim3((unsigned)(p_last + 1), kBtSplit), dim3(256), 0, st, ws, n,
                         last);
      hipLaunchKernelGGL(k_bt_prep, dim3((unsigned)(p_last + 1)), dim3(256), 0, st, ws, n, last);
    }
    for (int32_t j0 = 0; j0 < nvec; j0 += kMaxVec) {
      const int32_t nb = std::min<int32_t>(kMaxVec, nvec - j0);
      int32_t c0 = j0;  // earlier vectors within tol of eigenvalue j0 (descending order)
      while (c0 > 0 && std::fabs(ev[N - 1 - (c0 - 1)] - ev[N - 1 - j0]) <= tol) --c0;
      const int32_t nc = j0 - c0;
      for (int it = 0; it < 3; ++it) {
        const int32_t f = (it == 0 ? (kInvFactor | kInvStart) : 0) | kInvSolve;
        hipLaunchKernelGGL(k_invit, dim3(1), dim3(256), 0, st, ws, n, nb, j0, f, 1);
        for (int rep = 0; rep < 2 && nc > 0; ++rep) {
          const double *P = XT + (size_t)c0 * N;
          EIG_TRY(eig_gemm_f64(ctx, st, true, false, nc, nb, N, 1.0, P, N, 0, ws.X, N, 0, 0.0, Cc,
                               nc, 0, 1));
          EIG_TRY(eig_gemm_f64(ctx, st, false, false, N, nb, nc, -1.0, P, N, 0, Cc, nc, 0, 1.0,
                               ws.X, N, 0, 1));
        }
        hipLaunchKernelGGL(k_invit, dim3(1), dim3(256), 0, st, ws, n, nb, j0, (int32_t)kInvGs, 1);
      }
      EIG_HIP(ctx, hipMemcpyAsync(XT + (size_t)j0 * N, ws.X, sizeof(double) * N * nb,
                                  hipMemcpyDeviceToDevice, st));
      EIG_TRY(back_transform(nb));
      if (evecs)
        EIG_HIP(ctx, hipMemcpyAsync(evecs + (size_t)j0 * N, ws.X, sizeof(double) * N * nb,
                                    hipMemcpyDeviceToHost, st));
      if (!internal && j0 == 0) {
        EIG_TRY(eig_ensure(ctx, ctx->evecs_dev, sizeof(double) * N * kMaxVec));
        EIG_HIP(ctx, hipMemcpyAsync(ctx->evecs_dev.p, ws.X, sizeof(double) * N * kMaxVec,
                                    hipMemcpyDeviceToDevice, st));
      }
    }
    EIG_HIP(ctx, hipStreamSynchronize(st));
    tend("invit");
    tbeg();
  } else {
    tend("invit");
    tbeg();
  }
  tend("eigvec");
  std::vector<double> asc(N);
  int32_t hsync[2] = {0, 0};
  EIG_HIP(ctx, hipMemcpyAsync(hsync, ws.sync, sizeof(hsync), hipMemcpyDeviceToHost, st));
  EIG_HIP(ctx, hipMemcpyAsync(asc.data(), ws.eval, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  if (nvec > 0 && nvec <= kMaxVec) {
    if (evecs)
      EIG_HIP(ctx, hipMemcpyAsync(evecs, ws.X, sizeof(double) * N * nvec, hipMemcpyDeviceToHost, st));
    if (!internal) {
      EIG_TRY(eig_ensure(ctx, ctx->evecs_dev, sizeof(double) * N * std::max(nvec, 1)));
      EIG_HIP(ctx, hipMemcpyAsync(ctx->evecs_dev.p, ws.X, sizeof(double) * N * nvec,
                                  hipMemcpyDeviceToDevice, st));
    }
  }
  EIG_HIP(ctx, hipStreamSynchronize(st));
  if (hsync[1] != 0)
    return eig_set_err(ctx, EIG_E_CONVERGE, "eigh: tridiagonalisation hand-off timed out");
  for (int64_t i = 0; i < N; ++i) {
    if (std::isnan(asc[i])) return eig_set_err(ctx, EIG_E_CONVERGE, "eigenvalue is NaN");
    if (evals) evals[i] = asc[N - 1 - i];
  }
  if (nvec > 0 && evecs)
    for (int64_t i = 0; i < N * nvec; ++i)
      if (!std::isfinite(evecs[i]))
        return eig_set_err(ctx, EIG_E_CONVERGE, "eigenvector %lld is not finite",
                           (long long)(i / N));
  if (!internal) ctx->last_nvec = nvec;
  return EIG_OK;
}

extern "C" int eig_eigh_xtx(eig_ctx *ctx, int32_t nvec, double *evals, double *evecs) {
  if (!ctx) return EIG_E_ARG;
  if (!ctx->have_xtx) return eig_set_err(ctx, EIG_E_STATE, "eig_eigh_xtx before eig_xtx");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  return eig_run_eigh(ctx, (const double *)ctx->xtx.p, ctx->nrows, nvec, evals, evecs);
}

extern "C" int eig_eigh(eig_ctx *ctx, const double *a, int32_t n, int32_t nvec, double *evals,
                        double *evecs) {
  if (!ctx || !a || n <= 0) return eig_set_err(ctx, EIG_E_ARG, "eig_eigh: args");
  for (int64_t i = 0; i < (int64_t)n * n; ++i)
    if (!std::isfinite(a[i])) return eig_set_err(ctx, EIG_E_ARG, "eig_eigh: matrix is not finite");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  EIG_TRY(eig_ensure(ctx, ctx->scratch, sizeof(double) * (size_t)n * n));
  EIG_HIP(ctx, hipMemcpyAsync(ctx->scratch.p, a, sizeof(double) * (size_t)n * n,
                              hipMemcpyHostToDevice, ctx->stream));
  return eig_run_eigh(ctx, (const double *)ctx->scratch.p, n, nvec, evals, evecs);
}
