// FP64 symmetric eigensolver on the device.
//
//  1. Householder tridiagonalisation, blocked by panels of kNB columns
//     (the dsytrd/dlatrd scheme).  The matrix is row-major and symmetric, so
//     "column c below the diagonal" is read as row c right of the diagonal and
//     every vector access is contiguous.  Per column three kernels run:
//       k_col  : row c minus the panel's rank-2j correction, Householder scalars
//       k_symv : y = A_trail v, plus the panel dot products V v and W v
//       k_wvec : w = tau (y - V (W v) - W (V v)), then alpha = -tau/2 w.v
//     Cross-block reductions use the last-arriving block (atomic ticket), so no
//     kernel ever waits on another block.  After each panel k_syr2k applies
//     A -= V W^T + W V^T to the trailing matrix.  The Householder vector of row
//     c is left in A[c][c+1:].
//  2. All eigenvalues of the tridiagonal T by bisection on Sturm counts, one
//     thread per eigenvalue.
//  3. The top nvec eigenvectors of T by inverse iteration (LU with partial
//     pivoting), with Gram-Schmidt inside clusters of close eigenvalues.
//  4. Back-transform X = Q Y panel by panel with the compact-WY form
//     Q_p = I - V_p T_p V_p^T, last panel first.
#include <cmath>
#include <cstring>
#include <vector>

#include "eig_internal.h"

constexpr int kNB = 32;       // panel width
constexpr int kMaxVec = 64;   // max eigenvectors computed

// scalar slots of the solver state (doubles)
enum {
  S_TAU = 0,      // tau of the current column
  S_SCALE,        // 1/(alpha-beta) of the current column
  S_ALPHAW,       // alpha of the current w (w_final = w_pre + alpha v)
  S_GL, S_GU, S_PIVMIN, S_TNORM,
  S_PD = 8,       // V_q . v  (kNB)
  S_QD = 8 + kNB, // W_q . v  (kNB)
  S_COUNT = 8 + 2 * kNB
};

struct EigWs {
  double *a;       // n x n working matrix (destroyed)
  double *d, *e, *e2, *tau;
  double *V, *W;   // kNB x n panel vectors
  double *xtmp, *wtmp, *y;
  double *sc;      // scalars
  double *part;    // block partials
  unsigned *ticket;
  double *eval;    // ascending
  double *X;       // nvec x n eigenvectors
  double *lu;      // nvec x 5n scratch for inverse iteration
  double *Z;       // kNB x kMaxVec
};

__device__ inline bool last_block(unsigned *ticket) {
  __shared__ bool amlast;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned prev = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    amlast = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (amlast) __threadfence();
  return amlast;
}

__device__ inline double block_sum(double v, double *red) {
  // 256 threads
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------- k_col
// Row c after the panel correction -> xtmp[c..n); finalises W[j-1] = wtmp +
// alpha v on this block's range; last block forms the Householder reflector.
__global__ __launch_bounds__(256) void k_col(EigWs ws, int32_t n, int32_t c, int32_t j) {
  __shared__ double red[4];
  const int32_t s = c + blockIdx.x * 256 + threadIdx.x;
  double x = 0.0;
  double nrm = 0.0;
  if (s < n) {
    double aw = 0.0;
    if (j > 0) {
      // W[j-1] on the fly (race-free: wtmp and V are not written here)
      const double al = ws.sc[S_ALPHAW];
      const double *vj = ws.V + (int64_t)(j - 1) * n, *wj = ws.wtmp;
      double wc = wj[c] + al * vj[c];
      double ws_ = wj[s] + al * vj[s];
      ws.W[(int64_t)(j - 1) * n + s] = ws_;
      x = ws.a[(int64_t)c * n + s];
      for (int q = 0; q < j - 1; ++q) {
        const double *vq = ws.V + (int64_t)q * n, *wq = ws.W + (int64_t)q * n;
        x -= vq[c] * wq[s] + wq[c] * vq[s];
      }
      x -= vj[c] * ws_ + wc * vj[s];
      (void)aw;
    } else {
      x = ws.a[(int64_t)c * n + s];
    }
    ws.xtmp[s] = x;
    if (s == c) ws.d[c] = x;
    if (s >= c + 2) nrm = x * x;
  }
  double bs = block_sum(nrm, red);
  if (threadIdx.x == 0) ws.part[blockIdx.x] = bs;
  if (!last_block(ws.ticket)) return;
  if (threadIdx.x == 0) {
    double xn2 = 0.0;
    for (unsigned b = 0; b < gridDim.x; ++b) xn2 += ws.part[b];
    ws.ticket[0] = 0;
    if (c + 1 < n) {
      double alpha = ws.xtmp[c + 1];
      double tau = 0.0, scale = 0.0, beta = alpha;
      if (xn2 > 0.0) {
        double xn = sqrt(xn2);
        beta = -copysign(sqrt(alpha * alpha + xn * xn), alpha);
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      ws.e[c] = beta;
      ws.e2[c] = beta * beta;
      ws.tau[c] = tau;
      ws.sc[S_TAU] = tau;
      ws.sc[S_SCALE] = scale;
    }
  }
}

// ---------------------------------------------------------------- k_symv
// y[r] = sum_{s>c} A[r][s] v[s] for rows r > c (one wave per row, 4 rows per
// wave), with v[c+1] = 1 and v[s] = xtmp[s]*scale.  Writes v into V[j] and into
// A[c][r] for the block's own rows, and partial dots V_q.v, W_q.v (q < j).
constexpr int kSymvRows = 16;  // rows per block
__global__ __launch_bounds__(256) void k_symv(EigWs ws, int32_t n, int32_t c, int32_t j) {
  __shared__ double vred[kSymvRows][2 * kNB];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double scale = ws.sc[S_SCALE];
  const int32_t r0 = c + 1 + blockIdx.x * kSymvRows;
  const double *xt = ws.xtmp;
  for (int rr = 0; rr < kSymvRows / 4; ++rr) {
    const int lr = wv * (kSymvRows / 4) + rr;
    const int32_t r = r0 + lr;
    if (r >= n) continue;
    const double *row = ws.a + (int64_t)r * n;
    double acc = 0.0;
    for (int32_t s = c + 1 + lane; s < n; s += 64) {
      double v = (s == c + 1) ? 1.0 : xt[s] * scale;
      acc += row[s] * v;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    const double vr = (r == c + 1) ? 1.0 : xt[r] * scale;
    if (lane == 0) {
      ws.y[r] = acc;
      ws.V[(int64_t)j * n + r] = vr;
      ws.a[(int64_t)c * n + r] = vr;
    }
    if (lane < j) {
      vred[lr][lane] = ws.V[(int64_t)lane * n + r] * vr;
      vred[lr][kNB + lane] = ws.W[(int64_t)lane * n + r] * vr;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * kNB) {
    const int q = threadIdx.x & (kNB - 1), isw = threadIdx.x >= kNB;
    double s = 0.0;
    if (q < j)
      for (int lr = 0; lr < kSymvRows; ++lr)
        if (r0 + lr < n) s += vred[lr][isw * kNB + q];
    ws.part[(int64_t)blockIdx.x * (2 * kNB) + threadIdx.x] = s;
  }
  if (!last_block(ws.ticket)) return;
  if (threadIdx.x < 2 * kNB) {
    double s = 0.0;
    for (unsigned b = 0; b < gridDim.x; ++b) s += ws.part[(int64_t)b * (2 * kNB) + threadIdx.x];
    const int q = threadIdx.x & (kNB - 1), isw = threadIdx.x >= kNB;
    ws.sc[(isw ? S_QD : S_PD) + q] = s;
  }
  if (threadIdx.x == 0) ws.ticket[0] = 0;
}

// ---------------------------------------------------------------- k_wvec
__global__ __launch_bounds__(256) void k_wvec(EigWs ws, int32_t n, int32_t c, int32_t j) {
  __shared__ double red[4];
  const int32_t r = c + 1 + blockIdx.x * 256 + threadIdx.x;
  const double tau = ws.sc[S_TAU];
  double part = 0.0;
  if (r < n) {
    double y = ws.y[r];
    for (int q = 0; q < j; ++q)
      y -= ws.V[(int64_t)q * n + r] * ws.sc[S_QD + q] + ws.W[(int64_t)q * n + r] * ws.sc[S_PD + q];
    double w = tau * y;
    ws.wtmp[r] = w;
    part = w * ws.V[(int64_t)j * n + r];
  }
  double bs = block_sum(part, red);
  if (threadIdx.x == 0) ws.part[blockIdx.x] = bs;
  if (!last_block(ws.ticket)) return;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (unsigned b = 0; b < gridDim.x; ++b) s += ws.part[b];
    ws.sc[S_ALPHAW] = -0.5 * tau * s;
    ws.ticket[0] = 0;
  }
}

// W[j] = wtmp + alpha V[j] on rows >= r0 (end of panel)
__global__ void k_wfin(EigWs ws, int32_t n, int32_t r0, int32_t j) {
  const int32_t r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  ws.W[(int64_t)j * n + r] = ws.wtmp[r] + ws.sc[S_ALPHAW] * ws.V[(int64_t)j * n + r];
}

// A[r][s] -= sum_q V[q][r] W[q][s] + W[q][r] V[q][s] for r, s >= k0 (64x64 tiles)
__global__ __launch_bounds__(256) void k_syr2k(EigWs ws, int32_t n, int32_t k0, int32_t nb) {
  __shared__ double sv[kNB][64], sw[kNB][64], tv[kNB][64], tw[kNB][64];
  const int32_t nt = (n - k0 + 63) / 64;
  const int32_t tr = blockIdx.x / nt, tc = blockIdx.x % nt;
  const int32_t rb = k0 + tr * 64, cb = k0 + tc * 64;
  for (int idx = threadIdx.x; idx < nb * 64; idx += 256) {
    int q = idx / 64, o = idx % 64;
    int32_t r = rb + o, s = cb + o;
    sv[q][o] = r < n ? ws.V[(int64_t)q * n + r] : 0.0;
    sw[q][o] = r < n ? ws.W[(int64_t)q * n + r] : 0.0;
    tv[q][o] = s < n ? ws.V[(int64_t)q * n + s] : 0.0;
    tw[q][o] = s < n ? ws.W[(int64_t)q * n + s] : 0.0;
  }
  __syncthreads();
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int q = 0; q < nb; ++q) {
    double vr[4], wr[4], vs[4], wsv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      vr[a] = sv[q][ty + 16 * a];
      wr[a] = sw[q][ty + 16 * a];
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      vs[b] = tv[q][tx + 16 * b];
      wsv[b] = tw[q][tx + 16 * b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += vr[a] * wsv[b] + wr[a] * vs[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int32_t r = rb + ty + 16 * a, s = cb + tx + 16 * b;
      if (r < n && s < n) ws.a[(int64_t)r * n + s] -= acc[a][b];
    }
}

// ---------------------------------------------------------------- bisection
__global__ void k_gersh(EigWs ws, int32_t n) {
  __shared__ double lo[256], hi[256], mx[256];
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
  mx[threadIdx.x] = fmax(emax, 0.0);
  __syncthreads();
  __shared__ double tnr[256];
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

__device__ inline int32_t sturm(const double *__restrict__ d, const double *__restrict__ e2,
                                int32_t n, double x, double pivmin) {
  int32_t cnt = 0;
  double q = d[0] - x;
  if (q <= pivmin) {
    ++cnt;
    q = fmin(q, -pivmin);
  }
  for (int32_t i = 1; i < n; ++i) {
    q = d[i] - e2[i - 1] / q - x;
    if (q <= pivmin) {
      ++cnt;
      q = fmin(q, -pivmin);
    }
  }
  return cnt;
}

// thread k -> k-th smallest eigenvalue
__global__ void k_bisect(EigWs ws, int32_t n) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double lo = ws.sc[S_GL], hi = ws.sc[S_GU];
  const double pivmin = ws.sc[S_PIVMIN];
  const double eps = 2.220446049250313e-16;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    double tol = 2.0 * eps * fmax(fabs(lo), fabs(hi)) + 4.0 * pivmin;
    if (hi - lo <= tol || mid == lo || mid == hi) break;
    if (sturm(ws.d, ws.e2, n, mid, pivmin) > k) hi = mid;
    else lo = mid;
  }
  ws.eval[k] = 0.5 * (lo + hi);
}

// ---------------------------------------------------------------- inverse iteration
// One block.  Vectors kv = 0..nvec-1 for the eigenvalues eval[n-1-kv] (largest
// first).  Thread kv factors T - lambda I (LU, partial pivoting) and solves;
// the block orthogonalises against earlier vectors of the same cluster.
__global__ __launch_bounds__(256) void k_invit(EigWs ws, int32_t n, int32_t nvec) {
  __shared__ double red[4];
  __shared__ double dots[kMaxVec];
  const double tnorm = ws.sc[S_TNORM];
  const double eps = 2.220446049250313e-16;
  const double tiny = eps * tnorm;
  const int tid = threadIdx.x;
  // factor + first solve in parallel, one thread per vector
  if (tid < nvec) {
    const int kv = tid;
    const double lam = ws.eval[n - 1 - kv];
    double *u0 = ws.lu + (int64_t)kv * 5 * n, *u1 = u0 + n, *u2 = u1 + n, *ml = u2 + n,
           *pv = ml + n;
    double acur = ws.d[0] - lam, bcur = (n > 1) ? ws.e[0] : 0.0;
    for (int32_t i = 0; i < n - 1; ++i) {
      double l = ws.e[i];
      double an = ws.d[i + 1] - lam, bn = (i + 1 < n - 1) ? ws.e[i + 1] : 0.0;
      if (fabs(acur) >= fabs(l)) {
        double a0 = acur;
        if (fabs(a0) < tiny) a0 = copysign(tiny, a0 == 0.0 ? 1.0 : a0);
        double m = l / a0;
        u0[i] = a0; u1[i] = bcur; u2[i] = 0.0; ml[i] = m; pv[i] = 0.0;
        acur = an - m * bcur;
        bcur = bn;
      } else {
        double m = acur / l;
        u0[i] = l; u1[i] = an; u2[i] = bn; ml[i] = m; pv[i] = 1.0;
        acur = bcur - m * an;
        bcur = -m * bn;
      }
    }
    if (fabs(acur) < tiny) acur = copysign(tiny, acur == 0.0 ? 1.0 : acur);
    u0[n - 1] = acur;
    u1[n - 1] = 0.0; u2[n - 1] = 0.0; ml[n - 1] = 0.0; pv[n - 1] = 0.0;
    // starting vector: deterministic pseudo-random in (0.5, 1.5)
    double *x = ws.X + (int64_t)kv * n;
    for (int32_t i = 0; i < n; ++i) {
      uint32_t h = (uint32_t)(i + 1) * 2654435761u ^ (uint32_t)(kv + 7) * 40503u;
      h ^= h >> 15; h *= 2246822519u; h ^= h >> 13;
      x[i] = 0.5 + (double)(h & 0xFFFFFF) / 16777216.0;
    }
  }
  __syncthreads();
  for (int it = 0; it < 3; ++it) {
    if (tid < nvec) {
      const int kv = tid;
      double *u0 = ws.lu + (int64_t)kv * 5 * n, *u1 = u0 + n, *u2 = u1 + n, *ml = u2 + n,
             *pv = ml + n;
      double *x = ws.X + (int64_t)kv * n;
      for (int32_t i = 0; i < n - 1; ++i) {
        if (pv[i] != 0.0) {
          double t = x[i]; x[i] = x[i + 1]; x[i + 1] = t;
        }
        x[i + 1] -= ml[i] * x[i];
      }
      x[n - 1] = x[n - 1] / u0[n - 1];
      if (n > 1) x[n - 2] = (x[n - 2] - u1[n - 2] * x[n - 1]) / u0[n - 2];
      for (int32_t i = n - 3; i >= 0; --i)
        x[i] = (x[i] - u1[i] * x[i + 1] - u2[i] * x[i + 2]) / u0[i];
    }
    __syncthreads();
    // orthogonalise in order, then normalise
    for (int kv = 0; kv < nvec; ++kv) {
      double *x = ws.X + (int64_t)kv * n;
      const double lam = ws.eval[n - 1 - kv];
      for (int pv2 = 0; pv2 < kv; ++pv2) {
        const double lp = ws.eval[n - 1 - pv2];
        if (fabs(lam - lp) > 1e-3 * tnorm) continue;
        const double *p = ws.X + (int64_t)pv2 * n;
        double s = 0.0;
        for (int32_t i = tid; i < n; i += 256) s += x[i] * p[i];
        s = block_sum(s, red);
        if (tid == 0) dots[0] = s;
        __syncthreads();
        s = dots[0];
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
// For panel rows c0..c0+nb-1 (Householder vectors in A[c][c+1:], v[c+1] = 1):
//   Z = V^T X  (nb x nvec) and G = V^T V (nb x nb), reduced by the last block,
//   which also forms T (forward, columnwise) and Z <- T Z.
__global__ __launch_bounds__(256) void k_bt_dots(EigWs ws, int32_t n, int32_t c0, int32_t nb,
                                                 int32_t nvec) {
  // each block: 256 rows; partial dots for (nb x (nvec + nb)) pairs
  const int32_t nd = nb * (nvec + nb);
  const int32_t rbase = c0 + 1 + blockIdx.x * 256;
  for (int32_t p = threadIdx.x; p < nd; p += 256) {
    const int q = p / (nvec + nb), o = p % (nvec + nb);
    const int32_t cq = c0 + q;
    double s = 0.0;
    for (int32_t r = rbase; r < min(n, rbase + 256); ++r) {
      double vq = (r <= cq) ? 0.0 : ws.a[(int64_t)cq * n + r];
      if (vq == 0.0) continue;
      double other;
      if (o < nvec) other = ws.X[(int64_t)o * n + r];
      else {
        int32_t co = c0 + (o - nvec);
        other = (r <= co) ? 0.0 : ws.a[(int64_t)co * n + r];
      }
      s += vq * other;
    }
    ws.part[(int64_t)blockIdx.x * nd + p] = s;
  }
  if (!last_block(ws.ticket)) return;
  __shared__ double G[kNB][kNB], T[kNB][kNB], Zs[kNB][kMaxVec];
  for (int32_t p = threadIdx.x; p < nd; p += 256) {
    double s = 0.0;
    for (unsigned b = 0; b < gridDim.x; ++b) s += ws.part[(int64_t)b * nd + p];
    const int q = p / (nvec + nb), o = p % (nvec + nb);
    if (o < nvec) Zs[q][o] = s;
    else G[q][o - nvec] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ws.ticket[0] = 0;
    // dlarft forward/columnwise: T[j][j] = tau_j; T[0:j, j] = -tau_j T[0:j,0:j] G[0:j, j]
    for (int jj = 0; jj < nb; ++jj) {
      const double tj = ws.tau[c0 + jj];
      for (int i = 0; i < jj; ++i) {
        double s = 0.0;
        for (int k = i; k < jj; ++k) s += T[i][k] * G[k][jj];
        T[i][jj] = -tj * s;
      }
      T[jj][jj] = tj;
      for (int i = jj + 1; i < nb; ++i) T[i][jj] = 0.0;
    }
  }
  __syncthreads();
  // Z <- T Z
  for (int32_t p = threadIdx.x; p < nb * nvec; p += 256) {
    const int q = p / nvec, o = p % nvec;
    double s = 0.0;
    for (int k = q; k < nb; ++k) s += T[q][k] * Zs[k][o];
    ws.Z[q * kMaxVec + o] = s;
  }
}

// X[o][r] -= sum_q v_q[r] Z[q][o]
__global__ void k_bt_apply(EigWs ws, int32_t n, int32_t c0, int32_t nb, int32_t nvec) {
  const int32_t r = c0 + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double v[kNB];
  for (int q = 0; q < nb; ++q) v[q] = (r <= c0 + q) ? 0.0 : ws.a[(int64_t)(c0 + q) * n + r];
  for (int o = 0; o < nvec; ++o) {
    double s = 0.0;
    for (int q = 0; q < nb; ++q) s += v[q] * ws.Z[q * kMaxVec + o];
    ws.X[(int64_t)o * n + r] -= s;
  }
}

// normalise + sign rule (largest |.| component positive, first on ties)
__global__ __launch_bounds__(256) void k_finish_vec(EigWs ws, int32_t n) {
  __shared__ double red[4];
  __shared__ double bv[256];
  __shared__ int32_t bi[256];
  double *x = ws.X + (int64_t)blockIdx.x * n;
  double s = 0.0, best = -1.0;
  int32_t besti = 0x7fffffff;
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
  __shared__ double fac;
  if (threadIdx.x == 0) {
    double f = 1.0 / sqrt(s);
    if (x[bi[0]] < 0.0) f = -f;
    fac = f;
  }
  __syncthreads();
  for (int32_t i = threadIdx.x; i < n; i += 256) x[i] *= fac;
}

int eig_run_eigh(eig_ctx *ctx, const double *d_a, int32_t n, int32_t nvec, double *evals,
                 double *evecs) {
  if (n <= 0) return eig_set_err(ctx, EIG_E_ARG, "eigh: n <= 0");
  if (nvec < 0 || nvec > kMaxVec || nvec > n)
    return eig_set_err(ctx, EIG_E_ARG, "eigh: nvec must be in [0, min(n, %d)]", kMaxVec);
  const int64_t N = n;
  const int nblk_max = (int)((N + 255) / 256) + 8;
  const int nsym_max = (int)((N + kSymvRows - 1) / kSymvRows) + 8;
  size_t part_n = std::max<size_t>((size_t)nsym_max * 2 * kNB,
                                   (size_t)nblk_max * kNB * (kMaxVec + kNB));
  size_t need = 0;
  auto add = [&](size_t cnt) {
    size_t o = need;
    need += ((cnt * 8 + 255) / 256) * 256;
    return o;
  };
  size_t o_a = add(N * N), o_d = add(N), o_e = add(N), o_e2 = add(N), o_tau = add(N),
         o_V = add(kNB * N), o_W = add(kNB * N), o_x = add(N), o_w = add(N), o_y = add(N),
         o_sc = add(S_COUNT), o_part = add(part_n), o_tk = add(8), o_ev = add(N),
         o_X = add((size_t)std::max(nvec, 1) * N), o_lu = add((size_t)std::max(nvec, 1) * 5 * N),
         o_Z = add(kNB * kMaxVec);
  EIG_TRY(eig_ensure(ctx, ctx->ework, need));
  char *base = (char *)ctx->ework.p;
  EigWs ws;
  ws.a = (double *)(base + o_a);
  ws.d = (double *)(base + o_d);
  ws.e = (double *)(base + o_e);
  ws.e2 = (double *)(base + o_e2);
  ws.tau = (double *)(base + o_tau);
  ws.V = (double *)(base + o_V);
  ws.W = (double *)(base + o_W);
  ws.xtmp = (double *)(base + o_x);
  ws.wtmp = (double *)(base + o_w);
  ws.y = (double *)(base + o_y);
  ws.sc = (double *)(base + o_sc);
  ws.part = (double *)(base + o_part);
  ws.ticket = (unsigned *)(base + o_tk);
  ws.eval = (double *)(base + o_ev);
  ws.X = (double *)(base + o_X);
  ws.lu = (double *)(base + o_lu);
  ws.Z = (double *)(base + o_Z);
  hipStream_t st = ctx->stream;
  EIG_HIP(ctx, hipMemcpyAsync(ws.a, d_a, sizeof(double) * N * N, hipMemcpyDeviceToDevice, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.ticket, 0, 64, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.sc, 0, sizeof(double) * S_COUNT, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.e, 0, sizeof(double) * N, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.e2, 0, sizeof(double) * N, st));
  EIG_HIP(ctx, hipMemsetAsync(ws.tau, 0, sizeof(double) * N, st));

  eig_time_begin(ctx);
  // ---- tridiagonalisation
  for (int32_t k0 = 0; k0 < n; k0 += kNB) {
    const int32_t nb = std::min<int32_t>(kNB, n - k0);
    for (int32_t j = 0; j < nb; ++j) {
      const int32_t c = k0 + j;
      unsigned gcol = (unsigned)((n - c + 255) / 256);
      hipLaunchKernelGGL(k_col, dim3(gcol), dim3(256), 0, st, ws, n, c, j);
      if (c + 1 >= n) continue;  // last row: only d
      unsigned gsym = (unsigned)((n - c - 1 + kSymvRows - 1) / kSymvRows);
      hipLaunchKernelGGL(k_symv, dim3(gsym), dim3(256), 0, st, ws, n, c, j);
      unsigned gw = (unsigned)((n - c - 1 + 255) / 256);
      hipLaunchKernelGGL(k_wvec, dim3(gw), dim3(256), 0, st, ws, n, c, j);
    }
    const int32_t k1 = k0 + nb;
    if (k1 < n) {
      const int32_t clast = k0 + nb - 1;
      unsigned gf = (unsigned)((n - clast - 1 + 255) / 256);
      hipLaunchKernelGGL(k_wfin, dim3(gf), dim3(256), 0, st, ws, n, clast + 1, nb - 1);
      int32_t nt = (n - k1 + 63) / 64;
      hipLaunchKernelGGL(k_syr2k, dim3((unsigned)(nt * nt)), dim3(256), 0, st, ws, n, k1, nb);
    }
  }
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "tridiag");
  // ---- eigenvalues
  eig_time_begin(ctx);
  hipLaunchKernelGGL(k_gersh, dim3(1), dim3(256), 0, st, ws, n);
  hipLaunchKernelGGL(k_bisect, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, ws, n);
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "bisect");
  // ---- eigenvectors
  eig_time_begin(ctx);
  if (nvec > 0) {
    hipLaunchKernelGGL(k_invit, dim3(1), dim3(256), 0, st, ws, n, nvec);
    // back-transform, last panel first; panel columns c with c+1 < n carry reflectors
    int32_t last = n - 2;  // last reflector row
    if (last >= 0) {
      int32_t p_last = last / kNB;
      for (int32_t p = p_last; p >= 0; --p) {
        const int32_t c0 = p * kNB;
        const int32_t nb = std::min<int32_t>(kNB, last - c0 + 1);
        unsigned g = (unsigned)((n - c0 - 1 + 255) / 256);
        hipLaunchKernelGGL(k_bt_dots, dim3(g), dim3(256), 0, st, ws, n, c0, nb, nvec);
        hipLaunchKernelGGL(k_bt_apply, dim3(g), dim3(256), 0, st, ws, n, c0, nb, nvec);
      }
    }
    hipLaunchKernelGGL(k_finish_vec, dim3((unsigned)nvec), dim3(256), 0, st, ws, n);
    EIG_HIP(ctx, hipGetLastError());
  }
  eig_time_end(ctx, "eigvec");
  // ---- results: descending eigenvalues
  std::vector<double> asc(N);
  EIG_HIP(ctx, hipMemcpyAsync(asc.data(), ws.eval, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  if (nvec > 0 && evecs)
    EIG_HIP(ctx, hipMemcpyAsync(evecs, ws.X, sizeof(double) * N * nvec, hipMemcpyDeviceToHost, st));
  EIG_TRY(eig_ensure(ctx, ctx->evecs_dev, sizeof(double) * N * std::max(nvec, 1)));
  if (nvec > 0)
    EIG_HIP(ctx, hipMemcpyAsync(ctx->evecs_dev.p, ws.X, sizeof(double) * N * nvec,
                                hipMemcpyDeviceToDevice, st));
  EIG_HIP(ctx, hipStreamSynchronize(st));
  for (int64_t i = 0; i < N; ++i) {
    if (std::isnan(asc[i])) return eig_set_err(ctx, EIG_E_CONVERGE, "eigenvalue is NaN");
    if (evals) evals[i] = asc[N - 1 - i];
  }
  ctx->last_nvec = nvec;
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
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  EIG_TRY(eig_ensure(ctx, ctx->scratch, sizeof(double) * (size_t)n * n));
  EIG_HIP(ctx, hipMemcpyAsync(ctx->scratch.p, a, sizeof(double) * (size_t)n * n,
                              hipMemcpyHostToDevice, ctx->stream));
  return eig_run_eigh(ctx, (const double *)ctx->scratch.p, n, nvec, evals, evecs);
}
