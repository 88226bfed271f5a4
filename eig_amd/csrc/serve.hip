// lsqproject serving: a request loop over a device-resident reference panel.
//
// The panel (per SNP: the K projection rows e_m = fxscal .* ffvecs[:,m], the
// fancy scale, -1 for an ignored SNP, and xmean = mean*fancy) is built once on
// the device from a context whose loadings are current.  Requests of 1..max_batch
// query samples arrive as sample-major packed rows (SNP m of a sample at byte m/4,
// bits 7-6, 5-4, 3-2, 1-0 for m%4 = 0..3, code 3 = missing).  A worker thread
// admits pending requests into free lanes: each lane is a stream with its own
// descriptor, row, partial-sum and output memory and a captured graph of three
// persistent kernels.  Requests that arrive while other lanes are on the device
// are batched into the next free lane, so their work overlaps.
//
// Per sample the arithmetic is lsqproj + regressit + solvit (smartpca.c:4606-4757,
// regsubs.c:8-74, linsubs.c:296-393): over the sample's observed, kept SNPs the
// Gram co = sum e_m e_m^T and rhs rr = sum e_m x_m with x = g*fancy - xmean
// (fixxrow, qpsubs.c:338-349), then Cholesky (choldc/cholsl order).  The SNP
// range is cut into fixed chunks of `chunk` SNPs, each summed in SNP order, and
// the chunk sums are added in chunk order: a sample's result depends on its own
// row and the server's chunk only, never on the other samples of its batch.
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "eig_internal.h"

constexpr int kSrvMaxK = 16;
constexpr int kSrvMaxBatch = 1024;  // samples per lane (descriptor capacity)
constexpr int kSrvMaxLanes = 8;

struct SrvDesc {
  int32_t n;
  int32_t pad_[3];
  const uint8_t *rows[kSrvMaxBatch];
};

constexpr int panel_stride(int K) { return (K + 3) & ~1; }  // K + 2 rounded up to even

// ------------------------------------------------------------------ kernels
__global__ void k_srv_panel(const double *__restrict__ ff, const double *__restrict__ fxscal,
                            const double *__restrict__ fancy, const double *__restrict__ xmean,
                            const uint8_t *__restrict__ ignore, int64_t M, int32_t K, int32_t S,
                            double *__restrict__ panel) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  double *p = panel + m * S;
  for (int j = 0; j < K; ++j) p[j] = fxscal[j] * ff[(int64_t)j * M + m];
  p[K] = ignore[m] ? -1.0 : fancy[m];
  p[K + 1] = xmean[m];
  for (int j = K + 2; j < S; ++j) p[j] = 0.0;
}

// Random query rows for the launch-shape tuner (any byte is a valid row).
__global__ void k_srv_fill(uint8_t *__restrict__ rows, int64_t bytes, uint64_t seed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bytes) return;
  uint64_t x = seed + (uint64_t)i * 0x9E3779B97F4A7C15ull;
  x ^= x >> 31;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 29;
  rows[i] = (uint8_t)(x >> 24);
}

// Per (64-sample block, SNP chunk) work item: the Gram upper triangle, the rhs
// and the observed count of each sample over the chunk.  Persistent: a fixed
// grid strides over the n-dependent item count read from the descriptor.
// partial layout: [(c * W + p) * nmax + q].
template <int K, bool kGuard>
__device__ __forceinline__ void srv_group(uint4 w, const double *__restrict__ pm, int64_t mg,
                                          int64_t m1, double *G, double *r, double &kk) {
  // built without contraction: x = g*fancy - xmean rounds as the reference's two
  // operations; the Gram and rhs accumulations are explicit fused multiply-adds
  constexpr int S = panel_stride(K);
#pragma unroll 1
  for (int wi = 0; wi < 4; ++wi) {
    const uint32_t wv = wi == 0 ? w.x : (wi == 1 ? w.y : (wi == 2 ? w.z : w.w));
#pragma unroll 2
    for (int t = 0; t < 16; ++t) {
      const int s = 16 * wi + t;
      if (kGuard && mg + s >= m1) break;
      const uint32_t code = (wv >> (8 * (t >> 2) + 2 * (3 - (t & 3)))) & 3u;
      const double *pe = pm + (int64_t)s * S;
      const double fz = pe[K], xm = pe[K + 1];
      const bool obs = code != 3u && fz >= 0.0;
      const double wobs = obs ? 1.0 : 0.0;
      const double y = (double)code * fz;
      const double xv = obs ? y - xm : 0.0;
      // e * 1 = e and e * 0 = +-0 exactly: the masked row adds nothing to G
      double e[K], ew[K];
#pragma unroll
      for (int a = 0; a < K; ++a) {
        e[a] = pe[a];
        ew[a] = e[a] * wobs;
      }
#pragma unroll
      for (int a = 0; a < K; ++a) r[a] = fma(ew[a], xv, r[a]);
#pragma unroll
      for (int a = 0; a < K; ++a)
#pragma unroll
        for (int c = a; c < K; ++c) {
          double &g = G[a * K - a * (a - 1) / 2 + (c - a)];
          g = fma(ew[a], e[c], g);
        }
      kk += wobs;
    }
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_srv_gram(const SrvDesc *__restrict__ desc,
                                                 const double *__restrict__ panel, int64_t M,
                                                 int32_t chunk, int32_t nchunk, int32_t nmax,
                                                 double *__restrict__ partial) {
  constexpr int S = panel_stride(K);
  constexpr int nP = K * (K + 1) / 2, W = nP + K + 1;
  const int32_t n = desc->n;
  const int32_t nsb = (n + 63) >> 6;
  const int64_t items = (int64_t)nsb * nchunk;
  const int lane = threadIdx.x;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int32_t sb = (int32_t)(it % nsb), c = (int32_t)(it / nsb);
    const int32_t q = sb * 64 + lane;
    const bool act = q < n;
    const uint8_t *row = desc->rows[act ? q : 0];
    double G[nP], r[K], kk = 0.0;
#pragma unroll
    for (int p = 0; p < nP; ++p) G[p] = 0.0;
#pragma unroll
    for (int a = 0; a < K; ++a) r[a] = 0.0;
    const int64_t m0 = (int64_t)c * chunk, m1 = min(M, m0 + chunk);
    for (int64_t mg = m0; mg < m1; mg += 64) {
      uint4 w = make_uint4(~0u, ~0u, ~0u, ~0u);  // inactive lanes read as all missing
      if (act) w = *(const uint4 *)(row + (mg >> 2));
      const double *pm = panel + mg * S;
      if (mg + 64 <= m1)
        srv_group<K, false>(w, pm, mg, m1, G, r, kk);
      else
        srv_group<K, true>(w, pm, mg, m1, G, r, kk);
    }
    if (act) {
      double *out = partial + (int64_t)c * W * nmax + q;
#pragma unroll
      for (int p = 0; p < nP; ++p) out[(int64_t)p * nmax] = G[p];
#pragma unroll
      for (int a = 0; a < K; ++a) out[(int64_t)(nP + a) * nmax] = r[a];
      out[(int64_t)(nP + K) * nmax] = kk;
    }
  }
}

// sums[p * nmax + q] = sum over chunks in chunk order; one thread per (p, q).
__global__ __launch_bounds__(256) void k_srv_reduce(const SrvDesc *__restrict__ desc,
                                                    const double *__restrict__ partial,
                                                    int32_t nchunk, int32_t W, int32_t nmax,
                                                    double *__restrict__ sums) {
  const int32_t n = desc->n;
  const int64_t total = (int64_t)W * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int32_t p = (int32_t)(idx / n), q = (int32_t)(idx % n);
    const double *src = partial + (int64_t)p * nmax + q;
    double s = 0.0;
    for (int32_t c = 0; c < nchunk; ++c) s += src[(int64_t)c * W * nmax];
    sums[(int64_t)p * nmax + q] = s;
  }
}

// Cholesky solve per sample (choldc + cholsl order) on the packed triangle.
// out[q * (K + 2) + j]: coefficients, then status (0 solved, 1 <= K observed,
// 2 Cholesky failed or NaN), then the observed-SNP count.
template <int K>
__global__ __launch_bounds__(64) void k_srv_solve(const SrvDesc *__restrict__ desc,
                                                  const double *__restrict__ sums, int32_t nmax,
                                                  double *__restrict__ out) {
  constexpr int nP = K * (K + 1) / 2;
  const int32_t n = desc->n;
  for (int32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    double A[nP], b[K], p[K], x[K];
#pragma unroll
    for (int i = 0; i < nP; ++i) A[i] = sums[(int64_t)i * nmax + q];
#pragma unroll
    for (int i = 0; i < K; ++i) b[i] = sums[(int64_t)(nP + i) * nmax + q];
    const double kk = sums[(int64_t)(nP + K) * nmax + q];
    double *o = out + (int64_t)q * (K + 2);
    o[K + 1] = kk;
    if (kk <= (double)K) {
#pragma unroll
      for (int j = 0; j < K; ++j) o[j] = 0.0;
      o[K] = 1.0;
      continue;
    }
    bool bad = false;
#pragma unroll
    for (int i = 0; i < nP; ++i) bad |= isnan(A[i]);
#pragma unroll
    for (int i = 0; i < K; ++i) bad |= isnan(b[i]);
    // P(i, j), i <= j, holds the input a[i][j] until step (i, j) overwrites it with L[j][i]
#define SRV_P(i, j) A[(i) * K - (i) * ((i) - 1) / 2 + ((j) - (i))]
#pragma unroll
    for (int i = 0; i < K; ++i) {
#pragma unroll
      for (int j = i; j < K; ++j) {
        double s = SRV_P(i, j);
#pragma unroll
        for (int k = i - 1; k >= 0; --k) s -= SRV_P(k, i) * SRV_P(k, j);
        if (i == j) {
          bad |= !(s > 0.0);
          p[i] = sqrt(s > 0.0 ? s : 1.0);
        } else {
          SRV_P(i, j) = s / p[i];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
      double s = b[i];
#pragma unroll
      for (int k = i - 1; k >= 0; --k) s -= SRV_P(k, i) * x[k];
      x[i] = s / p[i];
    }
#pragma unroll
    for (int i = K - 1; i >= 0; --i) {
      double s = x[i];
#pragma unroll
      for (int k = i + 1; k < K; ++k) s -= SRV_P(i, k) * x[k];
      x[i] = s / p[i];
    }
#undef SRV_P
#pragma unroll
    for (int j = 0; j < K; ++j) o[j] = bad ? 0.0 : x[j];
    o[K] = bad ? 2.0 : 0.0;
  }
}

// ------------------------------------------------------------------ host side
using Clock = std::chrono::steady_clock;

struct SrvReq {
  int64_t id = 0;
  int32_t nq = 0;
  const uint8_t *rows = nullptr;
  int64_t stride = 0;
  bool on_device = false;
  Clock::time_point t_submit, t_done;
  bool done = false;
  int err = EIG_OK;
  std::vector<double> coords;  // nq x K
  std::vector<int32_t> status;
};

struct SrvLane {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipGraphExec_t gexec = nullptr;
  SrvDesc *desc_h = nullptr, *desc_d = nullptr;
  uint8_t *rows_d = nullptr;
  double *partial_d = nullptr, *sums_d = nullptr, *out_d = nullptr, *out_h = nullptr;
  std::vector<std::shared_ptr<SrvReq>> reqs;
  std::vector<int32_t> offs;
  bool busy = false;
};

struct eig_server {
  int device = 0;
  int32_t K = 0, S = 0, nP = 0, W = 0;
  int64_t M = 0, rowbytes = 0, rowpad = 0;
  int32_t nlanes = 0, max_batch = 0;
  int32_t chunk = 0, nchunk = 0, grid_gram = 0, grid_red = 0, grid_solve = 0;
  int num_cu = 0;
  double eigscale[kSrvMaxK];
  double *panel_d = nullptr;
  SrvLane lane[kSrvMaxLanes];

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<SrvReq>> pending;
  std::unordered_map<int64_t, std::shared_ptr<SrvReq>> reqs;
  int64_t next_id = 1;
  bool stopping = false;
  std::thread worker;
  std::string err;
  int64_t nbatches = 0, nsamples = 0;
};

static int srv_err(eig_server *s, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (s) s->err = buf;
  return code;
}

#define SRV_HIP(s, call)                                                                      \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return srv_err((s), EIG_E_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,               \
                     hipGetErrorString(e_));                                                  \
  } while (0)

#define SRV_KCASE(n, ...) \
  case n: {               \
    constexpr int KC = n; \
    __VA_ARGS__;          \
  } break;
#define SRV_KSWITCH(s, k, ...)                                                              \
  switch (k) {                                                                              \
    SRV_KCASE(1, __VA_ARGS__) SRV_KCASE(2, __VA_ARGS__) SRV_KCASE(3, __VA_ARGS__)            \
    SRV_KCASE(4, __VA_ARGS__) SRV_KCASE(5, __VA_ARGS__) SRV_KCASE(6, __VA_ARGS__)            \
    SRV_KCASE(7, __VA_ARGS__) SRV_KCASE(8, __VA_ARGS__) SRV_KCASE(9, __VA_ARGS__)            \
    SRV_KCASE(10, __VA_ARGS__) SRV_KCASE(11, __VA_ARGS__) SRV_KCASE(12, __VA_ARGS__)         \
    SRV_KCASE(13, __VA_ARGS__) SRV_KCASE(14, __VA_ARGS__) SRV_KCASE(15, __VA_ARGS__)         \
    SRV_KCASE(16, __VA_ARGS__)                                                              \
    default:                                                                                \
      return srv_err(s, EIG_E_ARG, "K out of range");                                       \
  }

// The three kernels of one batch on `st`, reading the batch from desc_d.
static int srv_launch(eig_server *s, hipStream_t st, const SrvDesc *desc_d, double *partial_d,
                      double *sums_d, double *out_d, int32_t chunk, int32_t nchunk,
                      int32_t grid_gram) {
  const int32_t nmax = s->max_batch;
  SRV_KSWITCH(s, s->K,
              hipLaunchKernelGGL((k_srv_gram<KC>), dim3(grid_gram), dim3(64), 0, st, desc_d,
                                 (const double *)s->panel_d, s->M, chunk, nchunk, nmax,
                                 partial_d));
  hipLaunchKernelGGL(k_srv_reduce, dim3(s->grid_red), dim3(256), 0, st, desc_d,
                     (const double *)partial_d, nchunk, s->W, nmax, sums_d);
  SRV_KSWITCH(s, s->K,
              hipLaunchKernelGGL((k_srv_solve<KC>), dim3(s->grid_solve), dim3(64), 0, st, desc_d,
                                 (const double *)sums_d, nmax, out_d));
  SRV_HIP(s, hipGetLastError());
  return EIG_OK;
}

// Launch-shape tuner: times (chunk, persistent grid) candidates on random rows at
// a small and a large batch and keeps the fastest weighted sum.  The chunk it
// picks stays fixed for the server's lifetime (results depend on it).
static int srv_autotune(eig_server *s, int32_t chunk_req) {
  static const int32_t kChunks[] = {1024, 2048, 4096, 8192};
  static const int32_t kGridMul[] = {2, 4, 8};
  SrvLane &L = s->lane[0];
  hipStream_t st = L.stream;
  const int32_t ntest[2] = {32, std::min(512, s->max_batch)};
  const double weight[2] = {1.0, 0.25};
  uint8_t *rows = L.rows_d;
  hipLaunchKernelGGL(k_srv_fill, dim3((unsigned)((ntest[1] * s->rowpad + 255) / 256)), dim3(256),
                     0, st, rows, (int64_t)ntest[1] * s->rowpad, (uint64_t)0xE16);
  hipEvent_t e0, e1;
  SRV_HIP(s, hipEventCreate(&e0));
  SRV_HIP(s, hipEventCreate(&e1));
  double best = 1e30;
  int32_t best_c = chunk_req > 0 ? chunk_req : 2048, best_g = 4 * s->num_cu;
  for (int32_t ch : kChunks) {
    if (chunk_req > 0 && ch != chunk_req) continue;
    const int32_t nch = (int32_t)((s->M + ch - 1) / ch);
    if ((int64_t)nch > s->nchunk) continue;  // partial memory sized for the smallest chunk
    for (int32_t gm : kGridMul) {
      const int32_t g = gm * s->num_cu;
      double score = 0.0;
      for (int t = 0; t < 2; ++t) {
        L.desc_h->n = ntest[t];
        for (int32_t q = 0; q < ntest[t]; ++q) L.desc_h->rows[q] = rows + (int64_t)q * s->rowpad;
        SRV_HIP(s, hipMemcpyAsync(L.desc_d, L.desc_h, sizeof(SrvDesc), hipMemcpyHostToDevice, st));
        double tmin = 1e30;
        for (int rep = 0; rep < 3; ++rep) {
          SRV_HIP(s, hipEventRecord(e0, st));
          int rc = srv_launch(s, st, L.desc_d, L.partial_d, L.sums_d, L.out_d, ch, nch, g);
          if (rc != EIG_OK) return rc;
          SRV_HIP(s, hipEventRecord(e1, st));
          SRV_HIP(s, hipEventSynchronize(e1));
          float ms = 0.f;
          SRV_HIP(s, hipEventElapsedTime(&ms, e0, e1));
          if (rep > 0) tmin = std::min(tmin, (double)ms);
        }
        score += weight[t] * tmin;
      }
      if (score < best) {
        best = score;
        best_c = ch;
        best_g = g;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  s->chunk = best_c;
  s->nchunk = (int32_t)((s->M + best_c - 1) / best_c);
  s->grid_gram = best_g;
  return EIG_OK;
}

static int srv_capture(eig_server *s, SrvLane &L) {
  hipGraph_t graph = nullptr;
  SRV_HIP(s, hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
  int rc = srv_launch(s, L.stream, L.desc_d, L.partial_d, L.sums_d, L.out_d, s->chunk, s->nchunk,
                      s->grid_gram);
  hipError_t e = hipStreamEndCapture(L.stream, &graph);
  if (rc != EIG_OK) return rc;
  if (e != hipSuccess) return srv_err(s, EIG_E_HIP, "capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&L.gexec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) return srv_err(s, EIG_E_HIP, "graph: %s", hipGetErrorString(e));
  return EIG_OK;
}

// Launch one batch on a lane (worker thread, no lock held).
static int srv_dispatch(eig_server *s, SrvLane &L) {
  hipStream_t st = L.stream;
  int32_t off = 0;
  L.offs.clear();
  for (auto &r : L.reqs) {
    L.offs.push_back(off);
    if (r->on_device) {
      for (int32_t i = 0; i < r->nq; ++i) L.desc_h->rows[off + i] = r->rows + (int64_t)i * r->stride;
    } else {
      uint8_t *dst = L.rows_d + (int64_t)off * s->rowpad;
      SRV_HIP(s, hipMemcpy2DAsync(dst, (size_t)s->rowpad, r->rows, (size_t)r->stride,
                                  (size_t)s->rowbytes, (size_t)r->nq, hipMemcpyHostToDevice, st));
      for (int32_t i = 0; i < r->nq; ++i) L.desc_h->rows[off + i] = dst + (int64_t)i * s->rowpad;
    }
    off += r->nq;
  }
  L.desc_h->n = off;
  SRV_HIP(s, hipMemcpyAsync(L.desc_d, L.desc_h, 16 + sizeof(void *) * (size_t)off,
                            hipMemcpyHostToDevice, st));
  SRV_HIP(s, hipGraphLaunch(L.gexec, st));
  SRV_HIP(s, hipMemcpyAsync(L.out_h, L.out_d, sizeof(double) * (size_t)off * (s->K + 2),
                            hipMemcpyDeviceToHost, st));
  SRV_HIP(s, hipEventRecord(L.done, st));
  return EIG_OK;
}

// Results of a finished lane (worker thread, lock held).  Within a request, a
// failed Cholesky keeps the previous sample's answer (regsubs.c:51-53, regans
// zeroed per call at smartpca.c:4656); coordinates are eigscale * a
// (smartpca.c:1556-1564).
static void srv_finish(eig_server *s, SrvLane &L, int err) {
  const int K = s->K;
  const auto now = Clock::now();
  for (size_t r = 0; r < L.reqs.size(); ++r) {
    SrvReq &q = *L.reqs[r];
    q.err = err;
    q.coords.assign((size_t)q.nq * K, 0.0);
    q.status.assign((size_t)q.nq, 0);
    double last[kSrvMaxK] = {0.0};
    for (int32_t i = 0; i < q.nq && err == EIG_OK; ++i) {
      const double *o = L.out_h + (size_t)(L.offs[r] + i) * (K + 2);
      const int stt = (int)o[K];
      q.status[i] = stt;
      if (stt == 0)
        for (int j = 0; j < K; ++j) last[j] = o[j];
      if (stt != 1)
        for (int j = 0; j < K; ++j) q.coords[(size_t)i * K + j] = s->eigscale[j] * last[j];
    }
    q.t_done = now;
    q.done = true;
  }
  s->nbatches += 1;
  L.reqs.clear();
  L.busy = false;
}

static void srv_worker(eig_server *s) {
  hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progressed = false;
    // completions
    for (int l = 0; l < s->nlanes; ++l) {
      SrvLane &L = s->lane[l];
      if (!L.busy) continue;
      hipError_t q = hipEventQuery(L.done);
      if (q == hipErrorNotReady) continue;
      srv_finish(s, L, q == hipSuccess ? EIG_OK : EIG_E_HIP);
      if (q != hipSuccess) s->err = hipGetErrorString(q);
      progressed = true;
    }
    if (progressed) s->cv_done.notify_all();
    // admissions: FIFO requests into free lanes, as many as a lane holds
    for (int l = 0; l < s->nlanes && !s->pending.empty(); ++l) {
      SrvLane &L = s->lane[l];
      if (L.busy) continue;
      int32_t tot = 0;
      while (!s->pending.empty() && tot + s->pending.front()->nq <= s->max_batch) {
        tot += s->pending.front()->nq;
        L.reqs.push_back(s->pending.front());
        s->pending.pop_front();
      }
      L.busy = true;
      s->nsamples += tot;
      lk.unlock();
      int rc = srv_dispatch(s, L);
      lk.lock();
      if (rc != EIG_OK) {
        srv_finish(s, L, rc);
        s->cv_done.notify_all();
      }
      progressed = true;
    }
    bool any_busy = false;
    for (int l = 0; l < s->nlanes; ++l) any_busy |= s->lane[l].busy;
    if (s->stopping && !any_busy && s->pending.empty()) break;
    if (progressed) continue;
    if (any_busy)
      s->cv_work.wait_for(lk, std::chrono::microseconds(20));
    else
      s->cv_work.wait(lk, [s] { return s->stopping || !s->pending.empty(); });
  }
}

static void srv_free(eig_server *s) {
  for (int l = 0; l < kSrvMaxLanes; ++l) {
    SrvLane &L = s->lane[l];
    if (L.gexec) hipGraphExecDestroy(L.gexec);
    if (L.done) hipEventDestroy(L.done);
    if (L.stream) hipStreamDestroy(L.stream);
    if (L.desc_h) hipHostFree(L.desc_h);
    if (L.out_h) hipHostFree(L.out_h);
    if (L.desc_d) hipFree(L.desc_d);
    if (L.rows_d) hipFree(L.rows_d);
    if (L.partial_d) hipFree(L.partial_d);
    if (L.sums_d) hipFree(L.sums_d);
    if (L.out_d) hipFree(L.out_d);
  }
  if (s->panel_d) hipFree(s->panel_d);
  delete s;
}

static int srv_alloc_lane(eig_server *s, SrvLane &L, size_t partial_bytes) {
  SRV_HIP(s, hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
  SRV_HIP(s, hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
  SRV_HIP(s, hipHostMalloc((void **)&L.desc_h, sizeof(SrvDesc), hipHostMallocDefault));
  SRV_HIP(s, hipHostMalloc((void **)&L.out_h, sizeof(double) * s->max_batch * (s->K + 2),
                           hipHostMallocDefault));
  SRV_HIP(s, hipMalloc((void **)&L.desc_d, sizeof(SrvDesc)));
  SRV_HIP(s, hipMalloc((void **)&L.rows_d, (size_t)s->max_batch * s->rowpad));
  SRV_HIP(s, hipMalloc((void **)&L.partial_d, partial_bytes));
  SRV_HIP(s, hipMalloc((void **)&L.sums_d, sizeof(double) * s->W * s->max_batch));
  SRV_HIP(s, hipMalloc((void **)&L.out_d, sizeof(double) * s->max_batch * (s->K + 2)));
  std::memset(L.desc_h, 0, sizeof(SrvDesc));
  return EIG_OK;
}

extern "C" int eig_server_start(eig_ctx *ctx, const double *eigscale, int32_t nlanes,
                                int32_t max_batch, int32_t chunk, eig_server **out) {
  if (!ctx || !out) return eig_set_err(ctx, EIG_E_ARG, "eig_server_start: null argument");
  *out = nullptr;
  if (!ctx->have_load)
    return eig_set_err(ctx, EIG_E_STATE, "eig_server_start before eig_loadings");
  if (nlanes <= 0) nlanes = 4;
  if (max_batch <= 0) max_batch = kSrvMaxBatch;
  if (nlanes > kSrvMaxLanes || max_batch > kSrvMaxBatch)
    return eig_set_err(ctx, EIG_E_ARG, "eig_server_start: nlanes <= %d, max_batch <= %d",
                       kSrvMaxLanes, kSrvMaxBatch);
  if (chunk < 0 || chunk % 64 != 0)
    return eig_set_err(ctx, EIG_E_ARG, "eig_server_start: chunk must be a multiple of 64");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  EIG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  eig_server *s = new eig_server();
  s->device = ctx->device;
  s->num_cu = ctx->num_cu;
  s->K = ctx->last_K;
  s->S = panel_stride(s->K);
  s->nP = s->K * (s->K + 1) / 2;
  s->W = s->nP + s->K + 1;
  s->M = ctx->nsnp;
  s->rowbytes = (s->M + 3) / 4;
  s->rowpad = ((s->M + 63) / 64) * 16;
  s->nlanes = nlanes;
  s->max_batch = max_batch;
  s->grid_red = 4 * s->num_cu;
  s->grid_solve = (max_batch + 63) / 64;
  for (int j = 0; j < kSrvMaxK; ++j) s->eigscale[j] = (eigscale && j < s->K) ? eigscale[j] : 1.0;
  // partial sums sized for the smallest chunk the tuner may pick
  const int32_t min_chunk = chunk > 0 ? chunk : 1024;
  s->nchunk = (int32_t)((s->M + min_chunk - 1) / min_chunk);
  auto fail = [&](int rc) {
    ctx->err = s->err;
    srv_free(s);
    return rc;
  };
  if (hipMalloc((void **)&s->panel_d, sizeof(double) * s->S * s->M) != hipSuccess)
    return fail(srv_err(s, EIG_E_NOMEM, "panel allocation"));
  {
    const double *fx = (const double *)ctx->fxscal.p;
    hipLaunchKernelGGL(k_srv_panel, dim3((unsigned)((s->M + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const double *)ctx->ffvecs.p, fx, (const double *)ctx->fancy.p, fx + kMaxK,
                       (const uint8_t *)ctx->ignore_dev.p, s->M, s->K, s->S, s->panel_d);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
      return fail(srv_err(s, EIG_E_HIP, "panel build failed"));
  }
  // lane 0 first with the largest partial buffer (tuner), the rest after tuning
  int rc = srv_alloc_lane(s, s->lane[0], sizeof(double) * s->nchunk * s->W * (size_t)max_batch);
  if (rc != EIG_OK) return fail(rc);
  rc = srv_autotune(s, chunk);
  if (rc != EIG_OK) return fail(rc);
  for (int l = 1; l < nlanes; ++l) {
    rc = srv_alloc_lane(s, s->lane[l], sizeof(double) * s->nchunk * s->W * (size_t)max_batch);
    if (rc != EIG_OK) return fail(rc);
  }
  for (int l = 0; l < nlanes; ++l) {
    rc = srv_capture(s, s->lane[l]);
    if (rc != EIG_OK) return fail(rc);
  }
  s->worker = std::thread(srv_worker, s);
  *out = s;
  return EIG_OK;
}

extern "C" int eig_server_submit(eig_server *s, const uint8_t *rows, int64_t stride, int32_t nq,
                                 int32_t on_device, int64_t *ticket) {
  if (!s || !rows || !ticket) return srv_err(s, EIG_E_ARG, "eig_server_submit: null argument");
  if (nq < 1 || nq > s->max_batch)
    return srv_err(s, EIG_E_ARG, "eig_server_submit: nq must be in [1, %d]", s->max_batch);
  if (on_device) {
    if (stride < s->rowpad || stride % 16 != 0 || ((uintptr_t)rows & 15) != 0)
      return srv_err(s, EIG_E_ARG,
                     "eig_server_submit: device rows need 16-byte alignment and stride >= %lld",
                     (long long)s->rowpad);
  } else if (stride < s->rowbytes) {
    return srv_err(s, EIG_E_ARG, "eig_server_submit: stride < %lld", (long long)s->rowbytes);
  }
  auto r = std::make_shared<SrvReq>();
  r->nq = nq;
  r->rows = rows;
  r->stride = stride;
  r->on_device = on_device != 0;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stopping) return srv_err(s, EIG_E_STATE, "eig_server_submit after stop");
    r->id = s->next_id++;
    s->reqs[r->id] = r;
    s->pending.push_back(r);
  }
  s->cv_work.notify_one();
  *ticket = r->id;
  return EIG_OK;
}

extern "C" int eig_server_poll(eig_server *s, int64_t ticket, int32_t wait, double *coords,
                               int32_t *status, double *latency_ms) {
  if (!s) return EIG_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return srv_err(s, EIG_E_ARG, "eig_server_poll: unknown ticket");
  std::shared_ptr<SrvReq> r = it->second;
  if (!r->done) {
    if (!wait) return EIG_PENDING;
    s->cv_done.wait(lk, [&] { return r->done; });
  }
  s->reqs.erase(it);
  if (r->err != EIG_OK) return srv_err(s, r->err, "request failed: %s", s->err.c_str());
  if (coords) std::memcpy(coords, r->coords.data(), sizeof(double) * r->coords.size());
  if (status) std::memcpy(status, r->status.data(), sizeof(int32_t) * r->status.size());
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  return EIG_OK;
}

extern "C" int eig_server_info(eig_server *s, int32_t *K, int64_t *nsnp, int64_t *rowpad,
                               int32_t *chunk, int32_t *grid, int64_t *batches,
                               int64_t *samples) {
  if (!s) return EIG_E_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  if (K) *K = s->K;
  if (nsnp) *nsnp = s->M;
  if (rowpad) *rowpad = s->rowpad;
  if (chunk) *chunk = s->chunk;
  if (grid) *grid = s->grid_gram;
  if (batches) *batches = s->nbatches;
  if (samples) *samples = s->nsamples;
  return EIG_OK;
}

extern "C" const char *eig_server_error(const eig_server *s) {
  return s ? s->err.c_str() : "null server";
}

extern "C" void eig_server_stop(eig_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  hipSetDevice(s->device);
  for (int l = 0; l < s->nlanes; ++l)
    if (s->lane[l].stream) hipStreamSynchronize(s->lane[l].stream);
  srv_free(s);
}

// ------------------------------------------------------------------ synthetic queries
__device__ __host__ inline uint64_t srv_mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

// One thread per output byte: sample q, SNPs 4b..4b+3 (sample-major rows).
__global__ void k_synth_queries(uint8_t *__restrict__ rows, int64_t stride, int64_t nsnp,
                                int32_t nq, const uint32_t *__restrict__ thr,
                                const int32_t *__restrict__ pop,
                                const uint32_t *__restrict__ miss_thr, uint64_t seed) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= stride * nq) return;
  const int32_t q = (int32_t)(idx / stride);
  const int64_t b = idx - (int64_t)q * stride;
  uint32_t byte = 0;
  for (int k = 0; k < 4; ++k) {
    const int64_t m = 4 * b + k;
    uint32_t code = 3;
    if (m < nsnp) {
      const uint64_t key = (uint64_t)q * (uint64_t)nsnp + (uint64_t)m;
      const uint64_t h1 = srv_mix64(seed + (key + 1) * 0x9E3779B97F4A7C15ull);
      const uint64_t h2 = srv_mix64(h1 ^ 0x5851F42D4C957F2Dull);
      const uint32_t t = thr[(int64_t)pop[q] * nsnp + m];
      const uint32_t g = ((uint32_t)h1 < t ? 1u : 0u) + ((uint32_t)(h1 >> 32) < t ? 1u : 0u);
      code = ((uint32_t)h2 < miss_thr[q]) ? 3u : g;
    }
    byte |= code << (6 - 2 * k);
  }
  rows[idx] = (uint8_t)byte;
}

extern "C" int eig_synth_queries(eig_ctx *ctx, int64_t nsnp, int32_t nq, int32_t npop,
                                 const uint32_t *thr, const int32_t *pop,
                                 const uint32_t *miss_thr, uint64_t seed, uint8_t *rows_dev,
                                 int64_t stride) {
  if (!ctx || nsnp <= 0 || nq <= 0 || npop <= 0 || !thr || !pop || !miss_thr || !rows_dev ||
      stride < (nsnp + 3) / 4)
    return eig_set_err(ctx, EIG_E_ARG, "eig_synth_queries: bad args");
  for (int32_t q = 0; q < nq; ++q)
    if (pop[q] < 0 || pop[q] >= npop) return eig_set_err(ctx, EIG_E_ARG, "pop out of range");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  const size_t tb = sizeof(uint32_t) * (size_t)npop * nsnp;
  const size_t pb = ((sizeof(int32_t) * (size_t)nq + 255) / 256) * 256;
  EIG_TRY(eig_ensure(ctx, ctx->scratch, tb + 2 * pb));
  uint32_t *d_thr = (uint32_t *)ctx->scratch.p;
  int32_t *d_pop = (int32_t *)((char *)ctx->scratch.p + tb);
  uint32_t *d_miss = (uint32_t *)((char *)ctx->scratch.p + tb + pb);
  hipStream_t st = ctx->stream;
  EIG_HIP(ctx, hipMemcpyAsync(d_thr, thr, tb, hipMemcpyHostToDevice, st));
  EIG_HIP(ctx, hipMemcpyAsync(d_pop, pop, sizeof(int32_t) * nq, hipMemcpyHostToDevice, st));
  EIG_HIP(ctx, hipMemcpyAsync(d_miss, miss_thr, sizeof(uint32_t) * nq, hipMemcpyHostToDevice, st));
  const int64_t total = stride * nq;
  hipLaunchKernelGGL(k_synth_queries, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     rows_dev, stride, nsnp, nq, (const uint32_t *)d_thr, (const int32_t *)d_pop,
                     (const uint32_t *)d_miss, seed);
  EIG_HIP(ctx, hipGetLastError());
  EIG_HIP(ctx, hipStreamSynchronize(st));
  return EIG_OK;
}
