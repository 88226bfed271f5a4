// Context, device memory, genotype upload, synthetic generator, sample selection
// and per-SNP column statistics.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "eig_internal.h"

int eig_set_err(eig_ctx *ctx, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

int eig_ensure(eig_ctx *ctx, DevBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.bytes >= bytes) return EIG_OK;
  if (b.p) {
    hipStreamSynchronize(ctx->stream);
    hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return eig_set_err(ctx, EIG_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  b.bytes = bytes;
  return EIG_OK;
}

void eig_time_begin(eig_ctx *ctx) { hipEventRecord(ctx->ev0, ctx->stream); }

void eig_time_end(eig_ctx *ctx, const char *what) {
  hipEventRecord(ctx->ev1, ctx->stream);
  hipEventSynchronize(ctx->ev1);
  float ms = -1.f;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  ctx->ms[what] = ms;
}

static void free_buf(DevBuf &b) {
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// Workspace of eig_lsqproject_noalloc, the entry point for callers that cannot
// allocate device memory: one 64 MiB array per GPU, held until the process ends.
// Host code uses only the device address from hipGetSymbolAddress (eig_open stores
// it in the context); calls that overlap on one GPU share it and race.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kNoAllocBytes];

extern "C" int eig_open(int device, eig_ctx **out) {
  if (!out) return EIG_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return EIG_E_HIP;
  if (device < 0 || device >= n) return EIG_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return EIG_E_HIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return EIG_E_HIP;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    fprintf(stderr, "eig_open: device %d is %s, this library is built for gfx950\n", device,
            prop.gcnArchName);
    return EIG_E_HIP;
  }
  void *ws = nullptr;
  if (hipGetSymbolAddress(&ws, HIP_SYMBOL(scratch)) != hipSuccess || ws == nullptr)
    return EIG_E_HIP;
  eig_ctx *c = new eig_ctx();
  c->device = device;
  c->num_cu = prop.multiProcessorCount;
  c->noalloc = ws;
  c->noalloc_bytes = kNoAllocBytes;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return EIG_E_HIP;
  }
  *out = c;
  return EIG_OK;
}

extern "C" void eig_close(eig_ctx *ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  DevBuf *bufs[] = {&ctx->geno,   &ctx->xindex,  &ctx->sel,        &ctx->c0,       &ctx->c1,
                    &ctx->nmiss,  &ctx->status,  &ctx->mean,       &ctx->fancy,    &ctx->lev,
                    &ctx->levbf,  &ctx->live,    &ctx->xtx,        &ctx->part,     &ctx->ework,
                    &ctx->evecs_dev, &ctx->ffvecs, &ctx->fxscal,   &ctx->projw,    &ctx->proj_out,
                    &ctx->use_dev, &ctx->ignore_dev, &ctx->scratch, &ctx->scratch2,
                    &ctx->selg,   &ctx->selb,   &ctx->fill,     &ctx->evall};
  for (DevBuf *b : bufs) free_buf(*b);
  if (ctx->hostpin.p) hipHostFree(ctx->hostpin.p);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char *eig_last_error(const eig_ctx *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" double eig_last_ms(const eig_ctx *ctx, const char *what) {
  if (!ctx || !what) return -1.0;
  auto it = ctx->ms.find(what);
  return it == ctx->ms.end() ? -1.0 : it->second;
}

static void invalidate(eig_ctx *ctx) {
  ctx->have_sel = ctx->have_selg = ctx->have_selb = ctx->have_stats = ctx->have_xtx = ctx->have_load = false;
}

extern "C" int eig_load_packed(eig_ctx *ctx, const uint8_t *rows, int64_t rlen, int64_t nsnp,
                               int32_t nind) {
  if (!ctx || !rows || nsnp <= 0 || nind <= 0 || rlen < (nind + 3) / 4)
    return eig_set_err(ctx, EIG_E_ARG, "eig_load_packed: bad sizes");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  size_t bytes = (size_t)rlen * (size_t)nsnp;
  EIG_TRY(eig_ensure(ctx, ctx->geno, bytes));
  EIG_HIP(ctx, hipMemcpyAsync(ctx->geno.p, rows, bytes, hipMemcpyHostToDevice, ctx->stream));
  EIG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->rlen = rlen;
  ctx->nsnp = nsnp;
  ctx->nind = nind;
  invalidate(ctx);
  return EIG_OK;
}

// ---- caller memory on the context's device -----------------------------------
extern "C" int eig_dev_alloc(eig_ctx *ctx, size_t bytes, void **ptr) {
  if (!ctx || !ptr || bytes == 0) return eig_set_err(ctx, EIG_E_ARG, "eig_dev_alloc: args");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    *ptr = nullptr;
    return eig_set_err(ctx, EIG_E_NOMEM, "eig_dev_alloc(%zu)", bytes);
  }
  return EIG_OK;
}

extern "C" int eig_dev_free(eig_ctx *ctx, void *ptr) {
  if (!ctx) return EIG_E_ARG;
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  EIG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  EIG_HIP(ctx, hipFree(ptr));
  return EIG_OK;
}

extern "C" int eig_host_alloc(eig_ctx *ctx, size_t bytes, void **ptr) {
  if (!ctx || !ptr || bytes == 0) return eig_set_err(ctx, EIG_E_ARG, "eig_host_alloc: args");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
    *ptr = nullptr;
    return eig_set_err(ctx, EIG_E_NOMEM, "eig_host_alloc(%zu)", bytes);
  }
  return EIG_OK;
}

extern "C" int eig_host_free(eig_ctx *ctx, void *ptr) {
  if (!ctx) return EIG_E_ARG;
  EIG_HIP(ctx, hipHostFree(ptr));
  return EIG_OK;
}

extern "C" int eig_memcpy(eig_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return eig_set_err(ctx, EIG_E_ARG, "eig_memcpy: args");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  EIG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
  EIG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EIG_OK;
}

extern "C" int eig_fetch_packed(eig_ctx *ctx, int64_t m0, int64_t count, uint8_t *rows) {
  if (!ctx || !rows || m0 < 0 || count < 0 || m0 + count > ctx->nsnp)
    return eig_set_err(ctx, EIG_E_ARG, "eig_fetch_packed: range");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  EIG_HIP(ctx, hipMemcpyAsync(rows, (uint8_t *)ctx->geno.p + m0 * ctx->rlen,
                              (size_t)(count * ctx->rlen), hipMemcpyDeviceToHost, ctx->stream));
  EIG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EIG_OK;
}

// ---- synthetic genotypes ------------------------------------------------------
__device__ __host__ inline uint64_t eig_mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

// One thread per output byte (4 samples) of one SNP row.
__global__ void k_synth(uint8_t *__restrict__ geno, int64_t rlen, int64_t nsnp, int32_t nind,
                        const uint32_t *__restrict__ thr, const int32_t *__restrict__ anc,
                        const uint32_t *__restrict__ cum, uint32_t miss_thr, uint64_t seed) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = rlen * nsnp;
  if (idx >= total) return;
  int64_t m = idx / rlen;
  int32_t b = (int32_t)(idx - m * rlen);
  uint32_t byte = 0;
  for (int q = 0; q < 4; ++q) {
    int32_t i = 4 * b + q;
    uint32_t code = 0;
    if (i < nind) {
      uint64_t key = (uint64_t)m * (uint64_t)nind + (uint64_t)i;
      uint64_t h1 = eig_mix64(seed + (key + 1) * 0x9E3779B97F4A7C15ull);
      uint64_t h2 = eig_mix64(h1 ^ 0x5851F42D4C957F2Dull);
      uint64_t h3 = eig_mix64(h2 + 0x2545F4914F6CDD1Dull);
      uint32_t g = 0;
      for (int c = 0; c < 2; ++c) {
        uint32_t ua = (uint32_t)(c ? (h1 >> 32) : h1);
        uint32_t up = (uint32_t)(c ? (h2 >> 32) : h2);
        int32_t pop = up < cum[2 * i] ? anc[3 * i] : (up < cum[2 * i + 1] ? anc[3 * i + 1] : anc[3 * i + 2]);
        g += ua < thr[(int64_t)pop * nsnp + m] ? 1u : 0u;
      }
      code = ((uint32_t)h3 < miss_thr) ? 3u : g;
    }
    byte |= code << (6 - 2 * q);
  }
  geno[idx] = (uint8_t)byte;
}

extern "C" int eig_synth_packed(eig_ctx *ctx, int64_t nsnp, int32_t nind, int32_t npop,
                                const uint32_t *thr, const int32_t *anc, const uint32_t *cum,
                                uint32_t miss_thr, uint64_t seed) {
  if (!ctx || nsnp <= 0 || nind <= 0 || npop <= 0 || !thr || !anc || !cum)
    return eig_set_err(ctx, EIG_E_ARG, "eig_synth_packed: bad args");
  for (int64_t i = 0; i < 3 * (int64_t)nind; ++i)
    if (anc[i] < 0 || anc[i] >= npop) return eig_set_err(ctx, EIG_E_ARG, "anc out of range");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  int64_t rlen = std::max<int64_t>(48, (2 * (int64_t)nind + 7) / 8);
  EIG_TRY(eig_ensure(ctx, ctx->geno, (size_t)(rlen * nsnp)));
  DevBuf &a = ctx->scratch, &b = ctx->scratch2;
  size_t tb = sizeof(uint32_t) * (size_t)npop * nsnp;
  size_t ab = sizeof(int32_t) * 3 * (size_t)nind, cb = sizeof(uint32_t) * 2 * (size_t)nind;
  EIG_TRY(eig_ensure(ctx, a, tb));
  EIG_TRY(eig_ensure(ctx, b, ab + cb + 256));
  EIG_HIP(ctx, hipMemcpyAsync(a.p, thr, tb, hipMemcpyHostToDevice, ctx->stream));
  EIG_HIP(ctx, hipMemcpyAsync(b.p, anc, ab, hipMemcpyHostToDevice, ctx->stream));
  uint32_t *dcum = (uint32_t *)((char *)b.p + ((ab + 255) / 256) * 256);
  EIG_HIP(ctx, hipMemcpyAsync(dcum, cum, cb, hipMemcpyHostToDevice, ctx->stream));
  int64_t total = rlen * nsnp;
  int64_t blocks = (total + 255) / 256;
  hipLaunchKernelGGL(k_synth, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                     (uint8_t *)ctx->geno.p, rlen, nsnp, nind, (const uint32_t *)a.p,
                     (const int32_t *)b.p, (const uint32_t *)dcum, miss_thr, seed);
  EIG_HIP(ctx, hipGetLastError());
  EIG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->rlen = rlen;
  ctx->nsnp = nsnp;
  ctx->nind = nind;
  invalidate(ctx);
  return EIG_OK;
}

// ---- selection: GENO rows -> tile layout --------------------------------------
// A block re-packs up to 8 consecutive SNPs (fewer when 8 rows exceed 64 KB): their
// GENO rows are staged in LDS, then wave w
// builds tiles st = w, w+4, ...; lane (s = lane/8, word = lane%8) forms the 32-bit
// word of 16 selected samples of SNP m0+s (codes LSB-first, padding = 3), so one
// wave store writes the 8 SNPs' 32-byte records of a tile, 256 contiguous bytes.
constexpr int kSelSnps = 8;
__global__ __launch_bounds__(256) void k_select(const uint8_t *__restrict__ geno, int64_t rlen,
                                                int64_t nsnp, const int32_t *__restrict__ xindex,
                                                int32_t nrows, int32_t nst, int32_t nsb,
                                                int32_t ident, uint32_t *__restrict__ sel) {
  extern __shared__ __align__(16) uint8_t rows[];  // nsb (<= kSelSnps) x rlen
  const int64_t m0 = (int64_t)blockIdx.x * nsb;
  const int ns = (int)min((int64_t)nsb, nsnp - m0);
  const uint8_t *src = geno + m0 * rlen;
  const int64_t nbytes = (int64_t)ns * rlen;
  // 16-byte copies when the block's rows start 16-byte aligned (rlen * nsb a multiple
  // of 16, as for the 1,250-byte rows of 5,000 samples), bytes for the rest
  int64_t b16 = 0;
  if (((uintptr_t)src & 15) == 0) {
    b16 = nbytes & ~(int64_t)15;
    const uint4 *s4 = (const uint4 *)src;
    uint4 *r4 = (uint4 *)rows;
    for (int64_t b = threadIdx.x; b < b16 / 16; b += blockDim.x) r4[b] = s4[b];
  }
  for (int64_t b = b16 + threadIdx.x; b < nbytes; b += blockDim.x) rows[b] = src[b];
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = lane >> 3, ww = lane & 7;
  const uint8_t *row = rows + (int64_t)(s < ns ? s : 0) * rlen;
  for (int32_t st = w; st < nst; st += 4) {
    uint32_t word = 0;
    if (ident) {  // xindex[j] == j: the word's 16 samples are 4 consecutive GENO bytes
      const int32_t j0 = st * kTileS + ww * 16;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int32_t jb = j0 + 4 * t;
        const uint32_t byte = jb < nrows ? row[jb >> 2] : 0xFFu;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t code = (jb + i < nrows) ? ((byte >> (6 - 2 * i)) & 3u) : 3u;
          word |= code << (2 * (4 * t + i));
        }
      }
    } else {
      // the word's 16 sample indices are loaded together (clamped; past nrows unused)
      int32_t kk[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) kk[i] = xindex[min(st * kTileS + ww * 16 + i, nrows - 1)];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int32_t j = st * kTileS + ww * 16 + i;
        uint32_t code = 3;
        if (j < nrows) {
          const int32_t k = kk[i];
          code = (row[k >> 2] >> ((3 - (k & 3)) << 1)) & 3u;
        }
        word |= code << (2 * i);
      }
    }
    if (s < ns) sel[((int64_t)st * nsnp + m0 + s) * kTileWords + ww] = word;
  }
}

extern "C" int eig_select(eig_ctx *ctx, const int32_t *xindex, int32_t nrows) {
  if (!ctx || !xindex || nrows <= 0) return eig_set_err(ctx, EIG_E_ARG, "eig_select: args");
  if (!ctx->geno.p || ctx->nsnp <= 0) return eig_set_err(ctx, EIG_E_STATE, "no genotypes");
  if (nrows > ctx->nind) return eig_set_err(ctx, EIG_E_ARG, "nrows > nind");
  int32_t ident = 1;
  for (int32_t i = 0; i < nrows; ++i) {
    if (xindex[i] < 0 || xindex[i] >= ctx->nind || (i > 0 && xindex[i] <= xindex[i - 1]))
      return eig_set_err(ctx, EIG_E_ARG, "xindex must be increasing and in [0,nind)");
    ident &= xindex[i] == i;
  }
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  int32_t nst = (nrows + kTileS - 1) / kTileS;
  EIG_TRY(eig_ensure(ctx, ctx->xindex, sizeof(int32_t) * (size_t)nrows));
  EIG_TRY(eig_ensure(ctx, ctx->sel, (size_t)nst * ctx->nsnp * kTileWords * 4));
  EIG_HIP(ctx, hipMemcpyAsync(ctx->xindex.p, xindex, sizeof(int32_t) * (size_t)nrows,
                              hipMemcpyHostToDevice, ctx->stream));
  eig_time_begin(ctx);
  const int32_t nsb = (int32_t)std::min<int64_t>(kSelSnps, 65536 / ctx->rlen);
  if (nsb < 1) return eig_set_err(ctx, EIG_E_ARG, "eig_select: rows longer than 64 KB");
  hipLaunchKernelGGL(k_select, dim3((unsigned)((ctx->nsnp + nsb - 1) / nsb)), dim3(256),
                     (size_t)ctx->rlen * nsb, ctx->stream, (const uint8_t *)ctx->geno.p,
                     ctx->rlen, ctx->nsnp, (const int32_t *)ctx->xindex.p, nrows, nst, nsb,
                     ident, (uint32_t *)ctx->sel.p);
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "select");
  ctx->nrows = nrows;
  ctx->nst = nst;
  ctx->have_sel = true;
  ctx->have_selg = ctx->have_selb = ctx->have_stats = ctx->have_xtx = ctx->have_load = false;
  return EIG_OK;
}

// ---- column statistics ---------------------------------------------------------
// One thread per SNP; reads the SNP's 32 bytes in every sample tile and counts
// codes with popcounts.  FP64 arithmetic follows fvadjust_binary operation by
// operation (the file is built without FMA contraction).
__global__ void k_snpstats(const uint32_t *__restrict__ sel, int64_t nsnp, int32_t nst,
                           int32_t nrows, const uint8_t *__restrict__ live, int32_t altnorm,
                           int32_t fancynorm, int32_t minallelecnt, int32_t maxmissing,
                           int32_t *__restrict__ c0o, int32_t *__restrict__ c1o,
                           int32_t *__restrict__ nmo, int32_t *__restrict__ sto,
                           double *__restrict__ meano, double *__restrict__ fancyo,
                           double *__restrict__ lev, uint32_t *__restrict__ levbf) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nsnp) return;
  int32_t n1 = 0, n2 = 0, n3 = 0;
  if (live == nullptr || live[m]) {
    for (int32_t st = 0; st < nst; ++st) {
      const uint4 *p = (const uint4 *)(sel + ((int64_t)st * nsnp + m) * kTileWords);
      uint4 a = p[0], b = p[1];
      uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        uint32_t lo = w[k] & 0x55555555u, hi = (w[k] >> 1) & 0x55555555u;
        n1 += __popc(lo & ~hi);
        n2 += __popc(hi & ~lo);
        n3 += __popc(hi & lo);
      }
    }
  }
  double *lv = lev + 4 * m;
  uint32_t *lb = levbf + 4 * m;
  lv[3] = 0.0;
  lb[3] = 0u;
  if (live != nullptr && !live[m]) {
    c0o[m] = c1o[m] = nmo[m] = 0;
    sto[m] = EIG_SNP_DEAD;
    meano[m] = fancyo[m] = 0.0;
    lv[0] = lv[1] = lv[2] = 0.0;
    lb[0] = lb[1] = lb[2] = 0u;
    return;
  }
  int32_t nmiss = n3 - (nst * kTileS - nrows);
  int32_t c0 = n1 + 2 * n2;
  int32_t c1 = 2 * (nrows - nmiss) - c0;
  c0o[m] = c0;
  c1o[m] = c1;
  nmo[m] = nmiss;
  if (nmiss == nrows) {
    sto[m] = EIG_SNP_ALLMISS;
    meano[m] = fancyo[m] = 0.0;
    lv[0] = lv[1] = lv[2] = 0.0;
    lb[0] = lb[1] = lb[2] = 0u;
    return;
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
  int32_t t = c0 < c1 ? c0 : c1;
  int32_t st = ((t < minallelecnt) || (nmiss > maxmissing) || (t == 0)) ? EIG_SNP_DROPPED
                                                                        : EIG_SNP_KEPT;
  sto[m] = st;
  meano[m] = ymean;
  fancyo[m] = yfancy;
  double l0 = (-ymean) * yfancy, l1 = (1.0 - ymean) * yfancy, l2 = (2.0 - ymean) * yfancy;
  if (st != EIG_SNP_KEPT) l0 = l1 = l2 = 0.0;  // dropped SNPs take no part in XTX
  lv[0] = l0;
  lv[1] = l1;
  lv[2] = l2;
  double ls[3] = {l0, l1, l2};
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    uint16_t hi = eig_bf16_rne((float)ls[g]);
    uint16_t lo = eig_bf16_rne((float)(ls[g] - (double)eig_bf16_to_f32(hi)));
    lb[g] = (uint32_t)hi | ((uint32_t)lo << 16);
  }
}

extern "C" int eig_snpstats(eig_ctx *ctx, const uint8_t *live, int32_t altnorm,
                            int32_t fancynorm, int32_t minallelecnt, int32_t maxmissing,
                            int32_t *c0, int32_t *c1, int32_t *nmiss, double *mean,
                            double *fancy, int32_t *status) {
  if (!ctx) return EIG_E_ARG;
  if (!ctx->have_sel) return eig_set_err(ctx, EIG_E_STATE, "eig_snpstats before eig_select");
  EIG_HIP(ctx, hipSetDevice(ctx->device));
  size_t M = (size_t)ctx->nsnp;
  EIG_TRY(eig_ensure(ctx, ctx->c0, 4 * M));
  EIG_TRY(eig_ensure(ctx, ctx->c1, 4 * M));
  EIG_TRY(eig_ensure(ctx, ctx->nmiss, 4 * M));
  EIG_TRY(eig_ensure(ctx, ctx->status, 4 * M));
  EIG_TRY(eig_ensure(ctx, ctx->mean, 8 * M));
  EIG_TRY(eig_ensure(ctx, ctx->fancy, 8 * M));
  EIG_TRY(eig_ensure(ctx, ctx->lev, 32 * M));
  EIG_TRY(eig_ensure(ctx, ctx->levbf, 16 * M));
  const uint8_t *dlive = nullptr;
  if (live) {
    EIG_TRY(eig_ensure(ctx, ctx->live, M));
    EIG_HIP(ctx, hipMemcpyAsync(ctx->live.p, live, M, hipMemcpyHostToDevice, ctx->stream));
    dlive = (const uint8_t *)ctx->live.p;
  }
  ctx->have_live = live != nullptr;
  ctx->altnorm = altnorm;
  ctx->fancynorm = fancynorm;
  eig_time_begin(ctx);
  unsigned blocks = (unsigned)((M + 255) / 256);
  hipLaunchKernelGGL(k_snpstats, dim3(blocks), dim3(256), 0, ctx->stream,
                     (const uint32_t *)ctx->sel.p, ctx->nsnp, ctx->nst, ctx->nrows, dlive,
                     altnorm, fancynorm, minallelecnt, maxmissing, (int32_t *)ctx->c0.p,
                     (int32_t *)ctx->c1.p, (int32_t *)ctx->nmiss.p, (int32_t *)ctx->status.p,
                     (double *)ctx->mean.p, (double *)ctx->fancy.p, (double *)ctx->lev.p,
                     (uint32_t *)ctx->levbf.p);
  EIG_HIP(ctx, hipGetLastError());
  eig_time_end(ctx, "snpstats");
  struct {
    void *h;
    DevBuf *d;
    size_t b;
  } outs[] = {{c0, &ctx->c0, 4 * M},     {c1, &ctx->c1, 4 * M},     {nmiss, &ctx->nmiss, 4 * M},
              {mean, &ctx->mean, 8 * M}, {fancy, &ctx->fancy, 8 * M}, {status, &ctx->status, 4 * M}};
  for (auto &o : outs)
    if (o.h) EIG_HIP(ctx, hipMemcpyAsync(o.h, o.d->p, o.b, hipMemcpyDeviceToHost, ctx->stream));
  EIG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->have_stats = true;
  ctx->have_xtx = ctx->have_load = false;
  return EIG_OK;
}
