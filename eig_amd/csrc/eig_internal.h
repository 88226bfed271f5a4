// Internal definitions shared by the HIP translation units of libeig_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>

#include "../../include/eig_abi.h"

// Sample tile of the selected-genotype layout: 128 samples = 8 x 32-bit words of
// 2-bit codes (sample 16w+i of the tile at bits 2i..2i+1 of word w).
constexpr int kTileS = 128;
constexpr int kTileWords = kTileS / 16;

// Largest K of loadings/lsqproject; the fxscal buffer holds K scales in its first
// kMaxK slots and xmean = mean*fancy after them.
constexpr int kMaxK = 16;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct eig_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  std::map<std::string, double> ms;
  int num_cu = 0;

  // genotypes as loaded (GENO row layout)
  DevBuf geno;
  int64_t rlen = 0, nsnp = 0;
  int32_t nind = 0;

  // selection of axis samples and its tile layout sel[(st*nsnp + m)*8 + w]
  DevBuf xindex, sel;
  int32_t nrows = 0, nst = 0;
  bool have_sel = false;
  // the same codes regrouped per 64-SNP group for the split-BF16 XTX (built on
  // first use after each selection)
  DevBuf selg;
  bool have_selg = false;
  // the split-BF16 SYRK's operand codes: per 64-SNP group, sample and k-slot half,
  // 16 bytes, each the LDS byte offset of a table entry (built on first use)
  DevBuf selb;
  bool have_selb = false;

  // per-SNP statistics (device)
  DevBuf c0, c1, nmiss, status, mean, fancy;
  DevBuf lev;    // double [nsnp][4] : levels of g = 0,1,2 and 0 for missing
  DevBuf levbf;  // uint32 [nsnp][4] : bf16 hi (low half) | bf16 lo (high half)
  DevBuf live;   // uint8 [nsnp]
  bool have_live = false;          // live mask of the last eig_snpstats
  int32_t altnorm = 1, fancynorm = 1;  // normalisation of the last eig_snpstats
  bool have_stats = false;
  // lsqproject fill table of usepopsformissing: [nsnp][npop] doubles, then nind types
  DevBuf fill;

  // XTX (normalised, row-major nrows x nrows) and partial sums
  DevBuf xtx, part;
  bool have_xtx = false;

  // eigensolver
  DevBuf ework;
  DevBuf evecs_dev;  // nvec x nrows of the last eig_eigh_xtx
  DevBuf evall;      // all-vector eigensolves: every batch's tridiagonal-basis vectors
  int32_t last_nvec = 0;

  // loadings / projection
  DevBuf ffvecs, fxscal, projw, proj_out, use_dev, ignore_dev;
  int32_t last_K = 0;
  // device address of this GPU's static workspace of eig_lsqproject_noalloc (set by
  // eig_open through hipGetSymbolAddress; never the array's host-side name)
  void *noalloc = nullptr;
  size_t noalloc_bytes = 0;
  bool have_load = false;

  // generic scratch
  DevBuf scratch, scratch2, hostpin;
};

int eig_set_err(eig_ctx *ctx, int code, const char *fmt, ...);
int eig_ensure(eig_ctx *ctx, DevBuf &b, size_t bytes);
void eig_time_begin(eig_ctx *ctx);
void eig_time_end(eig_ctx *ctx, const char *what);

#define EIG_HIP(ctx, call)                                                        \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return eig_set_err((ctx), EIG_E_HIP, "%s:%d %s: %s", __FILE__, __LINE__,    \
                         #call, hipGetErrorString(e_));                           \
  } while (0)

#define EIG_TRY(expr)           \
  do {                          \
    int r_ = (expr);            \
    if (r_ != EIG_OK) return r_; \
  } while (0)

// ---- device helpers ---------------------------------------------------------
__device__ __host__ inline uint16_t eig_bf16_rne(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}

__device__ __host__ inline float eig_bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// internal launchers implemented in the kernel translation units
int eig_run_syrk(eig_ctx *ctx, int form, int bf16_block, double *trace_scale);
// builds ctx->selg: per sample tile st, 64-SNP group g and sample i, 16 bytes at
// [(st * ceil(M/64) + g) * 128 + i]; word 2h + (s >> 1), bits 16 (s & 1) + 2 j hold the
// code of SNP 16 s + k of the group, h = (k >> 2) & 1, j = (k & 3) | ((k >> 3) << 2)
constexpr int kGrp = 64;  // SNPs per group of the grouped code layout
int eig_ensure_selg(eig_ctx *ctx);
// internal: a helper's own small problem -- the context's published eigenvectors
// (used by the loadings) and its cached tridiagonal graph are left untouched
int eig_run_eigh(eig_ctx *ctx, const double *d_a, int32_t n, int32_t nvec, double *evals,
                 double *evecs,
                 bool internal = false);

// Column-major FP64 GEMM on MFMA with dgemm's conventions, batched by strides:
// C_b = alpha op(A_b) op(B_b) + beta C_b (C is not read when beta == 0).  lower: only
// the 64 x 64 tiles on or below the diagonal are formed (m == n).  Launched on `st`.
int eig_gemm_f64(eig_ctx *ctx, hipStream_t st, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                 double alpha, const double *A, int64_t lda, int64_t sA, const double *B,
                 int64_t ldb, int64_t sB, double beta, double *C, int64_t ldc, int64_t sC,
                 int32_t batch, bool lower = false);

// lsqproject's population fill table (usepopsformissing) from the loaded genotypes and
// the current ignore mask: types itype[q] (nind host entries, < 0 = none), npop types,
// use[q] the individuals projected, K the axes.  *fill = device table [nsnp][npop] (NaN:
// no data), *d_itype = the types on the device.
int eig_lsq_filltab(eig_ctx *ctx, const int32_t *use, const int32_t *itype, int32_t npop,
                    int32_t K, double **fill, const int32_t **d_itype);
