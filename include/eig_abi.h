/*
 * eig_abi.h -- C-ABI of the MI355X smartpca engine (libeig_amd.so).
 *
 * Plain C: pointers, sizes and error codes only.  Every entry point returns
 * EIG_OK (0) or a negative EIG_E_* code; eig_last_error() gives the message.
 * A context is used from one host thread at a time; all device work of a
 * context runs on its own HIP stream (default priority).
 *
 * The reference (EIGENSOFT 8.0 smartpca) has no plugin or FFI API: its only
 * public contract is the `smartpca -p parfile` process.  The entry points below
 * replace the internal seams of that process; each cites the reference code it
 * stands in for.  Genotype codes follow the reference: 0/1/2 = count of the
 * first allele, 3 = missing (admutils.c:575-592).
 */
#ifndef EIG_ABI_H
#define EIG_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EIG_OK 0
#define EIG_E_ARG (-1)     /* bad argument / size                                  */
#define EIG_E_HIP (-2)     /* HIP runtime error                                    */
#define EIG_E_NOMEM (-3)   /* device or host allocation failed                     */
#define EIG_E_NUMERIC (-4) /* NaN or non-positive XTX trace (smartpca.c:1230-1235) */
#define EIG_E_STATE (-5)   /* call order (e.g. eig_xtx before eig_select)          */
#define EIG_E_CONVERGE (-6)/* eigensolver did not converge                          */

/* XTX accumulation forms */
#define EIG_FORM_FP64 0   /* FP64 end to end (MFMA f64)                             */
#define EIG_FORM_BF16X3 1 /* split-BF16 hi+lo, 3 MFMA products (hi.hi, hi.lo, lo.hi),
                             FP32 per SNP block, FP64 across blocks                 */
#define EIG_FORM_BF16X4 2 /* as BF16X3 plus the lo.lo product                       */

/* per-SNP status written by eig_snpstats */
#define EIG_SNP_KEPT 0
#define EIG_SNP_DROPPED 1  /* min(c0,c1) < minallelecnt, == 0, or nmiss > maxmissing */
#define EIG_SNP_ALLMISS 2  /* no observed genotype among the axis samples           */
#define EIG_SNP_DEAD 3     /* not live in this pass (caller's mask)                 */

typedef struct eig_ctx eig_ctx;

/* Open a context on HIP device `device`. */
int eig_open(int device, eig_ctx **out);
void eig_close(eig_ctx *ctx);
const char *eig_last_error(const eig_ctx *ctx);
/* Wall time in ms of the last device phase named by `what` ("xtx", "eigh",
 * "tridiag", "syrk", ...), measured with HIP events on the context stream; -1 if
 * not measured. */
double eig_last_ms(const eig_ctx *ctx, const char *what);

/* Device and pinned host memory on the context's device, and a synchronous copy
 * between any two of them (direction from the addresses).  Plumbing for callers
 * that hold no other HIP runtime: query pools for the server, staging buffers.
 * No reference counterpart (the reference has no device). */
int eig_dev_alloc(eig_ctx *ctx, size_t bytes, void **ptr);
int eig_dev_free(eig_ctx *ctx, void *ptr);
int eig_host_alloc(eig_ctx *ctx, size_t bytes, void **ptr);
int eig_host_free(eig_ctx *ctx, void *ptr);
int eig_memcpy(eig_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Copy packed genotype rows (GENO layout: nsnp records of rlen bytes, sample k in
 * byte k/4 bits 7-6,5-4,3-2,1-0) to the device.  Replaces inpack's in-memory
 * `packgenos` buffer (mcio.c:2769-2879).  rlen >= ceil(nind/4). */
int eig_load_packed(eig_ctx *ctx, const uint8_t *rows, int64_t rlen, int64_t nsnp,
                    int32_t nind);

/* Generate seeded synthetic genotypes directly in device memory (Balding-Nichols
 * frequencies prepared by the caller).  thr[pop*nsnp + m] is the allele-frequency
 * threshold p*2^32 of population pop at SNP m; sample i draws each allele copy
 * from population anc[3i+c] with cumulative ancestry weights cum[2i], cum[2i+1]
 * (uint32 fractions of 2^32); missing when a third draw falls below miss_thr.
 * The same integer recipe is restated in eig_amd/synth.py for CPU checks. */
int eig_synth_packed(eig_ctx *ctx, int64_t nsnp, int32_t nind, int32_t npop,
                     const uint32_t *thr, const int32_t *anc, const uint32_t *cum,
                     uint32_t miss_thr, uint64_t seed);
/* Copy device genotype rows [m0, m0+count) back to the host (rlen bytes each). */
int eig_fetch_packed(eig_ctx *ctx, int64_t m0, int64_t count, uint8_t *rows);

/* Choose the axis samples of a pass: xindex[0..nrows) indexes the nind samples of
 * the loaded rows, in increasing order (loadindx, qpsubs.c:202-219).  Builds the
 * device tile layout the column statistics and XTX kernels read. */
int eig_select(eig_ctx *ctx, const int32_t *xindex, int32_t nrows);

/* Column statistics of every SNP over the selected samples: getcolxz_binary1 +
 * fvadjust_binary (smartpca.c:3219-3301, 2282-2313) and the keep rule
 * (smartpca.c:1131-1144).  live[m] == 0 marks a SNP already removed (status
 * EIG_SNP_DEAD, contributes nothing); live may be NULL.  Output arrays have nsnp
 * entries and may each be NULL.  Leaves the per-SNP level tables on the device
 * for eig_xtx / eig_loadings / eig_lsqproject. */
int eig_snpstats(eig_ctx *ctx, const uint8_t *live, int32_t altnorm, int32_t fancynorm,
                 int32_t minallelecnt, int32_t maxmissing, int32_t *c0, int32_t *c1,
                 int32_t *nmiss, double *mean, double *fancy, int32_t *status);

/* XTX over the kept SNPs and selected samples, then symmetric unpack and division
 * by trace/(nrows-1): domult_increment_lookup + symit2 + trace scaling
 * (smartpca.c:3426-3495, 480-508, 1222-1236).  The normalised matrix stays on the
 * device for eig_eigh_xtx; xtx_out (nrows*nrows, row-major) may be NULL.
 * *trace_scale receives trace/(nrows-1) before scaling.  bf16_block is the FP32
 * accumulation length in SNPs for the split-BF16 forms (multiple of 16; 0 = 4096). */
int eig_xtx(eig_ctx *ctx, int32_t form, int32_t bf16_block, double *trace_scale,
            double *xtx_out);

/* The non-lookup column path: XTX from explicit FP64 columns, for the reference's
 * usepopsformissing, weightname and ldregress options (getcolxz + fvadjust,
 * smartpca.c:3129-3216, 2236-2279; weights 1150-1152; ldregx + bumpldvv 2691-2826,
 * 1155-1170; domult_increment_normal 3498-3561; symit2 + trace scaling 480-508,
 * 1222-1236).  Call after eig_snpstats (its live mask and normalisation are reused);
 * replaces that call's statistics with this path's:
 *   xtypes (nrows, may be NULL): population index in [0, npop) of each selected
 *     sample -- a missing genotype takes its population's mean genotype at that SNP
 *     when the population has one (usepopsformissing: YES);
 *   weight (nsnp, may be NULL): each SNP's column is scaled by its weight;
 *   ldregress > 0: each column is replaced by its least-squares residual on the
 *     preceding kept SNPs of its chromosome in the reference's buffer of at most
 *     ldregress SNPs within ldposlimit bases (ppos) and ldlimit Morgans (gpos),
 *     using the sources whose squared correlation exceeds ldr2lo; a SNP whose squared
 *     correlation with one of them exceeds ldr2hi contributes nothing.  chrom, ppos,
 *     gpos: nsnp entries (required when ldregress > 0).
 * The keep rule counts missing genotypes after the fill.  c0/c1/nmiss/status (nsnp,
 * each may be NULL) as eig_snpstats; the per-SNP mean and normalisation kept on the
 * device (used by eig_loadings and eig_lsqproject) are the filled column's, the levels
 * used by eig_loadings the plain column's (getcolxf, smartpca.c:3564-3595).
 * *ynumsnps: the GRM's SNP count (kept SNPs, or with ldregress the sum of residual to
 * column squared norms, smartpca.c:1157-1162).  Always FP64. */
int eig_xtx_normal(eig_ctx *ctx, const int32_t *xtypes, int32_t npop, const double *weight,
                   int32_t ldregress, const int32_t *chrom, const double *ppos, const double *gpos,
                   int32_t ldposlimit, double ldlimit, double ldr2lo, double ldr2hi,
                   int32_t minallelecnt, int32_t maxmissing, int32_t *c0, int32_t *c1,
                   int32_t *nmiss, int32_t *status, double *trace_scale, double *ynumsnps,
                   double *xtx_out);

/* Allele counts per SNP and population over the selected samples, for the Fst
 * statistics (fstcolyy + getrawcolx, qpsubs.c:1205-1262, 251-283): counts[(m * npop + t)
 * * 2] = sum of g and counts[... + 1] = sum of 2 - g over the observed samples of
 * population t = xtypes[i] (nrows entries; < 0 = none).  On a SNP with xsnp[m] != 0
 * a sample with haploid[i] != 0 counts g/2 and 1 - g/2, and its heterozygous call is
 * skipped.  haploid and xsnp may be NULL.  counts: nsnp * npop * 2 entries. */
int eig_popcounts(eig_ctx *ctx, const int32_t *xtypes, int32_t npop, const uint8_t *haploid,
                  const uint8_t *xsnp, int32_t *counts);

/* Eigen-decomposition of the device-resident normalised XTX: eigvecs
 * (eigsubs.c:39-55) without its LAPACK call.  evals (nrows, descending) get every
 * eigenvalue; evecs (nvec x nrows, row j = eigenvector j, unit norm) the top
 * nvec eigenvectors, each with its largest-|.| component made positive. */
int eig_eigh_xtx(eig_ctx *ctx, int32_t nvec, double *evals, double *evecs);

/* Same, for a caller's symmetric row-major n x n matrix a (host memory). */
int eig_eigh(eig_ctx *ctx, const double *a, int32_t n, int32_t nvec, double *evals,
             double *evecs);

/* SNP loadings and projection scales from the top-K eigenvectors of the last
 * eig_eigh_xtx: ffvecs[j][m] = sum_i 10 v_j[i] z_m[i] over SNPs with ignore[m]==0
 * (setfvecs qpsubs.c:489-508, smartpca.c:1438-1494, getcolxf 3564-3595), and
 * fxscal_j = 1/sqrt(sum_i <x_i, ffvecs_j>^2) over the selected samples
 * (smartpca.c:1504-1525).  ignore has nsnp entries (NULL = use status != KEPT).
 * evecs: K x nrows host array.  ffvecs (K x nsnp) and fxscal (K) may be NULL. */
int eig_loadings(eig_ctx *ctx, const double *evecs, int32_t K, const uint8_t *ignore,
                 double *ffvecs, double *fxscal);

/* lsqproject of every individual with use[q] != 0 (q over the nind loaded
 * samples): lsqproj + regressit + solvit (smartpca.c:4606-4757, regsubs.c:8-74,
 * linsubs.c:296-393).  acoeffs / bcoeffs: K x nind (row j at j*nind).  status[q]:
 * 0 solved, 1 ignored (<= K observed SNPs, coefficients untouched), 2 Cholesky
 * failed (previous individual's answer kept, as the reference does), 3 not used. */
int eig_lsqproject(eig_ctx *ctx, const int32_t *use, double *acoeffs, double *bcoeffs,
                   int32_t *status);

/* eig_lsqproject with usepopsformissing's fill (lsqproj, smartpca.c:4655-4745): itype[q]
 * (nind) is individual q's population type (< 0: none).  A missing genotype of an
 * individual with a type takes the mean genotype of that type at the SNP, over the
 * individuals in use with that type (those with <= K observed genotypes excluded),
 * when there is one; such an individual is also ignored (status 1) when its observed
 * genotypes alone number <= K. */
int eig_lsqproject_fill(eig_ctx *ctx, const int32_t *use, const int32_t *itype, int32_t npop,
                        double *acoeffs, double *bcoeffs, int32_t *status);

/* eig_lsqproject that allocates nothing: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it) once the context is set up.  Same inputs, outputs,
 * status codes and bits as eig_lsqproject; its partial sums live in a static 64 MiB
 * device array per GPU (loadings and genotypes stay where the earlier calls put
 * them).  Returns EIG_E_ARG when nind x (K(K+1)/2 + K + 2) doubles plus the outputs do
 * not fit it.  Every call on one GPU uses the same array: calls that overlap on a GPU
 * race, from any contexts or threads. */
int eig_lsqproject_noalloc(eig_ctx *ctx, const int32_t *use, double *acoeffs, double *bcoeffs,
                           int32_t *status);

/* fastmode: randomized block-Krylov PCA (kjg_fpca, ksrc/kjg_fpca.c:24-101) of the
 * selected samples over every SNP not marked dead in the last eig_snpstats (the
 * keep rule does not apply, as in the reference), normalised as setgval does
 * (gval.c:31-87: ((g - xmean/xfancy) * xfancy) / sqrt(2), missing 0).  K
 * eigenpairs from L = fastdim > K columns and I = fastiter > 0 iterations,
 * (I+1)*L <= 256.  omega: nrows x L host Gaussian start panel, or NULL for the
 * built-in generator seeded by `seed` (the reference draws it from GSL's default
 * generator, kjg_gsl.c:96-186).  evals (K): S^2 / ncols (kjg_fpca.c:91-94);
 * evecs (K x nrows, may be NULL): unit norm, largest-|.| component positive. */
int eig_fpca(eig_ctx *ctx, int32_t K, int32_t L, int32_t I, uint64_t seed, const double *omega,
             double *evals, double *evecs);

/* eig_fpca with the panel-product form chosen: EIG_FORM_FP64 (as eig_fpca) or
 * EIG_FORM_BF16X4 -- genotype levels and the dense panels split into bf16 hi + lo,
 * four MFMA products in FP32 within 512-sample / 512-SNP blocks, FP64 across blocks
 * (L <= 64).  Same kjg_fpca call (ksrc/kjg_fpca.c:24-101); only the precision of the
 * products X A and X^T B (kjg_fpca.c:104-178) differs. */
int eig_fpca_form(eig_ctx *ctx, int32_t K, int32_t L, int32_t I, uint64_t seed,
                  const double *omega, int32_t form, double *evals, double *evecs);

/* ---- lsqproject serving ------------------------------------------------------
 * The reference projects samples once per run, inside lsqproj (smartpca.c:
 * 4606-4757); the server keeps that step's inputs (loadings, fxscal, fancy, xmean,
 * kept-SNP mask, eigscale) resident and answers projection requests as they
 * arrive.  Query rows are sample-major: SNP m of a sample at byte m/4, bits 7-6,
 * 5-4, 3-2, 1-0 for m%4 = 0..3 (the GENO bit order), code 3 = missing, over the
 * same nsnp SNPs as the panel.  Per sample the result is the reference's
 * lsqproj + regressit + solvit over its observed kept SNPs, times eigscale
 * (smartpca.c:1556-1564).  status per sample: 0 solved, 1 ignored (<= K observed
 * SNPs; coordinates 0), 2 Cholesky failed (the previous sample of the same
 * request's answer, as regsubs.c:51-53).  A sample's result depends only on its
 * own row: it is the same alone and batched with any other requests. */
#define EIG_PENDING 1 /* eig_server_poll: not finished yet */
typedef struct eig_server eig_server;

/* Start a server on the context's device from its current loadings (after
 * eig_loadings with K eigenvectors).  eigscale: K factors (NULL = 1).  nlanes:
 * batches in flight at once (0 = 4, at most 8); max_batch: samples per batch
 * (0 = 1024, at most 1024) -- together they bound the device memory admitted
 * requests may use.  chunk: SNPs per partial sum (multiple of 64; 0 = chosen by
 * the built-in launch-shape tuner at start-up).  The context may be used or
 * closed afterwards; the server owns its copies. */
int eig_server_start(eig_ctx *ctx, const double *eigscale, int32_t nlanes, int32_t max_batch,
                     int32_t chunk, eig_server **out);
/* Queue nq (1..max_batch) query rows of `stride` bytes.  on_device != 0: rows is
 * device memory (16-byte aligned, stride a multiple of 16 and >= the padded row
 * size eig_server_info reports); it is read in place and must stay valid and
 * unchanged until the request is polled.  on_device == 0: host memory (stride >=
 * ceil(nsnp/4)), copied to the device when the request is admitted and read
 * until then.  Thread-safe. */
int eig_server_submit(eig_server *s, const uint8_t *rows, int64_t stride, int32_t nq,
                      int32_t on_device, int64_t *ticket);
/* Result of a ticket: EIG_OK with coords (nq x K, row per sample), status (nq) and
 * the submit-to-completion latency; EIG_PENDING if not done and wait == 0 (wait
 * != 0 blocks).  A finished ticket is released by the poll that returns it. */
int eig_server_poll(eig_server *s, int64_t ticket, int32_t wait, double *coords,
                    int32_t *status, double *latency_ms);
/* Shape and counters: K, nsnp, padded device row bytes, the chunk and persistent
 * grid in use, batches launched and samples admitted so far.  Any may be NULL. */
int eig_server_info(eig_server *s, int32_t *K, int64_t *nsnp, int64_t *rowpad, int32_t *chunk,
                    int32_t *grid, int64_t *batches, int64_t *samples);
const char *eig_server_error(const eig_server *s);
/* Finish queued and running requests, then free the server. */
void eig_server_stop(eig_server *s);

/* Seeded synthetic query rows written to device memory rows_dev (nq rows of
 * `stride` bytes, sample-major as above): sample q draws both allele copies from
 * population pop[q] (thr as for eig_synth_packed) and is missing at a SNP when a
 * draw falls below miss_thr[q].  Restated in eig_amd/synth.py for CPU checks. */
int eig_synth_queries(eig_ctx *ctx, int64_t nsnp, int32_t nq, int32_t npop, const uint32_t *thr,
                      const int32_t *pop, const uint32_t *miss_thr, uint64_t seed,
                      uint8_t *rows_dev, int64_t stride);

#ifdef __cplusplus
}
#endif
#endif /* EIG_ABI_H */
