/*
 * smartpca_main.c -- `smartpca -p parfile`, the drop-in for EIGENSOFT's smartpca
 * (src/eigensrc/smartpca.c) on one MI355X.
 *
 * Same parfile keys, inputs and outputs as the reference (readcommands,
 * smartpca.c:2055-2233; main, smartpca.c:537-1591).  Host code does what the
 * reference's main does around its hot loops -- reading, filtering, the outlier
 * decisions, formatting -- and every pass over genotypes or the N x N matrix
 * runs on the GPU through the C-ABI in eig_abi.h: column statistics, XTX,
 * eigensolve, loadings, projection scales and lsqproject.
 *
 * Extra keys (absent from the reference, ignored by it):
 *   xtxform:  fp64 (default) | bf16x4 | bf16x3   XTX accumulation form
 *   bf16block: FP32 accumulation length in SNPs for the split-BF16 forms (256)
 *   gpu:      HIP device index (0)
 */
#include <ctype.h>
#include <getopt.h>
#include <math.h>
#include <time.h>
#include <unistd.h>

#include "smartpca.h"
#include "stats.h"

void die(const char *fmt, ...) {
  va_list ap;
  fflush(stdout);
  fprintf(stderr, "fatalx:\n");
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fprintf(stderr, "\n");
  exit(1);
}

#define CHK(call)                                                            \
  do {                                                                       \
    int rc_ = (call);                                                        \
    if (rc_ != EIG_OK) die("%s failed (%d): %s", #call, rc_, eig_last_error(ctx)); \
  } while (0)

typedef struct {
  char *genotypename, *snpname, *indivname, *poplistname, *outputname, *outputvname;
  char *outliername, *snpeigname, *grmoutname, *deletesnpoutname, *xtxform, *weightname;
  char *phylipname, *fstdetailsname, *megaoutname;
  int nostatslim, fsthiprec, fstverbose, fstz, pubmean;
  double blgsize;
  int numeigs, numoutliter, numoutleigs, outliermode, altnormstyle, fancynorm;
  int minallelecnt, maxmissing, noxdata, numchrom, hiprec, regmode, hashcheck;
  int familynames, doproject, malexhet, grmbinary, xchrom, zchrom, lopos, hipos;
  int pordercheck, fastmode, fastdim, fastiter, bf16block, gpu, verbose;
  int usepops, ldregress, ldposlimit;
  double outlthresh, ldlimit, ldr2lo, ldr2hi;
  long seed;
} Settings;

static void defaults(Settings *s) {
  memset(s, 0, sizeof *s);
  s->numeigs = 10;            /* smartpca.c:164 */
  s->numoutliter = 5;         /* smartpca.c:211 */
  s->numoutleigs = 10;
  s->outlthresh = 6.0;        /* smartpca.c:212 */
  s->altnormstyle = 1;        /* smartpca.c:186 */
  s->fancynorm = 1;           /* include/globals.h:5 */
  s->minallelecnt = 1;        /* smartpca.c:187 */
  s->maxmissing = 9999999;    /* smartpca.c:188 */
  s->noxdata = 1;             /* smartpca.c:178 */
  s->numchrom = 22;           /* include/globals.h:4 */
  s->regmode = 1;             /* smartpca.c:208 */
  s->hashcheck = 1;
  s->familynames = 1;
  s->doproject = 1;
  s->xchrom = -1;
  s->zchrom = -1;
  s->lopos = -999999999;
  s->hipos = 999999999;
  s->pordercheck = 1;
  s->fastdim = -1;
  s->fastiter = -1;
  s->bf16block = 256;
  s->ldlimit = 9999.0;          /* smartpca.c:289-293 */
  s->ldr2lo = 0.01;
  s->ldr2hi = 0.95;
  s->ldposlimit = 1000 * 1000 * 1000;
  s->nostatslim = 10;           /* smartpca.c:183 */
  s->blgsize = 0.05;            /* smartpca.c:251, Morgans */
  s->pubmean = 1;               /* smartpca.c:275 */
}

static void read_settings(const Params *pp, Settings *s) {
  par_string(pp, "genotypename:", &s->genotypename);
  par_string(pp, "snpname:", &s->snpname);
  par_string(pp, "indivname:", &s->indivname);
  par_string(pp, "poplistname:", &s->poplistname);
  par_string(pp, "snpeigname:", &s->snpeigname);
  par_string(pp, "snpweightoutname:", &s->snpeigname);
  par_string(pp, "output:", &s->outputname);
  par_string(pp, "outputvecs:", &s->outputname);
  par_string(pp, "evecoutname:", &s->outputname);
  par_string(pp, "outputvals:", &s->outputvname);
  par_string(pp, "evaloutname:", &s->outputvname);
  par_string(pp, "outliername:", &s->outliername);
  par_string(pp, "outlieroutname:", &s->outliername);
  par_string(pp, "deletsnpoutname:", &s->deletesnpoutname);
  par_string(pp, "grmoutname:", &s->grmoutname);
  par_int(pp, "grmbinary:", &s->grmbinary);
  par_int(pp, "numeigs:", &s->numeigs);
  par_int(pp, "numoutevec:", &s->numeigs);
  par_int(pp, "fancynorm:", &s->fancynorm);
  par_int(pp, "usenorm:", &s->fancynorm);
  par_int(pp, "altnormstyle:", &s->altnormstyle);
  par_int(pp, "popgenmode:", &s->altnormstyle);
  par_int(pp, "hashcheck:", &s->hashcheck);
  par_int(pp, "noxdata:", &s->noxdata);
  int t = -1;
  par_int(pp, "xdata:", &t);
  if (t >= 0) s->noxdata = 1 - t;
  par_int(pp, "fastmode:", &s->fastmode);
  par_int(pp, "fastdim:", &s->fastdim);
  par_int(pp, "fastiter:", &s->fastiter);
  t = 0;
  par_int(pp, "seed:", &t);
  s->seed = t;
  par_int(pp, "regmode:", &s->regmode);
  par_int(pp, "lsqproject:", &s->regmode);
  par_int(pp, "doproject:", &s->doproject);
  par_int(pp, "minallelecnt:", &s->minallelecnt);
  par_int(pp, "chrom:", &s->xchrom);
  par_int(pp, "nochrom:", &s->zchrom);
  par_int(pp, "maxmissing:", &s->maxmissing);
  par_int(pp, "lopos:", &s->lopos);
  par_int(pp, "hipos:", &s->hipos);
  par_int(pp, "hiprecision:", &s->hiprec);
  par_int(pp, "hiprec:", &s->hiprec);
  par_int(pp, "malexhet:", &s->malexhet);
  par_int(pp, "nomalexhet:", &s->malexhet);
  par_int(pp, "familynames:", &s->familynames);
  par_int(pp, "numoutliter:", &s->numoutliter);
  par_int(pp, "numoutlieriter:", &s->numoutliter);
  par_int(pp, "numoutleigs", &s->numoutleigs); /* sic: no colon in the reference */
  par_int(pp, "numoutlierevec:", &s->numoutleigs);
  par_dbl(pp, "outlthresh:", &s->outlthresh);
  par_dbl(pp, "outliersigmathresh:", &s->outlthresh);
  par_int(pp, "outliermode:", &s->outliermode);
  par_int(pp, "numchrom:", &s->numchrom);
  par_int(pp, "pordercheck:", &s->pordercheck);
  par_string(pp, "weightname:", &s->weightname);
  par_int(pp, "usepopsformissing:", &s->usepops);
  par_int(pp, "ldregress:", &s->ldregress);
  par_int(pp, "nsnpldregress:", &s->ldregress);
  par_dbl(pp, "ldlimit:", &s->ldlimit);
  par_int(pp, "ldposlimit:", &s->ldposlimit);
  par_dbl(pp, "ldr2lo:", &s->ldr2lo);
  par_dbl(pp, "ldr2hi:", &s->ldr2hi);
  par_dbl(pp, "maxdistldregress:", &s->ldlimit);
  par_string(pp, "phylipname:", &s->phylipname);
  par_string(pp, "phylipoutname:", &s->phylipname);
  par_string(pp, "fstdetailsname:", &s->fstdetailsname);
  par_string(pp, "megaoutname:", &s->megaoutname);
  par_int(pp, "nostatslim:", &s->nostatslim);
  par_int(pp, "minleneig:", &s->nostatslim);
  par_int(pp, "fsthiprecision:", &s->fsthiprec);
  par_int(pp, "fstverbose:", &s->fstverbose);
  par_int(pp, "fstz:", &s->fstz);
  par_int(pp, "pubmean:", &s->pubmean);
  par_dbl(pp, "blgsize:", &s->blgsize);
  par_string(pp, "xtxform:", &s->xtxform);
  par_int(pp, "bf16block:", &s->bf16block);
  par_int(pp, "gpu:", &s->gpu);
}

/* Reference keys that change the SNP set, the samples, the normalisation or the
 * outputs, and that this build does not implement (readcommands, smartpca.c:2055-2233;
 * defaults smartpca.c:154-300).  A parfile that sets one of them to anything but the
 * reference default stops with an error instead of producing a different PCA. */
static void reject_unsupported(const Params *pp) {
  static const char *const any_value[] = {
      "badsnpname:", "xregionname:", "elllistname:", "topright:", "id2pops:",
      "snpoutfilename:", "snpoutname:", "indoutfilename:", "indivoutname:",
      "genooutfilename:", "genotypeoutname:", NULL};
  static const char *const nonzero[] = {
      "killr2:", "shrinkmode:", "newshrink:", "fastshrink:", "autoshrink:", "mpshrink:",
      "missingmode:", "easymode:", "fstonly:", "qtmode:", "rounakmode:", "inbreed:",
      "markerscore:", "pcorrmode:", "dotpopsmode:", "chisqmode:", NULL};
  for (int i = 0; any_value[i]; ++i)
    if (par_str(pp, any_value[i]))
      die("parameter %s is not supported by this build", any_value[i]);
  for (int i = 0; nonzero[i]; ++i) {
    int v = 0;
    par_int(pp, nonzero[i], &v);
    if (v != 0) die("parameter %s is not supported by this build", nonzero[i]);
  }
  int v = -1;
  par_int(pp, "popsizelimit:", &v);
  if (v > 0) die("parameter popsizelimit: is not supported by this build");
  double h = -1.0;
  par_dbl(pp, "hwfilter:", &h);
  if (h > 0.0) die("parameter hwfilter: is not supported by this build");
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void log_deleted(const Settings *s, const char *id, const char *why) {
  if (!s->deletesnpoutname) return;
  FILE *f = fopen(s->deletesnpoutname, "a");
  if (!f) return;
  fprintf(f, "%-40s  %-40s\n", id, why); /* logdeletedsnp, mcio.c:4893-4900 */
  fclose(f);
}

/* ridoutlier (smartsubs.c:18-93) */
static int ridoutlier(const double *evecs, int n, int neigs, double thresh, int mode, int *bad,
                      int *vecno, double *score) {
  if (mode > 1 || n < 3) return 0;
  int *flag = calloc(n, sizeof(int)), *vno = malloc(sizeof(int) * n);
  double *sc = calloc(n, sizeof(double)), *w = malloc(sizeof(double) * n),
         *w2 = malloc(sizeof(double) * n);
  for (int j = 0; j < n; ++j) vno[j] = -1;
  for (int e = 0; e < neigs; ++e) {
    memcpy(w, evecs + (int64_t)e * n, sizeof(double) * n);
    if (mode == 0) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += w[j];
      double mu = s / (double)n, s2 = 0.0;
      for (int j = 0; j < n; ++j) w[j] += -mu;
      for (int j = 0; j < n; ++j) s2 += w[j] * w[j];
      double r = 1.0 / sqrt(s2 / (double)n);
      for (int j = 0; j < n; ++j) {
        w[j] *= r;
        if (fabs(w[j]) > thresh) {
          flag[j] = 1;
          if (vno[j] < 0) { vno[j] = e; sc[j] = w[j]; }
        }
      }
    } else {
      for (int j = 0; j < n; ++j) {
        double yy = w[j], s = 0.0, s2 = 0.0;
        w[j] = 0.0;
        for (int i = 0; i < n; ++i) s += w[i];
        double mu = s / (double)(n - 1);
        for (int i = 0; i < n; ++i) w2[i] = w[i] + (-mu);
        w2[j] = 0.0;
        for (int i = 0; i < n; ++i) s2 += w2[i] * w2[i];
        double z = (yy - mu) / sqrt(s2 / (double)n);
        if (fabs(z) > thresh) {
          flag[j] = 1;
          if (vno[j] < 0) { vno[j] = e; sc[j] = z; }
        }
        w[j] = yy;
      }
    }
  }
  int nb = 0;
  for (int j = 0; j < n; ++j)
    if (flag[j]) { bad[nb] = j; vecno[nb] = vno[j]; score[nb] = sc[j]; ++nb; }
  free(flag); free(vno); free(sc); free(w); free(w2);
  return nb;
}

static int form_of(const char *x) {
  if (!x || !strcasecmp(x, "fp64")) return EIG_FORM_FP64;
  if (!strcasecmp(x, "bf16x4")) return EIG_FORM_BF16X4;
  if (!strcasecmp(x, "bf16x3")) return EIG_FORM_BF16X3;
  die("unknown xtxform: %s (fp64 | bf16x4 | bf16x3)", x);
  return -1;
}

int main(int argc, char **argv) {
  char *parname = NULL;
  int opt, verbose = 0;
  while ((opt = getopt(argc, argv, "p:vV")) != -1) {
    if (opt == 'p') parname = optarg;
    else if (opt == 'v') printf("version: eig-mi355x smartpca 8.0.0-compatible\n");
    else if (opt == 'V') verbose = 1;
    else die("bad params");
  }
  if (!parname) {
    fprintf(stderr, "no parameters\n");
    return 0;
  }
  double t_start = now_s();
  Params pp;
  par_read(parname, &pp);
  Settings S;
  defaults(&S);
  read_settings(&pp, &S);
  reject_unsupported(&pp);
  S.verbose = verbose;
  printf("parameter file: %s\n### THE INPUT PARAMETERS\n##PARAMETER NAME: VALUE\n", parname);
  for (int i = 0; i < pp.n; ++i) printf("%s %s\n", pp.key[i], pp.val[i]);
  if (!S.genotypename || !S.snpname || !S.indivname)
    die("genotypename, snpname and indivname are required");
  /* the loadings and the projection keep K x K normal equations in registers */
  if (S.numeigs > 16 && !S.fastmode)
    die("numoutevec: %d -- this build writes at most 16 output eigenvectors", S.numeigs);
  if (S.xchrom == S.numchrom + 1) S.noxdata = 0;
  if (S.ldlimit <= 0 || S.ldposlimit <= 0) S.ldregress = 0; /* smartpca.c:743-744 */
  /* usepopsformissing, ldregress and SNP weights take the non-lookup column path
     (smartpca.c:995-1015) */
  const int normal = S.usepops || S.ldregress > 0 || S.weightname != NULL;
  const int form = form_of(S.xtxform);
  if (S.deletesnpoutname) remove(S.deletesnpoutname);

  /* ---- inputs */
  Snp *snps;
  Ind *inds;
  int is_map = 0;
  int64_t M = read_snps(S.snpname, &snps, S.numchrom, &is_map);
  int32_t N = read_inds(S.indivname, &inds, S.familynames);
  int64_t rlen;
  uint8_t *rows = read_genotypes(S.genotypename, snps, M, inds, N, S.hashcheck,
                                 S.pordercheck && !is_map, &rlen);
  printf("numsnps input: %lld  numindivs: %d\n", (long long)M, N);
  /* male X heterozygotes read as missing (checkxval, mcio.c:1606-1618) */
  if (!S.malexhet)
    for (int64_t m = 0; m < M; ++m) {
      if (snps[m].chrom != S.numchrom + 1) continue;
      uint8_t *row = rows + m * rlen;
      for (int32_t k = 0; k < N; ++k)
        if (inds[k].sex == 'M' && geno_code(row, k) == 1) {
          int sh = (3 - (k & 3)) << 1;
          row[k >> 2] |= (uint8_t)(3 << sh);
        }
    }

  /* ---- axis populations (setstatus / seteglist, mcio.c:1297-1331, egsubs.c:86-109) */
  int neg = 0;
  char **eglist = malloc(sizeof(char *) * (N + 1));
  if (S.poplistname) {
    FILE *f = fopen(S.poplistname, "r");
    if (!f) die("can't open %s", S.poplistname);
    char line[1024];
    while (fgets(line, sizeof line, f)) {
      char *p = line;
      while (*p && isspace((unsigned char)*p)) ++p;
      if (!*p || *p == '#') continue;
      char *e = p;
      while (*e && !isspace((unsigned char)*e)) ++e;
      *e = 0;
      eglist[neg++] = strdup(p);
    }
    fclose(f);
    for (int32_t i = 0; i < N; ++i) {
      inds[i].aff = 0;
      if (inds[i].ignore) continue;
      for (int k = 0; k < neg; ++k)
        if (!strcmp(inds[i].pop, eglist[k])) inds[i].aff = 1;
    }
  } else {
    for (int32_t i = 0; i < N; ++i) {
      inds[i].aff = !inds[i].ignore;
      if (inds[i].ignore) continue;
      int k = 0;
      for (; k < neg; ++k)
        if (!strcmp(inds[i].pop, eglist[k])) break;
      if (k == neg) eglist[neg++] = inds[i].pop;
    }
  }

  /* ---- SNP filters (smartpca.c:801-852): the chromosome tests first, then the
     position tests, each logged on its own, also for SNPs already dropped */
  for (int64_t m = 0; m < M; ++m) {
    Snp *c = &snps[m];
    const char *why = NULL;
    if (S.noxdata && c->chrom == S.numchrom + 1) why = "chrom-X";
    else if (c->chrom == 0) why = "chrom-0";
    else if (c->chrom > S.numchrom + 1) why = "chrom-big";
    else if (c->chrom == S.zchrom) why = "chrom-deltet";
    if (why) {
      c->ignore = 1;
      log_deleted(&S, c->id, why);
    }
  }
  /* nodata: <= 1 valid genotype among non-ignored axis individuals */
  int64_t nodata = 0;
  for (int64_t m = 0; m < M; ++m) {
    Snp *c = &snps[m];
    const long pos = lrint(c->ppos); /* nnint, nicksrc/vsubs.c:849-854 */
    if (S.xchrom > 0 && c->chrom != S.xchrom) {
      c->ignore = 1;
      log_deleted(&S, c->id, "not-chrom");
    }
    if (S.xchrom > 0 && pos < S.lopos) {
      c->ignore = 1;
      log_deleted(&S, c->id, "lopos");
    }
    if (S.xchrom > 0 && pos > S.hipos) {
      c->ignore = 1;
      log_deleted(&S, c->id, "hipos");
    }
    if (snps[m].ignore) continue;
    const uint8_t *row = rows + m * rlen;
    int nv = 0;
    for (int32_t k = 0; k < N && nv <= 1; ++k)
      if (!inds[k].ignore && inds[k].aff && geno_code(row, k) != 3) ++nv;
    if (nv <= 1) {
      snps[m].ignore = 1;
      ++nodata;
      log_deleted(&S, snps[m].id, "nodata");
    }
  }
  printf("snps deleted (nodata): %lld\n", (long long)nodata);
  /* individuals with no valid genotype over the kept SNPs (smartpca.c:885-897) */
  {
    int64_t *vv = calloc(N, sizeof(int64_t));
    for (int64_t m = 0; m < M; ++m) {
      if (snps[m].ignore) continue;
      const uint8_t *row = rows + m * rlen;
      for (int32_t k = 0; k < N; ++k) vv[k] += geno_code(row, k) != 3;
    }
    for (int32_t i = 0; i < N; ++i) {
      if (vv[i] == 0) inds[i].ignore = 1;
      if (!S.doproject) {
        int k = 0;
        for (; k < neg; ++k)
          if (!strcmp(inds[i].pop, eglist[k])) break;
        if (k == neg) inds[i].ignore = 1;
      }
    }
    free(vv);
  }
  /* rmsnps: keep only the SNPs that survived */
  int64_t Mk = 0;
  for (int64_t m = 0; m < M; ++m)
    if (!snps[m].ignore) {
      if (Mk != m) {
        memmove(rows + Mk * rlen, rows + m * rlen, (size_t)rlen);
        snps[Mk] = snps[m];
      }
      ++Mk;
    }
  if (Mk == 0) die("no data");
  /* SNP weights (getweights, mcio.c:2298-2336): "ID weight" lines, default 1 */
  double *weight = NULL;
  if (S.weightname) {
    weight = malloc(sizeof(double) * (size_t)Mk);
    for (int64_t m = 0; m < Mk; ++m) weight[m] = 1.0;
    FILE *f = fopen(S.weightname, "r");
    if (!f) die("can't open %s", S.weightname);
    char line[4096], id[4096], wv[4096];
    while (fgets(line, sizeof line, f)) {
      int nf = sscanf(line, "%4095s %4095s", id, wv);
      if (nf < 1 || id[0] == '#' || !strcmp(id, "SNP_ID") || !strcmp(id, "Indiv_ID") ||
          !strcmp(id, "Chr"))
        continue;
      if (nf < 2) continue;
      for (int64_t m = 0; m < Mk; ++m)
        if (!strcmp(snps[m].id, id)) {
          weight[m] = atof(wv);
          break;
        }
    }
    fclose(f);
  }
  /* population index (in eglist) of every individual, -1 outside it */
  int32_t *popidx = malloc(sizeof(int32_t) * N);
  for (int32_t i = 0; i < N; ++i) {
    popidx[i] = -1;
    for (int k = 0; k < neg; ++k)
      if (!strcmp(inds[i].pop, eglist[k])) { popidx[i] = k; break; }
  }
  /* xtypes as lsqproj reads it: indexed by individual number but filled by row
     (smartpca.c:853, 1096-1100, 4664), -2 where no pass wrote it */
  int32_t *xt_quirk = malloc(sizeof(int32_t) * N);
  for (int32_t i = 0; i < N; ++i) xt_quirk[i] = -2;
  int32_t *xtypes = malloc(sizeof(int32_t) * N);
  int32_t *chromv = NULL;
  double *pposv = NULL, *gposv = NULL;
  if (S.ldregress > 0) {
    chromv = malloc(sizeof(int32_t) * (size_t)Mk);
    pposv = malloc(sizeof(double) * (size_t)Mk);
    gposv = malloc(sizeof(double) * (size_t)Mk);
    for (int64_t m = 0; m < Mk; ++m) {
      chromv[m] = snps[m].chrom;
      pposv[m] = snps[m].ppos;
      gposv[m] = snps[m].gpos;
    }
  }
  double ynumsnps = 0.0;

  /* ---- engine */
  eig_ctx *ctx = NULL;
  if (eig_open(S.gpu, &ctx) != EIG_OK) die("eig_open(%d) failed: no usable gfx950 device", S.gpu);
  CHK(eig_load_packed(ctx, rows, rlen, Mk, N));

  int32_t *xindex = malloc(sizeof(int32_t) * N);
  uint8_t *dead = calloc((size_t)Mk, 1), *live = malloc((size_t)Mk);
  int32_t *c0 = malloc(4 * Mk), *c1 = malloc(4 * Mk), *nmiss = malloc(4 * Mk),
          *status = malloc(4 * Mk);
  double *mean = malloc(8 * Mk), *fancy = malloc(8 * Mk);
  double *evals = NULL, *evecs = NULL;
  int32_t nrows = 0;
  int64_t ncols = 0, nused = 0;
  FILE *outl = stdout;
  if (S.outliername) {
    outl = fopen(S.outliername, "w");
    if (!outl) die("can't open %s", S.outliername);
  }
  /* ---- fastmode (smartpca.c:955-980): randomized PCA, .evec only, then stop */
  if (S.fastmode && S.numeigs > 0) {
    const int K = S.numeigs;
    const int L = S.fastdim < 0 ? 2 * K : S.fastdim;  /* smartpca.c:650-656 */
    const int I = S.fastiter < 0 ? K : S.fastiter;
    if (K >= L || I == 0) exit(1);                     /* kjg_fpca.c:26-29 */
    printf("seed: %ld\n", S.seed);
    nrows = 0;
    for (int32_t i = 0; i < N; ++i)
      if (!inds[i].ignore && inds[i].aff) xindex[nrows++] = i;
    for (int64_t m = 0; m < Mk; ++m) live[m] = 1;
    printf("number of samples used: %d number of snps used: %lld\n", nrows, (long long)Mk);
    double t0 = now_s();
    CHK(eig_select(ctx, xindex, nrows));
    CHK(eig_snpstats(ctx, live, S.altnormstyle, S.fancynorm, S.minallelecnt, S.maxmissing, NULL,
                     NULL, NULL, NULL, NULL, NULL));
    double *fev = malloc(sizeof(double) * K), *fvec = malloc(sizeof(double) * K * (size_t)nrows);
    /* xtxform: bf16x4 selects the split-BF16 panel products (fp64 otherwise) */
    CHK(eig_fpca_form(ctx, K, L, I, (uint64_t)S.seed, NULL,
                      form == EIG_FORM_BF16X4 ? EIG_FORM_BF16X4 : EIG_FORM_FP64, fev, fvec));
    if (S.verbose) printf("fastmode: %.3f s on the device\n", now_s() - t0);
    /* printevecs (smartpca.c:3811-3875): header, then rows only when every
       individual is an axis sample */
    FILE *of = stdout;
    if (S.outputname) {
      of = fopen(S.outputname, "w");
      if (!of) die("can't open %s", S.outputname);
    }
    fprintf(of, "%20s ", "#eigvals:");
    for (int j = 0; j < K; ++j) fprintf(of, "%9.3f ", fev[j]);
    fprintf(of, "\n");
    if (nrows != N) {
      if (of != stdout) fclose(of);
      die("... not yet implemented!");
    }
    printf("writing eigenvecotrs for  %d samples\n", nrows);
    for (int j = 0; j < K; ++j) { /* setfvecs (x10) then unit norm */
      double *x = fvec + (size_t)j * nrows, y = 0.0;
      for (int32_t i = 0; i < nrows; ++i) {
        x[i] *= 10.0;
        y += x[i] * x[i];
      }
      const double f = 1.0 / sqrt(y);
      for (int32_t i = 0; i < nrows; ++i) x[i] *= f;
    }
    for (int32_t r = 0; r < nrows; ++r) {
      const Ind *ip = &inds[xindex[r]];
      fprintf(of, "%20s ", ip->id);
      for (int j = 0; j < K; ++j)
        fprintf(of, S.hiprec ? "%12.6f  " : "%10.4f  ", fvec[(size_t)j * nrows + r]);
      fprintf(of, "%15s\n", ip->pop);
    }
    if (of != stdout) fclose(of);
    printf("end of smartpca(fastmode)\n");
    eig_close(ctx);
    return 0;
  }
  int numoutiter = S.numoutliter >= 1 ? S.numoutliter + 1 : 1;
  int omode = S.numoutliter >= 1 ? S.outliermode : 2;
  /* the eigenvectors a pass needs: the output axes and, when outliers are removed,
     the outlier axes; the device computes all of them, so none is ever dropped */
  int nvec_pass = S.numeigs;
  if (S.numoutliter >= 1 && S.numoutleigs > nvec_pass) nvec_pass = S.numoutleigs;
  double t_xtx = 0, t_eig = 0;
  for (int outliter = 1; outliter <= numoutiter; ++outliter) {
    nrows = 0;
    for (int32_t i = 0; i < N; ++i)
      if (!inds[i].ignore && inds[i].aff) xindex[nrows++] = i;
    if (nrows < 2) die("too few samples: %d", nrows);
    if (nrows >= 46341) die("size of GRM matrix too large: nrows: %d\n Run with fastmode: YES perhaps", nrows);
    ncols = 0;
    for (int64_t m = 0; m < Mk; ++m) {
      live[m] = !dead[m];
      ncols += live[m];
    }
    double t0 = now_s();
    for (int32_t r = 0; r < nrows; ++r) xt_quirk[r] = xtypes[r] = popidx[xindex[r]];
    CHK(eig_select(ctx, xindex, nrows));
    CHK(eig_snpstats(ctx, live, S.altnormstyle, S.fancynorm, S.minallelecnt, S.maxmissing, c0,
                     c1, nmiss, mean, fancy, status));
    double yscale = 0.0;
    int rc;
    if (normal)
      rc = eig_xtx_normal(ctx, S.usepops ? xtypes : NULL, neg, weight, S.ldregress, chromv, pposv,
                          gposv, S.ldposlimit, S.ldlimit, S.ldr2lo, S.ldr2hi, S.minallelecnt,
                          S.maxmissing, c0, c1, nmiss, status, &yscale, &ynumsnps, NULL);
    else
      rc = eig_xtx(ctx, form, S.bf16block, &yscale, NULL);
    int nkill = 0;
    nused = 0;
    for (int64_t m = 0; m < Mk; ++m) {
      if (status[m] == EIG_SNP_DEAD) continue;
      if (status[m] != EIG_SNP_KEPT) {
        dead[m] = 1;
        log_deleted(&S, snps[m].id, "minallelecnt");
        if (nkill < 10) {
          int tt = status[m] == EIG_SNP_ALLMISS ? 0 : (c0[m] < c1[m] ? c0[m] : c1[m]);
          printf(" snp %20s ignored . allelecnt: %5d  missing: %5d\n", snps[m].id, tt,
                 status[m] == EIG_SNP_ALLMISS ? 0 : nmiss[m]);
        }
        ++nkill;
      } else {
        ++nused;
      }
    }
    printf("number of samples used: %d number of snps used: %lld\n", nrows, (long long)nused);
    if (!normal) ynumsnps = (double)nused;
    if (rc == EIG_E_NUMERIC) die("%s", eig_last_error(ctx));
    CHK(rc);
    double t1 = now_s();
    free(evals);
    free(evecs);
    evals = malloc(sizeof(double) * nrows);
    int nv = nvec_pass < nrows ? nvec_pass : nrows;
    evecs = malloc(sizeof(double) * (size_t)nrows * (nv > 0 ? nv : 1));
    CHK(eig_eigh_xtx(ctx, nv, evals, evecs));
    double t2 = now_s();
    t_xtx += t1 - t0;
    t_eig += t2 - t1;
    if (outliter > S.numoutliter) break;
    int nle = S.numoutleigs < nrows - 1 ? S.numoutleigs : nrows - 1;
    if (nle > nv) nle = nv;
    int *bad = malloc(sizeof(int) * nrows), *vno = malloc(sizeof(int) * nrows);
    double *sc = malloc(sizeof(double) * nrows);
    int nbad = ridoutlier(evecs, nrows, nle, S.outlthresh, omode, bad, vno, sc);
    for (int b = 0; b < nbad; ++b) {
      Ind *x = &inds[xindex[bad[b]]];
      fprintf(outl, "REMOVED outlier %s iter %d evec %d sigmage %.3f pop: %s\n", x->id, outliter,
              vno[b], sc[b], x->pop);
      x->ignore = 1;
    }
    free(bad);
    free(vno);
    free(sc);
    if (nbad == 0) break;
    int nr = 0;
    for (int32_t i = 0; i < N; ++i) nr += !inds[i].ignore && inds[i].aff;
    printf("number of samples after outlier removal: %d\n", nr);
  }
  if (outl != stdout) fclose(outl);

  /* ---- GRM (dumpgrm, smartpca.c:3687-3808) */
  if (S.grmoutname) {
    double *xtx = malloc(sizeof(double) * (size_t)nrows * nrows);
    double ys, yn;
    if (normal)
      CHK(eig_xtx_normal(ctx, S.usepops ? xtypes : NULL, neg, weight, S.ldregress, chromv, pposv,
                         gposv, S.ldposlimit, S.ldlimit, S.ldr2lo, S.ldr2hi, S.minallelecnt,
                         S.maxmissing, NULL, NULL, NULL, NULL, &ys, &yn, xtx));
    else
      CHK(eig_xtx(ctx, form, S.bf16block, &ys, xtx));
    /* dumpgrm takes ynumsnps as an int (smartpca.c:1278, 3770) */
    const int grm_nsnp = (int)ynumsnps;
    char nm[4096];
    snprintf(nm, sizeof nm, "%s.id", S.grmoutname);
    FILE *f = fopen(nm, "w");
    if (!f) die("can't open %s", nm);
    for (int32_t a = 0; a < nrows; ++a) fprintf(f, "%s\t%s\n", "NA", inds[xindex[a]].id);
    fclose(f);
    double tr = 0.0;
    for (int32_t a = 0; a < nrows; ++a) tr += xtx[(int64_t)a * nrows + a];
    if (S.grmbinary) {
      snprintf(nm, sizeof nm, "%s.N.bin", S.grmoutname);
      f = fopen(nm, "wb");
      if (!f) die("can't open %s", nm);
      int32_t w = (int32_t)grm_nsnp;
      for (int64_t a = 0; a < (int64_t)nrows * (nrows + 1) / 2; ++a) fwrite(&w, 4, 1, f);
      fclose(f);
      snprintf(nm, sizeof nm, "%s.bin", S.grmoutname);
      f = fopen(nm, "wb");
      if (!f) die("can't open %s", nm);
      double yn = tr / (double)nrows;
      for (int32_t a = 0; a < nrows; ++a)
        for (int32_t b = 0; b <= a; ++b) {
          float v = (float)(xtx[(int64_t)a * nrows + b] / yn);
          fwrite(&v, 4, 1, f);
        }
      fclose(f);
    } else {
      f = fopen(S.grmoutname, "w");
      if (!f) die("can't open %s", S.grmoutname);
      double rec = (double)nrows / tr;
      for (int32_t a = 0; a < nrows; ++a)
        for (int32_t b = 0; b <= a; ++b)
          fprintf(f, "%d %d %d %0.6f\n", a + 1, b + 1, grm_nsnp,
                  xtx[(int64_t)a * nrows + b] * rec);
      fclose(f);
    }
    free(xtx);
    printf("grm dumped\n");
  }

  /* ---- outputs (smartpca.c:1300-1591) */
  int K = S.numeigs;
  if (K > nrows) K = nrows;
  if (K > ncols) K = (int)ncols;
  int nvec_got = nvec_pass < nrows ? nvec_pass : nrows;
  if (K > nvec_got) K = nvec_got;
  const int mpos = st_numgtz(evals, nrows);
  if (mpos == 0) die("no data");
  /* Tracy-Widom statistics of the positive eigenvalues (smartpca.c:1312-1368) */
  printf("\n## Tracy-Widom statistics: rows: %d  cols: %lld\n", nrows, (long long)ncols);
  st_print_tw(stdout, evals, mpos, -1.0, (double)nrows, S.nostatslim < 3 ? 3 : S.nostatslim);
  if (S.outputvname) {
    FILE *f = fopen(S.outputvname, "w");
    if (!f) die("can't open %s", S.outputvname);
    for (int32_t i = 0; i < nrows; ++i) fprintf(f, "%12.6f\n", evals[i]);
    fclose(f);
  }
  FILE *ofile = stdout;
  if (S.outputname) {
    ofile = fopen(S.outputname, "w");
    if (!ofile) die("can't open %s", S.outputname);
  }
  uint8_t *ign = malloc((size_t)Mk);
  for (int64_t m = 0; m < Mk; ++m) ign[m] = dead[m] || status[m] != EIG_SNP_KEPT;
  double *ffvecs = NULL;
  if (K > 0) {
    fprintf(ofile, "%20s ", "#eigvals:");
    for (int j = 0; j < K; ++j) fprintf(ofile, "%9.3f ", evals[j]);
    fprintf(ofile, "\n");
    /* loadings + projection scales over the SNPs that survived the final pass */
    ffvecs = malloc(sizeof(double) * (size_t)K * Mk);
    double fxscal[64];
    CHK(eig_loadings(ctx, evecs, K, ign, ffvecs, fxscal));
    /* kurtosis of each axis's SNP loadings (over the pass's SNPs) and sample
       coordinates (smartpca.c:1497-1501, 1531-1538) */
    {
      double *w = malloc(sizeof(double) * (size_t)(ncols > nrows ? ncols : nrows));
      printf("%12s %4s %9s %9s\n", "kurtosis", "", "snps", "indivs");
      for (int j = 0; j < K; ++j) {
        int64_t c = 0;
        for (int64_t m = 0; m < Mk; ++m)
          if (live[m]) w[c++] = ffvecs[(int64_t)j * Mk + m];
        const double ks = st_kurtosis(w, (int)c);
        const double ki = st_kurtosis(evecs + (int64_t)j * nrows, nrows);
        printf("%12s %4d %9.3f %9.3f\n", "eigenvector", j + 1, ks, ki);
      }
      free(w);
    }
    double *acoeffs = calloc((size_t)K * N, sizeof(double));
    double *bcoeffs = calloc((size_t)K * N, sizeof(double));
    if (S.regmode) {
      int32_t *use = malloc(4 * (size_t)N), *lst = malloc(4 * (size_t)N);
      for (int32_t i = 0; i < N; ++i) use[i] = !inds[i].ignore;
      if (S.usepops)
        CHK(eig_lsqproject_fill(ctx, use, xt_quirk, neg, acoeffs, bcoeffs, lst));
      else
        CHK(eig_lsqproject(ctx, use, acoeffs, bcoeffs, lst));
      for (int32_t i = 0; i < N; ++i)
        if (lst[i] == 1) {
          inds[i].ignore = 1;
          printf("%s ignored (insufficient data\n", inds[i].id);
        }
      /* seteigscale (smartpca.c:4758-4782) over the axis samples */
      for (int j = 0; j < K; ++j) {
        double ab = 0.0, aa = 0.0;
        for (int32_t r = 0; r < nrows; ++r) {
          double a = acoeffs[(int64_t)j * N + xindex[r]], b = bcoeffs[(int64_t)j * N + xindex[r]];
          ab += a * b;
          aa += a * a;
        }
        double es = ab / aa;
        for (int32_t i = 0; i < N; ++i) acoeffs[(int64_t)j * N + i] *= es;
      }
      free(use);
      free(lst);
    } else {
      printf("lsqproject off:  deprecated\n");
    }
    for (int32_t i = 0; i < N; ++i) {
      if (inds[i].ignore) continue;
      fprintf(ofile, "%20s ", inds[i].id);
      for (int j = 0; j < K; ++j)
        fprintf(ofile, S.hiprec ? "%12.6f  " : "%10.4f  ", acoeffs[(int64_t)j * N + i]);
      fprintf(ofile, "%15s\n", inds[i].pop);
    }
    free(acoeffs);
    free(bcoeffs);
  }
  if (ofile != stdout) fclose(ofile);

  /* ---- population statistics (smartpca.c:1779-1994) */
  int *xpopsize = calloc(neg > 0 ? neg : 1, sizeof(int));
  for (int32_t r = 0; r < nrows; ++r) ++xpopsize[xtypes[r]];
  for (int k = 0; k < neg; ++k) {
    printf("population: %3d %20s %4d", k, eglist[k], xpopsize[k]);
    if (xpopsize[k] == 0) printf(" ***");
    printf("\n");
  }
  {
    /* Fst between the axis populations with a weighted block jackknife over blocks of
       blgsize Morgans (dofstxx, smartpca.c:2830-2861; dofstnumx, qpsubs.c:2488-2740) */
    const int G2 = neg * neg;
    int32_t *cnt = malloc(sizeof(int32_t) * (size_t)Mk * neg * 2);
    uint8_t *hap = malloc((size_t)nrows), *xs = malloc((size_t)Mk);
    for (int32_t r = 0; r < nrows; ++r) hap[r] = inds[xindex[r]].sex == 'M';
    for (int64_t m = 0; m < Mk; ++m) xs[m] = snps[m].chrom == 23;
    CHK(eig_popcounts(ctx, xtypes, neg, hap, xs, cnt));
    int *tag = malloc(sizeof(int) * (size_t)Mk), *cm = malloc(sizeof(int) * (size_t)Mk);
    double *gp = malloc(sizeof(double) * (size_t)Mk);
    for (int64_t m = 0; m < Mk; ++m) { cm[m] = snps[m].chrom; gp[m] = snps[m].gpos; }
    const int nblocks = st_setblocks(cm, gp, ign, Mk, S.blgsize, tag);
    printf("number of blocks for block jackknife: %d\n", nblocks);
    double *fmean = malloc(sizeof(double) * G2), *fans = malloc(sizeof(double) * G2),
           *fsd = malloc(sizeof(double) * G2), *fzz = malloc(sizeof(double) * G2);
    int *fnum = malloc(sizeof(int) * G2);
    st_fst(cnt, ign, tag, Mk, neg, nblocks, fmean, fans, fsd, fnum);
    for (int k = 0; k < G2; ++k) fzz[k] = fans[k];
    for (int k1 = 0; k1 < neg; ++k1)
      for (int k2 = k1 + 1; k2 < neg; ++k2) {
        double z = fans[k1 * neg + k2] / (fsd[k1 * neg + k2] + 1.0e-10);
        z = z < -9.999999 ? -9.999999 : (z > 9.999999 ? 9.999999 : z);
        fzz[k2 * neg + k1] = z;
      }
    if (S.megaoutname) { /* printmega (smartpca.c:4785-4812) */
      FILE *f = fopen(S.megaoutname, "w");
      if (!f) die("can't open %s", S.megaoutname);
      fprintf(f, "#mega\n!TITLE distance for %d samples ;\n\n", neg);
      for (int k = 0; k < neg; ++k) fprintf(f, "#%s\n", eglist[k]);
      fprintf(f, "\n\n");
      for (int a = 0; a < neg; ++a) {
        for (int b = 0; b < a; ++b) fprintf(f, "  %9.6f", fans[a * neg + b] > 0 ? fans[a * neg + b] : 0.0);
        fprintf(f, "\n");
      }
      fclose(f);
    }
    if ((!S.phylipname && neg > 10) || S.fstverbose) {
      printf("## Fst statistics between populations:         fst       std error      Z\n");
      for (int k1 = 0; k1 < neg; ++k1)
        for (int k2 = k1 + 1; k2 < neg; ++k2) {
          const double z = fans[k1 * neg + k2] / (fsd[k1 * neg + k2] + 1.0e-10);
          if (!S.fsthiprec)
            printf(" %20s %20s %9.3f %10.4f", eglist[k1], eglist[k2], fans[k1 * neg + k2],
                   fsd[k1 * neg + k2]);
          else
            printf(" %20s %20s %12.6f %12.6f", eglist[k1], eglist[k2], fans[k1 * neg + k2],
                   fsd[k1 * neg + k2]);
          printf("  %9.3f ", z);
          printf(" %12d\n", fnum[k1 * neg + k2]);
        }
      printf("\n");
    }
    printf("## end of Fst statistics between populations\n");
    if (S.fstdetailsname) {
      FILE *f = fopen(S.fstdetailsname, "w");
      if (!f) die("can't open %s", S.fstdetailsname);
      fprintf(f, "## Fst statistics between populations:     fst     fstjack       std error      Z    snpnumber\n");
      for (int k1 = 0; k1 < neg; ++k1)
        for (int k2 = k1 + 1; k2 < neg; ++k2) {
          const int k = k1 * neg + k2;
          fprintf(f, S.fsthiprec ? "F_st %20s %20s %15.9f %15.9f %15.9f"
                                 : "F_st %20s %20s %12.6f %12.6f %12.6f",
                  eglist[k1], eglist[k2], fmean[k], fans[k], fsd[k]);
          fprintf(f, " %9.3f", fans[k] / (fsd[k] + 1.0e-10));
          fprintf(f, " %6d\n", fnum[k]);
        }
      fclose(f);
    }
    if (S.fstz) {
      printf("\n");
      printf(S.fsthiprec ? "Fst: Z/F *1000000 [clipped to -10,10]\n"
                         : "Fst: Z/F *1000 [clipped to -10,10]\n");
      st_printmatz(fzz, eglist, neg, S.fsthiprec ? 1e6 : 1e3, S.fsthiprec ? 10 : 5);
    }
    printf("\n");
    if (S.phylipname) {
      FILE *f = fopen(S.phylipname, "w");
      if (!f) die("can't open %s", S.phylipname);
      fprintf(f, "%6d\n", neg);
      for (int k1 = 0; k1 < neg; ++k1) {
        char nm[11];
        snprintf(nm, sizeof nm, "%s", eglist[k1]);
        fprintf(f, "%10s", nm);
        for (int k2 = 0; k2 < neg; ++k2)
          fprintf(f, "%6.3f", 0.5 * (fans[k1 * neg + k2] + fans[k2 * neg + k1]));
        fprintf(f, "\n");
      }
      fclose(f);
    }
    if (neg <= 10) {
      printf(S.fsthiprec ? "fst *1000000:\n" : "fst *1000:\n");
      st_printmatz(fans, eglist, neg, S.fsthiprec ? 1e6 : 1e3, S.fsthiprec ? 10 : 5);
      printf("\n");
    }
    printf("std. err.  * 1000000:\n");
    st_printmatz(fsd, eglist, neg, 1e6, 5);
    printf("\n");
    free(cnt); free(hap); free(xs); free(tag); free(cm); free(gp);
    free(fmean); free(fans); free(fsd); free(fzz); free(fnum);
  }
  /* anova of each eigenvector by population, and the population differences over the
     K axes (dottest, smartpca.c:1957-1994, 2315-2413) */
  double *chitot = calloc((size_t)(neg > 0 ? neg * neg : 1), sizeof(double));
  for (int j = 0; j < K; ++j) {
    char sss[64];
    snprintf(sss, sizeof sss, "eigenvector %d", j + 1);
    st_dottest(sss, evecs + (int64_t)j * nrows, eglist, neg, xtypes, nrows, chitot, S.pubmean);
  }
  printf("\n## Statistical significance of differences beween populations:\n");
  printf("                                pop1                  pop2      chisq          p-value   |pop1|   |pop2|\n");
  for (int k1 = 0; k1 < neg; ++k1)
    for (int k2 = k1 + 1; k2 < neg; ++k2) {
      const double ychi = chitot[k1 * neg + k2];
      printf("popdifference:  %20s  %20s  %12.3f  %12.6g", eglist[k1], eglist[k2], ychi,
             st_chitail(K, ychi));
      printf("   %5d", xpopsize[k1]);
      printf("   %5d\n", xpopsize[k2]);
    }
  printf("\n");
  free(chitot);
  free(xpopsize);

  if (K > 0) {
    /* snp weights (writesnpeigs, smartpca.c:3038-3104): loadings scaled to mean square
       1 over the pass's SNPs, the 10 largest |weights| of each axis on stdout, then the
       weights file */
    for (int j = 0; j < K; ++j) {
      double y = 0.0;
      for (int64_t m = 0; m < Mk; ++m) y += ffvecs[(int64_t)j * Mk + m] * ffvecs[(int64_t)j * Mk + m];
      double sc = sqrt((double)ncols / y);
      for (int64_t m = 0; m < Mk; ++m) ffvecs[(int64_t)j * Mk + m] *= sc;
    }
    {
      double *snpsc = malloc(sizeof(double) * (size_t)Mk);
      for (int64_t m = 0; m < Mk; ++m) snpsc[m] = -99999.0;
      for (int j = 0; j < K; ++j) {
        for (int64_t m = 0; m < Mk; ++m)
          if (!ign[m]) snpsc[m] = fabs(ffvecs[(int64_t)j * Mk + m]);
        for (int k = 0; k < 10 && ncols > 10; ++k) {
          int64_t kmax = 0; /* first maximum (vlmaxmin, nicksrc/vsubs.c:304-331) */
          for (int64_t m = 1; m < Mk; ++m)
            if (snpsc[m] > snpsc[kmax]) kmax = m;
          if (snpsc[kmax] < 0) break;
          printf("eigbestsnp %4d %20s %2d %12d %9.3f\n", j + 1, snps[kmax].id, snps[kmax].chrom,
                 (int)lrint(snps[kmax].ppos), snpsc[kmax]);
          snpsc[kmax] = -1.0;
        }
      }
      free(snpsc);
    }
    if (S.snpeigname) {
      FILE *f = fopen(S.snpeigname, "w");
      if (!f) die("can't open %s", S.snpeigname);
      for (int64_t m = 0; m < Mk; ++m) {
        if (ign[m]) continue;
        fprintf(f, "%20s %2d %12d", snps[m].id, snps[m].chrom, (int)lrint(snps[m].ppos));
        for (int j = 0; j < K; ++j)
          fprintf(f, S.hiprec ? " %12.6f" : " %9.3f", ffvecs[(int64_t)j * Mk + m]);
        fprintf(f, "\n");
      }
      fclose(f);
    }
  }
  free(ign);
  free(ffvecs);
  printf("##end of smartpca: %12.3f seconds wall (xtx %.3f, eigensolve %.3f)\n",
         now_s() - t_start, t_xtx, t_eig);
  eig_close(ctx);
  return 0;
}
