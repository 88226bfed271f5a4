/*
 * stats.c -- the reference smartpca's statistics after the eigendecomposition, on the
 * host: Tracy-Widom tests of the eigenvalues, Fst between populations with a block
 * jackknife, the anova of each eigenvector by population, the pairwise population
 * differences, kurtosis.  They print to stdout (and the Fst side files) in the
 * reference's formats; the per-SNP population allele counts that Fst needs come from
 * the device (eig_popcounts).
 *
 * Reference: smartpca.c:1312-1368 (Tracy-Widom), 1497-1501 + 1531-1538 (kurtosis),
 * 1779-1994 (population sizes, Fst, anova, popdifference), 2315-2524 (dottest,
 * anova, anova1, publishit); nicksrc/statsubs.c:1590-1726 (twtail, twnorm, dotwcalc),
 * 1793-1900 (gettw and its cubic interpolation); qpsubs.c:430-455 (calcmean),
 * 1205-1350 (fstcolyy), 1578-1637 (setblocks), 2488-2740 (dofstnumx), 2793-2860
 * (wjackest); admutils.c:995-1017 (kurtosis), 1185-1238 (printmatz5, get3).
 *
 * The Tracy-Widom (beta = 1) distribution is computed here from its definition: the
 * Hastings-McLeod solution q of Painleve II, q'' = s q + 2 q^3 with q ~ Ai(s) as
 * s -> +inf, integrated downward from s = 16 by classical 4th-order Runge-Kutta,
 * F1(s) = exp(-1/2 int_s^inf q - 1/2 int_s^inf (x-s) q^2) (Tracy & Widom 1996).
 * Tail and density are tabulated on the reference's grid (-8 .. 8 by 0.1) and
 * rounded to 9 decimals as its published table is; between grid points they are
 * interpolated by the same cubic Hermite scheme, beyond the grid the
 * Margetis-Edelman asymptotics take over exactly as in the reference (whose code
 * returns the asymptotic density in place of the tail there).
 *
 * The F and chi-square tails are the regularized incomplete beta function (by the
 * Lentz continued fraction, DLMF 8.17.22) and the regularized upper incomplete gamma
 * function (series DLMF 8.7.1 below a+1, Legendre continued fraction DLMF 8.9.2
 * above).
 */
#include "stats.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ special functions */
static double lbeta(double a, double b) { return lgamma(a) + lgamma(b) - lgamma(a + b); }

/* continued fraction of I_x(a, b), modified Lentz (DLMF 8.17.22) */
static double betacf(double a, double b, double x) {
  const double tiny = 1e-300, eps = 1e-16;
  double c = 1.0, d = 1.0 - (a + b) * x / (a + 1.0);
  if (fabs(d) < tiny) d = tiny;
  d = 1.0 / d;
  double f = d;
  for (int m = 1; m <= 10000; ++m) {
    const double m2 = 2.0 * m;
    double num = m * (b - m) * x / ((a + m2 - 1.0) * (a + m2));
    d = 1.0 + num * d;
    if (fabs(d) < tiny) d = tiny;
    c = 1.0 + num / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    f *= d * c;
    num = -(a + m) * (a + b + m) * x / ((a + m2) * (a + m2 + 1.0));
    d = 1.0 + num * d;
    if (fabs(d) < tiny) d = tiny;
    c = 1.0 + num / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    f *= del;
    if (fabs(del - 1.0) < eps) break;
  }
  return f;
}

/* regularized incomplete beta I_x(a, b) */
double st_ibeta(double a, double b, double x) {
  if (x <= 0.0) return 0.0;
  if (x >= 1.0) return 1.0;
  const double front = exp(a * log(x) + b * log1p(-x) - lbeta(a, b));
  if (x < (a + 1.0) / (a + b + 2.0)) return front * betacf(a, b, x) / a;
  return 1.0 - front * betacf(b, a, 1.0 - x) / b;
}

/* regularized upper incomplete gamma Q(a, x) */
double st_gammaq(double a, double x) {
  if (x <= 0.0) return 1.0;
  const double lpre = a * log(x) - x - lgamma(a);
  if (x < a + 1.0) { /* series for P */
    double sum = 1.0 / a, term = sum;
    for (int n = 1; n < 100000; ++n) {
      term *= x / (a + n);
      sum += term;
      if (fabs(term) < fabs(sum) * 1e-17) break;
    }
    return 1.0 - sum * exp(lpre);
  }
  /* continued fraction for Q, modified Lentz */
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; ++i) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < 1e-17) break;
  }
  return exp(lpre) * h;
}

/* upper tail of F(df1, df2) at F (rtlf) */
double st_ftail(int df1, int df2, double F) {
  if (F < 0.000001 || df1 < 1 || df2 < 1) return 1.0;
  return st_ibeta(0.5 * df2, 0.5 * df1, df2 / (df2 + df1 * F));
}

/* upper tail of chi-square(df) at z (rtlchsq) */
double st_chitail(int df, double z) {
  if (df <= 0) return 1.0;
  if (df == 1) return erfc(sqrt(0.5 * z));
  if (df == 2) return exp(-0.5 * z);
  return st_gammaq(0.5 * df, 0.5 * z);
}

/* ------------------------------------------------------------------ Tracy-Widom */
#define TW_N 161
static double tw_x[TW_N], tw_tail_t[TW_N], tw_pdf_t[TW_N];
static int tw_ready = 0;

/* Ai and Ai' for large x by their asymptotic series (DLMF 9.7.5, 9.7.6) */
static void airy_large(double x, double *ai, double *aip) {
  const double z = 2.0 / 3.0 * x * sqrt(x);
  double u = 1.0, su = 1.0, sv = 1.0, zk = 1.0;
  for (int k = 1; k < 10; ++k) {
    /* u_k = u_{k-1} (6k-5)(6k-3)(6k-1) / ((2k-1) 216 k) */
    u *= (6.0 * k - 5.0) * (6.0 * k - 3.0) * (6.0 * k - 1.0) / ((2.0 * k - 1.0) * 216.0 * k);
    const double v = -(6.0 * k + 1.0) / (6.0 * k - 1.0) * u;
    zk *= -z;
    su += u / zk;
    sv += v / zk;
  }
  const double e = exp(-z) / (2.0 * sqrt(M_PI));
  *ai = e / pow(x, 0.25) * su;
  *aip = -e * pow(x, 0.25) * sv;
}

static void tw_rhs(double s, const double *y, double *f) {
  f[0] = y[1];
  f[1] = s * y[0] + 2.0 * y[0] * y[0] * y[0];
  f[2] = -y[0];        /* I = int_s^inf q          */
  f[3] = -y[0] * y[0]; /* J = int_s^inf q^2        */
  f[4] = -y[3];        /* K = int_s^inf (x-s) q^2  */
}

static void tw_build(void) {
  const double s0 = 16.0, h = -1.0e-4;
  double y[5] = {0, 0, 0, 0, 0};
  airy_large(s0, &y[0], &y[1]);
  const long nstep = 240000; /* s0 -> -8 */
  int next = TW_N - 1;       /* grid point 8.0 first */
  for (long i = 0; i <= nstep && next >= 0; ++i) {
    const double s = s0 + i * h;
    const long gi = lround((s + 8.0) * 10.0);
    if (gi == next && fabs(s - (-8.0 + 0.1 * gi)) < 1e-7) {
      const double F1 = exp(-0.5 * y[2] - 0.5 * y[4]);
      tw_x[next] = (double)(next - 80) / 10.0; /* the grid value, correctly rounded */
      tw_tail_t[next] = round((1.0 - F1) * 1e9) / 1e9;
      tw_pdf_t[next] = round(F1 * 0.5 * (y[0] + y[3]) * 1e9) / 1e9;
      --next;
    }
    double k1[5], k2[5], k3[5], k4[5], t[5];
    tw_rhs(s, y, k1);
    for (int j = 0; j < 5; ++j) t[j] = y[j] + 0.5 * h * k1[j];
    tw_rhs(s + 0.5 * h, t, k2);
    for (int j = 0; j < 5; ++j) t[j] = y[j] + 0.5 * h * k2[j];
    tw_rhs(s + 0.5 * h, t, k3);
    for (int j = 0; j < 5; ++j) t[j] = y[j] + h * k3[j];
    tw_rhs(s + h, t, k4);
    for (int j = 0; j < 5; ++j) y[j] += h / 6.0 * (k1[j] + 2.0 * k2[j] + 2.0 * k3[j] + k4[j]);
  }
  tw_ready = 1;
}

int st_twtable(double *x, double *tail, double *pdf) {
  if (!tw_ready) tw_build();
  memcpy(x, tw_x, sizeof tw_x);
  memcpy(tail, tw_tail_t, sizeof tw_tail_t);
  memcpy(pdf, tw_pdf_t, sizeof tw_pdf_t);
  return TW_N;
}

/* first grid index by the reference's recursive search (statsubs.c:1757-1780) */
static int fgtx(const double *tab, int lo, int hi, double val) {
  if (val >= tab[hi]) return hi + 1;
  if (val < tab[lo]) return lo;
  const int k = (lo + hi) / 2;
  if (val <= tab[k]) return fgtx(tab, lo + 1, k, val);
  return fgtx(tab, k, hi - 1, val);
}

static double tw_asym(double tw, double p) { /* Margetis-Edelman, exponent -p on tw */
  return exp(-p * log(tw) - 2.0 * pow(tw, 1.5) / 3.0 - log(4.0 * sqrt(M_PI)));
}

double st_twtail(double x) {
  if (!tw_ready) tw_build();
  const int k = fgtx(tw_x, 0, TW_N - 1, x);
  if (k <= 0) return 1.0;
  if (k >= TW_N) return x <= 0.0 ? 0.0 : tw_asym(x, 0.25); /* the reference's tail here */
  const double x0 = tw_x[k - 1], x1 = tw_x[k];
  const double f0 = tw_tail_t[k - 1], f1 = tw_tail_t[k];
  const double f0p = -tw_pdf_t[k - 1], f1p = -tw_pdf_t[k];
  /* cubic Hermite on [x0, x1] */
  const double inc = x1 - x0, t = (x - x0) / inc;
  const double a0 = f0, a1 = f0p * inc;
  const double cc0 = f1 - (a0 + a1), cc1 = f1p * inc - a1;
  const double a2 = 3.0 * cc0 - cc1, a3 = cc1 - 2.0 * cc0;
  return ((a3 * t + a2) * t + a1) * t + a0;
}

double st_twnorm(double lam, double p, double n) {
  if (n < 0.0 || p < 0.0) return -10.0;
  if (n < p) return st_twnorm(lam, n, p);
  const double y1 = sqrt(n - 1) + sqrt(p);
  const double mu = y1 * y1;
  const double y2 = 1.0 / sqrt(n - 1) + 1.0 / sqrt(p);
  const double phi = y1 * pow(y2, 1.0 / 3.0);
  return (lam - mu) / phi;
}

static double seqsum(const double *a, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i];
  return s;
}

double st_dotwcalc(const double *lambda, int m, double *ptw, double *pzn, int minm) {
  if (m < minm) {
    *pzn = *ptw = -1;
    return -1.0;
  }
  const double lsum = seqsum(lambda, m);
  if (lsum <= 0.0) {
    *pzn = *ptw = -1;
    return -1.0;
  }
  const double tzn = *pzn, tm = (double)m;
  double y = (double)m / lsum;
  if (tzn > 0.0) {
    const double tw = st_twnorm(lambda[0] * y * tzn, tm, tzn);
    *ptw = tw;
    return st_twtail(tw);
  }
  double s2 = 0.0;
  for (int i = 0; i < m; ++i) {
    const double e = lambda[i] * y;
    s2 += e * e;
  }
  const double zn = (double)(m * (m + 2)) / (s2 - (double)m);
  const double tw = st_twnorm(lambda[0] * y * zn, tm, zn);
  *pzn = zn;
  *ptw = tw;
  return st_twtail(tw);
}

int st_numgtz(const double *a, int n) {
  int k = 0;
  for (int i = 0; i < n; ++i) k += a[i] > 0.000001;
  return k;
}

void st_print_tw(FILE *f, const double *lambda, int m, double znval, double nrows, int minm) {
  fprintf(f, "%4s  %12s", "#N", "eigenvalue");
  fprintf(f, "%12s", "difference");
  fprintf(f, " %9s %12s", "twstat", "p-value");
  fprintf(f, " %9s", "effect. n");
  fprintf(f, "\n");
  for (int i = 0; i < m; ++i) {
    double zn = znval, tw;
    if (zn > 0 && nrows > zn) zn = nrows;
    const double tail = st_dotwcalc(lambda + i, m - i, &tw, &zn, minm);
    fprintf(f, "%4d  %12.6f", i + 1, lambda[i]);
    if (i == 0) fprintf(f, "%12s", "NA");
    else fprintf(f, "%12.6f", lambda[i] - lambda[i - 1]);
    if (tail >= 0.0) fprintf(f, " %9.3f %12.6g", tw, tail);
    else fprintf(f, " %9s %12s", "NA", "NA");
    if (zn > 0.0) fprintf(f, " %9.3f", zn);
    else fprintf(f, " %9s", "NA");
    fprintf(f, "\n");
  }
}

/* ------------------------------------------------------------------ kurtosis */
double st_kurtosis(const double *a, int n) {
  const double y1 = seqsum(a, n) / (double)n;
  double y2 = 0.0;
  for (int i = 0; i < n; ++i) y2 += (a[i] - y1) * (a[i] - y1);
  y2 /= (double)n;
  const double r = 1.0 / sqrt(y2);
  double y4 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double w = (a[i] - y1) * r;
    y4 += (w * w) * (w * w);
  }
  return y4 / (double)n;
}

/* ------------------------------------------------------------------ anova */
static void calcmean(double *wmean, const double *vec, int len, const int *xt, int numeg) {
  double *pop = calloc(numeg, sizeof(double));
  const double y1 = seqsum(vec, len) / (double)len;
  for (int k = 0; k < numeg; ++k) wmean[k] = 0.0;
  for (int i = 0; i < len; ++i) {
    pop[xt[i]] += 1.0;
    wmean[xt[i]] += vec[i] - y1;
  }
  for (int k = 0; k < numeg; ++k) wmean[k] /= pop[k] + 1.0e-12;
  free(pop);
}

/* deviations from the grand mean, the population means of those, and the residuals */
static void anova_parts(const double *vec, int len, const int *xt, int numeg, double *w0,
                        double *wmean, double *pop, double *w1) {
  const double y1 = seqsum(vec, len) / (double)len;
  for (int k = 0; k < numeg; ++k) wmean[k] = pop[k] = 0.0;
  for (int i = 0; i < len; ++i) {
    w0[i] = vec[i] + (-y1);
    pop[xt[i]] += 1.0;
    wmean[xt[i]] += w0[i];
  }
  for (int k = 0; k < numeg; ++k) {
    pop[k] += 1.0e-12;
    wmean[k] /= pop[k];
  }
  for (int i = 0; i < len; ++i) w1[i] = w0[i] - wmean[xt[i]];
}

double st_anova(const double *vec, int len, const int *xt, int numeg) {
  if (numeg >= len) {
    printf("*** warning: bad anova popsizes too small\n");
    return 0.0;
  }
  double *w0 = malloc(sizeof(double) * len), *w1 = malloc(sizeof(double) * len);
  double *wm = malloc(sizeof(double) * numeg), *pop = malloc(sizeof(double) * numeg);
  anova_parts(vec, len, xt, numeg, w0, wm, pop, w1);
  double top = 0.0, bot = 0.0;
  for (int k = 0; k < numeg; ++k) top += (wm[k] * wm[k]) * pop[k];
  for (int i = 0; i < len; ++i) bot += w1[i] * w1[i];
  bot /= (double)(len - numeg);
  bot *= (double)(numeg - 1);
  const double t = st_ftail(numeg - 1, len - numeg, top / bot);
  free(w0); free(w1); free(wm); free(pop);
  return t;
}

double st_anova1(const double *vec, int len, const int *xt, int numeg) {
  double *w0 = malloc(sizeof(double) * len), *w1 = malloc(sizeof(double) * len);
  double *wm = malloc(sizeof(double) * numeg), *pop = malloc(sizeof(double) * numeg);
  anova_parts(vec, len, xt, numeg, w0, wm, pop, w1);
  double y1 = 0.0, y2 = 0.0;
  for (int i = 0; i < len; ++i) {
    y1 += w0[i] * w0[i];
    y2 += w1[i] * w1[i];
  }
  y1 /= (double)len;
  y2 /= (double)len;
  free(w0); free(w1); free(wm); free(pop);
  return 0.5 * (double)len * log(y1 / y2);
}

static int published = 0;

static void publishit(const char *sss, double tail) {
  char s2[4096];
  snprintf(s2, sizeof s2, "%s", sss);
  for (char *p = s2; *p; ++p)
    if (*p == ' ') *p = '_';
  if (published++ == 0) {
    printf("## Anova statistics for population differences along each eigenvector:\n");
    printf("%40s %12s\n", "", "p-value");
  }
  printf("%40s ", s2);
  printf(" %12.6g", tail);
  const char *hit = tail < 1.0e-6 ? "+++" : (tail < 0.001 ? "***" : "");
  printf(" %s\n", hit);
}

static const double *sort_key;
static int by_key(const void *a, const void *b) {
  const double x = sort_key[*(const int *)a], y = sort_key[*(const int *)b];
  return x < y ? -1 : (x > y ? 1 : 0);
}

void st_dottest(const char *sss, const double *vec, char *const *eglist, int numeg,
                const int *xt, int len, double *chitot, int pubmean) {
  double *wm = malloc(sizeof(double) * numeg), *w1 = malloc(sizeof(double) * (len + numeg));
  int *isort = malloc(sizeof(int) * numeg), *x2 = malloc(sizeof(int) * len);
  calcmean(wm, vec, len, xt, numeg);
  for (int k = 0; k < numeg; ++k) isort[k] = k;
  sort_key = wm;
  qsort(isort, numeg, sizeof(int), by_key);
  if (pubmean) {
    printf("%s:means\n", sss);
    for (int i = 0; i < numeg; ++i) {
      printf("%20s ", eglist[isort[i]]);
      printf(" %9.3f\n", wm[isort[i]]);
    }
  }
  int imax = 0, imin = 0; /* first extreme (vlmaxmin) */
  for (int k = 1; k < numeg; ++k) {
    if (wm[k] > wm[imax]) imax = k;
    if (wm[k] < wm[imin]) imin = k;
  }
  const double ans = st_anova(vec, len, xt, numeg);
  char buf[4096];
  if (numeg > 2) {
    snprintf(buf, sizeof buf, "%s %s ", sss, "overall");
    publishit(buf, ans);
    printf(" %20s minv: %9.3f %20s maxv: %9.3f\n", eglist[imin], wm[imin], eglist[imax], wm[imax]);
  }
  for (int k1 = 0; k1 < numeg; ++k1)
    for (int k2 = k1 + 1; k2 < numeg; ++k2) {
      int n = 0, c1 = 0, c2 = 0;
      for (int i = 0; i < len; ++i) {
        if (xt[i] == k1) { w1[n] = vec[i]; x2[n++] = 0; ++c1; }
        if (xt[i] == k2) { w1[n] = vec[i]; x2[n++] = 1; ++c2; }
      }
      if (c1 <= 1 || c2 <= 1) continue;
      chitot[k1 * numeg + k2] += 2.0 * st_anova1(w1, n, x2, 2);
      const double t = st_anova(w1, n, x2, 2);
      snprintf(buf, sizeof buf, "%s %s %s ", sss, eglist[k1], eglist[k2]);
      publishit(buf, t);
    }
  free(wm); free(w1); free(isort); free(x2);
}

/* ------------------------------------------------------------------ Fst */
int st_setblocks(const int *chrom, const double *gpos, const uint8_t *ignore, int64_t n,
                 double blocklen, int *tag) {
  int nb = 0, lchrom = -1, xsize = 0;
  double fpos = -1.0e20;
  for (int64_t i = 0; i < n; ++i) {
    tag[i] = -1;
    if (ignore[i]) continue;
    const double dis = gpos[i] - fpos;
    if (chrom[i] != lchrom || dis >= blocklen) {
      if (xsize > 0) ++nb;
      lchrom = chrom[i];
      fpos = gpos[i];
      xsize = 0;
    }
    tag[i] = nb;
    ++xsize;
  }
  if (xsize > 0) ++nb;
  return nb;
}

/* weighted block jackknife (Busing et al.; the reference's wjackestx) over the blocks
   of positive weight */
static void wjack(double *est, double *sig, double mean, const double *jm, const double *jw,
                  int g) {
  double *m = malloc(sizeof(double) * g), *w = malloc(sizeof(double) * g);
  int n = 0;
  for (int i = 0; i < g; ++i)
    if (jw[i] >= 1.0e-6) { m[n] = jm[i]; w[n] = jw[i]; ++n; }
  double yn = 0.0, td = 0.0, wm = 0.0;
  for (int i = 0; i < n; ++i) yn += w[i];
  for (int i = 0; i < n; ++i) td += -(m[i] + (-mean));
  for (int i = 0; i < n; ++i) wm += w[i] * m[i];
  const double jk = td + wm / yn;
  double v = 0.0;
  for (int i = 0; i < n; ++i) {
    const double hh = yn / w[i];
    double x = hh * mean - (hh + (-1.0)) * m[i];
    x = x + (-jk);
    v += (x * x) / (hh + (-1.0));
  }
  *est = jk;
  *sig = sqrt(v / (double)n);
  free(m); free(w);
}

void st_fst(const int32_t *counts, const uint8_t *ignore, const int *tag, int64_t ncols,
            int numeg, int nblocks, double *fstmean, double *fstest, double *fstsig, int *fstnum) {
  const int P = numeg * numeg;
  double *btop = calloc((size_t)nblocks * P, sizeof(double));
  double *bbot = calloc((size_t)nblocks * P, sizeof(double));
  double *wj = calloc(nblocks, sizeof(double)), *dj = calloc(nblocks, sizeof(double));
  double *gtop = calloc(P, sizeof(double)), *gbot = calloc(P, sizeof(double));
  for (int k = 0; k < P; ++k) { fstest[k] = fstsig[k] = 0.0; fstnum[k] = 0; }
  if (nblocks == 1) printf("number of blocks 1: no standard error\n");
  for (int64_t c = 0; c < ncols; ++c) {
    if (ignore[c] || tag[c] < 0) continue;
    const int b = tag[c];
    wj[b] += 1.0;
    const int32_t *cc = counts + c * numeg * 2;
    for (int a = 0; a < numeg; ++a)
      for (int d = a + 1; d < numeg; ++d) {
        /* fstcolyy: Hudson-style numerator and denominator per pair */
        const double ya = cc[2 * a], yb = cc[2 * a + 1], yaa = cc[2 * d], ybb = cc[2 * d + 1];
        const double z = ya + yb, zz = yaa + ybb;
        if (z < 1.5 || zz < 1.5) continue;
        const double p1 = ya / z, h1 = ya * yb / (z * (z - 1.0));
        const double p2 = yaa / zz, h2 = yaa * ybb / (zz * (zz - 1.0));
        double en = (p1 - p2) * (p1 - p2);
        en -= h1 / z;
        en -= h2 / zz;
        double ed = en;
        ed += h1;
        ed += h2;
        const int k = a * numeg + d;
        ++fstnum[k];
        btop[(size_t)b * P + k] += en;
        bbot[(size_t)b * P + k] += ed;
      }
  }
  for (int a = 0; a < numeg; ++a)
    for (int d = a + 1; d < numeg; ++d) {
      fstnum[d * numeg + a] = fstnum[a * numeg + d];
      for (int b = 0; b < nblocks; ++b) {
        btop[(size_t)b * P + d * numeg + a] = btop[(size_t)b * P + a * numeg + d];
        bbot[(size_t)b * P + d * numeg + a] = bbot[(size_t)b * P + a * numeg + d];
      }
    }
  for (int b = 0; b < nblocks; ++b)
    for (int k = 0; k < P; ++k) {
      gtop[k] += btop[(size_t)b * P + k];
      gbot[k] += bbot[(size_t)b * P + k];
    }
  for (int b = 0; b < nblocks; ++b) /* delete-block estimates */
    for (int k = 0; k < P; ++k) {
      const double wt = gtop[k] - btop[(size_t)b * P + k];
      const double wb = (gbot[k] - bbot[(size_t)b * P + k]) + 1.0e-10;
      btop[(size_t)b * P + k] = wt / wb;
    }
  for (int k = 0; k < P; ++k) fstmean[k] = gtop[k] = gtop[k] / (gbot[k] + 1.0e-10);
  for (int i = 0; i < numeg; ++i)
    for (int j = i + 1; j < numeg; ++j) {
      const double mean = gtop[i * numeg + j];
      double est = mean, sg = 0.0;
      for (int b = 0; b < nblocks; ++b) dj[b] = btop[(size_t)b * P + i * numeg + j];
      if (nblocks > 1) wjack(&est, &sg, mean, dj, wj, nblocks);
      fstest[i * numeg + j] = fstest[j * numeg + i] = est;
      fstsig[i * numeg + j] = fstsig[j * numeg + i] = sg;
    }
  free(btop); free(bbot); free(wj); free(dj); free(gtop); free(gbot);
}

/* printmatz5 (width 5, x1000) and printmatz10 (width 10, x1e6) with get3's names: the
   first character and a blank */
void st_printmatz(const double *ww, char *const *eglist, int n, double scale, int width) {
  char s3[8];
  printf(" %5s", "   ");
  for (int i = 0; i < n; ++i) {
    snprintf(s3, sizeof s3, "%c ", eglist[i][0]);
    printf(" %*s", width, s3);
  }
  printf("\n");
  for (int i = 0; i < n; ++i) {
    snprintf(s3, sizeof s3, "%c ", eglist[i][0]);
    printf("%5s", s3);
    for (int j = 0; j < n; ++j) printf(" %*d", width, (int)lrint(scale * ww[i * n + j]));
    printf("\n");
  }
}
