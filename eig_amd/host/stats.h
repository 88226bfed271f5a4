/* stats.h -- the reference smartpca's post-eigendecomposition statistics (stats.c). */
#ifndef EIG_HOST_STATS_H
#define EIG_HOST_STATS_H

#include <stdint.h>
#include <stdio.h>

double st_ibeta(double a, double b, double x);
double st_gammaq(double a, double x);
double st_ftail(int df1, int df2, double F);
double st_chitail(int df, double z);

int st_twtable(double *x, double *tail, double *pdf); /* 161 grid points */
double st_twtail(double x);
double st_twnorm(double lam, double p, double n);
double st_dotwcalc(const double *lambda, int m, double *ptw, double *pzn, int minm);
int st_numgtz(const double *a, int n);
void st_print_tw(FILE *f, const double *lambda, int m, double znval, double nrows, int minm);

double st_kurtosis(const double *a, int n);
double st_anova(const double *vec, int len, const int *xt, int numeg);
double st_anova1(const double *vec, int len, const int *xt, int numeg);
void st_dottest(const char *sss, const double *vec, char *const *eglist, int numeg,
                const int *xt, int len, double *chitot, int pubmean);

int st_setblocks(const int *chrom, const double *gpos, const uint8_t *ignore, int64_t n,
                 double blocklen, int *tag);
void st_fst(const int32_t *counts, const uint8_t *ignore, const int *tag, int64_t ncols,
            int numeg, int nblocks, double *fstmean, double *fstest, double *fstsig, int *fstnum);
void st_printmatz(const double *ww, char *const *eglist, int n, double scale, int width);

#endif
