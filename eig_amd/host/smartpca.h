/* smartpca.h -- host side of the MI355X smartpca drop-in (C11). */
#ifndef SMARTPCA_HOST_H
#define SMARTPCA_HOST_H

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "eig_abi.h"

#define MAXPARS 512
#define IDLEN 64

typedef struct {
  char *key[MAXPARS];
  char *val[MAXPARS];
  int n;
} Params;

typedef struct {
  char id[IDLEN];
  int chrom;
  double gpos, ppos;
  int ignore;
} Snp;

typedef struct {
  char id[IDLEN];
  char sex;
  char *pop;
  int ignore;
  int aff; /* defines the axes (affstatus) */
} Ind;

void die(const char *fmt, ...);

void par_read(const char *path, Params *pp);
const char *par_str(const Params *pp, const char *key);
void par_int(const Params *pp, const char *key, int *dst);
void par_dbl(const Params *pp, const char *key, double *dst);
void par_string(const Params *pp, const char *key, char **dst);

/* readers (io.c) */
int64_t read_snps(const char *path, Snp **out, int numchrom, int *is_map);
int32_t read_inds(const char *path, Ind **out, int familynames);
int32_t geno_hash(const char *const *ids, int64_t n);
uint8_t *read_genotypes(const char *path, const Snp *snps, int64_t nsnp, const Ind *inds,
                        int32_t nind, int hashcheck, int pordercheck, int64_t *rlen_out);

static inline int geno_code(const uint8_t *row, int32_t k) {
  return (row[k >> 2] >> ((3 - (k & 3)) << 1)) & 3;
}

#endif
