/*
 * io.c -- the smartpca input formats, read into one host buffer of GENO rows.
 *
 *  .snp   ID chrom gpos ppos [ref var]    (readsnpdata, mcio.c:415-540)
 *  .bim / .map  chrom ID gpos ppos [a1 a2] (readsnpmapdata, mcio.c:545-664)
 *  .ind   ID sex pop, pop "Ignore" drops   (readinddata, mcio.c:982-1035)
 *  .fam / .ped  fam ID pat mat sex pheno   (readindpeddata, mcio.c:1040-1160)
 *  genotypes: packed GENO with header hash check (inpack, mcio.c:2769-2879),
 *  PLINK .bed (getbedgenos + ancval, mcio.c:3893-4009) or EIGENSTRAT text
 *  (ineigenstrat, mcio.c:2528-2656).  Every format ends up as GENO rows: rlen =
 *  max(48, ceil(N/4)) bytes per SNP, sample k at byte k/4, bits 7-6 .. 1-0,
 *  3 = missing; male heterozygotes on chromosome X read as missing (checkxval,
 *  mcio.c:1606-1618).
 */
#include <ctype.h>
#include <errno.h>
#include <math.h>

#include "smartpca.h"

static int ends_with(const char *s, const char *suf) {
  size_t a = strlen(s), b = strlen(suf);
  return a >= b && strcmp(s + a - b, suf) == 0;
}

static int split(char *line, char **tok, int maxt) {
  int n = 0;
  char *p = line;
  while (*p && n < maxt) {
    while (*p && isspace((unsigned char)*p)) ++p;
    if (!*p) break;
    tok[n++] = p;
    while (*p && !isspace((unsigned char)*p)) ++p;
    if (*p) *p++ = 0;
  }
  return n;
}

static int skip_word(const char *s) {
  return s[0] == '#' || strcmp(s, "SNP_ID") == 0 || strcmp(s, "Indiv_ID") == 0 ||
         strcmp(s, "Chr") == 0;
}

/* str2chrom (mcio.c:4545-4572) */
static int str2chrom(const char *s, int numchrom) {
  char ss[16];
  if (strlen(s) > 5) die("bad chrom: %s", s);
  const char *t = strstr(s, "chr") ? s + 3 : s;
  snprintf(ss, sizeof ss, "%s", t);
  for (char *p = ss; *p; ++p) *p = (char)toupper((unsigned char)*p);
  if (strcmp(ss, "X") == 0) return numchrom + 1;
  if (strcmp(ss, "Y") == 0) return numchrom + 2;
  if (strcmp(ss, "MT") == 0) return 90;
  if (strcmp(ss, "XY") == 0) return 91;
  if (strcmp(ss, "2A") == 0 || strcmp(ss, "2B") == 0) return 2;
  for (const char *p = ss; *p; ++p)
    if (!isdigit((unsigned char)*p)) return -1;
  return atoi(ss);
}

int64_t read_snps(const char *path, Snp **out, int numchrom, int *is_map) {
  FILE *f = fopen(path, "r");
  if (!f) die("can't open %s", path);
  *is_map = ends_with(path, ".map") || ends_with(path, ".bim") || ends_with(path, ".pedsnp");
  int64_t cap = 1024, n = 0;
  Snp *s = malloc(sizeof(Snp) * cap);
  char line[4096];
  char *tok[16];
  int usecm = 0;
  double maxg = -9999.0;
  while (fgets(line, sizeof line, f)) {
    int nt = split(line, tok, 16);
    if (nt == 0 || skip_word(tok[0])) continue;
    if (n == cap) {
      cap *= 2;
      s = realloc(s, sizeof(Snp) * cap);
      if (!s) die("out of memory");
    }
    Snp *p = &s[n];
    memset(p, 0, sizeof *p);
    const char *id, *ch;
    double gp, pp;
    if (*is_map) {
      if (nt < 4) die("(readsnpmapdata) bad line in %s", path);
      ch = tok[0];
      id = tok[1];
      gp = atof(tok[2]);
      pp = atof(tok[3]);
    } else {
      if (nt < 6) die("(readsnpdata) bad line: %s 6 columns required", tok[0]);
      id = tok[0];
      ch = tok[1];
      gp = atof(tok[2]);
      pp = atof(tok[3]);
    }
    if (strlen(id) >= IDLEN) die("snp ID too long: %s", id);
    strcpy(p->id, id);
    int c = str2chrom(ch, numchrom);
    if (c >= 100 || c <= 0) {
      p->ignore = 1;
      c = c <= 0 ? 0 : 99;
    }
    p->chrom = c;
    if (gp > 100) {
      if (gp > 1.0e6) die("absurd genetic distance: %s", id);
      usecm = 1;
    }
    if (gp > maxg) maxg = gp;
    p->gpos = gp;
    p->ppos = pp;
    ++n;
  }
  fclose(f);
  if (n == 0) die("no snps found: snpfname: %s", path);
  for (int64_t i = 0; i < n; ++i) {
    if (maxg <= 0.00001) s[i].gpos = 1.0e-8 * s[i].ppos;
    else if (usecm) s[i].gpos /= 100.0;
  }
  *out = s;
  return n;
}

int32_t read_inds(const char *path, Ind **out, int familynames) {
  FILE *f = fopen(path, "r");
  if (!f) die("can't open %s", path);
  int ped = ends_with(path, ".ped") || ends_with(path, ".fam") || ends_with(path, ".pedind");
  int32_t cap = 1024, n = 0;
  Ind *v = malloc(sizeof(Ind) * cap);
  char *line = malloc(1 << 20);
  char *tok[16];
  while (fgets(line, 1 << 20, f)) {
    int nt = split(line, tok, 16);
    if (nt == 0 || (ped ? tok[0][0] == '#' : skip_word(tok[0]))) continue;
    if (n == cap) {
      cap *= 2;
      v = realloc(v, sizeof(Ind) * cap);
      if (!v) die("out of memory");
    }
    Ind *p = &v[n];
    memset(p, 0, sizeof *p);
    if (ped) {
      if (nt < 6) die("%s bad line", path);
      char buf[2 * IDLEN];
      if (familynames) snprintf(buf, sizeof buf, "%s:%s", tok[0], tok[1]);
      else snprintf(buf, sizeof buf, "%s", tok[1]);
      if (strlen(buf) >= IDLEN) die("idnames too long: %s", buf);
      strcpy(p->id, buf);
      int k = atoi(tok[4]);
      p->sex = k == 1 ? 'M' : (k == 2 ? 'F' : 'U');
      const char *ph = tok[5];
      if (!strcmp(ph, "-9") || !strcmp(ph, "9") || !strcmp(ph, "0")) {
        p->ignore = 1;
        p->pop = strdup("Ignore");
        printf("%s ignored\n", p->id);
      } else if (!strcmp(ph, "1")) {
        p->pop = strdup("Control");
      } else if (!strcmp(ph, "2")) {
        p->pop = strdup("Case");
      } else {
        p->pop = strdup(ph);
      }
    } else {
      if (nt < 3) die("%s bad line", path);
      if (strlen(tok[0]) >= IDLEN) die("Indiv ID too long: %s", tok[0]);
      strcpy(p->id, tok[0]);
      p->sex = tok[1][0];
      p->pop = strdup(tok[2]);
      if (strcmp(tok[2], "Ignore") == 0) p->ignore = 1;
    }
    ++n;
  }
  free(line);
  fclose(f);
  if (n == 0) die("no indivs found: indivname: %s", path);
  *out = v;
  return n;
}

/* hasharr / hashit (admutils.c:651-682), 32-bit wraparound */
int32_t geno_hash(const char *const *ids, int64_t n) {
  uint32_t h = 0;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t t = 0;
    for (const char *c = ids[i]; *c; ++c) t = t * 23u + (uint32_t)(int32_t)(signed char)*c;
    h = h * 17u ^ t;
  }
  return (int32_t)h;
}

static void put_code(uint8_t *row, int32_t k, int g) {
  int sh = (3 - (k & 3)) << 1;
  row[k >> 2] = (uint8_t)((row[k >> 2] & ~(3 << sh)) | (g << sh));
}

/* SNP order required for packed input (mcio.c:180-226, 2805-2806) */
static void check_order(const Snp *s, int64_t n) {
  for (int64_t i = 1; i < n; ++i) {
    long a0 = s[i - 1].chrom, a1 = lround(s[i - 1].gpos * 1e6), a2 = lround(s[i - 1].ppos);
    long b0 = s[i].chrom, b1 = lround(s[i].gpos * 1e6), b2 = lround(s[i].ppos);
    int lt = (b0 < a0) || (b0 == a0 && (b1 < a1 || (b1 == a1 && b2 < a2)));
    if (lt) die("snps out of order and packed format.  Run convertf with pordercheck: NO");
  }
}

uint8_t *read_genotypes(const char *path, const Snp *snps, int64_t nsnp, const Ind *inds,
                        int32_t nind, int hashcheck, int pordercheck, int64_t *rlen_out) {
  const int64_t rlen = (2 * (int64_t)nind + 7) / 8 > 48 ? (2 * (int64_t)nind + 7) / 8 : 48;
  *rlen_out = rlen;
  FILE *f = fopen(path, "rb");
  if (!f) die("(ispack) bad open %s", path);
  uint8_t magic[8] = {0};
  size_t got = fread(magic, 1, 8, f);
  uint8_t *rows = calloc((size_t)(rlen * nsnp), 1);
  if (!rows) die("out of memory for %lld genotype rows", (long long)nsnp);
  if (got >= 4 && memcmp(magic, "GENO", 4) == 0) {
    /* packed GENO */
    if (pordercheck) check_order(snps, nsnp);
    fseek(f, 0, SEEK_SET);
    char *hdr = calloc((size_t)rlen + 1, 1);
    if (fread(hdr, 1, (size_t)rlen, f) != (size_t)rlen) die("(inpack) bad read");
    if (hashcheck) {
      int xn = 0, xs = 0;
      unsigned xih = 0, xsh = 0;
      sscanf(hdr, "GENO %d %d %x %x", &xn, &xs, &xih, &xsh);
      if (xn != nind) die("OOPS number of individuals %d != %d in input files", nind, xn);
      if (xs != nsnp) die("OOPS number of SNPs %lld != %d in input file: %s", (long long)nsnp, xs, path);
      const char **ids = malloc(sizeof(char *) * (size_t)(nind > nsnp ? nind : nsnp));
      for (int32_t i = 0; i < nind; ++i) ids[i] = inds[i].id;
      if ((uint32_t)geno_hash(ids, nind) != xih)
        die("OOPS indiv file has changed since genotype file was created");
      for (int64_t i = 0; i < nsnp; ++i) ids[i] = snps[i].id;
      if ((uint32_t)geno_hash(ids, nsnp) != xsh)
        die("OOPS snp file has changed since genotype file was created");
      free(ids);
    }
    free(hdr);
    size_t want = (size_t)(rlen * nsnp);
    if (fread(rows, 1, want, f) != want) die("(inpack) bad data read (length mismatch)");
    printf("packed geno read OK\n");
  } else if (got >= 3 && magic[0] == 0x6c && magic[1] == 0x1b && magic[2] == 0x01) {
    /* PLINK .bed, SNP-major, LSB-first; ancval: 00->2, 01->missing, 10->1, 11->0 */
    static const int map[4] = {2, 3, 1, 0};
    const int64_t blen = ((int64_t)nind + 3) / 4;
    fseek(f, 3, SEEK_SET);
    uint8_t *buf = malloc((size_t)blen);
    for (int64_t m = 0; m < nsnp; ++m) {
      if (fread(buf, 1, (size_t)blen, f) != (size_t)blen) die("(getbedgenos) premature EOF");
      uint8_t *row = rows + m * rlen;
      for (int32_t k = 0; k < nind; ++k) put_code(row, k, map[(buf[k >> 2] >> ((k & 3) << 1)) & 3]);
    }
    free(buf);
  } else {
    if (got >= 3 && magic[0] == 0x6c && magic[1] == 0x1b)
      die("bad .bed magic in %s (only SNP-major .bed is supported)", path);
    /* EIGENSTRAT text: one line per SNP, one digit per individual, 9 = missing */
    fseek(f, 0, SEEK_SET);
    char *line = malloc((size_t)nind + 4096);
    int64_t m = 0;
    while (m < nsnp && fgets(line, nind + 4096, f)) {
      size_t L = strlen(line);
      while (L && isspace((unsigned char)line[L - 1])) line[--L] = 0;
      if (L == 0) continue;
      if ((int32_t)L < nind) die("bad genotype line %lld in %s", (long long)m, path);
      uint8_t *row = rows + m * rlen;
      for (int32_t k = 0; k < nind; ++k) {
        int c = line[k] - '0';
        put_code(row, k, (c >= 0 && c <= 2) ? c : 3);
      }
      ++m;
    }
    free(line);
    if (m != nsnp) die("genotype file %s has %lld SNP lines, expected %lld", path, (long long)m,
                       (long long)nsnp);
  }
  fclose(f);
  return rows;
}
