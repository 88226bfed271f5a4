/*
 * twstats_main.c -- `twstats -i evals -o out [-n nval] [-m minleneig] [-p parfile]`, the
 * reference's Tracy-Widom tool (src/eigensrc/twstats.c): reads eigenvalues (first
 * field of each line, '#' lines skipped), sorts them in decreasing order and prints
 * the Tracy-Widom statistic, p-value and effective number of markers of each
 * eigenvalue above 1e-6 in smartpca's format.  -t names the reference's table file;
 * the distribution is computed here (stats.c), so the option is accepted and unused.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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

static int desc(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return x > y ? -1 : (x < y ? 1 : 0);
}

int main(int argc, char **argv) {
  char *iname = NULL, *oname = NULL, *parname = NULL;
  double nval = -1.0;
  int minleneig = 10, opt;
  while ((opt = getopt(argc, argv, "i:o:p:n:m:t:V")) != -1) {
    switch (opt) {
      case 'i': iname = optarg; break;
      case 'o': oname = optarg; break;
      case 'p': parname = optarg; break;
      case 'n': nval = atof(optarg); break;
      case 'm': minleneig = atoi(optarg); break;
      case 't': case 'V': break;
      default: die("bad params");
    }
  }
  if (parname) {
    Params pp;
    par_read(parname, &pp);
    printf("parameter file: %s\n", parname);
    par_string(&pp, "input:", &iname);
    par_string(&pp, "output:", &oname);
    par_dbl(&pp, "nval:", &nval);
    par_int(&pp, "minleneig:", &minleneig);
  }
  if (!iname) die("i paraameter compulsory");
  FILE *f = fopen(iname, "r");
  if (!f) die("can't open %s", iname);
  size_t cap = 1024, n = 0;
  double *lam = malloc(sizeof(double) * cap);
  char line[4096];
  while (fgets(line, sizeof line, f)) {
    char *p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '#' || *p == '\n' || *p == 0) continue;
    if (n == cap) lam = realloc(lam, sizeof(double) * (cap *= 2));
    lam[n++] = atof(p);
  }
  fclose(f);
  qsort(lam, n, sizeof(double), desc);
  FILE *of = stdout;
  if (oname && !(of = fopen(oname, "w"))) die("can't open %s", oname);
  st_print_tw(of, lam, st_numgtz(lam, (int)n), nval, -1.0, minleneig);
  if (of != stdout) fclose(of);
  free(lam);
  return 0;
}
