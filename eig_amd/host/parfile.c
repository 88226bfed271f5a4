/*
 * parfile.c -- smartpca parameter files: `key: value` lines.
 *
 * Follows nicksrc/getpars.c: '#' starts a comment, a duplicate key is fatal, a key
 * that does not end in ':' draws a warning, and keys written in upper case are
 * macros substituted (longest first) into the values of the other keys
 * (dostrsub, getpars.c:378-449).  YES/NO read as 1/0 (getint, getpars.c:173-210).
 */
#include "smartpca.h"

#include <ctype.h>

static char *dupstr(const char *s) {
  char *d = strdup(s);
  if (!d) die("out of memory");
  return d;
}

static int is_upper_key(const char *k) {
  int any = 0;
  for (const char *p = k; *p && *p != ':'; ++p) {
    if (islower((unsigned char)*p)) return 0;
    if (isalpha((unsigned char)*p)) any = 1;
  }
  return any;
}

static char *replace_all(const char *s, const char *pat, const char *rep) {
  size_t lp = strlen(pat), lr = strlen(rep), n = 0;
  if (lp == 0) return dupstr(s);
  for (const char *p = strstr(s, pat); p; p = strstr(p + lp, pat)) ++n;
  char *out = malloc(strlen(s) + n * (lr + 1) + 1);
  if (!out) die("out of memory");
  char *o = out;
  const char *p = s;
  for (const char *q = strstr(p, pat); q; q = strstr(p, pat)) {
    memcpy(o, p, (size_t)(q - p));
    o += q - p;
    memcpy(o, rep, lr);
    o += lr;
    p = q + lp;
  }
  strcpy(o, p);
  return out;
}

void par_read(const char *path, Params *pp) {
  FILE *f = fopen(path, "r");
  if (!f) die("can't open %s", path);
  char line[8192];
  pp->n = 0;
  while (fgets(line, sizeof line, f)) {
    char *p = line;
    while (*p && isspace((unsigned char)*p)) ++p;
    if (!*p || *p == '#') continue;
    char *k = p;
    while (*p && !isspace((unsigned char)*p)) ++p;
    char *kend = p;
    while (*p && isspace((unsigned char)*p)) ++p;
    char *hash = strchr(p, '#');
    if (hash) *hash = 0;
    size_t vl = strlen(p);
    while (vl && isspace((unsigned char)p[vl - 1])) p[--vl] = 0;
    *kend = 0;
    if (kend[-1] != ':')
      printf("**warning: dubious parameter, please check the parameter name %s\n", k);
    for (int i = 0; i < pp->n; ++i)
      if (strcmp(pp->key[i], k) == 0) die("duplicate parameter: %s", k);
    if (pp->n >= MAXPARS) die("too many parameters in %s", path);
    pp->key[pp->n] = dupstr(k);
    pp->val[pp->n] = dupstr(p);
    ++pp->n;
  }
  fclose(f);
  /* upper-case keys are macros: substitute into the other values, longest first */
  for (;;) {
    int best = -1;
    size_t bl = 0;
    static int used[MAXPARS];
    for (int i = 0; i < pp->n; ++i)
      if (!used[i] && is_upper_key(pp->key[i]) && strlen(pp->key[i]) > bl) {
        best = i;
        bl = strlen(pp->key[i]);
      }
    if (best < 0) break;
    used[best] = 1;
    char name[512];
    snprintf(name, sizeof name, "%s", pp->key[best]);
    size_t ln = strlen(name);
    if (ln && name[ln - 1] == ':') name[ln - 1] = 0;
    for (int i = 0; i < pp->n; ++i) {
      if (i == best) continue;
      char *nv = replace_all(pp->val[i], name, pp->val[best]);
      free(pp->val[i]);
      pp->val[i] = nv;
    }
  }
}

const char *par_str(const Params *pp, const char *key) {
  for (int i = 0; i < pp->n; ++i)
    if (strcmp(pp->key[i], key) == 0) return pp->val[i];
  return NULL;
}

void par_int(const Params *pp, const char *key, int *dst) {
  const char *v = par_str(pp, key);
  if (!v) return;
  if (strcasecmp(v, "YES") == 0 || strcasecmp(v, "Y") == 0 || strcasecmp(v, "TRUE") == 0)
    *dst = 1;
  else if (strcasecmp(v, "NO") == 0 || strcasecmp(v, "N") == 0 || strcasecmp(v, "FALSE") == 0)
    *dst = 0;
  else
    *dst = atoi(v);
}

void par_dbl(const Params *pp, const char *key, double *dst) {
  const char *v = par_str(pp, key);
  if (v) *dst = atof(v);
}

void par_string(const Params *pp, const char *key, char **dst) {
  const char *v = par_str(pp, key);
  if (v) *dst = dupstr(v);
}
