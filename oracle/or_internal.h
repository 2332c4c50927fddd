/* or_internal.h -- oracle (test infrastructure only): shared internals. */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H

#include "ig_oracle.h"

/* A shell as one ERI engine call sees it: the angular momentum may differ from
 * the basis shell's (gradient raise/lower) and the coefficients may be the
 * 2-alpha weighted ones (direct_grad.rs:554-620). */
typedef struct {
  int l;
  const double *center;
  int nprim;
  const double *exps;
  const double *coefs; /* effective coefficients */
} or_shref;

typedef struct {
  int l, kind, nprim;
  double center[3];
  const double *exps, *coefs, *eff, *eff2a;
  int nfunc, ncart, off, atom;
  double *c2s; /* nfunc x ncart (already scaled by monomial->raw) or NULL */
} or_shell;

struct or_basis {
  int nsh, nao, nao_cart, natom;
  or_shell *sh;
  double *store;
  double *atoms; /* natom x 3 */
};

static inline int or_ncart(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) / 2; }
/* number of Cartesian triples of degree 0..l (os_eri.rs:96-99) */
static inline int or_naddr(int l) { return (l + 1) * (l + 2) * (l + 3) / 6; }
/* am.rs:51-55 */
static inline int or_cart_index(const int *c) {
  int d = c[1] + c[2];
  return d * (d + 1) / 2 + c[2];
}
/* os_eri.rs:108-117 */
static inline int or_addr(const int *c) {
  int n = c[0] + c[1] + c[2];
  return n * (n + 1) * (n + 2) / 6 + or_cart_index(c);
}
const int (*or_cart_comps(int l))[3];

int or_select_engine(int ne, int nf, int degree);
void or_cart_quartet(int engine, const or_shref *s, double *out, int64_t *nprimq);
void or_set_error(const char *msg);

#endif
