/*
 * ig_oracle.h -- CPU restatement of the two-electron hot path of the Rust
 * library `integral` (v0.4.3), used ONLY as test infrastructure: the parity
 * checker for the HIP implementation and the timed CPU baseline.  Nothing in
 * the product (integral_amd/) includes, links or calls this code.
 *
 * Each routine cites the reference file:line it restates.  Conventions
 * (Cartesian order, normalisation, block layout, canonical enumeration) follow
 * crates/integral/src/{math/am.rs, math/norm.rs, shell.rs, integrals.rs}.
 *
 * Pinning (what checks each function independently of this code):
 *   or_boys*            mpmath golden vectors (tests/golden/boys_mpmath.json),
 *                       F_m(0)=1/(2m+1), F_0(1), large-T closed form.
 *   or_rys*             mpmath Gauss quadrature goldens (tests/golden/rys_mpmath.json),
 *                       moment identity sum w x^k = F_k, n=1 closed form.
 *   or_c2s_matrix       the reference's 17-digit mpmath c2s tables
 *                       (tests/golden/c2s_reference.json, extracted as data).
 *   or_eri_block        (ss|ss) closed form, numpy McMurchie-Davidson brute
 *                       force (tests/golden/md_eri.json), OS vs Rys cross-engine,
 *                       8-fold symmetry.
 *   or_direct_jk        dense contraction of or_eri_dense; linearity; symmetry.
 *   or_direct_grad      central finite differences of the contracted energy,
 *                       dense or_eri_grad_contract, translational invariance,
 *                       single-atom zero.
 */
#ifndef IG_ORACLE_H
#define IG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_L 6
#define OR_MAX_GRAD_L 5
#define OR_MAX_RYS 13
#define OR_BOYS_MAX_M 24

/* engine selector, integrals.rs:330-338 */
enum { OR_ENGINE_AUTO = 0, OR_ENGINE_OS = 1, OR_ENGINE_RYS = 2 };

/* status codes (mirror the library's) */
enum {
  OR_OK = 0,
  OR_ERR_EMPTY_CONTRACTION = 1,
  OR_ERR_MISMATCHED_CONTRACTION = 2,
  OR_ERR_L_TOO_HIGH = 3,
  OR_ERR_L_TOO_HIGH_FOR_GRADIENT = 4,
  OR_ERR_INVALID_ARGUMENT = 10,
};

typedef struct or_basis or_basis;

/* ---- math ---------------------------------------------------------------- */
void or_boys(int m_max, double t, double *out);           /* boys.rs:146-193 */
void or_boys_reference(int m_max, double t, double *out); /* boys.rs:366-384 */
void or_rys(int n, double t, double *roots, double *weights);           /* rys.rs:174-185 */
void or_rys_reference(int n, double t, double *roots, double *weights); /* rys.rs:197-211 */
int or_rys_table_nsub(int n);
double or_rys_table_worst(int n);
void or_c2s_matrix(int l, double *out); /* solid_harmonics.rs:218-250, (2l+1) x ncart */
double or_cart_norm(double alpha, int lx, int ly, int lz); /* norm.rs:42-49 */

/* ---- basis --------------------------------------------------------------- */
/* kinds: 0 Cartesian, 1 spherical.  exps/coefs flattened in shell order. */
int or_basis_create(int nsh, const int *l, const int *kind, const double *centers,
                    const int *nprim, const double *exps, const double *coefs,
                    or_basis **out);
void or_basis_destroy(or_basis *b);
int or_basis_nao(const or_basis *b);
int or_basis_nao_cart(const or_basis *b);
int or_basis_natom(const or_basis *b);
int or_basis_nsh(const or_basis *b);

/* ---- ERI ----------------------------------------------------------------- */
/* function-space block (ij|kl), integrals.rs:625-639 */
int or_eri_block(const or_basis *b, int engine, int i, int j, int k, int l, double *out);
/* Cartesian primitive-level blocks for engine tests: contracted, no c2s */
int or_eri_block_cart(const or_basis *b, int engine, int i, int j, int k, int l, double *out);
int or_eri_dense(const or_basis *b, int engine, double *out, int nthreads); /* integrals.rs:671-754 */
int or_schwarz(const or_basis *b, int engine, double *q, int nthreads);     /* integrals.rs:842-862 */
/* integrals.rs:874-983; stats[0]=total, stats[1]=skipped */
int or_eri_screened(const or_basis *b, int engine, double tau, double *out, int64_t *stats,
                    int nthreads);

/* ---- direct J/K (direct.rs) --------------------------------------------- */
/* bounds: ndens x nsh x nsh; direct.rs:103-125 */
int or_density_shell_bounds(const or_basis *b, int ndens, const double *const *dens,
                            double *bounds);
/* Accumulates (+=) J and K (K may be NULL when want_exchange==0) for all
 * canonical bra pairs in [bra_begin, bra_end).  schwarz may be NULL (computed).
 * stats (optional): [0]=quartets evaluated, [1]=primitive quartets evaluated.
 * direct.rs:131-218, 284-404 */
int or_direct_jk(const or_basis *b, int engine, const double *schwarz, int ndens,
                 const double *const *dens, const double *bounds, double threshold,
                 int want_exchange, double *const *J, double *const *K, int64_t bra_begin,
                 int64_t bra_end, int nthreads, int64_t *stats);

/* ---- gradient (direct_grad.rs, grad.rs) ---------------------------------- */
/* terms: factor[t], left[t], right[t]; out natom x 3 accumulated (+=).
 * direct_grad.rs:272-485 */
int or_direct_grad(const or_basis *b, int engine, const double *schwarz, int ncoul,
                   const double *cfac, const double *const *cleft,
                   const double *const *cright, int nexch, const double *xfac,
                   const double *const *xleft, const double *const *xright,
                   double threshold, double *out, int64_t bra_begin, int64_t bra_end,
                   int nthreads, int64_t *stats);
/* dense reference over all ordered quartets, grad.rs:426-466 */
int or_eri_grad_contract(const or_basis *b, int engine, const double *gamma, double *out,
                         int nthreads);

const char *or_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
