/*
 * integral_hip.h -- C ABI of the MI355X (gfx950) implementation of the
 * two-electron hot path of the Rust library `integral` (v0.4.3).
 *
 * Plain C types only (pointers, sizes, doubles); no HIP or torch types cross
 * this boundary except the opaque `void *stream` (a hipStream_t, or NULL for
 * the default stream).  Every entry point returns an `int` status
 * (IG_OK == 0) and never unwinds or aborts across the ABI; the message of the
 * last failure on the calling thread is available from ig_last_error_message().
 *
 * Conventions are the reference's (crates/integral/src):
 *   - Cartesian order lx descending then ly descending (math/am.rs:1-31);
 *   - primitive normalisation N(alpha; l,0,0), contractions not renormalised,
 *     zero exponents exempt (shell.rs:34-47, :182-187, math/norm.rs:42-49);
 *   - spherical components in m order -l..l, p = (x,y,z)
 *     (math/solid_harmonics.rs:145-172, spherical.rs:183-194);
 *   - AO offsets by n_func in shell order; atoms = distinct bitwise-equal
 *     centres in order of first appearance (shell.rs:226-275);
 *   - dense tensors row-major ((i*n+j)*n+k)*n+l (integrals.rs:641-662);
 *   - canonical shell pairs (i,j), i>=j, index i(i+1)/2+j (integrals.rs:1252-1260).
 * J/K/gradient outputs are ACCUMULATED (+=) as in direct.rs:398-402 and
 * direct_grad.rs:476-484; the caller zeroes them.
 */
#ifndef INTEGRAL_HIP_H
#define INTEGRAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: one per IntegralError variant (lib.rs:83-123) and one per
 * panic category of the direct contractors (direct.rs:141-159,
 * direct_grad.rs:337-346, integrals.rs:316-321). */
enum {
  IG_OK = 0,
  IG_ERR_EMPTY_CONTRACTION = 1,        /* IntegralError::EmptyContraction */
  IG_ERR_MISMATCHED_CONTRACTION = 2,   /* IntegralError::MismatchedContraction */
  IG_ERR_L_TOO_HIGH = 3,               /* IntegralError::AngularMomentumTooHigh */
  IG_ERR_L_TOO_HIGH_FOR_GRADIENT = 4,  /* IntegralError::AngularMomentumTooHighForGradient */
  IG_ERR_GAMMA_LENGTH = 5,             /* IntegralError::GammaLengthMismatch */
  IG_ERR_NO_DENSITY = 10,              /* "needs at least one density" */
  IG_ERR_TOO_MANY_DENSITIES = 11,      /* "at most 64 simultaneous densities" */
  IG_ERR_LENGTH_MISMATCH = 12,         /* assert_eq! on buffer lengths */
  IG_ERR_NON_CANONICAL_BRA = 13,       /* bra not a pair from bra_pairs() */
  IG_ERR_FOREIGN_DENSITY = 14,         /* density prepared by another plan */
  IG_ERR_INVALID_OMEGA = 15,           /* EriKernel::Erf with omega <= 0 / NaN */
  IG_ERR_INVALID_ARGUMENT = 16,        /* null pointer, bad enum, bad range */
  IG_ERR_UNSUPPORTED = 17,             /* valid input outside what this build implements */
  IG_ERR_DEVICE = 20,                  /* HIP runtime failure (allocation, launch) */
};

/* Engine knob, integrals.rs:330-338.  The GPU evaluates every class with its
 * own Rys-quadrature kernels; results are engine-transparent to the
 * reference's cross-engine tolerance (integrals.rs:323-329). */
enum { IG_ENGINE_AUTO = 0, IG_ENGINE_OS_HGP = 1, IG_ENGINE_RYS = 2 };

/* ShellKind, shell.rs:17-27 */
enum { IG_CARTESIAN = 0, IG_SPHERICAL = 1 };

/* Returns the ABI version; replaces integral-sys/src/lib.rs:19-23 (which
 * returns 0).  This build's ABI is version 1. */
uint32_t integral_abi_version(void);

/* Message for the last non-OK status returned on this thread. */
const char *ig_last_error_message(void);

/* Number of visible HIP devices (0 when no GPU); never fails. */
int ig_device_count(void);

/* ---- the quadrature building blocks, evaluated on the device ------------- */
/* boys_array (math/boys.rs:146-193): F_0..F_m_max (m_max <= 25) at each of the
 * n points T (host arrays); out receives n*(m_max+1) doubles, point-major. */
int ig_boys_eval(int m_max, const double *T, size_t n, double *out);
/* rys_roots_weights (math/rys.rs:174-185): the nroots-point rule (1..13) at
 * each T; roots (squared, ascending) and weights receive n*nroots doubles. */
int ig_rys_eval(int nroots, const double *T, size_t n, double *roots, double *weights);

/* ---- basis (shell.rs:64-118 Shell::with_kind, :205 Basis::new) ----------- */
typedef struct ig_shell_desc {
  int32_t l;             /* angular momentum, 0..=6 (engine/os.rs:34 MAX_L) */
  int32_t kind;          /* IG_CARTESIAN or IG_SPHERICAL */
  double center[3];      /* bohr */
  int32_t n_exponents;   /* number of exponents */
  int32_t n_coefficients;/* number of raw coefficients (must equal n_exponents) */
  const double *exponents;
  const double *coefficients; /* raw, unnormalised contraction coefficients */
} ig_shell_desc;

typedef struct ig_basis ig_basis;

/* Validates every shell like Shell::with_kind and copies all data. */
int ig_basis_create(const ig_shell_desc *shells, size_t n_shells, ig_basis **out);
void ig_basis_destroy(ig_basis *basis);
size_t ig_basis_nao(const ig_basis *basis);      /* Basis::nao, shell.rs:236-240 */
size_t ig_basis_nao_cart(const ig_basis *basis); /* Basis::nao_cart, shell.rs:229-233 */
size_t ig_basis_natom(const ig_basis *basis);    /* Basis::atoms().len(), shell.rs:258-266 */
size_t ig_basis_nshell(const ig_basis *basis);

/* ---- dense ERI tensor and Schwarz bounds --------------------------------- */
/* Basis::eri_with, integrals.rs:671-754: out has nao^4 doubles (host memory),
 * overwritten. */
int ig_eri_dense(const ig_basis *basis, int engine, double *out, size_t out_len, void *stream);
/* Basis::eri_screened_with, integrals.rs:874-983: quartets with
 * Q_ij*Q_kl < tau are left zero; stats[0]=total canonical quartets,
 * stats[1]=skipped (may be NULL). */
int ig_eri_screened(const ig_basis *basis, int engine, double tau, double *out, size_t out_len,
                    int64_t *stats, void *stream);
/* EriBuilder::fill (eri_builder.rs:337-454) for the canonical bra pairs
 * [bra_begin, bra_end) (bra_end < 0: all): out is the caller's nao^4 host
 * tensor (zeroed, as EriBuilder::partition requires); only the rows (mu,nu)
 * those bra pairs own -- mu in shell i and nu in shell j, or the reverse --
 * are written, every element equal to Basis::eri's to round-off (the
 * reference's contract, eri_builder.rs:419-446).  Filling every range gives
 * the whole tensor; EriBuilder::build is ig_eri_dense.  Coulomb operator;
 * ig_eri_builder_fill_kernel takes the EriKernel. */
int ig_eri_builder_fill(const ig_basis *basis, int engine, int64_t bra_begin, int64_t bra_end,
                        double *out, size_t out_len, void *stream);
/* EriBuilder::kernel(k) then fill (eri_builder.rs:280-305, 459-470): as
 * ig_eri_builder_fill with the operator `kernel` (IG_KERNEL_COULOMB, or
 * IG_KERNEL_ERF: erf(omega r)/r by the attenuated Rys transform, omega finite
 * and > 0, else IG_ERR_INVALID_OMEGA). */
int ig_eri_builder_fill_kernel(const ig_basis *basis, int engine, int kernel, double omega,
                               int64_t bra_begin, int64_t bra_end, double *out, size_t out_len,
                               void *stream);
/* Basis::eri_kernel (integrals.rs:773-830): the dense nao^4 tensor of the
 * operator `kernel` (Coulomb: ig_eri_dense), host memory, overwritten. */
int ig_eri_kernel(const ig_basis *basis, int kernel, double omega, double *out, size_t out_len,
                  void *stream);
/* Basis::schwarz_bounds_with, integrals.rs:842-862: nshell^2 doubles, host. */
int ig_schwarz_bounds(const ig_basis *basis, int engine, double *out, size_t out_len,
                      void *stream);

/* ---- density fitting (df.rs): 2- and 3-center Coulomb integrals ----------
 * Evaluated by the 4-center kernels with a unit-s dummy (zero exponent,
 * coefficient 1: the constant function) on the aux index (df.rs:1-39). */
/* Basis::eri_2c_with (df.rs:56-90): the [naux, naux] metric over `aux`,
 * exactly symmetric (each canonical aux shell pair once, mirrored). */
int ig_eri_2c(const ig_basis *aux, int engine, double *out, size_t len, void *stream);
/* Basis::schwarz_aux_bounds_with (df.rs:257-280): sqrt(max |(P|P)|) per aux
 * shell, len = number of aux shells. */
int ig_schwarz_aux_bounds(const ig_basis *aux, int engine, double *out, size_t len,
                          void *stream);
/* Eri3cBuilder::fill (df.rs:463-512) for the canonical main-basis bra pairs
 * [bra_begin, bra_end) (bra_end < 0: all, i.e. Eri3cBuilder::build): out is
 * the caller's [nao, nao, naux] host tensor, P fastest (zeroed, as
 * Eri3cBuilder::partition requires); the rows (mu in i, nu in j) and, for
 * i != j, (nu, mu) of those pairs are written. */
int ig_eri_3c_fill(const ig_basis *main, const ig_basis *aux, int engine, int64_t bra_begin,
                   int64_t bra_end, double *out, size_t len, void *stream);
/* Basis::eri_2c_kernel (df.rs:95-128): ig_eri_2c with the operator `kernel`
 * (erf: omega finite and > 0, else IG_ERR_INVALID_OMEGA). */
int ig_eri_2c_kernel(const ig_basis *aux, int kernel, double omega, double *out, size_t len,
                     void *stream);
/* Basis::eri_3c_block_kernel (df.rs:190-240) over bra-pair ranges: ig_eri_3c_fill
 * with the operator `kernel`. */
int ig_eri_3c_fill_kernel(const ig_basis *main, const ig_basis *aux, int engine, int kernel,
                          double omega, int64_t bra_begin, int64_t bra_end, double *out,
                          size_t len, void *stream);

/* ---- packed symmetry-unique ERIs: PackedEriBuilder (packed_eri.rs) -------
 * AO pairs (p, q), p >= q, numbered shell-pair blocked: canonical shell pairs
 * (i >= j) row-major, inside a pair the function pairs (a, b) row-major with
 * b <= a when i == j (packed_eri.rs:10-21, :99-110).  The output is the
 * row-major lower triangle over pair indices: (t|u), u <= t, at t(t+1)/2 + u.
 * Canonical bra pair k owns the contiguous rows of its AO pairs. */
/* Number of AO pairs m (= nao(nao+1)/2) and PackedEriBuilder::output_len,
 * m(m+1)/2 (packed_eri.rs:140-144). */
size_t ig_packed_eri_n_ao_pairs(const ig_basis *basis);
size_t ig_packed_eri_output_len(const ig_basis *basis);
/* PackedEriBuilder::ao_pairs (packed_eri.rs:133-138): n must be the AO-pair
 * count; pairs receives (p, q) as 2n ints in storage order. */
int ig_packed_eri_ao_pairs(const ig_basis *basis, int32_t *pairs, size_t n);
/* The slice [offset, offset+len) of the output owned by canonical bra pairs
 * [bra_begin, bra_end) (bra_end < 0: all) -- PackedEriBuilder::partition
 * (packed_eri.rs:146-171) for a range of tasks. */
int ig_packed_eri_rows(const ig_basis *basis, int64_t bra_begin, int64_t bra_end, size_t *offset,
                       size_t *len);
/* PackedEriBuilder::fill for the bra pairs [bra_begin, bra_end)
 * (packed_eri.rs:173-220): every canonical quartet with that bra and a ket
 * pair <= it; with threshold > 0, quartets with Q_ij*Q_kl < threshold are
 * skipped.  rows (host, len = the range's length from ig_packed_eri_rows) is
 * overwritten: skipped and zero integrals are 0. */
int ig_packed_eri_fill(const ig_basis *basis, int engine, int64_t bra_begin, int64_t bra_end,
                       double threshold, double *rows, size_t len, void *stream);
/* PackedEriBuilder::build (packed_eri.rs:222-230): all rows; out (host) has
 * ig_packed_eri_output_len doubles, overwritten. */
int ig_packed_eri_build(const ig_basis *basis, int engine, double threshold, double *out,
                        size_t len, void *stream);

/* ---- integral-direct J/K: DirectContractor (direct.rs:39-218) ------------ */
typedef struct ig_direct ig_direct;

/* DirectContractor::with_engine (direct.rs:59-87): caches shell data, the
 * primitive-pair tables (sorted by angular class) and the Schwarz bounds on
 * the device.  Basis::direct_contractor (direct.rs:331) is engine AUTO. */
int ig_direct_create(const ig_basis *basis, int engine, ig_direct **out);
void ig_direct_destroy(ig_direct *direct);
size_t ig_direct_nao(const ig_direct *direct);
/* DirectContractor::bra_pairs: writes n_pairs (i,j) pairs as 2*n ints. */
size_t ig_direct_n_bra_pairs(const ig_direct *direct);
int ig_direct_bra_pairs(const ig_direct *direct, int32_t *pairs, size_t n_pairs);
/* DirectContractor::density_shell_bounds (direct.rs:103-125): densities are
 * ndens host pointers to nao*nao row-major matrices; bounds receives
 * ndens*nshell*nshell doubles (max |D| per ordered shell block). */
int ig_direct_density_shell_bounds(const ig_direct *direct, const double *const *densities,
                                   size_t ndens, double *bounds, size_t bounds_len);
/* DirectContractor::accumulate_bra_pair (direct.rs:131-218), one canonical
 * bra pair (i,j) per call; host pointers; J (and K if want_exchange) of each
 * density accumulated.  `bounds` as produced above (ndens*nshell^2). */
int ig_direct_accumulate_bra_pair(const ig_direct *direct, int32_t i, int32_t j,
                                  const double *const *densities, const double *bounds,
                                  size_t ndens, double threshold, int want_exchange,
                                  double *const *coulomb, double *const *exchange,
                                  void *stream);
/* Bulk entry (the performance path): every canonical bra pair with index in
 * [bra_begin, bra_end) (bra_end < 0 means all), same result as calling
 * accumulate_bra_pair on each.  Host pointers; exchange may be NULL when
 * want_exchange == 0.  Density bounds are computed internally. */
int ig_direct_jk(const ig_direct *direct, const double *const *densities, size_t ndens,
                 double threshold, int want_exchange, double *const *coulomb,
                 double *const *exchange, int64_t bra_begin, int64_t bra_end, void *stream);
/* Device-resident variant of ig_direct_jk for the full bra range: d_dens,
 * d_J, d_K are DEVICE pointers to ndens contiguous nao*nao matrices each
 * (d_K may be NULL when want_exchange == 0).  The densities' exact symmetry
 * is tested on the device and its 4-byte flag read back, which synchronises
 * `stream` once per call; the kernels then run asynchronously on `stream`.
 * ig_direct_jk_device_ws takes the symmetry from the caller and never
 * synchronises. */
int ig_direct_jk_device(const ig_direct *direct, const double *d_dens, size_t ndens,
                        double threshold, int want_exchange, double *d_J, double *d_K,
                        void *stream);
/* Number of canonical quartets accepted by the screening of the last call on
 * this plan ([0]) and primitive quartets evaluated ([1]). */
int ig_direct_last_stats(const ig_direct *direct, int64_t *stats2);

/* DirectWorkspace (direct.rs:19-33, ::new at :30): per-worker scratch.  The
 * reference lets many workers call accumulate_bra_pair(&self, &mut
 * DirectWorkspace, ...) on one shared contractor; here a workspace holds its
 * own device scratch over the contractor's read-only plan (shell data, pair
 * tables, Schwarz bounds), so calls through different workspaces run
 * concurrently, each on its own stream.  Calls through one workspace (or
 * with workspace NULL, through the contractor's own scratch) are ordered on
 * the device: a call waits for the previous call's device work even when it
 * came on another stream.  The schedule and precision policy are the
 * contractor's, read at each call.  A workspace is tied to the contractor it
 * was created for (IG_ERR_INVALID_ARGUMENT otherwise) and must be destroyed
 * before it. */
typedef struct ig_direct_workspace ig_direct_workspace;
int ig_direct_workspace_create(const ig_direct *direct, ig_direct_workspace **out);
/* Waits for the workspace's last device work, then frees it. */
void ig_direct_workspace_destroy(ig_direct_workspace *workspace);
/* DirectContractor::accumulate_bra_pair (direct.rs:131-218) with an explicit
 * workspace; as ig_direct_accumulate_bra_pair otherwise (NULL: the
 * contractor's own scratch). */
int ig_direct_accumulate_bra_pair_ws(const ig_direct *direct, ig_direct_workspace *workspace,
                                     int32_t i, int32_t j, const double *const *densities,
                                     const double *bounds, size_t ndens, double threshold,
                                     int want_exchange, double *const *coulomb,
                                     double *const *exchange, void *stream);
/* ig_direct_jk with an explicit workspace. */
int ig_direct_jk_ws(const ig_direct *direct, ig_direct_workspace *workspace,
                    const double *const *densities, size_t ndens, double threshold,
                    int want_exchange, double *const *coulomb, double *const *exchange,
                    int64_t bra_begin, int64_t bra_end, void *stream);
/* Fully asynchronous device-resident J/K (the performance path): as
 * ig_direct_jk_device, with nothing read back to the host.  dens_symmetric
 * != 0 declares every density symmetric (the caller's guarantee; the
 * antisymmetric part is then not formed); 0 runs the general path, exact for
 * any density (direct_contractor.rs:30-91).  workspace may be NULL (the
 * contractor's own scratch). */
int ig_direct_jk_device_ws(const ig_direct *direct, ig_direct_workspace *workspace,
                           const double *d_dens, size_t ndens, double threshold,
                           int want_exchange, double *d_J, double *d_K, int dens_symmetric,
                           void *stream);
/* ig_direct_last_stats of the last call through a workspace (synchronises
 * that call's stream). */
int ig_direct_workspace_last_stats(ig_direct_workspace *workspace, int64_t *stats2);

/* J/K that allocates no device memory.  No counterpart in the reference: for
 * callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  As
 * ig_direct_jk_device_ws (the same kernels and results), with all call
 * scratch -- the split and Cartesian densities, block bounds, J'/K'
 * accumulators (one copy), quartet-set rows, counters and runtime-l block
 * slots -- carved from one fixed device array of 64 MiB per GPU (a
 * file-scope __device__ array, held by each GPU the library runs on until
 * the process ends).  A call needing more fails with IG_ERR_UNSUPPORTED and
 * the size in the message (a nao 520 spherical s,s,p,p,d basis with two
 * symmetric densities needs about 25 MiB).  Every call on one GPU shares the
 * array and its host uploads go on `stream`: calls that overlap on one GPU
 * (other streams, other threads, other contractors) race.  The contractor's
 * plan must exist (ig_direct_create allocates); no statistics are recorded. */
int ig_direct_jk_device_noalloc(const ig_direct *direct, const double *d_dens, size_t ndens,
                                double threshold, int want_exchange, double *d_J, double *d_K,
                                int dens_symmetric, void *stream);

/* Per-class attribution of the last J/K call (the analogue of the
 * reference's `eri_profile` classes mode, benches/examples/eri_profile.rs:
 * 396-531).  With profiling enabled each quartet-class launch is bracketed by
 * HIP events on the call's stream; ig_direct_profile_read fills up to `cap`
 * records and sets *n to the number available. */
typedef struct ig_class_profile {
  int32_t la, lb, lc, ld;   /* kernel orientation, pair class (la,lb) >= (lc,ld) */
  int32_t fp32;             /* 1: the FP32 part of a mixed-precision call */
  int32_t reserved;
  int64_t quartets;         /* canonical quartets accepted (density mask != 0) */
  int64_t prim_quartets;    /* primitive quartets evaluated */
  double ms;                /* kernel time from HIP events (0 when not profiling) */
} ig_class_profile;
int ig_direct_set_profiling(ig_direct *direct, int enable);

/* Precision policy of the whole-list J/K calls (ig_direct_jk over all bra
 * pairs, ig_direct_jk_device).  Quartets whose Schwarz product Q_ij*Q_kl is
 * below fp32_below (and at or above the threshold) run with FP32 2D tables and
 * contracted block; roots and weights, prefactors, the screening and the J/K
 * accumulation stay FP64.  0 (the default) = FP64 only.  Per-bra-pair and
 * bra-range calls are always FP64.  No counterpart in the reference. */
int ig_direct_set_mixed_precision(ig_direct *direct, double fp32_below);
int ig_direct_profile_read(const ig_direct *direct, ig_class_profile *out, size_t cap, size_t *n);

/* Class-to-kernel schedule of a plan.  Every compiled quartet class has these
 * J/K kernels: variant 0 runs one quartet per team of lanes with its 2D
 * tables in LDS, variant 1 one quartet per lane in registers, variant 2 the
 * register kernel with the bra row's Coulomb and exchange sums kept in LDS
 * (used for the whole-list calls; per-bra-pair calls fall back to 1),
 * variant 3, for total angular momentum <= 2, variant 2 with an Obara-Saika /
 * Head-Gordon-Pople block from the Boys function instead of Rys roots, and
 * variant 4 the team kernel with only the vertical recurrence per primitive
 * and the horizontal transfers once per contracted quartet (classes of more
 * than 162 block elements with a p or higher shell in position b or d).  Classes with f shells have the
 * team kernels (up to 3600 block elements) and, up to 270, the lane ones.
 * The reference
 * has no counterpart (its CPU engines pick by angular momentum only,
 * integrals.rs:323-329); this replaces a fixed dispatch on the GPU.
 *
 * ig_direct_autotune times every variant of every class pair with work, on
 * the given device densities and threshold (min over `reps` launches each),
 * and keeps the faster in the plan.  Trial results go to plan scratch; no
 * caller buffer is written.  Synchronises `stream`. */
typedef struct ig_schedule_entry {
  int32_t la, lb, lc, ld;  /* kernel orientation, pair class (la,lb) >= (lc,ld) */
  int32_t variant;         /* 0 team/LDS, 1 lane/registers, 2 registers + LDS row sums,
                              3 as 2 with an OS/HGP block, 4 team with contracted
                              transfers; -1 on write = default */
  int32_t tuned;           /* read: 1 if set by the autotuner or a write */
  double ms;               /* read: the autotuner's time for the chosen variant */
} ig_schedule_entry;
int ig_direct_autotune(ig_direct *direct, const double *d_dens, size_t ndens, double threshold,
                       int want_exchange, int reps, void *stream);
/* All compiled class pairs present in the plan's basis, with the variant in
 * use; *n = the number available, up to `cap` written. */
int ig_direct_schedule_read(const ig_direct *direct, ig_schedule_entry *out, size_t cap,
                            size_t *n);
/* Load a schedule (e.g. saved from an earlier autotune of the same basis). */
int ig_direct_schedule_write(ig_direct *direct, const ig_schedule_entry *in, size_t n);

/* ---- Batched J/K over many molecules -------------------------------------
 * No counterpart in the reference, whose DirectJkContractor serves one basis
 * per object (direct.rs:75-131); a caller with many small molecules (a
 * screening or dataset workload) would loop over contractors.  Here one
 * object holds a plan per basis and runs every molecule's J/K in one HIP
 * graph: per quartet class one persistent launch whose workgroups take
 * chunks of the concatenated quartet lists of all molecules from a device
 * work queue.  The graph is captured on the first call and replayed while
 * threshold, exchange flag, precision policy and the device pointers stay the
 * same; any change re-captures it. */
typedef struct ig_batch ig_batch;
/* One plan per basis (as ig_direct_create with IG_ENGINE_AUTO). */
int ig_batch_create(const ig_basis *const *bases, size_t n, ig_batch **out);
void ig_batch_destroy(ig_batch *batch);
size_t ig_batch_size(const ig_batch *batch);
/* As ig_direct_set_mixed_precision, for every molecule of the batch. */
int ig_batch_set_mixed_precision(ig_batch *batch, double fp32_below);
/* J (and K when want_exchange) of one density per molecule: d_dens[m],
 * d_J[m], d_K[m] are DEVICE pointers to nao_m*nao_m matrices; J and K are
 * accumulated into them.  Densities are taken as symmetric: the result is
 * that of (D + D^T)/2.  Asynchronous on
 * `stream` (the graph runs on the batch's own stream, ordered after the work
 * queued on `stream` before the call and before the work queued after it).
 * Calls on one batch are serialised; a batch is not shared between threads
 * that need them to overlap. */
int ig_batch_jk_device(ig_batch *batch, const double *const *d_dens, double threshold,
                       int want_exchange, double *const *d_J, double *const *d_K, void *stream);
/* Class-to-kernel schedule for the whole batch: ig_direct_autotune on
 * molecule `molecule` (d_dens: its nao*nao device density), then the chosen
 * variants copied to every molecule's plan.  The next ig_batch_jk_device
 * captures the graph again.  Synchronises `stream`. */
int ig_batch_autotune(ig_batch *batch, size_t molecule, const double *d_dens, double threshold,
                      int want_exchange, int reps, void *stream);
/* Load a schedule (as read from an ig_direct plan, e.g. one autotuned on a
 * molecule like the batch's) into every molecule's plan; the next call
 * captures the graph again. */
int ig_batch_schedule_write(ig_batch *batch, const ig_schedule_entry *in, size_t n);
/* Quartets accepted ([0]) and primitive quartets ([1]) of the last call,
 * summed over the molecules.  Synchronises the batch's stream. */
int ig_batch_last_stats(ig_batch *batch, int64_t *stats2);
/* Diagnostic with no counterpart in the reference: prepares the device views
 * and work lists that ig_batch_jk_device would capture into its graph for
 * (threshold, want_exchange, the batch's precision policy), runs no kernel,
 * and walks every quartet of every batched row-sum launch on the host with
 * that kernel's indexing (chunk -> molecule decode, row cursor, pair, shell
 * and matrix indices, LDS row-sum columns), checking each baked pointer
 * against its plan buffer.  IG_OK when every access is in range; otherwise
 * IG_ERR_DEVICE with the first offending access, also written to msg (cap
 * bytes, NUL-terminated; msg may be null).  For tests of the batched path. */
int ig_batch_check(ig_batch *batch, double threshold, int want_exchange, char *msg, size_t cap);

/* ---- ERI-derivative contraction: DirectGradContractor (direct_grad.rs) --- */
typedef struct ig_grad ig_grad;
typedef struct ig_grad_density ig_grad_density;

/* DirectGradContractor::with_engine (direct_grad.rs:180-247); fails with
 * IG_ERR_L_TOO_HIGH_FOR_GRADIENT when a shell has l > 5 (grad.rs:50). */
int ig_grad_create(const ig_basis *basis, int engine, ig_grad **out);
void ig_grad_destroy(ig_grad *grad);
size_t ig_grad_natom(const ig_grad *grad);
size_t ig_grad_nao(const ig_grad *grad);
/* DirectGradContractor::density (direct_grad.rs:272-290): Coulomb-type terms
 * factor*left[mn]*right[ls] and exchange-type factor*left[ml]*right[ns];
 * left/right are symmetric nao*nao host matrices, copied to the device. */
int ig_grad_density_create(const ig_grad *grad, size_t n_coulomb, const double *coulomb_factors,
                           const double *const *coulomb_left, const double *const *coulomb_right,
                           size_t n_exchange, const double *exchange_factors,
                           const double *const *exchange_left,
                           const double *const *exchange_right, ig_grad_density **out);
void ig_grad_density_destroy(ig_grad_density *density);
/* EriKernel (integrals.rs:304-312): IG_KERNEL_ERF takes omega, which must be
 * finite and > 0 (integrals.rs:316-321), else IG_ERR_INVALID_OMEGA. */
enum { IG_KERNEL_COULOMB = 0, IG_KERNEL_ERF = 1 };
/* DirectGradContractor::accumulate_bra_pair (direct_grad.rs:328-373):
 * out is natom*3 doubles (host), accumulated. */
int ig_grad_accumulate_bra_pair(const ig_grad *grad, const ig_grad_density *density, int32_t i,
                                int32_t j, int kernel, double omega, double threshold,
                                double *out, size_t out_len, void *stream);
/* DirectGradContractor::contract (direct_grad.rs:380-392): all bra pairs,
 * accumulated into out (natom*3, host). */
int ig_grad_contract(const ig_grad *grad, const ig_grad_density *density, int kernel,
                     double omega, double threshold, double *out, size_t out_len, void *stream);
int ig_grad_last_stats(const ig_grad *grad, int64_t *stats2);
/* Per-class attribution of the last gradient contraction, as
 * ig_direct_set_profiling / ig_direct_profile_read for J/K (quartets and
 * primitive quartets per class launch; kernel time from HIP events when
 * profiling is enabled). */
int ig_grad_set_profiling(ig_grad *grad, int enable);
int ig_grad_profile_read(const ig_grad *grad, ig_class_profile *out, size_t cap, size_t *n);
/* DirectGradWorkspace (direct_grad.rs:51-67): per-worker scratch over a
 * gradient contractor's read-only plan, as ig_direct_workspace. */
typedef struct ig_grad_workspace ig_grad_workspace;
int ig_grad_workspace_create(const ig_grad *grad, ig_grad_workspace **out);
void ig_grad_workspace_destroy(ig_grad_workspace *workspace);
/* ig_grad_accumulate_bra_pair / ig_grad_contract through a workspace (NULL:
 * the contractor's own scratch). */
int ig_grad_accumulate_bra_pair_ws(const ig_grad *grad, ig_grad_workspace *workspace,
                                   const ig_grad_density *density, int32_t i, int32_t j,
                                   int kernel, double omega, double threshold, double *out,
                                   size_t out_len, void *stream);
int ig_grad_contract_ws(const ig_grad *grad, ig_grad_workspace *workspace,
                        const ig_grad_density *density, int kernel, double omega,
                        double threshold, double *out, size_t out_len, void *stream);

/* ---- dense two-particle-density gradient: Basis::eri_grad_contract ------ */
/* Basis::eri_grad_contract / eri_grad_contract_with / eri_grad_contract_kernel
 * (grad.rs:417-466, 493-540): per atom F_c = sum over all ordered quartets of
 * gamma[mu,nu,lam,sig] * d(mu nu|lam sig)/dR_c.  d_gamma: the dense nao^4
 * row-major two-particle density on the device (gamma_len must be nao^4, else
 * IG_ERR_GAMMA_LENGTH); any gamma, symmetric or not.  out: natom*3 doubles
 * (host), overwritten, atoms in Basis::atoms order.  Fails with
 * IG_ERR_L_TOO_HIGH_FOR_GRADIENT for l > 5 and IG_ERR_INVALID_OMEGA as
 * ig_grad_contract.  Allocates two nao_cart^4 device buffers (the Cartesian
 * gamma); synchronises `stream`. */
int ig_eri_grad_contract_device(const ig_basis *basis, int engine, const double *d_gamma,
                                size_t gamma_len, int kernel, double omega, double *out,
                                size_t out_len, void *stream);

#ifdef __cplusplus
}
#endif
#endif
