/*
 * meanas_fdfd.h -- C ABI of the MI355X (gfx950) FDFD solve library libmeanas_fdfd.so.
 *
 * Drop-in boundary: meanas.fdfd.solvers.generic(omega, dxes, J, epsilon, mu, *, pec,
 * pmc, adjoint, matrix_solver, matrix_solver_opts, E_guess) -> E
 * (reference meanas/fdfd/solvers.py:66-141).  The Python package meanas_amd.fdfd.solvers
 * binds these entry points with ctypes; a C/C++ caller can bind them directly.
 *
 * Conventions
 *  - Every field vector is complex, interleaved (re, im), in the C-order layout
 *    (3, nx, ny, nz) with z fastest: meanas.fdmath.vec (fdmath/vectorization.py:51-68).
 *  - All pointers inside the descriptors and all vector arguments are DEVICE pointers
 *    owned by the caller; the library keeps no reference past the call (plans and
 *    sessions keep the descriptors they were created with until destroyed).
 *  - hipStream_t arguments are passed as void*; NULL is the default stream.
 *  - Return value: 0 on success, < 0 on error; mx_last_error() describes the last
 *    error of the calling thread.  Entry points are not reentrant on one plan/session.
 *
 * The operator: A = Pl e_full Pr (solvers.py:118-125) with p factored per axis
 * (p_x = sqrt(dxe_x[i]) sqrt(dxh_y[j]) sqrt(dxh_z[k]), ...), equal to the reference's
 * sqrt(p^2) (fdfd/operators.py:116-120) up to a per-cell sign that cancels in E.
 * The per-axis coefficient table holds, for each axis a and index n,
 *   fa = 1/(sqrt(dxe_a[n]) sqrt(dxh_a[(n+1) mod N_a])),  fb = 1/(sqrt(dxe_a[n]) sqrt(dxh_a[n])),
 *   se = sqrt(dxe_a[n]),  sh = sqrt(dxh_a[n])
 * as 4 consecutive complex values, x axis entries first, then y, then z.
 */
#ifndef MEANAS_FDFD_H
#define MEANAS_FDFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MX_ABI_VERSION 1

/* precision / mode */
#define MX_MIXED 0   /* complex64 COCG + FP64 residual replacement, FP64 solution */
#define MX_FP64 1    /* complex128 COCG throughout */

/* operator flags (mx_operator.flags) */
#define MX_EPS_COMPLEX 1
#define MX_EPS_ANISO 2   /* eps holds 3N values (per component); else N (isotropic) */
#define MX_MU 4          /* inv_mu present (1/mu); else mu = 1 */
#define MX_MU_COMPLEX 8
#define MX_MU_ANISO 16
#define MX_MASK 32       /* mask: one byte per cell, bits 0-2 PEC x,y,z, bits 3-5 PMC x,y,z */

/* solve status (mx_stats.status) */
#define MX_CONVERGED 0
#define MX_MAXITER 1
#define MX_BREAKDOWN 2   /* repeated COCG breakdown, stopped at the iteration limit */
#define MX_NONFINITE 3
#define MX_STALLED 4     /* FP64 residual replacements stopped reducing the true residual */

/* One operator (one structure).  Replaces the sparse matrix that
 * meanas/fdfd/operators.py:40-96 (e_full) and :99-123 (e_full_preconditioners)
 * assemble; consumed matrix-free.  Device pointers; the *32 fields feed the complex64
 * kernels and the *64 fields the complex128 kernels. */
typedef struct mx_operator {
  int32_t nx, ny, nz;
  int32_t flags;
  double omega2_re, omega2_im;  /* omega**2 */
  const void* coef32;           /* (nx+ny+nz) x 4 complex64  (fa, fb, se, sh) */
  const void* coef64;           /* (nx+ny+nz) x 4 complex128 */
  const void* eps32;            /* float32 or complex64, N or 3N */
  const void* eps64;            /* float64 or complex128 */
  const void* inv_mu32;         /* NULL unless MX_MU */
  const void* inv_mu64;
  const uint8_t* mask;          /* NULL unless MX_MASK */
} mx_operator;

/* One linear system A x = b (one right-hand side of one operator). */
typedef struct mx_system {
  int32_t op;        /* index into the operator array of the call */
  int32_t maxiter;   /* inner iteration limit (reference default: 10 * 3N, scipy qmr) */
  int32_t slab;      /* x-planes per apply block for this system (0 = mx_options.slab / automatic) */
  int32_t tile_rows; /* apply tile height for this system: 2, 4 or 8 (0 = mx_options.tile_rows) */
  double rtol, atol; /* stop when ||b - A x|| <= max(rtol ||b||, atol) (scaled system) */
  const void* b;     /* complex128 3N, scaled right-hand side */
  void* x;           /* complex128 3N: in = initial guess (scaled), out = solution (scaled) */
  void* work;        /* mx_work_bytes(op, mode) bytes of device memory */
} mx_system;

typedef struct mx_options {
  int32_t mode;        /* MX_MIXED or MX_FP64 */
  int32_t tile_rows;   /* apply tile height: 2, 4 or 8 (0 = automatic) */
  int32_t slab;        /* x-planes per apply block (0 = automatic) */
  int32_t chunk;       /* iterations per captured graph (0 = automatic) */
  int32_t check_every; /* iterations between residual-replacement checks (0 = automatic) */
  int32_t use_graph;   /* 1: replay hipGraph-captured iteration chunks, 0: plain launches */
  double delta;        /* replacement trigger: recursive ||r|| < delta * max ||r|| (0 = automatic) */
  int32_t stall_iters; /* stall window: the solve stops with MX_STALLED when its best true residual
                          has not halved within this many iterations (0 = automatic: 2000 for
                          MX_MIXED, never for MX_FP64; < 0 = never).  An FP64 residual replacement
                          is forced every min(stall_iters, 2000) iterations without one. */
  int32_t stall_limit; /* stop with MX_STALLED after this many consecutive replacements at which the
                          recursive residual had drifted >= 100x below the true one (0 = automatic: 3;
                          < 0 = never) */
  int32_t form;        /* iteration form: MX_FORM_AUTO (fused for MX_FP64, classic for MX_MIXED),
                          MX_FORM_FUSED (MX_FP64 only) or MX_FORM_CLASSIC */
  int32_t reserved;    /* 0 */
} mx_options;

/* COCG iteration forms.  Fused: the direction update rides in the x/r pass with beta
 * predicted from q^T q (9 vector streams per iteration); a system whose recursion loses
 * track of its true residual falls back to the classic form by itself.  Classic: beta
 * from the exact r^T r, the direction updated in its own pass (11 streams). */
#define MX_FORM_AUTO 0
#define MX_FORM_FUSED 1
#define MX_FORM_CLASSIC 2

typedef struct mx_stats {
  int32_t status;        /* MX_CONVERGED, ... */
  int32_t iterations;    /* inner COCG iterations */
  int32_t replacements;  /* FP64 residual replacements (incl. the initial residual) */
  int32_t flags;         /* raw device state flags (MX_FLAG_SWITCHED below) */
  double bnorm;          /* ||b|| */
  double rel_residual;   /* ||b - A x|| / ||b|| in FP64 at exit (scaled system) */
} mx_stats;
/* mx_stats.flags bit of a fused-form session system that finished because its fused
 * recursion lost track (true and recursive residuals parted): continue it from its x in
 * a classic-form session (mx_solve does this itself). */
#define MX_FLAG_SWITCHED 8192

int mx_abi_version(void);
const char* mx_last_error(void);

/* Device bytes of per-system workspace (Krylov vectors, partial sums, state). */
size_t mx_work_bytes(const mx_operator* op, int32_t mode);

/* y = A x for one operator; x, y complex64 (mode MX_MIXED) or complex128 (MX_FP64).
 * Replaces the CSR product A @ x that scipy performs inside qmr
 * (scipy sparse/linalg/_isolve/iterative.py:996, called from fdfd/solvers.py:62). */
int mx_apply(const mx_operator* op, int32_t mode, const void* x, void* y, void* stream);

/* Reusable apply plan (tables built once; no device allocation per apply). */
typedef struct mx_plan mx_plan;
mx_plan* mx_plan_create(const mx_operator* op, int32_t mode, int32_t tile_rows, int32_t slab);
int mx_plan_apply(mx_plan* plan, const void* x, void* y, void* stream);
int mx_plan_blocks(const mx_plan* plan);
void mx_plan_destroy(mx_plan* plan);

/* r = b - A x in FP64 (complex128 in and out); *norm2_out (host) = ||r||^2.
 * The residual the reference's logging callback prints (fdfd/solvers.py:46-48). */
int mx_residual(const mx_operator* op, const void* b, const void* x, void* r, double* norm2_out, void* stream);

/* mx_residual without any device allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and bits as mx_residual; its workspace is a fixed
 * device array of 192 KiB per GPU (enough for every grid mx_residual accepts; a larger
 * need is refused with an error), held until the process ends.  Every call on one GPU
 * uses that GPU's single array, so calls that overlap on one GPU (other host threads or
 * streams) race; the call returns after its stream has finished with the array. */
int mx_residual_noalloc(const mx_operator* op, const void* b, const void* x, void* r, double* norm2_out,
                        void* stream);

/* Diagonal maps of generic() (fdfd/solvers.py:115,120-139), complex128 in/out:
 *   MX_MAP_RHS      out = p * (c * in)              b = Pl (-i w J): c = -i w
 *   MX_MAP_RHS_ADJ  out = conj(c * in) / p          conj of Pr^H (-i w J) (adjoint)
 *   MX_MAP_E        out = in / p                    E = Pr x
 *   MX_MAP_E_ADJ    out = conj(p * in)              E = Pl^H z, z = conj(x) (adjoint)
 *   MX_MAP_SIGN     out = s * (c * in),  s = +-1    the factored p times s is the reference's
 *                                                  principal root sqrt(p^2) (operators.py:120);
 *                                                  maps a scaled vector between the two conventions
 * With c = 1, MX_MAP_RHS / MX_MAP_RHS_ADJ also scale an initial guess E_guess. */
#define MX_MAP_RHS 0
#define MX_MAP_RHS_ADJ 1
#define MX_MAP_E 2
#define MX_MAP_E_ADJ 3
#define MX_MAP_SIGN 4
int mx_map(const mx_operator* op, int32_t map, double c_re, double c_im, const void* in, void* out, void* stream);

/* Solve a batch of systems (one or many right-hand sides, one or many operators of
 * uneven size) in one set of launches per iteration.  Replaces the qmr call of
 * fdfd/solvers.py:62 (scipy sparse/linalg/_isolve/iterative.py:844-1045).
 * stats: nsys entries (host memory).
 * Each host thread keeps the plan of its last successful call (device tables, block
 * maps, stream, pinned buffers, captured iteration graph; a few KB plus the graph) and
 * reuses it when the next call passes the same operators, systems (rtol, atol and
 * maxiter aside) and options: no allocation, upload or capture then.  Environment
 * MEANAS_AMD_PLAN_CACHE=0 builds every plan afresh.
 * Several host threads may call it at once.  No entry point issues work on the legacy
 * (null) stream; a caller's own work on it while another thread's mx_solve captures its
 * iteration graph is refused by HIP and invalidates that capture, so threads that solve
 * at once queue their work on streams of their own. */
int mx_solve(const mx_operator* ops, int32_t nops, const mx_system* systems, int32_t nsys,
             const mx_options* opt, mx_stats* stats, void* stream);
/* Number of mx_solve calls of the calling thread that reused a kept plan (diagnostics;
 * no counterpart in the reference). */
int64_t mx_solve_plan_reuses(void);
/* Free the calling thread's kept plan (its device tables, stream, pinned buffers and
 * graph); no counterpart in the reference.  A thread that ends keeps its plan's memory
 * until the process ends unless it calls this first.  Freeing device memory may
 * synchronise the device: call it when no other thread is inside mx_solve. */
int mx_solve_plan_release(void);

/* ---- fields after the solve ------------------------------------------------------
 * Grid of a post-processing call: the 1D cell widths dxes[0] (E grid) and dxes[1]
 * (H grid) as device arrays of complex128, lengths nx, ny, nz (fdmath/types.py:62-71). */
typedef struct mx_grid {
  int32_t nx, ny, nz, pad;
  const void* dxe[3];
  const void* dxh[3];
} mx_grid;

/* H from E: H = curl_forward(E) / (-i omega), times 1/mu, zero where pmc is set.
 * Replaces fdfd/operators.py:253-282 (e2h, sparse) and fdfd/functional.py:89-117
 * (e2h, matrix-free).  inv_mu: 3N complex128 (1/mu per component) or NULL; pmc: 3N bytes
 * (nonzero = PMC) or NULL; e, h: 3N complex128 in vec() order. */
int mx_e2h(const mx_grid* g, double omega_re, double omega_im, const void* inv_mu, const uint8_t* pmc,
           const void* e, void* h, void* stream);

/* J from a magnetic current M: J = curl_back(M / mu) / (-i omega), backward differences on
 * dxes[1].  Replaces fdfd/functional.py:120-151 (m2j).  The sparse form
 * (fdfd/operators.py:285-307) divides by +i omega instead, so it equals -J.  mu: 3N
 * complex128 or NULL (mu = 1); m, j: 3N complex128 in vec() order. */
int mx_m2j(const mx_grid* g, double omega_re, double omega_im, const void* mu, const void* m, void* j,
           void* stream);

/* The Poynting cross product with the E field shifted one cell forward, both fields
 * weighted by their cell widths: s = E x H as fdfd/functional.py:187-221 computes it
 * (poynting_e_cross_h), equal to poynting_e_cross(e, dxes) @ h of fdfd/operators.py:310-336.
 * The time-averaged Poynting vector is Re(s)/2 with h replaced by conj(H); conj_h = 1 does
 * that conjugation in the kernel.  e, h, s: 3N complex128. */
int mx_poynting_e_cross_h(const mx_grid* g, const void* e, const void* h, int32_t conj_h, void* s, void* stream);

/* poynting_h_cross(h, dxes) @ e (fdfd/operators.py:339-363, which builds the sparse
 * matrix): s = -(E(+1) x H) with the same width weights as mx_poynting_e_cross_h, e
 * conjugated first when conj_e is nonzero.  h, e, s: 3N complex128 in vec() order. */
int mx_poynting_h_cross_e(const mx_grid* g, const void* h, const void* e, int32_t conj_e, void* s, void* stream);

/* ---- time domain (FDTD) ------------------------------------------------------------
 * Real fields (3, nx, ny, nz) in vec() order, float32 (MX_REAL32) or float64 (MX_REAL64).
 * Grid: real cell widths dxes[0] (E grid) and dxes[1] (H grid), float64 device arrays of
 * lengths nx, ny, nz; NULL means unit widths (the reference's dxes=None). */
#define MX_REAL32 0
#define MX_REAL64 1
typedef struct mx_fdtd_grid {
  int32_t nx, ny, nz, pad;
  const double* dxe[3];
  const double* dxh[3];
} mx_fdtd_grid;

/* h -= dt curl_forward(e) / mu (fdtd/base.py:67-125, maxwell_h).  mu: 3N values in the
 * field precision, or NULL to divide by mu_scalar. */
int mx_fdtd_h(const mx_fdtd_grid* g, int32_t prec, double dt, const void* e, void* h, const void* mu,
              double mu_scalar, void* stream);
/* e += dt curl_back(h) / epsilon (fdtd/base.py:13-64, maxwell_e).  eps: 3N values or NULL
 * to divide by eps_scalar. */
int mx_fdtd_e(const mx_fdtd_grid* g, int32_t prec, double dt, void* e, const void* h, const void* eps,
              double eps_scalar, void* stream);
/* `steps` leapfrog steps: maxwell_h then maxwell_e, as the reference's update loop
 * (meanas/test/test_fdtd.py:215-233) runs them, without sources.  Each step is one fused
 * pass that gives the same bits as mx_fdtd_h followed by mx_fdtd_e; the passes alternate
 * between (e, h) and a second pair of fields that the call allocates and frees in stream
 * order (hipMallocAsync, 2 x 3N values), and the result ends in (e, h). */
int mx_fdtd_run(const mx_fdtd_grid* g, int32_t prec, double dt, void* e, void* h, const void* eps,
                double eps_scalar, const void* mu, double mu_scalar, int32_t steps, void* stream);

/* Convolutional PML (fdtd/pml.py).  One layer covers cells [0, thickness) (polarity -1) or
 * [n - thickness, n) (polarity +1) along `axis`.  p_e / p_h: device float64 arrays of
 * 3 x thickness values, the profiles p0, p1, p2 at the E and H sample points, indexed
 * by the cell's position in the layer (pml.py:23-93, cpml_params).  psi_e / psi_h: device
 * arrays of 2 x (cells of the layer) auxiliary values (psi_u then psi_v, each in the
 * layer's C-order sub-grid), float32 or float64 (psi_prec), zeroed by the caller before
 * the first step and kept between steps. */
typedef struct mx_cpml_layer {
  int32_t axis, polarity, thickness, reserved;
  const double* p_e;
  const double* p_h;
  void* psi_e;
  void* psi_h;
} mx_cpml_layer;
#define MX_CPML_MAX_LAYERS 6
/* update_E of updates_with_cpml (fdtd/pml.py:139-183): e_c += dt / eps_c (curl_c + pE_c),
 * the derivatives along each layer's axis scaled by p2 inside it, psi advanced with p0
 * and p1.  Every difference uses g->dxh (pml.py:106-107; g->dxe is not read).  Layers in
 * increasing (axis, polarity) order, at most one per face.  eps: 3N values or NULL for
 * eps_scalar. */
int mx_fdtd_cpml_e(const mx_fdtd_grid* g, int32_t prec, int32_t psi_prec, double dt, void* e, const void* h,
                   const void* eps, double eps_scalar, const mx_cpml_layer* layers, int32_t nlayers, void* stream);
/* update_H of updates_with_cpml (fdtd/pml.py:185-229): h_c -= dt / mu_c (curl_c + pH_c).
 * mu: 3N values or NULL for mu_scalar (the reference's default mu = (1, 1, 1)). */
int mx_fdtd_cpml_h(const mx_fdtd_grid* g, int32_t prec, int32_t psi_prec, double dt, const void* e, void* h,
                   const void* mu, double mu_scalar, const mx_cpml_layer* layers, int32_t nlayers, void* stream);

/* Conducting boundary of one face (fdtd/boundaries.py:12-78, conducting_boundary): the
 * in-place update that its en/hn (polarity -1) or ep/hp (polarity +1) make to e
 * (field = MX_FIELD_E) or h (MX_FIELD_H), with the evident tuple indexing (the
 * reference indexes with a list of slices, which numpy >= 1.23 rejects).  Polarity -1
 * rewrites plane 0 from plane 1; +1 rewrites planes n-1 and n-2 from n-2 and n-3, so the
 * axis needs at least 2 or 3 cells.  g's widths are not read. */
#define MX_FIELD_E 0
#define MX_FIELD_H 1
int mx_fdtd_conducting_boundary(const mx_fdtd_grid* g, int32_t prec, int32_t direction, int32_t polarity,
                                int32_t field, void* f, void* stream);

/* Cell energies of fdtd/energy.py:128-295, in float64 (the reference's float64 widths
 * promote): out = sum_c(ee_c eps_c) dV_e + sum_c(hh_c mu_c) dV_h with dV the products of
 * g's widths (dxe for E, dxh for H; unit when NULL) and ee, hh formed from the fields
 * f[0..3] (each 3N values in the field precision) by `form`:
 *   MX_ENERGY_DXMUL (dxmul)             f = ee, hh
 *   MX_ENERGY_HSTEP (energy_hstep)      f = e0, h1, e2:     ee = e0 e2, hh = h1 h1
 *   MX_ENERGY_ESTEP (energy_estep)      f = h0, e1, h2:     ee = e1 e1, hh = h0 h2
 *   MX_ENERGY_H2E (delta_energy_h2e)    f = e0, h1, e2, h3: ee = e2 (e2 - e0) / dt, hh = h1 (h3 - h1) / dt
 *   MX_ENERGY_E2H (delta_energy_e2h)    f = h0, e1, h2, e3: ee = e1 (e3 - e1) / dt, hh = h2 (h2 - h0) / dt
 *   MX_ENERGY_J (delta_energy_j)        f = j0, e1:         ee = j0 e1, no H term, eps unused
 * eps / mu: 3N values in the field precision, or NULL for eps_scalar / mu_scalar. */
#define MX_ENERGY_DXMUL 0
#define MX_ENERGY_HSTEP 1
#define MX_ENERGY_ESTEP 2
#define MX_ENERGY_H2E 3
#define MX_ENERGY_E2H 4
#define MX_ENERGY_J 5
int mx_fdtd_energy(const mx_fdtd_grid* g, int32_t prec, int32_t form, double dt, const void* const* f,
                   const void* eps, double eps_scalar, const void* mu, double mu_scalar, double* out, void* stream);
/* Poynting vector of fdtd/energy.py:10-87 (poynting): s_x = ey(i+1) hz - ez(i+1) hy and
 * cyclically, E and H multiplied by their widths, E taken at the next cell (periodic);
 * computed in double and stored in the field precision (3N values). */
int mx_fdtd_poynting(const mx_fdtd_grid* g, int32_t prec, const void* e, const void* h, void* s, void* stream);
/* Divergence of s (fdtd/energy.py:90-125, poynting_divergence): N values
 * sum_a (s_a - s_a at the previous cell along a), unit spacing as the reference's
 * deriv_back() without widths; g's widths are not read. */
int mx_fdtd_poynting_divergence(const mx_fdtd_grid* g, int32_t prec, const void* s, void* ds, void* stream);

/* Kernel timing of solver iterations (measurement; no counterpart in the reference).
 * Sets up `sys` as mx_solve would (r = b - A x, p = r) with every stopping rule off, runs
 * `reps` iterations on `stream` exactly as a solve launches them, with a HIP event
 * between consecutive kernels, and reports the average microseconds per launch:
 * us_out[0] operator apply (COCG epilogue), [1] its finishing kernel, [2] the vector
 * update, [3] its finishing kernel, [4] the direction update of the classic form
 * (a no-op launch for MX_FP64 systems in the fused form).  Fills min(n_out, 5) entries.
 * x and the workspace are overwritten. */
int mx_time_kernels(const mx_operator* op, const mx_system* sys, const mx_options* opt, int32_t reps,
                    double* us_out, int32_t n_out, void* stream);

/* Continuous-batching session: systems are admitted into slots while others iterate;
 * no counterpart in the reference (its solves are one at a time).  A session owns its
 * device tables; vectors stay caller-owned. */
typedef struct mx_session mx_session;
mx_session* mx_session_create(int32_t capacity, const mx_options* opt);
/* admit one system (operator copied into the session); returns the slot or < 0 */
int mx_session_admit(mx_session* s, const mx_operator* op, const mx_system* sys, void* stream);
/* run `iterations` COCG iterations on every active slot.  Slots run in launch groups
 * (by tile height, material kind and Infinity Cache budget), one group after the other;
 * a group of smaller systems repeats its `iterations` up to 16 times, so each group gets
 * about the same time per step (a slot's results do not depend on it).  Fused-form
 * sessions launch as mx_solve's fused batches do (apply, one finishing kernel, update
 * per iteration): a slot that loses track finishes with MX_FLAG_SWITCHED instead of
 * switching form in place. */
int mx_session_step(mx_session* s, int32_t iterations, void* stream);
/* host copy of every slot's stats; done[i] = 1 when slot i finished (free to retire) */
int mx_session_poll(mx_session* s, mx_stats* stats, int32_t* done, void* stream);
int mx_session_retire(mx_session* s, int32_t slot);
void mx_session_destroy(mx_session* s);

#ifdef __cplusplus
}
#endif
#endif /* MEANAS_FDFD_H */
