/*
 * gp_simp.h -- C ABI of the MI355X SIMP inner-loop library (libgpsimp.so).
 *
 * Plain C: every argument is a POD, a pointer or a size. Device pointers are
 * caller-owned HIP allocations (e.g. torch tensor data_ptr()), contiguous, of
 * the dtype named in the comment. `stream` is a hipStream_t passed as void*
 * (NULL = the null stream). Every function returns GP_OK (0) or an error code;
 * gp_last_error() returns a thread-local message for the last failure.
 *
 * Numbering conventions (identical to genpipeline/topology/simp_solver.py):
 *   element e = i*ny*nz + j*nz + k                      (simp_solver.py:276)
 *   node    n = i*(ny+1)*(nz+1) + j*(nz+1) + k          (simp_solver.py:277,96)
 *   dof       = 3*n + c, c in {x,y,z}                   (simp_solver.py:106)
 *   element corner order (C3D8): (0,0,0) (1,0,0) (1,1,0) (0,1,0)
 *                                (0,0,1) (1,0,1) (1,1,1) (0,1,1)  (:255-267)
 */
#ifndef GP_SIMP_H
#define GP_SIMP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define GP_OK 0
#define GP_ERR_ARG 1          /* bad argument / shape                       */
#define GP_ERR_OOM 2          /* device allocation failed                   */
#define GP_ERR_HIP 3          /* HIP runtime error                          */
#define GP_BREAKDOWN 4        /* PCG: p^T A p <= 0 (indefinite operator)    */
#define GP_NOT_CONVERGED 5    /* PCG: max_iter reached                      */
#define GP_ERR_NAN 6          /* NaN detected in a result                   */
#define GP_ERR_STATE 7        /* call out of order (e.g. no BCs set)        */

/* ---- enums ----------------------------------------------------------- */
#define GP_F32 0              /* PCG vectors float32, FP64 dots              */
#define GP_F64 1              /* PCG vectors float64                          */
#define GP_F16 2              /* the PCG iterate x and search direction p    */
                              /* stored fp16; the recurrence vectors r, w, s, */
                              /* M^-1 and the eight element coefficients stay */
                              /* float; fp32 math, FP64 dots; runs only as    */
                              /* the inner solve of FP64 iterative refinement */
#define GP_BF16 3             /* as GP_F16 with bfloat16 storage (the fp32   */
                              /* exponent range: safe for Emin = 1e-9)        */

#define GP_LAW_CLAMP 0        /* E = max(x,1e-3)^p (simp_solver.py:201-203) */
#define GP_LAW_EMIN 1         /* E = Emin + x^p (E0-Emin)                   */

#define GP_OC_CPU 0           /* l2=1e9, <=50 steps, tol 1e-4, floor 1e-12  */
                              /* (simp_solver.py:225-249)                    */
#define GP_OC_GPU 1           /* l2=2max(max(-dc),1), <=30 steps, tol 1e-5,  */
                              /* floor 1e-15 (simp_solver_gpu.py:801-841)    */

#define GP_KE_HEX8 0          /* exact 2x2x2-Gauss trilinear hex (default)   */
#define GP_KE_CUSTOM 1        /* caller-supplied 24x24 table (gp_spec.ke)    */

#define GP_FILTER_F64 0       /* filter weights rmin - |d| in float64        */
                              /* (SIMPSolver._build_filter, simp_solver.py:  */
                              /* 1545-1567)                                  */
#define GP_FILTER_F32 1       /* the same weights rounded once to float32,   */
                              /* as SIMPSolverGPU._build_filter stores them  */
                              /* whatever its dtype (simp_solver_gpu.py:915) */

/* ---- problem specification ------------------------------------------ */
typedef struct gp_spec {
  int32_t nx, ny, nz;         /* elements per axis                           */
  int32_t law;                /* GP_LAW_*                                    */
  int32_t oc_variant;         /* GP_OC_*                                     */
  int32_t ke_kind;            /* GP_KE_*                                     */
  double E0, Emin, nu, penal, rmin, move;
  const double* ke;           /* host 24x24 row-major if GP_KE_CUSTOM        */
  int32_t filter_weights;     /* GP_FILTER_*                                 */
} gp_spec;

typedef struct gp_problem gp_problem;   /* opaque */

typedef struct gp_pcg_opts {
  int32_t precision;          /* GP_F32 / GP_F64 / GP_F16 / GP_BF16          */
  int32_t max_iter;           /* inner PCG iteration cap                     */
  double tol;                 /* ||r|| <= tol ||f_free||  (outer when refine) */
  int32_t use_x0;             /* 1: u holds the initial guess               */
  int32_t refine;             /* 1: FP64 iterative refinement around the     */
                              /*    low-precision PCG (u must be float64)    */
  double inner_tol;           /* inner PCG tol when refine=1                 */
  int32_t max_outer;          /* refinement sweeps cap                       */
} gp_pcg_opts;

typedef struct gp_pcg_stats {
  int32_t status;             /* GP_OK / GP_BREAKDOWN / GP_NOT_CONVERGED     */
  int32_t iters;              /* total inner PCG iterations                  */
  int32_t outer;              /* refinement sweeps (0 without refine)        */
  double relres;              /* final ||r||/||f|| (FP64 true residual when  */
                              /* refine=1, recursive residual otherwise)     */
} gp_pcg_stats;

typedef struct gp_oc_stats {
  double lmid;                /* final Lagrange multiplier                   */
  int32_t steps;              /* bisection steps taken                       */
} gp_oc_stats;

typedef struct gp_simp_opts {
  int32_t n_iters;            /* OC iterations                               */
  double volfrac;
  gp_pcg_opts pcg;
  int32_t tol_schedule;       /* 1: SIMPSolverGPU tolerance schedule         */
                              /*    1e-4/1e-5/1e-6 (simp_solver_gpu.py:97-103);*/
                              /*    pcg.max_iter stays the iteration cap     */
  int32_t early_stop;         /* 1: stop after 3 consecutive dC<1%           */
                              /*    (simp_solver_gpu.py:118-125)             */
  int32_t compliance_each_iter; /* 1: compliance solve every OC iteration     */
  int32_t warm_start;         /* 1: PCG starts from the previous u           */
  int32_t warm_compliance;    /* 1: with compliance_each_iter and warm_start, */
                              /*    the compliance solve also starts from the */
                              /*    previous u (the reference solves it from  */
                              /*    zero, simp_solver_gpu.py:747-751)         */
  int32_t direct;             /* 1: every solve is gp_direct's exact FP64     */
                              /*    banded LU (the reference CPU solver's     */
                              /*    SuperLU, simp_solver.py:213); pcg ignored */
  int32_t recycle;            /* 0: a warm-started solve starts from the last */
                              /*    u (simp_solver_gpu.py:593,704). 1..8: it  */
                              /*    starts from the Galerkin combination of   */
                              /*    the last `recycle` solutions of the run   */
                              /*    (gp_recycle_guess, float copies); any PCG */
                              /*    mode, not with direct                     */
} gp_simp_opts;

typedef struct gp_simp_stats {
  int32_t iters_done;         /* OC iterations executed                      */
  int32_t pcg_iters_total;
  int32_t bisect_steps_total;
  int32_t status;             /* worst PCG status seen                       */
  double compliance;          /* final f.u (reference: last_compliance)      */
  double seconds;             /* wall time of the call                       */
} gp_simp_stats;

/* ---- library --------------------------------------------------------- */
const char* gp_last_error(void);
int gp_version(void);                                 /* 100*major + minor */

/* Device allocator of the library's workspaces (problem, batch, scheduler
 * and CSR-solver buffers). The reference's kernels allocate through
 * PyTorch (torch::zeros_like, simp_sensitivity_ptx.cu:405,461); this routes
 * the build's allocations through the caller's allocator in the same way,
 * e.g. a framework's caching allocator. alloc(bytes, device, ctx) returns
 * device memory on `device` (the current HIP device) or NULL; release(ptr,
 * device, ctx) takes it back. The library waits for the device's queued work
 * before each release, as hipFree does. NULL, NULL restores hipMalloc and
 * hipFree. Returns GP_ERR_STATE while any allocation made through the
 * current pair is alive (set it before creating handles, and not while
 * another thread may be creating one).                                   */
typedef void* (*gp_alloc_fn)(size_t bytes, int32_t device, void* ctx);
typedef void (*gp_free_fn)(void* ptr, int32_t device, void* ctx);
int gp_set_allocator(gp_alloc_fn alloc, gp_free_fn release, void* ctx);

/* Exact unit-cube hex8 element stiffness (E = 1) for Poisson ratio nu into
 * ke_out (host, 24x24 row-major). Replaces SIMPSolver._get_Ke
 * (simp_solver.py:284-1543), whose literal table is not a valid hex8 KE.  */
int gp_hex8_ke(double nu, double* ke_out);

/* Device bytes a problem of this spec allocates (admission by footprint). */
int gp_footprint_bytes(const gp_spec* spec, int32_t precision, size_t* bytes);

/* Device bytes a problem of this spec adds for recycled starts over
 * `recycle` stored solutions (0..8; 0: none). No counterpart in the
 * reference, which keeps no solution history; for admission by footprint. */
int gp_recycle_bytes(const gp_spec* spec, int32_t recycle, size_t* bytes);

/* ---- problem lifecycle ----------------------------------------------- *
 * Replaces the per-solver set-up of SIMPSolver.__init__ / SIMPSolverGPU.__init__
 * (simp_solver.py:24-54, simp_solver_gpu.py:16-69): KE, edof (derived
 * arithmetically in-kernel, no table), filter stencil and Hs.              */
int gp_problem_create(const gp_spec* spec, gp_problem** out);
int gp_problem_destroy(gp_problem* p);

/* Fixed DOFs (host int64 list, any order, duplicates allowed) and load
 * vector f (host float64, n_dof). Replaces _get_bcs (simp_solver.py:81-157,
 * simp_solver_gpu.py:138-248); the caller computes the lists.             */
int gp_problem_set_bcs(gp_problem* p, const int64_t* fixed_dofs, int64_t n_fixed,
                       const double* f_host);
/* Preserved (non-design) elements: host uint8 per element or NULL.         */
int gp_problem_set_preserved(gp_problem* p, const uint8_t* mask_host);
/* Sampled kernel timing: with every > 0, HIP events bracket every every-th
 * chunk of back-to-back PCG iteration launches on the caller's stream
 * (0 = off). get_profile returns the summed milliseconds and the number of
 * launches in the chunks whose launches all executed (chunks holding early
 * exits after convergence are dropped).
 * No counterpart in the reference; measurement hook for benchmarks.        */
int gp_problem_set_profiling(gp_problem* p, int32_t every);
int gp_problem_get_profile(gp_problem* p, double* ku_ms_sum, int64_t* ku_samples,
                           int32_t reset);
/* Workgroup shape of the modal K.u sweep for one precision (GP_F32/GP_F64):
 * tj x tk element columns (one of 8x32 16x16 14x18 6x34 14x34 4x64 10x26 7x34),
 * ni node planes per workgroup; tj = 0 selects the dense gather kernel.
 * Used by the autotuner. No counterpart in the reference.                */
int gp_problem_set_ku_tile(gp_problem* p, int32_t precision, int32_t tj, int32_t tk,
                           int32_t ni);
/* Current shape (tj = 0: gather kernel) and whether KE has the isotropic
 * hex8 modal structure the sweep kernel needs (modal = 1).                 */
int gp_problem_get_ku_tile(gp_problem* p, int32_t precision, int32_t* tj, int32_t* tk,
                           int32_t* ni, int32_t* modal);
/* Copy of the device load vector into a float64 device buffer (n_dof).      */
int gp_problem_load(gp_problem* p, double* f_dev, void* stream);

/* ---- operator pieces (stream-ordered, no host sync) ------------------ */
/* E[e] = law(xPhys[e]) as float32. xPhys float64 (n_elem).                   */
int gp_stiffness(gp_problem* p, const double* xPhys, float* E, void* stream);
/* y = K(E) u with fixed DOFs of y zeroed when mask != 0. u,y dtype =
 * GP_F32 (float) or GP_F64 (double). Replaces the assembled K SpMV
 * (simp_solver.py:201-213, simp_solver_gpu.py:313-326,341) and fused_spmv
 * (cuda_kernels/__init__.py:329-368) on the structured grid.              */
int gp_ku(gp_problem* p, const float* E, const void* u, void* y, int32_t dtype,
          int32_t mask, void* stream);
/* diag(K(E)) (dtype as gp_ku), fixed DOFs report their raw diagonal.        */
int gp_diag(gp_problem* p, const float* E, void* diag, int32_t dtype, void* stream);

/* ---- solve ------------------------------------------------------------ *
 * Jacobi-PCG on the free DOFs, u_fixed = 0. Replaces SIMPSolverGPU._solve /
 * _pcg / _pcg_ic (simp_solver_gpu.py:328-475,571-705) and the SuperLU solve
 * of SIMPSolver (simp_solver.py:213). f float64 device (n_dof). u dtype:
 * float64 when opts->refine or precision==GP_F64, else float32. Blocks the
 * host until the solve finishes (the iteration count is a host value).     */
int gp_pcg(gp_problem* p, const float* E, const double* f, void* u,
           const gp_pcg_opts* opts, gp_pcg_stats* stats, void* stream);

/* Exact solve of K(xPhys) u = f on the free DOFs (u = 0 on the fixed ones)
 * by banded LU with partial pivoting in FP64, the moduli E(xPhys) taken in
 * FP64 as well. Replaces SciPy SuperLU's spsolve in SIMPSolver._solve
 * (simp_solver.py:195-213); like it, it needs no symmetry or definiteness of
 * the element table (the reference's own table is neither). xPhys (n_elem),
 * f, u (n_dof): float64 device. Band of n_free * (6 (sI + sJ + 1) + 6)
 * doubles, sI = (ny+1)(nz+1), sJ = nz+1: grids up to (ny+1)(nz+1) ~ 1300
 * and a band of 16 GiB; larger grids are refused (GP_ERR_ARG). A zero pivot
 * (singular system) returns GP_BREAKDOWN. One launch per free DOF; blocks.
 * stats: status only (iters = outer = 0, relres = 0).                      */
int gp_direct(gp_problem* p, const double* xPhys, const double* f, double* u,
              gp_pcg_stats* stats, void* stream);

/* ---- SIMP pieces ------------------------------------------------------ */
/* dc[e] = -dE/dx(xPhys[e]) * u_e^T KE u_e (float64 out); ce_sum (device
 * double scalar or NULL) receives sum_e E_e u_e^T KE u_e.
 * Replaces SIMPSolver._sensitivity's energy part (simp_solver.py:215-218)
 * and simp_sensitivity (cuda_kernels/__init__.py:286-326).                 */
int gp_sens(gp_problem* p, const double* xPhys, const void* u, int32_t dtype,
            double* dc, double* ce_sum, void* stream);
/* mode 0: out = (H in)/Hs (density filter, simp_solver.py:240)
 * mode 1: out = H^T (in/Hs) (sensitivity chain rule, simp_solver.py:221-223) */
int gp_filter(gp_problem* p, const double* in, double* out, int32_t mode, void* stream);
/* OC bisection update (simp_solver.py:225-249 / simp_solver_gpu.py:801-841).
 * x, dc: float64 in; xnew, xPhys_new: float64 out. Blocks for stats.       */
int gp_oc(gp_problem* p, const double* x, const double* dc, double volfrac,
          double* xnew, double* xPhys_new, gp_oc_stats* stats, void* stream);
/* out_host = f . u  (FP64 sum). Blocks.                                     */
int gp_compliance(gp_problem* p, const void* u, int32_t dtype, double* out_host,
                  void* stream);

/* ---- whole SIMP loop ---------------------------------------------------- *
 * SIMPSolver.run / SIMPSolverGPU.run (simp_solver.py:56-79,
 * simp_solver_gpu.py:71-136). x, xPhys: float64 device (n_elem) in/out;
 * u: float64 device (n_dof) warm-start state in/out.                       */
int gp_simp_run(gp_problem* p, const gp_simp_opts* opts, double* x, double* xPhys,
                double* u, gp_simp_stats* stats, void* stream);

/* Recycled starting guess: u0 = V c with (V^T K V) c = V^T f, K the
 * stiffness at moduli E (float32 device, n_elem) with the supports applied
 * and f the problem's load, over n (1..8) float32 device vectors V[i]
 * (n_dof each, zero on fixed DOFs; V is a host array of device pointers).
 * It minimises the K-norm of the error over span(V); nearly dependent
 * vectors (Cholesky pivots below 1e-12 of the largest diagonal of V^T K V)
 * are left out. u0: float32 device (n_dof). coef: NULL, or host doubles
 * receiving c (8 entries) and the number of vectors kept (after a stream
 * sync). The start gp_simp_run's recycle option uses; no counterpart in the
 * reference, which starts from the previous u (simp_solver_gpu.py:593,704). */
int gp_recycle_guess(gp_problem* p, const float* E, const float* const* V, int32_t n,
                     float* u0, double* coef, void* stream);

/* ---- batches ----------------------------------------------------------- *
 * Many independent problems (any grid sizes, hex8 KE, BCs set) run as one:
 * every phase of the SIMP loop is one launch over all problems, a problem
 * whose PCG has converged idles inside that launch. tj x tk is the modal
 * K.u workgroup shape shared by the batch (each problem keeps its own sweep
 * depth from gp_problem_set_ku_tile). No counterpart in the reference,
 * which solves one problem per call (topo_data_gen.py:235-307 threads them). */
typedef struct gp_batch gp_batch;
int gp_batch_create(gp_problem* const* problems, int32_t n, int32_t tj, int32_t tk,
                    gp_batch** out);
int gp_batch_destroy(gp_batch* b);
/* One SIMP run per problem (opts shared; opts->pcg.precision must be GP_F32,
 * refine 0). volfracs: host array (n). x, xPhys, u: host arrays of n device
 * pointers (float64 n_elem / n_elem / n_dof, in/out as in gp_simp_run).
 * stats: host array (n). Blocks until all problems finish.              */
int gp_batch_simp_run(gp_batch* b, const gp_simp_opts* opts, const double* volfracs,
                      double* const* x, double* const* xPhys, double* const* u,
                      gp_simp_stats* stats, void* stream);

/* ---- device scheduler: independent problems, continuous batching ------ *
 * No counterpart in the reference (it runs one solve per call from Python
 * threads: voxel_fem.py:559-611, topo_data_gen.py:235-307). A scheduler has
 * n slots; each slot runs one job (a SIMP run or a single solve) on one
 * problem. One tick launch advances every occupied slot by one phase of its
 * job (material law, Jacobi set-up, one PCG iteration, sensitivity, filter,
 * one OC bisection step, commit, compliance), so slots progress
 * independently and a finished slot can be released and refilled between
 * ticks while the others keep running. Results are bitwise those of
 * gp_simp_run with the same K.u tile (tj, tk, ni), whatever else runs.   */
typedef struct gp_sched gp_sched;

typedef struct gp_job {
  int32_t n_iters;            /* OC iterations; 0: one solve (compliance)   */
  double volfrac;
  double tol;                 /* PCG relative tolerance (FP32 vectors)      */
  int32_t max_iter;           /* PCG iteration cap per solve                */
  int32_t tol_schedule;       /* as gp_simp_opts                            */
  int32_t early_stop;
  int32_t compliance_each_iter;
  int32_t warm_start;         /* first solve starts from u; later solves    */
                              /* from the previous solution                */
  int32_t refine;             /* 1: FP64 iterative refinement around FP32   */
                              /*    inner solves, as gp_pcg with refine = 1 */
                              /*    (tol is then the outer FP64 residual    */
                              /*    tolerance); u is the FP64 iterate       */
  double inner_tol;           /* inner PCG tolerance (0: 1e-3)              */
  int32_t max_outer;          /* refinement sweeps cap (0: 50)              */
  int32_t warm_compliance;    /* as gp_simp_opts.warm_compliance            */
  int32_t recycle;            /* as gp_simp_opts.recycle (0..8). The ring is */
                              /* the problem's own: a problem runs in one    */
                              /* such job at a time, and not in gp_simp_run  */
                              /* meanwhile                                   */
} gp_job;

typedef struct gp_slot_status {
  int32_t phase;              /* 0 idle, 18 done, else running (1-17,       */
                              /* 19-22 in the refinement sweeps, 23-27 in   */
                              /* recycled starts)                           */
  int32_t done;
  int32_t iters_done;         /* OC iterations completed                    */
  int32_t pcg_iters_total;
  int32_t bisect_steps_total;
  int32_t status;             /* worst PCG status, or GP_ERR_NAN            */
  double compliance;          /* final f.u once done                        */
} gp_slot_status;

/* slot_dims: 4 ints per slot (nx, ny, nz, ni): the largest grid the slot
 * accepts and the modal sweep depth its jobs use. (tj, tk) = (8, 32),
 * (16, 16), (14, 18), (7, 34) or (6, 34). The device is the current one. */
int gp_sched_create(int32_t n_slots, const int32_t* slot_dims, int32_t tj, int32_t tk,
                    gp_sched** out);
int gp_sched_destroy(gp_sched* s);
/* Blocks of one slot's range (slot >= 0) or of the whole tick (slot -1). */
int gp_sched_blocks(gp_sched* s, int32_t slot, int32_t* blocks);
/* Starts a job in a free slot. x, xPhys (float64 n_elem) and u (float64
 * n_dof) are device buffers the job reads and updates in place, as in
 * gp_simp_run; they and the problem must stay alive until the slot is
 * released. Stream-ordered: takes effect at the next tick on `stream`.   */
int gp_sched_admit(gp_sched* s, int32_t slot, gp_problem* p, const gp_job* job, double* x,
                   double* xPhys, double* u, void* stream);
/* Replaces the load of a single-solve job (n_iters = 0) just admitted to
 * `slot` with the device vector f (n_dof float64): the job then solves
 * K d = f from d = 0, and its reported compliance is f.d. No counterpart in
 * the reference, which solves each evaluator field alone
 * (GPUVoxelFEM.solve, genpipeline/fem/gpu_fem_solver.py:76-133): it carries
 * the residual right-hand sides of the batched evaluator's FP64 correction
 * passes. Call it before the next tick; f must stay valid until the slot is
 * released. Stream-ordered.                                               */
int gp_sched_set_rhs(gp_sched* s, int32_t slot, const double* f, void* stream);
/* Frees a slot (its job need not be done). Stream-ordered.               */
int gp_sched_release(gp_sched* s, int32_t slot, void* stream);
/* Captures `ticks` (even) tick launches in a HIP graph that later ticks
 * replay; 0 goes back to one launch per tick.                            */
int gp_sched_set_graph(gp_sched* s, int32_t ticks);
/* Enqueues n ticks on `stream` (asynchronous).                           */
int gp_sched_tick(gp_sched* s, int64_t n_ticks, void* stream);
/* Copies every slot's status (n_slots entries) after the ticks enqueued
 * so far; synchronises `stream`.                                          */
int gp_sched_status(gp_sched* s, gp_slot_status* out, void* stream);
/* Ticks until every occupied slot is done or max_ticks have run; out
 * (n_slots entries, may be NULL) receives the final statuses.            */
int gp_sched_run(gp_sched* s, int64_t max_ticks, int64_t* ticks_run, gp_slot_status* out,
                 void* stream);

/* ---- generic-table kernels (reference kernel API) --------------------- *
 * Explicit KE (24x24 row-major) and edof (int32 n_elem x 24, every entry in
 * [0, n_dof)) as the reference kernels take them; edof may be arbitrary (no
 * grid assumption). dtype GP_F32 or GP_F64 names the type of every float
 * array (the reference kernels take double); element products are FP64.
 * simp_sensitivity: genpipeline/cuda_kernels/__init__.py:286-326,
 *   kernel simp_sensitivity_ptx.cu:163-270 -> dc = -p max(x,1e-3)^(p-1) u_e^T KE u_e
 * fused_spmv:       __init__.py:329-368, kernel simp_sensitivity_ptx.cu:301-373
 *   -> y = sum_e max(x,1e-3)^p KE u_e scattered by edof. The reference scatters
 *   with atomic adds (summation order varies run to run); here every DOF
 *   gathers its terms in a fixed order (element, then local index), so the
 *   result is deterministic. workspace: gp_generic_spmv_workspace ints.     */
int gp_generic_sens(const void* xPhys, const void* u, const void* ke, const int32_t* edof,
                    int64_t n_elem, double penal, int32_t dtype, void* dc, void* stream);
int gp_generic_spmv_workspace(int64_t n_elem, int64_t n_dof, int64_t* n_ints);
int gp_generic_spmv(const void* xPhys, const void* u, const void* ke, const int32_t* edof,
                    int64_t n_elem, int64_t n_dof, double penal, int32_t dtype,
                    int32_t* workspace, void* y, void* stream);
/* Jacobi-PCG on a CSR matrix (float32 values, int64 row_ptr/col_idx), fixed
 * DOFs (int64 list) held at zero. Replaces pcg_solve_cuda
 * (__init__.py:443-473; simp_pcg_cuda.cu:66-124). x: in x0, out solution.  */
int gp_csr_pcg(const int64_t* row_ptr, const int64_t* col_idx, const float* values,
               const float* diagonal, const float* b, float* x, const int64_t* fixed,
               int64_t n_fixed, int64_t n, int32_t max_iter, double tol,
               gp_pcg_stats* stats, void* stream);

/* gp_csr_pcg for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. The same arguments, kernels and results as
 * gp_csr_pcg; its four PCG vectors, scalars and partial sums live in a
 * static device array of the library (16 MiB + 8 KiB per GPU, held until
 * the process ends), so n is at most 1048576 (GP_ERR_ARG beyond). Calls
 * that overlap on one GPU share that array and race; calls in sequence do
 * not (the scalars are cleared by a kernel at each call).                  */
int gp_csr_pcg_noalloc(const int64_t* row_ptr, const int64_t* col_idx, const float* values,
                       const float* diagonal, const float* b, float* x, const int64_t* fixed,
                       int64_t n_fixed, int64_t n, int32_t max_iter, double tol,
                       gp_pcg_stats* stats, void* stream);

/* ---- isosurface ---------------------------------------------------------- *
 * Triangles of the surface density = level of an nx x ny x nz float64 field
 * (C order), the marching-cubes step of density_to_stl
 * (genpipeline/topology/mesh_export.py:7-67, which calls
 * skimage.measure.marching_cubes and writes the faces with numpy-stl).
 * A sample is inside when > level; each 2x2x2 cube's triangles come from a
 * generated 256-case table (inside on the normal's back side); a vertex is
 * the linear level crossing of its grid edge. Record per triangle: 12
 * float32 (unit normal, then 3 vertices in index units times sx, sy, sz),
 * in C order of the cubes.
 * gp_mc_workspace: int32 entries of device workspace the other two need.
 * gp_mc_count: fills the workspace, returns the triangle count (synchronises
 *   the stream); gp_mc_emit: writes n_tri * 12 floats to tris.
 * gp_mc_tables: the generated table (ntri[256], tri[256][15], -1 padded). */
int gp_mc_workspace(int32_t nx, int32_t ny, int32_t nz, int64_t* n_ints);
int gp_mc_count(const double* density, int32_t nx, int32_t ny, int32_t nz, double level,
                int32_t* workspace, int64_t* n_tri, void* stream);
int gp_mc_emit(const double* density, int32_t nx, int32_t ny, int32_t nz, double level,
               double sx, double sy, double sz, const int32_t* workspace, float* tris,
               void* stream);
int gp_mc_tables(int8_t* ntri, int8_t* tri);

/* ---- C3D8 hex mesh of a voxel field ------------------------------------ *
 * The mesh the CalculiX path writes to its .inp file
 * (VoxelHexMesher.voxels_to_inp, genpipeline/fem/voxel_fem.py:120-146, with
 * the solid-voxel scan get_solid_voxels_simd, cuda_kernels/__init__.py:
 * 371-388 and mesher_simd.cpp:11-52): one element per voxel > 0.5 of a
 * D x H x W float32 field (C order), elements in C order of their voxels;
 * grid corners numbered from 1 in the order the elements first touch them,
 * each element's corners in C3D8 order (0,0,0) (1,0,0) (1,1,0) (0,1,0)
 * (0,0,1) (1,0,1) (1,1,1) (0,1,1).
 * gp_hex_mesh_workspace: int32 entries of device workspace;
 * gp_hex_mesh_count: fills it, returns the element and node counts
 *   (synchronises the stream);
 * gp_hex_mesh_emit: elems (n_elem x 8 node numbers), node_ijk (n_node x 3
 *   grid corner indices, in node-number order), elem_voxel (n_elem voxel
 *   C indices, or NULL).                                                    */
int gp_hex_mesh_workspace(int32_t D, int32_t H, int32_t W, int64_t* n_ints);
int gp_hex_mesh_count(const float* voxels, int32_t D, int32_t H, int32_t W, int32_t* workspace,
                      int64_t* n_elem, int64_t* n_node, void* stream);
int gp_hex_mesh_emit(const float* voxels, int32_t D, int32_t H, int32_t W, int32_t* workspace,
                     int32_t* elems, int32_t* node_ijk, int32_t* elem_voxel, void* stream);

/* ---- evaluator pre-processing ------------------------------------------ *
 * Decoded fields -> FEM voxel densities, on the device. Replaces the CPU
 * chain of VoxelFEMEvaluator.evaluate (genpipeline/fem/voxel_fem.py:452-462):
 * scipy.ndimage.zoom(v, 0.5, order=1), then smooth_voxels
 * (genpipeline/pipeline_utils.py:55-76): gaussian_filter(v, sigma) and
 * 1 / (1 + exp(-sharpness (v - 0.5))). in: batch fields of d x h x w float32
 * (C order); out and work: batch * od * oh * ow float32 (od = round(d * zoom)
 * as scipy rounds it). weights: host array of 2 radius + 1 symmetric taps
 * (the Gaussian of gaussian_filter), radius <= 16. Order-1 zoom and the
 * filter follow scipy's operation order and rounding (float32 after the zoom
 * and after each axis); the logistic is float arithmetic as numpy does it on
 * float32.                                                                  */
int gp_voxel_prep(const float* in, int32_t batch, int32_t d, int32_t h, int32_t w, int32_t od,
                  int32_t oh, int32_t ow, const double* weights, int32_t radius,
                  double sharpness, float* work, float* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GP_SIMP_H */
