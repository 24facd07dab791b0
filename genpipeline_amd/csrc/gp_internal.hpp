// Internal declarations shared by the kernel and API translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cstddef>
#include "../../include/gp_simp.h"

namespace gp {

constexpr int kBlock = 256;          // threads per block of the 1-D kernels
constexpr int kNumXcd = 8;           // MI355X: workgroups go round-robin over 8 XCDs
constexpr int kVecBlocks = 512;      // cap on blocks of reducing vector kernels
constexpr int kMaxOffsets = 343;     // filter stencil cap: (2*3+1)^3, rmin <= 3
constexpr int kMaxTaps = 33;         // 1-D voxel filter cap: radius <= 16

// Structured hex8 grid. Node strides: sJ = nz+1, sI = (ny+1)(nz+1).
struct Grid {
  int nx, ny, nz;
  int sJ, sI;          // node strides
  int eJ, eI;          // element strides: nz, ny*nz
  int nn, ne;          // nodes, elements
  long long ndof;
};

// Density-filter stencil: offsets (di,dj,dk) in lexicographic order and
// weights rmin - |d| (only strictly positive weights are kept).
struct Stencil {
  int n;
  int di[kMaxOffsets], dj[kMaxOffsets], dk[kMaxOffsets];
  double w[kMaxOffsets];
};

// Modal coefficients of an isotropic hex8 KE (Khat/64, see gp_ku_modal.hip).
template <typename T>
struct ModalCoef {
  T An, Ln, S, B1, B2, T1, T2, X;
};

// Workgroup shape of the modal K.u sweep: TJ x TK element columns, NI node
// planes per workgroup.
struct KuTile {
  int tj, tk, ni;
};

// FP64 partial sums, one region per reduction kind. Every reduction is
// finished by its consumers: each consuming block sums the producer's
// partials in the same fixed order, so all blocks agree bit for bit and no
// kernel needs device-scope fences or atomics.
// OC bisection: steps evaluated per launch (body_oc_step), candidate
// volumes per launch, and the launches that cover oc_max steps.
constexpr int kOcLevels = 3;
constexpr int kOcCand = (1 << kOcLevels) - 1;
__host__ __device__ constexpr int oc_launches(int oc_max) { return (oc_max - 1) / kOcLevels + 2; }

struct Parts {
  double* A;       // p.Ap per K.u block
  double* RZ[2];   // r.z per vector block, by iteration parity
  double* RR;      // r.r per vector block
  double* BB;      // b.b per vector block
  double* O[2];    // OC candidate volume sums per block (candidate c at c * nb), by launch parity
  double* M;       // OC max(-dc) per block
  double* S;       // one-shot sums (energy, dot, residual)
  double* G[2];    // single-kernel PCG: (r,z), (z,Az), (r,r) per block, by parity
  double* D[2];
  double* R[2];
  int cap;         // entries per region
};

// Scalars of one PCG solve / OC update, resident in device memory. Fields
// are written by block 0 of a kernel and read by later kernels only.
struct DevScalars {
  double bb, tol2, rr, rz, pAp;
  double ocl1[2], ocl2[2];   // bisection interval at the start of a launch's steps, by parity
  double lmid;               // final multiplier
  double volfrac, move, oc_eps, oc_tol, oc_floor;
  double sum;                // one-shot reduction result
  double cgg[2], cga[2];     // single-kernel PCG: gamma and alpha by parity
  double cgb[2];             // and beta (read by the two-step x and p update)
  double ir_rr, ir_bb;       // FP64 refinement: |r|^2 and |f_free|^2 of the last residual
  int iter, max_iter, done, status;
  int oc_steps, oc_max, oc_done, nan_flag;
};

// One problem of a batch as the batched kernels see it (device memory).
struct BatchProb {
  Grid g;
  ModalCoef<float> mc;
  ModalCoef<double> mc64;
  int ni, ntj, ntk;             // modal sweep geometry for the batch tile
  int law, oc_variant;
  double penal, E0, Emin, move, volfrac, n_pres;
  const double* ke64;
  const float* ke32;
  const Stencil* st;
  const double* Hs;
  const double* w;
  const uint8_t* pres;
  const uint8_t* fixmask;
  const double* f;
  double *x, *xPhys, *xnew, *xphys_new, *dc, *dcf;
  float* E;
  float* u;                     // float32 PCG iterate (FP64-refined jobs: the correction)
  double* u64;                  // caller's float64 state
  double* r64;                  // FP64 refinement residual
  float *rb[2], *wb[2], *sb[2], *pc, *minv;
  DevScalars* sc;               // PCG scalars
  DevScalars* sc_oc;            // OC / one-shot scalars
  Parts parts;
  // recycled starts (scheduler jobs with recycle > 0): the problem's ring of
  // stored float solutions, their products K v, Gram partials, coefficients
  float* rc_v;
  float* rc_kv;
  double* rc_part;
  double* rc_coef;
};

// Launch description of a batch: problem table and, per kernel kind, the
// block offsets off[0..n] of each problem inside one launch.
struct BatchLaunch {
  const BatchProb* probs;
  int n;
  int tj, tk;
  const int *off_cg, *off_node, *off_elem, *off_vec, *off_oc;
  int tot_cg, tot_node, tot_elem, tot_vec, tot_oc;
};

struct BatchStatus {
  int done, iter, status, oc_steps;
  double rr, bb, sum, lmid;
  int nan_flag, pad;
};

void batch_stiffness(const BatchLaunch& bl, hipStream_t st);
void batch_pcg_setup(const BatchLaunch& bl, bool use_x0, double tol, int max_iter,
                     hipStream_t st);
void batch_cg(const BatchLaunch& bl, int k, hipStream_t st);
void batch_post_solve(const BatchLaunch& bl, int oc_max_steps, hipStream_t st);
void batch_compliance(const BatchLaunch& bl, hipStream_t st);
void batch_cast(const BatchLaunch& bl, int dir, hipStream_t st);
void batch_status(const BatchLaunch& bl, BatchStatus* out, hipStream_t st);

// ------------------------------------------------------------ scheduler
// A slot holds one problem and a phase machine; one tick launch advances every
// slot by one phase. Phases run the same device bodies, with the same block
// counts per kind, as the single-problem kernels, so a problem's results do
// not depend on the scheduler or on the other slots.
enum SchedPhase : int {
  PH_IDLE = 0,   // free slot
  PH_CAST_IN,    // u32 <- (float) u64 (warm start from the caller's u)
  PH_STIFF,      // E = law(xPhys)
  PH_DIAG,       // Jacobi M^-1
  PH_KX0,        // w0 = K u (warm start)
  PH_CGINIT,     // r = f - K x0, PCG vectors
  PH_INITFIN,    // set-up sums, initial convergence test
  PH_CG,         // one PCG iteration (k)
  PH_SENS,       // dc
  PH_FILT,       // dcf = H^T (dc / Hs); NaN check
  PH_OCMAX,      // max(-dcf) partials (GPU OC constants)
  PH_OCINIT,     // bisection interval
  PH_OCSTEP,     // bisection step (sub)
  PH_OCFINAL,    // xnew at the final multiplier
  PH_COMMIT,     // xPhys_new = H xnew / Hs, preserved, x <- xnew, xPhys <- xPhys_new
  PH_DOT,        // f.u partials
  PH_SUMFIN,     // compliance; early-stop test
  PH_CAST_OUT,   // u64 <- (double) u32
  PH_DONE,       // (18)
  // FP64 iterative refinement around FP32 inner solves (jobs with refine)
  PH_Z64,        // u64 <- 0
  PH_RES,        // r64 = f - K u64 on free DOFs (FP64), |r|^2 and |f|^2 partials
  PH_RESFIN,     // residual sums; converged, capped, broken down, or an inner solve
  PH_CORR,       // u64 += correction
  // recycled starts (jobs with recycle > 0; bodies in gp_recycle.hpp)
  PH_RC_KV,      // K v for stored solution sub (oldest first)
  PH_RC_GRAM,    // Gram partials v_i.(K v_j), f.v_i
  PH_RC_SOLVE,   // Gram sums and coefficients (the slot's first block)
  PH_RC_COMB,    // start = V c (refined jobs: widened into u64)
  PH_RC_STORE,   // the finished solve's solution into the ring
  // x and p two iterations at a time (see k_tick)
  PH_ALIGN,      // one tick's wait so that PCG launch k runs on a tick of k's parity
  PH_XPFIX       // pending x and p updates of an odd last PCG iteration
};
enum SolveKind : int { SK_OC = 0, SK_COMPL = 1, SK_FINAL = 2 };

struct SchedCfg {
  int n_iters, max_iter, tol_schedule, compl_each, early_stop, warm_start;
  double tol;
  int refine, max_outer;  // FP64 refinement (outer tolerance: tol, as scheduled)
  double inner_tol;
  int warm_compl;         // compliance solves warm-started too
  int recycle;            // warm solves start from the Galerkin combination of
                          // the job's last `recycle` solutions (0: the last u)
};

struct SchedState {
  int phase, k, oc_it, sub;
  int solve, small, status, pcg_total;
  int bis_total, nan, outer, ir_brk;  // refinement sweep; inner breakdown seen
  int rc_n, rc_head;                  // recycled starts: solutions stored, next slot
  double prev_c, curr_c;
};

// A slot's block range holds as many blocks as its job's PCG sweep uses
// (nb_cg). Element/node phases spread over those blocks; phases whose block
// count fixes a partial-sum layout (nb_vec, nb_oc) run that many virtual
// blocks on them, so partial sums keep the single-problem layout.
struct SchedSlot {
  BatchProb P;
  SchedCfg cfg;
  int nb_cg, nb_node, nb_elem, nb_vec, nb_oc, cap;
  SchedState st[2];  // by tick parity: tick t reads st[t&1], writes st[(t+1)&1]
};

struct SchedTable {
  SchedSlot* slots;
  int n;
  const int* off;      // off[0..n]: slot s owns blocks [off[s], off[s+1])
  const int* b2s;      // slot of every block of the tick
  SchedState* status;  // state after the last tick, one per slot
  const int* act;      // null: the launch covers every block; else the
                       // blocks it covers (launch block i runs block act[i])
};

// Enqueues one tick with parity t & 1 (tj x tk must be 256 threads).
bool sched_tick(const SchedTable& tab, int tj, int tk, int blocks, int t, hipStream_t st);

// ------------------------------------------------------ storage types
// PCG vectors may be stored in a 16-bit type (bf16: FP32 exponent range;
// _Float16: 11-bit significand, range to 65504) with FP32 arithmetic.
using bf16 = __hip_bfloat16;
template <typename T, typename S>
__device__ __host__ inline T ld_s(S v) {
  return (T)v;
}
template <>
__device__ __host__ inline float ld_s<float, bf16>(bf16 v) {
  return __bfloat162float(v);
}
template <typename S, typename T>
__device__ __host__ inline S st_s(T v) {
  return (S)v;
}
template <>
__device__ __host__ inline bf16 st_s<bf16, float>(float v) {
  return __float2bfloat16(v);
}

// Logical block of hardware block b out of n. Workgroups are dispatched
// round-robin over the XCDs, each with its own L2; this bijection gives every
// XCD a contiguous range of logical blocks, so neighbouring tiles -- which
// share halo nodes and partly used cache lines -- run under the same L2.
__device__ inline int xcd_block(int b, int n) {
  const int x = b % kNumXcd, i = b / kNumXcd;
  const int q = n / kNumXcd, r = n - q * kNumXcd;
  return x * q + (x < r ? x : r) + i;
}

// Sum over a wave by an xor butterfly of lane shuffles; every lane gets the
// same bits. Lanes missing from a partial wave read 0.
__device__ inline double wave_sum64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// v from the lane that DPP control CTRL names, 0 where there is none.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp((int)__double2loint(v), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)__double2hiint(v), CTRL, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// Sum over a wave whose 64 lanes are all active, returned to every lane: a
// running sum within each 16-lane row (shifts by 1, 2, 4, 8), so lane 16r+15
// holds row r; then lane 15 of each row into the next row (row_bcast:15),
// so lane 63 holds rows 3 + 2 and lane 31 rows 1 + 0; then lane 31 into rows
// 2 and 3 (row_bcast:31): lane 63 holds (row3 + row2) + (row1 + row0) and is
// broadcast. Other lanes end with partial sums that are not read. Fixed
// order, and no LDS traffic: six DPP moves and adds instead of six shuffle
// round trips.
__device__ inline double wave_sum64_full(double x) {
  x = x + dpp_f64<0x111>(x);  // row_shr:1
  x = x + dpp_f64<0x112>(x);  // row_shr:2
  x = x + dpp_f64<0x114>(x);  // row_shr:4
  x = x + dpp_f64<0x118>(x);  // row_shr:8
  x = x + dpp_f64<0x142>(x);  // row_bcast:15
  x = x + dpp_f64<0x143>(x);  // row_bcast:31
  const int lo = __builtin_amdgcn_readlane((int)__double2loint(x), 63);
  const int hi = __builtin_amdgcn_readlane((int)__double2hiint(x), 63);
  return __hiloint2double(hi, lo);
}

// Threads of a block of n lanes rounded up to whole waves.
constexpr int whole_waves(int n) { return (n + 63) / 64 * 64; }

// Wave sum for a block of NB threads: the DPP form when every wave is full.
template <int NB>
__device__ inline double wave_sum_nb(double v) {
  if constexpr (NB % 64 == 0)
    return wave_sum64_full(v);
  else
    return wave_sum64(v);
}

// ---------------------------------------------------------------- launchers
// All launchers are asynchronous on `st`.
template <typename T>
void launch_ku(const Grid& g, const T* ke, const float* E, const T* u, T* y,
               const uint8_t* fixmask, const DevScalars* sc, double* partA, bool fuse_pap,
               hipStream_t st);
int gather_blocks(const Grid& g);
// mode 0: y = K u. mode 1: PCG front half of iteration `it` (u = z,
// u2 = p_old, pout = p_new, y = A p_new, p.Ap partials into parts.A).
bool launch_ku_modal(const Grid& g, const KuTile& t, const void* kc, int prec, const float* E,
                     const void* u, const void* u2, void* y, void* pout, const uint8_t* fixmask,
                     DevScalars* sc, Parts parts, int it, int nV, int mode, hipStream_t st);
int modal_blocks(const Grid& g, const KuTile& t);

template <typename T>
void launch_diag_minv(const Grid& g, const T* ke, const float* E, const uint8_t* fixmask,
                      T* diag_or_minv, bool invert, hipStream_t st);
// Jacobi M^-1 with one value per node (n_nodes floats, unmasked), the form
// the float single-kernel PCG sweep reads.
void launch_diag_node(const Grid& g, const float* ke, const float* E, float* minv_node,
                      hipStream_t st);
template <typename T>
void launch_mask_fixed(const Grid& g, T* x, const uint8_t* fixmask, hipStream_t st);
int vec_blocks(const Grid& g);
template <typename T>
void launch_pcg_init(const Grid& g, const double* f, const T* Ax, bool use_x0, T* x, T* r,
                     T* z, T* p, const T* minv, Parts parts, DevScalars* sc, double tol,
                     int max_iter, hipStream_t st);
void launch_pcg_init_fin(Parts parts, int nV, double tol, int max_iter, DevScalars* sc,
                         hipStream_t st);
// S: storage of x and p (T, or _Float16 / bf16 with T = float)
template <typename T, typename S = T>
void launch_cg_init(const Grid& g, const double* f, const T* Ax, bool use_x0, S* x, T* r, T* s0,
                    T* w1, S* p, const T* minv, const uint8_t* fixmask, Parts parts,
                    DevScalars* sc, double tol, int max_iter, hipStream_t st);
// One Jacobi-PCG iteration k (Chronopoulos-Gear form) on the modal sweep
// (prec: GP_F32 / GP_F64, or GP_F16 / GP_BF16 16-bit storage of x and p):
// reads r_{k-1} (rp), w_{k-1} (wp), s_{k-2} (sp); writes r_k, w_k, s_{k-1}
// (rn, wn, sn); updates x and p in place. two_step: odd launches leave x and
// p alone and even launches k >= 2 apply both updates (GP_F32 / GP_F64:
// bitwise the same x at every even launch; 16-bit storage rounds once per
// pair); after the last launch launch_cg_xp_fix completes an odd final
// iteration.
bool launch_cg_modal(const Grid& g, const KuTile& t, const void* kc, int prec, const float* E,
                     const void* rp, const void* wp, const void* sp, const void* minv, void* rn,
                     void* wn, void* sn, void* x, void* p, const uint8_t* fixmask,
                     DevScalars* sc, Parts parts, int k, hipStream_t st, bool two_step = false);
// After a two-step solve: when the last iteration K (sc->iter) is odd, x_K =
// x_{K-1} + alpha_{K-1} p_{K-1} with p_{K-1} = M r_{K-1} + beta_{K-1} p_{K-2},
// as iteration K's launch would have stored them (r0: the r buffer of even
// iterations). S: storage of x and p.
template <typename T, typename S = T>
void launch_cg_xp_fix(const Grid& g, const DevScalars* sc, const T* r0, const T* minv, S* x,
                      S* p, hipStream_t st);
template <typename T>
void launch_pcg_b(const Grid& g, int it, int nA, T* x, T* r, T* z, const T* p, const T* Ap,
                  const T* minv, Parts parts, DevScalars* sc, hipStream_t st);
template <typename T>
void launch_pcg_c(const Grid& g, int it, T* p, const T* z, Parts parts, DevScalars* sc,
                  hipStream_t st);

// FP64 refinement helpers
void launch_residual64(const Grid& g, const double* f, const double* Ku, const uint8_t* fixmask,
                       double* r, Parts parts, DevScalars* sc, hipStream_t st);
// r = f - K(E) u on the free DOFs in FP64 (24x24 table ke), sc->ir_rr = |r|^2,
// sc->ir_bb = |f_free|^2; vec_blocks(g) partial sums, fixed order.
void launch_res64(const Grid& g, const double* ke, const float* E, const double* u,
                  const double* f, const uint8_t* fixmask, double* r, Parts parts, DevScalars* sc,
                  hipStream_t st);
template <typename T>
void launch_add_correction(const Grid& g, double* u, const T* d, hipStream_t st);
// u += scale * d for 16-bit stored corrections (S = _Float16 or bf16)
template <typename S>
void launch_add_correction_lp(const Grid& g, double* u, const S* d, double scale, hipStream_t st);
// v *= a over all DOFs
void launch_scale64(const Grid& g, double* v, double a, hipStream_t st);
void launch_cast_f64_to_f32(const double* a, float* b, long long n, hipStream_t st);
void launch_cast_f32_to_f64(const float* a, double* b, long long n, hipStream_t st);

void launch_stiffness(const Grid& g, const double* xPhys, float* E, int law, double penal,
                      double E0, double Emin, hipStream_t st);
// mk: modal coefficients of an isotropic hex8 table (nullptr: dense ke64)
template <typename T>
void launch_sens(const Grid& g, const double* ke64, const ModalCoef<double>* mk,
                 const double* xPhys, const T* u, int law, double penal, double E0, double Emin,
                 double* dc, double* ce_sum, Parts parts, DevScalars* sc, hipStream_t st);
// `s` points to a Stencil in device memory.
void launch_filter(const Grid& g, const Stencil* s, const double* Hs, const double* in,
                   double* out, int mode, hipStream_t st);
void launch_hs(const Grid& g, const Stencil* s, double* Hs, hipStream_t st);
void launch_oc_weights(const Grid& g, const Stencil* s, const double* Hs, const uint8_t* pres,
                       double* w, hipStream_t st);
int oc_blocks(const Grid& g);
void launch_oc_init(const Grid& g, const double* dc, int variant, double volfrac, double move,
                    Parts parts, DevScalars* sc, hipStream_t st);
void launch_oc_step(const Grid& g, int s, const double* x, const double* dc, const double* w,
                    const uint8_t* pres, double n_pres, Parts parts, DevScalars* sc,
                    hipStream_t st);
void launch_oc_final(const Grid& g, const double* x, const double* dc, const uint8_t* pres,
                     const DevScalars* sc, double* xnew, hipStream_t st);
void launch_set_preserved(const Grid& g, const uint8_t* pres, double* xPhys, hipStream_t st);
template <typename TA, typename T>
void launch_dot(const TA* a, const T* b, long long n, Parts parts, DevScalars* sc,
                hipStream_t st);
void launch_reset_flags(DevScalars* sc, hipStream_t st);
// every field of *sc to zero (a fill kernel, so it can sit in a captured graph)
void launch_zero_scalars(DevScalars* sc, hipStream_t st);
void launch_zero_list(float* v, const int64_t* idx, long long n, hipStream_t st);
void launch_vec(float* y, const float* x, const float* d, float a, long long n, int mode,
                hipStream_t st);

// generic-table kernels (arbitrary edof)
// Marching cubes over an nx x ny x nz float64 sample grid (cubes in C order,
// 256 a block): mc_prepare copies the generated case table to the current
// device once (error text or null); launch_mc_count leaves per-block
// triangle offsets in off[0..nb) and the total in off[nb]; launch_mc_emit
// writes 12 floats (unit normal, 3 vertices scaled by sx, sy, sz) per
// triangle. mc_tables_host gives the table (ntri[256], tri[256][15]).
int mc_blocks(int nx, int ny, int nz);
const char* mc_prepare(int device);
void mc_tables_host(int8_t* ntri, int8_t* tri);
void launch_mc_count(const double* d, int nx, int ny, int nz, double level, int* off,
                     hipStream_t st);
void launch_mc_emit(const double* d, int nx, int ny, int nz, double level, double sx, double sy,
                    double sz, const int* off, float* out, hipStream_t st);

// C3D8 hex mesh of the solid voxels (> 0.5) of a D x H x W float32 field:
// launch_hex_count fills the workspace (2 (nb + 1) offsets, then the
// (D+1)(H+1)(W+1) node grid) and leaves the element total at ws[nb] and the
// node total at ws[2 nb + 1]; launch_hex_emit writes 8 node numbers
// (1-based) per element, 3 grid ints per node and, when elem_voxel is not
// null, each element's voxel index.
int hex_blocks(int D, int H, int W);
void launch_hex_count(const float* v, int D, int H, int W, int* ws, hipStream_t st);
void launch_hex_emit(const float* v, int D, int H, int W, int* ws, int* elems, int* node_ijk,
                     int* elem_voxel, hipStream_t st);

// Evaluator pre-processing of decoded fields: order-1 zoom to od x oh x ow,
// a separable symmetric (2 radius + 1)-tap filter with reflect boundaries
// along each axis, logistic sharpening. work and out: batch*od*oh*ow floats.
void launch_voxel_prep(const float* in, int batch, int d, int h, int w, int od, int oh, int ow,
                       const double* weights, int radius, double sharpness, float* work,
                       float* out, hipStream_t st);
template <typename T>
void launch_generic_sens(const T* xPhys, const T* u, const T* ke, const int32_t* edof,
                         long long ne, double penal, T* dc, hipStream_t st);
// deterministic gather-form fused SpMV; ws: generic_spmv_ws_ints(ne, ndof) ints
long long generic_spmv_ws_ints(long long ne, long long ndof);
template <typename T>
void launch_generic_spmv(const T* xPhys, const T* u, const T* ke, const int32_t* edof,
                         long long ne, long long ndof, double penal, int* ws, T* y,
                         hipStream_t st);
void launch_csr_spmv(const int64_t* rp, const int64_t* ci, const float* v, const float* x, float* y,
                     long long n, hipStream_t st);

// Exact FP64 banded-LU solve of the free-DOF system (device buffers of one
// problem; layout described with the kernels)
struct DirectWs {
  int nf = 0;               // free DOFs
  int bw = 0;               // lower bandwidth bound, 3 (sI + sJ + 1) + 2
  int W = 0;                // band row slot, 2 bw + 2 doubles
  int* fi = nullptr;        // ndof: free index of each DOF, -1 when fixed
  double* band = nullptr;   // nf * W
  double* b = nullptr;      // nf, right-hand side per physical slot
  double* x = nullptr;      // nf, solution per free index
  int* map = nullptr;       // 2 nf, logical row -> physical slot (two copies)
  int* urow = nullptr;      // nf, physical slot of U row k
  double* cand = nullptr;   // 2 (bw + 1), pivot candidates of the next column
  int* info = nullptr;      // 0, or k + 1 for a zero pivot in column k
};
int direct_band_width(const Grid& g);
void launch_direct(const Grid& g, const double* ke, const double* xPhys, int law, double penal,
                   double E0, double Emin, const double* f, const DirectWs& w, double* u,
                   hipStream_t st);

inline int grid_blocks(long long n, int per_thread = 1) {
  long long b = (n + (long long)kBlock * per_thread - 1) / ((long long)kBlock * per_thread);
  if (b < 1) b = 1;
  return (int)b;
}

// ------------------------------------------------------- recycled starts
// Galerkin start u0 = V c, (V^T K V) c = V^T f, over up to kRecycleMax stored
// solutions v_i with their products K v_i at the current moduli
// (gp_recycle.hip). coef receives c (kRecycleMax entries, 0 for solutions
// left out as nearly dependent) and then the number kept.
constexpr int kRecycleMax = 8;
struct RecycleSet {
  const float* v[kRecycleMax];
  const float* kv[kRecycleMax];
  int n;
};
size_t recycle_part_doubles();
bool launch_recycle_guess(const RecycleSet& rs, const double* f, long long nd, double* part,
                          double* coef, float* u0, hipStream_t st);

}  // namespace gp
