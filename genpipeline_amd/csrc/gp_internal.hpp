// Internal declarations shared by the kernel and API translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <cstddef>
#include "../../include/gp_simp.h"

namespace gp {

constexpr int kBlock = 256;          // threads per block for all 1-D kernels
constexpr int kMaxPartials = 2048;   // max blocks of a reducing grid-stride kernel
constexpr int kMaxOffsets = 343;     // filter stencil cap: (2*3+1)^3, rmin <= 3

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

// Scalars of one PCG solve / one OC update, resident in device memory.
struct DevScalars {
  double rz, pAp, rr, bb, alpha, beta, tol2, rr0;
  double l1, l2, lmid, oc_eps, oc_tol, oc_floor, volfrac, move;
  double dot;          // generic reduction result
  int iter, max_iter, done, status;
  int oc_steps, oc_max, oc_done, nan_flag;
  unsigned int cnt[8]; // last-block counters, one per reducing kernel kind
};

enum CounterSlot { CNT_A = 0, CNT_B, CNT_INIT, CNT_OC, CNT_DOT, CNT_MAX, CNT_SENS, CNT_RES };

// ---------------------------------------------------------------- launchers
// All launchers are asynchronous on `st`.
template <typename T>
void launch_ku(const Grid& g, const T* ke, const float* E, const T* u, T* y,
               const uint8_t* fixmask, DevScalars* sc, double* partials, bool fuse_pap,
               hipStream_t st);
template <typename T>
void launch_diag_minv(const Grid& g, const T* ke, const float* E, const uint8_t* fixmask,
                      T* diag_or_minv, bool invert, hipStream_t st);
template <typename T>
void launch_mask_fixed(const Grid& g, T* x, const uint8_t* fixmask, hipStream_t st);
template <typename T>
void launch_pcg_init(const Grid& g, const double* f, const T* Ax, bool use_x0, T* x, T* r,
                     T* z, T* p, const T* minv, DevScalars* sc, double* partials, double tol,
                     int max_iter, hipStream_t st);
template <typename T>
void launch_pcg_b(const Grid& g, T* x, T* r, T* z, const T* p, const T* Ap, const T* minv,
                  DevScalars* sc, double* partials, hipStream_t st);
template <typename T>
void launch_pcg_c(const Grid& g, T* p, const T* z, DevScalars* sc, hipStream_t st);

// FP64 refinement helpers
void launch_residual64(const Grid& g, const double* f, const double* Ku, const uint8_t* fixmask,
                       double* r, DevScalars* sc, double* partials, hipStream_t st);
template <typename T>
void launch_add_correction(const Grid& g, double* u, const T* d, hipStream_t st);
void launch_cast_f64_to_f32(const double* a, float* b, long long n, hipStream_t st);
void launch_cast_f32_to_f64(const float* a, double* b, long long n, hipStream_t st);

void launch_stiffness(const Grid& g, const double* xPhys, float* E, int law, double penal,
                      double E0, double Emin, hipStream_t st);
template <typename T>
void launch_sens(const Grid& g, const double* ke64, const double* xPhys, const T* u, int law,
                 double penal, double E0, double Emin, double* dc, double* ce_sum,
                 DevScalars* sc, double* partials, hipStream_t st);
// `s` points to a Stencil in device memory.
void launch_filter(const Grid& g, const Stencil* s, const double* Hs, const double* in,
                   double* out, int mode, hipStream_t st);
void launch_hs(const Grid& g, const Stencil* s, double* Hs, hipStream_t st);
void launch_oc_weights(const Grid& g, const Stencil* s, const double* Hs, const uint8_t* pres,
                       double* w, hipStream_t st);
void launch_oc_init(const Grid& g, const double* dc, int variant, double volfrac, double move,
                    DevScalars* sc, double* partials, hipStream_t st);
void launch_oc_step(const Grid& g, const double* x, const double* dc, const double* w,
                    const uint8_t* pres, double n_pres, DevScalars* sc, double* partials,
                    hipStream_t st);
void launch_oc_final(const Grid& g, const double* x, const double* dc, const uint8_t* pres,
                     DevScalars* sc, double* xnew, hipStream_t st);
void launch_set_preserved(const Grid& g, const uint8_t* pres, double* xPhys, hipStream_t st);
template <typename TA, typename T>
void launch_dot(const TA* a, const T* b, long long n, DevScalars* sc, double* partials,
                hipStream_t st);
void launch_reset_flags(DevScalars* sc, hipStream_t st);
void launch_zero_list(float* v, const int64_t* idx, long long n, hipStream_t st);
void launch_vec(float* y, const float* x, const float* d, float a, long long n, int mode,
                hipStream_t st);

// generic-table kernels (arbitrary edof)
void launch_generic_sens(const float* xPhys, const float* u, const float* ke, const int32_t* edof,
                         long long ne, double penal, float* dc, hipStream_t st);
void launch_generic_spmv(const float* xPhys, const float* u, const float* ke, const int32_t* edof,
                         long long ne, double penal, float* y, hipStream_t st);
void launch_csr_spmv(const int64_t* rp, const int64_t* ci, const float* v, const float* x, float* y,
                     long long n, hipStream_t st);

inline int grid_blocks(long long n, int per_thread = 1) {
  long long b = (n + (long long)kBlock * per_thread - 1) / ((long long)kBlock * per_thread);
  if (b < 1) b = 1;
  return (int)b;
}
inline int reduce_blocks(long long n) {
  int b = grid_blocks(n);
  return b > kMaxPartials ? kMaxPartials : b;
}

}  // namespace gp
