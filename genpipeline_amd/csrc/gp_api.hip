// C ABI and host-side drivers of the SIMP inner loop (libgpsimp.so).
//
// Host code here owns per-problem device state and sequences the kernels in
// gp_kernels.hip on the caller's stream. PCG runs without a host round trip
// per iteration: iterations are issued in chunks, each chunk's device-side
// status is copied to pinned memory, and the next chunk is already queued
// while the host inspects the previous one (kernels of a finished solve
// return immediately).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "gp_internal.hpp"

using namespace gp;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      return fail(GP_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),    \
                  __FILE__, __LINE__);                                                  \
  } while (0)

#define LAUNCHCHK()                                                                     \
  do {                                                                                  \
    hipError_t _e = hipGetLastError();                                                  \
    if (_e != hipSuccess)                                                               \
      return fail(GP_ERR_HIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString(_e), \
                  __FILE__, __LINE__);                                                  \
  } while (0)

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Exact trilinear hex stiffness of a unit cube, 2x2x2 Gauss, isotropic
// material with E=1 and Poisson ratio nu; engineering shear strains; DOF
// order 3*corner + component with the C3D8 corner order.
void hex8_ke(double nu, double* K) {
  static const int C[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const double lam = nu / ((1 + nu) * (1 - 2 * nu));
  const double mu = 1.0 / (2 * (1 + nu));
  double D[6][6] = {};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) D[a][b] = lam + (a == b ? 2 * mu : 0.0);
  for (int a = 3; a < 6; ++a) D[a][a] = mu;
  std::memset(K, 0, sizeof(double) * 576);
  const double gp1 = 1.0 / std::sqrt(3.0);
  for (int qa = 0; qa < 2; ++qa)
    for (int qb = 0; qb < 2; ++qb)
      for (int qc = 0; qc < 2; ++qc) {
        const double xi[3] = {qa ? gp1 : -gp1, qb ? gp1 : -gp1, qc ? gp1 : -gp1};
        double B[6][24] = {};
        for (int n = 0; n < 8; ++n) {
          double s[3], dN[3];
          for (int d = 0; d < 3; ++d) s[d] = 2.0 * C[n][d] - 1.0;
          // N = prod_d (1 + s_d xi_d) / 8 ; d/dx = 2 d/dxi on the unit cube
          for (int d = 0; d < 3; ++d) {
            double v = s[d] / 8.0;
            for (int o = 0; o < 3; ++o)
              if (o != d) v *= (1.0 + s[o] * xi[o]);
            dN[d] = 2.0 * v;
          }
          B[0][3 * n] = dN[0];
          B[1][3 * n + 1] = dN[1];
          B[2][3 * n + 2] = dN[2];
          B[3][3 * n] = dN[1];
          B[3][3 * n + 1] = dN[0];
          B[4][3 * n + 1] = dN[2];
          B[4][3 * n + 2] = dN[1];
          B[5][3 * n] = dN[2];
          B[5][3 * n + 2] = dN[0];
        }
        double DB[6][24];
        for (int a = 0; a < 6; ++a)
          for (int c = 0; c < 24; ++c) {
            double v = 0;
            for (int b = 0; b < 6; ++b) v += D[a][b] * B[b][c];
            DB[a][c] = v;
          }
        for (int r = 0; r < 24; ++r)
          for (int c = 0; c < 24; ++c) {
            double v = 0;
            for (int a = 0; a < 6; ++a) v += B[a][r] * DB[a][c];
            K[r * 24 + c] += v * 0.125;  // det J * weights = (1/2)^3
          }
      }
}

}  // namespace

// --------------------------------------------------------------- problem
struct gp_problem {
  gp_spec spec{};
  Grid g{};
  int device = 0;
  double ke64_h[576];
  Stencil st_h{};
  // device constants: the set-up buffers below (element tables, stencil,
  // Hs, w, supports, load, scalars, partials) are 256-byte aligned pieces of
  // one allocation, so a problem costs one hipMalloc and one hipFree for them
  void* arena = nullptr;
  double* ke64 = nullptr;
  float* ke32 = nullptr;
  Stencil* st = nullptr;
  double* Hs = nullptr;
  double* w = nullptr;          // OC volume weights
  uint8_t* fixmask = nullptr;   // per node, bit c = DOF c fixed
  uint8_t* pres = nullptr;      // per element or null
  double n_pres = 0;
  double* f = nullptr;          // load vector (float64)
  bool has_bcs = false;
  // scalars
  DevScalars* sc = nullptr;     // [0] PCG, [1] OC/dots, [2] refinement
  DevScalars* h_sc = nullptr;   // pinned mirrors, 3 slots
  double* partials = nullptr;   // backing store of `parts`
  size_t n_partials = 0;
  Parts parts{};
  // modal K.u: coefficients projected from KE, valid when the table has the
  // isotropic hex8 modal structure (always for GP_KE_HEX8)
  bool modal_ok = false;
  ModalCoef<float> mc32{};
  ModalCoef<double> mc64{};
  KuTile tile32{8, 32, 4};
  KuTile tile64{16, 16, 4};
  bool use_modal32 = true, use_modal64 = true;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // PCG workspace per precision
  void* ws32 = nullptr;         // PCG vectors (float, 8*ndof), layout in pcg_core
  void* ws16 = nullptr;         // 16-bit PCG: M^-1 (float) + 8 vectors (layout in pcg_core_lp)
  void* ws64 = nullptr;         // PCG vectors (double, 8*ndof)
  double* r64 = nullptr;        // refinement residual
  double* y64 = nullptr;        // refinement K u
  float* d32 = nullptr;         // refinement correction
  // SIMP loop workspace
  float* E = nullptr;
  double* dc = nullptr;
  double* dcf = nullptr;
  double* xnew = nullptr;
  double* xphys_new = nullptr;
  double* ce_sum = nullptr;
  // recycled starts: a ring of the last rc_cap FP32 solutions, their K v at
  // the current moduli, the Gram partials and the coefficients (one
  // allocation); rc_n solutions stored, the next one goes to slot rc_head
  void* rc_mem = nullptr;
  float* rc_v = nullptr;
  float* rc_kv = nullptr;
  double* rc_part = nullptr;
  double* rc_coef = nullptr;
  int rc_cap = 0, rc_n = 0, rc_head = 0;
  // exact direct solve: host copy of the supports (the free numbering is
  // built from it) and the banded-LU workspace, one allocation
  std::vector<uint8_t> fixmask_h;
  std::vector<int> fi_h;
  DirectWs dws{};
  void* dws_mem = nullptr;
  size_t dws_bytes = 0;
  int* dws_info_h = nullptr;    // pinned
  bool dws_valid = false;
  // sampled iteration timing: events bracket every `prof_every`-th chunk of
  // PCG iteration launches
  int iter_hint[3] = {0, 0, 0};  // last PCG iteration count (fp32, fp64, 16-bit)
  int prof_every = 0;
  long long prof_counter = 0;
  double prof_ms = 0.0;
  long long prof_n = 0;
  std::vector<hipEvent_t> prof_pool;
  struct ProfSample {
    hipEvent_t a, b;
    int launch;  // index of the chunk's first iteration launch within its solve
    int n;       // launches in the chunk
  };
  std::vector<ProfSample> prof_pending;
};

namespace {

// Device allocator: hipMalloc / hipFree, or the caller's pair set by
// gp_set_allocator (e.g. a framework's caching allocator). g_live counts the
// allocations alive, so the pair changes only while none is.
std::atomic<gp_alloc_fn> g_alloc_fn{nullptr};
std::atomic<gp_free_fn> g_free_fn{nullptr};
std::atomic<void*> g_alloc_ctx{nullptr};
std::atomic<long long> g_live{0};
std::mutex g_alloc_mu;

int alloc(void** p, size_t bytes) {
  if (bytes == 0) {  // nothing to hold (dev_free ignores NULL)
    *p = nullptr;
    return GP_OK;
  }
  const gp_alloc_fn fn = g_alloc_fn.load();
  if (fn) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    *p = fn(bytes, dev, g_alloc_ctx.load());
    if (!*p) return fail(GP_ERR_OOM, "the caller's allocator returned NULL for %zu bytes", bytes);
  } else {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
      *p = nullptr;
      return fail(GP_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
  }
  ++g_live;
  return GP_OK;
}

// Frees an allocation made by alloc(). Like hipFree, it first waits for the
// device's queued work, which may still use the buffer: a caching allocator
// would otherwise hand the memory out again while a kernel reads it.
// synced: the caller has just waited (one wait for several frees).
hipError_t dev_free(void* p, bool synced = false) {
  if (!p) return hipSuccess;
  --g_live;
  const gp_free_fn fn = g_free_fn.load();
  if (!fn) return hipFree(p);
  const hipError_t e = synced ? hipSuccess : hipDeviceSynchronize();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  fn(p, dev, g_alloc_ctx.load());
  return e;
}

// One wait before several dev_free(p, true) calls (hipFree waits by itself).
void sync_for_free() {
  if (g_free_fn.load()) hipDeviceSynchronize();
}
#define ALLOC(ptr, bytes)                                      \
  do {                                                         \
    int _r = alloc(reinterpret_cast<void**>(&(ptr)), (bytes)); \
    if (_r) return _r;                                         \
  } while (0)

const KuTile kTiles[] = {{8, 32, 1}, {16, 16, 1}, {14, 18, 1}, {6, 34, 1},
                         {14, 34, 1}, {4, 64, 1}, {10, 26, 1}, {7, 34, 1}};

// entries per partials region: enough for every reducing launch shape
size_t partial_cap(const Grid& g) {
  size_t n = (size_t)grid_blocks(g.ndof);
  n = std::max(n, (size_t)grid_blocks(g.nn));
  n = std::max(n, (size_t)grid_blocks(g.ne));
  for (const KuTile& t : kTiles) n = std::max(n, (size_t)modal_blocks(g, t));
  return n + 16;
}
// 13 regions of cap entries, then the OC bisection's candidate volumes (two
// parities of kOcCand x cap)
constexpr int kRegions = 13 + 2 * kOcCand;

// Khat entry (mode mA comp qA, mode mB comp qB) of a 24x24 table, /64.
double khat(const double* ke, int mA, int qA, int mB, int qB) {
  static const int C[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  auto W = [&](int c, int m) {
    double w = 1.0;
    for (int d = 0; d < 3; ++d)
      if ((m >> (2 - d)) & 1) w *= (2.0 * C[c][d] - 1.0);
    return w;
  };
  double s = 0.0;
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) s += W(a, mA) * ke[(3 * a + qA) * 24 + 3 * b + qB] * W(b, mB);
  return s / 64.0;
}

// Modes as bit masks x=4, y=2, z=1 (bit of mi, mj, mk).
enum { MX = 4, MY = 2, MZ = 1, MXY = 6, MYZ = 3, MXZ = 5, MXYZ = 7 };

// v = Khat b / 64 with the 8-coefficient isotropic form, exactly as the
// modal kernel applies it (b, v indexed [mode mask][component]).
void modal_apply(const ModalCoef<double>& k, const double (&b)[8][3], double (&v)[8][3]) {
  for (auto& r : v) r[0] = r[1] = r[2] = 0.0;
  const double tr = b[MX][0] + b[MY][1] + b[MZ][2];
  v[MX][0] = k.An * b[MX][0] + k.Ln * tr;
  v[MY][1] = k.An * b[MY][1] + k.Ln * tr;
  v[MZ][2] = k.An * b[MZ][2] + k.Ln * tr;
  v[MX][1] = v[MY][0] = k.S * (b[MX][1] + b[MY][0]);
  v[MX][2] = v[MZ][0] = k.S * (b[MX][2] + b[MZ][0]);
  v[MY][2] = v[MZ][1] = k.S * (b[MY][2] + b[MZ][1]);
  v[MXY][0] = k.B1 * b[MXY][0] + k.B2 * b[MYZ][2];
  v[MYZ][2] = k.B2 * b[MXY][0] + k.B1 * b[MYZ][2];
  v[MXY][1] = k.B1 * b[MXY][1] + k.B2 * b[MXZ][2];
  v[MXZ][2] = k.B2 * b[MXY][1] + k.B1 * b[MXZ][2];
  v[MYZ][1] = k.B1 * b[MYZ][1] + k.B2 * b[MXZ][0];
  v[MXZ][0] = k.B2 * b[MYZ][1] + k.B1 * b[MXZ][0];
  const double s3 = b[MXY][2] + b[MYZ][0] + b[MXZ][1];
  v[MXY][2] = k.T1 * b[MXY][2] + k.T2 * s3;
  v[MYZ][0] = k.T1 * b[MYZ][0] + k.T2 * s3;
  v[MXZ][1] = k.T1 * b[MXZ][1] + k.T2 * s3;
  for (int q = 0; q < 3; ++q) v[MXYZ][q] = k.X * b[MXYZ][q];
}

// Projects KE onto the modal basis; true when the 8-coefficient form
// reproduces every Khat entry to 1e-10 relative.
bool modal_coefs(const double* ke, ModalCoef<double>* out) {
  ModalCoef<double> k;
  k.Ln = khat(ke, MX, 0, MY, 1);
  k.An = khat(ke, MX, 0, MX, 0) - k.Ln;
  k.S = khat(ke, MX, 1, MX, 1);
  k.B1 = khat(ke, MXY, 0, MXY, 0);
  k.B2 = khat(ke, MXY, 0, MYZ, 2);
  k.T2 = khat(ke, MXY, 2, MYZ, 0);
  k.T1 = khat(ke, MXY, 2, MXY, 2) - k.T2;
  k.X = khat(ke, MXYZ, 0, MXYZ, 0);
  double mx = 0.0, err = 0.0;
  for (int mB = 0; mB < 8; ++mB)
    for (int qB = 0; qB < 3; ++qB) {
      double b[8][3] = {}, v[8][3];
      b[mB][qB] = 1.0;
      modal_apply(k, b, v);
      for (int mA = 0; mA < 8; ++mA)
        for (int qA = 0; qA < 3; ++qA) {
          const double ref = khat(ke, mA, qA, mB, qB);
          mx = std::max(mx, std::fabs(ref));
          err = std::max(err, std::fabs(ref - v[mA][qA]));
        }
    }
  *out = k;
  return mx > 0 && err <= 1e-10 * mx;
}

// K.u on the problem's grid in precision T; returns the number of blocks
// (= p.Ap partials written when fuse).
template <typename T>
int ku_dispatch(gp_problem* P, const float* E, const T* u, T* y, bool mask, bool fuse,
                hipStream_t s);

int validate_spec(const gp_spec* s) {
  if (!s) return fail(GP_ERR_ARG, "spec is NULL");
  if (s->nx < 1 || s->ny < 1 || s->nz < 1) return fail(GP_ERR_ARG, "grid must be >= 1 per axis");
  if (s->nx > (1 << 20) || s->ny > (1 << 20) || s->nz > (1 << 20))
    return fail(GP_ERR_ARG, "grid axis above 2^20 elements");
  const long long nn = (long long)(s->nx + 1) * (s->ny + 1) * (s->nz + 1);
  if (nn > (1ll << 29)) return fail(GP_ERR_ARG, "grid too large (%lld nodes)", nn);
  if (s->law != GP_LAW_CLAMP && s->law != GP_LAW_EMIN) return fail(GP_ERR_ARG, "bad law");
  if (s->oc_variant != GP_OC_CPU && s->oc_variant != GP_OC_GPU) return fail(GP_ERR_ARG, "bad oc_variant");
  if (s->ke_kind == GP_KE_CUSTOM && s->ke == nullptr) return fail(GP_ERR_ARG, "custom KE is NULL");
  if (s->ke_kind != GP_KE_HEX8 && s->ke_kind != GP_KE_CUSTOM) return fail(GP_ERR_ARG, "bad ke_kind");
  if (!(s->rmin > 0) || s->rmin > 3.0) return fail(GP_ERR_ARG, "rmin must be in (0, 3]");
  if (s->filter_weights != GP_FILTER_F64 && s->filter_weights != GP_FILTER_F32)
    return fail(GP_ERR_ARG, "bad filter_weights");
  if (s->ke_kind == GP_KE_HEX8 && !(s->nu > -1.0 && s->nu < 0.5)) return fail(GP_ERR_ARG, "nu out of range");
  return GP_OK;
}

Grid make_grid(const gp_spec* s) {
  Grid g;
  g.nx = s->nx;
  g.ny = s->ny;
  g.nz = s->nz;
  g.sJ = s->nz + 1;
  g.sI = (s->ny + 1) * (s->nz + 1);
  g.eJ = s->nz;
  g.eI = s->ny * s->nz;
  g.nn = (s->nx + 1) * g.sI;
  g.ne = s->nx * s->ny * s->nz;
  g.ndof = 3ll * g.nn;
  return g;
}

// f32: each weight rounded once to float32 (SIMPSolverGPU._build_filter
// fills its value arrays with np.float32, simp_solver_gpu.py:915). Zero
// weights (|d| = rmin) are left out: they add nothing to a product or a sum.
void make_stencil(double rmin, bool f32, Stencil* st) {
  const int r = (int)std::ceil(rmin);
  st->n = 0;
  for (int di = -r; di <= r; ++di)
    for (int dj = -r; dj <= r; ++dj)
      for (int dk = -r; dk <= r; ++dk) {
        const double dist = std::sqrt((double)(di * di + dj * dj + dk * dk));
        if (dist > rmin) continue;
        const double w = f32 ? (double)(float)(rmin - dist) : rmin - dist;
        if (!(w > 0.0)) continue;
        st->di[st->n] = di;
        st->dj[st->n] = dj;
        st->dk[st->n] = dk;
        st->w[st->n] = w;
        st->n++;
      }
}

int sync_read(gp_problem* P, int slot, hipStream_t s) {
  HIPCHK(hipMemcpyAsync(&P->h_sc[slot], &P->sc[slot], sizeof(DevScalars), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return GP_OK;
}

int ensure_pcg_ws(gp_problem* P, int prec) {
  const size_t nd = (size_t)P->g.ndof;
  if (prec == GP_F64) {
    if (!P->ws64) ALLOC(P->ws64, 8 * nd * sizeof(double));
  } else {
    if (!P->ws32) ALLOC(P->ws32, 8 * nd * sizeof(float));
  }
  return GP_OK;
}

int ensure_lp_ws(gp_problem* P) {
  const size_t nd = (size_t)P->g.ndof;
  if (!P->ws16) ALLOC(P->ws16, 7 * nd * sizeof(float) + 2 * nd * sizeof(uint16_t));
  return GP_OK;
}

// the 16-bit correction vector d of the refinement (after p in ws16)
template <typename S>
S* lp_correction(gp_problem* P) {
  const size_t nd = (size_t)P->g.ndof;
  return reinterpret_cast<S*>(reinterpret_cast<float*>(P->ws16) + 7 * nd) + nd;
}

int ensure_refine_ws(gp_problem* P) {
  const size_t nd = (size_t)P->g.ndof;
  if (!P->r64) ALLOC(P->r64, nd * sizeof(double));
  if (!P->y64) ALLOC(P->y64, nd * sizeof(double));
  if (!P->d32) ALLOC(P->d32, nd * sizeof(float));
  return GP_OK;
}

// Bytes of the ring for m solutions of nd DOFs: m solutions and their
// products K v in float (rounded up to 256 B), then the Gram partials and the
// coefficients in FP64.
size_t recycle_vec_bytes(size_t nd, int m) {
  return (2 * (size_t)m * nd * sizeof(float) + 255) & ~size_t(255);
}
size_t recycle_ws_bytes(size_t nd, int m) {
  return recycle_vec_bytes(nd, m) + recycle_part_doubles() * sizeof(double);
}

// Recycled-start workspace for up to m stored solutions (kept when large
// enough; the ring restarts empty).
int ensure_recycle_ws(gp_problem* P, int m) {
  if (P->rc_cap >= m) return GP_OK;
  if (P->rc_mem) {
    HIPCHK(dev_free(P->rc_mem));
    P->rc_mem = nullptr;
  }
  const size_t nd = (size_t)P->g.ndof;
  const size_t vb = recycle_vec_bytes(nd, m);
  ALLOC(P->rc_mem, recycle_ws_bytes(nd, m));
  char* b = static_cast<char*>(P->rc_mem);
  P->rc_v = reinterpret_cast<float*>(b);
  P->rc_kv = P->rc_v + (size_t)m * nd;
  P->rc_part = reinterpret_cast<double*>(b + vb);
  P->rc_coef = P->rc_part + recycle_part_doubles() - (kRecycleMax + 1);
  P->rc_cap = m;
  P->rc_n = P->rc_head = 0;
  return GP_OK;
}

// Galerkin start over the n vectors V[i] (device, n_dof floats each) at the
// moduli E, into u0; K V[i] go to the workspace's product slots.
int recycle_guess(gp_problem* P, const float* E, const float* const* V, int n, float* u0,
                  hipStream_t s) {
  if (n < 1 || n > kRecycleMax) return fail(GP_ERR_ARG, "recycle: 1 to %d vectors", kRecycleMax);
  int rc = ensure_recycle_ws(P, n);
  if (rc) return rc;
  const size_t nd = (size_t)P->g.ndof;
  RecycleSet rs{};
  rs.n = n;
  for (int i = 0; i < n; ++i) {
    rs.v[i] = V[i];
    rs.kv[i] = P->rc_kv + i * nd;
    ku_dispatch<float>(P, E, V[i], P->rc_kv + i * nd, true, false, s);
  }
  launch_recycle_guess(rs, P->f, (long long)nd, P->rc_part, P->rc_coef, u0, s);
  LAUNCHCHK();
  return GP_OK;
}

// Largest band slot (the LDS of the LU kernels holds a slot's worth of
// doubles) and largest band (device bytes) of an exact direct solve.
constexpr int kDirectMaxW = 8000;
constexpr size_t kDirectMaxBytes = size_t(16) << 30;

int ensure_direct_ws(gp_problem* P, hipStream_t s) {
  if (P->dws_valid) return GP_OK;
  const Grid& g = P->g;
  const int bw = direct_band_width(g);
  const int W = 2 * bw + 2;
  if (W > kDirectMaxW)
    return fail(GP_ERR_ARG,
                "direct solve: band slot of %d doubles exceeds %d (node strides %d, %d); "
                "grids with (ny+1)(nz+1) up to about 1300 are supported",
                W, kDirectMaxW, g.sI, g.sJ);
  P->fi_h.assign(g.ndof, -1);
  int nf = 0;
  for (long long d = 0; d < g.ndof; ++d)
    if (!((P->fixmask_h[d / 3] >> (d % 3)) & 1)) P->fi_h[d] = nf++;
  const size_t band = (size_t)nf * W * sizeof(double);
  if (band > kDirectMaxBytes)
    return fail(GP_ERR_ARG, "direct solve: band of %zu bytes exceeds %zu", band, kDirectMaxBytes);
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t o_fi = 0, o_band = up(o_fi + g.ndof * sizeof(int));
  const size_t o_b = up(o_band + band), o_x = up(o_b + nf * sizeof(double));
  const size_t o_map = up(o_x + nf * sizeof(double)), o_urow = up(o_map + 2 * nf * sizeof(int));
  const size_t o_cand = up(o_urow + nf * sizeof(int));
  const size_t o_info = up(o_cand + 2 * (bw + 1) * sizeof(double));
  const size_t bytes = up(o_info + sizeof(int));
  if (bytes > P->dws_bytes) {
    if (P->dws_mem) {
      HIPCHK(hipStreamSynchronize(s));  // the old workspace may still be in use on s
      HIPCHK(dev_free(P->dws_mem));
      P->dws_mem = nullptr;
      P->dws_bytes = 0;
    }
    ALLOC(P->dws_mem, bytes);
    P->dws_bytes = bytes;
  }
  if (!P->dws_info_h) HIPCHK(hipHostMalloc((void**)&P->dws_info_h, sizeof(int)));
  char* base = (char*)P->dws_mem;
  DirectWs& w = P->dws;
  w.nf = nf;
  w.bw = bw;
  w.W = W;
  w.fi = (int*)(base + o_fi);
  w.band = (double*)(base + o_band);
  w.b = (double*)(base + o_b);
  w.x = (double*)(base + o_x);
  w.map = (int*)(base + o_map);
  w.urow = (int*)(base + o_urow);
  w.cand = (double*)(base + o_cand);
  w.info = (int*)(base + o_info);
  // fi_h stays alive with the problem, so the copy may complete later
  HIPCHK(hipMemcpyAsync(w.fi, P->fi_h.data(), g.ndof * sizeof(int), hipMemcpyHostToDevice, s));
  P->dws_valid = true;
  return GP_OK;
}

// u = K(xPhys)^-1 f on the free DOFs (0 on the fixed ones), exact in FP64.
// Blocks for the pivot check.
int direct_solve(gp_problem* P, const double* xPhys, const double* f, double* u,
                 gp_pcg_stats* st, hipStream_t s) {
  std::memset(st, 0, sizeof(*st));
  int rc = ensure_direct_ws(P, s);
  if (rc) return rc;
  const gp_spec& sp = P->spec;
  launch_direct(P->g, P->ke64, xPhys, sp.law, sp.penal, sp.E0, sp.Emin, f, P->dws, u, s);
  LAUNCHCHK();
  *P->dws_info_h = 0;
  if (P->dws.nf > 0)
    HIPCHK(hipMemcpyAsync(P->dws_info_h, P->dws.info, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (*P->dws_info_h) {
    st->status = GP_BREAKDOWN;
    return fail(GP_BREAKDOWN, "direct solve: zero pivot in column %d of %d (singular system)",
                *P->dws_info_h - 1, P->dws.nf);
  }
  return GP_OK;
}

int ensure_simp_ws(gp_problem* P) {
  const size_t ne = (size_t)P->g.ne;
  if (!P->E) ALLOC(P->E, ne * sizeof(float));
  if (!P->dc) ALLOC(P->dc, ne * sizeof(double));
  if (!P->dcf) ALLOC(P->dcf, ne * sizeof(double));
  if (!P->xnew) ALLOC(P->xnew, ne * sizeof(double));
  if (!P->xphys_new) ALLOC(P->xphys_new, ne * sizeof(double));
  if (!P->ce_sum) ALLOC(P->ce_sum, sizeof(double));
  return GP_OK;
}

template <typename T>
int ku_dispatch(gp_problem* P, const float* E, const T* u, T* y, bool mask, bool fuse,
                hipStream_t s) {
  const bool f32 = std::is_same<T, float>::value;
  const uint8_t* fm = mask ? P->fixmask : nullptr;
  const DevScalars* sc = &P->sc[0];
  if (!fuse && P->modal_ok && (f32 ? P->use_modal32 : P->use_modal64)) {
    const KuTile& t = f32 ? P->tile32 : P->tile64;
    const void* kc = f32 ? (const void*)&P->mc32 : (const void*)&P->mc64;
    if (launch_ku_modal(P->g, t, kc, f32 ? GP_F32 : GP_F64, E, u, nullptr, y, nullptr, fm,
                        &P->sc[0], P->parts, 0, 0, 0, s))
      return modal_blocks(P->g, t);
  }
  const T* ke = reinterpret_cast<const T*>(f32 ? (const void*)P->ke32 : (const void*)P->ke64);
  launch_ku<T>(P->g, ke, E, u, y, fm, sc, P->parts.A, fuse, s);
  return gather_blocks(P->g);
}

// Issues PCG iteration launches after the set-up: chunked and speculative (one
// chunk always queued behind the one whose status the host inspects), until
// the device reports the solve finished or the cap is passed. issue_one(m)
// enqueues iteration launch m. Keeps the sampled timing and the iteration
// hint of slot hint_slot; fills iters / status / relres.
template <typename IssueOne>
int pcg_drive(gp_problem* P, int hint_slot, int max_iter, bool modal, IssueOne issue_one,
              gp_pcg_stats* st, hipStream_t s) {
  DevScalars* sc = &P->sc[0];
  auto take_event = [&]() {
    hipEvent_t e = nullptr;
    if (!P->prof_pool.empty()) {
      e = P->prof_pool.back();
      P->prof_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  };
  int launch_idx = 0;
  auto issue = [&](int n_it) {
    // sampled timing: one event pair brackets a whole chunk of back-to-back
    // iteration launches (no other work in between)
    const bool sample = P->prof_every > 0 && (P->prof_counter++ % P->prof_every) == 0;
    hipEvent_t ea = sample ? take_event() : nullptr;
    hipEvent_t eb = sample ? take_event() : nullptr;
    const int first = launch_idx;
    if (ea && eb) hipEventRecord(ea, s);
    for (int it = 0; it < n_it; ++it, ++launch_idx) issue_one(launch_idx);
    if (ea && eb) {
      hipEventRecord(eb, s);
      P->prof_pending.push_back({ea, eb, first, n_it});
    } else {
      if (ea) P->prof_pool.push_back(ea);
      if (eb) P->prof_pool.push_back(eb);
    }
  };
  // Chunk sizes: the previous solve's iteration count (consecutive solves of
  // one problem need similar counts) lets most launches go out in large
  // chunks; near that count, and without a hint, chunks stay small so few
  // launches are issued after convergence.
  const int hint = P->iter_hint[hint_slot];
  int issued = 0;
  auto next_chunk = [&](int prev) {
    int c = std::min(prev * 2, 64);
    if (hint > 0) {
      const int target = (int)(0.8 * hint) - issued;
      c = target >= 32 ? std::min(target, 256) : 16;
    }
    // never far past the iteration cap
    return std::max(8, std::min(c, max_iter + 2 - issued));
  };
  int chunk = next_chunk(4);
  int slot = 0;
  // the init status doubles as the first checkpoint
  HIPCHK(hipMemcpyAsync(&P->h_sc[0], sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(P->ev[0], s));
  issue(chunk);
  issued += chunk;
  HIPCHK(hipMemcpyAsync(&P->h_sc[1], sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(P->ev[1], s));
  LAUNCHCHK();
  while (true) {
    HIPCHK(hipEventSynchronize(P->ev[slot]));
    const DevScalars& h = P->h_sc[slot];
    if (h.done || issued >= max_iter + 2 * chunk) break;
    chunk = next_chunk(chunk);
    issue(chunk);
    issued += chunk;
    HIPCHK(hipMemcpyAsync(&P->h_sc[slot], sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(P->ev[slot], s));
    LAUNCHCHK();
    slot ^= 1;
  }
  int rc = sync_read(P, 0, s);
  if (rc) return rc;
  const DevScalars& h = P->h_sc[0];
  // Launch m ran the kernel body iff m < iterations completed (later launches
  // found the solve done and returned at once); only chunks whose launches
  // all ran count.
  for (auto& pr : P->prof_pending) {
    float ms = 0.f;
    const int last = pr.launch + pr.n - 1;
    const bool ran = modal ? last <= h.iter : last < h.iter;
    if (ran && hipEventElapsedTime(&ms, pr.a, pr.b) == hipSuccess) {
      P->prof_ms += ms;
      P->prof_n += pr.n;
    }
    P->prof_pool.push_back(pr.a);
    P->prof_pool.push_back(pr.b);
  }
  P->prof_pending.clear();
  P->iter_hint[hint_slot] = h.iter;
  st->status = h.status;
  st->iters += h.iter;
  st->relres = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
  if (!h.done) st->status = GP_NOT_CONVERGED;
  return GP_OK;
}


// Plain PCG in precision T on the free DOFs. u: T*, in x0 / out solution.
template <typename T>
int pcg_core(gp_problem* P, const float* E, const double* f, T* u, double tol, int max_iter,
             bool use_x0, gp_pcg_stats* st, hipStream_t s) {
  const Grid& g = P->g;
  const size_t nd = (size_t)g.ndof;
  T* base = reinterpret_cast<T*>(std::is_same<T, float>::value ? P->ws32 : P->ws64);
  T* r = base;
  T* z = base + nd;
  T* p = base + 2 * nd;
  T* Ap = base + 3 * nd;
  T* minv = base + 4 * nd;
  const bool modal = P->modal_ok && (std::is_same<T, float>::value ? P->use_modal32 : P->use_modal64);
  const KuTile tile = std::is_same<T, float>::value ? P->tile32 : P->tile64;
  const void* kc = std::is_same<T, float>::value ? (const void*)&P->mc32 : (const void*)&P->mc64;
  const int prec = std::is_same<T, float>::value ? GP_F32 : GP_F64;
  const T* ke = reinterpret_cast<const T*>(std::is_same<T, float>::value ? (const void*)P->ke32
                                                                          : (const void*)P->ke64);
  DevScalars* sc = &P->sc[0];
  // single-kernel (Chronopoulos-Gear) iteration buffers on the modal path:
  // r, w, s ping-pong by parity; p in place
  T* rb[2] = {base, base + nd};
  T* wb[2] = {base + 2 * nd, base + 3 * nd};
  T* sb[2] = {base + 5 * nd, base + 6 * nd};
  T* pc = base + 7 * nd;
  // the float sweep reads M^-1 per node; FP64 and the gather path per DOF
  if constexpr (std::is_same<T, float>::value) {
    if (modal)
      launch_diag_node(g, P->ke32, E, minv, s);
    else
      launch_diag_minv<T>(g, ke, E, P->fixmask, minv, true, s);
  } else {
    launch_diag_minv<T>(g, ke, E, P->fixmask, minv, true, s);
  }
  if (use_x0) {
    launch_mask_fixed<T>(g, u, P->fixmask, s);
    ku_dispatch<T>(P, E, u, modal ? wb[0] : Ap, true, false, s);
  }
  if (modal)
    launch_cg_init<T>(g, f, wb[0], use_x0, u, rb[1], sb[0], wb[1], pc, minv, P->fixmask,
                      P->parts, sc, tol, max_iter, s);
  else
    launch_pcg_init<T>(g, f, Ap, use_x0, u, r, z, p, minv, P->parts, sc, tol, max_iter, s);
  LAUNCHCHK();

  const int hint_slot = std::is_same<T, float>::value ? 0 : 1;
  if (modal) {
    // x and p updated two iterations at a time (bitwise the every-iteration
    // x), then the pending update of an odd last iteration
    const int rc = pcg_drive(
        P, hint_slot, max_iter, true,
        [&](int k) {
          const int c = k & 1, o = c ^ 1;
          launch_cg_modal(g, tile, kc, prec, E, rb[o], wb[o], sb[c], minv, rb[c], wb[c], sb[o], u,
                          pc, P->fixmask, sc, P->parts, k, s, true);
        },
        st, s);
    if (rc) return rc;
    launch_cg_xp_fix<T>(g, sc, rb[0], minv, u, pc, s);
    LAUNCHCHK();
    return GP_OK;
  }
  return pcg_drive(
      P, hint_slot, max_iter, false,
      [&](int k) {
        const int nA = ku_dispatch<T>(P, E, p, Ap, true, true, s);
        launch_pcg_b<T>(g, k, nA, u, r, z, p, Ap, minv, P->parts, sc, s);
        launch_pcg_c<T>(g, k, p, z, P->parts, sc, s);
      },
      st, s);
}

// Inner solve of FP64 refinement with the iterate d and the search direction
// held in 16 bits (S = bf16 or _Float16): K d = rhs from d = 0, float
// arithmetic and float r, w, s, M^-1, FP64 dot products. Modal operator only.
// The eight element coefficients stay in float: rounding them one by one
// breaks the relations that give the element its rotational null space, and
// the perturbed operator is then indefinite on soft regions.
template <typename S>
int pcg_core_lp(gp_problem* P, const float* E, const double* rhs, S* d, double tol, int max_iter,
                gp_pcg_stats* st, hipStream_t s) {
  const Grid& g = P->g;
  const size_t nd = (size_t)g.ndof;
  // ws16: 7 float vectors (M^-1, r, w, s ping-pong), then p and d in S
  float* minv = reinterpret_cast<float*>(P->ws16);
  float* rb[2] = {minv + nd, minv + 2 * nd};
  float* wb[2] = {minv + 3 * nd, minv + 4 * nd};
  float* sb[2] = {minv + 5 * nd, minv + 6 * nd};
  S* pc = reinterpret_cast<S*>(minv + 7 * nd);
  DevScalars* sc = &P->sc[0];
  const ModalCoef<float> kc = P->mc32;
  const int prec = std::is_same<S, bf16>::value ? GP_BF16 : GP_F16;
  const KuTile tile = P->tile32.tj ? P->tile32 : KuTile{8, 32, 2};
  launch_diag_node(g, P->ke32, E, minv, s);
  launch_cg_init<float, S>(g, rhs, nullptr, false, d, rb[1], sb[0], wb[1], pc, minv, P->fixmask,
                           P->parts, sc, tol, max_iter, s);
  LAUNCHCHK();
  // x and p rounded to 16 bits once per pair of iterations
  const int rc = pcg_drive(
      P, 2, max_iter, true,
      [&](int k) {
        const int c = k & 1, o = c ^ 1;
        launch_cg_modal(g, tile, &kc, prec, E, rb[o], wb[o], sb[c], minv, rb[c], wb[c], sb[o], d,
                        pc, P->fixmask, sc, P->parts, k, s, true);
      },
      st, s);
  if (rc) return rc;
  launch_cg_xp_fix<float, S>(g, sc, rb[0], minv, d, pc, s);
  LAUNCHCHK();
  return GP_OK;
}

// FP64 iterative refinement: r = f - K u in FP64, correction solved in FP32.
int pcg_refine(gp_problem* P, const float* E, const double* f, double* u, const gp_pcg_opts* o,
               gp_pcg_stats* st, hipStream_t s) {
  const Grid& g = P->g;
  int rc = ensure_refine_ws(P);
  if (rc) return rc;
  const bool lp = o->precision == GP_F16 || o->precision == GP_BF16;
  if (lp && !P->modal_ok)
    return fail(GP_ERR_ARG, "16-bit PCG storage needs an isotropic hex8 element table");
  rc = lp ? ensure_lp_ws(P) : ensure_pcg_ws(P, GP_F32);
  if (rc) return rc;
  DevScalars* sc2 = &P->sc[2];
  if (!o->use_x0) HIPCHK(hipMemsetAsync(u, 0, g.ndof * sizeof(double), s));
  launch_mask_fixed<double>(g, u, P->fixmask, s);
  st->status = GP_OK;
  const int max_outer = o->max_outer > 0 ? o->max_outer : 50;
  double rr = 0.0, bb = 0.0;
  bool brk = false;
  // The same sequence as a scheduler job with refinement (PH_RES, PH_RESFIN,
  // inner solve, PH_CORR): every sweep forms the FP64 residual with both
  // sums, then stops on convergence, a broken-down inner solve or the cap.
  for (int k = 0;; ++k) {
    launch_res64(g, P->ke64, E, u, f, P->fixmask, P->r64, P->parts, sc2, s);
    LAUNCHCHK();
    rc = sync_read(P, 2, s);
    if (rc) return rc;
    rr = P->h_sc[2].ir_rr;
    bb = P->h_sc[2].ir_bb;
    st->outer = k;
    if (bb == 0.0) {
      HIPCHK(hipMemsetAsync(u, 0, g.ndof * sizeof(double), s));
      st->relres = 0.0;
      return sync_read(P, 2, s);
    }
    if (rr <= o->tol * o->tol * bb || !std::isfinite(rr)) break;
    if (brk) {
      st->status = GP_BREAKDOWN;
      break;
    }
    if (k >= max_outer) {
      st->status = GP_NOT_CONVERGED;
      break;
    }
    gp_pcg_stats in{};
    const double itol = o->inner_tol > 0 ? o->inner_tol : 1e-3;
    // 16-bit inner solves see the residual rescaled to the norm of f, so the
    // correction keeps the magnitude of the solution instead of shrinking
    // with the residual into the 16-bit underflow range.
    const double up = std::sqrt(bb / rr), down = std::sqrt(rr / bb);
    if (lp) launch_scale64(g, P->r64, up, s);
    if (o->precision == GP_BF16) {
      bf16* d = lp_correction<bf16>(P);
      rc = pcg_core_lp<bf16>(P, E, P->r64, d, itol, o->max_iter, &in, s);
      if (!rc) launch_add_correction_lp<bf16>(g, u, d, down, s);
    } else if (o->precision == GP_F16) {
      _Float16* d = lp_correction<_Float16>(P);
      rc = pcg_core_lp<_Float16>(P, E, P->r64, d, itol, o->max_iter, &in, s);
      if (!rc) launch_add_correction_lp<_Float16>(g, u, d, down, s);
    } else {
      rc = pcg_core<float>(P, E, P->r64, P->d32, itol, o->max_iter, false, &in, s);
      if (!rc) launch_add_correction<float>(g, u, P->d32, s);
    }
    if (rc) return rc;
    st->iters += in.iters;
    // a broken-down inner solve still adds its correction; one more
    // residual is formed for the report
    if (in.status == GP_BREAKDOWN) brk = true;
  }
  st->relres = std::sqrt(rr / bb);
  if (!std::isfinite(rr)) st->status = GP_ERR_NAN;
  return GP_OK;
}

int run_oc(gp_problem* P, const double* x, const double* dc, double volfrac, double* xnew,
           double* xphys, hipStream_t s) {
  const Grid& g = P->g;
  DevScalars* sc = &P->sc[1];
  launch_oc_init(g, dc, P->spec.oc_variant, volfrac, P->spec.move, P->parts, sc, s);
  const int max_steps = P->spec.oc_variant == GP_OC_GPU ? 30 : 50;
  for (int i = 0; i < oc_launches(max_steps); ++i)
    launch_oc_step(g, i, x, dc, P->w, P->pres, P->n_pres, P->parts, sc, s);
  launch_oc_final(g, x, dc, P->pres, sc, xnew, s);
  launch_filter(g, P->st, P->Hs, xnew, xphys, 0, s);
  launch_set_preserved(g, P->pres, xphys, s);
  LAUNCHCHK();
  return GP_OK;
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

const char* gp_last_error(void) { return g_err.c_str(); }

int gp_hex8_ke(double nu, double* ke_out) {
  if (!ke_out) return fail(GP_ERR_ARG, "ke_out is NULL");
  if (!(nu > -1.0 && nu < 0.5)) return fail(GP_ERR_ARG, "nu out of range");
  hex8_ke(nu, ke_out);
  return GP_OK;
}
int gp_version(void) { return 100; }

int gp_set_allocator(gp_alloc_fn alloc_fn, gp_free_fn free_fn, void* ctx) {
  if ((alloc_fn == nullptr) != (free_fn == nullptr))
    return fail(GP_ERR_ARG, "give both allocator functions, or neither");
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  if (g_live.load() != 0)
    return fail(GP_ERR_STATE, "%lld library allocations are alive; the allocator changes only "
                "while none is", g_live.load());
  g_alloc_fn.store(alloc_fn);
  g_free_fn.store(free_fn);
  g_alloc_ctx.store(ctx);
  return GP_OK;
}

int gp_recycle_bytes(const gp_spec* spec, int32_t recycle, size_t* bytes) {
  int rc = validate_spec(spec);
  if (rc) return rc;
  if (!bytes) return fail(GP_ERR_ARG, "bytes is NULL");
  if (recycle < 0 || recycle > kRecycleMax)
    return fail(GP_ERR_ARG, "recycle must be 0..%d", kRecycleMax);
  *bytes = recycle ? recycle_ws_bytes((size_t)make_grid(spec).ndof, recycle) : 0;
  return GP_OK;
}

int gp_footprint_bytes(const gp_spec* spec, int32_t precision, size_t* bytes) {
  int rc = validate_spec(spec);
  if (rc) return rc;
  if (!bytes) return fail(GP_ERR_ARG, "bytes is NULL");
  const Grid g = make_grid(spec);
  const size_t nd = (size_t)g.ndof, ne = (size_t)g.ne, nn = (size_t)g.nn;
  size_t b = 576 * (8 + 4) + sizeof(Stencil) + 2 * ne * 8 /*Hs,w*/ + nn /*fixmask*/ + ne /*pres*/ +
             nd * 8 /*f*/ + 3 * sizeof(DevScalars) + kRegions * partial_cap(g) * 8;
  b += ne * (4 + 4 * 8) + 8;                                   // SIMP loop workspace
  if (precision == GP_F64) b += 8 * nd * 8;
  else if (precision == GP_F16 || precision == GP_BF16) b += nd * (28 + 4) + nd * (8 + 8 + 4);
  else b += 8 * nd * 4 + nd * (8 + 8 + 4);                     // FP32 PCG + refinement
  *bytes = b;
  return GP_OK;
}

int gp_problem_create(const gp_spec* spec, gp_problem** out) {
  int rc = validate_spec(spec);
  if (rc) return rc;
  if (!out) return fail(GP_ERR_ARG, "out is NULL");
  *out = nullptr;
  gp_problem* P = new gp_problem();
  P->spec = *spec;
  P->spec.ke = nullptr;
  P->g = make_grid(spec);
  HIPCHK(hipGetDevice(&P->device));
  if (spec->ke_kind == GP_KE_HEX8)
    hex8_ke(spec->nu, P->ke64_h);
  else
    std::memcpy(P->ke64_h, spec->ke, sizeof(double) * 576);
  float ke32[576];
  for (int i = 0; i < 576; ++i) ke32[i] = (float)P->ke64_h[i];
  make_stencil(spec->rmin, spec->filter_weights == GP_FILTER_F32, &P->st_h);
  const Grid& g = P->g;
  auto cleanup_fail = [&](int code) {
    gp_problem_destroy(P);
    return code;
  };
#define CALLOC(ptr, bytes)                                                \
  do {                                                                    \
    int _r = alloc(reinterpret_cast<void**>(&(ptr)), (bytes));            \
    if (_r) return cleanup_fail(_r);                                      \
  } while (0)
  const size_t cap = partial_cap(g);
  P->n_partials = kRegions * cap;
  size_t arena_bytes = 0;
  auto carve = [&](size_t bytes) {
    const size_t at = arena_bytes;
    arena_bytes += (bytes + 255) & ~size_t(255);
    return at;
  };
  const size_t o_ke64 = carve(576 * sizeof(double)), o_ke32 = carve(576 * sizeof(float));
  const size_t o_st = carve(sizeof(Stencil)), o_hs = carve(g.ne * sizeof(double));
  const size_t o_w = carve(g.ne * sizeof(double)), o_fix = carve(g.nn);
  const size_t o_f = carve(g.ndof * sizeof(double)), o_sc = carve(3 * sizeof(DevScalars));
  const size_t o_part = carve(P->n_partials * sizeof(double));
  CALLOC(P->arena, arena_bytes);
  {
    char* a0 = static_cast<char*>(P->arena);
    P->ke64 = reinterpret_cast<double*>(a0 + o_ke64);
    P->ke32 = reinterpret_cast<float*>(a0 + o_ke32);
    P->st = reinterpret_cast<Stencil*>(a0 + o_st);
    P->Hs = reinterpret_cast<double*>(a0 + o_hs);
    P->w = reinterpret_cast<double*>(a0 + o_w);
    P->fixmask = reinterpret_cast<uint8_t*>(a0 + o_fix);
    P->f = reinterpret_cast<double*>(a0 + o_f);
    P->sc = reinterpret_cast<DevScalars*>(a0 + o_sc);
    P->partials = reinterpret_cast<double*>(a0 + o_part);
  }
  {
    double* b = P->partials;
    P->parts.cap = (int)cap;
    P->parts.A = b;
    P->parts.RZ[0] = b + cap;
    P->parts.RZ[1] = b + 2 * cap;
    P->parts.RR = b + 3 * cap;
    P->parts.BB = b + 4 * cap;
    P->parts.M = b + 5 * cap;
    P->parts.S = b + 6 * cap;
    for (int q = 0; q < 2; ++q) {
      P->parts.G[q] = b + (7 + q) * cap;
      P->parts.D[q] = b + (9 + q) * cap;
      P->parts.R[q] = b + (11 + q) * cap;
      P->parts.O[q] = b + (13 + q * kOcCand) * cap;
    }
  }
  {
    ModalCoef<double> mc;
    P->modal_ok = modal_coefs(P->ke64_h, &mc);
    // The float sweeps keep one Jacobi value per node, built from each
    // corner's x-component diagonal (body_diag invert 2). That equals the
    // per-DOF M^-1 only when the three float diagonal entries of every corner
    // are equal, as they are for the isotropic hex8 element; a custom table
    // that matches the modal form only to rounding may not satisfy it, and
    // then keeps the gather kernels with M^-1 per DOF.
    for (int c = 0; c < 8 && P->modal_ok; ++c) {
      const float d0 = ke32[(3 * c) * 24 + 3 * c];
      if (ke32[(3 * c + 1) * 24 + 3 * c + 1] != d0 || ke32[(3 * c + 2) * 24 + 3 * c + 2] != d0)
        P->modal_ok = false;
    }
    P->mc64 = mc;
    P->mc32 = {(float)mc.An, (float)mc.Ln, (float)mc.S, (float)mc.B1,
               (float)mc.B2, (float)mc.T1, (float)mc.T2, (float)mc.X};
    // default sweep depth: enough workgroups to cover the device a few times
    const int ntj = (g.ny + 1 + 6) / 7, ntk = (g.nz + 1 + 30) / 31;
    int ni = (ntj * ntk * (g.nx + 1)) / 1024;
    P->tile32.ni = std::max(2, std::min(16, ni));
    P->tile64.ni = P->tile32.ni;
  }
#undef CALLOC
  if (hipHostMalloc(reinterpret_cast<void**>(&P->h_sc), 3 * sizeof(DevScalars)) != hipSuccess)
    return cleanup_fail(fail(GP_ERR_OOM, "hipHostMalloc failed"));
  if (hipEventCreateWithFlags(&P->ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&P->ev[1], hipEventDisableTiming) != hipSuccess)
    return cleanup_fail(fail(GP_ERR_HIP, "hipEventCreate failed"));
  hipError_t e = hipSuccess;
  // zero the whole set-up allocation (scalars, supports, load and partials
  // start at 0), then the tables
  e = e != hipSuccess ? e : hipMemset(P->arena, 0, arena_bytes);
  e = e != hipSuccess ? e : hipMemcpy(P->ke64, P->ke64_h, sizeof(double) * 576, hipMemcpyHostToDevice);
  e = e != hipSuccess ? e : hipMemcpy(P->ke32, ke32, sizeof(float) * 576, hipMemcpyHostToDevice);
  e = e != hipSuccess ? e : hipMemcpy(P->st, &P->st_h, sizeof(Stencil), hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup_fail(fail(GP_ERR_HIP, "init copy: %s", hipGetErrorString(e)));
  launch_hs(g, P->st, P->Hs, nullptr);
  launch_oc_weights(g, P->st, P->Hs, nullptr, P->w, nullptr);
  e = hipGetLastError();
  // the set-up ran on the null stream: wait for that stream only
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) return cleanup_fail(fail(GP_ERR_HIP, "init kernels: %s", hipGetErrorString(e)));
  *out = P;
  return GP_OK;
}

int gp_problem_destroy(gp_problem* P) {
  if (!P) return GP_OK;
  hipSetDevice(P->device);
  // one wait for the work that may still use these buffers, then the frees
  void* dev[] = {P->arena, P->pres, P->ws32, P->ws64, P->ws16, P->r64, P->y64, P->d32, P->E,
                 P->dc, P->dcf, P->xnew, P->xphys_new, P->ce_sum, P->dws_mem, P->rc_mem};
  sync_for_free();
  for (void* p : dev)
    if (p) dev_free(p, true);
  if (P->h_sc) hipHostFree(P->h_sc);
  if (P->dws_info_h) hipHostFree(P->dws_info_h);
  for (auto& ev : P->ev)
    if (ev) hipEventDestroy(ev);
  for (auto& pr : P->prof_pending) {
    hipEventDestroy(pr.a);
    hipEventDestroy(pr.b);
  }
  for (auto e : P->prof_pool) hipEventDestroy(e);
  delete P;
  return GP_OK;
}

int gp_problem_set_bcs(gp_problem* P, const int64_t* fixed, int64_t n_fixed, const double* f_host) {
  if (!P) return fail(GP_ERR_ARG, "problem is NULL");
  if (n_fixed < 0 || (n_fixed > 0 && !fixed)) return fail(GP_ERR_ARG, "bad fixed list");
  if (!f_host) return fail(GP_ERR_ARG, "f is NULL");
  HIPCHK(hipSetDevice(P->device));
  const Grid& g = P->g;
  std::vector<uint8_t> m(g.nn, 0);
  for (int64_t q = 0; q < n_fixed; ++q) {
    const int64_t d = fixed[q];
    if (d < 0 || d >= g.ndof) return fail(GP_ERR_ARG, "fixed dof %lld out of range", (long long)d);
    m[d / 3] |= (uint8_t)(1u << (d % 3));
  }
  HIPCHK(hipMemcpy(P->fixmask, m.data(), g.nn, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P->f, f_host, g.ndof * sizeof(double), hipMemcpyHostToDevice));
  P->fixmask_h = std::move(m);
  P->dws_valid = false;  // the free numbering follows the supports
  P->has_bcs = true;
  return GP_OK;
}

int gp_problem_set_preserved(gp_problem* P, const uint8_t* mask) {
  if (!P) return fail(GP_ERR_ARG, "problem is NULL");
  HIPCHK(hipSetDevice(P->device));
  const Grid& g = P->g;
  if (mask == nullptr) {
    if (P->pres) HIPCHK(dev_free(P->pres));
    P->pres = nullptr;
    P->n_pres = 0;
  } else {
    if (!P->pres) {
      int rc = alloc(reinterpret_cast<void**>(&P->pres), g.ne);
      if (rc) return rc;
    }
    std::vector<uint8_t> m(mask, mask + g.ne);
    long long c = 0;
    for (auto& v : m) {
      v = v ? 1 : 0;
      c += v;
    }
    P->n_pres = (double)c;
    HIPCHK(hipMemcpy(P->pres, m.data(), g.ne, hipMemcpyHostToDevice));
  }
  launch_oc_weights(g, P->st, P->Hs, P->pres, P->w, nullptr);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(nullptr));  // the weights were formed on the null stream
  return GP_OK;
}

int gp_problem_set_profiling(gp_problem* P, int32_t every) {
  if (!P || every < 0) return fail(GP_ERR_ARG, "bad argument");
  P->prof_every = every;
  P->prof_counter = 0;
  return GP_OK;
}

int gp_problem_get_profile(gp_problem* P, double* ku_ms_sum, int64_t* ku_samples, int32_t reset) {
  if (!P || !ku_ms_sum || !ku_samples) return fail(GP_ERR_ARG, "NULL argument");
  *ku_ms_sum = P->prof_ms;
  *ku_samples = P->prof_n;
  if (reset) {
    P->prof_ms = 0.0;
    P->prof_n = 0;
  }
  return GP_OK;
}

int gp_problem_set_ku_tile(gp_problem* P, int32_t precision, int32_t tj, int32_t tk, int32_t ni) {
  if (!P) return fail(GP_ERR_ARG, "problem is NULL");
  if (precision != GP_F32 && precision != GP_F64) return fail(GP_ERR_ARG, "bad precision");
  bool& use = precision == GP_F32 ? P->use_modal32 : P->use_modal64;
  KuTile& t = precision == GP_F32 ? P->tile32 : P->tile64;
  if (tj == 0) {
    use = false;  // dense gather kernel
    return GP_OK;
  }
  bool known = false;
  for (const KuTile& k : kTiles) known = known || (k.tj == tj && k.tk == tk);
  if (!known) return fail(GP_ERR_ARG, "unsupported K.u tile %dx%d", tj, tk);
  if (ni < 1 || ni > 256) return fail(GP_ERR_ARG, "ni must be in [1, 256]");
  t = {tj, tk, ni};
  use = true;
  return GP_OK;
}

int gp_problem_get_ku_tile(gp_problem* P, int32_t precision, int32_t* tj, int32_t* tk, int32_t* ni,
                           int32_t* modal) {
  if (!P || !tj || !tk || !ni || !modal) return fail(GP_ERR_ARG, "NULL argument");
  const bool f32 = precision == GP_F32;
  const KuTile& t = f32 ? P->tile32 : P->tile64;
  const bool use = P->modal_ok && (f32 ? P->use_modal32 : P->use_modal64);
  *tj = use ? t.tj : 0;
  *tk = use ? t.tk : 0;
  *ni = use ? t.ni : 0;
  *modal = P->modal_ok ? 1 : 0;
  return GP_OK;
}

int gp_problem_load(gp_problem* P, double* f_dev, void* stream) {
  if (!P || !f_dev) return fail(GP_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(P->device));
  HIPCHK(hipMemcpyAsync(f_dev, P->f, P->g.ndof * sizeof(double), hipMemcpyDeviceToDevice, S(stream)));
  return GP_OK;
}

int gp_stiffness(gp_problem* P, const double* xPhys, float* E, void* stream) {
  if (!P || !xPhys || !E) return fail(GP_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(P->device));
  launch_stiffness(P->g, xPhys, E, P->spec.law, P->spec.penal, P->spec.E0, P->spec.Emin, S(stream));
  LAUNCHCHK();
  return GP_OK;
}

int gp_ku(gp_problem* P, const float* E, const void* u, void* y, int32_t dtype, int32_t mask,
          void* stream) {
  if (!P || !E || !u || !y) return fail(GP_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(P->device));
  if (dtype == GP_F32)
    ku_dispatch<float>(P, E, (const float*)u, (float*)y, mask != 0, false, S(stream));
  else if (dtype == GP_F64)
    ku_dispatch<double>(P, E, (const double*)u, (double*)y, mask != 0, false, S(stream));
  else
    return fail(GP_ERR_ARG, "gp_ku: dtype must be GP_F32 or GP_F64");
  LAUNCHCHK();
  return GP_OK;
}

int gp_diag(gp_problem* P, const float* E, void* diag, int32_t dtype, void* stream) {
  if (!P || !E || !diag) return fail(GP_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(P->device));
  if (dtype == GP_F32)
    launch_diag_minv<float>(P->g, P->ke32, E, P->fixmask, (float*)diag, false, S(stream));
  else if (dtype == GP_F64)
    launch_diag_minv<double>(P->g, P->ke64, E, P->fixmask, (double*)diag, false, S(stream));
  else
    return fail(GP_ERR_ARG, "gp_diag: dtype must be GP_F32 or GP_F64");
  LAUNCHCHK();
  return GP_OK;
}

int gp_pcg(gp_problem* P, const float* E, const double* f, void* u, const gp_pcg_opts* o,
           gp_pcg_stats* st, void* stream) {
  if (!P || !E || !f || !u || !o || !st) return fail(GP_ERR_ARG, "NULL argument");
  if (o->max_iter < 1 || !(o->tol >= 0)) return fail(GP_ERR_ARG, "bad max_iter/tol");
  HIPCHK(hipSetDevice(P->device));
  std::memset(st, 0, sizeof(*st));
  hipStream_t s = S(stream);
  int rc;
  const bool lp = o->precision == GP_F16 || o->precision == GP_BF16;
  if (lp && !o->refine)
    return fail(GP_ERR_ARG, "16-bit PCG storage runs inside FP64 refinement (refine = 1)");
  if (o->refine) {
    rc = pcg_refine(P, E, f, (double*)u, o, st, s);
  } else if (o->precision == GP_F64) {
    rc = ensure_pcg_ws(P, GP_F64);
    if (!rc) rc = pcg_core<double>(P, E, f, (double*)u, o->tol, o->max_iter, o->use_x0 != 0, st, s);
  } else if (o->precision == GP_F32) {
    rc = ensure_pcg_ws(P, GP_F32);
    if (!rc) rc = pcg_core<float>(P, E, f, (float*)u, o->tol, o->max_iter, o->use_x0 != 0, st, s);
  } else {
    return fail(GP_ERR_ARG, "bad precision %d", o->precision);
  }
  if (rc) return rc;
  if (st->status == GP_BREAKDOWN) fail(GP_BREAKDOWN, "PCG breakdown (p^T A p <= 0)");
  return GP_OK;
}

int gp_direct(gp_problem* P, const double* xPhys, const double* f, double* u, gp_pcg_stats* st,
              void* stream) {
  if (!P || !xPhys || !f || !u || !st) return fail(GP_ERR_ARG, "NULL argument");
  if (!P->has_bcs) return fail(GP_ERR_STATE, "boundary conditions not set");
  HIPCHK(hipSetDevice(P->device));
  return direct_solve(P, xPhys, f, u, st, S(stream));
}

int gp_sens(gp_problem* P, const double* xPhys, const void* u, int32_t dtype, double* dc,
            double* ce_sum, void* stream) {
  if (!P || !xPhys || !u || !dc) return fail(GP_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(P->device));
  const gp_spec& sp = P->spec;
  if (dtype == GP_F32)
    launch_sens<float>(P->g, P->ke64, P->modal_ok ? &P->mc64 : nullptr, xPhys, (const float*)u, sp.law, sp.penal, sp.E0, sp.Emin, dc,
                       ce_sum, P->parts, &P->sc[1], S(stream));
  else if (dtype == GP_F64)
    launch_sens<double>(P->g, P->ke64, P->modal_ok ? &P->mc64 : nullptr, xPhys, (const double*)u, sp.law, sp.penal, sp.E0, sp.Emin,
                        dc, ce_sum, P->parts, &P->sc[1], S(stream));
  else
    return fail(GP_ERR_ARG, "gp_sens: dtype must be GP_F32 or GP_F64");
  LAUNCHCHK();
  return GP_OK;
}

int gp_filter(gp_problem* P, const double* in, double* out, int32_t mode, void* stream) {
  if (!P || !in || !out) return fail(GP_ERR_ARG, "NULL argument");
  if (mode != 0 && mode != 1) return fail(GP_ERR_ARG, "mode must be 0 or 1");
  HIPCHK(hipSetDevice(P->device));
  launch_filter(P->g, P->st, P->Hs, in, out, mode, S(stream));
  LAUNCHCHK();
  return GP_OK;
}

int gp_oc(gp_problem* P, const double* x, const double* dc, double volfrac, double* xnew,
          double* xphys, gp_oc_stats* stats, void* stream) {
  if (!P || !x || !dc || !xnew || !xphys) return fail(GP_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(P->device));
  int rc = run_oc(P, x, dc, volfrac, xnew, xphys, S(stream));
  if (rc) return rc;
  rc = sync_read(P, 1, S(stream));
  if (rc) return rc;
  if (stats) {
    stats->lmid = P->h_sc[1].lmid;
    stats->steps = P->h_sc[1].oc_steps;
  }
  return GP_OK;
}

int gp_compliance(gp_problem* P, const void* u, int32_t dtype, double* out, void* stream) {
  if (!P || !u || !out) return fail(GP_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(P->device));
  if (dtype == GP_F32)
    launch_dot<double, float>(P->f, (const float*)u, P->g.ndof, P->parts, &P->sc[1], S(stream));
  else if (dtype == GP_F64)
    launch_dot<double, double>(P->f, (const double*)u, P->g.ndof, P->parts, &P->sc[1], S(stream));
  else
    return fail(GP_ERR_ARG, "gp_compliance: dtype must be GP_F32 or GP_F64");
  LAUNCHCHK();
  int rc = sync_read(P, 1, S(stream));
  if (rc) return rc;
  *out = P->h_sc[1].sum;
  return GP_OK;
}

int gp_simp_run(gp_problem* P, const gp_simp_opts* o, double* x, double* xPhys, double* u,
                gp_simp_stats* st, void* stream) {
  if (!P || !o || !x || !xPhys || !u || !st) return fail(GP_ERR_ARG, "NULL argument");
  if (!P->has_bcs) return fail(GP_ERR_STATE, "boundary conditions not set");
  HIPCHK(hipSetDevice(P->device));
  auto t0 = std::chrono::steady_clock::now();
  std::memset(st, 0, sizeof(*st));
  hipStream_t s = S(stream);
  int rc = ensure_simp_ws(P);
  if (rc) return rc;
  const Grid& g = P->g;
  const gp_spec& sp = P->spec;
  const bool direct = o->direct != 0;
  // u is the float64 state whenever the solves are FP64 (refined, FP64 PCG
  // or direct)
  const bool refine = direct || o->pcg.refine != 0 || o->pcg.precision == GP_F64 ||
                      o->pcg.precision == GP_F16 || o->pcg.precision == GP_BF16;
  const int recycle = o->recycle;
  if (recycle < 0 || recycle > kRecycleMax)
    return fail(GP_ERR_ARG, "recycle must be 0..%d", kRecycleMax);
  if (recycle > 0 && direct)
    return fail(GP_ERR_ARG, "recycled starts apply to PCG solves, not the direct solve");
  if (recycle > 0) {
    rc = ensure_recycle_ws(P, recycle);
    if (rc) return rc;
    P->rc_n = P->rc_head = 0;  // a run's history starts empty
  }
  // u is float64 state; plain FP32 PCG works on a float copy
  float* u32 = nullptr;
  if (!refine) {
    rc = ensure_refine_ws(P);  // d32 doubles as the FP32 iterate
    if (rc) return rc;
    u32 = P->d32;
    launch_cast_f64_to_f32(u, u32, g.ndof, s);
  }
  // Recycled starts: the Galerkin combination of the stored solutions
  // (oldest first) replaces the last solution as the start of a warm solve.
  // The ring holds float copies; FP64-state solves (refined, 16-bit inner,
  // FP64 PCG) form the start in the FP32 correction buffer and widen it.
  auto recycled_start = [&]() -> int {
    const float* V[kRecycleMax];
    const size_t nd = (size_t)g.ndof;
    for (int i = 0; i < P->rc_n; ++i)
      V[i] = P->rc_v + (size_t)((P->rc_head - P->rc_n + i + recycle) % recycle) * nd;
    if (!refine) return recycle_guess(P, P->E, V, P->rc_n, u32, s);
    int r = ensure_refine_ws(P);
    if (!r) r = recycle_guess(P, P->E, V, P->rc_n, P->d32, s);
    if (!r) launch_cast_f32_to_f64(P->d32, u, g.ndof, s);
    return r;
  };
  auto store_solution = [&]() -> int {
    float* slot = P->rc_v + (size_t)P->rc_head * g.ndof;
    if (refine)
      launch_cast_f64_to_f32(u, slot, g.ndof, s);
    else
      HIPCHK(hipMemcpyAsync(slot, u32, g.ndof * sizeof(float), hipMemcpyDeviceToDevice, s));
    P->rc_head = (P->rc_head + 1) % recycle;
    P->rc_n = std::min(P->rc_n + 1, recycle);
    return GP_OK;
  };
  auto solve = [&](double tol, int max_iter, bool warm, gp_pcg_stats* ps) -> int {
    gp_pcg_opts po = o->pcg;
    po.tol = tol;
    po.max_iter = max_iter;
    po.use_x0 = warm ? 1 : 0;
    int r = GP_OK;
    std::memset(ps, 0, sizeof(*ps));
    if (!direct && recycle > 0 && warm && P->rc_n > 0) r = recycled_start();
    if (r) return r;
    if (direct) {
      r = direct_solve(P, xPhys, P->f, u, ps, s);
    } else if (refine) {
      if (po.precision == GP_F64 && !po.refine) {
        r = ensure_pcg_ws(P, GP_F64);
        if (!r) r = pcg_core<double>(P, P->E, P->f, u, tol, max_iter, warm, ps, s);
      } else {
        r = pcg_refine(P, P->E, P->f, u, &po, ps, s);
      }
    } else {
      r = ensure_pcg_ws(P, GP_F32);
      if (!r) r = pcg_core<float>(P, P->E, P->f, u32, tol, max_iter, warm, ps, s);
    }
    if (!r && recycle > 0) r = store_solution();
    if (r) return r;
    st->pcg_iters_total += ps->iters;
    if (ps->status > st->status) st->status = ps->status;
    return GP_OK;
  };
  auto compliance = [&](double* c) -> int {
    if (refine)
      launch_dot<double, double>(P->f, u, g.ndof, P->parts, &P->sc[1], s);
    else
      launch_dot<double, float>(P->f, u32, g.ndof, P->parts, &P->sc[1], s);
    int r = sync_read(P, 1, s);
    if (r) return r;
    *c = P->h_sc[1].sum;
    return GP_OK;
  };

  double prev_c = INFINITY, curr_c = 0.0;
  int small_changes = 0;
  bool have_u = o->warm_start != 0;  // caller's u is a valid guess only when asked
  for (int it = 0; it < o->n_iters; ++it) {
    double tol = o->pcg.tol;
    const int max_iter = o->pcg.max_iter;
    // simp_solver_gpu.py:97-103 tolerance schedule. Its iteration caps
    // (200/400/800) belong to the reference's incomplete-Cholesky PCG; a
    // Jacobi-PCG needs more iterations for the same tolerance, so the caller's
    // cap applies instead.
    if (o->tol_schedule) tol = it < 3 ? 1e-4 : (it < 8 ? 1e-5 : 1e-6);
    launch_stiffness(g, xPhys, P->E, sp.law, sp.penal, sp.E0, sp.Emin, s);
    gp_pcg_stats ps;
    rc = solve(tol, max_iter, o->warm_start && have_u, &ps);
    if (rc) return rc;
    have_u = true;
    launch_reset_flags(&P->sc[1], s);
    if (refine)
      launch_sens<double>(g, P->ke64, P->modal_ok ? &P->mc64 : nullptr, xPhys, u, sp.law, sp.penal, sp.E0, sp.Emin, P->dc, nullptr,
                          P->parts, &P->sc[1], s);
    else
      launch_sens<float>(g, P->ke64, P->modal_ok ? &P->mc64 : nullptr, xPhys, u32, sp.law, sp.penal, sp.E0, sp.Emin, P->dc, nullptr,
                         P->parts, &P->sc[1], s);
    launch_filter(g, P->st, P->Hs, P->dc, P->dcf, 1, s);
    rc = run_oc(P, x, P->dcf, o->volfrac, P->xnew, P->xphys_new, s);
    if (rc) return rc;
    rc = sync_read(P, 1, s);
    if (rc) return rc;
    if (P->h_sc[1].nan_flag) {
      st->status = GP_ERR_NAN;
      break;  // reference aborts before the OC update is applied
    }
    st->bisect_steps_total += P->h_sc[1].oc_steps;
    HIPCHK(hipMemcpyAsync(x, P->xnew, g.ne * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(xPhys, P->xphys_new, g.ne * sizeof(double), hipMemcpyDeviceToDevice, s));
    st->iters_done = it + 1;
    if (o->compliance_each_iter) {
      // simp_solver_gpu.py:747-751 -- a fresh solve (zero start, tol 1e-6;
      // the caller's iteration cap as above); its u becomes the next warm
      // start (simp_solver_gpu.py:593,704)
      launch_stiffness(g, xPhys, P->E, sp.law, sp.penal, sp.E0, sp.Emin, s);
      rc = solve(1e-6 > o->pcg.tol ? o->pcg.tol : 1e-6, max_iter,
                 o->warm_compliance != 0 && o->warm_start != 0, &ps);
      if (rc) return rc;
      rc = compliance(&curr_c);
      if (rc) return rc;
      if (o->early_stop) {
        if (std::isfinite(prev_c)) {
          const double ch = std::fabs(prev_c - curr_c) / (prev_c + 1e-10);
          small_changes = ch < 0.01 ? small_changes + 1 : 0;
          if (small_changes >= 3) break;
        }
        prev_c = curr_c;
      }
    }
  }
  if (o->compliance_each_iter && st->iters_done > 0) {
    st->compliance = curr_c;
  } else {
    // simp_solver.py:78 -- a final solve at the returned density
    launch_stiffness(g, xPhys, P->E, sp.law, sp.penal, sp.E0, sp.Emin, s);
    gp_pcg_stats ps;
    rc = solve(o->pcg.tol, o->pcg.max_iter, o->warm_start != 0, &ps);
    if (rc) return rc;
    rc = compliance(&st->compliance);
    if (rc) return rc;
  }
  if (!refine) launch_cast_f32_to_f64(u32, u, g.ndof, s);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(s));
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return GP_OK;
}

int gp_recycle_guess(gp_problem* P, const float* E, const float* const* V, int32_t n, float* u0,
                     double* coef, void* stream) {
  if (!P || !E || !V || !u0) return fail(GP_ERR_ARG, "NULL argument");
  if (!P->has_bcs) return fail(GP_ERR_STATE, "boundary conditions not set");
  if (n < 1 || n > kRecycleMax) return fail(GP_ERR_ARG, "n must be 1..%d", kRecycleMax);
  for (int i = 0; i < n; ++i)
    if (!V[i]) return fail(GP_ERR_ARG, "V[%d] is NULL", i);
  HIPCHK(hipSetDevice(P->device));
  hipStream_t s = S(stream);
  int rc = recycle_guess(P, E, V, n, u0, s);
  if (rc) return rc;
  if (coef) {
    HIPCHK(hipMemcpyAsync(coef, P->rc_coef, (kRecycleMax + 1) * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return GP_OK;
}

// ------------------------------------------------------------ generic API
}  // extern "C"

namespace {

// Workspace of the CSR Jacobi-PCG: four vectors of n floats, the scalars
// and kVecBlocks partial sums.
struct CsrPcgWork {
  float *r, *z, *p, *Ap;
  DevScalars* sc;
  double* part;
};

// Jacobi-PCG on a CSR matrix with the reference's pcg_solve_cuda semantics
// (simp_pcg_cuda.cu:66-124: r starts at b, fixed DOFs zeroed in r and A p,
// ||r|| < tol ||b||). h: host copy of the scalars. Allocates nothing.
int csr_pcg_run(const int64_t* rp, const int64_t* ci, const float* v, const float* diagonal,
                const float* b, float* x, const int64_t* fixed, int64_t n_fixed, int64_t n,
                int32_t max_iter, double tol, const CsrPcgWork& w, DevScalars* h,
                gp_pcg_stats* st, hipStream_t s) {
  float *r = w.r, *z = w.z, *p = w.p, *Ap = w.Ap;
  auto dot = [&](const float* a, const float* c, double* out) -> int {
    Parts pp{};
    pp.S = w.part;
    launch_dot<float, float>(a, c, n, pp, w.sc, s);
    HIPCHK(hipMemcpyAsync(h, w.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *out = h->sum;
    return GP_OK;
  };
  // the scalars start at zero: a fill kernel, not a memset node
  launch_zero_scalars(w.sc, s);
  if (hipMemcpyAsync(r, b, n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return fail(GP_ERR_HIP, "init copy failed");
  launch_zero_list(r, fixed, n_fixed, s);
  launch_vec(z, r, diagonal, 0.f, n, 2, s);
  if (hipMemcpyAsync(p, z, n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return fail(GP_ERR_HIP, "init copy failed");
  int rc = GP_OK;
  double rz, bn2;
  if ((rc = dot(r, z, &rz)) || (rc = dot(b, b, &bn2))) return rc;
  const double bnorm = std::sqrt(bn2);
  int it = 0;
  double rr = 0;
  for (; it < max_iter; ++it) {
    launch_csr_spmv(rp, ci, v, p, Ap, n, s);
    launch_zero_list(Ap, fixed, n_fixed, s);
    double den;
    if ((rc = dot(p, Ap, &den))) return rc;
    if (std::fabs(den) < 1e-30) break;
    const float alpha = (float)(rz / den);
    launch_vec(x, p, nullptr, alpha, n, 1, s);
    launch_vec(r, Ap, nullptr, -alpha, n, 1, s);
    launch_zero_list(r, fixed, n_fixed, s);
    if ((rc = dot(r, r, &rr))) return rc;
    if (std::sqrt(rr) < tol * bnorm) {
      ++it;
      break;
    }
    launch_vec(z, r, diagonal, 0.f, n, 2, s);
    double rzn;
    if ((rc = dot(r, z, &rzn))) return rc;
    launch_vec(p, z, nullptr, (float)(rzn / rz), n, 0, s);
    rz = rzn;
  }
  st->iters = it;
  st->relres = bnorm > 0 ? std::sqrt(rr) / bnorm : 0.0;
  st->status = (it >= max_iter) ? GP_NOT_CONVERGED : GP_OK;
  if (hipStreamSynchronize(s) != hipSuccess) return fail(GP_ERR_HIP, "sync failed");
  return GP_OK;
}

// Static device scratch of the no-allocation CSR solve: the scalars, the
// partials and four vectors of kNoAllocMaxN floats (16 MiB + 8 KiB per GPU,
// held until the process ends). In host code `scratch` names a host-side
// stand-in; the device address comes from hipGetSymbolAddress, once per
// device, and only that address is used.
constexpr int64_t kNoAllocMaxN = 1 << 20;
constexpr size_t kScratchScalars = 256;
constexpr size_t kScratchParts = (kVecBlocks * sizeof(double) + 255) & ~size_t(255);
static_assert(sizeof(DevScalars) <= kScratchScalars, "scalars fit their slot");
static __device__ __align__(256) char scratch[kScratchScalars + kScratchParts +
                                              4 * kNoAllocMaxN * sizeof(float)];

std::mutex g_scratch_mu;
std::vector<char*> g_scratch_base;  // device address of `scratch` per device

int scratch_base(char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_base.size() <= dev) g_scratch_base.resize(dev + 1, nullptr);
  if (!g_scratch_base[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GP_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_base[dev] = static_cast<char*>(p);
  }
  *out = g_scratch_base[dev];
  return GP_OK;
}

}  // namespace

extern "C" {

int gp_generic_sens(const void* xPhys, const void* u, const void* ke, const int32_t* edof,
                    int64_t ne, double penal, int32_t dtype, void* dc, void* stream) {
  if (!xPhys || !u || !ke || !edof || !dc || ne < 0) return fail(GP_ERR_ARG, "bad argument");
  if (dtype != GP_F32 && dtype != GP_F64) return fail(GP_ERR_ARG, "dtype must be GP_F32 or GP_F64");
  if (ne == 0) return GP_OK;
  if (dtype == GP_F32)
    launch_generic_sens<float>((const float*)xPhys, (const float*)u, (const float*)ke, edof, ne,
                               penal, (float*)dc, S(stream));
  else
    launch_generic_sens<double>((const double*)xPhys, (const double*)u, (const double*)ke, edof,
                                ne, penal, (double*)dc, S(stream));
  LAUNCHCHK();
  return GP_OK;
}

int gp_generic_spmv_workspace(int64_t ne, int64_t ndof, int64_t* n_ints) {
  if (ne < 0 || ndof < 0 || !n_ints) return fail(GP_ERR_ARG, "bad argument");
  if (24 * ne + 3 * ndof >= (1LL << 31)) return fail(GP_ERR_ARG, "system too large");
  *n_ints = generic_spmv_ws_ints(ne, ndof);
  return GP_OK;
}

int gp_generic_spmv(const void* xPhys, const void* u, const void* ke, const int32_t* edof,
                    int64_t ne, int64_t ndof, double penal, int32_t dtype, int32_t* workspace,
                    void* y, void* stream) {
  if (!xPhys || !u || !ke || !edof || !y || !workspace || ne < 0 || ndof < 1)
    return fail(GP_ERR_ARG, "bad argument");
  if (dtype != GP_F32 && dtype != GP_F64) return fail(GP_ERR_ARG, "dtype must be GP_F32 or GP_F64");
  if (24 * ne + 3 * ndof >= (1LL << 31)) return fail(GP_ERR_ARG, "system too large");
  if (dtype == GP_F32)
    launch_generic_spmv<float>((const float*)xPhys, (const float*)u, (const float*)ke, edof, ne,
                               ndof, penal, workspace, (float*)y, S(stream));
  else
    launch_generic_spmv<double>((const double*)xPhys, (const double*)u, (const double*)ke, edof,
                                ne, ndof, penal, workspace, (double*)y, S(stream));
  LAUNCHCHK();
  return GP_OK;
}

// Isosurface triangles of a density field (mesh_export.py:7-67).
static int mc_check(int32_t nx, int32_t ny, int32_t nz) {
  if (nx < 2 || ny < 2 || nz < 2) return fail(GP_ERR_ARG, "grid %dx%dx%d has no cube", nx, ny, nz);
  if ((long long)(nx - 1) * (ny - 1) * (nz - 1) * 5 >= (1LL << 31))
    return fail(GP_ERR_ARG, "grid %dx%dx%d too large for 32-bit triangle offsets", nx, ny, nz);
  return GP_OK;
}

int gp_mc_workspace(int32_t nx, int32_t ny, int32_t nz, int64_t* n_ints) {
  if (!n_ints) return fail(GP_ERR_ARG, "bad argument");
  int rc = mc_check(nx, ny, nz);
  if (rc) return rc;
  *n_ints = (int64_t)mc_blocks(nx, ny, nz) + 1;
  return GP_OK;
}

int gp_mc_count(const double* density, int32_t nx, int32_t ny, int32_t nz, double level,
                int32_t* workspace, int64_t* n_tri, void* stream) {
  if (!density || !workspace || !n_tri) return fail(GP_ERR_ARG, "bad argument");
  int rc = mc_check(nx, ny, nz);
  if (rc) return rc;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (const char* e = mc_prepare(dev)) return fail(GP_ERR_HIP, "%s", e);
  hipStream_t s = S(stream);
  launch_mc_count(density, nx, ny, nz, level, workspace, s);
  LAUNCHCHK();
  int32_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, workspace + mc_blocks(nx, ny, nz), sizeof(total),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *n_tri = total;
  return GP_OK;
}

int gp_mc_emit(const double* density, int32_t nx, int32_t ny, int32_t nz, double level,
               double sx, double sy, double sz, const int32_t* workspace, float* tris,
               void* stream) {
  if (!density || !workspace || !tris) return fail(GP_ERR_ARG, "bad argument");
  int rc = mc_check(nx, ny, nz);
  if (rc) return rc;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (const char* e = mc_prepare(dev)) return fail(GP_ERR_HIP, "%s", e);
  launch_mc_emit(density, nx, ny, nz, level, sx, sy, sz, workspace, tris, S(stream));
  LAUNCHCHK();
  return GP_OK;
}

int gp_mc_tables(int8_t* ntri, int8_t* tri) {
  if (!ntri || !tri) return fail(GP_ERR_ARG, "bad argument");
  mc_tables_host(ntri, tri);
  return GP_OK;
}

// C3D8 mesh of the solid voxels (voxel_fem.py:120-146, mesher_simd.cpp:11-52).
static int hex_check(int32_t D, int32_t H, int32_t W) {
  if (D < 1 || H < 1 || W < 1) return fail(GP_ERR_ARG, "voxel grid must be >= 1 per axis");
  if ((long long)D * H * W > (1LL << 28) || (long long)(D + 1) * (H + 1) * (W + 1) >= (1LL << 30))
    return fail(GP_ERR_ARG, "voxel grid %dx%dx%d too large for 32-bit node numbers", D, H, W);
  return GP_OK;
}

int gp_hex_mesh_workspace(int32_t D, int32_t H, int32_t W, int64_t* n_ints) {
  if (!n_ints) return fail(GP_ERR_ARG, "bad argument");
  int rc = hex_check(D, H, W);
  if (rc) return rc;
  *n_ints = 2ll * (hex_blocks(D, H, W) + 1) + (int64_t)(D + 1) * (H + 1) * (W + 1);
  return GP_OK;
}

int gp_hex_mesh_count(const float* voxels, int32_t D, int32_t H, int32_t W, int32_t* workspace,
                      int64_t* n_elem, int64_t* n_node, void* stream) {
  if (!voxels || !workspace || !n_elem || !n_node) return fail(GP_ERR_ARG, "bad argument");
  int rc = hex_check(D, H, W);
  if (rc) return rc;
  hipStream_t s = S(stream);
  launch_hex_count(voxels, D, H, W, workspace, s);
  LAUNCHCHK();
  const int nb = hex_blocks(D, H, W);
  int32_t tot[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(&tot[0], workspace + nb, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&tot[1], workspace + 2 * nb + 1, sizeof(int32_t), hipMemcpyDeviceToHost,
                        s));
  HIPCHK(hipStreamSynchronize(s));
  *n_elem = tot[0];
  *n_node = tot[1];
  return GP_OK;
}

int gp_hex_mesh_emit(const float* voxels, int32_t D, int32_t H, int32_t W, int32_t* workspace,
                     int32_t* elems, int32_t* node_ijk, int32_t* elem_voxel, void* stream) {
  if (!voxels || !workspace || !elems || !node_ijk) return fail(GP_ERR_ARG, "bad argument");
  int rc = hex_check(D, H, W);
  if (rc) return rc;
  launch_hex_emit(voxels, D, H, W, workspace, elems, node_ijk, elem_voxel, S(stream));
  LAUNCHCHK();
  return GP_OK;
}

// Evaluator pre-processing (voxel_fem.py:452-462, pipeline_utils.py:55-76).
int gp_voxel_prep(const float* in, int32_t batch, int32_t d, int32_t h, int32_t w, int32_t od,
                  int32_t oh, int32_t ow, const double* weights, int32_t radius,
                  double sharpness, float* work, float* out, void* stream) {
  if (!in || !weights || !work || !out || batch < 0 || d < 1 || h < 1 || w < 1 || od < 1 ||
      oh < 1 || ow < 1)
    return fail(GP_ERR_ARG, "bad argument");
  if (radius < 0 || 2 * radius + 1 > kMaxTaps)
    return fail(GP_ERR_ARG, "filter radius %d outside [0, %d]", radius, (kMaxTaps - 1) / 2);
  for (int q = 1; q <= radius; ++q)
    if (weights[radius - q] != weights[radius + q])
      return fail(GP_ERR_ARG, "filter weights are not symmetric");
  if (batch == 0) return GP_OK;
  launch_voxel_prep(in, batch, d, h, w, od, oh, ow, weights, radius, sharpness, work, out,
                    S(stream));
  LAUNCHCHK();
  return GP_OK;
}

// simp_pcg_cuda.cu:66-124 restated: x starts at x0 but r starts at b (the
// reference never forms b - A x0); z = r / diagonal; ||b|| is the unmasked
// norm; |p^T A p| < 1e-30 stops the loop. Host-synchronous per iteration,
// like the reference.
int gp_csr_pcg(const int64_t* rp, const int64_t* ci, const float* v, const float* diagonal,
               const float* b, float* x, const int64_t* fixed, int64_t n_fixed, int64_t n,
               int32_t max_iter, double tol, gp_pcg_stats* st, void* stream) {
  if (!rp || !ci || !v || !diagonal || !b || !x || !st || n < 0 || n_fixed < 0 ||
      (n_fixed > 0 && !fixed))
    return fail(GP_ERR_ARG, "bad argument");
  std::memset(st, 0, sizeof(*st));
  if (n == 0) return GP_OK;
  hipStream_t s = S(stream);
  CsrPcgWork w{};
  DevScalars* h = nullptr;
  int rc = GP_OK;
  auto cleanup = [&]() {
    sync_for_free();
    for (void* q : {(void*)w.r, (void*)w.z, (void*)w.p, (void*)w.Ap, (void*)w.sc, (void*)w.part})
      if (q) dev_free(q, true);
    if (h) hipHostFree(h);
  };
  do {
    if ((rc = alloc((void**)&w.r, n * 4)) || (rc = alloc((void**)&w.z, n * 4)) ||
        (rc = alloc((void**)&w.p, n * 4)) || (rc = alloc((void**)&w.Ap, n * 4)) ||
        (rc = alloc((void**)&w.sc, sizeof(DevScalars))) ||
        (rc = alloc((void**)&w.part, kVecBlocks * sizeof(double))))
      break;
    if (hipHostMalloc((void**)&h, sizeof(DevScalars)) != hipSuccess) {
      rc = fail(GP_ERR_OOM, "hipHostMalloc failed");
      break;
    }
    rc = csr_pcg_run(rp, ci, v, diagonal, b, x, fixed, n_fixed, n, max_iter, tol, w, h, st, s);
  } while (false);
  cleanup();
  return rc;
}

// No counterpart in the reference: gp_csr_pcg for callers that cannot
// allocate (code under graph capture, tight serving loops). The four PCG
// vectors, the scalars and the partial sums live in the library's static
// device array; n is limited to kNoAllocMaxN.
int gp_csr_pcg_noalloc(const int64_t* rp, const int64_t* ci, const float* v,
                       const float* diagonal, const float* b, float* x, const int64_t* fixed,
                       int64_t n_fixed, int64_t n, int32_t max_iter, double tol,
                       gp_pcg_stats* st, void* stream) {
  if (!rp || !ci || !v || !diagonal || !b || !x || !st || n < 0 || n_fixed < 0 ||
      (n_fixed > 0 && !fixed))
    return fail(GP_ERR_ARG, "bad argument");
  if (n > kNoAllocMaxN)
    return fail(GP_ERR_ARG, "n = %lld exceeds the no-allocation limit of %lld unknowns",
                (long long)n, (long long)kNoAllocMaxN);
  std::memset(st, 0, sizeof(*st));
  if (n == 0) return GP_OK;
  char* base = nullptr;
  int rc = scratch_base(&base);
  if (rc) return rc;
  CsrPcgWork w{};
  w.sc = reinterpret_cast<DevScalars*>(base);
  w.part = reinterpret_cast<double*>(base + kScratchScalars);
  float* vec = reinterpret_cast<float*>(base + kScratchScalars + kScratchParts);
  w.r = vec;
  w.z = vec + kNoAllocMaxN;
  w.p = vec + 2 * kNoAllocMaxN;
  w.Ap = vec + 3 * kNoAllocMaxN;
  DevScalars h{};  // pageable: read back with a synchronising copy
  return csr_pcg_run(rp, ci, v, diagonal, b, x, fixed, n_fixed, n, max_iter, tol, w, &h, st,
                     S(stream));
}

}  // extern "C"

// ================================================================ batches
struct gp_batch {
  std::vector<gp_problem*> probs;
  int device = 0;
  KuTile tile{8, 32, 1};
  std::vector<BatchProb> h_probs;
  BatchProb* d_probs = nullptr;
  int* d_off = nullptr;
  BatchLaunch bl{};
  BatchStatus* d_status = nullptr;
  BatchStatus* h_status = nullptr;  // pinned, 2 slots of n
  hipEvent_t ev[2] = {nullptr, nullptr};
  int oc_max_steps = 50;
};

namespace {

// Device view of problem P for the batched and scheduled kernels: pointers
// into P's workspaces, modal sweep geometry for tile (tj, tk, ni).
int fill_batch_prob(gp_problem* P, int tj, int tk, int ni, BatchProb& q) {
  int rc = ensure_simp_ws(P);
  if (!rc) rc = ensure_refine_ws(P);
  if (!rc) rc = ensure_pcg_ws(P, GP_F32);
  if (rc) return rc;
  const Grid& g = P->g;
  const size_t nd = (size_t)g.ndof;
  float* base = reinterpret_cast<float*>(P->ws32);
  q = BatchProb{};
  q.g = g;
  q.mc = P->mc32;
  q.mc64 = P->mc64;
  q.ni = ni;
  q.ntj = (g.ny + 1 + (tj - 2)) / (tj - 1);
  q.ntk = (g.nz + 1 + (tk - 2)) / (tk - 1);
  q.law = P->spec.law;
  q.oc_variant = P->spec.oc_variant;
  q.penal = P->spec.penal;
  q.E0 = P->spec.E0;
  q.Emin = P->spec.Emin;
  q.move = P->spec.move;
  q.n_pres = P->n_pres;
  q.ke64 = P->ke64;
  q.ke32 = P->ke32;
  q.st = P->st;
  q.Hs = P->Hs;
  q.w = P->w;
  q.pres = P->pres;
  q.fixmask = P->fixmask;
  q.f = P->f;
  q.xnew = P->xnew;
  q.xphys_new = P->xphys_new;
  q.dc = P->dc;
  q.dcf = P->dcf;
  q.E = P->E;
  q.u = P->d32;
  q.r64 = P->r64;
  q.rb[0] = base;
  q.rb[1] = base + nd;
  q.wb[0] = base + 2 * nd;
  q.wb[1] = base + 3 * nd;
  q.minv = base + 4 * nd;
  q.sb[0] = base + 5 * nd;
  q.sb[1] = base + 6 * nd;
  q.pc = base + 7 * nd;
  q.sc = &P->sc[0];
  q.sc_oc = &P->sc[1];
  q.parts = P->parts;
  return GP_OK;
}

// Block counts of the five kernel kinds (cg, node, elem, vec, oc) for grid g.
void kind_counts(const Grid& g, const KuTile& t, int out[5]) {
  out[0] = modal_blocks(g, t);
  out[1] = grid_blocks(g.nn);
  out[2] = grid_blocks(g.ne);
  out[3] = vec_blocks(g);
  out[4] = oc_blocks(g);
}

int batch_wait_all_done(gp_batch* B, hipStream_t s, int max_iter, int* iters_out) {
  const int n = (int)B->probs.size();
  int chunk = 8, issued = 0, slot = 0, k = 0;
  auto issue = [&](int cnt) {
    for (int i = 0; i < cnt; ++i, ++k) batch_cg(B->bl, k, s);
  };
  auto snap = [&](int sl) -> int {
    batch_status(B->bl, B->d_status + sl * n, s);
    HIPCHK(hipMemcpyAsync(B->h_status + sl * n, B->d_status + sl * n, n * sizeof(BatchStatus),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(B->ev[sl], s));
    return GP_OK;
  };
  int rc = snap(0);
  if (rc) return rc;
  issue(chunk);
  issued += chunk;
  rc = snap(1);
  if (rc) return rc;
  LAUNCHCHK();
  while (true) {
    HIPCHK(hipEventSynchronize(B->ev[slot]));
    bool all = true;
    for (int b = 0; b < n; ++b) all = all && B->h_status[slot * n + b].done;
    if (all || issued >= max_iter + 2 * chunk + 2) break;
    chunk = std::min(chunk * 2, 64);
    issue(chunk);
    issued += chunk;
    rc = snap(slot);
    if (rc) return rc;
    LAUNCHCHK();
    slot ^= 1;
  }
  HIPCHK(hipStreamSynchronize(s));
  rc = snap(0);
  if (rc) return rc;
  HIPCHK(hipEventSynchronize(B->ev[0]));
  for (int b = 0; b < n; ++b) iters_out[b] = B->h_status[b].iter;
  return GP_OK;
}

}  // namespace

extern "C" {

int gp_batch_create(gp_problem* const* probs, int32_t n, int32_t tj, int32_t tk, gp_batch** out) {
  if (!probs || n < 1 || !out) return fail(GP_ERR_ARG, "bad argument");
  *out = nullptr;
  bool known = false;
  for (const KuTile& t : kTiles) known = known || (t.tj == tj && t.tk == tk);
  if (!known) return fail(GP_ERR_ARG, "unsupported K.u tile %dx%d", tj, tk);
  for (int b = 0; b < n; ++b) {
    if (!probs[b]) return fail(GP_ERR_ARG, "problem %d is NULL", b);
    if (!probs[b]->modal_ok) return fail(GP_ERR_ARG, "problem %d: KE is not an isotropic hex8 table", b);
    if (!probs[b]->has_bcs) return fail(GP_ERR_STATE, "problem %d: boundary conditions not set", b);
    if (probs[b]->device != probs[0]->device) return fail(GP_ERR_ARG, "problems on different devices");
  }
  HIPCHK(hipSetDevice(probs[0]->device));
  gp_batch* B = new gp_batch();
  B->probs.assign(probs, probs + n);
  B->device = probs[0]->device;
  B->tile = {tj, tk, 1};
  auto bail = [&](int code) {
    gp_batch_destroy(B);
    return code;
  };
  std::vector<int> off(5 * (n + 1), 0);
  B->h_probs.resize(n);
  bool any_cpu_oc = false;
  for (int b = 0; b < n; ++b) {
    gp_problem* P = probs[b];
    BatchProb& q = B->h_probs[b];
    const int ni = P->tile32.ni > 0 ? P->tile32.ni : 2;
    int rc = fill_batch_prob(P, tj, tk, ni, q);
    if (rc) return bail(rc);
    any_cpu_oc = any_cpu_oc || q.oc_variant == GP_OC_CPU;
    const Grid& g = P->g;
    const KuTile t{tj, tk, q.ni};
    const int counts[5] = {modal_blocks(g, t), grid_blocks(g.nn), grid_blocks(g.ne), vec_blocks(g),
                           oc_blocks(g)};
    for (int kk = 0; kk < 5; ++kk) {
      if (counts[kk] > P->parts.cap) return bail(fail(GP_ERR_ARG, "partials too small"));
      off[kk * (n + 1) + b + 1] = off[kk * (n + 1) + b] + counts[kk];
    }
  }
  B->oc_max_steps = any_cpu_oc ? 50 : 30;
  int rc = alloc(reinterpret_cast<void**>(&B->d_probs), n * sizeof(BatchProb));
  if (!rc) rc = alloc(reinterpret_cast<void**>(&B->d_off), off.size() * sizeof(int));
  if (!rc) rc = alloc(reinterpret_cast<void**>(&B->d_status), 2 * n * sizeof(BatchStatus));
  if (rc) return bail(rc);
  if (hipHostMalloc(reinterpret_cast<void**>(&B->h_status), 2 * n * sizeof(BatchStatus)) != hipSuccess)
    return bail(fail(GP_ERR_OOM, "hipHostMalloc failed"));
  if (hipEventCreateWithFlags(&B->ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&B->ev[1], hipEventDisableTiming) != hipSuccess)
    return bail(fail(GP_ERR_HIP, "hipEventCreate failed"));
  if (hipMemcpy(B->d_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(GP_ERR_HIP, "offset copy failed"));
  BatchLaunch& bl = B->bl;
  bl.probs = B->d_probs;
  bl.n = n;
  bl.tj = tj;
  bl.tk = tk;
  const int* d = B->d_off;
  bl.off_cg = d;
  bl.off_node = d + (n + 1);
  bl.off_elem = d + 2 * (n + 1);
  bl.off_vec = d + 3 * (n + 1);
  bl.off_oc = d + 4 * (n + 1);
  bl.tot_cg = off[n];
  bl.tot_node = off[(n + 1) + n];
  bl.tot_elem = off[2 * (n + 1) + n];
  bl.tot_vec = off[3 * (n + 1) + n];
  bl.tot_oc = off[4 * (n + 1) + n];
  *out = B;
  return GP_OK;
}

int gp_batch_destroy(gp_batch* B) {
  if (!B) return GP_OK;
  hipSetDevice(B->device);
  // one wait for the work that may still use these buffers, then the frees
  sync_for_free();
  if (B->d_probs) dev_free(B->d_probs, true);
  if (B->d_off) dev_free(B->d_off, true);
  if (B->d_status) dev_free(B->d_status, true);
  if (B->h_status) hipHostFree(B->h_status);
  for (auto& e : B->ev)
    if (e) hipEventDestroy(e);
  delete B;
  return GP_OK;
}

int gp_batch_simp_run(gp_batch* B, const gp_simp_opts* o, const double* volfracs, double* const* x,
                      double* const* xPhys, double* const* u, gp_simp_stats* stats, void* stream) {
  if (!B || !o || !volfracs || !x || !xPhys || !u || !stats) return fail(GP_ERR_ARG, "NULL argument");
  if (o->pcg.refine || o->pcg.precision != GP_F32)
    return fail(GP_ERR_ARG, "batched runs use float32 PCG vectors (precision GP_F32, refine 0)");
  if (o->recycle != 0)
    return fail(GP_ERR_ARG, "batched runs start each solve from the last u (recycle 0)");
  const int n = (int)B->probs.size();
  for (int b = 0; b < n; ++b)
    if (!x[b] || !xPhys[b] || !u[b]) return fail(GP_ERR_ARG, "NULL state pointer for problem %d", b);
  HIPCHK(hipSetDevice(B->device));
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = S(stream);
  for (int b = 0; b < n; ++b) {
    B->h_probs[b].volfrac = volfracs[b];
    B->h_probs[b].x = x[b];
    B->h_probs[b].xPhys = xPhys[b];
    B->h_probs[b].u64 = u[b];
    std::memset(&stats[b], 0, sizeof(gp_simp_stats));
  }
  HIPCHK(hipMemcpyAsync(B->d_probs, B->h_probs.data(), n * sizeof(BatchProb), hipMemcpyHostToDevice, s));
  const BatchLaunch& bl = B->bl;
  if (o->warm_start) batch_cast(bl, 0, s);
  std::vector<int> its(n, 0);
  for (int it = 0; it < o->n_iters; ++it) {
    batch_stiffness(bl, s);
    batch_pcg_setup(bl, o->warm_start != 0, o->pcg.tol, o->pcg.max_iter, s);
    LAUNCHCHK();
    int rc = batch_wait_all_done(B, s, o->pcg.max_iter, its.data());
    if (rc) return rc;
    for (int b = 0; b < n; ++b) {
      stats[b].pcg_iters_total += its[b];
      if (B->h_status[b].status > stats[b].status) stats[b].status = B->h_status[b].status;
    }
    batch_post_solve(bl, B->oc_max_steps, s);
    LAUNCHCHK();
    for (int b = 0; b < n; ++b) stats[b].iters_done = it + 1;
  }
  // final compliance solve at the returned densities (simp_solver.py:78)
  batch_stiffness(bl, s);
  batch_pcg_setup(bl, o->warm_start != 0, o->pcg.tol, o->pcg.max_iter, s);
  int rc = batch_wait_all_done(B, s, o->pcg.max_iter, its.data());
  if (rc) return rc;
  batch_compliance(bl, s);
  batch_cast(bl, 1, s);
  batch_status(bl, B->d_status, s);
  HIPCHK(hipMemcpyAsync(B->h_status, B->d_status, n * sizeof(BatchStatus), hipMemcpyDeviceToHost, s));
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int b = 0; b < n; ++b) {
    stats[b].pcg_iters_total += its[b];
    stats[b].compliance = B->h_status[b].sum;
    stats[b].seconds = secs;
  }
  return GP_OK;
}


}  // extern "C"

// ============================================================ scheduler
struct gp_sched {
  int device = 0;
  int tj = 8, tk = 32;
  int n = 0;
  std::vector<SchedSlot> h_slots;
  std::vector<int> dims;       // per slot: nx, ny, nz, ni (capacity)
  std::vector<int> off;        // n + 1 block offsets
  std::vector<char> busy;
  SchedSlot* d_slots = nullptr;
  int* d_off = nullptr;
  int* d_b2s = nullptr;
  SchedState* d_status = nullptr;
  SchedState* h_status = nullptr;  // pinned, 2 snapshots of n
  hipEvent_t ev[2] = {nullptr, nullptr};
  long long t = 0;                 // ticks issued
  hipStream_t cap = nullptr;       // capture-only stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_ticks = 0;
  // Blocks with work: while every slot runs a job that fills its range, or
  // the running jobs are small, the launch covers all blocks (nact = -1) and
  // ticks replay from the graph; otherwise only the first nb_cg blocks of
  // each running slot are launched, listed in d_act (a finished or empty
  // slot's blocks would only idle).
  int* d_act = nullptr;
  std::vector<int> h_act;
  int nact = -1;
  std::vector<char> done_seen;  // a status read showed the slot's job done
  SchedTable table() const { return SchedTable{d_slots, n, d_off, d_b2s, d_status, nullptr}; }
  SchedTable table_act() const { return SchedTable{d_slots, n, d_off, d_b2s, d_status, d_act}; }
};

namespace {

void sched_free_graph(gp_sched* Q) {
  if (Q->exec) hipGraphExecDestroy(Q->exec);
  if (Q->graph) hipGraphDestroy(Q->graph);
  Q->exec = nullptr;
  Q->graph = nullptr;
  Q->graph_ticks = 0;
}

SchedState idle_state() {
  SchedState z{};
  z.phase = PH_IDLE;
  z.prev_c = INFINITY;
  return z;
}

// Smallest job (in sweep blocks) whose ticks are launched without the
// blocks that would idle; measured on a 48-problem 32^3..64^3 batch (ticks of
// about 120 us, 64^3 jobs of 180 blocks) and 64 data-generation jobs of
// 15-60 blocks (ticks of a few us).
constexpr int kCompactMinBlocks = 96;

// Rebuilds the list of launched blocks after a job starts, ends or is seen
// finished; the upload is ordered on `s` after every tick issued before.
int sched_update_active(gp_sched* Q, hipStream_t s) {
  std::vector<int> a;
  bool full = true;
  int longest = 0;
  for (int i = 0; i < Q->n; ++i) {
    const bool run = Q->busy[i] && !Q->done_seen[i];
    const int nb = run ? Q->h_slots[i].nb_cg : 0;
    if (!run || nb != Q->off[i + 1] - Q->off[i]) full = false;
    longest = std::max(longest, nb);
    for (int k = 0; k < nb; ++k) a.push_back(Q->off[i] + k);
  }
  // A tick lasts about as long as the largest running job's sweep. Short
  // ticks (small jobs) are launch-bound and replay cheaper from the graph of
  // all blocks than they launch directly; long ticks gain from not
  // dispatching idle blocks.
  if (full || longest < kCompactMinBlocks) {
    Q->nact = -1;
    return GP_OK;
  }
  if (Q->nact == (int)a.size() && a == Q->h_act) return GP_OK;
  Q->h_act.swap(a);
  Q->nact = (int)Q->h_act.size();
  if (Q->nact > 0)
    HIPCHK(hipMemcpyAsync(Q->d_act, Q->h_act.data(), Q->nact * sizeof(int),
                          hipMemcpyHostToDevice, s));
  return GP_OK;
}

// Marks the jobs a status snapshot shows finished, then updates the list.
int sched_note_status(gp_sched* Q, const SchedState* st, hipStream_t s) {
  bool changed = false;
  for (int i = 0; i < Q->n; ++i)
    if (Q->busy[i] && !Q->done_seen[i] && st[i].phase == PH_DONE) {
      Q->done_seen[i] = 1;
      changed = true;
    }
  return changed ? sched_update_active(Q, s) : GP_OK;
}

int sched_issue(gp_sched* Q, long long n, hipStream_t s) {
  if (Q->nact >= 0) {
    // some blocks would only idle: launch just the listed ones (no graph,
    // whose launches cover every block); no job running, no launch
    const SchedTable tab = Q->table_act();
    for (; n > 0; --n, ++Q->t)
      if (Q->nact > 0) sched_tick(tab, Q->tj, Q->tk, Q->nact, (int)(Q->t & 1), s);
    LAUNCHCHK();
    return GP_OK;
  }
  const SchedTable tab = Q->table();
  const int blocks = Q->off[Q->n];
  while (n > 0) {
    if (Q->exec && (Q->t & 1) == 0 && n >= Q->graph_ticks) {
      HIPCHK(hipGraphLaunch(Q->exec, s));
      Q->t += Q->graph_ticks;
      n -= Q->graph_ticks;
    } else {
      sched_tick(tab, Q->tj, Q->tk, blocks, (int)(Q->t & 1), s);
      Q->t += 1;
      n -= 1;
    }
  }
  LAUNCHCHK();
  return GP_OK;
}

void to_public(const SchedState& st, gp_slot_status* o) {
  o->phase = st.phase;
  o->done = st.phase == PH_DONE ? 1 : 0;
  o->iters_done = st.oc_it;
  o->pcg_iters_total = st.pcg_total;
  o->bisect_steps_total = st.bis_total;
  o->status = st.status;
  o->compliance = st.curr_c;
}

}  // namespace

extern "C" {

int gp_sched_create(int32_t n_slots, const int32_t* slot_dims, int32_t tj, int32_t tk,
                    gp_sched** out) {
  if (n_slots < 1 || !slot_dims || !out) return fail(GP_ERR_ARG, "bad argument");
  if (!((tj == 8 && tk == 32) || (tj == 16 && tk == 16) || (tj == 14 && tk == 18) ||
        (tj == 6 && tk == 34) || (tj == 7 && tk == 34)))
    return fail(GP_ERR_ARG, "scheduler tiles are 8x32, 16x16, 14x18, 7x34 or 6x34, got %dx%d", tj,
                tk);
  *out = nullptr;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  gp_sched* Q = new gp_sched();
  Q->device = dev;
  Q->tj = tj;
  Q->tk = tk;
  Q->n = n_slots;
  Q->dims.assign(slot_dims, slot_dims + 4 * n_slots);
  Q->h_slots.assign(n_slots, SchedSlot{});
  Q->busy.assign(n_slots, 0);
  Q->done_seen.assign(n_slots, 0);
  Q->off.assign(n_slots + 1, 0);
  auto bail = [&](int code) {
    gp_sched_destroy(Q);
    return code;
  };
  for (int i = 0; i < n_slots; ++i) {
    gp_spec sp{};
    sp.nx = slot_dims[4 * i];
    sp.ny = slot_dims[4 * i + 1];
    sp.nz = slot_dims[4 * i + 2];
    const int ni = slot_dims[4 * i + 3];
    if (sp.nx < 1 || sp.ny < 1 || sp.nz < 1 || ni < 1)
      return bail(fail(GP_ERR_ARG, "slot %d: bad capacity dims", i));
    const Grid g = make_grid(&sp);
    const int cap = std::max(1, modal_blocks(g, KuTile{tj, tk, ni}));
    Q->off[i + 1] = Q->off[i] + cap;
    SchedSlot& q = Q->h_slots[i];
    q.cap = cap;
    q.st[0] = idle_state();
    q.st[1] = idle_state();
  }
  const int total = Q->off[n_slots];
  std::vector<int> b2s(total);
  for (int i = 0; i < n_slots; ++i)
    for (int k = Q->off[i]; k < Q->off[i + 1]; ++k) b2s[k] = i;
  int rc = alloc(reinterpret_cast<void**>(&Q->d_slots), n_slots * sizeof(SchedSlot));
  if (!rc) rc = alloc(reinterpret_cast<void**>(&Q->d_off), (n_slots + 1) * sizeof(int));
  if (!rc) rc = alloc(reinterpret_cast<void**>(&Q->d_b2s), total * sizeof(int));
  if (!rc) rc = alloc(reinterpret_cast<void**>(&Q->d_act), total * sizeof(int));
  if (!rc) rc = alloc(reinterpret_cast<void**>(&Q->d_status), n_slots * sizeof(SchedState));
  if (rc) return bail(rc);
  if (hipHostMalloc(reinterpret_cast<void**>(&Q->h_status), 2 * n_slots * sizeof(SchedState)) !=
      hipSuccess)
    return bail(fail(GP_ERR_OOM, "hipHostMalloc failed"));
  if (hipEventCreateWithFlags(&Q->ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&Q->ev[1], hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&Q->cap, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(GP_ERR_HIP, "event/stream creation failed"));
  std::vector<SchedState> st(n_slots, idle_state());
  if (hipMemcpy(Q->d_slots, Q->h_slots.data(), n_slots * sizeof(SchedSlot), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(Q->d_off, Q->off.data(), (n_slots + 1) * sizeof(int), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(Q->d_b2s, b2s.data(), total * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(Q->d_status, st.data(), n_slots * sizeof(SchedState), hipMemcpyHostToDevice) !=
          hipSuccess)
    return bail(fail(GP_ERR_HIP, "scheduler table upload failed"));
  *out = Q;
  return GP_OK;
}

int gp_sched_destroy(gp_sched* Q) {
  if (!Q) return GP_OK;
  hipSetDevice(Q->device);
  // tick graphs may still be executing on the caller's streams, which the
  // scheduler does not know: they end before the graphs are destroyed
  hipDeviceSynchronize();
  sched_free_graph(Q);
  if (Q->cap) hipStreamDestroy(Q->cap);
  if (Q->d_slots) dev_free(Q->d_slots, true);  // after the device-wide wait above
  if (Q->d_off) dev_free(Q->d_off, true);
  if (Q->d_b2s) dev_free(Q->d_b2s, true);
  if (Q->d_act) dev_free(Q->d_act, true);
  if (Q->d_status) dev_free(Q->d_status, true);
  if (Q->h_status) hipHostFree(Q->h_status);
  for (auto& e : Q->ev)
    if (e) hipEventDestroy(e);
  delete Q;
  return GP_OK;
}

int gp_sched_blocks(gp_sched* Q, int32_t slot, int32_t* blocks) {
  if (!Q || !blocks || slot < -1 || slot >= Q->n) return fail(GP_ERR_ARG, "bad argument");
  *blocks = slot < 0 ? Q->off[Q->n] : Q->off[slot + 1] - Q->off[slot];
  return GP_OK;
}

int gp_sched_admit(gp_sched* Q, int32_t slot, gp_problem* P, const gp_job* job, double* x,
                   double* xPhys, double* u, void* stream) {
  if (!Q || !P || !job || !x || !xPhys || !u) return fail(GP_ERR_ARG, "NULL argument");
  if (slot < 0 || slot >= Q->n) return fail(GP_ERR_ARG, "slot %d out of range", slot);
  if (Q->busy[slot]) return fail(GP_ERR_STATE, "slot %d is occupied", slot);
  if (!P->modal_ok) return fail(GP_ERR_ARG, "problem KE is not an isotropic hex8 table");
  if (!P->has_bcs) return fail(GP_ERR_STATE, "boundary conditions not set");
  if (P->device != Q->device) return fail(GP_ERR_ARG, "problem lives on another device");
  if (job->n_iters < 0 || job->max_iter < 1 || !(job->tol >= 0))
    return fail(GP_ERR_ARG, "bad job parameters");
  if (job->refine != 0 && job->refine != 1) return fail(GP_ERR_ARG, "refine must be 0 or 1");
  if (job->refine && (!(job->inner_tol >= 0) || job->max_outer < 0))
    return fail(GP_ERR_ARG, "bad refinement parameters");
  if (job->recycle < 0 || job->recycle > kRecycleMax)
    return fail(GP_ERR_ARG, "recycle must be 0..%d", kRecycleMax);
  HIPCHK(hipSetDevice(Q->device));
  // recycled starts use the problem's own ring (allocated on first use; the
  // job's history starts empty)
  if (job->recycle > 0) {
    int rr = ensure_recycle_ws(P, job->recycle);
    if (rr) return rr;
  }
  SchedSlot& q = Q->h_slots[slot];
  const int ni = Q->dims[4 * slot + 3];
  int rc = fill_batch_prob(P, Q->tj, Q->tk, ni, q.P);
  if (rc) return rc;
  int c[5];
  kind_counts(P->g, KuTile{Q->tj, Q->tk, ni}, c);
  if (c[0] > q.cap)
    return fail(GP_ERR_ARG, "problem %dx%dx%d does not fit slot %d", P->g.nx, P->g.ny, P->g.nz,
                slot);
  for (int kk = 0; kk < 5; ++kk)
    if (c[kk] > P->parts.cap) return fail(GP_ERR_ARG, "partials too small for slot tile");
  q.nb_cg = c[0];
  q.nb_node = c[1];
  q.nb_elem = c[2];
  q.nb_vec = c[3];
  q.nb_oc = c[4];
  q.P.volfrac = job->volfrac;
  q.P.x = x;
  q.P.xPhys = xPhys;
  q.P.u64 = u;
  q.cfg.n_iters = job->n_iters;
  q.cfg.max_iter = job->max_iter;
  q.cfg.tol_schedule = job->tol_schedule;
  q.cfg.compl_each = job->compliance_each_iter;
  q.cfg.early_stop = job->early_stop;
  q.cfg.warm_start = job->warm_start;
  q.cfg.tol = job->tol;
  // the defaults of pcg_refine for a zero inner tolerance or sweep cap
  q.cfg.refine = job->refine;
  q.cfg.inner_tol = job->inner_tol > 0 ? job->inner_tol : 1e-3;
  q.cfg.max_outer = job->max_outer > 0 ? job->max_outer : 50;
  q.cfg.warm_compl = job->warm_compliance != 0 ? 1 : 0;
  q.cfg.recycle = job->recycle;
  if (job->recycle > 0) {
    q.P.rc_v = P->rc_v;
    q.P.rc_kv = P->rc_kv;
    q.P.rc_part = P->rc_part;
    q.P.rc_coef = P->rc_coef;
  }
  SchedState s0 = idle_state();
  // an FP64-refined job works on the caller's u directly: no float copy in
  s0.phase = job->warm_start && !job->refine ? PH_CAST_IN : PH_STIFF;
  s0.solve = job->n_iters > 0 ? SK_OC : SK_FINAL;
  q.st[0] = s0;
  q.st[1] = s0;
  hipStream_t s = S(stream);
  HIPCHK(hipMemcpyAsync(Q->d_slots + slot, &q, sizeof(SchedSlot), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(Q->d_status + slot, &q.st[0], sizeof(SchedState), hipMemcpyHostToDevice, s));
  Q->busy[slot] = 1;
  Q->done_seen[slot] = 0;
  return sched_update_active(Q, s);
}

int gp_sched_set_rhs(gp_sched* Q, int32_t slot, const double* f, void* stream) {
  if (!Q || !f || slot < 0 || slot >= Q->n) return fail(GP_ERR_ARG, "bad argument");
  if (!Q->busy[slot]) return fail(GP_ERR_STATE, "slot %d holds no job", slot);
  SchedSlot& q = Q->h_slots[slot];
  if (q.cfg.n_iters != 0) return fail(GP_ERR_ARG, "a right-hand side applies to single solves only");
  HIPCHK(hipSetDevice(Q->device));
  q.P.f = f;
  HIPCHK(hipMemcpyAsync(&Q->d_slots[slot].P.f, &q.P.f, sizeof(q.P.f), hipMemcpyHostToDevice,
                        S(stream)));
  return GP_OK;
}

int gp_sched_release(gp_sched* Q, int32_t slot, void* stream) {
  if (!Q || slot < 0 || slot >= Q->n) return fail(GP_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(Q->device));
  SchedSlot& q = Q->h_slots[slot];
  q.st[0] = idle_state();
  q.st[1] = idle_state();
  hipStream_t s = S(stream);
  HIPCHK(hipMemcpyAsync(&Q->d_slots[slot].st[0], &q.st[0], 2 * sizeof(SchedState),
                        hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(Q->d_status + slot, &q.st[0], sizeof(SchedState), hipMemcpyHostToDevice, s));
  Q->busy[slot] = 0;
  Q->done_seen[slot] = 0;
  return sched_update_active(Q, s);
}

int gp_sched_set_graph(gp_sched* Q, int32_t ticks) {
  if (!Q || ticks < 0 || (ticks & 1)) return fail(GP_ERR_ARG, "graph ticks must be even and >= 0");
  HIPCHK(hipSetDevice(Q->device));
  sched_free_graph(Q);
  if (ticks == 0) return GP_OK;
  const SchedTable tab = Q->table();
  const int blocks = Q->off[Q->n];
  HIPCHK(hipStreamBeginCapture(Q->cap, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < ticks; ++i) sched_tick(tab, Q->tj, Q->tk, blocks, i & 1, Q->cap);
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamEndCapture(Q->cap, &g));
  Q->graph = g;
  if (hipGraphInstantiate(&Q->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    sched_free_graph(Q);
    return fail(GP_ERR_HIP, "hipGraphInstantiate failed");
  }
  Q->graph_ticks = ticks;
  return GP_OK;
}

int gp_sched_tick(gp_sched* Q, int64_t n_ticks, void* stream) {
  if (!Q || n_ticks < 0) return fail(GP_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(Q->device));
  return sched_issue(Q, n_ticks, S(stream));
}

int gp_sched_status(gp_sched* Q, gp_slot_status* out, void* stream) {
  if (!Q || !out) return fail(GP_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(Q->device));
  hipStream_t s = S(stream);
  HIPCHK(hipMemcpyAsync(Q->h_status, Q->d_status, Q->n * sizeof(SchedState), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int i = 0; i < Q->n; ++i) to_public(Q->h_status[i], &out[i]);
  return sched_note_status(Q, Q->h_status, s);
}

int gp_sched_run(gp_sched* Q, int64_t max_ticks, int64_t* ticks_run, gp_slot_status* out,
                 void* stream) {
  if (!Q || max_ticks < 0) return fail(GP_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(Q->device));
  hipStream_t s = S(stream);
  const int n = Q->n;
  const long long chunk = Q->graph_ticks > 0 ? Q->graph_ticks : 32;
  const long long t0 = Q->t;
  auto snap = [&](int sl) -> int {
    HIPCHK(hipMemcpyAsync(Q->h_status + sl * n, Q->d_status, n * sizeof(SchedState),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(Q->ev[sl], s));
    return GP_OK;
  };
  auto all_done = [&](int sl) {
    for (int i = 0; i < n; ++i) {
      const int ph = Q->h_status[sl * n + i].phase;
      if (Q->busy[i] && ph != PH_DONE && ph != PH_IDLE) return false;
    }
    return true;
  };
  // one chunk always queued behind the snapshot the host is inspecting
  int rc = sched_issue(Q, chunk, s);
  if (!rc) rc = snap(0);
  int slot = 0;
  while (!rc) {
    if (Q->t - t0 < max_ticks) {
      rc = sched_issue(Q, chunk, s);
      if (!rc) rc = snap(slot ^ 1);
      if (rc) break;
    }
    HIPCHK(hipEventSynchronize(Q->ev[slot]));
    if (all_done(slot) || Q->t - t0 >= max_ticks) break;
    rc = sched_note_status(Q, Q->h_status + slot * n, s);
    slot ^= 1;
  }
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(Q->h_status, Q->d_status, n * sizeof(SchedState), hipMemcpyDeviceToHost));
  if (ticks_run) *ticks_run = Q->t - t0;
  if (out)
    for (int i = 0; i < n; ++i) to_public(Q->h_status[i], &out[i]);
  return GP_OK;
}

}  // extern "C"
