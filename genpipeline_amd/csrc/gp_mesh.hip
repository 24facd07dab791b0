// Isosurface of a density field as triangles (marching cubes), on the device.
//
// The post-step of the data generator writes each design as an STL surface
// at density 0.5 (genpipeline/topology/mesh_export.py:7-67, which calls
// skimage.measure.marching_cubes). Here one thread takes one cube of 2x2x2
// samples: corner c is inside when its sample is > level, the 8 bits give the
// case, and the case's triangles join vertices on the crossing edges, each at
// the linear crossing point measured from the edge's lower-coordinate end.
// The 256-case triangle table is generated at first use from a rule (see
// build_tables) rather than typed in. Output is one STL record per triangle:
// unit normal (right-hand rule; 0 for a degenerate triangle), then the three
// vertices, scaled to physical units, in float64 arithmetic rounded once to
// float32. Triangles come in C order of their cube, each cube's in table
// order, so the result is deterministic.
//
// Three launches: per-block triangle counts (256 cubes a block), one block
// scanning those counts into offsets, and the emitting pass.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "gp_internal.hpp"

#pragma clang fp contract(off)

namespace gp {

namespace {

constexpr int kMcMaxTri = 5;  // the generated table needs at most 5 a cube

__constant__ int8_t c_mc_ntri[256];
__constant__ int8_t c_mc_tri[256][3 * kMcMaxTri];

// corner offsets (C3D8 order) and the 12 cube edges
constexpr int kCx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
constexpr int kCy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
constexpr int kCz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
constexpr int kEa[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
constexpr int kEb[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};

struct McTables {
  int8_t ntri[256];
  int8_t tri[256][3 * kMcMaxTri];
};

int edge_of(int a, int b) {
  for (int e = 0; e < 12; ++e)
    if ((kEa[e] == a && kEb[e] == b) || (kEa[e] == b && kEb[e] == a)) return e;
  return -1;
}

// The six faces as corner cycles, counter-clockwise seen from outside.
void build_faces(int (&face)[6][4]) {
  for (int axis = 0, f = 0; axis < 3; ++axis)
    for (int side = 0; side < 2; ++side, ++f) {
      int cs[4], n = 0;
      for (int c = 0; c < 8; ++c) {
        const int co = axis == 0 ? kCx[c] : (axis == 1 ? kCy[c] : kCz[c]);
        if (co == side) cs[n++] = c;
      }
      // angle of each corner around the outward normal, with u = next axis
      // and v = normal x u
      const int ua = (axis + 1) % 3, va = (axis + 2) % 3;
      const double sgn = side ? 1.0 : -1.0;
      double ang[4];
      for (int i = 0; i < 4; ++i) {
        const int c = cs[i];
        const double p[3] = {kCx[c] - 0.5, kCy[c] - 0.5, kCz[c] - 0.5};
        ang[i] = std::atan2(sgn * p[va], p[ua]);
      }
      for (int i = 0; i < 4; ++i) {  // sort by angle
        int best = i;
        for (int j = i + 1; j < 4; ++j)
          if (ang[j] < ang[best]) best = j;
        std::swap(ang[i], ang[best]);
        std::swap(cs[i], cs[best]);
      }
      for (int i = 0; i < 4; ++i) face[f][i] = cs[i];
    }
}

// Case triangulation: on each face the crossings are joined so that the
// inside corners lie on the left of each segment walking the face
// counter-clockwise from outside (a face with two diagonal inside corners
// keeps them apart); the segments close into loops, each fanned from its
// first crossing against the loop direction, so normals point from inside
// to outside. Neighbouring cubes join a shared face's crossings alike.
bool build_tables(McTables& t) {
  int face[6][4];
  build_faces(face);
  for (int cs = 0; cs < 256; ++cs) {
    bool in[8];
    for (int c = 0; c < 8; ++c) in[c] = (cs >> c) & 1;
    int nxt[12];
    for (int e = 0; e < 12; ++e) nxt[e] = -1;
    for (int f = 0; f < 6; ++f) {
      int ce[4], ent[4], m = 0;
      for (int i = 0; i < 4; ++i) {
        const int a = face[f][i], b = face[f][(i + 1) % 4];
        if (in[a] != in[b]) {
          ce[m] = edge_of(a, b);
          ent[m] = in[b] ? 1 : 0;
          ++m;
        }
      }
      for (int i = 0; i < m; ++i)
        if (ent[i]) {
          const int q = ce[(i + 1) % m];
          if (nxt[q] != -1) return false;
          nxt[q] = ce[i];
        }
    }
    int n = 0;
    bool seen[12] = {};
    for (int s = 0; s < 12; ++s) {
      if (nxt[s] < 0 || seen[s]) continue;
      int loop[12], len = 0;
      int e = s;
      do {
        loop[len++] = e;
        seen[e] = true;
        e = nxt[e];
      } while (e != s && len < 12);
      if (e != s) return false;
      for (int i = 1; i + 1 < len; ++i) {
        if (n >= kMcMaxTri) return false;
        t.tri[cs][3 * n] = (int8_t)loop[0];
        t.tri[cs][3 * n + 1] = (int8_t)loop[i + 1];
        t.tri[cs][3 * n + 2] = (int8_t)loop[i];
        ++n;
      }
    }
    for (int i = 3 * n; i < 3 * kMcMaxTri; ++i) t.tri[cs][i] = -1;
    t.ntri[cs] = (int8_t)n;
  }
  return true;
}

__device__ inline int cube_case(const double* __restrict__ d, int i, int j, int k, int ny,
                                int nz, double level, double (&v)[8]) {
  int cs = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    v[c] = d[((long long)(i + kCx[c]) * ny + (j + kCy[c])) * nz + (k + kCz[c])];
    if (v[c] > level) cs |= 1 << c;
  }
  return cs;
}

__device__ inline void cube_of(long long q, int cy, int cz, int& i, int& j, int& k) {
  k = (int)(q % cz);
  const long long r = q / cz;
  j = (int)(r % cy);
  i = (int)(r / cy);
}

// Triangle count of each 256-cube block.
__global__ __launch_bounds__(kBlock) void k_mc_count(const double* __restrict__ d, int nx, int ny,
                                                     int nz, double level,
                                                     int* __restrict__ block_cnt) {
  const int cy = ny - 1, cz = nz - 1;
  const long long ncube = (long long)(nx - 1) * cy * cz;
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  int n = 0;
  if (q < ncube) {
    int i, j, k;
    cube_of(q, cy, cz, i, j, k);
    double v[8];
    n = c_mc_ntri[cube_case(d, i, j, k, ny, nz, level, v)];
  }
  __shared__ int sh[kBlock];
  sh[threadIdx.x] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int t = 0; t < kBlock; ++t) s += sh[t];
    block_cnt[blockIdx.x] = s;
  }
}

// Exclusive scan of nb block counts (in place) into offsets; off[nb] = total.
// One block: thread t sums a contiguous chunk, the chunk sums are scanned in
// LDS, then each chunk is rewritten.
__global__ __launch_bounds__(kBlock) void k_mc_scan(int* __restrict__ off, int nb) {
  __shared__ int sh[kBlock];
  const int t = threadIdx.x;
  const int chunk = (nb + kBlock - 1) / kBlock;
  const int b0 = t * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
  int s = 0;
  for (int b = b0; b < b1; ++b) s += off[b];
  sh[t] = s;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int u = 0; u < kBlock; ++u) {
      const int x = sh[u];
      sh[u] = acc;
      acc += x;
    }
    off[nb] = acc;
  }
  __syncthreads();
  int acc = sh[t];
  for (int b = b0; b < b1; ++b) {
    const int x = off[b];
    off[b] = acc;
    acc += x;
  }
}

// Emit the triangles of each cube at its block offset plus its rank in the
// block.
__global__ __launch_bounds__(kBlock) void k_mc_emit(const double* __restrict__ d, int nx, int ny,
                                                    int nz, double level, double sx, double sy,
                                                    double sz, const int* __restrict__ off,
                                                    float* __restrict__ out) {
  const int cy = ny - 1, cz = nz - 1;
  const long long ncube = (long long)(nx - 1) * cy * cz;
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  int n = 0, cs = 0, i = 0, j = 0, k = 0;
  double v[8];
  if (q < ncube) {
    cube_of(q, cy, cz, i, j, k);
    cs = cube_case(d, i, j, k, ny, nz, level, v);
    n = c_mc_ntri[cs];
  }
  __shared__ int sh[kBlock];
  sh[threadIdx.x] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int t = 0; t < kBlock; ++t) {
      const int x = sh[t];
      sh[t] = acc;
      acc += x;
    }
  }
  __syncthreads();
  long long slot = (long long)off[blockIdx.x] + sh[threadIdx.x];
  for (int t = 0; t < n; ++t, ++slot) {
    double p[3][3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int e = c_mc_tri[cs][3 * t + u];
      int a = kEa[e], b = kEb[e];
      if (kCx[a] + kCy[a] + kCz[a] > kCx[b] + kCy[b] + kCz[b]) {  // from the lower end
        const int s = a;
        a = b;
        b = s;
      }
      const double s = (level - v[a]) / (v[b] - v[a]);
      const double pa[3] = {(double)(i + kCx[a]), (double)(j + kCy[a]), (double)(k + kCz[a])};
      const double pb[3] = {(double)(i + kCx[b]), (double)(j + kCy[b]), (double)(k + kCz[b])};
      const double sc[3] = {sx, sy, sz};
#pragma unroll
      for (int c = 0; c < 3; ++c) p[u][c] = (pa[c] + s * (pb[c] - pa[c])) * sc[c];
    }
    const double ax = p[1][0] - p[0][0], ay = p[1][1] - p[0][1], az = p[1][2] - p[0][2];
    const double bx = p[2][0] - p[0][0], by = p[2][1] - p[0][1], bz = p[2][2] - p[0][2];
    double nx_ = ay * bz - az * by, ny_ = az * bx - ax * bz, nz_ = ax * by - ay * bx;
    const double ln = sqrt(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
    if (ln > 0.0) {
      nx_ = nx_ / ln;
      ny_ = ny_ / ln;
      nz_ = nz_ / ln;
    } else {
      nx_ = ny_ = nz_ = 0.0;
    }
    float* r = out + 12 * slot;
    r[0] = (float)nx_;
    r[1] = (float)ny_;
    r[2] = (float)nz_;
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) r[3 + 3 * u + c] = (float)p[u][c];
  }
}

// ------------------------------------------------- voxel hex mesh (C3D8)
// The CalculiX path of the evaluator meshes every solid voxel (value > 0.5)
// as one C3D8 element, in C order of the voxels, and numbers the grid
// corners in the order the elements first touch them, corners in C3D8 order
// (genpipeline/fem/voxel_fem.py:120-146). A corner's number is therefore
// taken by the solid voxel of smallest C index among the up to eight around
// it, and within that voxel by its corner position. Each pass is one thread
// per voxel; counts become offsets by the same one-block scan as the
// triangles.

__device__ inline bool solid_at(const float* __restrict__ v, int D, int H, int W, int i, int j,
                                int k) {
  if ((unsigned)i >= (unsigned)D || (unsigned)j >= (unsigned)H || (unsigned)k >= (unsigned)W)
    return false;
  return v[((long long)i * H + j) * W + k] > 0.5f;
}

// Corner c of solid voxel (i, j, k) is numbered by this voxel when no solid
// voxel of smaller C index shares that corner.
__device__ inline bool owns_corner(const float* __restrict__ v, int D, int H, int W, int i, int j,
                                   int k, int c) {
  const int X = i + kCx[c], Y = j + kCy[c], Z = k + kCz[c];
  const long long self = ((long long)i * H + j) * W + k;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int ui = X - 1 + a, uj = Y - 1 + b, uk = Z - 1 + e;
        if (!solid_at(v, D, H, W, ui, uj, uk)) continue;
        if (((long long)ui * H + uj) * W + uk < self) return false;
      }
  return true;
}

__device__ inline void voxel_of(long long q, int H, int W, int& i, int& j, int& k) {
  k = (int)(q % W);
  const long long r = q / W;
  j = (int)(r % H);
  i = (int)(r / H);
}

// Per 256-voxel block: solid voxels (cnt_e) and corners they number (cnt_n).
__global__ __launch_bounds__(kBlock) void k_hex_count(const float* __restrict__ v, int D, int H,
                                                      int W, int* __restrict__ cnt_e,
                                                      int* __restrict__ cnt_n) {
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  int ne = 0, nn = 0;
  if (q < (long long)D * H * W) {
    int i, j, k;
    voxel_of(q, H, W, i, j, k);
    if (v[q] > 0.5f) {
      ne = 1;
#pragma unroll
      for (int c = 0; c < 8; ++c) nn += owns_corner(v, D, H, W, i, j, k, c) ? 1 : 0;
    }
  }
  __shared__ int sh[2][kBlock];
  sh[0][threadIdx.x] = ne;
  sh[1][threadIdx.x] = nn;
  __syncthreads();
  if (threadIdx.x == 0) {
    int se = 0, sn = 0;
    for (int t = 0; t < kBlock; ++t) {
      se += sh[0][t];
      sn += sh[1][t];
    }
    cnt_e[blockIdx.x] = se;
    cnt_n[blockIdx.x] = sn;
  }
}

// Exclusive in-block ranks of a and b; returns them through ra, rb.
__device__ inline void block_ranks2(int a, int b, int& ra, int& rb) {
  __shared__ int sh[2][kBlock];
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    int sa = 0, sb = 0;
    for (int t = 0; t < kBlock; ++t) {
      const int xa = sh[0][t], xb = sh[1][t];
      sh[0][t] = sa;
      sh[1][t] = sb;
      sa += xa;
      sb += xb;
    }
  }
  __syncthreads();
  ra = sh[0][threadIdx.x];
  rb = sh[1][threadIdx.x];
}

// Numbers the corners (1-based, in first-touch order) into the node grid
// (D+1)(H+1)(W+1) and writes each new node's grid corner (3 ints); records
// every solid voxel's C index in element order.
__global__ __launch_bounds__(kBlock) void k_hex_nodes(const float* __restrict__ v, int D, int H,
                                                      int W, const int* __restrict__ off_e,
                                                      const int* __restrict__ off_n,
                                                      int* __restrict__ node_grid,
                                                      int* __restrict__ node_ijk,
                                                      int* __restrict__ elem_voxel) {
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  int i = 0, j = 0, k = 0, ne = 0, nn = 0, own = 0;
  if (q < (long long)D * H * W) {
    voxel_of(q, H, W, i, j, k);
    if (v[q] > 0.5f) {
      ne = 1;
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (owns_corner(v, D, H, W, i, j, k, c)) {
          own |= 1 << c;
          ++nn;
        }
    }
  }
  int re, rn;
  block_ranks2(ne, nn, re, rn);
  if (!ne) return;
  if (elem_voxel) elem_voxel[off_e[blockIdx.x] + re] = (int)q;
  int id = off_n[blockIdx.x] + rn;  // 0-based
#pragma unroll
  for (int c = 0; c < 8; ++c)
    if (own & (1 << c)) {
      const int X = i + kCx[c], Y = j + kCy[c], Z = k + kCz[c];
      node_grid[((long long)X * (H + 1) + Y) * (W + 1) + Z] = id + 1;
      node_ijk[3ll * id] = X;
      node_ijk[3ll * id + 1] = Y;
      node_ijk[3ll * id + 2] = Z;
      ++id;
    }
}

// Eight node numbers per solid voxel, C3D8 corner order.
__global__ __launch_bounds__(kBlock) void k_hex_elems(const float* __restrict__ v, int D, int H,
                                                      int W, const int* __restrict__ off_e,
                                                      const int* __restrict__ node_grid,
                                                      int* __restrict__ elems) {
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  int i = 0, j = 0, k = 0, ne = 0;
  if (q < (long long)D * H * W) {
    voxel_of(q, H, W, i, j, k);
    ne = v[q] > 0.5f ? 1 : 0;
  }
  int re, unused;
  block_ranks2(ne, 0, re, unused);
  if (!ne) return;
  const long long e = off_e[blockIdx.x] + re;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    elems[8 * e + c] =
        node_grid[((long long)(i + kCx[c]) * (H + 1) + (j + kCy[c])) * (W + 1) + (k + kCz[c])];
}

std::mutex g_mc_mu;
std::vector<int> g_mc_ready;  // per device: tables copied

}  // namespace

int hex_blocks(int D, int H, int W) {
  return (int)(((long long)D * H * W + kBlock - 1) / kBlock);
}

// Workspace: off_e[nb + 1], off_n[nb + 1], node grid (D+1)(H+1)(W+1).
void launch_hex_count(const float* v, int D, int H, int W, int* ws, hipStream_t st) {
  const int nb = hex_blocks(D, H, W);
  int* off_e = ws;
  int* off_n = ws + nb + 1;
  hipLaunchKernelGGL(k_hex_count, dim3(nb), dim3(kBlock), 0, st, v, D, H, W, off_e, off_n);
  hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(kBlock), 0, st, off_e, nb);
  hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(kBlock), 0, st, off_n, nb);
}

void launch_hex_emit(const float* v, int D, int H, int W, int* ws, int* elems, int* node_ijk,
                     int* elem_voxel, hipStream_t st) {
  const int nb = hex_blocks(D, H, W);
  const int* off_e = ws;
  const int* off_n = ws + nb + 1;
  int* grid = ws + 2 * (nb + 1);
  hipLaunchKernelGGL(k_hex_nodes, dim3(nb), dim3(kBlock), 0, st, v, D, H, W, off_e, off_n, grid,
                     node_ijk, elem_voxel);
  hipLaunchKernelGGL(k_hex_elems, dim3(nb), dim3(kBlock), 0, st, v, D, H, W, off_e, grid, elems);
}

int mc_blocks(int nx, int ny, int nz) {
  const long long ncube = (long long)(nx - 1) * (ny - 1) * (nz - 1);
  return (int)((ncube + kBlock - 1) / kBlock);
}

const char* mc_prepare(int device) {
  std::lock_guard<std::mutex> lk(g_mc_mu);
  if ((int)g_mc_ready.size() <= device) g_mc_ready.resize(device + 1, 0);
  if (g_mc_ready[device]) return nullptr;
  static McTables t;
  static const bool ok = build_tables(t);
  if (!ok) return "marching-cubes table generation failed";
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_mc_ntri), t.ntri, sizeof(t.ntri)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_mc_tri), t.tri, sizeof(t.tri)) != hipSuccess)
    return "copying the marching-cubes tables to the device failed";
  g_mc_ready[device] = 1;
  return nullptr;
}

void mc_tables_host(int8_t* ntri, int8_t* tri) {
  McTables t;
  build_tables(t);
  std::memcpy(ntri, t.ntri, sizeof(t.ntri));
  std::memcpy(tri, t.tri, sizeof(t.tri));
}

void launch_mc_count(const double* d, int nx, int ny, int nz, double level, int* off,
                     hipStream_t st) {
  const int nb = mc_blocks(nx, ny, nz);
  hipLaunchKernelGGL(k_mc_count, dim3(nb), dim3(kBlock), 0, st, d, nx, ny, nz, level, off);
  hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(kBlock), 0, st, off, nb);
}

void launch_mc_emit(const double* d, int nx, int ny, int nz, double level, double sx, double sy,
                    double sz, const int* off, float* out, hipStream_t st) {
  const int nb = mc_blocks(nx, ny, nz);
  hipLaunchKernelGGL(k_mc_emit, dim3(nb), dim3(kBlock), 0, st, d, nx, ny, nz, level, sx, sy, sz,
                     off, out);
}

}  // namespace gp
