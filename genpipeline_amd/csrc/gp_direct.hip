// Exact FP64 direct solve of the free-DOF system K(xPhys) u = f on the device:
// the reference CPU solver's method (SciPy SuperLU, simp_solver.py:195-213).
// It works with element tables that no Krylov method converges on, such as
// the reference's own asymmetric, indefinite table (simp_solver.py:284-1543).
//
// Method: banded LU with partial pivoting (row interchanges), right-looking,
// one column per launch. Free DOFs are numbered in DOF order, so the free-free
// block of a structured grid has lower and upper bandwidth
// bw <= 3 (sI + sJ + 1) + 2 (sI, sJ: node strides). Interchanges widen the
// upper band of U to 2 bw.
//
// Storage: row r of the band lives in a physical slot of W = 2 bw + 2
// doubles; column j of that row sits at position j mod W. A row active at
// step k holds the live columns [k, k + 2 bw] (2 bw + 1 positions); the one
// spare position is that of column k - 1, which step k zeroes so that column
// k + 2 bw + 1 starts from zero when fill reaches it at step k + 1.
// Interchanges move no data: a map from logical row to physical slot is
// updated instead (two maps, read one / write the other per step, so blocks
// of a step never race on it). The right-hand side is eliminated alongside
// (it is stored per physical slot), so L is never kept: one factorisation per
// right-hand side, as SuperLU's spsolve does.
//
// Every value is produced by one thread in a fixed order: the result is the
// same bits on every call. Data in HBM for nf free DOFs: nf W doubles of band
// plus O(nf) vectors.
#include "gp_bodies.hpp"

namespace gp {

namespace {

constexpr int kStepBlock = 256;  // 4 waves, one band row each
constexpr int kSubBlock = 1024;  // back substitution: one workgroup
constexpr int kStepStage = 32;   // band columns a lane of the LU step holds at once
constexpr int kSubStage = 16;    // ... and of the back substitution

__global__ void k_zero_f64(double* __restrict__ v, long long n) {
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    v[i] = 0.0;
}

__global__ void k_iota2(int* __restrict__ m, int nf) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nf) {
    m[i] = i;
    m[nf + i] = i;
  }
}

// Entry (row (n,r), column (m,c)) of the free-free block, m in the 27-node
// neighbourhood of n: sum over the elements holding both nodes of
// E(xPhys_e) ke[3 Ln + r][3 Lm + c] (y_n = sum_e E_e ke u_e, the orientation
// of every other operator in the library), elements in a fixed order.
__global__ __launch_bounds__(kBlock) void k_band_assemble(Grid g, const double* __restrict__ ke,
                                                          const double* __restrict__ xPhys,
                                                          int law, double penal, double E0,
                                                          double Emin,
                                                          const int* __restrict__ fi, int W,
                                                          double* __restrict__ band) {
  const long long q = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (q >= (long long)g.nn * 243) return;
  const int rc = (int)(q % 9), o = (int)((q / 9) % 27);
  const int n = (int)(q / 243);
  const int r = rc / 3, c = rc % 3;
  const int row = fi[3ll * n + r];
  if (row < 0) return;
  const int i = n / g.sI, rem = n - i * g.sI, j = rem / g.sJ, k = rem - j * g.sJ;
  const int mi = i + o / 9 - 1, mj = j + (o / 3) % 3 - 1, mk = k + o % 3 - 1;
  if ((unsigned)mi > (unsigned)g.nx || (unsigned)mj > (unsigned)g.ny ||
      (unsigned)mk > (unsigned)g.nz)
    return;
  const int col = fi[3ll * (mi * g.sI + mj * g.sJ + mk) + c];
  if (col < 0) return;
  double s = 0.0;
  bool any = false;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int cc = 0; cc < 2; ++cc) {
        // element (i-1+a, j-1+b, k-1+cc) holds node n at corner (1-a, 1-b, 1-cc)
        const int ei = i - 1 + a, ej = j - 1 + b, ek = k - 1 + cc;
        if ((unsigned)ei >= (unsigned)g.nx || (unsigned)ej >= (unsigned)g.ny ||
            (unsigned)ek >= (unsigned)g.nz)
          continue;
        const int mx = mi - ei, my = mj - ej, mz = mk - ek;  // m's corner in that element
        if ((unsigned)mx > 1u || (unsigned)my > 1u || (unsigned)mz > 1u) continue;
        const int Ln = LOC[(1 - a) * 4 + (1 - b) * 2 + (1 - cc)];
        const int Lm = LOC[mx * 4 + my * 2 + mz];
        const double Ee = law_E(xPhys[ei * g.eI + ej * g.eJ + ek], law, penal, E0, Emin);
        s += Ee * ke[(3 * Ln + r) * 24 + 3 * Lm + c];
        any = true;
      }
  if (any) band[(long long)row * W + col % W] = s;
}

__global__ void k_gather_free(const double* __restrict__ f, const int* __restrict__ fi,
                              long long ndof, double* __restrict__ bf) {
  const long long d = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (d < ndof && fi[d] >= 0) bf[fi[d]] = f[d];
}

__global__ void k_scatter_free(const double* __restrict__ xf, const int* __restrict__ fi,
                               long long ndof, double* __restrict__ u) {
  const long long d = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (d < ndof) u[d] = fi[d] >= 0 ? xf[fi[d]] : 0.0;
}

// Column 0 of rows 0..min(bw, nf-1): the candidates of the first pivot.
__global__ void k_cand0(const double* __restrict__ band, int W, int nf, int bw,
                        double* __restrict__ cand, int* __restrict__ info) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t == 0) *info = 0;
  if (t <= bw && t < nf) cand[t] = band[(long long)t * W];
}

// Step k: pick the pivot among logical rows k..kend (the first largest
// |column k| value; cand[t] holds column k of row k + t), record the slot
// that now holds U row k, and subtract l_i times the pivot row from each
// row i in (k, kend] and from its right-hand side. One wave per row; wave gw
// always takes the rows i = gw (mod 4 gridDim.x), so a row stays with one
// workgroup (and one L2) for the steps it is active in. A wave loads its
// row's first kStepStage * 64 columns before the pivot search, so that load
// overlaps the search and the pivot-row load; the one row that trades places
// with the pivot row loads again from its new slot.
__global__ __launch_bounds__(kStepBlock) void k_lu_step(
    double* __restrict__ band, int W, int nf, int bw, int k, const int* __restrict__ sm,
    int* __restrict__ dm, int* __restrict__ urow, double* __restrict__ b,
    const double* __restrict__ cand, double* __restrict__ cand_next, int* __restrict__ info) {
  extern __shared__ double s_u[];  // pivot row, columns k+1 .. kendc
  __shared__ double s_v[kStepBlock / 64];
  __shared__ int s_t[kStepBlock / 64];
  const int uw = 2 * bw;
  const int kend = min(k + bw, nf - 1), kendc = min(k + uw, nf - 1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // columns k+1+c, c < ncols, sit at positions (k+1+c) mod W; positions
  // advance incrementally by (stride mod W), which wraps at most once
  const int ncols = kendc - k;
  const int wrap = 64 % W;
  const int pos0 = (k + 1 + lane) % W;
  const int nw = 4 * gridDim.x;
  const int gw = blockIdx.x * 4 + wave;
  const int i = k + 1 + ((gw - (k + 1)) % nw + nw) % nw;
  const bool own = i <= kend;
  const int slot_i = own ? sm[i] : 0;
  double v[kStepStage];
  if (own) {
    const double* row = band + (long long)slot_i * W;
    int pos = pos0;
#pragma unroll
    for (int t = 0; t < kStepStage; ++t) {
      v[t] = lane + 64 * t < ncols ? row[pos] : 0.0;
      pos += wrap;
      if (pos >= W) pos -= W;
    }
  }
  // pivot search (every workgroup, same order: the same answer)
  double best = -1.0;
  int bt = 0;
  for (int t = threadIdx.x; t <= kend - k; t += kStepBlock) {
    const double cv = fabs(cand[t]);
    if (cv > best) {
      best = cv;
      bt = t;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(best, off, 64);
    const int ot = __shfl_xor(bt, off, 64);
    if (ov > best || (ov == best && ot < bt)) {
      best = ov;
      bt = ot;
    }
  }
  if (lane == 0) {
    s_v[wave] = best;
    s_t[wave] = bt;
  }
  __syncthreads();
  best = s_v[0];
  bt = s_t[0];
  for (int w = 1; w < kStepBlock / 64; ++w)
    if (s_v[w] > best || (s_v[w] == best && s_t[w] < bt)) {
      best = s_v[w];
      bt = s_t[w];
    }
  const int p = k + bt;
  const double piv = cand[bt];
  const int slot_k = sm[k], slot_p = sm[p];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    urow[k] = slot_p;
    if (!(fabs(piv) > 0.0) && *info == 0) *info = k + 1;  // zero or NaN pivot
    if (k + 1 + bw <= nf - 1)  // the row entering the window at step k+1
      cand_next[bw] = band[(long long)(k + 1 + bw) * W + (k + 1) % W];
  }
  if (!(fabs(piv) > 0.0)) return;  // uniform: every thread read the same piv
  const double* urow_p = band + (long long)slot_p * W;
  {
    const int swrap = kStepBlock % W;
    int pos = (k + 1 + (int)threadIdx.x) % W;
    for (int c = threadIdx.x; c < ncols; c += kStepBlock) {
      s_u[c] = urow_p[pos];
      pos += swrap;
      if (pos >= W) pos -= W;
    }
  }
  const double bp = b[slot_p];
  __syncthreads();
  if (!own) return;
  const int slot = i == p ? slot_k : slot_i;
  double* row = band + (long long)slot * W;
  if (i == p) {  // wave-uniform
    int pos = pos0;
#pragma unroll
    for (int t = 0; t < kStepStage; ++t) {
      v[t] = lane + 64 * t < ncols ? row[pos] : 0.0;
      pos += wrap;
      if (pos >= W) pos -= W;
    }
  }
  const double a = i == p ? cand[0] : cand[i - k];
  const double l = a / piv;
  if (lane == 0 && k >= 1) row[(k - 1) % W] = 0.0;
  int pos = pos0;
#pragma unroll
  for (int t = 0; t < kStepStage; ++t) {
    const int c = lane + 64 * t;
    if (l != 0.0 && c < ncols) {
      v[t] = v[t] - l * s_u[c];
      row[pos] = v[t];
    }
    pos += wrap;
    if (pos >= W) pos -= W;
  }
  const double first = v[0];  // column k+1 after the update (lane 0)
  // columns beyond the prefetched ones (bands wider than 64 kStepStage):
  // loaded kStepStage at a time before any of them is stored
  for (int c0 = 64 * kStepStage; c0 < ncols; c0 += 64 * kStepStage) {
    int ps[kStepStage];
#pragma unroll
    for (int t = 0; t < kStepStage; ++t) {
      ps[t] = pos;
      v[t] = c0 + lane + 64 * t < ncols ? row[pos] : 0.0;
      pos += wrap;
      if (pos >= W) pos -= W;
    }
    if (l != 0.0) {
#pragma unroll
      for (int t = 0; t < kStepStage; ++t) {
        const int c = c0 + lane + 64 * t;
        if (c < ncols) row[ps[t]] = v[t] - l * s_u[c];
      }
    }
  }
  if (lane == 0) {
    if (l != 0.0) b[slot] = b[slot] - l * bp;
    dm[i] = slot;
    cand_next[i - k - 1] = first;
  }
}

// U x = y from the bottom row up, 64 rows (a tile [k0, k1]) at a time. For
// each row r of the tile, the part of its dot over the columns already solved
// (k1 < j <= r + 2 bw) is summed by the 16 waves in parallel (4 rows each;
// lane-strided, then a fixed shuffle tree); then wave 0 solves the tile's
// 64 x 64 upper triangle from LDS, lane q holding row k0 + q and adding
// U[r][j] x_j for j from k1 down to r + 1 as each x_j is found. x is read
// back from global memory by the same workgroup after a barrier.
__global__ __launch_bounds__(kSubBlock) void k_lu_backsub(const double* __restrict__ band, int W,
                                                          int nf, int bw,
                                                          const int* __restrict__ urow,
                                                          const double* __restrict__ b,
                                                          double* __restrict__ x) {
  __shared__ double s_tri[64][65];  // U[k0+q][k0+c], 0 outside the band or below the diagonal
  __shared__ double s_known[64];
  __shared__ int s_slot[64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int uw = 2 * bw;
  for (int k1 = nf - 1; k1 >= 0; k1 -= 64) {
    const int k0 = max(0, k1 - 63), nt = k1 - k0 + 1;
    if (threadIdx.x < nt) s_slot[threadIdx.x] = urow[k0 + threadIdx.x];
    __syncthreads();
    // the tile's triangle
    for (int e = threadIdx.x; e < 64 * 64; e += kSubBlock) {
      const int q = e >> 6, c = e & 63;
      double v = 0.0;
      if (q < nt && c < nt && c >= q && c - q <= uw)
        v = band[(long long)s_slot[q] * W + (k0 + c) % W];
      s_tri[q][c] = v;
    }
    // the solved part of each row's dot
    for (int q = wave; q < nt; q += kSubBlock / 64) {
      const int r = k0 + q;
      const int jend = min(r + uw, nf - 1);
      const double* row = band + (long long)s_slot[q] * W;
      double acc = 0.0;
      for (int j0 = k1 + 1; j0 <= jend; j0 += 64 * kSubStage) {
        double u[kSubStage], xv[kSubStage];
#pragma unroll
        for (int t = 0; t < kSubStage; ++t) {
          const int j = j0 + lane + 64 * t;
          u[t] = j <= jend ? row[j % W] : 0.0;
          xv[t] = j <= jend ? x[j] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < kSubStage; ++t) acc += u[t] * xv[t];
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
      if (lane == 0) s_known[q] = acc;
    }
    __syncthreads();
    if (wave == 0) {
      const int q = lane;
      double acc = q < nt ? s_known[q] : 0.0;
      const double y = q < nt ? b[s_slot[q]] : 0.0;
      double xq = 0.0;
      for (int c = nt - 1; c >= 0; --c) {
        if (q == c) xq = (y - acc) / s_tri[q][q];
        const double xc = __shfl(xq, c, 64);
        if (q < c) acc += s_tri[q][c] * xc;
      }
      if (q < nt) x[k0 + q] = xq;
    }
    __syncthreads();
  }
}

}  // namespace

int direct_band_width(const Grid& g) { return 3 * (g.sI + g.sJ + 1) + 2; }

void launch_direct(const Grid& g, const double* ke, const double* xPhys, int law, double penal,
                   double E0, double Emin, const double* f, const DirectWs& w, double* u,
                   hipStream_t st) {
  const int nf = w.nf, bw = w.bw, W = w.W;
  if (nf == 0) {
    hipLaunchKernelGGL(k_zero_f64, dim3(grid_blocks(g.ndof)), dim3(kBlock), 0, st, u,
                       (long long)g.ndof);
    return;
  }
  const long long nb = (long long)nf * W;
  hipLaunchKernelGGL(k_zero_f64, dim3(std::min<long long>(grid_blocks(nb), 8192)), dim3(kBlock), 0,
                     st, w.band, nb);
  hipLaunchKernelGGL(k_iota2, dim3(grid_blocks(nf)), dim3(kBlock), 0, st, w.map, nf);
  hipLaunchKernelGGL(k_band_assemble, dim3(grid_blocks((long long)g.nn * 243)), dim3(kBlock), 0,
                     st, g, ke, xPhys, law, penal, E0, Emin, w.fi, W, w.band);
  hipLaunchKernelGGL(k_gather_free, dim3(grid_blocks(g.ndof)), dim3(kBlock), 0, st, f, w.fi,
                     (long long)g.ndof, w.b);
  hipLaunchKernelGGL(k_cand0, dim3(grid_blocks(bw + 1)), dim3(kBlock), 0, st, w.band, W, nf, bw,
                     w.cand, w.info);
  const int blocks = (bw + 3) / 4;
  const size_t lds = sizeof(double) * (size_t)(2 * bw);
  for (int k = 0; k < nf; ++k) {
    const int* sm = w.map + (k & 1) * nf;
    int* dm = w.map + ((k + 1) & 1) * nf;
    const double* c0 = w.cand + (k & 1) * (bw + 1);
    double* c1 = w.cand + ((k + 1) & 1) * (bw + 1);
    hipLaunchKernelGGL(k_lu_step, dim3(blocks), dim3(kStepBlock), lds, st, w.band, W, nf, bw, k,
                       sm, dm, w.urow, w.b, c0, c1, w.info);
  }
  hipLaunchKernelGGL(k_lu_backsub, dim3(1), dim3(kSubBlock), 0, st, w.band, W, nf, bw, w.urow,
                     w.b, w.x);
  hipLaunchKernelGGL(k_scatter_free, dim3(grid_blocks(g.ndof)), dim3(kBlock), 0, st, w.x, w.fi,
                     (long long)g.ndof, u);
}

}  // namespace gp
