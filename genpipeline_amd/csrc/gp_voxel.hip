// Decoded-field pre-processing of the FEM evaluator, on the device.
//
// The evaluator turns a decoder output (B fields of d x h x w float32, C
// order) into the voxel densities it solves on (genpipeline/fem/
// voxel_fem.py:452-462, genpipeline/pipeline_utils.py:55-76):
//   v = scipy.ndimage.zoom(v, 0.5, order=1)            (float32 out)
//   v = scipy.ndimage.gaussian_filter(v, 0.5)          (float32 after each axis)
//   v = 1 / (1 + exp(-15 (v - 0.5)))                   (float32 arithmetic)
// Each kernel restates one of those steps with scipy's own order of
// operations (scipy 1.15, ndimage/src/ni_interpolation.c NI_ZoomShift and
// ni_filters.c NI_Correlate1D), in double where scipy works in double, and
// rounds to float where scipy stores float32. Contraction of a multiply and
// an add into one rounding is off in this file: scipy rounds each.
//
// The work is byte-streaming (a 64^3 field is 1 MiB in, 128 KiB out) and
// small; one thread per output voxel, last axis fastest, so every pass reads
// and writes whole cache lines.
#include <cmath>

#include "gp_internal.hpp"

#pragma clang fp contract(off)

namespace gp {

namespace {

// Index of position q on an axis of n samples under scipy's 'reflect'
// extension (d c b a | a b c d | d c b a): period 2n, mirrored halves.
__device__ __forceinline__ int reflect_index(int q, int n) {
  if (n == 1) return 0;
  const int p = 2 * n;
  q %= p;
  if (q < 0) q += p;
  return q >= n ? p - 1 - q : q;
}

// Order-1 zoom (grid_mode False, mode 'constant', cval 0). Output index o of
// an axis of m samples maps to c = o * ((n-1)/(m-1)); a c beyond n-1 (only
// possible by rounding of that product) gives cval; otherwise s = floor(c),
// t = c - s and the axis weights are (1-t, t). The eight samples are taken
// in footprint order (last axis fastest), each multiplied by the axis-0,
// axis-1, axis-2 weights in that order, and summed from 0.0 in double.
__global__ __launch_bounds__(kBlock) void k_zoom1(const float* __restrict__ in,
                                                  float* __restrict__ out, long long total, int d,
                                                  int h, int w, int od, int oh, int ow, double z0,
                                                  double z1, double z2) {
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(idx % ow);
    const long long r1 = idx / ow;
    const int j = (int)(r1 % oh);
    const long long r2 = r1 / oh;
    const int i = (int)(r2 % od);
    const long long b = r2 / od;
    const double c0 = (double)i * z0, c1 = (double)j * z1, c2 = (double)k * z2;
    if (c0 > (double)(d - 1) || c1 > (double)(h - 1) || c2 > (double)(w - 1)) {
      out[idx] = 0.0f;
      continue;
    }
    const int s0 = (int)floor(c0), s1 = (int)floor(c1), s2 = (int)floor(c2);
    const double t0 = c0 - s0, t1 = c1 - s1, t2 = c2 - s2;
    const double w0[2] = {1.0 - t0, t0}, w1[2] = {1.0 - t1, t1}, w2[2] = {1.0 - t2, t2};
    // the upper sample past the last plane only occurs with weight t = 0
    const int i0[2] = {s0, s0 + 1 < d ? s0 + 1 : d - 1};
    const int j0[2] = {s1, s1 + 1 < h ? s1 + 1 : h - 1};
    const int k0[2] = {s2, s2 + 1 < w ? s2 + 1 : w - 1};
    const float* src = in + b * (long long)d * h * w;
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          double coeff = (double)src[((long long)i0[a] * h + j0[c]) * w + k0[e]];
          coeff = coeff * w0[a];
          coeff = coeff * w1[c];
          coeff = coeff * w2[e];
          acc = acc + coeff;
        }
    out[idx] = (float)acc;
  }
}

// One axis of a separable symmetric filter of 2r+1 taps (the Gaussian of
// gaussian_filter), 'reflect' boundaries:
//   out = v[0] w[0] + sum_{j=-r..-1} (v[j] + v[-j]) w[j]
// accumulated in double in that order, rounded to float. With LOGISTIC the
// rounded value then goes through 1 / (1 + exp(a)), a = (-sharpness) (v - 0.5),
// every step in float as numpy computes it on a float32 array; exp is
// evaluated in double and rounded once.
struct Taps {
  double w[kMaxTaps];  // w[r + j], j = -r..r
  int r;
};

template <int AX, bool LOGISTIC>
__global__ __launch_bounds__(kBlock) void k_filter_axis(const float* __restrict__ in,
                                                        float* __restrict__ out, long long total,
                                                        int d, int h, int w, Taps tp,
                                                        float neg_sharp) {
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(idx % w);
    const long long r1 = idx / w;
    const int j = (int)(r1 % h);
    const long long r2 = r1 / h;
    const int i = (int)(r2 % d);
    const int pos = AX == 0 ? i : (AX == 1 ? j : k);
    const int n = AX == 0 ? d : (AX == 1 ? h : w);
    const long long stride = AX == 0 ? (long long)h * w : (AX == 1 ? (long long)w : 1LL);
    const float* line = in + idx - (long long)pos * stride;  // sample 0 of this line
    double acc = (double)line[(long long)pos * stride] * tp.w[tp.r];
    for (int jj = -tp.r; jj < 0; ++jj) {
      const double lo = (double)line[(long long)reflect_index(pos + jj, n) * stride];
      const double hi = (double)line[(long long)reflect_index(pos - jj, n) * stride];
      acc = acc + (lo + hi) * tp.w[tp.r + jj];
    }
    float v = (float)acc;
    if (LOGISTIC) {
      const float a = neg_sharp * (v - 0.5f);
      const float e = (float)exp((double)a);
      v = 1.0f / (1.0f + e);
    }
    out[idx] = v;
  }
}

int grid_for(long long total) {
  const long long nb = (total + kBlock - 1) / kBlock;
  return (int)(nb < 65535 ? (nb > 0 ? nb : 1) : 65535);
}

}  // namespace

void launch_voxel_prep(const float* in, int batch, int d, int h, int w, int od, int oh, int ow,
                       const double* weights, int radius, double sharpness, float* work,
                       float* out, hipStream_t st) {
  const long long total = (long long)batch * od * oh * ow;
  const double z0 = od > 1 ? (double)(d - 1) / (double)(od - 1) : 1.0;
  const double z1 = oh > 1 ? (double)(h - 1) / (double)(oh - 1) : 1.0;
  const double z2 = ow > 1 ? (double)(w - 1) / (double)(ow - 1) : 1.0;
  Taps tp{};
  tp.r = radius;
  for (int q = 0; q <= 2 * radius; ++q) tp.w[q] = weights[q];
  const float ns = (float)(-sharpness);
  const int nb = grid_for(total);
  // zoom -> work; axis 0: work -> out; axis 1: out -> work; axis 2 + logistic: work -> out
  hipLaunchKernelGGL(k_zoom1, dim3(nb), dim3(kBlock), 0, st, in, work, total, d, h, w, od, oh,
                     ow, z0, z1, z2);
  hipLaunchKernelGGL((k_filter_axis<0, false>), dim3(nb), dim3(kBlock), 0, st, work, out, total,
                     od, oh, ow, tp, ns);
  hipLaunchKernelGGL((k_filter_axis<1, false>), dim3(nb), dim3(kBlock), 0, st, out, work, total,
                     od, oh, ow, tp, ns);
  hipLaunchKernelGGL((k_filter_axis<2, true>), dim3(nb), dim3(kBlock), 0, st, work, out, total,
                     od, oh, ow, tp, ns);
}

}  // namespace gp
