// Device helpers shared by the SU(2), XYZ and TRI flow kernels: the frequency
// mesh (bracket searches, interpolation), the self-energy interpolation, the
// trapezoid evaluation points of ImplicitIntegrator, the per-cutoff tables of
// evaluation points and propagator weights, and the access-buffer descriptors
// (support rows, bilinear weights, density signs) of the SU(2) and XYZ vertex
// layouts.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace frg {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;

// ---------------------------------------------------------------------------
// frequency mesh (positive half m[0..n) ascending; signed index i in [-n, n))
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mesh_at(const double *m, int i) { return i >= 0 ? m[i] : -m[-i - 1]; }

__device__ inline int greater_pos(const double *m, int n, double w) {
  if (w <= m[0]) return 0;
  for (int i = 1; i < n; ++i)
    if (m[i] > w) return i;
  return n - 1;
}

__device__ inline int lesser_pos(const double *m, int n, double w) {
  if (w <= m[0]) return 0;
  for (int i = 1; i < n; ++i)
    if (m[i] > w) return i - 1;
  return n - 1;
}

// FrequencyDiscretization::lesser / greater (src/FrequencyDiscretization.hpp:253-296)
__device__ inline int mesh_lesser(const double *m, int n, double w) { return w < 0 ? -(greater_pos(m, n, -w) + 1) : lesser_pos(m, n, w); }
__device__ inline int mesh_greater(const double *m, int n, double w) { return w < 0 ? -(lesser_pos(m, n, -w) + 1) : greater_pos(m, n, w); }

// number of mesh points <= w (branch-free binary lifting)
__device__ __forceinline__ int count_le(const double *m, int n, double w) {
  int pos = 0;
  for (int step = 1 << (31 - __clz(n)); step > 0; step >>= 1) {
    const int p = pos + step;
    if (p <= n && m[p - 1] <= w) pos = p;
  }
  return pos;
}

// FrequencyDiscretization::interpolateOffset (src/FrequencyDiscretization.hpp:326-356)
// (written as !(w > m[0]) so that a NaN argument clamps to index 0 instead of
// producing index -1; identical to the reference for every non-NaN w)
__device__ __forceinline__ void interp(const double *m, int n, double w, int &lo, int &hi, double &bias) {
  if (!(w > m[0])) { lo = 0; hi = 0; bias = 0.0; return; }
  const int c = count_le(m, n, w);
  if (c >= n) { lo = n - 1; hi = n - 1; bias = 0.0; return; }
  lo = c - 1;
  hi = c;
  bias = (w - m[lo]) / (m[hi] - m[lo]);
}

// SU2VertexSingleParticle::getValue (src/SU2/SU2VertexSingleParticle.hpp:73-87)
__device__ inline double sigma_at(const double *m, int n, const double *data, double w) {
  double sign = 1.0;
  if (w < 0) { w = -w; sign = -1.0; }
  int lo, hi; double b;
  interp(m, n, w, lo, hi, b);
  return sign * ((1 - b) * data[lo] + b * data[hi]);
}

__device__ __forceinline__ int tri(int a, int b) { return a * (a + 1) / 2 + b; }

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// trapezoid windows: evaluation points and weights (0.5 folded in), in the
// order ImplicitIntegrator visits them (src/lib/Integrator.hpp:137-287)
// ---------------------------------------------------------------------------
struct PointList {
  double *wp;
  double *w;
  int cnt;
  int cap;
  __device__ inline void push(double x, double weight) {
    if (cnt < cap) { wp[cnt] = x; w[cnt] = weight; }
    ++cnt;
  }
};

__device__ inline void window_right(const double *m, int n, double max, PointList &o) {
  int umin = -n;
  const int umax = mesh_lesser(m, n, max);
  if (umin != umax) {
    o.push(mesh_at(m, umin), 0.5 * (mesh_at(m, umin + 1) - mesh_at(m, umin)));
    while (++umin != umax) o.push(mesh_at(m, umin), 0.5 * (mesh_at(m, umin + 1) - mesh_at(m, umin - 1)));
    o.push(mesh_at(m, umin), 0.5 * (max - mesh_at(m, umin - 1)));
    o.push(max, 0.5 * (max - mesh_at(m, umin)));
  } else {
    const double h = 0.5 * (max - mesh_at(m, umin));
    o.push(max, h);
    o.push(mesh_at(m, umin), h);
  }
}

__device__ inline void window_both(const double *m, int n, double min, double max, PointList &o) {
  int umin = mesh_greater(m, n, min);
  const int umax = mesh_lesser(m, n, max);
  if (umax >= umin) {
    o.push(min, 0.5 * (mesh_at(m, umin) - min));
    if (umax != umin) {
      o.push(mesh_at(m, umin), 0.5 * (mesh_at(m, umin + 1) - min));
      while (++umin != umax) o.push(mesh_at(m, umin), 0.5 * (mesh_at(m, umin + 1) - mesh_at(m, umin - 1)));
      o.push(mesh_at(m, umin), 0.5 * (max - mesh_at(m, umin - 1)));
    } else {
      o.push(mesh_at(m, umin), 0.5 * (max - min));
    }
    o.push(max, 0.5 * (max - mesh_at(m, umin)));
  } else {
    const double h = 0.5 * (max - min);
    o.push(max, h);
    o.push(min, h);
  }
}

__device__ inline void window_left(const double *m, int n, double min, PointList &o) {
  int umin = mesh_greater(m, n, min);
  const int max = n - 1;
  if (umin != max) {
    o.push(min, 0.5 * (mesh_at(m, umin) - min));
    o.push(mesh_at(m, umin), 0.5 * (mesh_at(m, umin + 1) - min));
    while (++umin != max) o.push(mesh_at(m, umin), 0.5 * (mesh_at(m, umin + 1) - mesh_at(m, umin - 1)));
    o.push(mesh_at(m, umin), 0.5 * (mesh_at(m, umin) - mesh_at(m, umin - 1)));
  } else {
    const double h = 0.5 * (mesh_at(m, umin) - min);
    o.push(min, h);
    o.push(mesh_at(m, umin), h);
  }
}

// the three Katanin windows of a channel with transfer frequency X
// (src/SU2/SU2FrgCore.cpp:378-392; with X = nu = 0 the correlation windows,
// src/SU2/SU2MeasurementCorrelation.cpp:145-159)
__device__ inline void katanin_windows(const double *m, int n, double X, double cut, PointList &o) {
  if (-(X + cut) > mesh_at(m, -n)) window_right(m, n, -(X + cut), o);
  if (X - cut > cut) window_both(m, n, cut - X, -cut, o);
  if (cut < mesh_at(m, n - 1)) window_left(m, n, cut, o);
}

// Evaluation points and weights of one channel with exact transfer frequency
// m[x] at cutoff cut: the conventional points +-cut with p(cut, cut +- X)
// (src/SU2/SU2FrgCore.cpp:349-371, src/TRI/TRIFrgCore.cpp:2917-2940), then the
// Katanin windows with dSigma(w') G(w')^2 G(X + w') folded into the trapezoid
// weight (src/SU2/SU2FrgCore.cpp:343-347,373-424).  Depends on (x, cut) only,
// so every item of a flow shares the table.
//
// weight_tables_wave: one 64-lane workgroup per table, x < nw a channel list,
// x == nw the plain trapezoid list of the correlation integrals at nu = 0
// (src/SU2/SU2MeasurementCorrelation.cpp:145-159,
// src/TRI/TRIMeasurementCorrelation.cpp:270-284).  s_m is LDS space for the
// nw mesh values; lane 0 builds the point list (the window walks are
// sequential), then the lanes fold the Katanin weights of its points in
// parallel.  *overflow is raised if a list exceeds emax entries.
__device__ inline void weight_tables_wave(int x, const double *mesh, int nw, double cut, const double *v2,
                                          const double *dv2, double *Wwp, double *Wc, int *Wn, double *Cwp,
                                          double *Cw, int *Cn, int emax, int *overflow, double *s_m) {
  __shared__ int s_k[2];
  const int lane = threadIdx.x;
  for (int i = lane; i < nw; i += 64) s_m[i] = mesh[i];
  __syncthreads();
  const double *m = s_m;
  if (x == nw) {
    if (lane == 0) {
      PointList o{Cwp, Cw, 0, emax};
      katanin_windows(m, nw, 0.0, cut, o);
      *Cn = o.cnt < emax ? o.cnt : emax;
      if (o.cnt > emax) atomicOr(overflow, 1);
    }
    return;
  }
  const double X = m[x];
  double *wp = Wwp + (size_t)x * emax, *w = Wc + (size_t)x * emax;
  if (lane == 0) {
    PointList o{wp, w, 0, emax};
    auto p = [&](double a, double b) { return 1.0 / ((a + sigma_at(m, nw, v2, a)) * (b + sigma_at(m, nw, v2, b))); };
    o.push(cut, p(cut, cut + X));
    if (X > 2.0 * cut) o.push(-cut, p(cut, cut - X));
    s_k[0] = o.cnt;
    katanin_windows(m, nw, X, cut, o);
    s_k[1] = o.cnt;
  }
  __syncthreads();  // the list in global memory is visible to the workgroup
  const int k0 = s_k[0], cnt = s_k[1];
  const int kend = cnt < emax ? cnt : emax;
  for (int k = k0 + lane; k < kend; k += 64) {
    const double e = wp[k];
    const double d1 = e + sigma_at(m, nw, v2, e);
    w[k] *= sigma_at(m, nw, dv2, e) / (d1 * d1 * (X + e + sigma_at(m, nw, v2, X + e)));
  }
  if (lane == 0) {
    Wn[x] = kend;
    if (cnt > emax) atomicOr(overflow, 1);
  }
}

// ---------------------------------------------------------------------------
// access-buffer descriptors (SU(2) and XYZ layouts: one element per (row, site))
// ---------------------------------------------------------------------------
// row index of a support point with the s <-> u exchange of
// _generateAccessBufferOffset (src/SU2/SU2VertexTwoParticle.hpp:616-632,
// src/XYZ/XYZVertexTwoParticle.hpp:637-653); neg = density sign flip
__device__ __forceinline__ int support_row(int so, int to, int uo, int nw, bool &neg) {
  if (so < uo) { neg = true; return tri(uo, so) * nw + to; }
  neg = false;
  return tri(so, uo) * nw + to;
}

struct Interp {
  int lo, hi;
  double b;
};

__device__ __forceinline__ Interp interp_of(const double *m, int nw, double x) {
  Interp r;
  interp(m, nw, x, r.lo, r.hi, r.b);
  return r;
}

// interpolateOffset from the bracket position c = count_le(m, n, w)
__device__ __forceinline__ void interp_finish(const double *m, int n, int c, double w, Interp &r) {
  if (!(w > m[0])) { r.lo = 0; r.hi = 0; r.b = 0.0; return; }
  if (c >= n) { r.lo = n - 1; r.hi = n - 1; r.b = 0.0; return; }
  r.lo = c - 1;
  r.hi = c;
  r.b = (w - m[r.lo]) / (m[r.hi] - m[r.lo]);
}

// two interpolateOffset calls whose bracket searches run in one binary-lifting
// loop, so that the two chains of dependent mesh reads overlap; each result
// equals interp_of(m, n, x)
__device__ __forceinline__ void interp_pair(const double *m, int n, double x0, double x1, Interp &a, Interp &b) {
  int p0 = 0, p1 = 0;
  for (int step = 1 << (31 - __clz(n)); step > 0; step >>= 1) {
    const int q0 = p0 + step, q1 = p1 + step;
    if (q0 <= n && m[q0 - 1] <= x0) p0 = q0;
    if (q1 <= n && m[q1 - 1] <= x1) p1 = q1;
  }
  interp_finish(m, n, p0, x0, a);
  interp_finish(m, n, p1, x1, b);
}

// Bracket positions on a geometric mesh (the exponential discretisation,
// m[i] close to m[0] r^i): the guess floor(log2(w) ga + gb) + 1 from the
// hardware FP32 logarithm is within one of count_le, and the comparisons with
// the mesh that follow make the position exact for any mesh.  ga = 1 / log2 r,
// gb = -log2(m[0]) / log2 r (geo_mesh_params).
__device__ __forceinline__ int count_le_geo(const double *m, int n, double w, float ga, float gb) {
  float t = fmaf(__log2f((float)w), ga, gb) + 1.0f;
  t = fminf(fmaxf(t, 0.0f), (float)n);  // NaN (w <= 0) -> 0
  int c = (int)t;
  while (c < n && m[c] <= w) ++c;
  while (c > 0 && m[c - 1] > w) --c;
  return c;
}

__device__ __forceinline__ void interp_pair_geo(const double *m, int n, double x0, double x1, float ga, float gb,
                                                Interp &a, Interp &b) {
  const int p0 = count_le_geo(m, n, x0, ga, gb), p1 = count_le_geo(m, n, x1, ga, gb);
  interp_finish(m, n, p0, x0, a);
  interp_finish(m, n, p1, x1, b);
}

// interpolateOffset with the bracket from count_le_geo (geo != 0) or from the
// binary search
__device__ __forceinline__ void interp_sel(const double *m, int n, double w, int geo, float ga, float gb, int &lo,
                                           int &hi, double &bias) {
  Interp r;
  interp_finish(m, n, geo ? count_le_geo(m, n, w, ga, gb) : count_le(m, n, w), w, r);
  lo = r.lo;
  hi = r.hi;
  bias = r.b;
}

// Host: parameters of count_le_geo when every log2(m[i] / m[0]) lies within a
// quarter step of i log2 r (on = 0 otherwise: use the binary search).
inline void geo_mesh_params(const double *m, int n, int &on, float &ga, float &gb) {
  on = 0;
  ga = gb = 0.0f;
  if (n < 2 || !(m[0] > 0.0)) return;
  const double l2r = std::log2(m[n - 1] / m[0]) / (n - 1);
  if (!(l2r > 0.0) || !std::isfinite(l2r)) return;
  for (int i = 0; i < n; ++i)
    if (!(std::fabs(std::log2(m[i] / m[0]) - i * l2r) <= 0.25)) return;
  on = 1;
  ga = (float)(1.0 / l2r);
  gb = (float)(-std::log2(m[0]) / l2r);
}

template <class V2> __device__ __forceinline__ V2 buf_load(__amdgpu_buffer_rsrc_t r, int byte_off);

template <> __device__ __forceinline__ double2 buf_load<double2>(__amdgpu_buffer_rsrc_t r, int byte_off) {
  auto v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0);
  double2 d;
  __builtin_memcpy(&d, &v, 16);
  return d;
}

template <> __device__ __forceinline__ float2 buf_load<float2>(__amdgpu_buffer_rsrc_t r, int byte_off) {
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0);
  float2 d;
  __builtin_memcpy(&d, &v, 8);
  return d;
}

// Access-buffer descriptor of generateAccessBuffer(s, t, u, channel)
// (src/SU2/SU2VertexTwoParticle.hpp:399-490, src/XYZ/XYZVertexTwoParticle.hpp:419-496)
// from the two interpolations it needs: chan 0 = s exact (I1 on t, I2 on u),
// 1 = t exact (I1 on s, I2 on u), 2 = u exact (I1 on s, I2 on t).  Support byte
// offsets are row * rowbytes; the density weight carries the sign of the
// s <-> u exchange.
struct Desc4 {
  int off[4];
  double ws[4];
  double wd[4];
  int neg;  // bit q: support q is stored s <-> u exchanged (density weight negated)
};

__device__ __forceinline__ Desc4 assemble(int chan, int ex, Interp i1, Interp i2, int nw, int rowbytes) {
  Desc4 d;
  const int l1[4] = {i1.lo, i1.hi, i1.lo, i1.hi};
  const int l2[4] = {i2.lo, i2.lo, i2.hi, i2.hi};
  const double b1 = i1.b, b2 = i2.b;
  const double w[4] = {(1 - b2) * (1 - b1), (1 - b2) * b1, b2 * (1 - b1), b2 * b1};
  d.neg = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int so, to, uo;
    if (chan == 0) { so = ex; to = l1[q]; uo = l2[q]; }
    else if (chan == 1) { so = l1[q]; to = ex; uo = l2[q]; }
    else { so = l1[q]; to = l2[q]; uo = ex; }
    bool neg;
    const int row = support_row(so, to, uo, nw, neg);
    d.off[q] = row * rowbytes;
    d.ws[q] = w[q];
    d.wd[q] = neg ? -w[q] : w[q];
    if (neg) d.neg |= 1 << q;
  }
  return d;
}

// one access buffer of a pair: the signed s and u arguments decide the site
// exchange (exactly one of them negative); swap says whether the
// pair's interpolations (I0, I1) feed the descriptor's (slot 1, slot 2) or
// (slot 2, slot 1)
struct PairBuf {
  int k;
  double a, c;
  bool swap;
  int chan;
};

}  // namespace frg
