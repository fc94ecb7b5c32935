// MI355X (gfx950) kernels and C ABI of the SU(2) pf-FRG flow engine, FP64 form.
//
// Device data layout
//   G, dG   vertex and its flow, interleaved channels: double2 {spin, density}
//           at [(tri(s,u) * Nw + t) * L + rid], tri(s,u) = s(s+1)/2 + u, s >= u.
//           One 16-byte load fetches both channels of a support point.
//   v2,dv2  self-energy and its flow, Nw doubles.
//   W       per exact-frequency index x: the evaluation points w' of the
//           frequency integral of a channel whose transfer frequency is
//           mesh[x] (conventional +-Lambda terms and the three trapezoid
//           windows of the Katanin term) with their coefficients
//           c = trapezoid weight * propagator product.  Depends only on
//           (x, Lambda), so it is built once per step for all items.
//
// Kernels (one derivative evaluation = k_weights, k_dsigma, k_weights, k_dgamma):
//   k_dsigma   one wavefront per mesh point; lanes split the lattice range sum.
//   k_weights  one thread per exact index: window points and coefficients.
//   k_dgamma   one 256-thread workgroup per frequency triple (s,t,u); threads
//              are (slot, rid) pairs, slots share the evaluation points of a
//              batch.  Per batch: (A) access-buffer descriptors (bracket search
//              on the LDS mesh, sign / site-exchange folding) one per thread,
//              (B) 16-byte vertex gathers and the channel algebra per (slot,
//              rid), (C, t channel only) the RPA lattice bubble accumulated as
//              C[a][b] += sum_e c_e A2_e[a] A3_e[b] in LDS, contracted with the
//              overlap lists once per item.
//   k_euler    in-place Gamma += dLambda * dGamma (streaming, HBM bound).
//   k_corr_*   static susceptibility, one workgroup per outer frequency.
//
// Reference semantics followed (paths relative to the SpinParser repository):
//   src/FrequencyDiscretization.hpp:253-356, src/lib/Integrator.hpp:137-287,
//   src/SU2/SU2VertexTwoParticle.hpp:185-632, src/SU2/SU2FrgCore.cpp:111-431,
//   src/SU2/SU2MeasurementCorrelation.cpp:77-177.
// Frequency arguments, window predicates and bracket comparisons are evaluated
// with the same IEEE operations as the reference (the file is compiled with
// -ffp-contract=off); only the accumulations use explicit fma().
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "spfrg.h"

namespace {

constexpr int kThreads = 256;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr int kMaxBatch = 32;

// ---------------------------------------------------------------------------
// frequency mesh (positive half m[0..n) ascending; signed index i in [-n, n))
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mesh_at(const double *m, int i) { return i >= 0 ? m[i] : -m[-i - 1]; }

__device__ int greater_pos(const double *m, int n, double w) {
  if (w <= m[0]) return 0;
  for (int i = 1; i < n; ++i)
    if (m[i] > w) return i;
  return n - 1;
}

__device__ int lesser_pos(const double *m, int n, double w) {
  if (w <= m[0]) return 0;
  for (int i = 1; i < n; ++i)
    if (m[i] > w) return i - 1;
  return n - 1;
}

// FrequencyDiscretization::lesser / greater (src/FrequencyDiscretization.hpp:253-296)
__device__ int mesh_lesser(const double *m, int n, double w) { return w < 0 ? -(greater_pos(m, n, -w) + 1) : lesser_pos(m, n, w); }
__device__ int mesh_greater(const double *m, int n, double w) { return w < 0 ? -(lesser_pos(m, n, -w) + 1) : greater_pos(m, n, w); }

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
__device__ double sigma_at(const double *m, int n, const double *data, double w) {
  double sign = 1.0;
  if (w < 0) { w = -w; sign = -1.0; }
  int lo, hi; double b;
  interp(m, n, w, lo, hi, b);
  return sign * ((1 - b) * data[lo] + b * data[hi]);
}

__device__ __forceinline__ int tri(int a, int b) { return a * (a + 1) / 2 + b; }

// row index of a support point with the s <-> u exchange of
// SU2VertexTwoParticle::_generateAccessBufferOffset (:616-632); neg = density sign flip
__device__ __forceinline__ int support_row(int so, int to, int uo, int nw, bool &neg) {
  if (so < uo) { neg = true; return tri(uo, so) * nw + to; }
  neg = false;
  return tri(so, uo) * nw + to;
}

// vertex value with exact mesh indices, SU2VertexTwoParticle::_directAccessMapFrequencyExchange
__device__ __forceinline__ double2 direct2(const double2 *G, int L, int nw, int rid, int so, int to, int uo) {
  if (so >= uo) return G[(size_t)(tri(so, uo) * nw + to) * L + rid];
  double2 v = G[(size_t)(tri(uo, so) * nw + to) * L + rid];
  v.y = -v.y;
  return v;
}

// SU2VertexTwoParticle::getValue(..., FrequencyChannel::None) for both channels.
// rid_fwd / rid_inv: symmetryTransform(i1, i2) and symmetryTransform(i2, i1).
__device__ double2 value_none(const double2 *G, const double *m, int nw, int L, int rid_fwd, int rid_inv, double s,
                              double t, double u) {
  bool swap = false;
  if (s < 0 && u < 0) { s = -s; u = -u; }
  else {
    if (s < 0) { s = -s; swap = true; }
    else if (u < 0) { u = -u; swap = true; }
  }
  if (t < 0) t = -t;
  const int rid = swap ? rid_inv : rid_fwd;
  int lS, uS, lT, uT, lU, uU;
  double bS, bT, bU;
  interp(m, nw, s, lS, uS, bS);
  interp(m, nw, t, lT, uT, bT);
  interp(m, nw, u, lU, uU, bU);
  double2 r;
  {
    double2 a000 = direct2(G, L, nw, rid, lS, lT, lU), a100 = direct2(G, L, nw, rid, uS, lT, lU);
    double2 a010 = direct2(G, L, nw, rid, lS, uT, lU), a110 = direct2(G, L, nw, rid, uS, uT, lU);
    double2 a001 = direct2(G, L, nw, rid, lS, lT, uU), a101 = direct2(G, L, nw, rid, uS, lT, uU);
    double2 a011 = direct2(G, L, nw, rid, lS, uT, uU), a111 = direct2(G, L, nw, rid, uS, uT, uU);
    r.x = (1 - bU) * ((1 - bT) * ((1 - bS) * a000.x + bS * a100.x) + bT * ((1 - bS) * a010.x + bS * a110.x)) +
          bU * ((1 - bT) * ((1 - bS) * a001.x + bS * a101.x) + bT * ((1 - bS) * a011.x + bS * a111.x));
    r.y = (1 - bU) * ((1 - bT) * ((1 - bS) * a000.y + bS * a100.y) + bT * ((1 - bS) * a010.y + bS * a110.y)) +
          bU * ((1 - bT) * ((1 - bS) * a001.y + bS * a101.y) + bT * ((1 - bS) * a011.y + bS * a111.y));
  }
  return r;
}

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
  __device__ void push(double x, double weight) {
    if (cnt < cap) { wp[cnt] = x; w[cnt] = weight; }
    ++cnt;
  }
};

__device__ void window_right(const double *m, int n, double max, PointList &o) {
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

__device__ void window_both(const double *m, int n, double min, double max, PointList &o) {
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

__device__ void window_left(const double *m, int n, double min, PointList &o) {
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
__device__ void katanin_windows(const double *m, int n, double X, double cut, PointList &o) {
  if (-(X + cut) > mesh_at(m, -n)) window_right(m, n, -(X + cut), o);
  if (X - cut > cut) window_both(m, n, cut - X, -cut, o);
  if (cut < mesh_at(m, n - 1)) window_left(m, n, cut, o);
}

// W table: x < nw -> channel lists with propagators; x == nw -> plain
// correlation list.  flags[1] is raised if a list overflows its capacity.
__global__ void k_weights(const double *__restrict__ m, int nw, const double *__restrict__ v2,
                          const double *__restrict__ dv2, double cut, int emax, double *Wwp, double *Wc, int *Wn,
                          double *Cwp, double *Cw, int *Cn, int corr_only, int *flags) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x > nw) return;
  if (x == nw) {
    PointList o{Cwp, Cw, 0, emax};
    katanin_windows(m, nw, 0.0, cut, o);
    *Cn = o.cnt < emax ? o.cnt : emax;
    if (o.cnt > emax) atomicOr(&flags[1], 1);
    return;
  }
  if (corr_only) return;
  const double X = m[x];
  PointList o{Wwp + (size_t)x * emax, Wc + (size_t)x * emax, 0, emax};
  // conventional contribution (src/SU2/SU2FrgCore.cpp:349-371): p(L, L +- X)
  auto p = [&](double a, double b) { return 1.0 / ((a + sigma_at(m, nw, v2, a)) * (b + sigma_at(m, nw, v2, b))); };
  o.push(cut, p(cut, cut + X));
  if (X > 2.0 * cut) o.push(-cut, p(cut, cut - X));
  const int k0 = o.cnt;
  katanin_windows(m, nw, X, cut, o);
  const int kend = o.cnt < emax ? o.cnt : emax;
  // Katanin propagator dSigma(w') G(w')^2 G(X + w') (src/SU2/SU2FrgCore.cpp:343-347)
  for (int k = k0; k < kend; ++k) {
    const double wp = o.wp[k];
    const double d1 = wp + sigma_at(m, nw, v2, wp);
    o.w[k] *= sigma_at(m, nw, dv2, wp) / (d1 * d1 * (X + wp + sigma_at(m, nw, v2, X + wp)));
  }
  Wn[x] = kend;
  if (o.cnt > emax) atomicOr(&flags[1], 1);
}

// ---------------------------------------------------------------------------
// dSigma/dLambda (src/SU2/SU2FrgCore.cpp:139-169): one wavefront per mesh point
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_dsigma(const double2 *__restrict__ G, const double *__restrict__ m,
                                               const double *__restrict__ v2, const int *__restrict__ rfwd,
                                               const int *__restrict__ rinv, int R, int nw, int L, double cut,
                                               double spin, double *dv2, int *flags) {
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const double w = m[i];
  double sum = 0.0;
  for (int j = lane; j < R; j += 64) {
    const double2 a = value_none(G, m, nw, L, rfwd[j], rinv[j], w + cut, 0.0, w - cut);
    const double2 b = value_none(G, m, nw, L, rfwd[j], rinv[j], w - cut, 0.0, w + cut);
    sum += a.y - b.y;
  }
  sum = wave_sum(sum);
  if (lane == 0) {
    double v = 0.0;
    v -= 4.0 * spin * sum;
    const double2 p = value_none(G, m, nw, L, 0, 0, w + cut, w - cut, 0.0);
    const double2 q = value_none(G, m, nw, L, 0, 0, w - cut, w + cut, 0.0);
    v += 0.75 * (p.x - q.x);
    v += (p.y - q.y);
    v /= (kTwoPi * (cut + sigma_at(m, nw, v2, cut)));
    dv2[i] = v;
    if (isnan(v)) atomicOr(&flags[0], 1);
  }
}

// ---------------------------------------------------------------------------
// dGamma/dLambda
// ---------------------------------------------------------------------------
struct GammaArgs {
  const double2 *__restrict__ G;
  double2 *__restrict__ dG;
  const double *__restrict__ mesh;
  const int *__restrict__ site_rid;
  const int *__restrict__ inv_rid;
  const int *__restrict__ ovl_off;
  const int *__restrict__ ovl_rid1;
  const int *__restrict__ ovl_rid2;
  const double *__restrict__ Wwp;
  const double *__restrict__ Wc;
  const int *__restrict__ Wn;
  int *flags;
  int nw, L, emax, nb, slots, item0;
  double two_s, eight_s;
};

// LDS carve-up (bytes), computed identically on host and device
struct LdsLayout {
  int mesh, wp, c, w, loc, P, Q, C, rpa, red, row, flag, site, inv, total;
};

__host__ __device__ inline LdsLayout lds_layout(int nw, int L, int nb, int slots) {
  LdsLayout l;
  int off = 0;
  auto take = [&](int bytes) { int o = off; off += (bytes + 15) & ~15; return o; };
  l.mesh = take(nw * 8);
  l.wp = take(nb * 8);
  l.c = take(nb * 8);
  l.w = take(nb * 4 * 4 * 8);
  l.loc = take(nb * 4 * 2 * 8);
  const int pq = 2 * nb * L * 8;
  const int red = slots * L * 2 * 8;
  l.P = take(pq > red ? pq : red);  // the final slot reduction reuses P
  l.red = l.P;
  l.Q = take(pq);
  l.C = take(2 * L * L * 8);
  l.rpa = take(2 * L * 8);
  l.row = take(nb * 4 * 4 * 4);
  l.flag = take(nb * 4 * 4);
  l.site = take(L * 4);
  l.inv = take(L * 4);
  l.total = off;
  return l;
}

// access-buffer descriptor: 4 support rows, weights, density-sign mask, exchange
struct Desc {
  int row[4];
  double w[4];
  int flags;  // bit i: negate density of support i; bit 4: site exchange
};

// SU2VertexTwoParticle::generateAccessBuffer(s, t, u, channel) (:399-490) with
// the exact argument's mesh index given: chan 0 = s exact, 1 = t, 2 = u
__device__ __forceinline__ Desc make_desc(const double *m, int nw, double a, double b, double c, int chan, int ex) {
  Desc d;
  int fl = 0;
  if (a < 0 && c < 0) { a = -a; c = -c; }
  else {
    if (a < 0) { a = -a; fl |= 16; }
    else if (c < 0) { c = -c; fl |= 16; }
  }
  if (b < 0) b = -b;
  int lo1, hi1, lo2, hi2;
  double b1, b2;
  bool n0, n1, n2, n3;
  if (chan == 0) {
    interp(m, nw, b, lo1, hi1, b1);
    interp(m, nw, c, lo2, hi2, b2);
    d.row[0] = support_row(ex, lo1, lo2, nw, n0);
    d.row[1] = support_row(ex, hi1, lo2, nw, n1);
    d.row[2] = support_row(ex, lo1, hi2, nw, n2);
    d.row[3] = support_row(ex, hi1, hi2, nw, n3);
  } else if (chan == 1) {
    interp(m, nw, a, lo1, hi1, b1);
    interp(m, nw, c, lo2, hi2, b2);
    d.row[0] = support_row(lo1, ex, lo2, nw, n0);
    d.row[1] = support_row(hi1, ex, lo2, nw, n1);
    d.row[2] = support_row(lo1, ex, hi2, nw, n2);
    d.row[3] = support_row(hi1, ex, hi2, nw, n3);
  } else {
    interp(m, nw, a, lo1, hi1, b1);
    interp(m, nw, b, lo2, hi2, b2);
    d.row[0] = support_row(lo1, lo2, ex, nw, n0);
    d.row[1] = support_row(hi1, lo2, ex, nw, n1);
    d.row[2] = support_row(lo1, hi2, ex, nw, n2);
    d.row[3] = support_row(hi1, hi2, ex, nw, n3);
  }
  d.w[0] = (1 - b2) * (1 - b1);
  d.w[1] = (1 - b2) * b1;
  d.w[2] = b2 * (1 - b1);
  d.w[3] = b2 * b1;
  d.flags = fl | (n0 ? 1 : 0) | (n1 ? 2 : 0) | (n2 ? 4 : 0) | (n3 ? 8 : 0);
  return d;
}

// frequency arguments (s, t, u) of access buffer k of channel ch
// (src/SU2/SU2FrgCore.cpp:199-334)
__device__ __forceinline__ void buffer_args(int ch, int k, double s, double t, double u, double w1p, double w1,
                                            double w2p, double w2, double wp, double &a, double &b, double &c,
                                            int &chan) {
  if (ch == 0) {
    chan = 0;
    a = s;
    switch (k) {
      case 0: b = -w1 - wp; c = -w2 - wp; break;
      case 1: b = w1p + wp; c = -w2p - wp; break;
      case 2: b = w2 + wp; c = w1 + wp; break;
      default: b = -w2p - wp; c = w1p + wp; break;
    }
  } else if (ch == 1) {
    if (k < 4) {
      chan = 1;
      b = t;
      switch (k) {
        case 0: a = w1 - wp; c = w1p + wp; break;
        case 1: a = w2p - wp; c = -w2 - wp; break;
        case 2: a = w1p + wp; c = w1 - wp; break;
        default: a = w2 + wp; c = wp - w2p; break;
      }
    } else {
      chan = 2;
      switch (k) {
        case 4: a = w2p - wp; b = -w2 - wp; c = t; break;
        case 5: a = w1 - wp; b = -w1p - wp; c = -t; break;
        case 6: a = w2 + wp; b = wp - w2p; c = t; break;
        default: a = w1p + wp; b = wp - w1; c = -t; break;
      }
    }
  } else {
    chan = 2;
    c = u;
    switch (k) {
      case 0: a = w1 + wp; b = wp - w2p; break;
      case 1: a = w1p + wp; b = w2 - wp; break;
      case 2: a = w2p - wp; b = -w1 - wp; break;
      default: a = w2 - wp; b = w1p + wp; break;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_dgamma(GammaArgs g) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int nw = g.nw, L = g.L, nb = g.nb, slots = g.slots;
  const LdsLayout lay = lds_layout(nw, L, nb, slots);
  double *s_mesh = reinterpret_cast<double *>(lds_raw + lay.mesh);
  double *s_wp = reinterpret_cast<double *>(lds_raw + lay.wp);
  double *s_c = reinterpret_cast<double *>(lds_raw + lay.c);
  double *s_w = reinterpret_cast<double *>(lds_raw + lay.w);
  double *s_loc = reinterpret_cast<double *>(lds_raw + lay.loc);
  double *s_P = reinterpret_cast<double *>(lds_raw + lay.P);
  double *s_Q = reinterpret_cast<double *>(lds_raw + lay.Q);
  double *s_C = reinterpret_cast<double *>(lds_raw + lay.C);
  double *s_rpa = reinterpret_cast<double *>(lds_raw + lay.rpa);
  double *s_red = reinterpret_cast<double *>(lds_raw + lay.red);
  int *s_row = reinterpret_cast<int *>(lds_raw + lay.row);
  int *s_flag = reinterpret_cast<int *>(lds_raw + lay.flag);
  int *s_site = reinterpret_cast<int *>(lds_raw + lay.site);
  int *s_inv = reinterpret_cast<int *>(lds_raw + lay.inv);

  const int tid = threadIdx.x;
  const int item = blockIdx.x + g.item0;
  // SU2VertexTwoParticle::expandIterator (:136-158): item = tri(s,u) * nw + t
  const int su = item / nw;
  const int to = item - su * nw;
  int so = (int)((sqrt(8.0 * su + 1.0) - 1.0) * 0.5);
  while ((so + 1) * (so + 2) / 2 <= su) ++so;
  while (so * (so + 1) / 2 > su) --so;
  const int uo = su - so * (so + 1) / 2;

  for (int i = tid; i < nw; i += kThreads) s_mesh[i] = g.mesh[i];
  for (int i = tid; i < L; i += kThreads) { s_site[i] = g.site_rid[i]; s_inv[i] = g.inv_rid[i]; }
  for (int i = tid; i < 2 * L * L; i += kThreads) s_C[i] = 0.0;
  __syncthreads();

  const double s = s_mesh[so], t = s_mesh[to], u = s_mesh[uo];
  const double w1p = 0.5 * (s + t + u);
  const double w1 = 0.5 * (s - t + u);
  const double w2p = 0.5 * (s - t - u);
  const double w2 = 0.5 * (s + t - u);

  const int slot = tid / L;
  const int rid = tid - slot * L;
  const bool active = slot < slots;
  double acc_s = 0.0, acc_d = 0.0;

  for (int ch = 0; ch < 3; ++ch) {
    const int xo = ch == 0 ? so : (ch == 1 ? to : uo);
    const int n = g.Wn[xo];
    const double *Wwp = g.Wwp + (size_t)xo * g.emax;
    const double *Wc = g.Wc + (size_t)xo * g.emax;
    const int nbuf = ch == 1 ? 8 : 4;
    const double sgn = ch == 2 ? -1.0 : 1.0;
    for (int e0 = 0; e0 < n; e0 += nb) {
      const int nvalid = n - e0 < nb ? n - e0 : nb;
      // (A) descriptors, one (evaluation, buffer) per thread
      if (tid < nvalid * nbuf) {
        const int e = tid / nbuf, k = tid - e * nbuf;
        const double wp = Wwp[e0 + e];
        double a, b, c;
        int chan;
        buffer_args(ch, k, s, t, u, w1p, w1, w2p, w2, wp, a, b, c, chan);
        const Desc d = make_desc(s_mesh, nw, a, b, c, chan, xo);
        if (k < 4) {
          const int base = (e * 4 + k) * 4;
          for (int i = 0; i < 4; ++i) { s_row[base + i] = d.row[i]; s_w[base + i] = d.w[i]; }
          s_flag[e * 4 + k] = d.flags;
        } else {
          // SU2VertexTwoParticle::getValueLocal (:347-360): site 0, no exchange
          double vs = 0.0, vd = 0.0;
          for (int i = 0; i < 4; ++i) {
            const double2 v = g.G[(size_t)d.row[i] * L];
            vs = fma(d.w[i], v.x, vs);
            vd = fma((d.flags >> i) & 1 ? -d.w[i] : d.w[i], v.y, vd);
          }
          s_loc[(e * 4 + (k - 4)) * 2 + 0] = vs;
          s_loc[(e * 4 + (k - 4)) * 2 + 1] = vd;
        }
        if (k == 0) { s_wp[e] = wp; s_c[e] = sgn * Wc[e0 + e]; }
      }
      __syncthreads();
      // (B) gathers and channel algebra per (slot, rid)
      if (active) {
        for (int e = slot; e < nvalid; e += slots) {
          double As[4], Ad[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int fl = s_flag[e * 4 + k];
            const int sr = (fl & 16) ? s_inv[rid] : s_site[rid];
            const int base = (e * 4 + k) * 4;
            double xs = 0.0, xd = 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const double2 v = g.G[(size_t)s_row[base + i] * L + sr];
              const double w = s_w[base + i];
              xs = fma(w, v.x, xs);
              xd = fma((fl >> i) & 1 ? -w : w, v.y, xd);
            }
            As[k] = xs;
            Ad[k] = xd;
          }
          const double c = s_c[e];
          double rs, rd;
          if (ch == 0) {
            // integralKernelS (src/SU2/SU2FrgCore.cpp:199-227)
            rs = -0.5 * As[0] * As[1] - 0.5 * As[2] * As[3] + Ad[0] * As[1] + Ad[2] * As[3] + As[0] * Ad[1] +
                 As[2] * Ad[3];
            rd = 0.1875 * (As[0] * As[1] + As[2] * As[3]) + Ad[0] * Ad[1] + Ad[2] * Ad[3];
          } else if (ch == 2) {
            // integralKernelU (:305-334)
            rs = 0.5 * As[0] * As[1] + 0.5 * As[2] * As[3] + As[0] * Ad[1] + As[2] * Ad[3] + Ad[0] * As[1] +
                 Ad[2] * As[3];
            rd = 0.1875 * (As[0] * As[1] + As[2] * As[3]) + Ad[0] * Ad[1] + Ad[2] * Ad[3];
          } else {
            // integralKernelT (:229-303) without the lattice bubble
            const double *lc = s_loc + e * 8;
            const double Cbs = lc[0], Cbd = lc[1], ICas = lc[2], ICad = lc[3];
            const double Cbs2 = lc[4], Cbd2 = lc[5], ICas2 = lc[6], ICad2 = lc[7];
            rs = (0.25 * Cbs - Cbd) * As[0] + (0.25 * ICas - ICad) * As[1] + (0.25 * Cbs2 - Cbd2) * As[2] +
                 (0.25 * ICas2 - ICad2) * As[3];
            rd = -(Cbd + 0.75 * Cbs) * Ad[0] - (ICad + 0.75 * ICas) * Ad[1] - (Cbd2 + 0.75 * Cbs2) * Ad[2] -
                 (ICad2 + 0.75 * ICas2) * Ad[3];
            // lattice-bubble factors: P = c A2, Q = A3 per channel
            s_P[(0 * nb + e) * L + rid] = c * As[2];
            s_P[(1 * nb + e) * L + rid] = c * Ad[2];
            s_Q[(0 * nb + e) * L + rid] = As[3];
            s_Q[(1 * nb + e) * L + rid] = Ad[3];
          }
          acc_s = fma(c, rs, acc_s);
          acc_d = fma(c, rd, acc_d);
        }
      }
      __syncthreads();
      if (ch == 1) {
        // (C) C[ch][a][b] += sum_e P[ch][e][a] Q[ch][e][b]
        const int LL = L * L;
        for (int j = tid; j < 2 * LL; j += kThreads) {
          const int cch = j / LL;
          const int r = j - cch * LL;
          const int a = r / L, b = r - a * L;
          const double *P = s_P + cch * nb * L + a;
          const double *Q = s_Q + cch * nb * L + b;
          double acc = s_C[j];
          for (int e = 0; e < nvalid; ++e) acc = fma(P[e * L], Q[e * L], acc);
          s_C[j] = acc;
        }
        __syncthreads();
      }
    }
    if (ch == 1) {
      // RPA(rid) = sum over getOverlap(rid) of C[rid1][rid2] (src/SU2/SU2FrgCore.cpp:250-266)
      const int wave = tid >> 6, lane = tid & 63;
      for (int r = wave; r < L; r += kThreads / 64) {
        double ps = 0.0, pd = 0.0;
        const int b = g.ovl_off[r], e = g.ovl_off[r + 1];
        for (int i = b + lane; i < e; i += 64) {
          const int idx = g.ovl_rid1[i] * L + g.ovl_rid2[i];
          ps += s_C[idx];
          pd += s_C[L * L + idx];
        }
        ps = wave_sum(ps);
        pd = wave_sum(pd);
        if (lane == 0) { s_rpa[2 * r] = g.two_s * ps; s_rpa[2 * r + 1] = g.eight_s * pd; }
      }
      __syncthreads();
    }
  }

  // reduce the slots; prefactor 1/(2 pi) (src/SU2/SU2FrgCore.cpp:426-430)
  if (active) { s_red[(slot * L + rid) * 2] = acc_s; s_red[(slot * L + rid) * 2 + 1] = acc_d; }
  __syncthreads();
  if (tid < L) {
    double ts = s_rpa[2 * tid], td = s_rpa[2 * tid + 1];
    for (int k = 0; k < slots; ++k) { ts += s_red[(k * L + tid) * 2]; td += s_red[(k * L + tid) * 2 + 1]; }
    ts *= (1.0 / kTwoPi);
    td *= (1.0 / kTwoPi);
    g.dG[(size_t)item * L + tid] = make_double2(ts, td);
    if (isnan(ts) || isnan(td)) atomicOr(&g.flags[0], 1);
  }
}

// in-place explicit Euler step (src/SU2/SU2FrgCore.cpp:111-137)
__global__ void k_euler(double2 *__restrict__ G, const double2 *__restrict__ dG, size_t n, double *__restrict__ v2,
                        const double *__restrict__ dv2, int nw, double step) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double2 a = G[i];
    const double2 d = dG[i];
    a.x = fma(step, d.x, a.x);
    a.y = fma(step, d.y, a.y);
    G[i] = a;
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < nw; i += blockDim.x) v2[i] = fma(step, dv2[i], v2[i]);
}

// ---------------------------------------------------------------------------
// static susceptibility (src/SU2/SU2MeasurementCorrelation.cpp:77-177), nu = 0
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_corr_partial(const double2 *__restrict__ G,
                                                           const double *__restrict__ m, const double *__restrict__ v2,
                                                           const int *__restrict__ site_rid,
                                                           const int *__restrict__ inv_rid,
                                                           const double *__restrict__ Cwp,
                                                           const double *__restrict__ Cw, const int *__restrict__ Cn,
                                                           int nw, int L, int slots, double spin, double *partial) {
  __shared__ double red[kThreads * 2];
  const int o = blockIdx.x;
  const int n = *Cn;
  if (o >= n) return;
  const int tid = threadIdx.x;
  const int slot = tid / L, rid = tid - (tid / L) * L;
  const bool active = slot < slots;
  const double nu = 0.0;
  const double w = Cwp[o];
  const double gw = (w + sigma_at(m, nw, v2, w)) * (w + nu + sigma_at(m, nw, v2, w + nu));
  double acc_s = 0.0, acc_d = 0.0;
  if (active) {
    for (int e = slot; e < n; e += slots) {
      const double wp = Cwp[e];
      // generateAccessBuffer(w + wp + nu, nu, w - wp): trilinear, 8 supports (:500-557)
      double a = w + wp + nu, b = nu, c = w - wp;
      bool exch = false;
      if (a < 0 && c < 0) { a = -a; c = -c; }
      else {
        if (a < 0) { a = -a; exch = true; }
        else if (c < 0) { c = -c; exch = true; }
      }
      if (b < 0) b = -b;
      int lS, uS, lT, uT, lU, uU;
      double bS, bT, bU;
      interp(m, nw, a, lS, uS, bS);
      interp(m, nw, b, lT, uT, bT);
      interp(m, nw, c, lU, uU, bU);
      const int sr = exch ? inv_rid[rid] : site_rid[rid];
      const int S3[8] = {lS, uS, lS, uS, lS, uS, lS, uS};
      const int T3[8] = {lT, lT, uT, uT, lT, lT, uT, uT};
      const int U3[8] = {lU, lU, lU, lU, uU, uU, uU, uU};
      const double wts[8] = {(1 - bT) * (1 - bS) * (1 - bU), (1 - bT) * bS * (1 - bU), bT * (1 - bS) * (1 - bU),
                             bT * bS * (1 - bU),             (1 - bT) * (1 - bS) * bU, (1 - bT) * bS * bU,
                             bT * (1 - bS) * bU,             bT * bS * bU};
      double sbs = 0.0, sbd = 0.0;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        bool neg;
        const int row = support_row(S3[i], T3[i], U3[i], nw, neg);
        const double2 v = G[(size_t)row * L + sr];
        sbs = fma(wts[i], v.x, sbs);
        sbd = fma(neg ? -wts[i] : wts[i], v.y, sbd);
      }
      double rs = -(spin * spin) * sbs;
      double rd = -(16.0 * spin * spin) * sbd;
      if (rid == 0) {
        const double2 vv = value_none(G, m, nw, L, 0, 0, w + wp + nu, w - wp, nu);
        rs += spin * (-vv.x / 4.0 + vv.y);
        rd += spin * (3.0 * vv.x + 4.0 * vv.y);
      }
      const double norm = 1.0 / (gw * (wp + sigma_at(m, nw, v2, wp)) * (wp + nu + sigma_at(m, nw, v2, wp + nu)) *
                                 (4.0 * kPi * kPi));
      const double cf = Cw[e] * norm;
      acc_s = fma(cf, rs, acc_s);
      acc_d = fma(cf, rd, acc_d);
    }
  }
  red[2 * tid] = active ? acc_s : 0.0;
  red[2 * tid + 1] = active ? acc_d : 0.0;
  __syncthreads();
  if (tid < L) {
    double ts = 0.0, td = 0.0;
    for (int k = 0; k < slots; ++k) { ts += red[2 * (k * L + tid)]; td += red[2 * (k * L + tid) + 1]; }
    if (tid == 0) {
      const double term1 = 1.0 / gw;
      ts += spin * term1 / kTwoPi;
      td += 2.0 * spin * term1 / kPi;
    }
    partial[((size_t)o * L + tid) * 2] = Cw[o] * ts;
    partial[((size_t)o * L + tid) * 2 + 1] = Cw[o] * td;
  }
}

__global__ void k_corr_final(const double *__restrict__ partial, const int *__restrict__ Cn, int L,
                             const int *__restrict__ corr_rid, int nout, double *out) {
  __shared__ double susc[2 * 256];
  const int n = *Cn;
  for (int r = threadIdx.x; r < L; r += blockDim.x) {
    double s = 0.0, d = 0.0;
    for (int o = 0; o < n; ++o) { s += partial[((size_t)o * L + r) * 2]; d += partial[((size_t)o * L + r) * 2 + 1]; }
    susc[2 * r] = s;
    susc[2 * r + 1] = d;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nout; k += blockDim.x) {
    out[k] = susc[2 * corr_rid[k]];
    out[nout + k] = susc[2 * corr_rid[k] + 1];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct sp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nw = 0, L = 0, R = 0, B = 0, Rb = 0, O = 0, emax = 0;
  long long nf = 0;
  double spin = 0.5;
  double cutoff = 0.0;
  bool have_flow = false;
  int slots = 0, nb = 0;
  size_t lds = 0;
  double *d_mesh = nullptr;
  int *d_site = nullptr, *d_inv = nullptr, *d_ovl_off = nullptr, *d_ovl1 = nullptr, *d_ovl2 = nullptr;
  int *d_rfwd = nullptr, *d_rinv = nullptr, *d_corr = nullptr;
  double2 *d_G = nullptr, *d_dG = nullptr;
  double *d_v2 = nullptr, *d_dv2 = nullptr;
  double *d_Wwp = nullptr, *d_Wc = nullptr, *d_Cwp = nullptr, *d_Cw = nullptr;
  int *d_Wn = nullptr, *d_Cn = nullptr;
  double *d_partial = nullptr, *d_corr_out = nullptr;
  int *d_flags = nullptr;
  std::vector<void *> allocs;
  // profiling
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[4];
  std::vector<hipEvent_t> free_events;
  double prof_ms[4] = {0, 0, 0, 0};
  long prof_n[4] = {0, 0, 0, 0};
  std::string err;
};

namespace {

int fail(sp_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define SP_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail(ctx, SP_EHIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T> int dalloc(sp_ctx *ctx, T **p, size_t count) {
  void *q = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) return fail(ctx, SP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return SP_OK;
}

hipEvent_t get_event(sp_ctx *ctx) {
  if (!ctx->free_events.empty()) {
    hipEvent_t e = ctx->free_events.back();
    ctx->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct ProfScope {
  sp_ctx *ctx;
  int k;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(sp_ctx *c, int kernel) : ctx(c), k(kernel) {
    if (ctx->prof) {
      a = get_event(ctx);
      b = get_event(ctx);
      if (a && b) hipEventRecord(a, ctx->stream);
    }
  }
  ~ProfScope() {
    if (ctx->prof && a && b) {
      hipEventRecord(b, ctx->stream);
      ctx->pending[k].push_back({a, b});
    }
  }
};

int check_problem(sp_ctx *ctx, const sp_su2_problem *p) {
  if (!p || p->nw < 2 || p->nw > 4096 || !p->mesh) return fail(ctx, SP_EARG, "invalid frequency mesh");
  for (int i = 0; i < p->nw; ++i)
    if (!(p->mesh[i] > 0.0f) || (i > 0 && !(p->mesh[i] > p->mesh[i - 1])))
      return fail(ctx, SP_EARG, "mesh must be positive and strictly ascending");
  if (p->L < 1 || p->L > 64) return fail(ctx, SP_EARG, "number of representative sites must be in [1, 64]");
  if (!p->site_rid || !p->inv_rid || !p->ovl_off || !p->range_fwd || !p->range_inv || !p->corr_rid)
    return fail(ctx, SP_EARG, "missing lattice table");
  if (p->R < 1 || p->B < 1 || p->Rb < 1) return fail(ctx, SP_EARG, "invalid range sizes");
  if (p->ovl_off[0] != 0) return fail(ctx, SP_EARG, "overlap offsets must start at 0");
  for (int r = 0; r < p->L; ++r)
    if (p->ovl_off[r + 1] < p->ovl_off[r]) return fail(ctx, SP_EARG, "overlap offsets must be non-decreasing");
  const int O = p->ovl_off[p->L];
  if (O > 0 && (!p->ovl_rid1 || !p->ovl_rid2)) return fail(ctx, SP_EARG, "missing overlap lists");
  auto in_range = [&](const int *a, long n) {
    for (long i = 0; i < n; ++i)
      if (a[i] < 0 || a[i] >= p->L) return false;
    return true;
  };
  if (!in_range(p->site_rid, p->L) || !in_range(p->inv_rid, p->L) || !in_range(p->ovl_rid1, O) ||
      !in_range(p->ovl_rid2, O) || !in_range(p->range_fwd, p->R) || !in_range(p->range_inv, p->R) ||
      !in_range(p->corr_rid, (long)p->B * p->Rb))
    return fail(ctx, SP_EARG, "lattice table entry outside [0, L)");
  if (p->precision != SP_PRECISION_FP64) return fail(ctx, SP_EARG, "unsupported precision mode");
  return SP_OK;
}

void launch_shape(int nw, int L, int &slots, int &nb, size_t &lds) {
  slots = kThreads / L;
  int per = kMaxBatch / slots;
  if (per < 1) per = 1;
  nb = slots * per;
  if (nb > kMaxBatch) nb = kMaxBatch;
  lds = (size_t)lds_layout(nw, L, nb, slots).total;
}

int weights(sp_ctx *ctx, int corr_only) {
  ProfScope ps(ctx, 2);
  const int threads = 64;
  const int blocks = (ctx->nw + 1 + threads - 1) / threads;
  hipLaunchKernelGGL(k_weights, dim3(blocks), dim3(threads), 0, ctx->stream, ctx->d_mesh, ctx->nw, ctx->d_v2,
                     ctx->d_dv2, ctx->cutoff, ctx->emax, ctx->d_Wwp, ctx->d_Wc, ctx->d_Wn, ctx->d_Cwp, ctx->d_Cw,
                     ctx->d_Cn, corr_only, ctx->d_flags);
  SP_HIP(hipGetLastError());
  return SP_OK;
}

}  // namespace

extern "C" {

int sp_ctx_footprint(const sp_su2_problem *p, long long *bytes) {
  if (!p || !bytes || p->nw < 2 || p->L < 1) return SP_EARG;
  const long long nf = (long long)p->nw * p->nw * (p->nw + 1) / 2;
  const long long emax = 2LL * p->nw + 8;
  long long b = 2 * nf * p->L * 16;            // Gamma, dGamma
  b += 2LL * p->nw * 8 + p->nw * 8;            // Sigma, dSigma, mesh
  b += 2 * (long long)p->nw * emax * 8 + 2 * emax * 8 + (p->nw + 1) * 4;  // weight tables
  b += emax * p->L * 16 + 2LL * p->B * p->Rb * 8;  // correlation scratch / output
  b += 4LL * (3 * p->L + 1 + 2 * (p->ovl_off ? p->ovl_off[p->L] : 0) + 2 * p->R + p->B * p->Rb);
  *bytes = b;
  return SP_OK;
}

int sp_ctx_create(const sp_su2_problem *p, sp_ctx **out) {
  if (!out) return SP_EARG;
  *out = nullptr;
  sp_ctx *ctx = new (std::nothrow) sp_ctx();
  if (!ctx) return SP_ENOMEM;
  int rc = check_problem(ctx, p);
  if (rc != SP_OK) { *out = ctx; return rc; }
  *out = ctx;
  SP_HIP(hipGetDevice(&ctx->device));
  SP_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  ctx->nw = p->nw;
  ctx->L = p->L;
  ctx->R = p->R;
  ctx->B = p->B;
  ctx->Rb = p->Rb;
  ctx->O = p->ovl_off[p->L];
  ctx->nf = (long long)p->nw * p->nw * (p->nw + 1) / 2;
  ctx->emax = 2 * p->nw + 8;
  ctx->spin = p->spin;
  launch_shape(ctx->nw, ctx->L, ctx->slots, ctx->nb, ctx->lds);
  if (ctx->lds > 64 * 1024) return fail(ctx, SP_EARG, "problem too large for the LDS plan of the dGamma kernel");
  if (ctx->nf * ctx->L >= (1LL << 31) / 2) return fail(ctx, SP_EARG, "vertex too large for 32-bit row indices");

  const size_t nv = (size_t)ctx->nf * ctx->L;
  const int emax = ctx->emax;
  if ((rc = dalloc(ctx, &ctx->d_mesh, ctx->nw)) || (rc = dalloc(ctx, &ctx->d_site, ctx->L)) ||
      (rc = dalloc(ctx, &ctx->d_inv, ctx->L)) || (rc = dalloc(ctx, &ctx->d_ovl_off, ctx->L + 1)) ||
      (rc = dalloc(ctx, &ctx->d_ovl1, ctx->O)) || (rc = dalloc(ctx, &ctx->d_ovl2, ctx->O)) ||
      (rc = dalloc(ctx, &ctx->d_rfwd, ctx->R)) || (rc = dalloc(ctx, &ctx->d_rinv, ctx->R)) ||
      (rc = dalloc(ctx, &ctx->d_corr, (size_t)ctx->B * ctx->Rb)) || (rc = dalloc(ctx, &ctx->d_G, nv)) ||
      (rc = dalloc(ctx, &ctx->d_dG, nv)) || (rc = dalloc(ctx, &ctx->d_v2, ctx->nw)) ||
      (rc = dalloc(ctx, &ctx->d_dv2, ctx->nw)) || (rc = dalloc(ctx, &ctx->d_Wwp, (size_t)ctx->nw * emax)) ||
      (rc = dalloc(ctx, &ctx->d_Wc, (size_t)ctx->nw * emax)) || (rc = dalloc(ctx, &ctx->d_Wn, ctx->nw)) ||
      (rc = dalloc(ctx, &ctx->d_Cwp, emax)) || (rc = dalloc(ctx, &ctx->d_Cw, emax)) ||
      (rc = dalloc(ctx, &ctx->d_Cn, 1)) || (rc = dalloc(ctx, &ctx->d_partial, (size_t)emax * ctx->L * 2)) ||
      (rc = dalloc(ctx, &ctx->d_corr_out, (size_t)2 * ctx->B * ctx->Rb)) || (rc = dalloc(ctx, &ctx->d_flags, 4)))
    return rc;

  std::vector<double> mesh(ctx->nw);
  for (int i = 0; i < ctx->nw; ++i) mesh[i] = (double)p->mesh[i];
  SP_HIP(hipMemcpy(ctx->d_mesh, mesh.data(), ctx->nw * 8, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(ctx->d_site, p->site_rid, ctx->L * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(ctx->d_inv, p->inv_rid, ctx->L * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(ctx->d_ovl_off, p->ovl_off, (ctx->L + 1) * 4, hipMemcpyHostToDevice));
  if (ctx->O > 0) {
    SP_HIP(hipMemcpy(ctx->d_ovl1, p->ovl_rid1, (size_t)ctx->O * 4, hipMemcpyHostToDevice));
    SP_HIP(hipMemcpy(ctx->d_ovl2, p->ovl_rid2, (size_t)ctx->O * 4, hipMemcpyHostToDevice));
  }
  SP_HIP(hipMemcpy(ctx->d_rfwd, p->range_fwd, (size_t)ctx->R * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(ctx->d_rinv, p->range_inv, (size_t)ctx->R * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(ctx->d_corr, p->corr_rid, (size_t)ctx->B * ctx->Rb * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemset(ctx->d_G, 0, nv * sizeof(double2)));
  SP_HIP(hipMemset(ctx->d_dG, 0, nv * sizeof(double2)));
  SP_HIP(hipMemset(ctx->d_v2, 0, ctx->nw * 8));
  SP_HIP(hipMemset(ctx->d_dv2, 0, ctx->nw * 8));
  SP_HIP(hipMemset(ctx->d_flags, 0, 16));
  SP_HIP(hipFuncSetAttribute((const void *)k_dgamma, hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024));
  return SP_OK;
}

int sp_ctx_destroy(sp_ctx *ctx) {
  if (!ctx) return SP_OK;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (int k = 0; k < 4; ++k)
    for (auto &pr : ctx->pending[k]) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
  for (auto e : ctx->free_events) hipEventDestroy(e);
  for (void *q : ctx->allocs) hipFree(q);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return SP_OK;
}

int sp_set_state(sp_ctx *ctx, double cutoff, const double *v2, const double *ss, const double *dd) {
  if (!ctx || !v2 || !ss || !dd) return fail(ctx, SP_EARG, "null argument");
  if (!std::isfinite(cutoff) || cutoff <= 0.0) return fail(ctx, SP_EARG, "cutoff must be finite and positive");
  const size_t nv = (size_t)ctx->nf * ctx->L;
  std::vector<double2> buf;
  try { buf.resize(nv); } catch (...) { return fail(ctx, SP_ENOMEM, "host buffer"); }
  for (size_t i = 0; i < nv; ++i) buf[i] = make_double2(ss[i], dd[i]);
  SP_HIP(hipMemcpyAsync(ctx->d_G, buf.data(), nv * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
  SP_HIP(hipMemcpyAsync(ctx->d_v2, v2, ctx->nw * 8, hipMemcpyHostToDevice, ctx->stream));
  SP_HIP(hipStreamSynchronize(ctx->stream));
  ctx->cutoff = cutoff;
  ctx->have_flow = false;
  return SP_OK;
}

static int download_pair(sp_ctx *ctx, const double2 *src, double *a, double *b) {
  const size_t nv = (size_t)ctx->nf * ctx->L;
  std::vector<double2> buf;
  try { buf.resize(nv); } catch (...) { return fail(ctx, SP_ENOMEM, "host buffer"); }
  SP_HIP(hipMemcpyAsync(buf.data(), src, nv * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
  SP_HIP(hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i < nv; ++i) {
    if (a) a[i] = buf[i].x;
    if (b) b[i] = buf[i].y;
  }
  return SP_OK;
}

int sp_get_state(sp_ctx *ctx, double *cutoff, double *v2, double *ss, double *dd) {
  if (!ctx) return SP_EARG;
  if (cutoff) *cutoff = ctx->cutoff;
  if (v2) {
    SP_HIP(hipMemcpyAsync(v2, ctx->d_v2, ctx->nw * 8, hipMemcpyDeviceToHost, ctx->stream));
    SP_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (ss || dd) return download_pair(ctx, ctx->d_G, ss, dd);
  return SP_OK;
}

int sp_get_flow(sp_ctx *ctx, double *dv2, double *dss, double *ddd) {
  if (!ctx) return SP_EARG;
  if (!ctx->have_flow) return fail(ctx, SP_ESTATE, "no flow computed since the last state change");
  if (dv2) {
    SP_HIP(hipMemcpyAsync(dv2, ctx->d_dv2, ctx->nw * 8, hipMemcpyDeviceToHost, ctx->stream));
    SP_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (dss || ddd) return download_pair(ctx, ctx->d_dG, dss, ddd);
  return SP_OK;
}

int sp_compute_step(sp_ctx *ctx) {
  if (!ctx) return SP_EARG;
  int rc;
  SP_HIP(hipMemsetAsync(ctx->d_flags, 0, 16, ctx->stream));
  {
    ProfScope ps(ctx, 1);
    hipLaunchKernelGGL(k_dsigma, dim3(ctx->nw), dim3(64), 0, ctx->stream, ctx->d_G, ctx->d_mesh, ctx->d_v2,
                       ctx->d_rfwd, ctx->d_rinv, ctx->R, ctx->nw, ctx->L, ctx->cutoff, ctx->spin, ctx->d_dv2,
                       ctx->d_flags);
    SP_HIP(hipGetLastError());
  }
  if ((rc = weights(ctx, 0)) != SP_OK) return rc;
  GammaArgs g;
  g.G = ctx->d_G;
  g.dG = ctx->d_dG;
  g.mesh = ctx->d_mesh;
  g.site_rid = ctx->d_site;
  g.inv_rid = ctx->d_inv;
  g.ovl_off = ctx->d_ovl_off;
  g.ovl_rid1 = ctx->d_ovl1;
  g.ovl_rid2 = ctx->d_ovl2;
  g.Wwp = ctx->d_Wwp;
  g.Wc = ctx->d_Wc;
  g.Wn = ctx->d_Wn;
  g.flags = ctx->d_flags;
  g.nw = ctx->nw;
  g.L = ctx->L;
  g.emax = ctx->emax;
  g.nb = ctx->nb;
  g.slots = ctx->slots;
  g.two_s = 2.0 * ctx->spin;
  g.eight_s = 8.0 * ctx->spin;
  {
    ProfScope ps(ctx, 0);
    const long long chunk = 1LL << 30;
    for (long long i0 = 0; i0 < ctx->nf; i0 += chunk) {
      const long long n = ctx->nf - i0 < chunk ? ctx->nf - i0 : chunk;
      g.item0 = (int)i0;
      hipLaunchKernelGGL(k_dgamma, dim3((unsigned)n), dim3(kThreads), ctx->lds, ctx->stream, g);
      SP_HIP(hipGetLastError());
    }
  }
  ctx->have_flow = true;
  return SP_OK;
}

int sp_finalize_step(sp_ctx *ctx, double new_cutoff) {
  if (!ctx) return SP_EARG;
  if (!ctx->have_flow) return fail(ctx, SP_ESTATE, "finalize_step without a computed flow");
  if (!std::isfinite(new_cutoff) || new_cutoff <= 0.0) return fail(ctx, SP_EARG, "cutoff must be finite and positive");
  const double step = new_cutoff - ctx->cutoff;
  const size_t nv = (size_t)ctx->nf * ctx->L;
  {
    ProfScope ps(ctx, 3);
    size_t blocks = (nv + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_euler, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, ctx->d_G, ctx->d_dG, nv,
                       ctx->d_v2, ctx->d_dv2, ctx->nw, step);
    SP_HIP(hipGetLastError());
  }
  ctx->cutoff = new_cutoff;
  ctx->have_flow = false;
  return SP_OK;
}

int sp_is_diverged(sp_ctx *ctx, int *diverged) {
  if (!ctx || !diverged) return SP_EARG;
  if (!ctx->have_flow) return fail(ctx, SP_ESTATE, "no flow computed");
  int flags[4];
  SP_HIP(hipMemcpyAsync(flags, ctx->d_flags, 16, hipMemcpyDeviceToHost, ctx->stream));
  SP_HIP(hipStreamSynchronize(ctx->stream));
  if (flags[1]) return fail(ctx, SP_EARG, "evaluation list overflow");
  *diverged = flags[0] ? 1 : 0;
  return SP_OK;
}

int sp_measure_correlation(sp_ctx *ctx, double *zz, double *dd) {
  if (!ctx || !zz || !dd) return SP_EARG;
  int rc;
  if ((rc = weights(ctx, 1)) != SP_OK) return rc;
  hipLaunchKernelGGL(k_corr_partial, dim3(ctx->emax), dim3(kThreads), 0, ctx->stream, ctx->d_G, ctx->d_mesh,
                     ctx->d_v2, ctx->d_site, ctx->d_inv, ctx->d_Cwp, ctx->d_Cw, ctx->d_Cn, ctx->nw, ctx->L,
                     kThreads / ctx->L, ctx->spin, ctx->d_partial);
  SP_HIP(hipGetLastError());
  const int nout = ctx->B * ctx->Rb;
  hipLaunchKernelGGL(k_corr_final, dim3(1), dim3(256), 0, ctx->stream, ctx->d_partial, ctx->d_Cn, ctx->L,
                     ctx->d_corr, nout, ctx->d_corr_out);
  SP_HIP(hipGetLastError());
  std::vector<double> buf(2 * (size_t)nout);
  SP_HIP(hipMemcpyAsync(buf.data(), ctx->d_corr_out, 2 * (size_t)nout * 8, hipMemcpyDeviceToHost, ctx->stream));
  SP_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(zz, buf.data(), (size_t)nout * 8);
  std::memcpy(dd, buf.data() + nout, (size_t)nout * 8);
  return SP_OK;
}

int sp_synchronize(sp_ctx *ctx) {
  if (!ctx) return SP_EARG;
  SP_HIP(hipStreamSynchronize(ctx->stream));
  return SP_OK;
}

const char *sp_last_error(const sp_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sp_profile_enable(sp_ctx *ctx, int enable) {
  if (!ctx) return SP_EARG;
  ctx->prof = enable != 0;
  return SP_OK;
}

int sp_profile_get(sp_ctx *ctx, int kernel, double *total_ms, long *launches) {
  if (!ctx || kernel < 0 || kernel > 3) return SP_EARG;
  for (auto &pr : ctx->pending[kernel]) {
    SP_HIP(hipEventSynchronize(pr.second));
    float ms = 0.0f;
    SP_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
    ctx->prof_ms[kernel] += ms;
    ctx->prof_n[kernel] += 1;
    ctx->free_events.push_back(pr.first);
    ctx->free_events.push_back(pr.second);
  }
  ctx->pending[kernel].clear();
  if (total_ms) *total_ms = ctx->prof_ms[kernel];
  if (launches) *launches = ctx->prof_n[kernel];
  return SP_OK;
}

}  // extern "C"
