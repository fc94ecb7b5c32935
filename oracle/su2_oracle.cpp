// CPU restatement of the reference SU(2) pf-FRG right-hand side -- TEST
// INFRASTRUCTURE ONLY.  It is the checker the HIP path is compared against and
// the scalar/OpenMP CPU baseline; nothing in the product links or calls it.
//
// Templated on the arithmetic type T:
//   T = float  reproduces the reference arithmetic (all-float state,
//              -ffp-contract=off, sequential per-item operation order) and is
//              pinned against the reference's golden .obs files;
//   T = double the FP64 oracle for the GPU FP64 kernels.  Mesh and cutoff
//              values are the float grids promoted to double.
//
// Reference files followed (paths relative to the reference repository):
//   src/FrequencyDiscretization.hpp:253-356   lesser / greater / offset / interpolateOffset
//   src/lib/Integrator.hpp:137-287            ImplicitIntegrator (trapezoid, off-mesh ends)
//   src/lib/ValueBundle.hpp:84-227            bundle multAdd / multSub order of operations
//   src/SU2/SU2VertexTwoParticle.hpp:185-632  storage, sign/site-exchange folding, access buffers
//   src/SU2/SU2VertexSingleParticle.hpp:73-87 self-energy interpolation
//   src/SU2/SU2FrgCore.cpp:111-431            finalizeStep, dSigma, dGamma (S/T/U, RPA, Katanin)
//   src/SU2/SU2MeasurementCorrelation.cpp:77-177  static susceptibility
//   src/lib/LoadManager.hpp:551-557           omp parallel for schedule(guided) over items
//
// Parity status: pinned end to end (dSigma, dGamma, Euler update, correlation)
// by the reference's golden susceptibility flows (T = float); the FP64
// instantiation is the same code path.  The integrators are pinned by the
// reference's integrator known answers.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Problem {
  int nw;
  const float *mesh;       // nw positive mesh points (float32 grid)
  int L;                   // representative sites
  const int *site_rid;     // getSites()[j].rid, j < L
  const int *inv_rid;      // getInvertedSites()[j].rid, j < L
  const int *ovl_off;      // L+1 CSR offsets of the overlap lists
  const int *ovl_rid1;
  const int *ovl_rid2;
  int R;                   // sites in range of the origin
  const int *range_fwd;    // symmetryTransform(0, j) for j in range(0)
  const int *range_inv;    // symmetryTransform(j, 0)
  int B;                   // basis sites
  int Rb;                  // sites in range of each basis site
  const int *corr_rid;     // B*Rb: symmetryTransform(basis b, j in range(b))
  double spin;             // spin length S
};

template <class T> struct Mesh {
  int n;
  std::vector<T> v;  // 2n values, v[n + i] = mesh point i, i in [-n, n)
  explicit Mesh(const Problem &p) : n(p.nw), v(2 * p.nw) {
    for (int i = 0; i < n; ++i) {
      v[n + i] = T(p.mesh[i]);
      v[n - 1 - i] = -T(p.mesh[i]);
    }
  }
  T at(int i) const { return v[n + i]; }
  int lesser(T w) const {
    if (w < 0) return -(greater(-w) + 1);
    if (w <= at(0)) return 0;
    for (int i = 1; i < n; ++i) if (at(i) > w) return i - 1;
    return n - 1;
  }
  int greater(T w) const {
    if (w < 0) return -(lesser(-w) + 1);
    if (w <= at(0)) return 0;
    for (int i = 1; i < n; ++i) if (at(i) > w) return i;
    return n - 1;
  }
  int offset(T w) const {
    if (w <= at(0)) return 0;
    for (int i = 1; i < n; ++i) if (at(i) >= w) return i;
    return n - 1;
  }
  void interp(T w, int &lo, int &hi, T &bias) const {
    if (w <= at(0)) { lo = 0; hi = 0; bias = T(0); return; }
    for (int i = 1; i < n; ++i) {
      if (at(i) > w) {
        hi = i; lo = i - 1;
        bias = (w - at(lo)) / (at(hi) - at(lo));
        return;
      }
    }
    lo = n - 1; hi = n - 1; bias = T(0);
  }
};

// L-length two-channel bundle (spin, density)
template <class T> struct Bundle {
  int L;
  std::vector<T> d;
  explicit Bundle(int L_) : L(L_), d(2 * L_, T(0)) {}
  T *s() { return d.data(); }
  T *dd() { return d.data() + L; }
  void reset() { std::fill(d.begin(), d.end(), T(0)); }
  void mult_add(T a, const Bundle &b) { for (int i = 0; i < 2 * L; ++i) d[i] += a * b.d[i]; }
  void add(const Bundle &b) { for (int i = 0; i < 2 * L; ++i) d[i] += b.d[i]; }
  void scale(T a) { for (int i = 0; i < 2 * L; ++i) d[i] *= a; }
  void div(T a) { for (int i = 0; i < 2 * L; ++i) d[i] /= a; }
};

template <int N, class T> struct AccessBuffer {
  int off[N];
  T w[N];
  int sign[N];
  bool exch = false;
};

template <class T> struct Ctx {
  const Problem &p;
  Mesh<T> mesh;
  int L, nw, LT;
  const T *v2;     // self-energy, nw
  const T *dv2;    // its flow (Katanin term), nw
  const T *ss;     // spin vertex
  const T *dd;     // density vertex
  T S;
  Ctx(const Problem &p_, const T *v2_, const T *dv2_, const T *ss_, const T *dd_)
      : p(p_), mesh(p_), L(p_.L), nw(p_.nw), LT(p_.L * p_.nw), v2(v2_), dv2(dv2_), ss(ss_), dd(dd_),
        S(T(p_.spin)) {}

  // SU2VertexSingleParticle::getValue
  T sigma(const T *data, T w) const {
    T sign = T(1);
    if (w < 0) { w = -w; sign = T(-1); }
    int lo, hi; T bias;
    mesh.interp(w, lo, hi, bias);
    return sign * ((1 - bias) * data[lo] + bias * data[hi]);
  }

  int buffer_offset(int so, int to, int uo, int &sign) const {
    if (so < uo) { sign = -1; return LT * (uo * (uo + 1) / 2 + so) + to * L; }
    sign = 1;
    return LT * (so * (so + 1) / 2 + uo) + to * L;
  }

  // generateAccessBuffer(s, t, u, channel); channel 0 = S, 1 = T, 2 = U
  AccessBuffer<4, T> ab4(T s, T t, T u, int channel) const {
    AccessBuffer<4, T> ab;
    if (s < 0 && u < 0) { s = -s; u = -u; }
    else {
      if (s < 0) { s = -s; ab.exch = true; }
      else if (u < 0) { u = -u; ab.exch = true; }
    }
    if (t < 0) t = -t;
    int lo1, hi1, lo2, hi2; T b1, b2;
    if (channel == 0) {
      int ex = mesh.offset(s);
      mesh.interp(t, lo1, hi1, b1);  // t
      mesh.interp(u, lo2, hi2, b2);  // u
      ab.w[0] = (1 - b2) * (1 - b1); ab.off[0] = buffer_offset(ex, lo1, lo2, ab.sign[0]);
      ab.w[1] = (1 - b2) * b1;       ab.off[1] = buffer_offset(ex, hi1, lo2, ab.sign[1]);
      ab.w[2] = b2 * (1 - b1);       ab.off[2] = buffer_offset(ex, lo1, hi2, ab.sign[2]);
      ab.w[3] = b2 * b1;             ab.off[3] = buffer_offset(ex, hi1, hi2, ab.sign[3]);
    } else if (channel == 1) {
      int ex = mesh.offset(t);
      mesh.interp(s, lo1, hi1, b1);  // s
      mesh.interp(u, lo2, hi2, b2);  // u
      ab.w[0] = (1 - b2) * (1 - b1); ab.off[0] = buffer_offset(lo1, ex, lo2, ab.sign[0]);
      ab.w[1] = (1 - b2) * b1;       ab.off[1] = buffer_offset(hi1, ex, lo2, ab.sign[1]);
      ab.w[2] = b2 * (1 - b1);       ab.off[2] = buffer_offset(lo1, ex, hi2, ab.sign[2]);
      ab.w[3] = b2 * b1;             ab.off[3] = buffer_offset(hi1, ex, hi2, ab.sign[3]);
    } else {
      int ex = mesh.offset(u);
      mesh.interp(s, lo1, hi1, b1);  // s
      mesh.interp(t, lo2, hi2, b2);  // t
      ab.w[0] = (1 - b2) * (1 - b1); ab.off[0] = buffer_offset(lo1, lo2, ex, ab.sign[0]);
      ab.w[1] = (1 - b2) * b1;       ab.off[1] = buffer_offset(hi1, lo2, ex, ab.sign[1]);
      ab.w[2] = b2 * (1 - b1);       ab.off[2] = buffer_offset(lo1, hi2, ex, ab.sign[2]);
      ab.w[3] = b2 * b1;             ab.off[3] = buffer_offset(hi1, hi2, ex, ab.sign[3]);
    }
    return ab;
  }

  // generateAccessBuffer(s, t, u): trilinear, 8 supports
  AccessBuffer<8, T> ab8(T s, T t, T u) const {
    AccessBuffer<8, T> ab;
    if (s < 0 && u < 0) { s = -s; u = -u; }
    else {
      if (s < 0) { s = -s; ab.exch = true; }
      else if (u < 0) { u = -u; ab.exch = true; }
    }
    if (t < 0) t = -t;
    int lS, uS, lT, uT, lU, uU; T bS, bT, bU;
    mesh.interp(s, lS, uS, bS);
    mesh.interp(t, lT, uT, bT);
    mesh.interp(u, lU, uU, bU);
    ab.w[0] = (1 - bT) * (1 - bS) * (1 - bU); ab.off[0] = buffer_offset(lS, lT, lU, ab.sign[0]);
    ab.w[1] = (1 - bT) * bS * (1 - bU);       ab.off[1] = buffer_offset(uS, lT, lU, ab.sign[1]);
    ab.w[2] = bT * (1 - bS) * (1 - bU);       ab.off[2] = buffer_offset(lS, uT, lU, ab.sign[2]);
    ab.w[3] = bT * bS * (1 - bU);             ab.off[3] = buffer_offset(uS, uT, lU, ab.sign[3]);
    ab.w[4] = (1 - bT) * (1 - bS) * bU;       ab.off[4] = buffer_offset(lS, lT, uU, ab.sign[4]);
    ab.w[5] = (1 - bT) * bS * bU;             ab.off[5] = buffer_offset(uS, lT, uU, ab.sign[5]);
    ab.w[6] = bT * (1 - bS) * bU;             ab.off[6] = buffer_offset(lS, uT, uU, ab.sign[6]);
    ab.w[7] = bT * bS * bU;                   ab.off[7] = buffer_offset(uS, uT, uU, ab.sign[7]);
    return ab;
  }

  template <int N> void superbundle(const AccessBuffer<N, T> &ab, Bundle<T> &out) const {
    out.reset();
    const int *sites = ab.exch ? p.inv_rid : p.site_rid;
    T *os = out.s(), *od = out.dd();
    for (int i = 0; i < N; ++i) {
      T weight = ab.w[i];
      T signed_weight = T(ab.sign[i]) * ab.w[i];
      int fo = ab.off[i];
      for (int j = 0; j < L; ++j) {
        os[j] += weight * ss[fo + sites[j]];
        od[j] += signed_weight * dd[fo + sites[j]];
      }
    }
  }

  T local_spin(const AccessBuffer<4, T> &ab) const {
    T v = 0;
    for (int i = 0; i < 4; ++i) v += ab.w[i] * ss[ab.off[i]];
    return v;
  }
  T local_density(const AccessBuffer<4, T> &ab) const {
    T v = 0;
    for (int i = 0; i < 4; ++i) v += T(ab.sign[i]) * ab.w[i] * dd[ab.off[i]];
    return v;
  }

  // _directAccessMapFrequencyExchange
  T direct(int rid, int so, int to, int uo, bool density) const {
    const T *data = density ? dd : ss;
    if (so >= uo) return data[LT * (so * (so + 1) / 2 + uo) + L * to + rid];
    T v = data[LT * (uo * (uo + 1) / 2 + so) + L * to + rid];
    return density ? -v : v;
  }

  // getValue(i1, i2, s, t, u, symmetry, FrequencyChannel::None); rid_fwd / rid_inv
  // are symmetryTransform(i1, i2) and symmetryTransform(i2, i1)
  T value_none(int rid_fwd, int rid_inv, T s, T t, T u, bool density) const {
    bool swap = false;
    if (s < 0 && u < 0) { s = -s; u = -u; }
    else {
      if (s < 0) { s = -s; swap = true; }
      else if (u < 0) { u = -u; swap = true; }
    }
    if (t < 0) t = -t;
    int rid = swap ? rid_inv : rid_fwd;
    int lS, uS, lT, uT, lU, uU; T bS, bT, bU;
    mesh.interp(s, lS, uS, bS);
    mesh.interp(t, lT, uT, bT);
    mesh.interp(u, lU, uU, bU);
    auto A = [&](int a, int b, int c) { return direct(rid, a, b, c, density); };
    return (1 - bU) * ((1 - bT) * ((1 - bS) * A(lS, lT, lU) + bS * A(uS, lT, lU)) +
                       bT * ((1 - bS) * A(lS, uT, lU) + bS * A(uS, uT, lU))) +
           bU * ((1 - bT) * ((1 - bS) * A(lS, lT, uU) + bS * A(uS, lT, uU)) +
                 bT * ((1 - bS) * A(lS, uT, uU) + bS * A(uS, uT, uU)));
  }
};

// ImplicitIntegrator (src/lib/Integrator.hpp:137-287).  f(w, buffer) fills buffer;
// calls counts the integrand evaluations.
template <class T, class F>
void integrate_left(const Mesh<T> &m, T min, int max, F &&f, Bundle<T> &ib, Bundle<T> &rb, long &calls) {
  rb.reset();
  int umin = m.greater(min);
  if (umin != max) {
    f(min, ib); ++calls; rb.mult_add(m.at(umin) - min, ib);
    f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - min, ib);
    while (++umin != max) { f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - m.at(umin - 1), ib); }
    f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin) - m.at(umin - 1), ib);
    rb.scale(T(0.5));
  } else {
    f(min, ib); ++calls; rb.add(ib);
    f(m.at(umin), ib); ++calls; rb.add(ib);
    rb.scale(T(0.5) * (m.at(umin) - min));
  }
}

template <class T, class F>
void integrate_right(const Mesh<T> &m, int min, T max, F &&f, Bundle<T> &ib, Bundle<T> &rb, long &calls) {
  rb.reset();
  int umin = min;
  int umax = m.lesser(max);
  if (umin != umax) {
    f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - m.at(umin), ib);
    while (++umin != umax) { f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - m.at(umin - 1), ib); }
    f(m.at(umin), ib); ++calls; rb.mult_add(max - m.at(umin - 1), ib);
    f(max, ib); ++calls; rb.mult_add(max - m.at(umin), ib);
    rb.scale(T(0.5));
  } else {
    f(max, ib); ++calls; rb.add(ib);
    f(m.at(umin), ib); ++calls; rb.add(ib);
    rb.scale(T(0.5) * (max - m.at(umin)));
  }
}

template <class T, class F>
void integrate_both(const Mesh<T> &m, T min, T max, F &&f, Bundle<T> &ib, Bundle<T> &rb, long &calls) {
  rb.reset();
  int umin = m.greater(min);
  int umax = m.lesser(max);
  if (umax >= umin) {
    f(min, ib); ++calls; rb.mult_add(m.at(umin) - min, ib);
    if (umax != umin) {
      f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - min, ib);
      while (++umin != umax) { f(m.at(umin), ib); ++calls; rb.mult_add(m.at(umin + 1) - m.at(umin - 1), ib); }
      f(m.at(umin), ib); ++calls; rb.mult_add(max - m.at(umin - 1), ib);
    } else {
      f(m.at(umin), ib); ++calls; rb.mult_add(max - min, ib);
    }
    f(max, ib); ++calls; rb.mult_add(max - m.at(umin), ib);
    rb.scale(T(0.5));
  } else {
    f(max, ib); ++calls; rb.add(ib);
    f(min, ib); ++calls; rb.add(ib);
    rb.scale(T(0.5) * (max - min));
  }
}

template <class T> T two_pi() { return T(2.0) * T(M_PI); }
template <> float two_pi<float>() { return 2.0f * (float)M_PI; }

// per-thread scratch for one dGamma item
template <class T> struct ItemScratch {
  Bundle<T> b1, b2, rpa, cur, sb0, sb1, sb2, sb3;
  explicit ItemScratch(int L)
      : b1(L), b2(L), rpa(L), cur(L), sb0(L), sb1(L), sb2(L), sb3(L) {}
};

// SU2FrgCore::_calculateVertexTwoParticle (src/SU2/SU2FrgCore.cpp:171-431)
template <class T>
void vertex_two_particle(const Ctx<T> &c, int it, ItemScratch<T> &sc, T cutoff, T *out_s, T *out_d,
                         long *calls) {
  const int L = c.L, nw = c.nw;
  const Mesh<T> &m = c.mesh;
  // expandIterator(iterator, s, t, u) (SU2VertexTwoParticle.hpp:136-158)
  int su = it / nw;
  T t = m.at(it % nw);
  int so = 0;
  while ((so + 1) * (so + 2) / 2 <= su) ++so;
  int uo = su - so * (so + 1) / 2;
  T s = m.at(so), u = m.at(uo);

  const T w1p = T(0.5) * (s + t + u);
  const T w1 = T(0.5) * (s - t + u);
  const T w2p = T(0.5) * (s - t - u);
  const T w2 = T(0.5) * (s + t - u);
  const T S = c.S;

  auto kS = [&](T wp, Bundle<T> &ret) {
    c.superbundle(c.ab4(s, -w1 - wp, -w2 - wp, 0), sc.sb0);
    c.superbundle(c.ab4(s, w1p + wp, -w2p - wp, 0), sc.sb1);
    c.superbundle(c.ab4(s, w2 + wp, w1 + wp, 0), sc.sb2);
    c.superbundle(c.ab4(s, -w2p - wp, w1p + wp, 0), sc.sb3);
    ret.reset();
    T *rs = ret.s(), *rd = ret.dd();
    T *a0s = sc.sb0.s(), *a0d = sc.sb0.dd(), *a1s = sc.sb1.s(), *a1d = sc.sb1.dd();
    T *a2s = sc.sb2.s(), *a2d = sc.sb2.dd(), *a3s = sc.sb3.s(), *a3d = sc.sb3.dd();
    const T h = T(0.5), c316 = T(3.0) / T(16.0);
    for (int i = 0; i < L; ++i) rs[i] -= h * a0s[i] * a1s[i];
    for (int i = 0; i < L; ++i) rs[i] -= h * a2s[i] * a3s[i];
    for (int i = 0; i < L; ++i) rs[i] += a0d[i] * a1s[i];
    for (int i = 0; i < L; ++i) rs[i] += a2d[i] * a3s[i];
    for (int i = 0; i < L; ++i) rs[i] += a0s[i] * a1d[i];
    for (int i = 0; i < L; ++i) rs[i] += a2s[i] * a3d[i];
    for (int i = 0; i < L; ++i) rd[i] += c316 * a0s[i] * a1s[i];
    for (int i = 0; i < L; ++i) rd[i] += c316 * a2s[i] * a3s[i];
    for (int i = 0; i < L; ++i) rd[i] += a0d[i] * a1d[i];
    for (int i = 0; i < L; ++i) rd[i] += a2d[i] * a3d[i];
  };

  auto kT = [&](T wp, Bundle<T> &ret) {
    c.superbundle(c.ab4(w1 - wp, t, w1p + wp, 1), sc.sb0);
    c.superbundle(c.ab4(w2p - wp, t, -w2 - wp, 1), sc.sb1);
    c.superbundle(c.ab4(w1p + wp, t, w1 - wp, 1), sc.sb2);
    c.superbundle(c.ab4(w2 + wp, t, wp - w2p, 1), sc.sb3);
    ret.reset();
    sc.rpa.reset();
    T *ps = sc.rpa.s(), *pd = sc.rpa.dd();
    T *a2s = sc.sb2.s(), *a2d = sc.sb2.dd(), *a3s = sc.sb3.s(), *a3d = sc.sb3.dd();
    for (int rid = 0; rid < L; ++rid) {
      const int b = c.p.ovl_off[rid], e = c.p.ovl_off[rid + 1];
      for (int i = b; i < e; ++i) ps[rid] += a2s[c.p.ovl_rid1[i]] * a3s[c.p.ovl_rid2[i]];
      for (int i = b; i < e; ++i) pd[rid] += a2d[c.p.ovl_rid1[i]] * a3d[c.p.ovl_rid2[i]];
    }
    T *rs = ret.s(), *rd = ret.dd();
    const T twoS = T(2.0) * S, eightS = T(8.0) * S;
    for (int i = 0; i < L; ++i) rs[i] += twoS * ps[i];
    for (int i = 0; i < L; ++i) rd[i] += eightS * pd[i];

    const auto ab4 = c.ab4(w2p - wp, -w2 - wp, t, 2);
    const auto ab5 = c.ab4(w1 - wp, -w1p - wp, -t, 2);
    const auto ab6 = c.ab4(w2 + wp, wp - w2p, t, 2);
    const auto ab7 = c.ab4(w1p + wp, wp - w1, -t, 2);
    const T valCbs = c.local_spin(ab4), valCbd = c.local_density(ab4);
    const T valICas = c.local_spin(ab5), valICad = c.local_density(ab5);
    const T valCbs2 = c.local_spin(ab6), valCbd2 = c.local_density(ab6);
    const T valICas2 = c.local_spin(ab7), valICad2 = c.local_density(ab7);
    T *a0s = sc.sb0.s(), *a0d = sc.sb0.dd(), *a1s = sc.sb1.s(), *a1d = sc.sb1.dd();
    const T q = T(0.25), tq = T(0.75);
    T f;
    for (int i = 0; i < L; ++i) rs[i] -= valCbd * a0s[i];
    f = q * valCbs;   for (int i = 0; i < L; ++i) rs[i] += f * a0s[i];
    for (int i = 0; i < L; ++i) rs[i] -= valICad * a1s[i];
    f = q * valICas;  for (int i = 0; i < L; ++i) rs[i] += f * a1s[i];
    for (int i = 0; i < L; ++i) rs[i] -= valCbd2 * a2s[i];
    f = q * valCbs2;  for (int i = 0; i < L; ++i) rs[i] += f * a2s[i];
    for (int i = 0; i < L; ++i) rs[i] -= valICad2 * a3s[i];
    f = q * valICas2; for (int i = 0; i < L; ++i) rs[i] += f * a3s[i];

    for (int i = 0; i < L; ++i) rd[i] -= valCbd * a0d[i];
    f = tq * valCbs;   for (int i = 0; i < L; ++i) rd[i] -= f * a0d[i];
    for (int i = 0; i < L; ++i) rd[i] -= valICad * a1d[i];
    f = tq * valICas;  for (int i = 0; i < L; ++i) rd[i] -= f * a1d[i];
    for (int i = 0; i < L; ++i) rd[i] -= valCbd2 * a2d[i];
    f = tq * valCbs2;  for (int i = 0; i < L; ++i) rd[i] -= f * a2d[i];
    for (int i = 0; i < L; ++i) rd[i] -= valICad2 * a3d[i];
    f = tq * valICas2; for (int i = 0; i < L; ++i) rd[i] -= f * a3d[i];
  };

  auto kU = [&](T wp, Bundle<T> &ret) {
    c.superbundle(c.ab4(w1 + wp, wp - w2p, u, 2), sc.sb0);
    c.superbundle(c.ab4(w1p + wp, w2 - wp, u, 2), sc.sb1);
    c.superbundle(c.ab4(w2p - wp, -w1 - wp, u, 2), sc.sb2);
    c.superbundle(c.ab4(w2 - wp, w1p + wp, u, 2), sc.sb3);
    ret.reset();
    T *rs = ret.s(), *rd = ret.dd();
    T *a0s = sc.sb0.s(), *a0d = sc.sb0.dd(), *a1s = sc.sb1.s(), *a1d = sc.sb1.dd();
    T *a2s = sc.sb2.s(), *a2d = sc.sb2.dd(), *a3s = sc.sb3.s(), *a3d = sc.sb3.dd();
    const T h = T(0.5), c316 = T(3.0) / T(16.0);
    for (int i = 0; i < L; ++i) rs[i] += h * a0s[i] * a1s[i];
    for (int i = 0; i < L; ++i) rs[i] += h * a2s[i] * a3s[i];
    for (int i = 0; i < L; ++i) rs[i] += a0s[i] * a1d[i];
    for (int i = 0; i < L; ++i) rs[i] += a2s[i] * a3d[i];
    for (int i = 0; i < L; ++i) rs[i] += a0d[i] * a1s[i];
    for (int i = 0; i < L; ++i) rs[i] += a2d[i] * a3s[i];
    for (int i = 0; i < L; ++i) rd[i] += c316 * a0s[i] * a1s[i];
    for (int i = 0; i < L; ++i) rd[i] += c316 * a2s[i] * a3s[i];
    for (int i = 0; i < L; ++i) rd[i] += a0d[i] * a1d[i];
    for (int i = 0; i < L; ++i) rd[i] += a2d[i] * a3d[i];
  };

  auto p = [&](T a, T b) { return T(1) / ((a + c.sigma(c.v2, a)) * (b + c.sigma(c.v2, b))); };
  auto pK = [&](T a, T b) {
    T d1 = a + c.sigma(c.v2, a);
    return c.sigma(c.dv2, a) / (d1 * d1 * (b + c.sigma(c.v2, b)));
  };

  Bundle<T> &cur = sc.cur;
  cur.reset();
  long nS = 0, nT = 0, nU = 0;
  // conventional contribution (SU2FrgCore.cpp:349-371)
  kS(cutoff, sc.b1); ++nS; cur.mult_add(p(cutoff, cutoff + s), sc.b1);
  if (s > T(2.0) * cutoff) { kS(-cutoff, sc.b1); ++nS; cur.mult_add(p(cutoff, cutoff - s), sc.b1); }
  kT(cutoff, sc.b1); ++nT; cur.mult_add(p(cutoff, cutoff + t), sc.b1);
  if (t > T(2.0) * cutoff) { kT(-cutoff, sc.b1); ++nT; cur.mult_add(p(cutoff, cutoff - t), sc.b1); }
  kU(cutoff, sc.b1); ++nU; cur.mult_add(-p(cutoff, cutoff + u), sc.b1);
  if (u > T(2.0) * cutoff) { kU(-cutoff, sc.b1); ++nU; cur.mult_add(-p(cutoff, cutoff - u), sc.b1); }

  // Katanin contribution (SU2FrgCore.cpp:373-424)
  auto fS = [&](T wp, Bundle<T> &r) { kS(wp, r); r.scale(pK(wp, s + wp)); };
  auto fT = [&](T wp, Bundle<T> &r) { kT(wp, r); r.scale(pK(wp, t + wp)); };
  auto fU = [&](T wp, Bundle<T> &r) { kU(wp, r); r.scale(-pK(wp, u + wp)); };
  const T wmin = m.at(-nw), wmax = m.at(nw - 1);
  auto windows = [&](T X, auto &&f, long &n) {
    if (-(X + cutoff) > wmin) { integrate_right(m, -nw, -(X + cutoff), f, sc.b1, sc.b2, n); cur.add(sc.b2); }
    if (X - cutoff > cutoff) { integrate_both(m, cutoff - X, -cutoff, f, sc.b1, sc.b2, n); cur.add(sc.b2); }
    if (cutoff < wmax) { integrate_left(m, cutoff, nw - 1, f, sc.b1, sc.b2, n); cur.add(sc.b2); }
  };
  windows(s, fS, nS);
  windows(t, fT, nT);
  windows(u, fU, nU);

  cur.div(two_pi<T>());
  for (int rid = 0; rid < L; ++rid) out_s[(long)it * L + rid] = cur.s()[rid];
  for (int rid = 0; rid < L; ++rid) out_d[(long)it * L + rid] = cur.dd()[rid];
  if (calls) { calls[0] += nS; calls[1] += nT; calls[2] += nU; }
}

// SU2FrgCore::_calculateVertexSingleParticle (src/SU2/SU2FrgCore.cpp:139-169)
template <class T> T vertex_single_particle(const Ctx<T> &c, int i, T cutoff) {
  const T w = c.mesh.at(i);
  T v = 0;
  T sum = 0;
  for (int j = 0; j < c.p.R; ++j) {
    int f = c.p.range_fwd[j], r = c.p.range_inv[j];
    sum += c.value_none(f, r, w + cutoff, T(0), w - cutoff, true);
    sum -= c.value_none(f, r, w - cutoff, T(0), w + cutoff, true);
  }
  v -= T(4.0) * c.S * sum;
  v += T(0.75) * (c.value_none(0, 0, w + cutoff, w - cutoff, T(0), false) -
                  c.value_none(0, 0, w - cutoff, w + cutoff, T(0), false));
  v += (c.value_none(0, 0, w + cutoff, w - cutoff, T(0), true) -
        c.value_none(0, 0, w - cutoff, w + cutoff, T(0), true));
  v /= (two_pi<T>() * (cutoff + c.sigma(c.v2, cutoff)));
  return v;
}

template <class T>
void compute_step(const Problem &p, T cutoff, const T *v2, const T *ss, const T *dd, T *dv2, T *dss, T *ddd,
                  int nthreads, long *calls) {
  // dSigma first; the Katanin term of dGamma reads it (SU2FrgCore.cpp:94-108)
  {
    Ctx<T> c(p, v2, nullptr, ss, dd);
    for (int i = 0; i < p.nw; ++i) dv2[i] = vertex_single_particle(c, i, cutoff);
  }
  Ctx<T> c(p, v2, dv2, ss, dd);
  const int nf = p.nw * p.nw * (p.nw + 1) / 2;
  long cS = 0, cT = 0, cU = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : cS, cT, cU)
  {
    ItemScratch<T> sc(p.L);
    long local[3] = {0, 0, 0};
#pragma omp for schedule(guided)
    for (int it = 0; it < nf; ++it) vertex_two_particle(c, it, sc, cutoff, dss, ddd, local);
    cS += local[0]; cT += local[1]; cU += local[2];
  }
  if (calls) { calls[0] = cS; calls[1] = cT; calls[2] = cU; }
}

// SU2MeasurementCorrelation::_calculateCorrelation (src/SU2/SU2MeasurementCorrelation.cpp:77-177)
template <class T> void correlation(const Problem &p, T cut, const T *v2, const T *ss, const T *dd, T *zz, T *ddo) {
  Ctx<T> c(p, v2, nullptr, ss, dd);
  const Mesh<T> &m = c.mesh;
  const int L = p.L, nw = p.nw;
  const T nu = T(0);
  const T S = c.S;
  Bundle<T> susc(L), sb(L), buf1(L), buf2(L), buf3(L), buf4(L);
  long dummy = 0;
  const T wmin = m.at(-nw), wmax = m.at(nw - 1);
  T two_pi_f, pi_f, four_pi2;
  if (sizeof(T) == sizeof(float)) {
    two_pi_f = T(float(2.0f * M_PI)); pi_f = T(float(M_PI)); four_pi2 = T(float(4.0f * M_PI * M_PI));
  } else {
    two_pi_f = T(2.0 * M_PI); pi_f = T(M_PI); four_pi2 = T(4.0 * M_PI * M_PI);
  }

  auto outer = [&](T w, Bundle<T> &ret) {
    ret.reset();
    const T term1 = T(1) / ((w + c.sigma(v2, w)) * (w + nu + c.sigma(v2, w + nu)));
    ret.s()[0] += S * term1 / two_pi_f;
    ret.dd()[0] += T(2.0) * S * term1 / pi_f;
    auto inner = [&](T wp, Bundle<T> &r) {
      r.reset();
      c.superbundle(c.ab8(w + wp + nu, nu, w - wp), sb);
      const T vs = c.value_none(0, 0, w + wp + nu, w - wp, nu, false);
      const T vd = c.value_none(0, 0, w + wp + nu, w - wp, nu, true);
      const T s2 = S * S, s16 = T(16.0) * S * S;
      for (int i = 0; i < L; ++i) r.s()[i] -= s2 * sb.s()[i];
      for (int i = 0; i < L; ++i) r.dd()[i] -= s16 * sb.dd()[i];
      r.s()[0] += S * (-vs / T(4.0) + vd);
      r.dd()[0] += S * (T(3.0) * vs + T(4.0) * vd);
      const T norm = T(1) / ((w + c.sigma(v2, w)) * (w + nu + c.sigma(v2, w + nu)) * (wp + c.sigma(v2, wp)) *
                             (wp + nu + c.sigma(v2, wp + nu)) * four_pi2);
      r.scale(norm);
    };
    if (-(nu + cut) > wmin) { integrate_right(m, -nw, -cut - nu, inner, buf3, buf4, dummy); ret.add(buf4); }
    if (nu - cut > cut) { integrate_both(m, -nu + cut, -cut, inner, buf3, buf4, dummy); ret.add(buf4); }
    if (cut < wmax) { integrate_left(m, cut, nw - 1, inner, buf3, buf4, dummy); ret.add(buf4); }
  };
  susc.reset();
  if (-(nu + cut) > wmin) { integrate_right(m, -nw, -cut - nu, outer, buf1, buf2, dummy); susc.add(buf2); }
  if (nu - cut > cut) { integrate_both(m, -nu + cut, -cut, outer, buf1, buf2, dummy); susc.add(buf2); }
  if (cut < wmax) { integrate_left(m, cut, nw - 1, outer, buf1, buf2, dummy); susc.add(buf2); }
  for (int k = 0; k < p.B * p.Rb; ++k) {
    zz[k] = susc.s()[p.corr_rid[k]];
    ddo[k] = susc.dd()[p.corr_rid[k]];
  }
}

// SU2FrgCore::finalizeStep (src/SU2/SU2FrgCore.cpp:111-137)
template <class T> void finalize_step(long nv2, long nv4, T cutoff, T new_cutoff, T *v2, T *ss, T *dd,
                                      const T *dv2, const T *dss, const T *ddd) {
  const T step = new_cutoff - cutoff;
  for (long i = 0; i < nv2; ++i) v2[i] += step * dv2[i];
  for (long i = 0; i < nv4; ++i) dd[i] += step * ddd[i];
  for (long i = 0; i < nv4; ++i) ss[i] += step * dss[i];
}

template <class T> bool has_nan(long n, const T *x) {
  for (long i = 0; i < n; ++i) if (std::isnan(x[i])) return true;
  return false;
}

// Full flow loop (src/SpinParser.cpp:141-172): computeStep, measurement, NaN
// check on the derivative, Euler update; final measurement.  Writes one
// correlation record per cutoff reached; returns the number of records.
template <class T>
int run_flow(const Problem &p, int ncut, const float *cutoffs, T *v2, T *ss, T *dd, T *zz, T *ddo, int nthreads) {
  const long nf = (long)p.nw * p.nw * (p.nw + 1) / 2;
  const long nv4 = nf * p.L;
  std::vector<T> dv2(p.nw), dss(nv4), ddd(nv4);
  const long rec = (long)p.B * p.Rb;
  int n = 0;
  int k = 0;
  for (; k < ncut - 1; ++k) {
    const T cut = T(cutoffs[k]);
    compute_step<T>(p, cut, v2, ss, dd, dv2.data(), dss.data(), ddd.data(), nthreads, nullptr);
    correlation<T>(p, cut, v2, ss, dd, zz + n * rec, ddo + n * rec);
    ++n;
    if (has_nan(p.nw, dv2.data()) || has_nan(nv4, dss.data()) || has_nan(nv4, ddd.data())) return -n;
    finalize_step<T>(p.nw, nv4, cut, T(cutoffs[k + 1]), v2, ss, dd, dv2.data(), dss.data(), ddd.data());
  }
  correlation<T>(p, T(cutoffs[k]), v2, ss, dd, zz + n * rec, ddo + n * rec);
  ++n;
  return n;
}

}  // namespace

extern "C" {

// Problem descriptor as passed from Python (field order mirrors Problem).
typedef Problem or_problem;

void or_su2_compute_step_f64(const or_problem *p, double cutoff, const double *v2, const double *ss,
                             const double *dd, double *dv2, double *dss, double *ddd, int nthreads, long *calls) {
  compute_step<double>(*p, cutoff, v2, ss, dd, dv2, dss, ddd, nthreads, calls);
}

void or_su2_compute_step_f32(const or_problem *p, float cutoff, const float *v2, const float *ss, const float *dd,
                             float *dv2, float *dss, float *ddd, int nthreads, long *calls) {
  compute_step<float>(*p, cutoff, v2, ss, dd, dv2, dss, ddd, nthreads, calls);
}

// Items [it0, it1) of dGamma only, with dSigma given (for bounded timing samples).
void or_su2_items_f64(const or_problem *p, double cutoff, const double *v2, const double *dv2, const double *ss,
                      const double *dd, int it0, int it1, double *dss, double *ddd, int nthreads, long *calls) {
  Ctx<double> c(*p, v2, dv2, ss, dd);
  long cS = 0, cT = 0, cU = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : cS, cT, cU)
  {
    ItemScratch<double> sc(p->L);
    long local[3] = {0, 0, 0};
#pragma omp for schedule(guided)
    for (int it = it0; it < it1; ++it) vertex_two_particle(c, it, sc, cutoff, dss, ddd, local);
    cS += local[0]; cT += local[1]; cU += local[2];
  }
  if (calls) { calls[0] = cS; calls[1] = cT; calls[2] = cU; }
}

void or_su2_items_f32(const or_problem *p, float cutoff, const float *v2, const float *dv2, const float *ss,
                      const float *dd, int it0, int it1, float *dss, float *ddd, int nthreads, long *calls) {
  Ctx<float> c(*p, v2, dv2, ss, dd);
  long cS = 0, cT = 0, cU = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : cS, cT, cU)
  {
    ItemScratch<float> sc(p->L);
    long local[3] = {0, 0, 0};
#pragma omp for schedule(guided)
    for (int it = it0; it < it1; ++it) vertex_two_particle(c, it, sc, cutoff, dss, ddd, local);
    cS += local[0]; cT += local[1]; cU += local[2];
  }
  if (calls) { calls[0] = cS; calls[1] = cT; calls[2] = cU; }
}

void or_su2_dsigma_f64(const or_problem *p, double cutoff, const double *v2, const double *ss, const double *dd,
                       double *dv2) {
  Ctx<double> c(*p, v2, nullptr, ss, dd);
  for (int i = 0; i < p->nw; ++i) dv2[i] = vertex_single_particle(c, i, cutoff);
}

void or_su2_correlation_f64(const or_problem *p, double cutoff, const double *v2, const double *ss,
                            const double *dd, double *zz, double *ddo) {
  correlation<double>(*p, cutoff, v2, ss, dd, zz, ddo);
}

void or_su2_correlation_f32(const or_problem *p, float cutoff, const float *v2, const float *ss, const float *dd,
                            float *zz, float *ddo) {
  correlation<float>(*p, cutoff, v2, ss, dd, zz, ddo);
}

int or_su2_run_flow_f32(const or_problem *p, int ncut, const float *cutoffs, float *v2, float *ss, float *dd,
                        float *zz, float *ddo, int nthreads) {
  return run_flow<float>(*p, ncut, cutoffs, v2, ss, dd, zz, ddo, nthreads);
}

int or_su2_run_flow_f64(const or_problem *p, int ncut, const float *cutoffs, double *v2, double *ss, double *dd,
                        double *zz, double *ddo, int nthreads) {
  return run_flow<double>(*p, ncut, cutoffs, v2, ss, dd, zz, ddo, nthreads);
}

void or_su2_finalize_f64(long nv2, long nv4, double cutoff, double new_cutoff, double *v2, double *ss, double *dd,
                         const double *dv2, const double *dss, const double *ddd) {
  finalize_step<double>(nv2, nv4, cutoff, new_cutoff, v2, ss, dd, dv2, dss, ddd);
}

// Integrator hook for the known-answer tests: integrand (x, x*x) on the mesh.
// kind 0 = obscure left boundary [min, last], 1 = obscure right boundary
// [first positive, max], 2 = both boundaries [min, max]
void or_integrate_test_f32(int nw, const float *mesh, int kind, float a, float b, float *out2) {
  Problem p{};
  p.nw = nw; p.mesh = mesh; p.L = 1;
  Mesh<float> m(p);
  Bundle<float> ib(1), rb(1);
  long calls = 0;
  auto f = [](float x, Bundle<float> &r) { r.s()[0] = x; r.dd()[0] = x * x; };
  if (kind == 0) integrate_left(m, a, nw - 1, f, ib, rb, calls);
  else if (kind == 1) integrate_right(m, 0, b, f, ib, rb, calls);
  else integrate_both(m, a, b, f, ib, rb, calls);
  out2[0] = rb.s()[0];
  out2[1] = rb.dd()[0];
}

}  // extern "C"
