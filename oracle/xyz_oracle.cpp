// CPU restatement of the reference XYZ pf-FRG right-hand side (diagonal
// couplings J_xx, J_yy, J_zz; four vertex components X, Y, Z, D) -- TEST
// INFRASTRUCTURE ONLY.  It is the checker the HIP XYZ path is compared
// against; nothing in the product links or calls it.
//
// Templated on T: float follows the reference arithmetic (all-float state,
// -ffp-contract=off, the reference's term order per output component); double
// is the FP64 oracle for the GPU kernels.
//
// Reference files followed (paths relative to the reference repository):
//   src/XYZ/XYZVertexTwoParticle.hpp:63-77,157-640  storage (four planes), access
//                                                   buffers, density sign, site spin
//                                                   permutations
//   src/XYZ/XYZVertexSingleParticle.hpp:73-87       self-energy interpolation
//   src/XYZ/XYZFrgCore.cpp:128-193                  finalizeStep, dSigma
//   src/XYZ/XYZFrgCore.cpp:195-571                  dGamma: S, T (lattice bubble, chalice,
//                                                   inverse chalice), U, conventional, Katanin
//   src/XYZ/XYZMeasurementCorrelation.cpp:98-212    static susceptibility, 4 outputs
//   src/XYZ/XYZEffectiveAction.hpp:38-66            initial condition
//
// State layout (the reference's): v4 holds four planes X, Y, Z, D of N_f * L
// values each, plane c at v4 + c * N_f * L, entry ((s(s+1)/2 + u) * N_w + t) * L + rid.
//
// Parity status: pinned end to end by the reference's XYZ golden flow
// (Heisenberg-Kitaev hyperhoneycomb range 3, all four correlation outputs,
// T = float), by XYZ(Heisenberg) = SU2 on the same lattice and by
// XYZ = TRI on diagonal couplings (the reference's own cross-core check).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "frg_mesh.hpp"

namespace {

using frg::Mesh;

// the same lattice tables as the TRI core (tri_oracle.cpp's TriProblem)
struct XyzProblem {
  int nw;
  const float *mesh;
  int L;
  const int *site_rid;
  const int *site_perm;  // L x 3
  const int *inv_rid;
  const int *inv_perm;   // L x 3
  const int *ovl_off;    // L+1
  const int *ovl_rid1;
  const int *ovl_rid2;
  const int *ovl_perm1;  // O x 3: transformedX1, Y1, Z1
  const int *ovl_perm2;  // O x 3: transformedX2, Y2, Z2
  int R;
  const int *range_fwd;
  const int *range_inv;
  int B;
  int Rb;
  const int *corr_rid;
  const int *corr_perm;  // B*Rb x 3
};

constexpr int X = 0, Y = 1, Z = 2, D = 3;

template <int N, class T> struct XyzAB {
  int off[N];  // row offset (elements) within a plane
  T w[N];
  int sgn[N];  // density sign of the s <-> u exchanged storage
  bool exch = false;
};

template <class T> struct Bundle4 {
  int L;
  std::vector<T> d;
  explicit Bundle4(int L_) : L(L_), d(4 * L_, T(0)) {}
  T *b(int c) { return d.data() + c * L; }
  const T *b(int c) const { return d.data() + c * L; }
  void reset() { std::fill(d.begin(), d.end(), T(0)); }
  void mult_add(T a, const Bundle4 &x) { for (size_t i = 0; i < d.size(); ++i) d[i] += a * x.d[i]; }
  void add(const Bundle4 &x) { for (size_t i = 0; i < d.size(); ++i) d[i] += x.d[i]; }
  void scale(T a) { for (size_t i = 0; i < d.size(); ++i) d[i] *= a; }
  void div(T a) { for (size_t i = 0; i < d.size(); ++i) d[i] /= a; }
};

template <class T> struct XyzCtx {
  const XyzProblem &p;
  Mesh<T> mesh;
  int L, nw;
  long nv;  // plane size N_f * L
  const T *v2, *dv2, *v4;
  XyzCtx(const XyzProblem &p_, const T *v2_, const T *dv2_, const T *v4_)
      : p(p_), mesh(p_.nw, p_.mesh), L(p_.L), nw(p_.nw), nv((long)p_.nw * p_.nw * (p_.nw + 1) / 2 * p_.L),
        v2(v2_), dv2(dv2_), v4(v4_) {}

  const T *plane(int c) const { return v4 + c * nv; }

  T sigma(const T *data, T w) const {
    T sign = T(1);
    if (w < 0) { w = -w; sign = T(-1); }
    int lo, hi; T bias;
    mesh.interp(w, lo, hi, bias);
    return sign * ((1 - bias) * data[lo] + bias * data[hi]);
  }

  // _generateAccessBufferOffset (XYZVertexTwoParticle.hpp:637-653)
  int support(int so, int to, int uo, int &sgn) const {
    sgn = so < uo ? -1 : 1;
    if (so < uo) std::swap(so, uo);
    return (int)(((long)(so * (so + 1) / 2 + uo) * nw + to) * L);
  }

  // frequency folding of generateAccessBuffer / getValue: both s and u
  // negative flips both, exactly one negative flips it and exchanges the sites;
  // t enters with its modulus
  static bool fold(T &s, T &t, T &u) {
    bool exch = false;
    if (s < 0 && u < 0) { s = -s; u = -u; }
    else if (s < 0) { s = -s; exch = true; }
    else if (u < 0) { u = -u; exch = true; }
    if (t < 0) t = -t;
    return exch;
  }

  // generateAccessBuffer(s, t, u, channel), channel 0 = S, 1 = T, 2 = U
  XyzAB<4, T> ab4(T s, T t, T u, int channel) const {
    XyzAB<4, T> ab;
    ab.exch = fold(s, t, u);
    int lo1, hi1, lo2, hi2; T b1, b2;
    int f[4][3];
    if (channel == 0) {
      const int ex = mesh.offset(s);
      mesh.interp(t, lo1, hi1, b1);
      mesh.interp(u, lo2, hi2, b2);
      int g[4][3] = {{ex, lo1, lo2}, {ex, hi1, lo2}, {ex, lo1, hi2}, {ex, hi1, hi2}};
      std::memcpy(f, g, sizeof g);
    } else if (channel == 1) {
      const int ex = mesh.offset(t);
      mesh.interp(s, lo1, hi1, b1);
      mesh.interp(u, lo2, hi2, b2);
      int g[4][3] = {{lo1, ex, lo2}, {hi1, ex, lo2}, {lo1, ex, hi2}, {hi1, ex, hi2}};
      std::memcpy(f, g, sizeof g);
    } else {
      const int ex = mesh.offset(u);
      mesh.interp(s, lo1, hi1, b1);
      mesh.interp(t, lo2, hi2, b2);
      int g[4][3] = {{lo1, lo2, ex}, {hi1, lo2, ex}, {lo1, hi2, ex}, {hi1, hi2, ex}};
      std::memcpy(f, g, sizeof g);
    }
    ab.w[0] = (1 - b2) * (1 - b1);
    ab.w[1] = (1 - b2) * b1;
    ab.w[2] = b2 * (1 - b1);
    ab.w[3] = b2 * b1;
    for (int i = 0; i < 4; ++i) ab.off[i] = support(f[i][0], f[i][1], f[i][2], ab.sgn[i]);
    return ab;
  }

  // generateAccessBuffer(s, t, u): trilinear, 8 supports
  XyzAB<8, T> ab8(T s, T t, T u) const {
    XyzAB<8, T> ab;
    ab.exch = fold(s, t, u);
    int lS, uS, lT, uT, lU, uU; T bS, bT, bU;
    mesh.interp(s, lS, uS, bS);
    mesh.interp(t, lT, uT, bT);
    mesh.interp(u, lU, uU, bU);
    const int f[8][3] = {{lS, lT, lU}, {uS, lT, lU}, {lS, uT, lU}, {uS, uT, lU},
                         {lS, lT, uU}, {uS, lT, uU}, {lS, uT, uU}, {uS, uT, uU}};
    ab.w[0] = (1 - bT) * (1 - bS) * (1 - bU);
    ab.w[1] = (1 - bT) * bS * (1 - bU);
    ab.w[2] = bT * (1 - bS) * (1 - bU);
    ab.w[3] = bT * bS * (1 - bU);
    ab.w[4] = (1 - bT) * (1 - bS) * bU;
    ab.w[5] = (1 - bT) * bS * bU;
    ab.w[6] = bT * (1 - bS) * bU;
    ab.w[7] = bT * bS * bU;
    for (int i = 0; i < 8; ++i) ab.off[i] = support(f[i][0], f[i][1], f[i][2], ab.sgn[i]);
    return ab;
  }

  // getValueSuperbundle: spin components of every representative site through
  // the site's spin permutation (inverted sites on an exchange); the density
  // carries the sign of the exchanged storage
  template <int N> void superbundle(const XyzAB<N, T> &ab, Bundle4<T> &out) const {
    out.reset();
    const int *rid = ab.exch ? p.inv_rid : p.site_rid;
    const int *perm = ab.exch ? p.inv_perm : p.site_perm;
    for (int i = 0; i < N; ++i) {
      const T w = ab.w[i];
      const T sw = T(ab.sgn[i]) * ab.w[i];
      const int o = ab.off[i];
      for (int j = 0; j < L; ++j) {
        out.b(X)[j] += w * plane(perm[3 * j + 0])[o + rid[j]];
        out.b(Y)[j] += w * plane(perm[3 * j + 1])[o + rid[j]];
        out.b(Z)[j] += w * plane(perm[3 * j + 2])[o + rid[j]];
        out.b(D)[j] += sw * plane(D)[o + rid[j]];
      }
    }
  }

  // getValueLocal: site 0, component as requested
  template <int N> T local(int c, const XyzAB<N, T> &ab) const {
    T v = T(0);
    if (c != D) {
      for (int i = 0; i < N; ++i) v += ab.w[i] * plane(c)[ab.off[i]];
    } else {
      for (int i = 0; i < N; ++i) v += T(ab.sgn[i]) * ab.w[i] * plane(D)[ab.off[i]];
    }
    return v;
  }

  // _directAccessMapFrequencyExchange
  T direct(int rid, int so, int to, int uo, int c) const {
    if (so >= uo) return plane(c)[((long)(so * (so + 1) / 2 + uo) * nw + to) * L + rid];
    const T v = plane(c)[((long)(uo * (uo + 1) / 2 + so) * nw + to) * L + rid];
    return c == D ? -v : v;
  }

  // getValue(i1, i2, s, t, u, c, FrequencyChannel::None) for pairs whose
  // transformation leaves the component unchanged (density, or the pair (0, 0));
  // rid_fwd / rid_inv = symmetryTransform(i1, i2) / (i2, i1)
  T value_none(int rid_fwd, int rid_inv, T s, T t, T u, int c) const {
    const bool exch = fold(s, t, u);
    const int rid = exch ? rid_inv : rid_fwd;
    int lS, uS, lT, uT, lU, uU; T bS, bT, bU;
    mesh.interp(s, lS, uS, bS);
    mesh.interp(t, lT, uT, bT);
    mesh.interp(u, lU, uU, bU);
    auto A = [&](int a, int b, int cc) { return direct(rid, a, b, cc, c); };
    return (1 - bU) * ((1 - bT) * ((1 - bS) * A(lS, lT, lU) + bS * A(uS, lT, lU)) +
                       bT * ((1 - bS) * A(lS, uT, lU) + bS * A(uS, uT, lU))) +
           bU * ((1 - bT) * ((1 - bS) * A(lS, lT, uU) + bS * A(uS, lT, uU)) +
                 bT * ((1 - bS) * A(lS, uT, uU) + bS * A(uS, uT, uU)));
  }
};

template <class T> struct XyzScratch {
  Bundle4<T> b1, b2, rpa, cur, sb[4];
  explicit XyzScratch(int L) : b1(L), b2(L), rpa(L), cur(L), sb{Bundle4<T>(L), Bundle4<T>(L), Bundle4<T>(L), Bundle4<T>(L)} {}
};

// r[j] += sg * a[j] * b[j] over the sites
template <class T> inline void prod_acc(T *r, const T *a, const T *b, int sg, int L) {
  if (sg > 0) for (int j = 0; j < L; ++j) r[j] += a[j] * b[j];
  else for (int j = 0; j < L; ++j) r[j] -= a[j] * b[j];
}

// the two spin components partnering component c in the particle-particle and
// particle-hole ladders, in the reference's term order
constexpr int kOther[3][2] = {{Z, Y}, {Z, X}, {X, Y}};

// ladder channel of the pairs (sb0, sb1) and (sb2, sb3): spin output c gets
// A_d B_c + A_c B_d - A_c1 B_c2 - A_c2 B_c1 (S) or all four subtracted (U),
// the density output sum_k A_k B_k (S) or its negative (U)
template <class T> void ladder(const Bundle4<T> *sb, bool u_channel, Bundle4<T> &ret) {
  const int L = ret.L;
  const int sp = u_channel ? -1 : 1, sm = -1;
  for (int c = 0; c < 3; ++c) {
    T *r = ret.b(c);
    const int c1 = kOther[c][0], c2 = kOther[c][1];
    for (int q = 0; q < 2; ++q) {
      const Bundle4<T> &A = sb[2 * q], &B = sb[2 * q + 1];
      prod_acc(r, A.b(D), B.b(c), sp, L);
      prod_acc(r, A.b(c), B.b(D), sp, L);
      prod_acc(r, A.b(c1), B.b(c2), sm, L);
      prod_acc(r, A.b(c2), B.b(c1), sm, L);
    }
  }
  T *r = ret.b(D);
  for (int q = 0; q < 2; ++q) {
    const Bundle4<T> &A = sb[2 * q], &B = sb[2 * q + 1];
    prod_acc(r, A.b(D), B.b(D), sp, L);
    prod_acc(r, A.b(X), B.b(X), sp, L);
    prod_acc(r, A.b(Y), B.b(Y), sp, L);
    prod_acc(r, A.b(Z), B.b(Z), sp, L);
  }
}

// r[j] += sg * x[j] * a over the sites
template <class T> inline void scal_acc(T *r, const T *x, T a, int sg, int L) {
  if (sg > 0) for (int j = 0; j < L; ++j) r[j] += x[j] * a;
  else for (int j = 0; j < L; ++j) r[j] -= x[j] * a;
}

// chalice partners of spin output c (added) in the reference's order, and the
// inverse chalice partners
constexpr int kChal[3][2] = {{Y, Z}, {X, Z}, {X, Y}};
constexpr int kIChal[3][2] = {{Z, Y}, {Z, X}, {X, Y}};

// XYZFrgCore::_calculateVertexTwoParticle (src/XYZ/XYZFrgCore.cpp:195-571)
template <class T>
void xyz_vertex_two_particle(const XyzCtx<T> &c, int it, XyzScratch<T> &sc, T cutoff, T *out, long *calls) {
  const int L = c.L, nw = c.nw;
  const Mesh<T> &m = c.mesh;
  int su = it / nw;
  const T t = m.at(it % nw);
  int so = 0;
  while ((so + 1) * (so + 2) / 2 <= su) ++so;
  const int uo = su - so * (so + 1) / 2;
  const T s = m.at(so), u = m.at(uo);

  const T w1p = T(0.5) * (s + t + u);
  const T w1 = T(0.5) * (s - t + u);
  const T w2p = T(0.5) * (s - t - u);
  const T w2 = T(0.5) * (s + t - u);

  auto kS = [&](T wp, Bundle4<T> &ret) {
    c.superbundle(c.ab4(s, -w1 - wp, -w2 - wp, 0), sc.sb[0]);
    c.superbundle(c.ab4(s, w1p + wp, -w2p - wp, 0), sc.sb[1]);
    c.superbundle(c.ab4(s, w2 + wp, w1 + wp, 0), sc.sb[2]);
    c.superbundle(c.ab4(s, -w2p - wp, w1p + wp, 0), sc.sb[3]);
    ret.reset();
    ladder(sc.sb, false, ret);
  };

  auto kT = [&](T wp, Bundle4<T> &ret) {
    c.superbundle(c.ab4(w1 - wp, t, w1p + wp, 1), sc.sb[0]);
    c.superbundle(c.ab4(w2p - wp, t, -w2 - wp, 1), sc.sb[1]);
    c.superbundle(c.ab4(w1p + wp, t, w1 - wp, 1), sc.sb[2]);
    c.superbundle(c.ab4(w2 + wp, t, wp - w2p, 1), sc.sb[3]);
    ret.reset();
    sc.rpa.reset();
    // lattice bubble of the first pair: entry i of getOverlap(rid) pairs site
    // rid1 of bundle 0 and rid2 of bundle 1, each spin component through the
    // entry's transformation
    for (int rid = 0; rid < L; ++rid) {
      const int b = c.p.ovl_off[rid], e = c.p.ovl_off[rid + 1];
      for (int k = 0; k < 3; ++k)
        for (int i = b; i < e; ++i)
          sc.rpa.b(k)[rid] += sc.sb[0].b(c.p.ovl_perm1[3 * i + k])[c.p.ovl_rid1[i]] *
                              sc.sb[1].b(c.p.ovl_perm2[3 * i + k])[c.p.ovl_rid2[i]];
      for (int i = b; i < e; ++i)
        sc.rpa.b(D)[rid] += sc.sb[0].b(D)[c.p.ovl_rid1[i]] * sc.sb[1].b(D)[c.p.ovl_rid2[i]];
    }
    ret.mult_add(T(4.0), sc.rpa);
    // site-local vertices: chalice (Cb, Cb2) and inverse chalice (ICa, ICa2)
    const auto a4 = c.ab4(w2p - wp, -w2 - wp, t, 2);
    const auto a5 = c.ab4(w1 - wp, -w1p - wp, -t, 2);
    const auto a6 = c.ab4(w2 + wp, wp - w2p, t, 2);
    const auto a7 = c.ab4(w1p + wp, wp - w1, -t, 2);
    T cb[2][4], ica[2][4];
    for (int k = 0; k < 4; ++k) {
      cb[0][k] = c.local(k, a4);
      ica[0][k] = c.local(k, a5);
      cb[1][k] = c.local(k, a6);
      ica[1][k] = c.local(k, a7);
    }
    for (int k = 0; k < 3; ++k) {
      T *r = ret.b(k);
      for (int q = 0; q < 2; ++q) {
        const T *A = sc.sb[2 * q].b(k), *B = sc.sb[2 * q + 1].b(k);
        scal_acc(r, A, cb[q][D], -1, L);
        scal_acc(r, A, cb[q][kChal[k][0]], 1, L);
        scal_acc(r, A, cb[q][kChal[k][1]], 1, L);
        scal_acc(r, A, cb[q][k], -1, L);
        scal_acc(r, B, ica[q][D], -1, L);
        scal_acc(r, B, ica[q][kIChal[k][0]], 1, L);
        scal_acc(r, B, ica[q][kIChal[k][1]], 1, L);
        scal_acc(r, B, ica[q][k], -1, L);
      }
    }
    T *r = ret.b(D);
    static const int dord[4] = {D, X, Y, Z};
    for (int q = 0; q < 2; ++q) {
      const T *A = sc.sb[2 * q].b(D), *B = sc.sb[2 * q + 1].b(D);
      for (int k = 0; k < 4; ++k) scal_acc(r, A, cb[q][dord[k]], -1, L);
      for (int k = 0; k < 4; ++k) scal_acc(r, B, ica[q][dord[k]], -1, L);
    }
  };

  auto kU = [&](T wp, Bundle4<T> &ret) {
    c.superbundle(c.ab4(w1 + wp, wp - w2p, u, 2), sc.sb[0]);
    c.superbundle(c.ab4(w1p + wp, w2 - wp, u, 2), sc.sb[1]);
    c.superbundle(c.ab4(w2p - wp, -w1 - wp, u, 2), sc.sb[2]);
    c.superbundle(c.ab4(w2 - wp, w1p + wp, u, 2), sc.sb[3]);
    ret.reset();
    ladder(sc.sb, true, ret);
  };

  auto p = [&](T a, T b) { return T(1) / ((a + c.sigma(c.v2, a)) * (b + c.sigma(c.v2, b))); };
  auto pK = [&](T a, T b) {
    T d1 = a + c.sigma(c.v2, a);
    return c.sigma(c.dv2, a) / (d1 * d1 * (b + c.sigma(c.v2, b)));
  };

  Bundle4<T> &cur = sc.cur;
  cur.reset();
  long nS = 0, nT = 0, nU = 0;
  // conventional contribution (XYZFrgCore.cpp:488-509)
  kS(cutoff, sc.b1); ++nS; cur.mult_add(p(cutoff, cutoff + s), sc.b1);
  if (s > T(2.0) * cutoff) { kS(-cutoff, sc.b1); ++nS; cur.mult_add(p(cutoff, cutoff - s), sc.b1); }
  kT(cutoff, sc.b1); ++nT; cur.mult_add(p(cutoff, cutoff + t), sc.b1);
  if (t > T(2.0) * cutoff) { kT(-cutoff, sc.b1); ++nT; cur.mult_add(p(cutoff, cutoff - t), sc.b1); }
  kU(cutoff, sc.b1); ++nU; cur.mult_add(p(cutoff, cutoff + u), sc.b1);
  if (u > T(2.0) * cutoff) { kU(-cutoff, sc.b1); ++nU; cur.mult_add(p(cutoff, cutoff - u), sc.b1); }

  // Katanin contribution (XYZFrgCore.cpp:511-562)
  auto fS = [&](T wp, Bundle4<T> &r) { kS(wp, r); r.scale(pK(wp, s + wp)); };
  auto fT = [&](T wp, Bundle4<T> &r) { kT(wp, r); r.scale(pK(wp, t + wp)); };
  auto fU = [&](T wp, Bundle4<T> &r) { kU(wp, r); r.scale(pK(wp, u + wp)); };
  frg::katanin_windows(m, s, cutoff, fS, sc.b1, sc.b2, cur, nS);
  frg::katanin_windows(m, t, cutoff, fT, sc.b1, sc.b2, cur, nT);
  frg::katanin_windows(m, u, cutoff, fU, sc.b1, sc.b2, cur, nU);

  cur.div(frg::two_pi<T>());
  for (int k = 0; k < 4; ++k) std::memcpy(out + k * c.nv + (long)it * L, cur.b(k), sizeof(T) * L);
  if (calls) { calls[0] += nS; calls[1] += nT; calls[2] += nU; }
}

// XYZFrgCore::_calculateVertexSingleParticle (src/XYZ/XYZFrgCore.cpp:164-193)
template <class T> T xyz_vertex_single_particle(const XyzCtx<T> &c, int i, T cutoff) {
  const T w = c.mesh.at(i);
  T v = 0;
  T sum = 0;
  for (int j = 0; j < c.p.R; ++j) {
    const int f = c.p.range_fwd[j], r = c.p.range_inv[j];
    sum += c.value_none(f, r, w + cutoff, T(0), w - cutoff, D);
    sum -= c.value_none(f, r, w - cutoff, T(0), w + cutoff, D);
  }
  v -= T(2.0) * sum;
  for (int k = 0; k < 4; ++k)
    v += c.value_none(0, 0, w + cutoff, w - cutoff, T(0), k) - c.value_none(0, 0, w - cutoff, w + cutoff, T(0), k);
  v /= (frg::two_pi<T>() * (cutoff + c.sigma(c.v2, cutoff)));
  return v;
}

template <class T>
void xyz_items(const XyzProblem &p, T cutoff, const T *v2, const T *dv2, const T *v4, int it0, int it1, T *dv4,
               int nthreads, long *calls) {
  XyzCtx<T> c(p, v2, dv2, v4);
  long cS = 0, cT = 0, cU = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : cS, cT, cU)
  {
    XyzScratch<T> sc(p.L);
    long local[3] = {0, 0, 0};
#pragma omp for schedule(guided)
    for (int it = it0; it < it1; ++it) xyz_vertex_two_particle(c, it, sc, cutoff, dv4, local);
    cS += local[0]; cT += local[1]; cU += local[2];
  }
  if (calls) { calls[0] = cS; calls[1] = cT; calls[2] = cU; }
}

template <class T>
void xyz_compute_step(const XyzProblem &p, T cutoff, const T *v2, const T *v4, T *dv2, T *dv4, int nthreads,
                      long *calls) {
  {
    XyzCtx<T> c(p, v2, nullptr, v4);
    for (int i = 0; i < p.nw; ++i) dv2[i] = xyz_vertex_single_particle(c, i, cutoff);
  }
  const int nf = p.nw * p.nw * (p.nw + 1) / 2;
  xyz_items<T>(p, cutoff, v2, dv2, v4, 0, nf, dv4, nthreads, calls);
}

// XYZMeasurementCorrelation::_calculateCorrelation (src/XYZ/XYZMeasurementCorrelation.cpp:98-212).
// Output: 4 groups of B*Rb values in the order XX, YY, ZZ, DD.
template <class T> void xyz_correlation(const XyzProblem &p, T cut, const T *v2, const T *v4, T *out) {
  XyzCtx<T> c(p, v2, nullptr, v4);
  const Mesh<T> &m = c.mesh;
  const int L = p.L;
  const T nu = T(0);
  Bundle4<T> susc(L), sb(L), buf1(L), buf2(L), buf3(L), buf4(L);
  long dummy = 0;
  T four_pi, pi, four_pi2;
  if (sizeof(T) == sizeof(float)) {
    four_pi = T(float(4.0f * M_PI)); pi = T(float(M_PI)); four_pi2 = T(float(4.0f * M_PI * M_PI));
  } else {
    four_pi = T(4.0 * M_PI); pi = T(M_PI); four_pi2 = T(4.0 * M_PI * M_PI);
  }

  auto outer = [&](T w, Bundle4<T> &ret) {
    ret.reset();
    const T term1 = T(1) / ((w + c.sigma(v2, w)) * (w + nu + c.sigma(v2, w + nu)));
    ret.b(X)[0] += term1 / four_pi;
    ret.b(Y)[0] += term1 / four_pi;
    ret.b(Z)[0] += term1 / four_pi;
    ret.b(D)[0] += term1 / pi;
    auto inner = [&](T wp, Bundle4<T> &r) {
      r.reset();
      c.superbundle(c.ab8(w + wp + nu, nu, w - wp), sb);
      const T vx = c.value_none(0, 0, w + wp + nu, w - wp, nu, X);
      const T vy = c.value_none(0, 0, w + wp + nu, w - wp, nu, Y);
      const T vz = c.value_none(0, 0, w + wp + nu, w - wp, nu, Z);
      const T vd = c.value_none(0, 0, w + wp + nu, w - wp, nu, D);
      // dumbbell diagram
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < L; ++j) r.b(k)[j] -= T(1.0) * sb.b(k)[j];
      for (int j = 0; j < L; ++j) r.b(D)[j] -= T(4.0) * sb.b(D)[j];
      // egg diagram (site 0)
      r.b(X)[0] += T(0.5) * (vx - vy - vz + vd);
      r.b(Y)[0] += T(0.5) * (-vx + vy - vz + vd);
      r.b(Z)[0] += T(0.5) * (-vx - vy + vz + vd);
      r.b(D)[0] += T(2.0) * (vx + vy + vz + vd);
      const T norm = T(1) / ((w + c.sigma(v2, w)) * (w + nu + c.sigma(v2, w + nu)) * (wp + c.sigma(v2, wp)) *
                             (wp + nu + c.sigma(v2, wp + nu)) * four_pi2);
      r.scale(norm);
    };
    frg::katanin_windows(m, nu, cut, inner, buf3, buf4, ret, dummy);
  };
  susc.reset();
  frg::katanin_windows(m, nu, cut, outer, buf1, buf2, susc, dummy);
  const long n = (long)p.B * p.Rb;
  for (long k = 0; k < n; ++k) {
    const int rid = p.corr_rid[k];
    for (int g = 0; g < 3; ++g) out[g * n + k] = susc.b(p.corr_perm[3 * k + g])[rid];
    out[3 * n + k] = susc.b(D)[rid];
  }
}

template <class T> bool has_nan(long n, const T *x) {
  for (long i = 0; i < n; ++i) if (std::isnan(x[i])) return true;
  return false;
}

// flow loop (src/SpinParser.cpp:141-172); 4*B*Rb values per record
template <class T>
int xyz_run_flow(const XyzProblem &p, int ncut, const float *cutoffs, T *v2, T *v4, T *rec, int nthreads) {
  const long nf = (long)p.nw * p.nw * (p.nw + 1) / 2;
  const long nv4 = 4 * nf * p.L;
  std::vector<T> dv2(p.nw), dv4(nv4);
  const long rsz = 4L * p.B * p.Rb;
  int n = 0, k = 0;
  for (; k < ncut - 1; ++k) {
    const T cut = T(cutoffs[k]);
    xyz_compute_step<T>(p, cut, v2, v4, dv2.data(), dv4.data(), nthreads, nullptr);
    xyz_correlation<T>(p, cut, v2, v4, rec + n * rsz);
    ++n;
    if (has_nan(p.nw, dv2.data()) || has_nan(nv4, dv4.data())) return -n;
    // XYZFrgCore::finalizeStep (src/XYZ/XYZFrgCore.cpp:128-162)
    const T step = T(cutoffs[k + 1]) - cut;
    for (int i = 0; i < p.nw; ++i) v2[i] += step * dv2[i];
    for (long i = 0; i < nv4; ++i) v4[i] += step * dv4[i];
  }
  xyz_correlation<T>(p, T(cutoffs[k]), v2, v4, rec + n * rsz);
  ++n;
  return n;
}

}  // namespace

extern "C" {

typedef XyzProblem or_xyz_problem;

void or_xyz_compute_step_f64(const or_xyz_problem *p, double cutoff, const double *v2, const double *v4, double *dv2,
                             double *dv4, int nthreads, long *calls) {
  xyz_compute_step<double>(*p, cutoff, v2, v4, dv2, dv4, nthreads, calls);
}

void or_xyz_compute_step_f32(const or_xyz_problem *p, float cutoff, const float *v2, const float *v4, float *dv2,
                             float *dv4, int nthreads, long *calls) {
  xyz_compute_step<float>(*p, cutoff, v2, v4, dv2, dv4, nthreads, calls);
}

void or_xyz_items_f64(const or_xyz_problem *p, double cutoff, const double *v2, const double *dv2, const double *v4,
                      int it0, int it1, double *dv4, int nthreads, long *calls) {
  xyz_items<double>(*p, cutoff, v2, dv2, v4, it0, it1, dv4, nthreads, calls);
}

void or_xyz_items_f32(const or_xyz_problem *p, float cutoff, const float *v2, const float *dv2, const float *v4,
                      int it0, int it1, float *dv4, int nthreads, long *calls) {
  xyz_items<float>(*p, cutoff, v2, dv2, v4, it0, it1, dv4, nthreads, calls);
}

void or_xyz_correlation_f64(const or_xyz_problem *p, double cutoff, const double *v2, const double *v4, double *out) {
  xyz_correlation<double>(*p, cutoff, v2, v4, out);
}

void or_xyz_correlation_f32(const or_xyz_problem *p, float cutoff, const float *v2, const float *v4, float *out) {
  xyz_correlation<float>(*p, cutoff, v2, v4, out);
}

int or_xyz_run_flow_f32(const or_xyz_problem *p, int ncut, const float *cutoffs, float *v2, float *v4, float *rec,
                        int nthreads) {
  return xyz_run_flow<float>(*p, ncut, cutoffs, v2, v4, rec, nthreads);
}

int or_xyz_run_flow_f64(const or_xyz_problem *p, int ncut, const float *cutoffs, double *v2, double *v4, double *rec,
                        int nthreads) {
  return xyz_run_flow<double>(*p, ncut, cutoffs, v2, v4, rec, nthreads);
}

}  // extern "C"
