// CPU restatement of the reference TRI pf-FRG right-hand side (general
// time-reversal-invariant two-spin couplings, 16 spin components) -- TEST
// INFRASTRUCTURE ONLY.  It is the checker the HIP TRI path is compared against;
// nothing in the product links or calls it.
//
// Templated on T: float follows the reference arithmetic (all-float state,
// -ffp-contract=off); double is the FP64 oracle for the GPU kernels.
//
// Reference files followed (paths relative to the reference repository):
//   src/TRI/TRIVertexTwoParticle.hpp:62-76,190-666  storage, access buffers, signs zeta
//   src/TRI/TRIVertexSingleParticle.hpp:73-87       self-energy interpolation
//   src/TRI/TRIFrgCore.cpp:98-151                   finalizeStep, dSigma
//   src/TRI/TRIFrgCore.cpp:153-3001                 dGamma: S, T (lattice bubble, chalice,
//                                                   inverse chalice), U, conventional, Katanin
//   src/TRI/TRIMeasurementCorrelation.cpp:147-316   static susceptibility, 10 outputs
//
// The channel algebra is not transcribed: every coefficient table is computed
// here from Pauli-matrix products (spin structure Gamma = sum Gamma^{mu nu}
// sigma^mu (x) sigma^nu, mixed density-spin components carrying a factor i).
// The frequency arguments of the gathered vertices, the access-buffer sign
// rules and the integration windows follow the reference line by line.
//
// Parity status: pinned end to end by the reference's TRI golden flow
// (Heisenberg-Kitaev-Gamma honeycomb, all 10 correlation outputs, T = float),
// by TRI(Heisenberg) = SU2 on the same lattice, and by the integrator known
// answers shared with the SU(2) oracle.  The antisymmetric (DM) couplings are
// exercised by no reference fixture.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "frg_mesh.hpp"

namespace {

using frg::Mesh;

struct TriProblem {
  int nw;
  const float *mesh;
  int L;
  const int *site_rid;   // getSites()[j].rid, j < L
  const int *site_perm;  // L x 3: getSites()[j].spinPermutation
  const int *inv_rid;    // getInvertedSites()[j].rid
  const int *inv_perm;   // L x 3
  const int *ovl_off;    // L+1
  const int *ovl_rid1;
  const int *ovl_rid2;
  const int *ovl_perm1;  // O x 3: transformedX1, Y1, Z1
  const int *ovl_perm2;  // O x 3: transformedX2, Y2, Z2
  int R;
  const int *range_fwd;  // symmetryTransform(0, j)
  const int *range_inv;  // symmetryTransform(j, 0)
  int B;
  int Rb;
  const int *corr_rid;   // B*Rb
  const int *corr_perm;  // B*Rb x 3
};

// ---------------------------------------------------------------------------
// spin algebra: coefficient tables from Pauli products
// ---------------------------------------------------------------------------
using cplx = std::complex<double>;
struct M2 {
  cplx m[2][2];
};

M2 pauli(int k) {  // component order x, y, z, d (d = identity)
  M2 r{};
  const cplx i(0.0, 1.0);
  if (k == 0) { r.m[0][1] = 1.0; r.m[1][0] = 1.0; }
  else if (k == 1) { r.m[0][1] = -i; r.m[1][0] = i; }
  else if (k == 2) { r.m[0][0] = 1.0; r.m[1][1] = -1.0; }
  else { r.m[0][0] = 1.0; r.m[1][1] = 1.0; }
  return r;
}

M2 mul(const M2 &a, const M2 &b) {
  M2 r{};
  for (int p = 0; p < 2; ++p)
    for (int q = 0; q < 2; ++q) r.m[p][q] = a.m[p][0] * b.m[0][q] + a.m[p][1] * b.m[1][q];
  return r;
}

cplx tr(const M2 &a) { return a.m[0][0] + a.m[1][1]; }
// coefficient of sigma^r in a: tr(sigma^r a) / 2
cplx proj(int r, const M2 &a) { return tr(mul(pauli(r), a)) * 0.5; }
M2 prod(std::initializer_list<int> ks) {
  M2 r = pauli(3);
  for (int k : ks) r = mul(r, pauli(k));
  return r;
}
// phase convention of the stored components: i for one density and one spin index
cplx phase(int c) {
  const int a = c / 4, b = c % 4;
  return ((a == 3) != (b == 3)) ? cplx(0.0, 1.0) : cplx(1.0, 0.0);
}

double real_coefficient(cplx v) {
  if (std::abs(v.imag()) > 1e-12) throw 1;  // a non-real coefficient means a broken table
  return v.real();
}

struct Term {
  int o, a, b;
  double c;
};

struct Tables {
  std::vector<Term> S, U, CH, ICH;
  std::vector<Term> RPA;   // a = rho*4 + nu, b = nu*4 + tau (before the site transforms)
  double egg[16][16];      // measurement: local vertex -> outputs
  Tables() {
    for (int o = 0; o < 16; ++o) {
      const int r = o / 4, t = o % 4;
      for (int a = 0; a < 16; ++a) {
        const int mu = a / 4, nu = a % 4;
        for (int b = 0; b < 16; ++b) {
          const int ka = b / 4, la = b % 4;
          const cplx g = phase(a) * phase(b) / phase(o);
          // particle-particle: both lines sigma^(2nd vertex) sigma^(1st vertex)
          double c = real_coefficient(proj(r, prod({ka, mu})) * proj(t, prod({la, nu})) * g);
          if (c != 0.0) S.push_back({o, a, b, c});
          // crossed particle-hole: second line reversed, overall minus
          c = real_coefficient(-proj(r, prod({ka, mu})) * proj(t, prod({nu, la})) * g);
          if (c != 0.0) U.push_back({o, a, b, c});
          // chalice: line 1 carries the bundle's first index, line 2 the local
          // vertex wrapped around the bundle's second index
          c = real_coefficient(-proj(r, prod({mu})) * proj(t, prod({la, nu, ka})) * g);
          if (c != 0.0) CH.push_back({o, a, b, c});
          // inverse chalice (a = local vertex, b = bundle)
          c = real_coefficient(-proj(r, prod({nu, ka, mu})) * proj(t, prod({la})) * g);
          if (c != 0.0) ICH.push_back({o, a, b, c});
        }
      }
      // lattice bubble: closed loop, trace over the internal index nu
      for (int nu = 0; nu < 4; ++nu) {
        const double c = real_coefficient(2.0 * phase(4 * r + nu) * phase(4 * nu + t) / phase(o));
        RPA.push_back({o, 4 * r + nu, 4 * nu + t, c});
      }
      // egg diagram of the measurement
      const double n = (o == 15) ? 2.0 : 0.5;
      for (int cl = 0; cl < 16; ++cl) {
        const int ka = cl / 4, la = cl % 4;
        egg[o][cl] = real_coefficient(n * tr(prod({r, la, t, ka})) * 0.5 * phase(cl) / phase(o));
      }
    }
  }
};

const Tables &tables() {
  static const Tables t;
  return t;
}

inline int zeta(int s) { return s <= 2 ? -1 : 1; }

// ---------------------------------------------------------------------------
// vertex access
// ---------------------------------------------------------------------------
template <int N, class T> struct TriAB {
  int off[N];
  T w[N];
  bool fx[N];      // support stored with s and u exchanged
  bool exch = false;
  bool tu = false;  // odd number of sign flips of t and u
  // sign[i][s1][s2] of the reference's access buffer
  T sign(int i, int s1, int s2) const {
    T v = T(1);
    if (tu) v *= T(zeta(s1) * zeta(s2));
    if (fx[i]) v *= T(-zeta(exch ? s1 : s2));
    return v;
  }
};

template <class T> struct Bundle16 {
  int L;
  std::vector<T> d;
  explicit Bundle16(int L_) : L(L_), d(16 * L_, T(0)) {}
  T *b(int c) { return d.data() + c * L; }
  const T *b(int c) const { return d.data() + c * L; }
  void reset() { std::fill(d.begin(), d.end(), T(0)); }
  void mult_add(T a, const Bundle16 &x) { for (size_t i = 0; i < d.size(); ++i) d[i] += a * x.d[i]; }
  void add(const Bundle16 &x) { for (size_t i = 0; i < d.size(); ++i) d[i] += x.d[i]; }
  void scale(T a) { for (size_t i = 0; i < d.size(); ++i) d[i] *= a; }
  void div(T a) { for (size_t i = 0; i < d.size(); ++i) d[i] /= a; }
};

template <class T> struct TriCtx {
  const TriProblem &p;
  Mesh<T> mesh;
  int L, nw;
  long FS;  // frequency stride 16 L
  const T *v2, *dv2, *v4;
  TriCtx(const TriProblem &p_, const T *v2_, const T *dv2_, const T *v4_)
      : p(p_), mesh(p_.nw, p_.mesh), L(p_.L), nw(p_.nw), FS(16L * p_.L), v2(v2_), dv2(dv2_), v4(v4_) {}

  T sigma(const T *data, T w) const {
    T sign = T(1);
    if (w < 0) { w = -w; sign = T(-1); }
    int lo, hi; T bias;
    mesh.interp(w, lo, hi, bias);
    return sign * ((1 - bias) * data[lo] + bias * data[hi]);
  }

  // _generateAccessBufferOffsetMapFrequencyExchange
  int support(int so, int to, int uo, bool &fx) const {
    fx = so < uo;
    if (fx) std::swap(so, uo);
    return (int)(((long)(so * (so + 1) / 2 + uo) * nw + to) * FS);
  }

  // generateAccessBuffer(s, t, u, channel), channel 0 = S, 1 = T, 2 = U
  TriAB<4, T> ab4(T s, T t, T u, int channel) const {
    TriAB<4, T> ab;
    if (s < 0) { s = -s; ab.exch = !ab.exch; }
    if (t < 0) { t = -t; ab.tu = !ab.tu; }
    if (u < 0) { u = -u; ab.exch = !ab.exch; ab.tu = !ab.tu; }
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
    for (int i = 0; i < 4; ++i) ab.off[i] = support(f[i][0], f[i][1], f[i][2], ab.fx[i]);
    return ab;
  }

  // generateAccessBuffer(s, t, u): trilinear, 8 supports
  TriAB<8, T> ab8(T s, T t, T u) const {
    TriAB<8, T> ab;
    if (s < 0) { s = -s; ab.exch = !ab.exch; }
    if (t < 0) { t = -t; ab.tu = !ab.tu; }
    if (u < 0) { u = -u; ab.exch = !ab.exch; ab.tu = !ab.tu; }
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
    for (int i = 0; i < 8; ++i) ab.off[i] = support(f[i][0], f[i][1], f[i][2], ab.fx[i]);
    return ab;
  }

  // getValueSuperbundle: all 16 components on all representative sites; the
  // stored components are the requested ones, swapped on a site exchange and
  // mapped through the site's spin permutation
  template <int N> void superbundle(const TriAB<N, T> &ab, Bundle16<T> &out) const {
    out.reset();
    const int *rid = ab.exch ? p.inv_rid : p.site_rid;
    const int *perm = ab.exch ? p.inv_perm : p.site_perm;
    for (int i = 0; i < N; ++i) {
      for (int s1 = 0; s1 < 4; ++s1) {
        for (int s2 = 0; s2 < 4; ++s2) {
          const T f = ab.sign(i, s1, s2) * ab.w[i];
          T *o = out.b(4 * s1 + s2);
          for (int j = 0; j < L; ++j) {
            int a = ab.exch ? s2 : s1, b = ab.exch ? s1 : s2;
            if (a < 3) a = perm[3 * j + a];
            if (b < 3) b = perm[3 * j + b];
            o[j] += f * v4[ab.off[i] + (4 * a + b) * L + rid[j]];
          }
        }
      }
    }
  }

  // getValueLocal(s1, s2): the site-0 value; the reference indexes the sign
  // table with the exchanged components here (TRIVertexTwoParticle.hpp:300-309)
  template <int N> void local16(const TriAB<N, T> &ab, T out[16]) const {
    for (int c = 0; c < 16; ++c) {
      int s1 = c / 4, s2 = c % 4;
      if (ab.exch) std::swap(s1, s2);
      T v = T(0);
      for (int i = 0; i < N; ++i) v += ab.sign(i, s1, s2) * ab.w[i] * v4[ab.off[i] + (4 * s1 + s2) * L];
      out[c] = v;
    }
  }

  // _directAccessMapFrequencyExchange with stored components (s1, s2)
  T direct(int rid, int so, int to, int uo, int s1, int s2) const {
    if (so < uo) {
      return T(-zeta(s2)) * v4[((long)(uo * (uo + 1) / 2 + so) * nw + to) * FS + (4 * s1 + s2) * L + rid];
    }
    return v4[((long)(so * (so + 1) / 2 + uo) * nw + to) * FS + (4 * s1 + s2) * L + rid];
  }

  // getValue(i1, i2, s, t, u, s1, s2, FrequencyChannel::None) for pairs whose
  // transformation leaves the requested components unchanged (density-density,
  // or the site pair (0, 0)); rid_fwd / rid_inv = symmetryTransform(i1, i2) / (i2, i1)
  T value_none(int rid_fwd, int rid_inv, T s, T t, T u, int s1, int s2) const {
    T sign = T(1);
    bool swap = false;
    if (s < 0) { s = -s; std::swap(s1, s2); swap = !swap; }
    if (t < 0) { t = -t; sign *= T(zeta(s1) * zeta(s2)); }
    if (u < 0) { u = -u; std::swap(s1, s2); swap = !swap; sign *= T(zeta(s1) * zeta(s2)); }
    const int rid = swap ? rid_inv : rid_fwd;
    int lS, uS, lT, uT, lU, uU; T bS, bT, bU;
    mesh.interp(s, lS, uS, bS);
    mesh.interp(t, lT, uT, bT);
    mesh.interp(u, lU, uU, bU);
    auto A = [&](int a, int b, int c) { return direct(rid, a, b, c, s1, s2); };
    return sign * ((1 - bU) * ((1 - bT) * ((1 - bS) * A(lS, lT, lU) + bS * A(uS, lT, lU)) +
                               bT * ((1 - bS) * A(lS, uT, lU) + bS * A(uS, uT, lU))) +
                   bU * ((1 - bT) * ((1 - bS) * A(lS, lT, uU) + bS * A(uS, lT, uU)) +
                         bT * ((1 - bS) * A(lS, uT, uU) + bS * A(uS, uT, uU))));
  }
};

template <class T> struct TriScratch {
  Bundle16<T> b1, b2, rpa, cur, sb0, sb1, sb2, sb3;
  explicit TriScratch(int L) : b1(L), b2(L), rpa(L), cur(L), sb0(L), sb1(L), sb2(L), sb3(L) {}
};

// pair products of two bundles through a coefficient table, both pairs
// (sb0, sb1) and (sb2, sb3)
template <class T>
void bundle_pairs(const std::vector<Term> &tab, const Bundle16<T> &x0, const Bundle16<T> &y0, const Bundle16<T> &x1,
                  const Bundle16<T> &y1, Bundle16<T> &ret) {
  const int L = ret.L;
  for (const Term &e : tab) {
    const T c = T(e.c);
    T *r = ret.b(e.o);
    const T *a0 = x0.b(e.a), *b0 = y0.b(e.b), *a1 = x1.b(e.a), *b1 = y1.b(e.b);
    for (int j = 0; j < L; ++j) r[j] += c * (a0[j] * b0[j]);
    for (int j = 0; j < L; ++j) r[j] += c * (a1[j] * b1[j]);
  }
}

// TRIFrgCore::_calculateVertexTwoParticle (src/TRI/TRIFrgCore.cpp:153-3001)
template <class T>
void tri_vertex_two_particle(const TriCtx<T> &c, int it, TriScratch<T> &sc, T cutoff, T *out, long *calls) {
  const int L = c.L, nw = c.nw;
  const Mesh<T> &m = c.mesh;
  const Tables &tb = tables();
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

  auto kS = [&](T wp, Bundle16<T> &ret) {
    c.superbundle(c.ab4(s, w2 + wp, w1 + wp, 0), sc.sb0);
    c.superbundle(c.ab4(s, -w2p - wp, w1p + wp, 0), sc.sb1);
    c.superbundle(c.ab4(s, -w1 - wp, -w2 - wp, 0), sc.sb2);
    c.superbundle(c.ab4(s, w1p + wp, -w2p - wp, 0), sc.sb3);
    ret.reset();
    bundle_pairs(tb.S, sc.sb0, sc.sb1, sc.sb2, sc.sb3, ret);
  };

  auto kT = [&](T wp, Bundle16<T> &ret) {
    c.superbundle(c.ab4(w1 - wp, t, w1p + wp, 1), sc.sb0);
    c.superbundle(c.ab4(w2p - wp, t, -w2 - wp, 1), sc.sb1);
    c.superbundle(c.ab4(w1p + wp, t, w1 - wp, 1), sc.sb2);
    c.superbundle(c.ab4(w2 + wp, t, -w2p + wp, 1), sc.sb3);
    ret.reset();
    sc.rpa.reset();
    // lattice bubble (TRIFrgCore.cpp:733-1254): the bundles' spin indices pass
    // through the overlap's spin transformations
    for (int rid = 0; rid < L; ++rid) {
      const int b = c.p.ovl_off[rid], e = c.p.ovl_off[rid + 1];
      for (int i = b; i < e; ++i) {
        const int *p1 = c.p.ovl_perm1 + 3 * i, *p2 = c.p.ovl_perm2 + 3 * i;
        auto t1 = [&](int k) { return k < 3 ? p1[k] : 3; };
        auto t2 = [&](int k) { return k < 3 ? p2[k] : 3; };
        const int r1 = c.p.ovl_rid1[i], r2 = c.p.ovl_rid2[i];
        for (const Term &q : tb.RPA) {
          const int rho = q.a / 4, nu = q.a % 4, tau = q.b % 4;
          const int ca = 4 * t1(rho) + t1(nu), cb = 4 * t2(nu) + t2(tau);
          T *r = sc.rpa.b(q.o);
          r[rid] += T(q.c) * sc.sb0.b(ca)[r1] * sc.sb1.b(cb)[r2];
          r[rid] += T(q.c) * sc.sb2.b(ca)[r1] * sc.sb3.b(cb)[r2];
        }
      }
    }
    ret.add(sc.rpa);
    // chalice (TRIFrgCore.cpp:1257-1811)
    T l4[16], l5[16], l6[16], l7[16];
    c.local16(c.ab4(w2p - wp, -w2 - wp, t, 2), l4);
    c.local16(c.ab4(w2 + wp, -w2p + wp, t, 2), l5);
    for (const Term &q : tb.CH) {
      T *r = ret.b(q.o);
      const T *a0 = sc.sb0.b(q.a), *a2 = sc.sb2.b(q.a);
      const T f4 = l4[q.b], f5 = l5[q.b], cc = T(q.c);
      for (int j = 0; j < L; ++j) r[j] += cc * (a0[j] * f4);
      for (int j = 0; j < L; ++j) r[j] += cc * (a2[j] * f5);
    }
    // inverse chalice (TRIFrgCore.cpp:1814-2368)
    c.local16(c.ab4(w1 - wp, -w1p - wp, -t, 2), l6);
    c.local16(c.ab4(w1p + wp, -w1 + wp, -t, 2), l7);
    for (const Term &q : tb.ICH) {
      T *r = ret.b(q.o);
      const T *b1 = sc.sb1.b(q.b), *b3 = sc.sb3.b(q.b);
      const T f6 = l6[q.a], f7 = l7[q.a], cc = T(q.c);
      for (int j = 0; j < L; ++j) r[j] += cc * (f6 * b1[j]);
      for (int j = 0; j < L; ++j) r[j] += cc * (f7 * b3[j]);
    }
  };

  auto kU = [&](T wp, Bundle16<T> &ret) {
    c.superbundle(c.ab4(w1 + wp, -w2p + wp, u, 2), sc.sb0);
    c.superbundle(c.ab4(w1p + wp, w2 - wp, u, 2), sc.sb1);
    c.superbundle(c.ab4(w2p - wp, -w1 - wp, u, 2), sc.sb2);
    c.superbundle(c.ab4(w2 - wp, w1p + wp, u, 2), sc.sb3);
    ret.reset();
    bundle_pairs(tb.U, sc.sb0, sc.sb1, sc.sb2, sc.sb3, ret);
  };

  auto p = [&](T a, T b) { return T(1) / ((a + c.sigma(c.v2, a)) * (b + c.sigma(c.v2, b))); };
  auto pK = [&](T a, T b) {
    T d1 = a + c.sigma(c.v2, a);
    return c.sigma(c.dv2, a) / (d1 * d1 * (b + c.sigma(c.v2, b)));
  };

  Bundle16<T> &cur = sc.cur;
  cur.reset();
  long nS = 0, nT = 0, nU = 0;
  // conventional contribution (TRIFrgCore.cpp:2917-2940)
  kS(cutoff, sc.b1); ++nS; cur.mult_add(p(cutoff, cutoff + s), sc.b1);
  if (s > T(2.0) * cutoff) { kS(-cutoff, sc.b1); ++nS; cur.mult_add(p(cutoff, cutoff - s), sc.b1); }
  kT(cutoff, sc.b1); ++nT; cur.mult_add(p(cutoff, cutoff + t), sc.b1);
  if (t > T(2.0) * cutoff) { kT(-cutoff, sc.b1); ++nT; cur.mult_add(p(cutoff, cutoff - t), sc.b1); }
  kU(cutoff, sc.b1); ++nU; cur.mult_add(p(cutoff, cutoff + u), sc.b1);
  if (u > T(2.0) * cutoff) { kU(-cutoff, sc.b1); ++nU; cur.mult_add(p(cutoff, cutoff - u), sc.b1); }

  // Katanin contribution (TRIFrgCore.cpp:2943-2993)
  auto fS = [&](T wp, Bundle16<T> &r) { kS(wp, r); r.scale(pK(wp, s + wp)); };
  auto fT = [&](T wp, Bundle16<T> &r) { kT(wp, r); r.scale(pK(wp, t + wp)); };
  auto fU = [&](T wp, Bundle16<T> &r) { kU(wp, r); r.scale(pK(wp, u + wp)); };
  frg::katanin_windows(m, s, cutoff, fS, sc.b1, sc.b2, cur, nS);
  frg::katanin_windows(m, t, cutoff, fT, sc.b1, sc.b2, cur, nT);
  frg::katanin_windows(m, u, cutoff, fU, sc.b1, sc.b2, cur, nU);

  cur.div(frg::two_pi<T>());
  std::memcpy(out + (long)it * 16 * L, cur.d.data(), sizeof(T) * 16 * L);
  if (calls) { calls[0] += nS; calls[1] += nT; calls[2] += nU; }
}

// TRIFrgCore::_calculateVertexSingleParticle (src/TRI/TRIFrgCore.cpp:122-151)
template <class T> T tri_vertex_single_particle(const TriCtx<T> &c, int i, T cutoff) {
  const T w = c.mesh.at(i);
  T v = 0;
  T sum = 0;
  for (int j = 0; j < c.p.R; ++j) {
    const int f = c.p.range_fwd[j], r = c.p.range_inv[j];
    sum += c.value_none(f, r, w + cutoff, T(0), w - cutoff, 3, 3);
    sum -= c.value_none(f, r, w - cutoff, T(0), w + cutoff, 3, 3);
  }
  v -= T(2.0) * sum;
  for (int k = 0; k < 3; ++k)
    v += c.value_none(0, 0, w + cutoff, w - cutoff, T(0), k, k) - c.value_none(0, 0, w - cutoff, w + cutoff, T(0), k, k);
  v += c.value_none(0, 0, w + cutoff, w - cutoff, T(0), 3, 3) - c.value_none(0, 0, w - cutoff, w + cutoff, T(0), 3, 3);
  v /= (frg::two_pi<T>() * (cutoff + c.sigma(c.v2, cutoff)));
  return v;
}

template <class T>
void tri_compute_step(const TriProblem &p, T cutoff, const T *v2, const T *v4, T *dv2, T *dv4, int nthreads,
                      long *calls) {
  {
    TriCtx<T> c(p, v2, nullptr, v4);
    for (int i = 0; i < p.nw; ++i) dv2[i] = tri_vertex_single_particle(c, i, cutoff);
  }
  TriCtx<T> c(p, v2, dv2, v4);
  const int nf = p.nw * p.nw * (p.nw + 1) / 2;
  long cS = 0, cT = 0, cU = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : cS, cT, cU)
  {
    TriScratch<T> sc(p.L);
    long local[3] = {0, 0, 0};
#pragma omp for schedule(guided)
    for (int it = 0; it < nf; ++it) tri_vertex_two_particle(c, it, sc, cutoff, dv4, local);
    cS += local[0]; cT += local[1]; cU += local[2];
  }
  if (calls) { calls[0] = cS; calls[1] = cT; calls[2] = cU; }
}

// TRIMeasurementCorrelation::_calculateCorrelation (src/TRI/TRIMeasurementCorrelation.cpp:147-316).
// Output: 10 groups of B*Rb values in the order DD, XX, XY, XZ, YX, YY, YZ, ZX, ZY, ZZ.
template <class T> void tri_correlation(const TriProblem &p, T cut, const T *v2, const T *v4, T *out) {
  TriCtx<T> c(p, v2, nullptr, v4);
  const Mesh<T> &m = c.mesh;
  const int L = p.L;
  const Tables &tb = tables();
  const T nu = T(0);
  Bundle16<T> susc(L), sb(L), buf1(L), buf2(L), buf3(L), buf4(L);
  long dummy = 0;
  T two_pi_f, four_pi2;
  if (sizeof(T) == sizeof(float)) {
    two_pi_f = T(float(2.0f * M_PI)); four_pi2 = T(float(4.0f * M_PI * M_PI));
  } else {
    two_pi_f = T(2.0 * M_PI); four_pi2 = T(4.0 * M_PI * M_PI);
  }
  static const int outs[10] = {15, 0, 1, 2, 4, 5, 6, 8, 9, 10};

  auto outer = [&](T w, Bundle16<T> &ret) {
    ret.reset();
    const T term1 = T(1) / ((w + c.sigma(v2, w)) * (w + nu + c.sigma(v2, w + nu)));
    ret.b(15)[0] += T(2.0) * term1 / two_pi_f;
    ret.b(0)[0] += T(0.5) * term1 / two_pi_f;
    ret.b(5)[0] += T(0.5) * term1 / two_pi_f;
    ret.b(10)[0] += T(0.5) * term1 / two_pi_f;
    auto inner = [&](T wp, Bundle16<T> &r) {
      r.reset();
      c.superbundle(c.ab8(w + wp + nu, nu, w - wp), sb);
      T v[16];
      c.local16(c.ab8(w + wp + nu, w - wp, nu), v);
      // dumbbell diagram: the density output carries weight 4, the spin outputs 1
      for (int k = 0; k < 10; ++k) {
        const int o = outs[k];
        const T f = o == 15 ? T(4.0) : T(1.0);
        for (int j = 0; j < L; ++j) r.b(o)[j] -= f * sb.b(o)[j];
      }
      // egg diagram (site 0)
      for (int k = 0; k < 10; ++k) {
        const int o = outs[k];
        for (int cl = 0; cl < 16; ++cl)
          if (tb.egg[o][cl] != 0.0) r.b(o)[0] += T(tb.egg[o][cl]) * v[cl];
      }
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
    const int sx = p.corr_perm[3 * k], sy = p.corr_perm[3 * k + 1], sz = p.corr_perm[3 * k + 2];
    const int comp[10] = {15, 4 * sx + sx, 4 * sx + sy, 4 * sx + sz, 4 * sy + sx,
                          4 * sy + sy, 4 * sy + sz, 4 * sz + sx, 4 * sz + sy, 4 * sz + sz};
    for (int g = 0; g < 10; ++g) out[g * n + k] = susc.b(comp[g])[rid];
  }
}

template <class T> bool has_nan(long n, const T *x) {
  for (long i = 0; i < n; ++i) if (std::isnan(x[i])) return true;
  return false;
}

// flow loop (src/SpinParser.cpp:141-172); 10*B*Rb values per record
template <class T>
int tri_run_flow(const TriProblem &p, int ncut, const float *cutoffs, T *v2, T *v4, T *rec, int nthreads) {
  const long nf = (long)p.nw * p.nw * (p.nw + 1) / 2;
  const long nv4 = nf * 16 * p.L;
  std::vector<T> dv2(p.nw), dv4(nv4);
  const long rsz = 10L * p.B * p.Rb;
  int n = 0, k = 0;
  for (; k < ncut - 1; ++k) {
    const T cut = T(cutoffs[k]);
    tri_compute_step<T>(p, cut, v2, v4, dv2.data(), dv4.data(), nthreads, nullptr);
    tri_correlation<T>(p, cut, v2, v4, rec + n * rsz);
    ++n;
    if (has_nan(p.nw, dv2.data()) || has_nan(nv4, dv4.data())) return -n;
    // TRIFrgCore::finalizeStep (src/TRI/TRIFrgCore.cpp:98-120)
    const T step = T(cutoffs[k + 1]) - cut;
    for (int i = 0; i < p.nw; ++i) v2[i] += step * dv2[i];
    for (long i = 0; i < nv4; ++i) v4[i] += step * dv4[i];
  }
  tri_correlation<T>(p, T(cutoffs[k]), v2, v4, rec + n * rsz);
  ++n;
  return n;
}

template <class T>
void tri_items(const TriProblem &p, T cutoff, const T *v2, const T *dv2, const T *v4, int it0, int it1, T *dv4,
               int nthreads, long *calls) {
  TriCtx<T> c(p, v2, dv2, v4);
  long cS = 0, cT = 0, cU = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : cS, cT, cU)
  {
    TriScratch<T> sc(p.L);
    long local[3] = {0, 0, 0};
#pragma omp for schedule(guided)
    for (int it = it0; it < it1; ++it) tri_vertex_two_particle(c, it, sc, cutoff, dv4, local);
    cS += local[0]; cT += local[1]; cU += local[2];
  }
  if (calls) { calls[0] = cS; calls[1] = cT; calls[2] = cU; }
}

}  // namespace

extern "C" {

typedef TriProblem or_tri_problem;

void or_tri_compute_step_f64(const or_tri_problem *p, double cutoff, const double *v2, const double *v4, double *dv2,
                             double *dv4, int nthreads, long *calls) {
  tri_compute_step<double>(*p, cutoff, v2, v4, dv2, dv4, nthreads, calls);
}

void or_tri_compute_step_f32(const or_tri_problem *p, float cutoff, const float *v2, const float *v4, float *dv2,
                             float *dv4, int nthreads, long *calls) {
  tri_compute_step<float>(*p, cutoff, v2, v4, dv2, dv4, nthreads, calls);
}

// dSigma/dLambda only (FP64), for checks on large problems
void or_tri_dsigma_f64(const or_tri_problem *p, double cutoff, const double *v2, const double *v4, double *dv2) {
  TriCtx<double> c(*p, v2, nullptr, v4);
  for (int i = 0; i < p->nw; ++i) dv2[i] = tri_vertex_single_particle(c, i, cutoff);
}

void or_tri_items_f64(const or_tri_problem *p, double cutoff, const double *v2, const double *dv2, const double *v4,
                      int it0, int it1, double *dv4, int nthreads, long *calls) {
  tri_items<double>(*p, cutoff, v2, dv2, v4, it0, it1, dv4, nthreads, calls);
}

void or_tri_items_f32(const or_tri_problem *p, float cutoff, const float *v2, const float *dv2, const float *v4,
                      int it0, int it1, float *dv4, int nthreads, long *calls) {
  tri_items<float>(*p, cutoff, v2, dv2, v4, it0, it1, dv4, nthreads, calls);
}

void or_tri_correlation_f64(const or_tri_problem *p, double cutoff, const double *v2, const double *v4, double *out) {
  tri_correlation<double>(*p, cutoff, v2, v4, out);
}

void or_tri_correlation_f32(const or_tri_problem *p, float cutoff, const float *v2, const float *v4, float *out) {
  tri_correlation<float>(*p, cutoff, v2, v4, out);
}

int or_tri_run_flow_f32(const or_tri_problem *p, int ncut, const float *cutoffs, float *v2, float *v4, float *rec,
                        int nthreads) {
  return tri_run_flow<float>(*p, ncut, cutoffs, v2, v4, rec, nthreads);
}

int or_tri_run_flow_f64(const or_tri_problem *p, int ncut, const float *cutoffs, double *v2, double *v4, double *rec,
                        int nthreads) {
  return tri_run_flow<double>(*p, ncut, cutoffs, v2, v4, rec, nthreads);
}

// Coefficient tables, for the tests: kind 0 S, 1 U, 2 chalice, 3 inverse
// chalice, 4 lattice bubble; out[16*16*16] (o, a, b) dense; returns #terms.
int or_tri_table(int kind, double *out) {
  const Tables &t = tables();
  const std::vector<Term> *v[5] = {&t.S, &t.U, &t.CH, &t.ICH, &t.RPA};
  if (kind < 0 || kind > 4) return -1;
  std::memset(out, 0, sizeof(double) * 4096);
  for (const Term &e : *v[kind]) out[(e.o * 16 + e.a) * 16 + e.b] += e.c;
  return (int)v[kind]->size();
}

}  // extern "C"
