// tri_flow.hip -- HIP kernels and C ABI of the TRI pf-FRG core (general
// time-reversal-invariant two-spin couplings; 16 spin components per vertex
// entry) for MI355X (gfx950), FP64.
//
// Reference behaviour followed (paths relative to the reference repository):
//   src/TRI/TRIVertexTwoParticle.hpp:62-76,190-666  storage, access buffers, signs zeta,
//                                                   site spin permutations
//   src/TRI/TRIFrgCore.cpp:98-151                   finalizeStep, dSigma
//   src/TRI/TRIFrgCore.cpp:153-3001                 dGamma: S, T (lattice bubble, chalice,
//                                                   inverse chalice), U, conventional, Katanin
//   src/TRI/TRIMeasurementCorrelation.cpp:147-316   static susceptibility, 10 outputs
//
// Device memory (per flow):
//   Gamma, dGamma  FP64 [N_f][16][L] in the reference order
//                  ((s(s+1)/2 + u) * N_w + t) * 16 L + (4 s1 + s2) L + rid
//   Sigma, dSigma  FP64 [N_w]
//   weight tables  per exact index x: evaluation points and weights at the cutoff
//
// The channel algebra is generated at compile time from Pauli products (the
// vertex spin structure sum Gamma^{mu nu} sigma^mu (x) sigma^nu with the mixed
// density-spin components carrying a factor i); no coefficient is spelled out
// by hand.  Each table expands into straight-line FMAs on register arrays.
//
// dGamma kernel: one workgroup per frequency triple (s, t, u), 128 threads for
// lattices of at most 32 sites and 256 above, items walked t-major in XCD
// runs; lattices of 13 to 32 sites run instantiations with the LDS bundle
// geometry compiled in (every component offset an immediate).  Per channel
// the evaluation points are processed in chunks of nb points:
//   A  one thread per (point, access buffer) builds the descriptor (4 supports,
//      weights, exchange / sign flags), the T channel's four site-local
//      vertices included;
//   B  one thread per (point, buffer, site) gathers the 16-component bundle
//      into LDS (4 supports x 16 loads, component offsets through the site's
//      spin permutation); in the T channel one thread per (point, local
//      vertex, component) evaluates the site-local vertices;
//   C  one thread per (point, site, buffer pair) applies the channel algebra:
//      S / U 256 products per pair; T the lattice bubble (per overlap entry
//      64 products; the entries of a site grouped by their spin
//      transformations, so that a group's permutations become row addresses
//      and one output permutation per group), the chalice and inverse
//      chalice (256 products each), accumulated with the point's weight.
//      Lanes take the sites in decreasing order of their overlap-entry
//      counts, so that a wave's sites cost about the same.
// The per-thread sums are reduced through LDS at the end of the item.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "frg_common.h"
#include "spfrg.h"

namespace {

using namespace frg;

constexpr int kTriThreads = 256;
// dGamma workgroups of 128 threads (three evaluation points per chunk at
// L = 19) for lattices of at most kTriSmallL sites: four resident workgroups
// per CU instead of two overlap one another's barrier-separated phases
// (kagome-DM r5: 100.0 against 106.8 ms at the smallest cutoff, 3.6 against
// 5.8 ms at the largest; 64 and 192 threads were slower)
constexpr int kTriSmallThreads = 128;
constexpr int kTriSmallL = 32;
constexpr int kTriMaxPoints = 16;  // evaluation points per chunk
#ifndef SPFRG_TRI_CHUNK
#define SPFRG_TRI_CHUNK 16
#endif
constexpr long long kTriChunk = SPFRG_TRI_CHUNK;
#ifndef SPFRG_TRI_MIN_BLOCKS
#define SPFRG_TRI_MIN_BLOCKS 1
#endif

// ---------------------------------------------------------------------------
// compile-time spin algebra (Gaussian integers; traces of Pauli products are
// 0, +-2 or +-2i, so every projection below is exact)
// ---------------------------------------------------------------------------
struct GI {
  int re, im;
};
constexpr GI gmul(GI a, GI b) { return GI{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
constexpr GI gadd(GI a, GI b) { return GI{a.re + b.re, a.im + b.im}; }
constexpr GI gneg(GI a) { return GI{-a.re, -a.im}; }

struct M2 {
  GI a, b, c, d;  // [[a, b], [c, d]]
};
// component order x, y, z, d (d = identity)
constexpr M2 pauli(int k) {
  return k == 0 ? M2{{0, 0}, {1, 0}, {1, 0}, {0, 0}}
       : k == 1 ? M2{{0, 0}, {0, -1}, {0, 1}, {0, 0}}
       : k == 2 ? M2{{1, 0}, {0, 0}, {0, 0}, {-1, 0}}
                : M2{{1, 0}, {0, 0}, {0, 0}, {1, 0}};
}
constexpr M2 mm(M2 x, M2 y) {
  return M2{gadd(gmul(x.a, y.a), gmul(x.b, y.c)), gadd(gmul(x.a, y.b), gmul(x.b, y.d)),
            gadd(gmul(x.c, y.a), gmul(x.d, y.c)), gadd(gmul(x.c, y.b), gmul(x.d, y.d))};
}
constexpr GI trace(M2 x) { return gadd(x.a, x.d); }
// coefficient of sigma^r in x
constexpr GI proj(int r, M2 x) {
  const GI t = trace(mm(pauli(r), x));
  return GI{t.re / 2, t.im / 2};
}
// sigma^a sigma^b = sum_c f[a][b][c] sigma^c
struct PauliProducts {
  GI f[4][4][4];
};
constexpr PauliProducts make_products() {
  PauliProducts p{};
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) p.f[a][b][c] = proj(c, mm(pauli(a), pauli(b)));
  return p;
}
constexpr PauliProducts kPP = make_products();
// coefficients of sigma^r in sigma^a sigma^b and in sigma^a sigma^b sigma^c
constexpr GI p2(int r, int a, int b) { return kPP.f[a][b][r]; }
constexpr GI p3(int r, int a, int b, int c) {
  GI v{0, 0};
  for (int d = 0; d < 4; ++d) v = gadd(v, gmul(kPP.f[a][b][d], kPP.f[d][c][r]));
  return v;
}
constexpr GI p1(int r, int a) { return r == a ? GI{1, 0} : GI{0, 0}; }
constexpr bool mixed(int c) { return (c / 4 == 3) != (c % 4 == 3); }
// stored components with one density and one spin index carry a factor i
constexpr GI times_phase(GI x, int c) { return mixed(c) ? GI{-x.im, x.re} : x; }
constexpr GI div_phase(GI x, int c) { return mixed(c) ? GI{x.im, -x.re} : x; }

// kind 0: particle-particle (S), 1: crossed particle-hole (U), 2: chalice,
// 3: inverse chalice.  o = output component, a / b = components of the two
// factors (chalice: a bundle, b site-local vertex; inverse chalice: a local,
// b bundle).
constexpr GI coefficient(int kind, int o, int a, int b) {
  const int r = o / 4, t = o % 4, mu = a / 4, nu = a % 4, ka = b / 4, la = b % 4;
  GI v{0, 0};
  if (kind == 0) v = gmul(p2(r, ka, mu), p2(t, la, nu));
  else if (kind == 1) v = gneg(gmul(p2(r, ka, mu), p2(t, nu, la)));
  else if (kind == 2) v = gneg(gmul(p1(r, mu), p3(t, la, nu, ka)));
  else v = gneg(gmul(p3(r, nu, ka, mu), p1(t, la)));
  return div_phase(times_phase(times_phase(v, a), b), o);
}

struct Term {
  int o, a, b, c;
};
struct TermTable {
  int n;
  bool real;
  Term t[256];
};
constexpr TermTable make_table(int kind) {
  TermTable tb{0, true, {}};
  for (int o = 0; o < 16; ++o)
    for (int a = 0; a < 16; ++a)
      for (int b = 0; b < 16; ++b) {
        const GI v = coefficient(kind, o, a, b);
        if (v.im != 0) tb.real = false;
        if (v.re != 0 && tb.n < 256) tb.t[tb.n++] = Term{o, a, b, v.re};
      }
  return tb;
}
constexpr TermTable kTab[4] = {make_table(0), make_table(1), make_table(2), make_table(3)};
static_assert(kTab[0].n == 256 && kTab[1].n == 256 && kTab[2].n == 256 && kTab[3].n == 256, "term count");
static_assert(kTab[0].real && kTab[1].real && kTab[2].real && kTab[3].real, "complex coefficient");

// lattice bubble: out(rho, tau) += 2 g A(rho, nu) B(nu, tau), closed loop over nu
constexpr int bubble_sign(int r, int nu, int t) {
  const GI v = div_phase(times_phase(times_phase(GI{1, 0}, 4 * r + nu), 4 * nu + t), 4 * r + t);
  return v.re;
}
// measurement egg diagram: out(rho, tau) += n tr(s_rho s_lambda s_tau s_kappa) / 2 (phase) v(kappa, lambda),
// n = 2 for the density output and 1/2 for the spin outputs
constexpr double egg_coefficient(int o, int cl) {
  const int r = o / 4, t = o % 4, ka = cl / 4, la = cl % 4;
  GI tr4{0, 0};  // tr(s_r s_la s_t s_ka) / 2 = identity coefficient of the product
  for (int d = 0; d < 4; ++d) tr4 = gadd(tr4, gmul(kPP.f[r][la][d], p3(3, d, t, ka)));
  const GI v = div_phase(times_phase(tr4, cl), o);
  return (o == 15 ? 2.0 : 0.5) * v.re;
}

template <int KIND, int K>
__device__ __forceinline__ void term(double (&acc)[16], const double (&x)[16], const double (&y)[16]) {
  constexpr Term e = kTab[KIND].t[K];
  if constexpr (e.c > 0) acc[e.o] = fma(x[e.a], y[e.b], acc[e.o]);
  else acc[e.o] = fma(-x[e.a], y[e.b], acc[e.o]);
}
template <int KIND, int... K>
__device__ __forceinline__ void apply_terms(double (&acc)[16], const double (&x)[16], const double (&y)[16],
                                            std::integer_sequence<int, K...>) {
  (term<KIND, K>(acc, x, y), ...);
}
template <int KIND>
__device__ __forceinline__ void apply_table(double (&acc)[16], const double (&x)[16], const double (&y)[16]) {
  apply_terms<KIND>(acc, x, y, std::make_integer_sequence<int, 256>{});
}

__device__ __forceinline__ int zeta(int s) { return s <= 2 ? -1 : 1; }

__device__ __forceinline__ double pick3(double u0, double u1, double u2, int i) {
  return i == 0 ? u0 : (i == 1 ? u1 : u2);
}

// ---------------------------------------------------------------------------
// per-flow descriptor (passed by value)
// ---------------------------------------------------------------------------
struct TriDev {
  double *G, *dG, *v2, *dv2;
  const double *mesh;
  const int *site_rid, *site_perm, *inv_rid, *inv_perm;
  // lattice-bubble tables: the overlap entries of each site grouped by their
  // spin transformations; bgrp_off[rid] .. bgrp_off[rid + 1] index bgrp =
  // {first entry, end entry, P1 | Q << 6 | P2 << 12 (2 bits per spin index), 0};
  // bent = rid1 | rid2 << 16 per entry
  const int *bgrp_off, *bent;
  const int4 *bgrp;
  // sites in decreasing order of their overlap-entry counts (channel-algebra
  // lanes take them in this order)
  const int *site_order;
  const int *range_fwd, *range_inv, *corr_rid, *corr_perm;
  double *Wwp, *Wc;
  int *Wn;
  double *Cwp, *Cw;
  int *Cn;
  double *corr_partial, *corr_out;
  int *flags;  // [0] NaN in the flow, [1] evaluation-list overflow
  int nw, L, R, B, Rb, emax, nb, euler;
  int nt;  // threads per workgroup of the dGamma kernel
  int cs;  // site stride of the LDS bundles compiled into the dGamma kernel (doubles); 0: run-time form
  long long nf, nv;
  unsigned g_bytes;
  double cut, step;
  int geo;  // geometric mesh: bracket guess from the logarithm (count_le_geo)
  float ga, gb;
};

// lattice bubble: load the next overlap entry's packed sites ahead (1) or
// right before use (0)
#ifndef SPFRG_TRI_PREFETCH
#define SPFRG_TRI_PREFETCH 1
#endif
// padding of the run-time bundle stride: bundle bases off a common bank (the
// lattice bubble's lanes read the same site of different bundles at once; a
// stride that is a multiple of 16 doubles puts every bundle on the same
// banks; measured 1 % faster at L = 19)
#ifndef SPFRG_TRI_BPAD
#define SPFRG_TRI_BPAD 6
#endif
// padding of the compiled bundle stride (doubles)
#ifndef SPFRG_TRI_CPAD
#define SPFRG_TRI_CPAD 2
#endif

// Geometry of the 16-component bundles in LDS: component (s1, s2) of site j
// of bundle d sits at d * bs + s1 * rs + s2 * ks + j (doubles).
//  * cs = 0 (site count L read at run time): rs = 4 L, ks = L, bundles of
//    16 L + pad doubles one after the other.
//  * cs > 0 (site stride cs >= L and nb compiled into the kernel): rs = cs,
//    bs = 4 cs + pad and the second spin index outermost, ks = 4 nb bs, made
//    odd.  All component offsets are then instruction immediates, and the
//    four components of one row s1 lie more than 255 doubles apart (4 nb bs >
//    255 for every compiled shape): the compiler cannot pair them into
//    ds_read2_b64, whose random-address reads cost the LDS about twice the
//    cycles of two single 8-byte reads on gfx950 (measured with the pairs:
//    SQ_LDS_IDX_ACTIVE +83 %, 8 % slower kernel).  ks is odd, so no two of the
//    offsets are multiples of 64 doubles either (no ds_read2st64_b64).
struct TriBundle {
  int rs, ks, bs, size;  // size: doubles of the whole bundle area
};
__host__ __device__ constexpr TriBundle tri_bundle(int L, int cs, int nb) {
  if (cs == 0) return TriBundle{4 * L, L, 16 * L + SPFRG_TRI_BPAD, nb * 4 * (16 * L + SPFRG_TRI_BPAD)};
  const int bs = 4 * cs + SPFRG_TRI_CPAD, ks = (nb * 4 * bs) | 1;
  return TriBundle{cs, ks, bs, 4 * ks};
}

// LDS carve-up of k_tri_dgamma (bytes)
struct TriLds {
  int mesh, cf, ci, off, w, fl, c, loc, A, total;
};
__host__ __device__ inline TriLds tri_lds(int nw, int L, int cs, int nb) {
  TriLds l;
  int o = 0;
  auto take = [&](int bytes) { int r = o; o += (bytes + 15) & ~15; return r; };
  const int ss = cs > 0 ? cs : L;  // site stride of the component-offset tables
  l.mesh = take(nw * 8);
  l.cf = take(16 * ss * 4);
  l.ci = take(16 * ss * 4);
  l.off = take(nb * 8 * 4 * 4);
  l.w = take(nb * 8 * 4 * 8);
  l.fl = take(nb * 8 * 4);
  l.c = take(nb * 8);
  l.loc = take(nb * 4 * 16 * 8);
  const int bun = tri_bundle(L, cs, nb).size, red = nb * 2 * 16 * L;
  l.A = take((bun > red ? bun : red) * 8);  // also holds the final reduction
  l.total = o;
  return l;
}

__device__ __forceinline__ double gload(__amdgpu_buffer_rsrc_t r, int byte_off) {
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0);
  double d;
  __builtin_memcpy(&d, &v, 8);
  return d;
}

// weight tables of one flow (shared with the SU(2) core's definition)
__global__ __launch_bounds__(64) void k_tri_weights(TriDev F, int corr_only) {
  extern __shared__ double s_wmesh[];
  const int x = corr_only ? F.nw : (int)blockIdx.x;
  weight_tables_wave(x, F.mesh, F.nw, F.cut, F.v2, F.dv2, F.Wwp, F.Wc, F.Wn, F.Cwp, F.Cw, F.Cn, F.emax, F.flags + 1,
                     s_wmesh);
}

// getValue(i1, i2, s, t, u, s1, s2, FrequencyChannel::None) for pairs that keep
// the requested components (density-density, or the pair (0, 0)); rid_fwd /
// rid_inv = symmetryTransform(i1, i2) / (i2, i1)
__device__ double tri_value_none(const double *G, const double *m, int nw, int L, int rid_fwd, int rid_inv, double s,
                                 double t, double u, int s1, int s2) {
  double sign = 1.0;
  bool swap = false;
  if (s < 0) { s = -s; const int q = s1; s1 = s2; s2 = q; swap = !swap; }
  if (t < 0) { t = -t; sign *= double(zeta(s1) * zeta(s2)); }
  if (u < 0) { u = -u; const int q = s1; s1 = s2; s2 = q; swap = !swap; sign *= double(zeta(s1) * zeta(s2)); }
  const int rid = swap ? rid_inv : rid_fwd;
  int lS, uS, lT, uT, lU, uU;
  double bS, bT, bU;
  interp(m, nw, s, lS, uS, bS);
  interp(m, nw, t, lT, uT, bT);
  interp(m, nw, u, lU, uU, bU);
  const size_t FS = 16 * (size_t)L;
  const int comp = (4 * s1 + s2) * L + rid;
  auto A = [&](int so, int to, int uo) {
    if (so < uo) return -double(zeta(s2)) * G[((size_t)tri(uo, so) * nw + to) * FS + comp];
    return G[((size_t)tri(so, uo) * nw + to) * FS + comp];
  };
  return sign * ((1 - bU) * ((1 - bT) * ((1 - bS) * A(lS, lT, lU) + bS * A(uS, lT, lU)) +
                             bT * ((1 - bS) * A(lS, uT, lU) + bS * A(uS, uT, lU))) +
                 bU * ((1 - bT) * ((1 - bS) * A(lS, lT, uU) + bS * A(uS, lT, uU)) +
                       bT * ((1 - bS) * A(lS, uT, uU) + bS * A(uS, uT, uU))));
}

// dSigma/dLambda (src/TRI/TRIFrgCore.cpp:122-151): one wavefront per mesh point
__global__ __launch_bounds__(64) void k_tri_dsigma(TriDev F) {
  const int i = blockIdx.x;
  if (i >= F.nw) return;
  const int lane = threadIdx.x;
  const int nw = F.nw, L = F.L;
  const double *m = F.mesh;
  const double cut = F.cut;
  const double w = m[i];
  double sum = 0.0;
  for (int j = lane; j < F.R; j += 64) {
    sum += tri_value_none(F.G, m, nw, L, F.range_fwd[j], F.range_inv[j], w + cut, 0.0, w - cut, 3, 3) -
           tri_value_none(F.G, m, nw, L, F.range_fwd[j], F.range_inv[j], w - cut, 0.0, w + cut, 3, 3);
  }
  sum = wave_sum(sum);
  if (lane == 0) {
    double v = -2.0 * sum;
    for (int k = 0; k < 4; ++k)
      v += tri_value_none(F.G, m, nw, L, 0, 0, w + cut, w - cut, 0.0, k, k) -
           tri_value_none(F.G, m, nw, L, 0, 0, w - cut, w + cut, 0.0, k, k);
    v /= (kTwoPi * (cut + sigma_at(m, nw, F.v2, cut)));
    F.dv2[i] = v;
    if (isnan(v)) atomicOr(&F.flags[0], 1);
  }
}

// frequency arguments of access buffer k in channel ch (0 S, 1 T, 2 U) and the
// channel of exactness (src/TRI/TRIFrgCore.cpp:181-188, 714-724, 1258-1260,
// 1815-1817, 2371-2379)
__device__ __forceinline__ void tri_args(int ch, int k, double s, double t, double u, double w1, double w1p, double w2,
                                         double w2p, double wp, double &a, double &b, double &c, int &chan) {
  if (ch == 0) {
    chan = 0;
    a = s;
    if (k == 0) { b = w2 + wp; c = w1 + wp; }
    else if (k == 1) { b = -w2p - wp; c = w1p + wp; }
    else if (k == 2) { b = -w1 - wp; c = -w2 - wp; }
    else { b = w1p + wp; c = -w2p - wp; }
  } else if (ch == 1) {
    if (k < 4) {
      chan = 1;
      b = t;
      if (k == 0) { a = w1 - wp; c = w1p + wp; }
      else if (k == 1) { a = w2p - wp; c = -w2 - wp; }
      else if (k == 2) { a = w1p + wp; c = w1 - wp; }
      else { a = w2 + wp; c = -w2p + wp; }
    } else {
      chan = 2;
      if (k == 4) { a = w2p - wp; b = -w2 - wp; c = t; }
      else if (k == 5) { a = w2 + wp; b = -w2p + wp; c = t; }
      else if (k == 6) { a = w1 - wp; b = -w1p - wp; c = -t; }
      else { a = w1p + wp; b = -w1 + wp; c = -t; }
    }
  } else {
    chan = 2;
    c = u;
    if (k == 0) { a = w1 + wp; b = -w2p + wp; }
    else if (k == 1) { a = w1p + wp; b = w2 - wp; }
    else if (k == 2) { a = w2p - wp; b = -w1 - wp; }
    else { a = w2 - wp; b = w1p + wp; }
  }
}

// support offset (elements) with the s <-> u exchange flag
__device__ __forceinline__ int tri_support(int so, int to, int uo, int nw, int FS, bool &fx) {
  fx = so < uo;
  const int row = fx ? tri(uo, so) * nw + to : tri(so, uo) * nw + to;
  return row * FS;
}

// CS, NB > 0: site stride of the bundles and evaluation points per chunk
// compiled in (F.nb == NB); CS = NB = 0: both read at run time
template <int NT, int CS, int NB>
__global__ __launch_bounds__(NT, SPFRG_TRI_MIN_BLOCKS) void k_tri_dgamma(TriDev F) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const long long total = F.nf;
  // workgroups are dispatched round-robin over the 8 XCDs: each XCD receives
  // runs of kTriChunk consecutive positions, and positions walk the items
  // t-major (every (s,u) row of one t before the next t), so the workgroups
  // resident on an XCD share the t-slab that all T-channel bundles read
  // (N_w (N_w + 1) / 2 support rows, 2.9 MB at L = 19, N_w = 48)
  const long long j0 = blockIdx.x >> 3;
  const long long pos = ((j0 / kTriChunk) * 8 + (blockIdx.x & 7)) * kTriChunk + (j0 % kTriChunk);
  if (pos >= total) return;

  const int nw = F.nw, L = F.L, nb = NB > 0 ? NB : F.nb;
  const long long nsu = total / nw;
  const long long item = (pos % nsu) * nw + pos / nsu;
  const int FS = 16 * L;
  // bundle geometry: compile-time constants for CS > 0, so that the
  // component offsets of every bundle read and write are immediates (the
  // lattice bubble's 32 reads per overlap entry then need 9 address
  // instructions instead of 38)
  const TriBundle geo = tri_bundle(L, CS, nb);
  const int RS = geo.rs, KS = geo.ks, BS = geo.bs;
  const int SS = CS > 0 ? CS : L;  // site stride of s_cf, s_ci
  // offset of component k = 4 s1 + s2 within a bundle
  auto comp = [&](int k) { return CS > 0 ? (k >> 2) * RS + (k & 3) * KS : k * L; };
  const TriLds lay = tri_lds(nw, L, CS, nb);
  double *s_mesh = reinterpret_cast<double *>(lds_raw + lay.mesh);
  int *s_cf = reinterpret_cast<int *>(lds_raw + lay.cf);
  int *s_ci = reinterpret_cast<int *>(lds_raw + lay.ci);
  int *s_off = reinterpret_cast<int *>(lds_raw + lay.off);
  double *s_w = reinterpret_cast<double *>(lds_raw + lay.w);
  int *s_fl = reinterpret_cast<int *>(lds_raw + lay.fl);
  double *s_c = reinterpret_cast<double *>(lds_raw + lay.c);
  double *s_loc = reinterpret_cast<double *>(lds_raw + lay.loc);
  double *s_A = reinterpret_cast<double *>(lds_raw + lay.A);

  const __amdgpu_buffer_rsrc_t rG = __builtin_amdgcn_make_buffer_rsrc((void *)F.G, (short)0, (int)F.g_bytes, 0x00020000);
  const int tid = threadIdx.x;

  // SU2VertexTwoParticle-style iterator: item = tri(s,u) * nw + t
  const int su = (int)(item / nw);
  const int to = (int)(item - (long long)su * nw);
  int so = (int)((sqrt(8.0 * su + 1.0) - 1.0) * 0.5);
  while ((so + 1) * (so + 2) / 2 <= su) ++so;
  while (so * (so + 1) / 2 > su) --so;
  const int uo = su - so * (so + 1) / 2;

  for (int i = tid; i < nw; i += NT) s_mesh[i] = F.mesh[i];
  // stored component offsets (component, site) for the plain and the exchanged
  // bundle: components swapped on exchange, spin indices through the site's
  // permutation (TRIVertexTwoParticle.hpp:368-388)
  for (int q = tid; q < 16 * L; q += NT) {
    const int c = q / L, j = q - (q / L) * L;
    const int s1 = c / 4, s2 = c % 4;
    const int *pf = F.site_perm + 3 * j, *pi = F.inv_perm + 3 * j;
    const int a = s1 < 3 ? pf[s1] : 3, b = s2 < 3 ? pf[s2] : 3;
    s_cf[c * SS + j] = (4 * a + b) * L + F.site_rid[j];
    const int ai = s2 < 3 ? pi[s2] : 3, bi = s1 < 3 ? pi[s1] : 3;
    s_ci[c * SS + j] = (4 * ai + bi) * L + F.inv_rid[j];
  }
  __syncthreads();

  const double s = s_mesh[so], t = s_mesh[to], u = s_mesh[uo];
  const double w1p = 0.5 * (s + t + u);
  const double w1 = 0.5 * (s - t + u);
  const double w2p = 0.5 * (s - t - u);
  const double w2 = 0.5 * (s + t - u);

  // phase C role: (point, site, pair).  Lanes come in runs of 2 nb, one run
  // per site, sites in decreasing order of their lattice-bubble entry counts:
  // a wave then holds sites of similar cost, and its bubble loop runs about as
  // long as its own lanes need (modelled at L = 19: 24 % fewer wave
  // iterations than sites in index order; measured 2.5 % faster)
  const int crank = tid / (2 * nb);
  const int crem = tid - crank * 2 * nb;
  const int ce = crem >> 1, cpair = crem & 1;
  const bool cactive = crank < L;
  const int cj = cactive ? F.site_order[crank] : 0;
  double acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.0;

  for (int ch = 0; ch < 3; ++ch) {
    const int xo = ch == 0 ? so : (ch == 1 ? to : uo);
    const int n = F.Wn[xo];
    const double *Wwp = F.Wwp + (size_t)xo * F.emax;
    const double *Wc = F.Wc + (size_t)xo * F.emax;
    const int nk = ch == 1 ? 8 : 4;
    for (int e0 = 0; e0 < n; e0 += nb) {
      const int nvalid = n - e0 < nb ? n - e0 : nb;
      // (A) descriptors; site-local vertices of the T channel
      if (tid < nvalid * nk) {
        const int e = tid / nk, k = tid - (tid / nk) * nk;
        const double wp = Wwp[e0 + e];
        double a, b, c;
        int chan;
        tri_args(ch, k, s, t, u, w1, w1p, w2, w2p, wp, a, b, c, chan);
        const bool exch = (a < 0) != (c < 0);
        const bool tu = (b < 0) != (c < 0);
        a = fabs(a); b = fabs(b); c = fabs(c);
        int lo1, hi1, lo2, hi2;
        double b1, b2;
        const double x1 = chan == 0 ? b : a, x2 = chan == 2 ? b : c;
        interp_sel(s_mesh, nw, x1, F.geo, F.ga, F.gb, lo1, hi1, b1);
        interp_sel(s_mesh, nw, x2, F.geo, F.ga, F.gb, lo2, hi2, b2);
        int off[4];
        double w[4];
        int fl = (exch ? 1 : 0) | (tu ? 2 : 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int l1 = (q & 1) ? hi1 : lo1, l2 = (q & 2) ? hi2 : lo2;
          w[q] = ((q & 2) ? b2 : (1 - b2)) * ((q & 1) ? b1 : (1 - b1));
          bool fx;
          if (chan == 0) off[q] = tri_support(xo, l1, l2, nw, FS, fx);
          else if (chan == 1) off[q] = tri_support(l1, xo, l2, nw, FS, fx);
          else off[q] = tri_support(l1, l2, xo, nw, FS, fx);
          if (fx) fl |= 4 << q;
        }
        // descriptor slot: bundles e * 4 + k, the T channel's site-local
        // vertices after all bundles (nb * 4 + e * 4 + k - 4)
        const int d = k < 4 ? e * 4 + k : nb * 4 + e * 4 + (k - 4);
#pragma unroll
        for (int q = 0; q < 4; ++q) { s_off[d * 4 + q] = off[q]; s_w[d * 4 + q] = w[q]; }
        s_fl[d] = fl;
        if (k == 0) s_c[e] = Wc[e0 + e];
      }
      __syncthreads();
      // (B) bundles: (point, buffer, site) -> 16 components into LDS; in the T
      // channel also the site-local vertices, one thread per (point, local, component)
      const int nbq = nvalid * 4 * L;
      const int nq = nbq + (ch == 1 ? nvalid * 4 * 16 : 0);
      for (int q = tid; q < nq; q += NT) {
        if (q >= nbq) {
          // getValueLocal (TRIVertexTwoParticle.hpp:300-309): site 0; the sign
          // table is indexed with the exchanged components
          const int lq = q - nbq;
          const int dl = lq >> 4, cc = lq & 15;  // dl = e * 4 + local
          const int d = nb * 4 + dl;
          const int fl = s_fl[d];
          const bool exch = fl & 1, tu = (fl >> 1) & 1;
          int s1 = cc / 4, s2 = cc % 4;
          if (exch) { const int x = s1; s1 = s2; s2 = x; }
          const double tsign = tu ? double(zeta(s1) * zeta(s2)) : 1.0;
          const double fsign = -double(zeta(exch ? s1 : s2));
          double v = 0.0;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double sg = ((fl >> (2 + r)) & 1) ? tsign * fsign : tsign;
            v = fma(sg * s_w[d * 4 + r], gload(rG, (s_off[d * 4 + r] + (4 * s1 + s2) * L) * 8), v);
          }
          s_loc[dl * 16 + cc] = v;
          continue;
        }
        const int d = q / L, j = q - (q / L) * L;  // d = e * 4 + k
        const int fl = s_fl[d];
        const bool exch = fl & 1, tu = (fl >> 1) & 1;
        const int *cofs = exch ? s_ci : s_cf;
        int off[4];
        double w[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) { off[r] = s_off[d * 4 + r]; w[r] = s_w[d * 4 + r]; }
        double *dst = s_A + (size_t)d * BS + j;
        // components per batch of loads: four (16 loads in flight) keep the
        // register footprint of the gather small next to the live
        // accumulators; the 128-thread shape, with four workgroups per CU,
        // gains up to 1 % from eight (256 threads: 3 % slower)
        constexpr int kComp = NT == kTriSmallThreads ? 2 : 1;  // rows s1 per batch
#pragma unroll 1
        for (int s1b = 0; s1b < 4; s1b += kComp) {
#pragma unroll
          for (int s1q = 0; s1q < kComp; ++s1q) {
            const int s1 = s1b + s1q;
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2) {
              const int cc = 4 * s1 + s2;
              const double tsign = tu ? double(zeta(s1) * zeta(s2)) : 1.0;
              const double fsign = -double(zeta(exch ? s1 : s2));
              const int co = cofs[cc * SS + j];
              double v = 0.0;
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const double sg = ((fl >> (2 + r)) & 1) ? tsign * fsign : tsign;
                v = fma(sg * w[r], gload(rG, (off[r] + co) * 8), v);
              }
              dst[comp(cc)] = v;
            }
          }
        }
      }
      __syncthreads();
      // (C) channel algebra for (point, site, pair)
      if (cactive && ce < nvalid) {
        const double cw = s_c[ce];
        const double *X = s_A + (size_t)(ce * 4 + 2 * cpair) * BS;      // first bundle of the pair
        const double *Y = s_A + (size_t)(ce * 4 + 2 * cpair + 1) * BS;  // second bundle
        double x[16], y[16];
        if (ch != 1) {
#pragma unroll
          for (int k = 0; k < 16; ++k) { x[k] = cw * X[comp(k) + cj]; y[k] = Y[comp(k) + cj]; }
          if (ch == 0) apply_table<0>(acc, x, y);
          else apply_table<1>(acc, x, y);
        } else {
          // lattice bubble (TRIFrgCore.cpp:733-1254): entry i of getOverlap(rid)
          // pairs site rid1 of the first bundle A and rid2 of the second B with
          // their spin transformations P1, P2 (density index fixed):
          //   out(rr, tt) += 2 g sum_nn sgn(rr, nn, tt) A(P1 rr, P1 nn) B(P2 nn, P2 tt).
          // The entries of a site come grouped by (P1, P2).  With m = P1 nn and
          // Q = P2 P1^-1 an entry adds
          //   o(rr, c) += sum_m sgn(rr, m, c) A(P1 rr, m) B(Q m, c)
          // (the sign depends only on which indices are the density, which
          // spin permutations keep), so rows are picked by address; the group
          // then adds 2 g o(rr, P2 tt) to out(rr, tt).
          const int gb = F.bgrp_off[cj], ge = F.bgrp_off[cj + 1];
          for (int gi = gb; gi < ge; ++gi) {
            const int4 gr = F.bgrp[gi];
            const int code = gr.z;
            const double c2 = 2.0 * cw * (double)gr.w;  // gr.w: multiplicity of the group's entries
            // component rows (4 components, stride L) of A and B
            const int xr[4] = {(code & 3) * RS, ((code >> 2) & 3) * RS, ((code >> 4) & 3) * RS, 3 * RS};
            const int yr[4] = {((code >> 6) & 3) * RS, ((code >> 8) & 3) * RS, ((code >> 10) & 3) * RS, 3 * RS};
            double o[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) o[k] = 0.0;
            // the next entry's packed sites are loaded while the current
            // entry is evaluated (the load is dependent: its result forms
            // the LDS addresses of the next 32 reads)
            int pk_next = (SPFRG_TRI_PREFETCH && gr.x < gr.y) ? F.bent[gr.x] : 0;
            for (int i = gr.x; i < gr.y; ++i) {
              const int pk = SPFRG_TRI_PREFETCH ? pk_next : F.bent[i];
              if (SPFRG_TRI_PREFETCH && i + 1 < gr.y) pk_next = F.bent[i + 1];
              const double *xa = X + (pk & 0xffff);
              const double *ya = Y + (pk >> 16);
              // all 32 operands are loaded first (the scheduling barrier keeps
              // the loads ahead of the products): with compiled offsets the
              // compiler would otherwise reuse the registers of A's rows and
              // wait for each row in turn (2 % slower at L = 19); with run-time
              // offsets the interleaved form keeps the address registers down
              double bv[16], av[16];
#pragma unroll
              for (int mm = 0; mm < 4; ++mm)
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) bv[4 * mm + cc] = ya[yr[mm] + cc * KS];
              if constexpr (CS > 0) {
#pragma unroll
                for (int rr = 0; rr < 4; ++rr)
#pragma unroll
                  for (int mm = 0; mm < 4; ++mm) av[4 * rr + mm] = xa[xr[rr] + mm * KS];
                __builtin_amdgcn_sched_barrier(0);
              }
#pragma unroll
              for (int rr = 0; rr < 4; ++rr) {
                if constexpr (CS == 0) {
#pragma unroll
                  for (int mm = 0; mm < 4; ++mm) av[4 * rr + mm] = xa[xr[rr] + mm * KS];
                }
#pragma unroll
                for (int mm = 0; mm < 4; ++mm)
#pragma unroll
                  for (int cc = 0; cc < 4; ++cc) {
                    const double a = av[4 * rr + mm];
                    o[4 * rr + cc] = fma(bubble_sign(rr, mm, cc) > 0 ? a : -a, bv[4 * mm + cc], o[4 * rr + cc]);
                  }
              }
            }
            const int p0 = (code >> 12) & 3, p1 = (code >> 14) & 3, p2 = (code >> 16) & 3;
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
              const double *orow = o + 4 * rr;
              acc[4 * rr + 0] = fma(c2, pick3(orow[0], orow[1], orow[2], p0), acc[4 * rr + 0]);
              acc[4 * rr + 1] = fma(c2, pick3(orow[0], orow[1], orow[2], p1), acc[4 * rr + 1]);
              acc[4 * rr + 2] = fma(c2, pick3(orow[0], orow[1], orow[2], p2), acc[4 * rr + 2]);
              acc[4 * rr + 3] = fma(c2, orow[3], acc[4 * rr + 3]);
            }
          }
          // chalice (pair 0: bundle 0 with local 4; pair 1: bundle 2 with local 5)
          const double *lc = s_loc + (ce * 4 + cpair) * 16;
#pragma unroll
          for (int k = 0; k < 16; ++k) { x[k] = cw * X[comp(k) + cj]; y[k] = lc[k]; }
          apply_table<2>(acc, x, y);
          // inverse chalice (pair 0: local 6 with bundle 1; pair 1: local 7 with bundle 3)
          const double *li = s_loc + (ce * 4 + 2 + cpair) * 16;
#pragma unroll
          for (int k = 0; k < 16; ++k) { x[k] = cw * li[k]; y[k] = Y[comp(k) + cj]; }
          apply_table<3>(acc, x, y);
        }
      }
      __syncthreads();
    }
  }
  // reduce the (point, pair) partial sums per (component, site); prefactor 1/(2 pi)
  double *red = s_A;
  if (cactive) {
    const int g = ce * 2 + cpair;
#pragma unroll
    for (int k = 0; k < 16; ++k) red[((size_t)g * 16 + k) * L + cj] = acc[k];
  }
  __syncthreads();
  const int ng = nb * 2;
  for (int q = tid; q < 16 * L; q += NT) {
    double v = 0.0;
    for (int g = 0; g < ng; ++g) v += red[(size_t)g * 16 * L + q];
    v *= (1.0 / kTwoPi);
    F.dG[(size_t)item * FS + q] = v;
    if (isnan(v)) atomicOr(&F.flags[0], 1);
  }
}

// in-place explicit Euler step (src/TRI/TRIFrgCore.cpp:98-120); skipped when
// the flow holds a NaN (src/SpinParser.cpp:151-159)
__global__ void k_tri_euler(TriDev F) {
  if (!F.euler || F.flags[0] != 0) return;
  const double step = F.step;
  const size_t n = (size_t)F.nv;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) F.G[i] = fma(step, F.dG[i], F.G[i]);
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < F.nw; i += blockDim.x) F.v2[i] = fma(step, F.dv2[i], F.v2[i]);
}

// ---------------------------------------------------------------------------
// static susceptibility (src/TRI/TRIMeasurementCorrelation.cpp:147-316), nu = 0
// ---------------------------------------------------------------------------
constexpr int kOut[10] = {15, 0, 1, 2, 4, 5, 6, 8, 9, 10};

template <int G>
__device__ __forceinline__ void egg_row(double (&r)[10], const double (&v)[16]) {
#pragma unroll
  for (int cl = 0; cl < 16; ++cl) {
    const double e = egg_coefficient(kOut[G], cl);
    if (e != 0.0) r[G] = fma(e, v[cl], r[G]);
  }
}
template <int... G>
__device__ __forceinline__ void egg_rows(double (&r)[10], const double (&v)[16], std::integer_sequence<int, G...>) {
  (egg_row<G>(r, v), ...);
}

// trilinear descriptor of generateAccessBuffer(s, t, u) (8 supports)
struct Ab8 {
  int off[8];
  double w[8];
  int fl;  // bit 0 exchange, bit 1 t/u sign flip, bits 2..9 s<->u exchange per support
};
__device__ __forceinline__ Ab8 ab8(const double *m, int nw, int FS, double a, double b, double c) {
  Ab8 r;
  const bool exch = (a < 0) != (c < 0);
  const bool tu = (b < 0) != (c < 0);
  a = fabs(a); b = fabs(b); c = fabs(c);
  int lS, uS, lT, uT, lU, uU;
  double bS, bT, bU;
  interp(m, nw, a, lS, uS, bS);
  interp(m, nw, b, lT, uT, bT);
  interp(m, nw, c, lU, uU, bU);
  r.fl = (exch ? 1 : 0) | (tu ? 2 : 0);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int S = (i & 1) ? uS : lS, T = (i & 2) ? uT : lT, U = (i & 4) ? uU : lU;
    r.w[i] = ((i & 2) ? bT : (1 - bT)) * ((i & 1) ? bS : (1 - bS)) * ((i & 4) ? bU : (1 - bU));
    bool fx;
    r.off[i] = tri_support(S, T, U, nw, FS, fx);
    if (fx) r.fl |= 4 << i;
  }
  return r;
}

__global__ __launch_bounds__(kTriThreads) void k_tri_corr_partial(TriDev F) {
  __shared__ double red[kTriThreads * 10];
  const int o = blockIdx.x;
  const int n = *F.Cn;
  if (o >= n) return;
  const int nw = F.nw, L = F.L, FS = 16 * L;
  const double *m = F.mesh, *v2 = F.v2;
  const double *G = F.G;
  const int slots = kTriThreads / L;
  const int tid = threadIdx.x;
  const int slot = tid / L, rid = tid - (tid / L) * L;
  const bool active = slot < slots;
  const double nu = 0.0;
  const double w = F.Cwp[o];
  const double gw = (w + sigma_at(m, nw, v2, w)) * (w + nu + sigma_at(m, nw, v2, w + nu));
  double acc[10];
#pragma unroll
  for (int g = 0; g < 10; ++g) acc[g] = 0.0;
  if (active) {
    for (int e = slot; e < n; e += slots) {
      const double wp = F.Cwp[e];
      double r[10];
      // dumbbell: bundle of generateAccessBuffer(w + wp + nu, nu, w - wp)
      const Ab8 d = ab8(m, nw, FS, w + wp + nu, nu, w - wp);
      const bool exch = d.fl & 1, tu = (d.fl >> 1) & 1;
      const int srid = exch ? F.inv_rid[rid] : F.site_rid[rid];
      const int *perm = exch ? F.inv_perm + 3 * rid : F.site_perm + 3 * rid;
#pragma unroll
      for (int g = 0; g < 10; ++g) {
        const int c = kOut[g], s1 = c / 4, s2 = c % 4;
        int a = exch ? s2 : s1, b = exch ? s1 : s2;
        if (a < 3) a = perm[a];
        if (b < 3) b = perm[b];
        const double tsign = tu ? double(zeta(s1) * zeta(s2)) : 1.0;
        const double fsign = -double(zeta(exch ? s1 : s2));
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const double sg = ((d.fl >> (2 + i)) & 1) ? tsign * fsign : tsign;
          v = fma(sg * d.w[i], G[(size_t)d.off[i] + (4 * a + b) * L + srid], v);
        }
        r[g] = (c == 15 ? -4.0 : -1.0) * v;
      }
      if (rid == 0) {
        // egg: site-local vertex of generateAccessBuffer(w + wp + nu, w - wp, nu)
        const Ab8 q = ab8(m, nw, FS, w + wp + nu, w - wp, nu);
        const bool qx = q.fl & 1, qt = (q.fl >> 1) & 1;
        double v[16];
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
          int s1 = cc / 4, s2 = cc % 4;
          if (qx) { const int z = s1; s1 = s2; s2 = z; }
          const double tsign = qt ? double(zeta(s1) * zeta(s2)) : 1.0;
          const double fsign = -double(zeta(qx ? s1 : s2));
          double x = 0.0;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const double sg = ((q.fl >> (2 + i)) & 1) ? tsign * fsign : tsign;
            x = fma(sg * q.w[i], G[(size_t)q.off[i] + (4 * s1 + s2) * L], x);
          }
          v[cc] = x;
        }
        egg_rows(r, v, std::make_integer_sequence<int, 10>{});
      }
      const double norm = 1.0 / (gw * (wp + sigma_at(m, nw, v2, wp)) * (wp + nu + sigma_at(m, nw, v2, wp + nu)) *
                                 (4.0 * kPi * kPi));
      const double cf = F.Cw[e] * norm;
#pragma unroll
      for (int g = 0; g < 10; ++g) acc[g] = fma(cf, r[g], acc[g]);
    }
  }
#pragma unroll
  for (int g = 0; g < 10; ++g) red[g * kTriThreads + tid] = active ? acc[g] : 0.0;
  __syncthreads();
  for (int q = tid; q < 10 * L; q += kTriThreads) {
    const int g = q / L, r = q - (q / L) * L;
    double tsum = 0.0;
    for (int k = 0; k < slots; ++k) tsum += red[g * kTriThreads + k * L + r];
    if (r == 0) {
      // bubble term (site 0): density 2, spin diagonal 1/2
      const double term1 = 1.0 / gw;
      if (g == 0) tsum += 2.0 * term1 / kTwoPi;
      else if (g == 1 || g == 5 || g == 9) tsum += 0.5 * term1 / kTwoPi;
    }
    F.corr_partial[((size_t)o * 10 + g) * L + r] = F.Cw[o] * tsum;
  }
}

// sum over outer points; output groups DD, XX, XY, XZ, YX, YY, YZ, ZX, ZY, ZZ
// with the components of each (basis, site) pair through its spin permutation
__global__ void k_tri_corr_final(TriDev F) {
  __shared__ double susc[10 * 64];
  const int n = *F.Cn;
  const int L = F.L;
  for (int q = threadIdx.x; q < 10 * L; q += blockDim.x) {
    double v = 0.0;
    for (int o = 0; o < n; ++o) v += F.corr_partial[(size_t)o * 10 * L + q];
    susc[q] = v;
  }
  __syncthreads();
  // position of component c among the 10 stored outputs
  auto slot_of = [](int c) {
    for (int g = 0; g < 10; ++g)
      if (kOut[g] == c) return g;
    return 0;
  };
  const int nout = F.B * F.Rb;
  for (int k = threadIdx.x; k < nout; k += blockDim.x) {
    const int rid = F.corr_rid[k];
    const int sx = F.corr_perm[3 * k], sy = F.corr_perm[3 * k + 1], sz = F.corr_perm[3 * k + 2];
    const int comp[10] = {15, 4 * sx + sx, 4 * sx + sy, 4 * sx + sz, 4 * sy + sx,
                          4 * sy + sy, 4 * sy + sz, 4 * sz + sx, 4 * sz + sy, 4 * sz + sz};
#pragma unroll
    for (int g = 0; g < 10; ++g) F.corr_out[(size_t)g * nout + k] = susc[slot_of(comp[g]) * L + rid];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct sp_tri_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  TriDev fd{};
  double cutoff = 0.0;
  bool have_flow = false;
  size_t lds = 0;
  std::vector<void *> allocs;
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[4];
  double prof_ms[4] = {0, 0, 0, 0};
  long prof_n[4] = {0, 0, 0, 0};
  std::string err;
};

namespace {

int tfail(sp_tri_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define SP_THIP(call)                                                                                      \
  do {                                                                                                     \
    hipError_t e_ = (call);                                                                                \
    if (e_ != hipSuccess) return tfail(ctx, SP_EHIP, std::string(#call ": ") + hipGetErrorString(e_));   \
  } while (0)

template <class T> int talloc(sp_tri_ctx *ctx, T **p, size_t count) {
  void *q = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) return tfail(ctx, SP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return SP_OK;
}

struct TriProf {
  sp_tri_ctx *ctx;
  int k;
  hipEvent_t a = nullptr, b = nullptr;
  TriProf(sp_tri_ctx *c, int kernel) : ctx(c), k(kernel) {
    if (ctx->prof && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, ctx->stream);
  }
  ~TriProf() {
    if (ctx->prof && a && b) {
      (void)hipEventRecord(b, ctx->stream);
      ctx->pending[k].push_back({a, b});
    }
  }
};

int check_tri_problem(sp_tri_ctx *ctx, const sp_tri_problem *p) {
  if (!p || p->nw < 2 || p->nw > 4096 || !p->mesh) return tfail(ctx, SP_EARG, "invalid frequency mesh");
  for (int i = 0; i < p->nw; ++i)
    if (!(p->mesh[i] > 0.0f) || (i > 0 && !(p->mesh[i] > p->mesh[i - 1])))
      return tfail(ctx, SP_EARG, "mesh must be positive and strictly ascending");
  if (p->L < 1 || p->L > 64) return tfail(ctx, SP_EARG, "number of representative sites must be in [1, 64]");
  if (!p->site_rid || !p->site_perm || !p->inv_rid || !p->inv_perm || !p->ovl_off || !p->range_fwd ||
      !p->range_inv || !p->corr_rid || !p->corr_perm)
    return tfail(ctx, SP_EARG, "missing lattice table");
  if (p->R < 1 || p->B < 1 || p->Rb < 1) return tfail(ctx, SP_EARG, "invalid range sizes");
  if (p->ovl_off[0] != 0) return tfail(ctx, SP_EARG, "overlap offsets must start at 0");
  for (int r = 0; r < p->L; ++r)
    if (p->ovl_off[r + 1] < p->ovl_off[r]) return tfail(ctx, SP_EARG, "overlap offsets must be non-decreasing");
  const int O = p->ovl_off[p->L];
  if (O > 0 && (!p->ovl_rid1 || !p->ovl_rid2 || !p->ovl_perm1 || !p->ovl_perm2))
    return tfail(ctx, SP_EARG, "missing overlap lists");
  auto in_range = [&](const int *a, long n, int hi) {
    for (long i = 0; i < n; ++i)
      if (a[i] < 0 || a[i] >= hi) return false;
    return true;
  };
  const long nc = (long)p->B * p->Rb;
  if (!in_range(p->site_rid, p->L, p->L) || !in_range(p->inv_rid, p->L, p->L) || !in_range(p->ovl_rid1, O, p->L) ||
      !in_range(p->ovl_rid2, O, p->L) || !in_range(p->range_fwd, p->R, p->L) || !in_range(p->range_inv, p->R, p->L) ||
      !in_range(p->corr_rid, nc, p->L))
    return tfail(ctx, SP_EARG, "lattice table entry outside [0, L)");
  if (!in_range(p->site_perm, 3L * p->L, 3) || !in_range(p->inv_perm, 3L * p->L, 3) ||
      !in_range(p->ovl_perm1, 3L * O, 3) || !in_range(p->ovl_perm2, 3L * O, 3) || !in_range(p->corr_perm, 3 * nc, 3))
    return tfail(ctx, SP_EARG, "spin permutation entry outside [0, 3)");
  // the overlap transformations are symmetry operations: permutations of x, y, z
  auto is_perm = [](const int *q) { return (1 << q[0] | 1 << q[1] | 1 << q[2]) == 7; };
  for (int i = 0; i < O; ++i)
    if (!is_perm(p->ovl_perm1 + 3 * i) || !is_perm(p->ovl_perm2 + 3 * i))
      return tfail(ctx, SP_EARG, "overlap spin transformation is not a permutation of x, y, z");
  return SP_OK;
}

// shapes (site stride, points per chunk) compiled into the 128-thread dGamma
// kernel, packed as cs * 32 + nb: L = 13..32 rounded up to a multiple of 4
// with nb = 64 / L (L = 21, (24, 3), exceeds the LDS budget).  A shape is used
// while its workgroup stays within kTriLdsBudget (four resident per CU);
// other lattices, and the 256-thread kernel, read L and nb at run time
constexpr int kTriLdsBudget = 160 * 1024 / 4;
constexpr int tri_shape(int cs, int nb) { return cs * 32 + nb; }
using TriShapes = std::integer_sequence<int, 0, tri_shape(16, 4), tri_shape(20, 3), tri_shape(24, 2), tri_shape(28, 2),
                                        tri_shape(32, 2)>;
static_assert(kTriSmallL == 32, "TriShapes covers the lattices up to kTriSmallL sites");
template <int... SH> constexpr bool tri_shapes_unpaired(std::integer_sequence<int, SH...>) {
  return ((SH == 0 || tri_bundle(1, SH / 32, SH % 32).ks > 255) && ...);
}
static_assert(tri_shapes_unpaired(TriShapes{}), "the columns of a compiled shape must lie over 255 doubles apart");

template <int... SH> constexpr bool tri_shape_compiled(int cs, int nb, std::integer_sequence<int, SH...>) {
  return ((SH != 0 && SH == tri_shape(cs, nb)) || ...);
}
template <int SH> bool launch_tri_small(const TriDev &F, dim3 grid, size_t lds, hipStream_t st) {
  if (SH != (F.cs ? tri_shape(F.cs, F.nb) : 0)) return false;
  hipLaunchKernelGGL((k_tri_dgamma<kTriSmallThreads, SH / 32, SH % 32>), grid, dim3(kTriSmallThreads), lds, st, F);
  return true;
}
template <int... SH>
void launch_tri_small(const TriDev &F, dim3 grid, size_t lds, hipStream_t st, std::integer_sequence<int, SH...>) {
  (launch_tri_small<SH>(F, grid, lds, st) || ...);
}
template <int... SH> hipError_t tri_allow_lds(int bytes, std::integer_sequence<int, SH...>) {
  hipError_t e = hipSuccess;
  ((e == hipSuccess ? (e = hipFuncSetAttribute((const void *)k_tri_dgamma<kTriSmallThreads, SH / 32, SH % 32>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes))
                    : e),
   ...);
  return e;
}

hipError_t tri_launch_compute(sp_tri_ctx *ctx) {
  TriDev &F = ctx->fd;
  F.cut = ctx->cutoff;
  {
    TriProf ps(ctx, 1);
    hipLaunchKernelGGL(k_tri_dsigma, dim3(F.nw), dim3(64), 0, ctx->stream, F);
  }
  {
    TriProf ps(ctx, 2);
    hipLaunchKernelGGL(k_tri_weights, dim3(F.nw + 1), dim3(64), F.nw * sizeof(double), ctx->stream, F, 0);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  TriProf ps(ctx, 0);
  const long long rounds = (F.nf + 8 * kTriChunk - 1) / (8 * kTriChunk);
  const dim3 grid((unsigned)(rounds * 8 * kTriChunk));
  if (F.nt == kTriSmallThreads)
    launch_tri_small(F, grid, ctx->lds, ctx->stream, TriShapes{});
  else
    hipLaunchKernelGGL((k_tri_dgamma<kTriThreads, 0, 0>), grid, dim3(kTriThreads), ctx->lds, ctx->stream, F);
  return hipGetLastError();
}

}  // namespace

extern "C" {

int sp_tri_ctx_footprint(const sp_tri_problem *p, long long *bytes) {
  if (!p || !bytes || p->nw < 2 || p->L < 1) return SP_EARG;
  const long long nf = (long long)p->nw * p->nw * (p->nw + 1) / 2;
  const long long emax = 2LL * p->nw + 8;
  const long long O = p->ovl_off ? p->ovl_off[p->L] : 0;
  long long b = 2 * nf * 16 * p->L * 8;                                       // Gamma, dGamma
  b += 3LL * p->nw * 8;                                                       // Sigma, dSigma, mesh
  b += 2 * (long long)p->nw * emax * 8 + 2 * emax * 8 + (p->nw + 1) * 4;      // weight tables
  b += emax * 10 * p->L * 8 + 10LL * p->B * p->Rb * 8 + 16;                   // correlation scratch, output, flags
  b += 4LL * (10 * p->L + 1 + 5 * O + 2 * p->R + 4 * p->B * p->Rb);           // lattice tables
  *bytes = b;
  return SP_OK;
}

int sp_tri_ctx_create(const sp_tri_problem *p, sp_tri_ctx **out) {
  if (!out) return SP_EARG;
  *out = nullptr;
  sp_tri_ctx *ctx = new (std::nothrow) sp_tri_ctx();
  if (!ctx) return SP_ENOMEM;
  *out = ctx;
  int rc = check_tri_problem(ctx, p);
  if (rc != SP_OK) return rc;
  SP_THIP(hipGetDevice(&ctx->device));
  SP_THIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  TriDev &f = ctx->fd;
  f.nw = p->nw;
  f.L = p->L;
  f.R = p->R;
  f.B = p->B;
  f.Rb = p->Rb;
  f.nf = (long long)p->nw * p->nw * (p->nw + 1) / 2;
  f.nv = f.nf * 16 * p->L;
  f.emax = 2 * p->nw + 8;
  f.nt = p->L <= kTriSmallL ? kTriSmallThreads : kTriThreads;
  int nb = f.nt / (2 * p->L);
  if (nb > kTriMaxPoints) nb = kTriMaxPoints;
  if (nb < 1) nb = 1;
  f.nb = nb;
  f.cs = 0;
  if (f.nt == kTriSmallThreads) {
    const int cs = (f.L + 3) / 4 * 4;
    if (tri_shape_compiled(cs, f.nb, TriShapes{}) && tri_lds(f.nw, f.L, cs, f.nb).total <= kTriLdsBudget) f.cs = cs;
  }
  ctx->lds = (size_t)tri_lds(f.nw, f.L, f.cs, f.nb).total;
  int max_lds = 0;
  SP_THIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
  if (ctx->lds > (size_t)max_lds) return tfail(ctx, SP_EARG, "problem too large for the LDS plan of the dGamma kernel");
  if (f.nv * 8 >= (1LL << 31)) return tfail(ctx, SP_EARG, "vertex too large for 32-bit byte offsets");
  f.g_bytes = (unsigned)(f.nv * 8);
  const int O = p->ovl_off[p->L];
  const size_t nv = (size_t)f.nv;
  const int emax = f.emax;
  const size_t nc = (size_t)f.B * f.Rb;
  double *mesh = nullptr;
  // lattice-bubble tables: per site, the overlap entries stably ordered by
  // their spin transformations (P1, P2) and cut into groups of equal (P1, P2),
  // the groups of a site by decreasing size (the lanes of a wave then run
  // their longest groups side by side).  An entry (rid1, rid2) that occurs m
  // times in a group is evaluated once, in a group of its own multiplicity m
  // (bgrp .w), whose sum the kernel scales by m: kagome-DM r5 has 31 such
  // repeats among 771 entries, in 3 of 76 groups
  std::vector<int> gofs(f.L + 1, 0), ent, order(f.L);
  std::vector<int4> grp;
  ent.reserve(O);
  for (int r = 0; r < f.L; ++r) {
    std::vector<std::pair<int, int>> keyed;  // (P1 code | P2 code << 6, entry)
    for (int i = p->ovl_off[r]; i < p->ovl_off[r + 1]; ++i) {
      const int *q1 = p->ovl_perm1 + 3 * i, *q2 = p->ovl_perm2 + 3 * i;
      keyed.push_back({(q1[0] | q1[1] << 2 | q1[2] << 4) | (q2[0] | q2[1] << 2 | q2[2] << 4) << 6, i});
    }
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
    for (size_t k0 = 0; k0 < keyed.size();) {
      size_t k1 = k0;
      while (k1 < keyed.size() && keyed[k1].first == keyed[k0].first) ++k1;
      // entries of the (P1, P2) group [k0, k1) with their multiplicities, in
      // order of first occurrence
      const int i0 = keyed[k0].second;
      std::vector<std::pair<int, int>> uniq;  // (packed sites, multiplicity)
      for (size_t k = k0; k < k1; ++k) {
        const int i = keyed[k].second;
        const int pk = p->ovl_rid1[i] | p->ovl_rid2[i] << 16;
        size_t u = 0;
        while (u < uniq.size() && uniq[u].first != pk) ++u;
        if (u == uniq.size()) uniq.push_back({pk, 1});
        else ++uniq[u].second;
      }
      const int *q1 = p->ovl_perm1 + 3 * i0, *q2 = p->ovl_perm2 + 3 * i0;
      int inv1[3];
      for (int c = 0; c < 3; ++c) inv1[q1[c]] = c;
      const int Q[3] = {q2[inv1[0]], q2[inv1[1]], q2[inv1[2]]};  // Q = P2 P1^-1
      const int code = (q1[0] | q1[1] << 2 | q1[2] << 4) | (Q[0] | Q[1] << 2 | Q[2] << 4) << 6 |
                       (q2[0] | q2[1] << 2 | q2[2] << 4) << 12;
      int mmax = 0;
      for (const auto &e : uniq) mmax = e.second > mmax ? e.second : mmax;
      for (int m = 1; m <= mmax; ++m) {
        const int b = (int)ent.size();
        for (const auto &e : uniq)
          if (e.second == m) ent.push_back(e.first);
        if ((int)ent.size() > b) grp.push_back(make_int4(b, (int)ent.size(), code, m));
      }
      k0 = k1;
    }
    gofs[r + 1] = (int)grp.size();
    std::stable_sort(grp.begin() + gofs[r], grp.end(),
                     [](const int4 &a, const int4 &b) { return a.y - a.x > b.y - b.x; });
  }
  for (int r = 0; r < f.L; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return p->ovl_off[a + 1] - p->ovl_off[a] > p->ovl_off[b + 1] - p->ovl_off[b];
  });
  int *site = nullptr, *sitep = nullptr, *inv = nullptr, *invp = nullptr, *goff = nullptr, *bent = nullptr,
      *rf = nullptr, *ri = nullptr, *corr = nullptr, *corrp = nullptr, *sorder = nullptr;
  int4 *bgrp = nullptr;
  if ((rc = talloc(ctx, &mesh, f.nw)) || (rc = talloc(ctx, &site, f.L)) || (rc = talloc(ctx, &sitep, 3 * f.L)) ||
      (rc = talloc(ctx, &inv, f.L)) || (rc = talloc(ctx, &invp, 3 * f.L)) || (rc = talloc(ctx, &goff, f.L + 1)) ||
      (rc = talloc(ctx, &sorder, f.L)) ||
      (rc = talloc(ctx, &bent, ent.size())) || (rc = talloc(ctx, &bgrp, grp.size())) ||
      (rc = talloc(ctx, &rf, f.R)) || (rc = talloc(ctx, &ri, f.R)) ||
      (rc = talloc(ctx, &corr, nc)) || (rc = talloc(ctx, &corrp, 3 * nc)) || (rc = talloc(ctx, &f.G, nv)) ||
      (rc = talloc(ctx, &f.dG, nv)) || (rc = talloc(ctx, &f.v2, f.nw)) || (rc = talloc(ctx, &f.dv2, f.nw)) ||
      (rc = talloc(ctx, &f.Wwp, (size_t)f.nw * emax)) || (rc = talloc(ctx, &f.Wc, (size_t)f.nw * emax)) ||
      (rc = talloc(ctx, &f.Wn, f.nw)) || (rc = talloc(ctx, &f.Cwp, emax)) || (rc = talloc(ctx, &f.Cw, emax)) ||
      (rc = talloc(ctx, &f.Cn, 1)) || (rc = talloc(ctx, &f.corr_partial, (size_t)emax * 10 * f.L)) ||
      (rc = talloc(ctx, &f.corr_out, 10 * nc)) || (rc = talloc(ctx, &f.flags, 4)))
    return rc;
  f.mesh = mesh;
  f.site_rid = site;
  f.site_perm = sitep;
  f.inv_rid = inv;
  f.inv_perm = invp;
  f.bgrp_off = goff;
  f.bent = bent;
  f.site_order = sorder;
  f.bgrp = bgrp;
  f.range_fwd = rf;
  f.range_inv = ri;
  f.corr_rid = corr;
  f.corr_perm = corrp;
  std::vector<double> m(f.nw);
  for (int i = 0; i < f.nw; ++i) m[i] = (double)p->mesh[i];
  geo_mesh_params(m.data(), f.nw, f.geo, f.ga, f.gb);
  SP_THIP(hipMemcpy(mesh, m.data(), f.nw * 8, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(site, p->site_rid, f.L * 4, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(sitep, p->site_perm, 3 * f.L * 4, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(inv, p->inv_rid, f.L * 4, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(invp, p->inv_perm, 3 * f.L * 4, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(goff, gofs.data(), (f.L + 1) * 4, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(sorder, order.data(), f.L * 4, hipMemcpyHostToDevice));
  if (!ent.empty()) {
    SP_THIP(hipMemcpy(bent, ent.data(), ent.size() * 4, hipMemcpyHostToDevice));
    SP_THIP(hipMemcpy(bgrp, grp.data(), grp.size() * sizeof(int4), hipMemcpyHostToDevice));
  }
  SP_THIP(hipMemcpy(rf, p->range_fwd, (size_t)f.R * 4, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(ri, p->range_inv, (size_t)f.R * 4, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(corr, p->corr_rid, nc * 4, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(corrp, p->corr_perm, 3 * nc * 4, hipMemcpyHostToDevice));
  SP_THIP(hipMemset(f.G, 0, nv * 8));
  SP_THIP(hipMemset(f.dG, 0, nv * 8));
  SP_THIP(hipMemset(f.v2, 0, f.nw * 8));
  SP_THIP(hipMemset(f.dv2, 0, f.nw * 8));
  SP_THIP(hipMemset(f.flags, 0, 16));
  SP_THIP(hipMemset(f.Wn, 0, f.nw * 4));
  SP_THIP(hipMemset(f.Cn, 0, 4));
  SP_THIP(hipFuncSetAttribute((const void *)k_tri_dgamma<kTriThreads, 0, 0>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  SP_THIP(tri_allow_lds(max_lds, TriShapes{}));
  return SP_OK;
}

int sp_tri_ctx_destroy(sp_tri_ctx *ctx) {
  if (!ctx) return SP_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (int k = 0; k < 4; ++k)
    for (auto &pr : ctx->pending[k]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
  for (void *q : ctx->allocs) (void)hipFree(q);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SP_OK;
}

int sp_tri_set_state(sp_tri_ctx *ctx, double cutoff, const double *v2, const double *v4) {
  if (!ctx || !v2 || !v4) return SP_EARG;
  if (!std::isfinite(cutoff) || !(cutoff > 0.0)) return tfail(ctx, SP_EARG, "cutoff must be finite and positive");
  SP_THIP(hipStreamSynchronize(ctx->stream));
  SP_THIP(hipMemcpy(ctx->fd.v2, v2, ctx->fd.nw * 8, hipMemcpyHostToDevice));
  SP_THIP(hipMemcpy(ctx->fd.G, v4, (size_t)ctx->fd.nv * 8, hipMemcpyHostToDevice));
  SP_THIP(hipMemset(ctx->fd.flags, 0, 16));
  ctx->cutoff = cutoff;
  ctx->have_flow = false;
  return SP_OK;
}

int sp_tri_get_state(sp_tri_ctx *ctx, double *cutoff, double *v2, double *v4) {
  if (!ctx) return SP_EARG;
  SP_THIP(hipStreamSynchronize(ctx->stream));
  if (cutoff) *cutoff = ctx->cutoff;
  if (v2) SP_THIP(hipMemcpy(v2, ctx->fd.v2, ctx->fd.nw * 8, hipMemcpyDeviceToHost));
  if (v4) SP_THIP(hipMemcpy(v4, ctx->fd.G, (size_t)ctx->fd.nv * 8, hipMemcpyDeviceToHost));
  return SP_OK;
}

int sp_tri_get_flow(sp_tri_ctx *ctx, double *dv2, double *dv4) {
  if (!ctx) return SP_EARG;
  if (!ctx->have_flow) return tfail(ctx, SP_ESTATE, "no flow computed at the current cutoff");
  SP_THIP(hipStreamSynchronize(ctx->stream));
  if (dv2) SP_THIP(hipMemcpy(dv2, ctx->fd.dv2, ctx->fd.nw * 8, hipMemcpyDeviceToHost));
  if (dv4) SP_THIP(hipMemcpy(dv4, ctx->fd.dG, (size_t)ctx->fd.nv * 8, hipMemcpyDeviceToHost));
  return SP_OK;
}

int sp_tri_compute_step(sp_tri_ctx *ctx) {
  if (!ctx) return SP_EARG;
  SP_THIP(hipMemsetAsync(ctx->fd.flags, 0, 16, ctx->stream));
  SP_THIP(tri_launch_compute(ctx));
  ctx->have_flow = true;
  return SP_OK;
}

int sp_tri_finalize_step(sp_tri_ctx *ctx, double new_cutoff) {
  if (!ctx) return SP_EARG;
  if (!std::isfinite(new_cutoff) || !(new_cutoff > 0.0)) return tfail(ctx, SP_EARG, "cutoff must be finite and positive");
  if (!ctx->have_flow) return tfail(ctx, SP_ESTATE, "finalize_step without a computed flow");
  TriDev F = ctx->fd;
  F.euler = 1;
  F.step = new_cutoff - ctx->cutoff;
  long long blocks = (F.nv + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  {
    TriProf ps(ctx, 3);
    hipLaunchKernelGGL(k_tri_euler, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, F);
  }
  SP_THIP(hipGetLastError());
  ctx->cutoff = new_cutoff;
  ctx->have_flow = false;
  return SP_OK;
}

int sp_tri_is_diverged(sp_tri_ctx *ctx, int *diverged) {
  if (!ctx || !diverged) return SP_EARG;
  if (!ctx->have_flow) return tfail(ctx, SP_ESTATE, "no flow computed at the current cutoff");
  int fl[2] = {0, 0};
  SP_THIP(hipMemcpyAsync(fl, ctx->fd.flags, 8, hipMemcpyDeviceToHost, ctx->stream));
  SP_THIP(hipStreamSynchronize(ctx->stream));
  if (fl[1]) return tfail(ctx, SP_EARG, "evaluation list overflow");
  *diverged = fl[0] ? 1 : 0;
  return SP_OK;
}

int sp_tri_measure_correlation(sp_tri_ctx *ctx, double *out) {
  if (!ctx || !out) return SP_EARG;
  TriDev F = ctx->fd;
  F.cut = ctx->cutoff;
  hipLaunchKernelGGL(k_tri_weights, dim3(1), dim3(64), F.nw * sizeof(double), ctx->stream, F, 1);
  hipLaunchKernelGGL(k_tri_corr_partial, dim3(F.emax), dim3(kTriThreads), 0, ctx->stream, F);
  hipLaunchKernelGGL(k_tri_corr_final, dim3(1), dim3(256), 0, ctx->stream, F);
  SP_THIP(hipGetLastError());
  const size_t n = 10 * (size_t)F.B * F.Rb;
  SP_THIP(hipMemcpyAsync(out, F.corr_out, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  SP_THIP(hipStreamSynchronize(ctx->stream));
  return SP_OK;
}

int sp_tri_synchronize(sp_tri_ctx *ctx) {
  if (!ctx) return SP_EARG;
  SP_THIP(hipStreamSynchronize(ctx->stream));
  return SP_OK;
}

const char *sp_tri_last_error(const sp_tri_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sp_tri_profile_enable(sp_tri_ctx *ctx, int enable) {
  if (!ctx) return SP_EARG;
  ctx->prof = enable != 0;
  return SP_OK;
}

int sp_tri_profile_get(sp_tri_ctx *ctx, int kernel, double *total_ms, long *launches) {
  if (!ctx || kernel < 0 || kernel > 3 || !total_ms || !launches) return SP_EARG;
  SP_THIP(hipStreamSynchronize(ctx->stream));
  for (auto &pr : ctx->pending[kernel]) {
    float ms = 0.f;
    SP_THIP(hipEventElapsedTime(&ms, pr.first, pr.second));
    ctx->prof_ms[kernel] += ms;
    ctx->prof_n[kernel] += 1;
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  ctx->pending[kernel].clear();
  *total_ms = ctx->prof_ms[kernel];
  *launches = ctx->prof_n[kernel];
  return SP_OK;
}

}  // extern "C"
