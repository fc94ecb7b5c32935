// xyz_flow.hip -- HIP kernels and C ABI of the XYZ pf-FRG core (diagonal
// two-spin couplings J_xx, J_yy, J_zz; four vertex components X, Y, Z, D) for
// MI355X (gfx950), FP64.
//
// Reference behaviour followed (paths relative to the reference repository):
//   src/XYZ/XYZVertexTwoParticle.hpp:63-77,188-653  storage (four planes), access buffers,
//                                                   density sign of the s <-> u exchange,
//                                                   site spin permutations
//   src/XYZ/XYZFrgCore.cpp:128-193                  finalizeStep, dSigma
//   src/XYZ/XYZFrgCore.cpp:195-571                  dGamma: S, T (lattice bubble, chalice,
//                                                   inverse chalice), U, conventional, Katanin
//   src/XYZ/XYZMeasurementCorrelation.cpp:98-212    static susceptibility, 4 outputs
//
// Device memory (per flow):
//   Gamma, dGamma  FP64, components interleaved: the element of (row, site) is
//                  {X, Y, Z, D} at ((row * L) + rid) * 4, row = (s(s+1)/2 + u) * N_w + t,
//                  so one support of one site is two 16-byte loads.  The C ABI
//                  converts from / to the reference's four planes
//                  (_dataXX ... _dataDD) on upload and download.
//   Sigma, dSigma  FP64 [N_w]
//   weight tables  per exact index x: evaluation points and weights at the cutoff
//
// dGamma kernel (k_xyz_dgamma): one workgroup of 384 threads per frequency
// triple (s, t, u), XCD-chunked like the SU(2) kernel.  Waves 0..4 are
// workers, (slot, rid) pairs; wave 5 is the producer.  The evaluation points
// of the three channels run as one sequence of chunks:
//   A  (producer, one chunk ahead, double-buffered LDS) one lane per (point,
//      pair of access buffers): the two bracket searches the pair shares,
//      descriptors (support byte offsets, weights, density signs, exchange);
//      the T channel's four site-local vertices, stored as chalice factors;
//   B  (workers) per (slot, rid): 4 buffers x 4 supports x 2 loads, the site's
//      spin permutation, the S / U ladder or the chalice terms; in the T
//      channel the bundles of the first buffer pair go to LDS;
//   C  (T channel, all waves) the lattice bubble as rank updates
//      M[a][b] += sum_e c_e A0_e[a] (x) A1_e[b] for every spin-component pair
//      (a, b) the overlap transformations use (4x4 register tiles over LDS),
//      contracted with the overlap lists once per item.
// When the bubble matrices do not fit the LDS of a workgroup, k_xyz_dgamma_direct
// evaluates the bubble per point from the overlap lists instead.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "frg_common.h"
#include "spfrg.h"

namespace {

using namespace frg;

constexpr int kXyzThreads = 256;
constexpr int kXyzMaxPoints = 16;  // evaluation points per chunk of the direct kernel
// launch shape of the producer/worker dGamma kernel (compile-time tunables;
// the defaults are the tuned ones): threads per workgroup, evaluation points
// per worker slot and chunk, workgroups per CU the register budget is set for
#ifndef SPFRG_XYZ_THREADS
#define SPFRG_XYZ_THREADS 384
#endif
#ifndef SPFRG_XYZ_PER
#define SPFRG_XYZ_PER 1
#endif
#ifndef SPFRG_XYZ_MIN_BLOCKS
#define SPFRG_XYZ_MIN_BLOCKS 2
#endif
constexpr int kXyzPW = SPFRG_XYZ_THREADS;
constexpr int kXyzPer = SPFRG_XYZ_PER;
constexpr long long kXyzChunk = 16;
constexpr int kX = 0, kY = 1, kZ = 2, kD = 3;
constexpr int kMaxPairs = 10;  // 9 spin-component pairs + density

// ---------------------------------------------------------------------------
// per-flow descriptor (passed by value)
// ---------------------------------------------------------------------------
struct XyzDev {
  double *G, *dG, *v2, *dv2;
  const double *mesh;
  const int *site_rid, *site_perm, *inv_rid, *inv_perm;
  const int *ovl_off, *ovl_rid1, *ovl_rid2, *ovl_perm1, *ovl_perm2;
  const int4 *ovl_midx;  // per overlap entry: bubble-matrix index of X, Y, Z, D
  const int *range_fwd, *range_inv, *corr_rid, *corr_perm;
  double *Wwp, *Wc;
  int *Wn;
  double *Cwp, *Cw;
  int *Cn;
  double *corr_partial, *corr_out;
  int *flags;  // [0] NaN in the flow, [1] evaluation-list overflow
  int nw, L, R, B, Rb, emax, nb, slots, euler;
  int npairs;
  int pair[kMaxPairs];  // 4 a + b of each bubble matrix
  long long nf;
  long long np;  // N_f * L
  unsigned g_bytes;
  double cut, step;
  int geo;  // geometric mesh: bracket guess from the logarithm (count_le_geo)
  float ga, gb;
};

__device__ __forceinline__ double gload(__amdgpu_buffer_rsrc_t r, int byte_off) {
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0);
  double d;
  __builtin_memcpy(&d, &v, 8);
  return d;
}

// weight tables of one flow (the SU(2) core's definition)
__global__ __launch_bounds__(64) void k_xyz_weights(XyzDev F, int corr_only) {
  extern __shared__ double s_wmesh[];
  const int x = corr_only ? F.nw : (int)blockIdx.x;
  weight_tables_wave(x, F.mesh, F.nw, F.cut, F.v2, F.dv2, F.Wwp, F.Wc, F.Wn, F.Cwp, F.Cw, F.Cn, F.emax, F.flags + 1,
                     s_wmesh);
}

// getValue(i1, i2, s, t, u, c, FrequencyChannel::None) (XYZVertexTwoParticle.hpp:188-310)
// for pairs whose transformation keeps the component (the density, or the
// pair (0, 0)); rid_fwd / rid_inv = symmetryTransform(i1, i2) / (i2, i1)
__device__ double xyz_value_none(const double *G, const double *m, int nw, int L, int rid_fwd, int rid_inv, double s,
                                 double t, double u, int c) {
  bool exch = false;
  if (s < 0 && u < 0) { s = -s; u = -u; }
  else if (s < 0) { s = -s; exch = true; }
  else if (u < 0) { u = -u; exch = true; }
  if (t < 0) t = -t;
  const int rid = exch ? rid_inv : rid_fwd;
  int lS, uS, lT, uT, lU, uU;
  double bS, bT, bU;
  interp(m, nw, s, lS, uS, bS);
  interp(m, nw, t, lT, uT, bT);
  interp(m, nw, u, lU, uU, bU);
  const double *P = G + (size_t)rid * 4 + c;
  const size_t rs = (size_t)L * 4;
  auto A = [&](int so, int to, int uo) {
    if (so < uo) {
      const double v = P[((size_t)tri(uo, so) * nw + to) * rs];
      return c == kD ? -v : v;
    }
    return P[((size_t)tri(so, uo) * nw + to) * rs];
  };
  return (1 - bU) * ((1 - bT) * ((1 - bS) * A(lS, lT, lU) + bS * A(uS, lT, lU)) +
                     bT * ((1 - bS) * A(lS, uT, lU) + bS * A(uS, uT, lU))) +
         bU * ((1 - bT) * ((1 - bS) * A(lS, lT, uU) + bS * A(uS, lT, uU)) +
               bT * ((1 - bS) * A(lS, uT, uU) + bS * A(uS, uT, uU)));
}

// dSigma/dLambda (src/XYZ/XYZFrgCore.cpp:164-193): one wavefront per mesh point
__global__ __launch_bounds__(64) void k_xyz_dsigma(XyzDev F) {
  const int i = blockIdx.x;
  if (i >= F.nw) return;
  const int lane = threadIdx.x;
  const int nw = F.nw, L = F.L;
  const double *m = F.mesh;
  const double cut = F.cut;
  const double w = m[i];
  double sum = 0.0;
  for (int j = lane; j < F.R; j += 64) {
    sum += xyz_value_none(F.G, m, nw, L, F.range_fwd[j], F.range_inv[j], w + cut, 0.0, w - cut, kD) -
           xyz_value_none(F.G, m, nw, L, F.range_fwd[j], F.range_inv[j], w - cut, 0.0, w + cut, kD);
  }
  sum = wave_sum(sum);
  if (lane == 0) {
    double v = -2.0 * sum;
    for (int k = 0; k < 4; ++k)
      v += xyz_value_none(F.G, m, nw, L, 0, 0, w + cut, w - cut, 0.0, k) -
           xyz_value_none(F.G, m, nw, L, 0, 0, w - cut, w + cut, 0.0, k);
    v /= (kTwoPi * (cut + sigma_at(m, nw, F.v2, cut)));
    F.dv2[i] = v;
    if (isnan(v)) atomicOr(&F.flags[0], 1);
  }
}

// chalice / inverse chalice factor of spin output c: the two other spin
// components of the local vertex enter with +, the component itself and the
// density with - (XYZFrgCore.cpp:333-416)
__device__ __forceinline__ double chal(const double *l, int c) {
  return (l[0] + l[1] + l[2]) - 2.0 * l[c] - l[kD];
}

// S (sx = +1) and U (sx = -1) ladders of one bundle pair (a pre-scaled by the
// point's weight) (XYZFrgCore.cpp:223-275, 419-470): spin output c gets
// sx (a_d b_c + a_c b_d) - (a_c1 b_c2 + a_c2 b_c1), the density sx sum_k a_k b_k
__device__ __forceinline__ void ladder(double (&acc)[4], const double *a, const double *b, double sx) {
  acc[kX] = fma(sx, fma(a[kD], b[kX], a[kX] * b[kD]), acc[kX]);
  acc[kX] = fma(-1.0, fma(a[kZ], b[kY], a[kY] * b[kZ]), acc[kX]);
  acc[kY] = fma(sx, fma(a[kD], b[kY], a[kY] * b[kD]), acc[kY]);
  acc[kY] = fma(-1.0, fma(a[kZ], b[kX], a[kX] * b[kZ]), acc[kY]);
  acc[kZ] = fma(sx, fma(a[kD], b[kZ], a[kZ] * b[kD]), acc[kZ]);
  acc[kZ] = fma(-1.0, fma(a[kX], b[kY], a[kY] * b[kX]), acc[kZ]);
  acc[kD] = fma(sx, fma(a[kD], b[kD], fma(a[kX], b[kX], fma(a[kY], b[kY], a[kZ] * b[kZ]))), acc[kD]);
}

// chalice (local lc with bundle a) and inverse chalice (local li with bundle
// b), both already scaled by the point's weight
__device__ __forceinline__ void chalices(double (&acc)[4], const double *a, const double *b, const double *lc,
                                         const double *li) {
#pragma unroll
  for (int k = 0; k < 3; ++k) acc[k] = fma(a[k], chal(lc, k), fma(b[k], chal(li, k), acc[k]));
  const double sc = (lc[kD] + lc[kX]) + (lc[kY] + lc[kZ]);
  const double si = (li[kD] + li[kX]) + (li[kY] + li[kZ]);
  acc[kD] = fma(-a[kD], sc, fma(-b[kD], si, acc[kD]));
}

__device__ __forceinline__ double pick3(double u0, double u1, double u2, int i) {
  return i == 0 ? u0 : (i == 1 ? u1 : u2);
}

// LDS carve-up of k_xyz_dgamma (bytes), identical on host and device.  Lp = L
// rounded up to a multiple of 4 (4x4 register tiles of the bubble matrices).  The per-chunk
// descriptors (c, off, ws, ex, loc) are double-buffered: the producer wave
// fills one buffer while the worker waves read the other.
struct XyzLds {
  int mesh, c, off, ws, ex, loc, dbuf, P, Q, M, rpa, red, total;  // buffer b at + b * dbuf
};
__host__ __device__ inline XyzLds xyz_lds(int nw, int L, int nb, int slots, int npairs) {
  XyzLds l;
  const int Lp = (L + 3) & ~3;
  int o = 0;
  auto take = [&](int bytes) { int r = o; o += (bytes + 15) & ~15; return r; };
  l.mesh = take(nw * 8);
  const int d0 = o;
  l.c = take(nb * 8);
  l.off = take(nb * 4 * 16);
  l.ws = take(nb * 4 * 4 * 8);
  l.ex = take(nb * 4 * 4);
  l.loc = take(nb * 16 * 8);
  l.dbuf = o - d0;
  o += l.dbuf;  // second descriptor buffer
  const int pq = 4 * nb * Lp * 8;
  const int red = slots * L * 4 * 8;
  l.P = take(pq > red ? pq : red);  // the final slot reduction reuses P
  l.red = l.P;
  l.Q = take(pq);
  l.M = take(npairs * Lp * Lp * 8);
  l.rpa = take(4 * L * 8);
  l.total = o;
  return l;
}

// dGamma kernel.  Waves 0 .. kXyzPW/64 - 2 are workers, (slot, rid) pairs with
// slots = floor((kXyzPW - 64) / L); the last wave is the producer.  The
// evaluation points of the three channels form one sequence of chunks of nb
// points; while the workers gather and evaluate chunk j, the producer builds
// the descriptors of chunk j + 1 (bracket searches, weights, the t channel's
// site-local vertices) into the other LDS buffer.
__global__ __launch_bounds__(kXyzPW, SPFRG_XYZ_MIN_BLOCKS) void k_xyz_dgamma(XyzDev F) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const long long total = F.nf;
  const long long j0 = blockIdx.x >> 3;
  const long long item = ((j0 / kXyzChunk) * 8 + (blockIdx.x & 7)) * kXyzChunk + (j0 % kXyzChunk);
  if (item >= total) return;

  const int nw = F.nw, L = F.L, nb = F.nb, slots = F.slots, npairs = F.npairs;
  const int Lp = (L + 3) & ~3;
  const XyzLds lay = xyz_lds(nw, L, nb, slots, npairs);
  double *s_mesh = reinterpret_cast<double *>(lds_raw + lay.mesh);
  double *s_P = reinterpret_cast<double *>(lds_raw + lay.P);
  double *s_Q = reinterpret_cast<double *>(lds_raw + lay.Q);
  double *s_M = reinterpret_cast<double *>(lds_raw + lay.M);
  double *s_rpa = reinterpret_cast<double *>(lds_raw + lay.rpa);
  double *s_red = reinterpret_cast<double *>(lds_raw + lay.red);

  const __amdgpu_buffer_rsrc_t rG = __builtin_amdgcn_make_buffer_rsrc((void *)F.G, (short)0, (int)F.g_bytes, 0x00020000);
  const int rowbytes = L * 32;
  const int tid = threadIdx.x;

  // iterator: item = tri(s,u) * nw + t
  const int su = (int)(item / nw);
  const int to = (int)(item - (long long)su * nw);
  int so = (int)((sqrt(8.0 * su + 1.0) - 1.0) * 0.5);
  while ((so + 1) * (so + 2) / 2 <= su) ++so;
  while (so * (so + 1) / 2 > su) --so;
  const int uo = su - so * (so + 1) / 2;

  for (int i = tid; i < nw; i += kXyzPW) s_mesh[i] = F.mesh[i];
  for (int i = tid; i < npairs * Lp * Lp; i += kXyzPW) s_M[i] = 0.0;
  for (int i = tid; i < 4 * nb * Lp; i += kXyzPW) { s_P[i] = 0.0; s_Q[i] = 0.0; }
  for (int i = tid; i < 4 * L; i += kXyzPW) s_rpa[i] = 0.0;
  __syncthreads();

  const int lane = tid & 63;
  const bool producer = tid >= kXyzPW - 64;
  const int slot = tid / L;
  const int rid = tid - slot * L;
  const bool active = slot < slots;  // slots * L <= kXyzPW - 64: workers only
  // byte offsets of this thread's site (plain / exchanged) and the packed spin
  // permutations (2 bits per component) of getSites / getInvertedSites
  int sr_fwd = 0, sr_inv = 0, pf = 0, pi = 0;
  if (active) {
    sr_fwd = F.site_rid[rid] * 32;
    sr_inv = F.inv_rid[rid] * 32;
    pf = F.site_perm[3 * rid] | (F.site_perm[3 * rid + 1] << 2) | (F.site_perm[3 * rid + 2] << 4);
    pi = F.inv_perm[3 * rid] | (F.inv_perm[3 * rid + 1] << 2) | (F.inv_perm[3 * rid + 2] << 4);
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};

  // chunk schedule: chunk j covers points [e0, e0 + nb) of channel ch
  const int nS = F.Wn[so], nT = F.Wn[to], nU = F.Wn[uo];
  const int cS = (nS + nb - 1) / nb, cT = (nT + nb - 1) / nb, cU = (nU + nb - 1) / nb;
  const int nchunks = cS + cT + cU;

  // iteration jc: the producer builds chunk jc + 1 while the workers evaluate
  // chunk jc (jc = -1: the first chunk's descriptors only)
  for (int jc = -1; jc < nchunks; ++jc) {
    const int cur = jc & 1;
    const int ch = jc < cS ? 0 : (jc < cS + cT ? 1 : 2);
    const int e0 = (jc - (ch == 0 ? 0 : (ch == 1 ? cS : cS + cT))) * nb;
    const int n = ch == 0 ? nS : (ch == 1 ? nT : nU);
    const int nvalid = n - e0 < nb ? n - e0 : nb;
    if (producer) {
      const int jj = jc + 1, bf = jj & 1;
      if (jj < nchunks) {
        const int ch = jj < cS ? 0 : (jj < cS + cT ? 1 : 2);
        const int e0 = (jj - (ch == 0 ? 0 : (ch == 1 ? cS : cS + cT))) * nb;
        const int xo = ch == 0 ? so : (ch == 1 ? to : uo);
        const int n = ch == 0 ? nS : (ch == 1 ? nT : nU);
        const int nvalid = n - e0 < nb ? n - e0 : nb;
        // (A) four tasks per point, one access buffer of a (point, pair) each
        // (in the t channel one gathered and one site-local buffer), with the
        // buffer's arguments picked arithmetically so that every lane runs the
        // same straight-line code (XYZFrgCore.cpp:226-230, 281-287, 325-331,
        // 423-427)
        const double s = s_mesh[so], t = s_mesh[to], u = s_mesh[uo];
        const double w1p = 0.5 * (s + t + u);
        const double w1 = 0.5 * (s - t + u);
        const double w2p = 0.5 * (s - t - u);
        const double w2 = 0.5 * (s + t - u);
        double *s_c = reinterpret_cast<double *>(lds_raw + lay.c + bf * lay.dbuf);
        int4 *s_off = reinterpret_cast<int4 *>(lds_raw + lay.off + bf * lay.dbuf);
        double *s_ws = reinterpret_cast<double *>(lds_raw + lay.ws + bf * lay.dbuf);
        int *s_ex = reinterpret_cast<int *>(lds_raw + lay.ex + bf * lay.dbuf);
        double *s_loc = reinterpret_cast<double *>(lds_raw + lay.loc + bf * lay.dbuf);
        for (int task = lane; task < 4 * nvalid; task += 64) {
          const int e = task >> 2, pr = (task >> 1) & 1, h = task & 1;
          const double wp = F.Wwp[(size_t)xo * F.emax + e0 + e];
          const double c = F.Wc[(size_t)xo * F.emax + e0 + e];
          // buffer h of pair pr: |bracket arguments| x0 / x1 of the pair, and
          // whether exactly one of the buffer's signed s / u arguments is
          // negative (site exchange); buffer 1 takes the interpolations in
          // the other order
          double x0, x1;
          int exb;
          if (ch == 0) {
            x0 = fabs(pr ? w1p + wp : w1 + wp);
            x1 = fabs(pr ? w2p + wp : w2 + wp);
            const double cc = h ? (pr ? w1p + wp : w1 + wp) : (pr ? -w2p - wp : -w2 - wp);
            exb = (s < 0) != (cc < 0);
          } else if (ch == 1) {
            const double A = pr ? w2p - wp : w1 - wp, B = pr ? w2 + wp : w1p + wp;
            x0 = fabs(A);
            x1 = fabs(B);
            const double cc = h ? (pr ? wp - w2p : w1 - wp) : (pr ? -w2 - wp : w1p + wp);
            exb = ((h ? B : A) < 0) != (cc < 0);
          } else {
            x0 = fabs(pr ? w1p + wp : w1 + wp);
            x1 = fabs(pr ? w2 - wp : w2p - wp);
            const double a = h ? (pr ? w2 - wp : w2p - wp) : (pr ? w1p + wp : w1 + wp);
            exb = (a < 0) != (u < 0);
          }
          const double xa = h ? x1 : x0, xb = h ? x0 : x1;
          Interp ia, ib;
          if (F.geo) interp_pair_geo(s_mesh, nw, xa, xb, F.ga, F.gb, ia, ib);
          else interp_pair(s_mesh, nw, xa, xb, ia, ib);
          {
            const Desc4 d = assemble(ch, xo, ia, ib, nw, rowbytes);
            const int sk = e * 4 + 2 * h + pr;
            s_off[sk] = make_int4(d.off[0], d.off[1], d.off[2], d.off[3]);
#pragma unroll
            for (int r = 0; r < 4; ++r) s_ws[sk * 4 + r] = d.ws[r];
            // bit 0: site exchange; bits 1..4: density weight of support r negated
            s_ex[sk] = exb | (d.neg << 1);
          }
          if (ch == 1) {
            // site-local buffer (k = 5, 7 for pair 0; 4, 6 for pair 1) with
            // the same interpolations: getValueLocal
            // (XYZVertexTwoParticle.hpp:355-376), site 0, no exchange
            const Desc4 d = assemble(2, xo, ia, ib, nw, rowbytes);
            double l[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const double2 v01 = buf_load<double2>(rG, d.off[r]);
              const double2 v23 = buf_load<double2>(rG, d.off[r] + 16);
              l[kX] = fma(d.ws[r], v01.x, l[kX]);
              l[kY] = fma(d.ws[r], v01.y, l[kY]);
              l[kZ] = fma(d.ws[r], v23.x, l[kZ]);
              l[kD] = fma(d.wd[r], v23.y, l[kD]);
            }
            // chalice / inverse chalice factors of this local vertex, scaled
            // by the point's weight (XYZFrgCore.cpp:333-416): spin output k
            // gets the two other spin components with +, the component
            // itself and the density with -; the density output -sum
            double *fo = s_loc + e * 16 + ((pr ? 0 : 1) + 2 * h) * 4;
#pragma unroll
            for (int k = 0; k < 3; ++k) fo[k] = c * chal(l, k);
            fo[kD] = -(c * ((l[kD] + l[kX]) + (l[kY] + l[kZ])));
          }
          if (pr == 0 && h == 0) s_c[e] = c;
        }
      }
    } else if (active && jc >= 0) {
      // (B) gathers and channel algebra per (slot, rid)
      const double *s_c = reinterpret_cast<const double *>(lds_raw + lay.c + cur * lay.dbuf);
      const int4 *s_off = reinterpret_cast<const int4 *>(lds_raw + lay.off + cur * lay.dbuf);
      const double *s_ws = reinterpret_cast<const double *>(lds_raw + lay.ws + cur * lay.dbuf);
      const int *s_ex = reinterpret_cast<const int *>(lds_raw + lay.ex + cur * lay.dbuf);
      const double *s_loc = reinterpret_cast<const double *>(lds_raw + lay.loc + cur * lay.dbuf);
      for (int e = slot; e < nvalid; e += slots) {
        double A[4][4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int sk = e * 4 + k;
          const int exw = s_ex[sk];
          const bool ex = (exw & 1) != 0;
          const int sr = ex ? sr_inv : sr_fwd;
          const int pm = ex ? pi : pf;
          const int4 o = s_off[sk];
          const int oo[4] = {o.x, o.y, o.z, o.w};
          double u0 = 0.0, u1 = 0.0, u2 = 0.0, u3 = 0.0;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double2 v01 = buf_load<double2>(rG, oo[r] + sr);
            const double2 v23 = buf_load<double2>(rG, oo[r] + sr + 16);
            const double ws = s_ws[sk * 4 + r];
            const double wd = (exw >> (1 + r)) & 1 ? -ws : ws;
            u0 = fma(ws, v01.x, u0);
            u1 = fma(ws, v01.y, u1);
            u2 = fma(ws, v23.x, u2);
            u3 = fma(wd, v23.y, u3);
          }
          // bundle component c of the site is the stored component perm[c]
          // (getValueSuperbundle, XYZVertexTwoParticle.hpp:385-407)
          A[k][kX] = pick3(u0, u1, u2, pm & 3);
          A[k][kY] = pick3(u0, u1, u2, (pm >> 2) & 3);
          A[k][kZ] = pick3(u0, u1, u2, (pm >> 4) & 3);
          A[k][kD] = u3;
        }
        const double c = s_c[e];
        if (ch != 1) {
          double a0[4], a2[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) { a0[q] = c * A[0][q]; a2[q] = c * A[2][q]; }
          const double sx = ch == 0 ? 1.0 : -1.0;
          ladder(acc, a0, A[1], sx);
          ladder(acc, a2, A[3], sx);
        } else {
          // chalice with Cb / Cb2, inverse chalice with ICa / ICa2: bundle k
          // meets the factors of local vertex k
          const double *fl = s_loc + e * 16;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            acc[q] = fma(A[3][q], fl[12 + q], fma(A[2][q], fl[8 + q], fma(A[1][q], fl[4 + q], fma(A[0][q], fl[q], acc[q]))));
          // lattice-bubble factors of the first pair: P = c A0, Q = A1
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            s_P[(q * nb + e) * Lp + rid] = c * A[0][q];
            s_Q[(q * nb + e) * Lp + rid] = A[1][q];
          }
        }
      }
    }
    __syncthreads();
    if (jc >= 0 && ch == 1) {
      // (C) M[k] += sum_e P[a_k][e] (x) Q[b_k][e], 4x4 register tiles
      const int hq = Lp >> 2;
      const int tiles = npairs * hq * hq;
      for (int jt = tid; jt < tiles; jt += kXyzPW) {
        const int k = jt / (hq * hq);
        const int r = jt - k * hq * hq;
        const int a = 4 * (r / hq), b = 4 * (r - (r / hq) * hq);
        const int pa = F.pair[k] >> 2, pbq = F.pair[k] & 3;
        const double *P = s_P + pa * nb * Lp + a;
        const double *Q = s_Q + pbq * nb * Lp + b;
        double *M = s_M + k * Lp * Lp + a * Lp + b;
        double2 m[4][2];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          m[i][0] = *reinterpret_cast<const double2 *>(M + i * Lp);
          m[i][1] = *reinterpret_cast<const double2 *>(M + i * Lp + 2);
        }
        for (int e = 0; e < nvalid; ++e) {
          const double2 p01 = *reinterpret_cast<const double2 *>(P + e * Lp);
          const double2 p23 = *reinterpret_cast<const double2 *>(P + e * Lp + 2);
          const double2 q01 = *reinterpret_cast<const double2 *>(Q + e * Lp);
          const double2 q23 = *reinterpret_cast<const double2 *>(Q + e * Lp + 2);
          const double pv[4] = {p01.x, p01.y, p23.x, p23.y};
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            m[i][0].x = fma(pv[i], q01.x, m[i][0].x);
            m[i][0].y = fma(pv[i], q01.y, m[i][0].y);
            m[i][1].x = fma(pv[i], q23.x, m[i][1].x);
            m[i][1].y = fma(pv[i], q23.y, m[i][1].y);
          }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          *reinterpret_cast<double2 *>(M + i * Lp) = m[i][0];
          *reinterpret_cast<double2 *>(M + i * Lp + 2) = m[i][1];
        }
      }
      __syncthreads();
      if (jc == cS + cT - 1) {
        // lattice bubble (XYZFrgCore.cpp:295-322): RPA_c(rid) = sum over
        // getOverlap(rid) of M[pair of the entry's transforms of c][rid1][rid2], x 4
        const int wave = tid >> 6;
        for (int r = wave; r < L; r += kXyzPW / 64) {
          double sx = 0.0, sy = 0.0, sz = 0.0, sd = 0.0;
          const int b = F.ovl_off[r], e = F.ovl_off[r + 1];
          for (int i = b + lane; i < e; i += 64) {
            const int4 m = F.ovl_midx[i];
            sx += s_M[m.x];
            sy += s_M[m.y];
            sz += s_M[m.z];
            sd += s_M[m.w];
          }
          sx = wave_sum(sx);
          sy = wave_sum(sy);
          sz = wave_sum(sz);
          sd = wave_sum(sd);
          if (lane == 0) {
            s_rpa[4 * r + 0] = 4.0 * sx;
            s_rpa[4 * r + 1] = 4.0 * sy;
            s_rpa[4 * r + 2] = 4.0 * sz;
            s_rpa[4 * r + 3] = 4.0 * sd;
          }
        }
        __syncthreads();
      }
    }
  }

  // reduce the slots; prefactor 1/(2 pi) (XYZFrgCore.cpp:564-570)
  if (active) {
#pragma unroll
    for (int k = 0; k < 4; ++k) s_red[(slot * L + rid) * 4 + k] = acc[k];
  }
  __syncthreads();
  for (int q = tid; q < 4 * L; q += kXyzPW) {
    const int r = q >> 2, k = q & 3;
    double v = s_rpa[q];
    for (int sl = 0; sl < slots; ++sl) v += s_red[(sl * L + r) * 4 + k];
    v *= (1.0 / kTwoPi);
    F.dG[(size_t)item * 4 * L + q] = v;
    if (isnan(v)) atomicOr(&F.flags[0], 1);
  }
}

// ---------------------------------------------------------------------------
// direct-bubble dGamma kernel, for lattices whose bubble matrices exceed the
// LDS of a workgroup.  One thread per (point, access buffer) for the
// descriptors, one per (point, buffer, site) for the gathers into LDS, one per
// (point, site, pair) for the algebra; the bubble is summed per point from the
// overlap lists, the entries of a site split between its two pair threads.
// ---------------------------------------------------------------------------
struct XyzLdsDirect {
  int mesh, cf, ci, off, w, fl, c, loc, A, total;
};
__host__ __device__ inline XyzLdsDirect xyz_lds_direct(int nw, int L, int nb) {
  XyzLdsDirect l;
  int o = 0;
  auto take = [&](int bytes) { int r = o; o += (bytes + 15) & ~15; return r; };
  l.mesh = take(nw * 8);
  l.cf = take(4 * L * 4);
  l.ci = take(4 * L * 4);
  l.off = take(nb * 8 * 4 * 4);
  l.w = take(nb * 8 * 4 * 8);
  l.fl = take(nb * 8 * 4);
  l.c = take(nb * 8);
  l.loc = take(nb * 4 * 4 * 8);
  l.A = take(nb * 4 * 4 * L * 8);  // also holds the final reduction (nb * 2 * 4 * L)
  l.total = o;
  return l;
}

// frequency arguments of access buffer k in channel ch (0 S, 1 T, 2 U) and the
// channel of exactness (src/XYZ/XYZFrgCore.cpp:226-230, 281-287, 325-331, 423-427).
// T: k 0..3 the bundles, 4..7 the site-local vertices Cb, ICa, Cb2, ICa2.
__device__ __forceinline__ void xyz_args(int ch, int k, double s, double t, double u, double w1, double w1p, double w2,
                                         double w2p, double wp, double &a, double &b, double &c, int &chan) {
  if (ch == 0) {
    chan = 0;
    a = s;
    if (k == 0) { b = -w1 - wp; c = -w2 - wp; }
    else if (k == 1) { b = w1p + wp; c = -w2p - wp; }
    else if (k == 2) { b = w2 + wp; c = w1 + wp; }
    else { b = -w2p - wp; c = w1p + wp; }
  } else if (ch == 1) {
    if (k < 4) {
      chan = 1;
      b = t;
      if (k == 0) { a = w1 - wp; c = w1p + wp; }
      else if (k == 1) { a = w2p - wp; c = -w2 - wp; }
      else if (k == 2) { a = w1p + wp; c = w1 - wp; }
      else { a = w2 + wp; c = wp - w2p; }
    } else {
      chan = 2;
      if (k == 4) { a = w2p - wp; b = -w2 - wp; c = t; }
      else if (k == 5) { a = w1 - wp; b = -w1p - wp; c = -t; }
      else if (k == 6) { a = w2 + wp; b = wp - w2p; c = t; }
      else { a = w1p + wp; b = wp - w1; c = -t; }
    }
  } else {
    chan = 2;
    c = u;
    if (k == 0) { a = w1 + wp; b = wp - w2p; }
    else if (k == 1) { a = w1p + wp; b = w2 - wp; }
    else if (k == 2) { a = w2p - wp; b = -w1 - wp; }
    else { a = w2 - wp; b = w1p + wp; }
  }
}

__global__ __launch_bounds__(kXyzThreads) void k_xyz_dgamma_direct(XyzDev F) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const long long total = F.nf;
  const long long j0 = blockIdx.x >> 3;
  const long long item = ((j0 / kXyzChunk) * 8 + (blockIdx.x & 7)) * kXyzChunk + (j0 % kXyzChunk);
  if (item >= total) return;

  const int nw = F.nw, L = F.L, nb = F.nb;
  const XyzLdsDirect lay = xyz_lds_direct(nw, L, nb);
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

  const int su = (int)(item / nw);
  const int to = (int)(item - (long long)su * nw);
  int so = (int)((sqrt(8.0 * su + 1.0) - 1.0) * 0.5);
  while ((so + 1) * (so + 2) / 2 <= su) ++so;
  while (so * (so + 1) / 2 > su) --so;
  const int uo = su - so * (so + 1) / 2;

  for (int i = tid; i < nw; i += kXyzThreads) s_mesh[i] = F.mesh[i];
  // element offsets (site, component) for the plain and the exchanged bundle:
  // spin components through the site's permutation
  for (int q = tid; q < 4 * L; q += kXyzThreads) {
    const int c = q / L, j = q - (q / L) * L;
    const int pf = c < 3 ? F.site_perm[3 * j + c] : kD, pi = c < 3 ? F.inv_perm[3 * j + c] : kD;
    s_cf[q] = F.site_rid[j] * 4 + pf;
    s_ci[q] = F.inv_rid[j] * 4 + pi;
  }
  __syncthreads();

  const double s = s_mesh[so], t = s_mesh[to], u = s_mesh[uo];
  const double w1p = 0.5 * (s + t + u);
  const double w1 = 0.5 * (s - t + u);
  const double w2p = 0.5 * (s - t - u);
  const double w2 = 0.5 * (s + t - u);

  const int ce = tid / (2 * L);
  const int crem = tid - ce * 2 * L;
  const int cpair = crem / L;
  const int cj = crem - cpair * L;
  const bool cactive = ce < nb;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int row4 = 4 * L;  // elements per row

  for (int ch = 0; ch < 3; ++ch) {
    const int xo = ch == 0 ? so : (ch == 1 ? to : uo);
    const int n = F.Wn[xo];
    const double *Wwp = F.Wwp + (size_t)xo * F.emax;
    const double *Wc = F.Wc + (size_t)xo * F.emax;
    const int nk = ch == 1 ? 8 : 4;
    for (int e0 = 0; e0 < n; e0 += nb) {
      const int nvalid = n - e0 < nb ? n - e0 : nb;
      if (tid < nvalid * nk) {
        const int e = tid / nk, k = tid - (tid / nk) * nk;
        const double wp = Wwp[e0 + e];
        double a, b, c;
        int chan;
        xyz_args(ch, k, s, t, u, w1, w1p, w2, w2p, wp, a, b, c, chan);
        const bool exch = (a < 0) != (c < 0);
        a = fabs(a); b = fabs(b); c = fabs(c);
        int lo1, hi1, lo2, hi2;
        double b1, b2;
        if (chan == 0) { interp(s_mesh, nw, b, lo1, hi1, b1); interp(s_mesh, nw, c, lo2, hi2, b2); }
        else if (chan == 1) { interp(s_mesh, nw, a, lo1, hi1, b1); interp(s_mesh, nw, c, lo2, hi2, b2); }
        else { interp(s_mesh, nw, a, lo1, hi1, b1); interp(s_mesh, nw, b, lo2, hi2, b2); }
        int off[4];
        double w[4];
        int fl = exch ? 1 : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int l1 = (q & 1) ? hi1 : lo1, l2 = (q & 2) ? hi2 : lo2;
          w[q] = ((q & 2) ? b2 : (1 - b2)) * ((q & 1) ? b1 : (1 - b1));
          bool fx;
          int rw;
          if (chan == 0) rw = support_row(xo, l1, l2, nw, fx);
          else if (chan == 1) rw = support_row(l1, xo, l2, nw, fx);
          else rw = support_row(l1, l2, xo, nw, fx);
          off[q] = rw * row4;
          if (fx) fl |= 4 << q;
        }
        if (k < 4) {
          const int d = e * 4 + k;
#pragma unroll
          for (int q = 0; q < 4; ++q) { s_off[d * 4 + q] = off[q]; s_w[d * 4 + q] = w[q]; }
          s_fl[d] = fl;
        } else {
#pragma unroll
          for (int cc = 0; cc < 4; ++cc) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const double sg = (cc == kD && ((fl >> (2 + q)) & 1)) ? -w[q] : w[q];
              v = fma(sg, gload(rG, (off[q] + cc) * 8), v);
            }
            s_loc[(e * 4 + (k - 4)) * 4 + cc] = v;
          }
        }
        if (k == 0) s_c[e] = Wc[e0 + e];
      }
      __syncthreads();
      for (int q = tid; q < nvalid * 4 * L; q += kXyzThreads) {
        const int d = q / L, j = q - (q / L) * L;
        const int fl = s_fl[d];
        const int *cofs = (fl & 1) ? s_ci : s_cf;
        int off[4];
        double w[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) { off[r] = s_off[d * 4 + r]; w[r] = s_w[d * 4 + r]; }
        int co[4];
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) co[cc] = cofs[cc * L + j];
        double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double wd = ((fl >> (2 + r)) & 1) ? -w[r] : w[r];
#pragma unroll
          for (int cc = 0; cc < 4; ++cc) v[cc] = fma(cc == kD ? wd : w[r], gload(rG, (off[r] + co[cc]) * 8), v[cc]);
        }
        double *dst = s_A + (size_t)d * 4 * L + j;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) dst[cc * L] = v[cc];
      }
      __syncthreads();
      if (cactive && ce < nvalid) {
        const double cw = s_c[ce];
        const double *Xb = s_A + (size_t)(ce * 4 + 2 * cpair) * 4 * L;
        const double *Yb = s_A + (size_t)(ce * 4 + 2 * cpair + 1) * 4 * L;
        double a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) { a[k] = cw * Xb[k * L + cj]; b[k] = Yb[k * L + cj]; }
        if (ch != 1) {
          ladder(acc, a, b, ch == 0 ? 1.0 : -1.0);
        } else {
          const double *B0 = s_A + (size_t)(ce * 4) * 4 * L;
          const double *B1 = s_A + (size_t)(ce * 4 + 1) * 4 * L;
          const int ib = F.ovl_off[cj], ie = F.ovl_off[cj + 1];
          const int im = ib + (ie - ib) / 2;
          const int lo = cpair == 0 ? ib : im, hi = cpair == 0 ? im : ie;
          const double c4 = 4.0 * cw;
          double r[4] = {0.0, 0.0, 0.0, 0.0};
          for (int i = lo; i < hi; ++i) {
            const int r1 = F.ovl_rid1[i], r2 = F.ovl_rid2[i];
#pragma unroll
            for (int k = 0; k < 3; ++k)
              r[k] = fma(B0[F.ovl_perm1[3 * i + k] * L + r1], B1[F.ovl_perm2[3 * i + k] * L + r2], r[k]);
            r[kD] = fma(B0[kD * L + r1], B1[kD * L + r2], r[kD]);
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) acc[k] = fma(c4, r[k], acc[k]);
          const double *lc = s_loc + (ce * 4 + 2 * cpair) * 4;
          const double *li = s_loc + (ce * 4 + 2 * cpair + 1) * 4;
          double bw[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) bw[k] = cw * b[k];
          chalices(acc, a, bw, lc, li);
        }
      }
      __syncthreads();
    }
  }
  double *red = s_A;
  if (cactive) {
    const int g = ce * 2 + cpair;
#pragma unroll
    for (int k = 0; k < 4; ++k) red[((size_t)g * 4 + k) * L + cj] = acc[k];
  }
  __syncthreads();
  const int ng = nb * 2;
  for (int q = tid; q < 4 * L; q += kXyzThreads) {
    double v = 0.0;
    for (int g = 0; g < ng; ++g) v += red[(size_t)g * 4 * L + q];
    v *= (1.0 / kTwoPi);
    const int c = q / L, j = q - (q / L) * L;
    F.dG[((size_t)item * L + j) * 4 + c] = v;
    if (isnan(v)) atomicOr(&F.flags[0], 1);
  }
}

// in-place explicit Euler step (src/XYZ/XYZFrgCore.cpp:128-162); skipped when
// the flow holds a NaN (src/SpinParser.cpp:151-159)
__global__ void k_xyz_euler(XyzDev F) {
  if (!F.euler || F.flags[0] != 0) return;
  const double step = F.step;
  const size_t n = 4 * (size_t)F.np;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) F.G[i] = fma(step, F.dG[i], F.G[i]);
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < F.nw; i += blockDim.x) F.v2[i] = fma(step, F.dv2[i], F.v2[i]);
}

// ---------------------------------------------------------------------------
// static susceptibility (src/XYZ/XYZMeasurementCorrelation.cpp:98-212), nu = 0
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kXyzThreads) void k_xyz_corr_partial(XyzDev F) {
  __shared__ double red[kXyzThreads * 4];
  const int o = blockIdx.x;
  const int n = *F.Cn;
  if (o >= n) return;
  const int nw = F.nw, L = F.L;
  const double *m = F.mesh, *v2 = F.v2;
  const double *G = F.G;
  const int slots = kXyzThreads / L;
  const int tid = threadIdx.x;
  const int slot = tid / L, rid = tid - (tid / L) * L;
  const bool active = slot < slots;
  const double nu = 0.0;
  const double w = F.Cwp[o];
  const double gw = (w + sigma_at(m, nw, v2, w)) * (w + nu + sigma_at(m, nw, v2, w + nu));
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (active) {
    for (int e = slot; e < n; e += slots) {
      const double wp = F.Cwp[e];
      double r[4];
      // dumbbell: bundle of generateAccessBuffer(w + wp + nu, nu, w - wp)
      {
        double a = w + wp + nu, b = nu, c = w - wp;
        bool exch = false;
        if (a < 0 && c < 0) { a = -a; c = -c; }
        else if (a < 0) { a = -a; exch = true; }
        else if (c < 0) { c = -c; exch = true; }
        if (b < 0) b = -b;
        int lS, uS, lT, uT, lU, uU;
        double bS, bT, bU;
        interp(m, nw, a, lS, uS, bS);
        interp(m, nw, b, lT, uT, bT);
        interp(m, nw, c, lU, uU, bU);
        const int srid = exch ? F.inv_rid[rid] : F.site_rid[rid];
        const int *perm = exch ? F.inv_perm + 3 * rid : F.site_perm + 3 * rid;
        const int comp[4] = {perm[0], perm[1], perm[2], kD};
        double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int S = (i & 1) ? uS : lS, T = (i & 2) ? uT : lT, U = (i & 4) ? uU : lU;
          const double wi = ((i & 2) ? bT : (1 - bT)) * ((i & 1) ? bS : (1 - bS)) * ((i & 4) ? bU : (1 - bU));
          const bool fx = S < U;
          const double *el = G + (((size_t)(fx ? tri(U, S) : tri(S, U)) * nw + T) * L + srid) * 4;
#pragma unroll
          for (int k = 0; k < 4; ++k) v[k] = fma((k == kD && fx) ? -wi : wi, el[comp[k]], v[k]);
        }
        r[kX] = -v[kX];
        r[kY] = -v[kY];
        r[kZ] = -v[kZ];
        r[kD] = -4.0 * v[kD];
      }
      if (rid == 0) {
        // egg: getValue(0, 0, w + wp + nu, w - wp, nu, c, None)
        double vv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) vv[k] = xyz_value_none(G, m, nw, L, 0, 0, w + wp + nu, w - wp, nu, k);
        r[kX] += 0.5 * (vv[kX] - vv[kY] - vv[kZ] + vv[kD]);
        r[kY] += 0.5 * (-vv[kX] + vv[kY] - vv[kZ] + vv[kD]);
        r[kZ] += 0.5 * (-vv[kX] - vv[kY] + vv[kZ] + vv[kD]);
        r[kD] += 2.0 * (vv[kX] + vv[kY] + vv[kZ] + vv[kD]);
      }
      const double norm = 1.0 / (gw * (wp + sigma_at(m, nw, v2, wp)) * (wp + nu + sigma_at(m, nw, v2, wp + nu)) *
                                 (4.0 * kPi * kPi));
      const double cf = F.Cw[e] * norm;
#pragma unroll
      for (int g = 0; g < 4; ++g) acc[g] = fma(cf, r[g], acc[g]);
    }
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) red[g * kXyzThreads + tid] = active ? acc[g] : 0.0;
  __syncthreads();
  for (int q = tid; q < 4 * L; q += kXyzThreads) {
    const int g = q / L, r = q - (q / L) * L;
    double tsum = 0.0;
    for (int k = 0; k < slots; ++k) tsum += red[g * kXyzThreads + k * L + r];
    if (r == 0) {
      // bubble term (site 0): spin components 1/(4 pi), density 1/pi
      const double term1 = 1.0 / gw;
      tsum += g == kD ? term1 / kPi : term1 / (4.0 * kPi);
    }
    F.corr_partial[((size_t)o * 4 + g) * L + r] = F.Cw[o] * tsum;
  }
}

// sum over outer points; output groups XX, YY, ZZ, DD, each spin component of
// a (basis, site) pair through its spin permutation (XYZMeasurementCorrelation.cpp:186-211)
__global__ void k_xyz_corr_final(XyzDev F) {
  __shared__ double susc[4 * 64];
  const int n = *F.Cn;
  const int L = F.L;
  for (int q = threadIdx.x; q < 4 * L; q += blockDim.x) {
    double v = 0.0;
    for (int o = 0; o < n; ++o) v += F.corr_partial[(size_t)o * 4 * L + q];
    susc[q] = v;
  }
  __syncthreads();
  const int nout = F.B * F.Rb;
  for (int k = threadIdx.x; k < nout; k += blockDim.x) {
    const int rid = F.corr_rid[k];
#pragma unroll
    for (int g = 0; g < 3; ++g) F.corr_out[(size_t)g * nout + k] = susc[F.corr_perm[3 * k + g] * L + rid];
    F.corr_out[(size_t)3 * nout + k] = susc[kD * L + rid];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct sp_xyz_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  XyzDev fd{};
  double cutoff = 0.0;
  bool have_flow = false;
  bool direct = false;  // bubble summed per point (k_xyz_dgamma_direct)
  size_t lds = 0;
  std::vector<void *> allocs;
  std::vector<double> host;  // plane <-> interleaved staging of the state
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[4];
  double prof_ms[4] = {0, 0, 0, 0};
  long prof_n[4] = {0, 0, 0, 0};
  std::string err;
};

namespace {

int xfail(sp_xyz_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define SP_XHIP(call)                                                                                      \
  do {                                                                                                     \
    hipError_t e_ = (call);                                                                                \
    if (e_ != hipSuccess) return xfail(ctx, SP_EHIP, std::string(#call ": ") + hipGetErrorString(e_));   \
  } while (0)

template <class T> int xalloc(sp_xyz_ctx *ctx, T **p, size_t count) {
  void *q = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) return xfail(ctx, SP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return SP_OK;
}

struct XyzProf {
  sp_xyz_ctx *ctx;
  int k;
  hipEvent_t a = nullptr, b = nullptr;
  XyzProf(sp_xyz_ctx *c, int kernel) : ctx(c), k(kernel) {
    if (ctx->prof && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, ctx->stream);
  }
  ~XyzProf() {
    if (ctx->prof && a && b) {
      (void)hipEventRecord(b, ctx->stream);
      ctx->pending[k].push_back({a, b});
    }
  }
};

int check_xyz_problem(sp_xyz_ctx *ctx, const sp_xyz_problem *p) {
  if (!p || p->nw < 2 || p->nw > 4096 || !p->mesh) return xfail(ctx, SP_EARG, "invalid frequency mesh");
  for (int i = 0; i < p->nw; ++i)
    if (!(p->mesh[i] > 0.0f) || (i > 0 && !(p->mesh[i] > p->mesh[i - 1])))
      return xfail(ctx, SP_EARG, "mesh must be positive and strictly ascending");
  if (p->L < 1 || p->L > 64) return xfail(ctx, SP_EARG, "number of representative sites must be in [1, 64]");
  if (!p->site_rid || !p->site_perm || !p->inv_rid || !p->inv_perm || !p->ovl_off || !p->range_fwd ||
      !p->range_inv || !p->corr_rid || !p->corr_perm)
    return xfail(ctx, SP_EARG, "missing lattice table");
  if (p->R < 1 || p->B < 1 || p->Rb < 1) return xfail(ctx, SP_EARG, "invalid range sizes");
  if (p->ovl_off[0] != 0) return xfail(ctx, SP_EARG, "overlap offsets must start at 0");
  for (int r = 0; r < p->L; ++r)
    if (p->ovl_off[r + 1] < p->ovl_off[r]) return xfail(ctx, SP_EARG, "overlap offsets must be non-decreasing");
  const int O = p->ovl_off[p->L];
  if (O > 0 && (!p->ovl_rid1 || !p->ovl_rid2 || !p->ovl_perm1 || !p->ovl_perm2))
    return xfail(ctx, SP_EARG, "missing overlap lists");
  auto in_range = [&](const int *a, long n, int hi) {
    for (long i = 0; i < n; ++i)
      if (a[i] < 0 || a[i] >= hi) return false;
    return true;
  };
  const long nc = (long)p->B * p->Rb;
  if (!in_range(p->site_rid, p->L, p->L) || !in_range(p->inv_rid, p->L, p->L) || !in_range(p->ovl_rid1, O, p->L) ||
      !in_range(p->ovl_rid2, O, p->L) || !in_range(p->range_fwd, p->R, p->L) || !in_range(p->range_inv, p->R, p->L) ||
      !in_range(p->corr_rid, nc, p->L))
    return xfail(ctx, SP_EARG, "lattice table entry outside [0, L)");
  if (!in_range(p->site_perm, 3L * p->L, 3) || !in_range(p->inv_perm, 3L * p->L, 3) ||
      !in_range(p->ovl_perm1, 3L * O, 3) || !in_range(p->ovl_perm2, 3L * O, 3) || !in_range(p->corr_perm, 3 * nc, 3))
    return xfail(ctx, SP_EARG, "spin permutation entry outside [0, 3)");
  return SP_OK;
}

hipError_t xyz_launch_compute(sp_xyz_ctx *ctx) {
  XyzDev &F = ctx->fd;
  F.cut = ctx->cutoff;
  {
    XyzProf ps(ctx, 1);
    hipLaunchKernelGGL(k_xyz_dsigma, dim3(F.nw), dim3(64), 0, ctx->stream, F);
  }
  {
    XyzProf ps(ctx, 2);
    hipLaunchKernelGGL(k_xyz_weights, dim3(F.nw + 1), dim3(64), F.nw * sizeof(double), ctx->stream, F, 0);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  XyzProf ps(ctx, 0);
  const long long rounds = (F.nf + 8 * kXyzChunk - 1) / (8 * kXyzChunk);
  const dim3 grid((unsigned)(rounds * 8 * kXyzChunk));
  if (ctx->direct)
    hipLaunchKernelGGL(k_xyz_dgamma_direct, grid, dim3(kXyzThreads), ctx->lds, ctx->stream, F);
  else
    hipLaunchKernelGGL(k_xyz_dgamma, grid, dim3(kXyzPW), ctx->lds, ctx->stream, F);
  return hipGetLastError();
}

// the reference's four planes X, Y, Z, D <-> the interleaved device layout
void planes_to_interleaved(const double *planes, double *out, size_t np) {
  for (size_t i = 0; i < np; ++i)
    for (int c = 0; c < 4; ++c) out[4 * i + c] = planes[c * np + i];
}
void interleaved_to_planes(const double *in, double *planes, size_t np) {
  for (size_t i = 0; i < np; ++i)
    for (int c = 0; c < 4; ++c) planes[c * np + i] = in[4 * i + c];
}

}  // namespace

extern "C" {

int sp_xyz_ctx_footprint(const sp_xyz_problem *p, long long *bytes) {
  if (!p || !bytes || p->nw < 2 || p->L < 1) return SP_EARG;
  const long long nf = (long long)p->nw * p->nw * (p->nw + 1) / 2;
  const long long emax = 2LL * p->nw + 8;
  const long long O = p->ovl_off ? p->ovl_off[p->L] : 0;
  long long b = 2 * nf * 4 * p->L * 8;                                        // Gamma, dGamma
  b += 3LL * p->nw * 8;                                                       // Sigma, dSigma, mesh
  b += 2 * (long long)p->nw * emax * 8 + 2 * emax * 8 + (p->nw + 1) * 4;      // weight tables
  b += emax * 4 * p->L * 8 + 4LL * p->B * p->Rb * 8 + 16;                     // correlation scratch, output, flags
  b += 4LL * (8 * p->L + 1 + 8 * O + 2 * p->R + 4 * p->B * p->Rb);            // lattice tables
  b += 16LL * O;                                                              // bubble-matrix indices
  *bytes = b;
  return SP_OK;
}

int sp_xyz_ctx_create(const sp_xyz_problem *p, sp_xyz_ctx **out) {
  if (!out) return SP_EARG;
  *out = nullptr;
  sp_xyz_ctx *ctx = new (std::nothrow) sp_xyz_ctx();
  if (!ctx) return SP_ENOMEM;
  *out = ctx;
  int rc = check_xyz_problem(ctx, p);
  if (rc != SP_OK) return rc;
  SP_XHIP(hipGetDevice(&ctx->device));
  SP_XHIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  XyzDev &f = ctx->fd;
  f.nw = p->nw;
  f.L = p->L;
  f.R = p->R;
  f.B = p->B;
  f.Rb = p->Rb;
  f.nf = (long long)p->nw * p->nw * (p->nw + 1) / 2;
  f.np = f.nf * p->L;
  f.emax = 2 * p->nw + 8;
  const int O = p->ovl_off[p->L];
  // bubble matrices: one per spin-component pair (a, b) that some overlap
  // transformation maps a bundle component onto, plus the density pair
  int slot_of[16];
  for (int k = 0; k < 16; ++k) slot_of[k] = -1;
  f.npairs = 0;
  auto use_pair = [&](int a, int b) {
    if (slot_of[4 * a + b] < 0) {
      slot_of[4 * a + b] = f.npairs;
      f.pair[f.npairs++] = 4 * a + b;
    }
  };
  use_pair(kD, kD);
  for (int i = 0; i < O; ++i)
    for (int c = 0; c < 3; ++c) use_pair(p->ovl_perm1[3 * i + c], p->ovl_perm2[3 * i + c]);
  const int Lp = (p->L + 3) & ~3;  // row stride of the bubble matrices
  std::vector<int4> midx(O > 0 ? O : 1);
  for (int i = 0; i < O; ++i) {
    const int base = p->ovl_rid1[i] * Lp + p->ovl_rid2[i];
    int v[4];
    for (int c = 0; c < 3; ++c) v[c] = slot_of[4 * p->ovl_perm1[3 * i + c] + p->ovl_perm2[3 * i + c]] * Lp * Lp + base;
    v[3] = slot_of[4 * kD + kD] * Lp * Lp + base;
    midx[i] = make_int4(v[0], v[1], v[2], v[3]);
  }
  int max_lds = 0;
  SP_XHIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
  // producer/worker kernel: (slot, rid) worker threads, kXyzPer points per
  // slot and chunk; the producer wave holds one lane per (point, pair)
  f.slots = (kXyzPW - 64) / p->L;
  f.nb = f.slots * kXyzPer;
  if (f.nb > 32) f.nb = 32;
  ctx->lds = (size_t)xyz_lds(f.nw, f.L, f.nb, f.slots, f.npairs).total;
  const char *force = std::getenv("SPFRG_XYZ_BUBBLE");
  ctx->direct = ctx->lds > (size_t)max_lds || (force && std::strcmp(force, "direct") == 0);
  if (ctx->direct) {
    int nb = kXyzThreads / (2 * p->L);
    if (nb > kXyzMaxPoints) nb = kXyzMaxPoints;
    if (nb < 1) nb = 1;
    f.nb = nb;
    ctx->lds = (size_t)xyz_lds_direct(f.nw, f.L, f.nb).total;
  }
  if (ctx->lds > (size_t)max_lds) return xfail(ctx, SP_EARG, "problem too large for the LDS plan of the dGamma kernel");
  // 32-bit byte offsets of the buffer loads cover all four planes
  if (4 * f.np * 8 >= (1LL << 31)) return xfail(ctx, SP_EARG, "vertex too large for 32-bit byte offsets");
  f.g_bytes = (unsigned)(4 * f.np * 8);
  const size_t nv = 4 * (size_t)f.np;
  const int emax = f.emax;
  const size_t nc = (size_t)f.B * f.Rb;
  double *mesh = nullptr;
  int *site = nullptr, *sitep = nullptr, *inv = nullptr, *invp = nullptr, *ooff = nullptr, *o1 = nullptr,
      *o2 = nullptr, *op1 = nullptr, *op2 = nullptr, *rf = nullptr, *ri = nullptr, *corr = nullptr, *corrp = nullptr;
  if ((rc = xalloc(ctx, &mesh, f.nw)) || (rc = xalloc(ctx, &site, f.L)) || (rc = xalloc(ctx, &sitep, 3 * f.L)) ||
      (rc = xalloc(ctx, &inv, f.L)) || (rc = xalloc(ctx, &invp, 3 * f.L)) || (rc = xalloc(ctx, &ooff, f.L + 1)) ||
      (rc = xalloc(ctx, &o1, O)) || (rc = xalloc(ctx, &o2, O)) || (rc = xalloc(ctx, &op1, 3 * (size_t)O)) ||
      (rc = xalloc(ctx, &op2, 3 * (size_t)O)) || (rc = xalloc(ctx, &rf, f.R)) || (rc = xalloc(ctx, &ri, f.R)) ||
      (rc = xalloc(ctx, &corr, nc)) || (rc = xalloc(ctx, &corrp, 3 * nc)) || (rc = xalloc(ctx, &f.G, nv)) ||
      (rc = xalloc(ctx, &f.dG, nv)) || (rc = xalloc(ctx, &f.v2, f.nw)) || (rc = xalloc(ctx, &f.dv2, f.nw)) ||
      (rc = xalloc(ctx, &f.Wwp, (size_t)f.nw * emax)) || (rc = xalloc(ctx, &f.Wc, (size_t)f.nw * emax)) ||
      (rc = xalloc(ctx, &f.Wn, f.nw)) || (rc = xalloc(ctx, &f.Cwp, emax)) || (rc = xalloc(ctx, &f.Cw, emax)) ||
      (rc = xalloc(ctx, &f.Cn, 1)) || (rc = xalloc(ctx, &f.corr_partial, (size_t)emax * 4 * f.L)) ||
      (rc = xalloc(ctx, &f.corr_out, 4 * nc)) || (rc = xalloc(ctx, &f.flags, 4)))
    return rc;
  int4 *midx_d = nullptr;
  if ((rc = xalloc(ctx, &midx_d, midx.size()))) return rc;
  SP_XHIP(hipMemcpy(midx_d, midx.data(), midx.size() * sizeof(int4), hipMemcpyHostToDevice));
  f.ovl_midx = midx_d;
  f.mesh = mesh;
  f.site_rid = site;
  f.site_perm = sitep;
  f.inv_rid = inv;
  f.inv_perm = invp;
  f.ovl_off = ooff;
  f.ovl_rid1 = o1;
  f.ovl_rid2 = o2;
  f.ovl_perm1 = op1;
  f.ovl_perm2 = op2;
  f.range_fwd = rf;
  f.range_inv = ri;
  f.corr_rid = corr;
  f.corr_perm = corrp;
  std::vector<double> m(f.nw);
  for (int i = 0; i < f.nw; ++i) m[i] = (double)p->mesh[i];
  geo_mesh_params(m.data(), f.nw, f.geo, f.ga, f.gb);
  SP_XHIP(hipMemcpy(mesh, m.data(), f.nw * 8, hipMemcpyHostToDevice));
  SP_XHIP(hipMemcpy(site, p->site_rid, f.L * 4, hipMemcpyHostToDevice));
  SP_XHIP(hipMemcpy(sitep, p->site_perm, 3 * f.L * 4, hipMemcpyHostToDevice));
  SP_XHIP(hipMemcpy(inv, p->inv_rid, f.L * 4, hipMemcpyHostToDevice));
  SP_XHIP(hipMemcpy(invp, p->inv_perm, 3 * f.L * 4, hipMemcpyHostToDevice));
  SP_XHIP(hipMemcpy(ooff, p->ovl_off, (f.L + 1) * 4, hipMemcpyHostToDevice));
  if (O > 0) {
    SP_XHIP(hipMemcpy(o1, p->ovl_rid1, (size_t)O * 4, hipMemcpyHostToDevice));
    SP_XHIP(hipMemcpy(o2, p->ovl_rid2, (size_t)O * 4, hipMemcpyHostToDevice));
    SP_XHIP(hipMemcpy(op1, p->ovl_perm1, 3 * (size_t)O * 4, hipMemcpyHostToDevice));
    SP_XHIP(hipMemcpy(op2, p->ovl_perm2, 3 * (size_t)O * 4, hipMemcpyHostToDevice));
  }
  SP_XHIP(hipMemcpy(rf, p->range_fwd, (size_t)f.R * 4, hipMemcpyHostToDevice));
  SP_XHIP(hipMemcpy(ri, p->range_inv, (size_t)f.R * 4, hipMemcpyHostToDevice));
  SP_XHIP(hipMemcpy(corr, p->corr_rid, nc * 4, hipMemcpyHostToDevice));
  SP_XHIP(hipMemcpy(corrp, p->corr_perm, 3 * nc * 4, hipMemcpyHostToDevice));
  SP_XHIP(hipMemset(f.G, 0, nv * 8));
  SP_XHIP(hipMemset(f.dG, 0, nv * 8));
  SP_XHIP(hipMemset(f.v2, 0, f.nw * 8));
  SP_XHIP(hipMemset(f.dv2, 0, f.nw * 8));
  SP_XHIP(hipMemset(f.flags, 0, 16));
  SP_XHIP(hipMemset(f.Wn, 0, f.nw * 4));
  SP_XHIP(hipMemset(f.Cn, 0, 4));
  SP_XHIP(hipFuncSetAttribute((const void *)k_xyz_dgamma, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  SP_XHIP(hipFuncSetAttribute((const void *)k_xyz_dgamma_direct, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  ctx->host.resize(nv);
  return SP_OK;
}

int sp_xyz_ctx_destroy(sp_xyz_ctx *ctx) {
  if (!ctx) return SP_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (int k = 0; k < 4; ++k)
    for (auto &pr : ctx->pending[k]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
  for (void *q : ctx->allocs) (void)hipFree(q);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SP_OK;
}

int sp_xyz_set_state(sp_xyz_ctx *ctx, double cutoff, const double *v2, const double *v4) {
  if (!ctx || !v2 || !v4) return SP_EARG;
  if (!std::isfinite(cutoff) || !(cutoff > 0.0)) return xfail(ctx, SP_EARG, "cutoff must be finite and positive");
  SP_XHIP(hipStreamSynchronize(ctx->stream));
  SP_XHIP(hipMemcpy(ctx->fd.v2, v2, ctx->fd.nw * 8, hipMemcpyHostToDevice));
  planes_to_interleaved(v4, ctx->host.data(), (size_t)ctx->fd.np);
  SP_XHIP(hipMemcpy(ctx->fd.G, ctx->host.data(), 4 * (size_t)ctx->fd.np * 8, hipMemcpyHostToDevice));
  SP_XHIP(hipMemset(ctx->fd.flags, 0, 16));
  ctx->cutoff = cutoff;
  ctx->have_flow = false;
  return SP_OK;
}

int sp_xyz_get_state(sp_xyz_ctx *ctx, double *cutoff, double *v2, double *v4) {
  if (!ctx) return SP_EARG;
  SP_XHIP(hipStreamSynchronize(ctx->stream));
  if (cutoff) *cutoff = ctx->cutoff;
  if (v2) SP_XHIP(hipMemcpy(v2, ctx->fd.v2, ctx->fd.nw * 8, hipMemcpyDeviceToHost));
  if (v4) {
    SP_XHIP(hipMemcpy(ctx->host.data(), ctx->fd.G, 4 * (size_t)ctx->fd.np * 8, hipMemcpyDeviceToHost));
    interleaved_to_planes(ctx->host.data(), v4, (size_t)ctx->fd.np);
  }
  return SP_OK;
}

int sp_xyz_get_flow(sp_xyz_ctx *ctx, double *dv2, double *dv4) {
  if (!ctx) return SP_EARG;
  if (!ctx->have_flow) return xfail(ctx, SP_ESTATE, "no flow computed at the current cutoff");
  SP_XHIP(hipStreamSynchronize(ctx->stream));
  if (dv2) SP_XHIP(hipMemcpy(dv2, ctx->fd.dv2, ctx->fd.nw * 8, hipMemcpyDeviceToHost));
  if (dv4) {
    SP_XHIP(hipMemcpy(ctx->host.data(), ctx->fd.dG, 4 * (size_t)ctx->fd.np * 8, hipMemcpyDeviceToHost));
    interleaved_to_planes(ctx->host.data(), dv4, (size_t)ctx->fd.np);
  }
  return SP_OK;
}

int sp_xyz_compute_step(sp_xyz_ctx *ctx) {
  if (!ctx) return SP_EARG;
  SP_XHIP(hipMemsetAsync(ctx->fd.flags, 0, 16, ctx->stream));
  SP_XHIP(xyz_launch_compute(ctx));
  ctx->have_flow = true;
  return SP_OK;
}

int sp_xyz_finalize_step(sp_xyz_ctx *ctx, double new_cutoff) {
  if (!ctx) return SP_EARG;
  if (!std::isfinite(new_cutoff) || !(new_cutoff > 0.0)) return xfail(ctx, SP_EARG, "cutoff must be finite and positive");
  if (!ctx->have_flow) return xfail(ctx, SP_ESTATE, "finalize_step without a computed flow");
  XyzDev F = ctx->fd;
  F.euler = 1;
  F.step = new_cutoff - ctx->cutoff;
  long long blocks = (4 * F.np + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  {
    XyzProf ps(ctx, 3);
    hipLaunchKernelGGL(k_xyz_euler, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, F);
  }
  SP_XHIP(hipGetLastError());
  ctx->cutoff = new_cutoff;
  ctx->have_flow = false;
  return SP_OK;
}

int sp_xyz_is_diverged(sp_xyz_ctx *ctx, int *diverged) {
  if (!ctx || !diverged) return SP_EARG;
  if (!ctx->have_flow) return xfail(ctx, SP_ESTATE, "no flow computed at the current cutoff");
  int fl[2] = {0, 0};
  SP_XHIP(hipMemcpyAsync(fl, ctx->fd.flags, 8, hipMemcpyDeviceToHost, ctx->stream));
  SP_XHIP(hipStreamSynchronize(ctx->stream));
  if (fl[1]) return xfail(ctx, SP_EARG, "evaluation list overflow");
  *diverged = fl[0] ? 1 : 0;
  return SP_OK;
}

int sp_xyz_measure_correlation(sp_xyz_ctx *ctx, double *out) {
  if (!ctx || !out) return SP_EARG;
  XyzDev F = ctx->fd;
  F.cut = ctx->cutoff;
  hipLaunchKernelGGL(k_xyz_weights, dim3(1), dim3(64), F.nw * sizeof(double), ctx->stream, F, 1);
  hipLaunchKernelGGL(k_xyz_corr_partial, dim3(F.emax), dim3(kXyzThreads), 0, ctx->stream, F);
  hipLaunchKernelGGL(k_xyz_corr_final, dim3(1), dim3(256), 0, ctx->stream, F);
  SP_XHIP(hipGetLastError());
  const size_t n = 4 * (size_t)F.B * F.Rb;
  SP_XHIP(hipMemcpyAsync(out, F.corr_out, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  SP_XHIP(hipStreamSynchronize(ctx->stream));
  return SP_OK;
}

int sp_xyz_synchronize(sp_xyz_ctx *ctx) {
  if (!ctx) return SP_EARG;
  SP_XHIP(hipStreamSynchronize(ctx->stream));
  return SP_OK;
}

const char *sp_xyz_last_error(const sp_xyz_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sp_xyz_profile_enable(sp_xyz_ctx *ctx, int enable) {
  if (!ctx) return SP_EARG;
  ctx->prof = enable != 0;
  return SP_OK;
}

int sp_xyz_profile_get(sp_xyz_ctx *ctx, int kernel, double *total_ms, long *launches) {
  if (!ctx || kernel < 0 || kernel > 3 || !total_ms || !launches) return SP_EARG;
  SP_XHIP(hipStreamSynchronize(ctx->stream));
  for (auto &pr : ctx->pending[kernel]) {
    float ms = 0.f;
    SP_XHIP(hipEventElapsedTime(&ms, pr.first, pr.second));
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
