// MI355X (gfx950) kernels and C ABI of the SU(2) pf-FRG flow engine.
//
// Device data layout (per flow)
//   G, dG   vertex and its flow, interleaved channels: double2 {spin, density}
//           at [(tri(s,u) * Nw + t) * L + rid], tri(s,u) = s(s+1)/2 + u, s >= u.
//           One 16-byte load fetches both channels of a support point.
//   G32     FP32 shadow of G (float2, same index) for the mixed-precision form,
//           G * 2^-S with a power-of-two scale S planned from the largest
//           exponent of G; rewritten by every Euler step from the FP64 master
//           copy.  Every dGamma term is quadratic in G, so the FP32 result is
//           rescaled by 2^(2S).
//   v2,dv2  self-energy and its flow, Nw doubles.
//   W       per exact-frequency index x: the evaluation points w' of the
//           frequency integral of a channel whose transfer frequency is
//           mesh[x] (conventional +-Lambda terms and the three trapezoid
//           windows of the Katanin term) with their coefficients
//           c = trapezoid weight * propagator product.  Depends only on
//           (x, Lambda), so it is built once per step for all items.
//   FlowDev per-flow descriptor (pointers, sizes, cutoff, Euler step).  Every
//           kernel takes an array of them, so one launch serves one flow or a
//           batch of resident flows of different sizes (the served sweep).
//
// Kernels (one derivative evaluation = k_dsigma, k_weights, k_dgamma):
//   k_dsigma   one wavefront per (flow, mesh point); lanes split the range sum.
//   k_weights  one thread per (flow, exact index): window points, coefficients.
//   k_dgamma   one workgroup per (flow, frequency triple (s,t,u)): 256 threads,
//              128 for lattices of at most 16 representative sites;
//              blocks mapped to XCDs in contiguous item ranges.  The evaluation
//              points of the three channels form one sequence of chunks; per chunk:
//              (A) producer wave (the last wave), one lane per (point, pair of access
//                  buffers): the two bracket searches a pair shares, sign /
//                  site-exchange folding, support byte offsets and weights ->
//                  one of two LDS descriptor buffers, one chunk ahead;
//              (B) worker waves (the others), (slot, rid) pairs: 4 x 4 vertex gathers
//                  (buffer loads with 32-bit offsets) and the channel algebra
//                  from the other descriptor buffer;
//              (C, t channel) the RPA lattice bubble accumulated as
//                  C[a][b] += sum_e c_e A2_e[a] A3_e[b] in FP64 in LDS (2x4
//                  register tiles, all waves), contracted with the overlap
//                  lists once per item.
//              Template MIXED: FP32 gathers and algebra with compensated
//              (two-sum) accumulation; the lattice-bubble sums, descriptors,
//              frequencies and the result stay FP64.
//   k_euler    in-place Gamma += dLambda * dGamma (streaming, HBM bound);
//              skipped for a flow whose derivative holds a NaN.
//   k_corr_*   static susceptibility in FP64, one workgroup per (flow, outer
//              frequency).
//
// Reference semantics followed (paths relative to the SpinParser repository):
//   src/FrequencyDiscretization.hpp:253-356, src/lib/Integrator.hpp:137-287,
//   src/SU2/SU2VertexTwoParticle.hpp:185-632, src/SU2/SU2FrgCore.cpp:111-431,
//   src/SU2/SU2MeasurementCorrelation.cpp:77-177.
// Frequency arguments, window predicates and bracket comparisons are evaluated
// with the same IEEE operations as the reference (the file is compiled with
// -ffp-contract=off); only the accumulations use explicit fma().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "spfrg.h"
#include "frg_common.h"

namespace {

// launch-shape tunables of k_dgamma (compile-time; the defaults are the tuned ones)
#ifndef SPFRG_THREADS
#define SPFRG_THREADS 256
#endif
// evaluation points per chunk (FP64 / mixed form), at most: the descriptor
// buffers and bubble factors of a chunk must leave LDS for the workgroups per
// CU that the kernel's registers allow (launch_shape shrinks the chunk)
#ifndef SPFRG_MAX_BATCH
#define SPFRG_MAX_BATCH 16
#endif
#ifndef SPFRG_MAX_BATCH_MIXED
#define SPFRG_MAX_BATCH_MIXED 32
#endif
#ifndef SPFRG_MIN_BLOCKS
#define SPFRG_MIN_BLOCKS 4
#endif
#ifndef SPFRG_MIN_BLOCKS_MIXED
#define SPFRG_MIN_BLOCKS_MIXED 5
#endif
constexpr int kThreads = SPFRG_THREADS;
// small lattices: the producer's work per evaluation point does not shrink
// with L while the workers' does, so flows with L <= kSmallL run the dGamma
// kernel in workgroups of kSmallThreads (one worker wave per producer wave,
// twice the workgroups per CU); measured 17-43 % faster for L = 7, 9, 11 and
// 16, slower for L = 23 and 31
constexpr int kSmallThreads = 128;
constexpr int kSmallL = 16;
inline int dgamma_threads(int L) { return L <= kSmallL ? kSmallThreads : kThreads; }
constexpr int kMaxBatch = SPFRG_MAX_BATCH;
constexpr int kMaxBatchMixed = SPFRG_MAX_BATCH_MIXED;
#ifndef SPFRG_UNROLL_FP64
#define SPFRG_UNROLL_FP64 1
#endif
#ifndef SPFRG_UNROLL_MIXED
#define SPFRG_UNROLL_MIXED 1
#endif
// register tile of the FP64 lattice-bubble rank update (rows x columns; both
// divide 4, the row stride of the bubble matrices).  2 x 4 keeps 144 of the
// 256 threads busy at L = 23 with 3 LDS reads per 8 FMAs (2 x 2: 2 rounds of
// 4 FMAs per 2 reads; 4 x 4: only 72 threads)
#ifndef SPFRG_TILE_R
#define SPFRG_TILE_R 2
#endif
#ifndef SPFRG_TILE_C
#define SPFRG_TILE_C 4
#endif
#ifndef SPFRG_CHUNK
#define SPFRG_CHUNK 16
#endif
// support gathers of one evaluation point when U == 1: 0 = per access buffer,
// 1 = the next buffer's four loads in flight while the current one is used,
// 2 = all sixteen loads issued before any use (the FP32 gathers of the mixed
// form fit this in the occupancy-4 register budget; FP64 does not).  FP64
// uses 1: 3 % faster than 0 at small cutoffs although the occupancy-4 budget
// then spills 12 bytes per lane; 2 needs 155 VGPRs (occupancy 3) and is 15 %
// slower
#ifndef SPFRG_GATHER
#define SPFRG_GATHER 1
#endif
#ifndef SPFRG_GATHER_MIXED
#define SPFRG_GATHER_MIXED 2
#endif
constexpr long long kChunk = SPFRG_CHUNK;

using namespace frg;


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
                              double t, double u, int geo = 0, float ga = 0.0f, float gb = 0.0f) {
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
  interp_sel(m, nw, s, geo, ga, gb, lS, uS, bS);
  interp_sel(m, nw, t, geo, ga, gb, lT, uT, bT);
  interp_sel(m, nw, u, geo, ga, gb, lU, uU, bU);
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

// ---------------------------------------------------------------------------
// per-flow descriptor
// ---------------------------------------------------------------------------
struct FlowDev {
  double2 *G;
  double2 *dG;
  float2 *G32;  // FP32 shadow of G (mixed precision), null otherwise
  double *v2, *dv2;
  const double *mesh;
  const int *site_rid, *inv_rid, *ovl_off, *ovl_rid1, *ovl_rid2, *range_fwd, *range_inv, *corr_rid;
  const unsigned short *ovl_idx;  // per overlap entry: rid1 * Lp + rid2 (Lp = L rounded up to a multiple of 4)
  double *Wwp, *Wc;
  int *Wn;
  double *Cwp, *Cw;
  int *Cn;
  double *corr_partial, *corr_out;
  int *flags;  // [0] NaN in the flow, [1] evaluation-list overflow, [2] stopped flow,
               // [4] scale exponent S of the FP32 shadow (G32 = G * 2^-S), [5] largest
               // binary exponent of G written since the last plan, [6] planned S
  int nw, L, R, B, Rb, emax, nb, slots, precision, euler;
  int nt;  // threads per workgroup of the dGamma kernel
  long long nf, item_base, nv;
  unsigned g_bytes, g32_bytes;
  double spin, cut, step;
  int pad_index;  // host bookkeeping: position of the flow in the caller's array
  int geo;        // geometric mesh: bracket guess from the logarithm (count_le_geo)
  float ga, gb;
};

__device__ __forceinline__ int flow_of(const FlowDev *__restrict__ fl, int nflows, long long item) {
  int lo = 0, hi = nflows - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fl[mid].item_base <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// W tables: workgroup x < nw -> channel list with propagators; x == nw ->
// plain correlation list (corr_only: that one alone, launched as workgroup 0).
// flags[1] is raised if a list overflows its capacity.  Dynamic LDS: the mesh.
__global__ __launch_bounds__(64) void k_weights(const FlowDev *__restrict__ fl, int corr_only) {
  extern __shared__ double s_wmesh[];
  const FlowDev &F = fl[blockIdx.y];
  const int x = corr_only ? F.nw : (int)blockIdx.x;
  if (x > F.nw) return;
  weight_tables_wave(x, F.mesh, F.nw, F.cut, F.v2, F.dv2, F.Wwp, F.Wc, F.Wn, F.Cwp, F.Cw, F.Cn, F.emax, F.flags + 1,
                     s_wmesh);
}

// ---------------------------------------------------------------------------
// dSigma/dLambda (src/SU2/SU2FrgCore.cpp:139-169): one 256-thread workgroup
// per mesh point, the range sum spread over its four waves; the mesh in LDS
// (dynamic, nw doubles)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dsigma(const FlowDev *__restrict__ fl) {
  extern __shared__ double s_dmesh[];
  __shared__ double part[4];
  const FlowDev &F = fl[blockIdx.y];
  const int i = blockIdx.x;
  if (i >= F.nw) return;
  const int tid = threadIdx.x, lane = tid & 63;
  const int nw = F.nw, L = F.L;
  for (int k = tid; k < nw; k += 256) s_dmesh[k] = F.mesh[k];
  __syncthreads();
  const double *m = s_dmesh;
  const double2 *G = F.G;
  const double cut = F.cut;
  const double w = m[i];
  double sum = 0.0;
  for (int j = tid; j < F.R; j += 256) {
    const double2 a = value_none(G, m, nw, L, F.range_fwd[j], F.range_inv[j], w + cut, 0.0, w - cut, F.geo, F.ga, F.gb);
    const double2 b = value_none(G, m, nw, L, F.range_fwd[j], F.range_inv[j], w - cut, 0.0, w + cut, F.geo, F.ga, F.gb);
    sum += a.y - b.y;
  }
  sum = wave_sum(sum);
  if (lane == 0) part[tid >> 6] = sum;
  __syncthreads();
  if (tid == 0) {
    sum = (part[0] + part[1]) + (part[2] + part[3]);
    double v = 0.0;
    v -= 4.0 * F.spin * sum;
    const double2 p = value_none(G, m, nw, L, 0, 0, w + cut, w - cut, 0.0, F.geo, F.ga, F.gb);
    const double2 q = value_none(G, m, nw, L, 0, 0, w - cut, w + cut, 0.0, F.geo, F.ga, F.gb);
    v += 0.75 * (p.x - q.x);
    v += (p.y - q.y);
    v /= (kTwoPi * (cut + sigma_at(m, nw, F.v2, cut)));
    F.dv2[i] = v;
    if (isnan(v)) atomicOr(&F.flags[0], 1);
  }
}

// ---------------------------------------------------------------------------
// dGamma/dLambda
// ---------------------------------------------------------------------------
// LDS carve-up (bytes), computed identically on host and device.  Lp = L
// rounded up to a multiple of 4 (register tiles of the lattice-bubble update); vs =
// size of the gather type (8 FP64, 4 mixed).  The per-chunk descriptors (c,
// off, wd, ex, loc) are double-buffered: the producer wave fills one buffer
// while the worker waves read the other.  Only the signed density weights
// are stored; the spin weights are their absolute values.
struct LdsLayout {
  int mesh, c, off, wd, ex, loc, dbuf, P, Q, C, rpa, red, site, inv, total;  // buffer b at + b * dbuf
};

__host__ __device__ inline LdsLayout lds_layout(int nw, int L, int nb, int slots, int vs) {
  LdsLayout l;
  const int Lp = (L + 3) & ~3;
  int off = 0;
  auto take = [&](int bytes) { int o = off; off += (bytes + 15) & ~15; return o; };
  l.mesh = take(nw * 8);
  const int d0 = off;
  l.c = take(nb * 8);
  l.off = take(nb * 4 * 16);
  l.wd = take(nb * 4 * 4 * vs);
  l.ex = take(nb * 4 * 4);
  l.loc = take(nb * 4 * 2 * vs);
  l.dbuf = off - d0;
  off += l.dbuf;  // second descriptor buffer
  const int pq = 2 * nb * Lp * vs;
  const int red = slots * L * 2 * 8;
  l.P = take(pq > red ? pq : red);  // the final slot reduction reuses P
  l.red = l.P;
  l.Q = take(pq);
  l.C = take(2 * Lp * Lp * 8);
  l.rpa = take(2 * L * 8);
  l.site = take(L * 4);
  l.inv = take(L * 4);
  l.total = off;
  return l;
}

template <class V2, class V> __device__ __forceinline__ V2 mk2(V a, V b) {
  V2 r;
  r.x = a;
  r.y = b;
  return r;
}

// compensated FP32 accumulator (Knuth two-sum per addition)
struct KSum {
  float s, c;
};

__device__ __forceinline__ void ksum_add(KSum &k, float x) {
  const float t = k.s + x;
  const float z = t - k.s;
  k.c += (k.s - (t - z)) + (x - z);
  k.s = t;
}

// dGamma kernel in workgroups of NT threads.  All waves but the last are
// workers, (slot, rid) pairs with slots = floor((NT - 64) / L); the last wave
// is the producer.  The evaluation points of the three
// channels form one sequence of chunks of nb points; while the workers gather
// and evaluate chunk j, the producer builds the descriptors of chunk j + 1 into
// the other LDS buffer, so the bracket searches and the t channel's site-local
// gathers overlap the workers' gathers instead of stalling them.
template <bool MIXED, int NT, int MINB>
__global__ __launch_bounds__(NT, MINB) void k_dgamma(const FlowDev *__restrict__ fl, int nflows, long long total,
                                                     long long per_xcd) {
  using V = typename std::conditional<MIXED, float, double>::type;
  using V2 = typename std::conditional<MIXED, float2, double2>::type;
  extern __shared__ __align__(16) unsigned char lds_raw[];

  // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs.
  // Hand out chunks of kChunk consecutive items (one (s,u) row of t values
  // shares the s and u vertex slabs) to the XCDs in turn, so each chunk stays
  // in one XCD's L2 while the XCDs still receive statistically equal work.
  const long long j = blockIdx.x >> 3;
  const long long lb = ((j / kChunk) * 8 + (blockIdx.x & 7)) * kChunk + (j % kChunk);
  if (lb >= total) return;
  const int f = flow_of(fl, nflows, lb);
  const FlowDev &F = fl[f];
  const int item = (int)(lb - F.item_base);

  const int nw = F.nw, L = F.L, nb = F.nb, slots = F.slots;
  const int Lp = (L + 3) & ~3;
  const LdsLayout lay = lds_layout(nw, L, nb, slots, (int)sizeof(V));
  double *s_mesh = reinterpret_cast<double *>(lds_raw + lay.mesh);
  V *s_P = reinterpret_cast<V *>(lds_raw + lay.P);
  V *s_Q = reinterpret_cast<V *>(lds_raw + lay.Q);
  double *s_C = reinterpret_cast<double *>(lds_raw + lay.C);  // FP64 lattice-bubble sums (both forms)
  double *s_rpa = reinterpret_cast<double *>(lds_raw + lay.rpa);
  double *s_red = reinterpret_cast<double *>(lds_raw + lay.red);
  int *s_site = reinterpret_cast<int *>(lds_raw + lay.site);
  int *s_inv = reinterpret_cast<int *>(lds_raw + lay.inv);

  const void *gbase = MIXED ? (const void *)F.G32 : (const void *)F.G;
  const __amdgpu_buffer_rsrc_t rG =
      __builtin_amdgcn_make_buffer_rsrc((void *)gbase, (short)0, (int)(MIXED ? F.g32_bytes : F.g_bytes), 0x00020000);
  const int esz = (int)sizeof(V2);
  const int rowbytes = L * esz;

  const int tid = threadIdx.x;
  // SU2VertexTwoParticle::expandIterator (:136-158): item = tri(s,u) * nw + t
  const int su = item / nw;
  const int to = item - su * nw;
  int so = (int)((sqrt(8.0 * su + 1.0) - 1.0) * 0.5);
  while ((so + 1) * (so + 2) / 2 <= su) ++so;
  while (so * (so + 1) / 2 > su) --so;
  const int uo = su - so * (so + 1) / 2;

  for (int i = tid; i < nw; i += NT) s_mesh[i] = F.mesh[i];
  for (int i = tid; i < L; i += NT) { s_site[i] = F.site_rid[i]; s_inv[i] = F.inv_rid[i]; }
  for (int i = tid; i < 2 * Lp * Lp; i += NT) s_C[i] = 0.0;
  for (int i = tid; i < 2 * nb * Lp; i += NT) { s_P[i] = V(0); s_Q[i] = V(0); }
  for (int i = tid; i < 2 * L; i += NT) s_rpa[i] = 0.0;
  __syncthreads();

  const int lane = tid & 63;
  const bool producer = tid >= NT - 64;
  const int slot = tid / L;
  const int rid = tid - slot * L;
  const bool active = slot < slots;  // slots * L <= NT - 64: workers only
  const int sr_fwd = active ? s_site[rid] * esz : 0;
  const int sr_inv = active ? s_inv[rid] * esz : 0;
  double acc_s = 0.0, acc_d = 0.0;   // FP64 form
  KSum ks{0.f, 0.f}, kd{0.f, 0.f};  // mixed form

  // chunk schedule: chunk j covers points [e0, e0 + nb) of channel ch
  const int nS = F.Wn[so], nT = F.Wn[to], nU = F.Wn[uo];
  const int cS = (nS + nb - 1) / nb, cT = (nT + nb - 1) / nb, cU = (nU + nb - 1) / nb;
  const int nchunks = cS + cT + cU;

  // (A) producer: one lane per (point, pair) of chunk jj builds the descriptors
  // into buffer bf -- the two bracket searches shared by the 2 (s, u channels)
  // or 4 (t channel) access buffers of the pair; the t channel's site-local
  // buffers are gathered right away
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
        if (!MIXED) {
          // FP64 form: four tasks per point, one access buffer of a pair each (in the t
          // channel one gathered and one site-local buffer), the buffer's
          // arguments picked arithmetically so that every lane runs the same
          // straight-line code: the wave issues one descriptor assembly per
          // lane instead of looping over the pair's buffers
          const double s = s_mesh[so], t = s_mesh[to], u = s_mesh[uo];
          const double w1p = 0.5 * (s + t + u);
          const double w1 = 0.5 * (s - t + u);
          const double w2p = 0.5 * (s - t - u);
          const double w2 = 0.5 * (s + t - u);
          double *s_c = reinterpret_cast<double *>(lds_raw + lay.c + bf * lay.dbuf);
          int4 *s_off = reinterpret_cast<int4 *>(lds_raw + lay.off + bf * lay.dbuf);
          V2 *s_wd = reinterpret_cast<V2 *>(lds_raw + lay.wd + bf * lay.dbuf);
          int *s_ex = reinterpret_cast<int *>(lds_raw + lay.ex + bf * lay.dbuf);
          V *s_loc = reinterpret_cast<V *>(lds_raw + lay.loc + bf * lay.dbuf);
          for (int task = lane; task < 4 * nvalid; task += 64) {
            const int e = task >> 2, pr = (task >> 1) & 1, h = task & 1;
            const double wp = F.Wwp[(size_t)xo * F.emax + e0 + e];
            // the (point, pair) buffers of SU2FrgCore.cpp:199-334 (k = access
            // buffer, a / c = signed s / u arguments, swap: interpolations fed
            // as (slot 2, slot 1)); buffer h of the pair
            double x0, x1;
            int exb;  // exactly one of the signed s / u arguments negative: site exchange
            if (ch == 0) {
              x0 = fabs(pr ? w1p + wp : w1 + wp);
              x1 = fabs(pr ? w2p + wp : w2 + wp);
              const double c = h ? (pr ? w1p + wp : w1 + wp) : (pr ? -w2p - wp : -w2 - wp);
              exb = (s < 0) != (c < 0);
            } else if (ch == 1) {
              const double A = pr ? w2p - wp : w1 - wp, B = pr ? w2 + wp : w1p + wp;
              x0 = fabs(A);
              x1 = fabs(B);
              const double c = h ? (pr ? wp - w2p : w1 - wp) : (pr ? -w2 - wp : w1p + wp);
              exb = ((h ? B : A) < 0) != (c < 0);
            } else {
              x0 = fabs(pr ? w1p + wp : w1 + wp);
              x1 = fabs(pr ? w2 - wp : w2p - wp);
              const double a = h ? (pr ? w2 - wp : w2p - wp) : (pr ? w1p + wp : w1 + wp);
              exb = (a < 0) != (u < 0);
            }
            // swap: the buffer takes the pair's interpolations in the other order
            const double xa = h ? x1 : x0, xb = h ? x0 : x1;
            Interp ia, ib;
            if (F.geo) interp_pair_geo(s_mesh, nw, xa, xb, F.ga, F.gb, ia, ib);
            else interp_pair(s_mesh, nw, xa, xb, ia, ib);
            {
              const Desc4 d = assemble(ch, xo, ia, ib, nw, rowbytes);
              const int sk = e * 4 + 2 * h + pr;
              s_off[sk] = make_int4(d.off[0], d.off[1], d.off[2], d.off[3]);
              s_wd[2 * sk] = mk2<V2>(V(d.wd[0]), V(d.wd[1]));
              s_wd[2 * sk + 1] = mk2<V2>(V(d.wd[2]), V(d.wd[3]));
              s_ex[sk] = exb;
            }
            if (ch == 1) {
              // site-local buffer (k = 5, 7 for pair 0; 4, 6 for pair 1) with
              // the same interpolations: SU2VertexTwoParticle::getValueLocal
              // (:347-360), site 0, no exchange
              const Desc4 d = assemble(2, xo, ia, ib, nw, rowbytes);
              V vs = V(0), vd = V(0);
  #pragma unroll
              for (int i = 0; i < 4; ++i) {
                const V2 v = buf_load<V2>(rG, d.off[i]);
                vs = fma(V(d.ws[i]), v.x, vs);
                vd = fma(V(d.wd[i]), v.y, vd);
              }
              const int kl = (pr ? 0 : 1) + 2 * h;  // local index k - 4
              s_loc[(e * 4 + kl) * 2 + 0] = fma(V(0.25), vs, -vd);
              s_loc[(e * 4 + kl) * 2 + 1] = -fma(V(0.75), vs, vd);
            }
            if (pr == 0 && h == 0) s_c[e] = (ch == 2 ? -1.0 : 1.0) * F.Wc[(size_t)xo * F.emax + e0 + e];
          }
        } else if (lane < 2 * nvalid) {
          // transfer frequencies (SU2FrgCore.cpp:193-197), recomputed here so
          // that they occupy no registers in the worker code
          const double s = s_mesh[so], t = s_mesh[to], u = s_mesh[uo];
          const double w1p = 0.5 * (s + t + u);
          const double w1 = 0.5 * (s - t + u);
          const double w2p = 0.5 * (s - t - u);
          const double w2 = 0.5 * (s + t - u);
          double *s_c = reinterpret_cast<double *>(lds_raw + lay.c + bf * lay.dbuf);
          int4 *s_off = reinterpret_cast<int4 *>(lds_raw + lay.off + bf * lay.dbuf);
          V2 *s_wd = reinterpret_cast<V2 *>(lds_raw + lay.wd + bf * lay.dbuf);
          int *s_ex = reinterpret_cast<int *>(lds_raw + lay.ex + bf * lay.dbuf);
          V *s_loc = reinterpret_cast<V *>(lds_raw + lay.loc + bf * lay.dbuf);
          const int e = lane >> 1, pr = lane & 1;
          const double wp = F.Wwp[(size_t)xo * F.emax + e0 + e];
          double x0, x1;
          PairBuf pb[4];
          int nbufs = 2;
          if (ch == 0) {
            if (pr == 0) {
              x0 = fabs(w1 + wp); x1 = fabs(w2 + wp);
              pb[0] = {0, s, -w2 - wp, false, 0};
              pb[1] = {2, s, w1 + wp, true, 0};
            } else {
              x0 = fabs(w1p + wp); x1 = fabs(w2p + wp);
              pb[0] = {1, s, -w2p - wp, false, 0};
              pb[1] = {3, s, w1p + wp, true, 0};
            }
          } else if (ch == 1) {
            nbufs = 4;
            if (pr == 0) {
              x0 = fabs(w1 - wp); x1 = fabs(w1p + wp);
              pb[0] = {0, w1 - wp, w1p + wp, false, 1};
              pb[1] = {2, w1p + wp, w1 - wp, true, 1};
              pb[2] = {5, w1 - wp, -t, false, 2};
              pb[3] = {7, w1p + wp, -t, true, 2};
            } else {
              x0 = fabs(w2p - wp); x1 = fabs(w2 + wp);
              pb[0] = {1, w2p - wp, -w2 - wp, false, 1};
              pb[1] = {3, w2 + wp, wp - w2p, true, 1};
              pb[2] = {4, w2p - wp, t, false, 2};
              pb[3] = {6, w2 + wp, t, true, 2};
            }
          } else {
            if (pr == 0) {
              x0 = fabs(w1 + wp); x1 = fabs(w2p - wp);
              pb[0] = {0, w1 + wp, u, false, 2};
              pb[1] = {2, w2p - wp, u, true, 2};
            } else {
              x0 = fabs(w1p + wp); x1 = fabs(w2 - wp);
              pb[0] = {1, w1p + wp, u, false, 2};
              pb[1] = {3, w2 - wp, u, true, 2};
            }
          }
          Interp i0, i1;
          if (F.geo) interp_pair_geo(s_mesh, nw, x0, x1, F.ga, F.gb, i0, i1);
          else interp_pair(s_mesh, nw, x0, x1, i0, i1);
          // fully unrolled with a guard so that pb[] stays in registers (a
          // dynamically indexed local array would live in scratch memory)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (q >= nbufs) break;
            const PairBuf &bb = pb[q];
            const Desc4 d = assemble(bb.chan, xo, bb.swap ? i1 : i0, bb.swap ? i0 : i1, nw, rowbytes);
            if (bb.k < 4) {
              const int sk = e * 4 + bb.k;
              s_off[sk] = make_int4(d.off[0], d.off[1], d.off[2], d.off[3]);
              s_wd[2 * sk] = mk2<V2>(V(d.wd[0]), V(d.wd[1]));
              s_wd[2 * sk + 1] = mk2<V2>(V(d.wd[2]), V(d.wd[3]));
              s_ex[sk] = ((bb.a < 0) != (bb.c < 0)) ? 1 : 0;
            } else {
              // SU2VertexTwoParticle::getValueLocal (:347-360): site 0, no exchange
              V vs = V(0), vd = V(0);
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                const V2 v = buf_load<V2>(rG, d.off[i]);
                vs = fma(V(d.ws[i]), v.x, vs);
                vd = fma(V(d.wd[i]), v.y, vd);
              }
              // the factors the t-channel algebra applies to this buffer's
              // partner bundle: spin 1/4 vs - vd, density -(3/4 vs + vd)
              s_loc[(e * 4 + (bb.k - 4)) * 2 + 0] = fma(V(0.25), vs, -vd);
              s_loc[(e * 4 + (bb.k - 4)) * 2 + 1] = -fma(V(0.75), vs, vd);
            }
          }
          if (pr == 0) s_c[e] = (ch == 2 ? -1.0 : 1.0) * F.Wc[(size_t)xo * F.emax + e0 + e];
        }
      }
    } else if (active && jc >= 0) {
      // (B) gathers and channel algebra per (slot, rid); U evaluation points
      // per iteration so that 16 U loads are in flight before the first use
      const double *s_c = reinterpret_cast<const double *>(lds_raw + lay.c + cur * lay.dbuf);
      const int4 *s_off = reinterpret_cast<const int4 *>(lds_raw + lay.off + cur * lay.dbuf);
      const V2 *s_wd = reinterpret_cast<const V2 *>(lds_raw + lay.wd + cur * lay.dbuf);
      const int *s_ex = reinterpret_cast<const int *>(lds_raw + lay.ex + cur * lay.dbuf);
      const V *s_loc = reinterpret_cast<const V *>(lds_raw + lay.loc + cur * lay.dbuf);
      constexpr int U = MIXED ? SPFRG_UNROLL_MIXED : SPFRG_UNROLL_FP64;
      for (int e = slot; e < nvalid; e += U * slots) {
        int ev[U];
        bool ok[U];
        // U == 1: loads and weights per access buffer, leaving the schedule
        // to the compiler; U > 1: all 16 U loads issued before any use
        constexpr int kGather = MIXED ? SPFRG_GATHER_MIXED : SPFRG_GATHER;
        constexpr bool kPre = U > 1 || kGather == 2;
        V2 val[kPre ? U : 1][16];
#pragma unroll
        for (int q = 0; q < U; ++q) {
          ev[q] = e + q * slots;
          ok[q] = ev[q] < nvalid;
          if (!ok[q]) ev[q] = e;
          if (kPre) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const int sk = ev[q] * 4 + k;
              const int sr = (s_ex[sk] & 1) ? sr_inv : sr_fwd;
              const int4 o = s_off[sk];
              val[q][4 * k + 0] = buf_load<V2>(rG, o.x + sr);
              val[q][4 * k + 1] = buf_load<V2>(rG, o.y + sr);
              val[q][4 * k + 2] = buf_load<V2>(rG, o.z + sr);
              val[q][4 * k + 3] = buf_load<V2>(rG, o.w + sr);
            }
          }
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
          if (!ok[q]) continue;
          const int ee = ev[q];
          V As[4], Ad[4];
          V2 vn[4];
          if (!kPre && kGather == 1) {
            const int sr = (s_ex[ee * 4] & 1) ? sr_inv : sr_fwd;
            const int4 o = s_off[ee * 4];
            vn[0] = buf_load<V2>(rG, o.x + sr);
            vn[1] = buf_load<V2>(rG, o.y + sr);
            vn[2] = buf_load<V2>(rG, o.z + sr);
            vn[3] = buf_load<V2>(rG, o.w + sr);
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int sk = ee * 4 + k;
            // density weights carry the sign of the s <-> u exchanged storage
            const V2 wd01 = s_wd[2 * sk], wd23 = s_wd[2 * sk + 1];
            // spin weight = |density weight|: the bilinear weights are >= 0, so
            // the absolute value (a free source modifier of the FMA) is exact,
            // and the workers read half the weight bytes from LDS
            const V2 ws01 = mk2<V2>(fabs(wd01.x), fabs(wd01.y)), ws23 = mk2<V2>(fabs(wd23.x), fabs(wd23.y));
            const int exf = s_ex[sk];
            V2 v[4];
            if (kPre) {
              v[0] = val[q][4 * k + 0]; v[1] = val[q][4 * k + 1];
              v[2] = val[q][4 * k + 2]; v[3] = val[q][4 * k + 3];
            } else if (kGather == 1) {
              v[0] = vn[0]; v[1] = vn[1]; v[2] = vn[2]; v[3] = vn[3];
              if (k < 3) {
                const int sr = (s_ex[sk + 1] & 1) ? sr_inv : sr_fwd;
                const int4 o = s_off[sk + 1];
                vn[0] = buf_load<V2>(rG, o.x + sr);
                vn[1] = buf_load<V2>(rG, o.y + sr);
                vn[2] = buf_load<V2>(rG, o.z + sr);
                vn[3] = buf_load<V2>(rG, o.w + sr);
              }
            } else {
              const int sr = (exf & 1) ? sr_inv : sr_fwd;
              const int4 o = s_off[sk];
              v[0] = buf_load<V2>(rG, o.x + sr);
              v[1] = buf_load<V2>(rG, o.y + sr);
              v[2] = buf_load<V2>(rG, o.z + sr);
              v[3] = buf_load<V2>(rG, o.w + sr);
            }
            As[k] = fma(ws23.y, v[3].x, fma(ws23.x, v[2].x, fma(ws01.y, v[1].x, ws01.x * v[0].x)));
            Ad[k] = fma(wd23.y, v[3].y, fma(wd23.x, v[2].y, fma(wd01.y, v[1].y, wd01.x * v[0].y)));
          }
          const double c = s_c[ee];
          V rs, rd;
          if (ch != 1) {
            // integralKernelS (src/SU2/SU2FrgCore.cpp:199-227), integralKernelU
            // (:305-334): with h = As0 As1 + As2 As3,
            // S: rs = -h/2 + Ad0 As1 + Ad2 As3 + As0 Ad1 + As2 Ad3, U: rs = h/2 + (same)
            // rd = 3/16 h + Ad0 Ad1 + Ad2 Ad3 for both
            const V h = fma(As[0], As[1], As[2] * As[3]);
            const V x = fma(Ad[0], As[1], fma(Ad[2], As[3], fma(As[0], Ad[1], As[2] * Ad[3])));
            rs = fma(ch == 0 ? V(-0.5) : V(0.5), h, x);
            rd = fma(V(0.1875), h, fma(Ad[0], Ad[1], Ad[2] * Ad[3]));
          } else {
            // integralKernelT (:229-303) without the lattice bubble: the
            // site-local vertices enter as the factors built by the producer
            const V *lc = s_loc + ee * 8;
            rs = fma(lc[0], As[0], fma(lc[2], As[1], fma(lc[4], As[2], lc[6] * As[3])));
            rd = fma(lc[1], Ad[0], fma(lc[3], Ad[1], fma(lc[5], Ad[2], lc[7] * Ad[3])));
            // lattice-bubble factors: P = c A2, Q = A3 per channel
            s_P[(0 * nb + ee) * Lp + rid] = V(c) * As[2];
            s_P[(1 * nb + ee) * Lp + rid] = V(c) * Ad[2];
            s_Q[(0 * nb + ee) * Lp + rid] = As[3];
            s_Q[(1 * nb + ee) * Lp + rid] = Ad[3];
          }
          if (MIXED) {
            ksum_add(ks, float(c) * float(rs));
            ksum_add(kd, float(c) * float(rd));
          } else {
            acc_s = fma(c, double(rs), acc_s);
            acc_d = fma(c, double(rd), acc_d);
          }
        }
      }
    }
    __syncthreads();
    if (jc >= 0 && ch == 1) {
      // (C) C[ch][a][b] += sum_e P[ch][e][a] Q[ch][e][b] in FP64, TR x TC
      // register tiles (2 x 4 in the mixed form, 4.5 % faster than 2 x 2 at
      // small cutoffs although the occupancy-4 budget then spills 12 bytes
      // per lane); the mixed form's FP32 factors are widened, so each product
      // is exact and the sums are FP64 as in the FP64 form
      if (MIXED) {
        const int nr = Lp / 2, nc = Lp / 4;
        const int tiles = 2 * nr * nc;
        for (int jt = tid; jt < tiles; jt += NT) {
          const int cch = jt / (nr * nc);
          const int r = jt - cch * nr * nc;
          const int a = 2 * (r / nc), b = 4 * (r - (r / nc) * nc);
          const float *P = reinterpret_cast<const float *>(s_P) + cch * nb * Lp + a;
          const float *Q = reinterpret_cast<const float *>(s_Q) + cch * nb * Lp + b;
          double *Cb = s_C + cch * Lp * Lp + a * Lp + b;
          double2 m[2][2];
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) m[i][jj] = *reinterpret_cast<const double2 *>(Cb + i * Lp + 2 * jj);
          for (int e = 0; e < nvalid; ++e) {
            const float2 pv = *reinterpret_cast<const float2 *>(P + e * Lp);
            const float4 qv = *reinterpret_cast<const float4 *>(Q + e * Lp);
            const double p0 = pv.x, p1 = pv.y;
            const double q0 = qv.x, q1 = qv.y, q2 = qv.z, q3 = qv.w;
            m[0][0].x = fma(p0, q0, m[0][0].x);
            m[0][0].y = fma(p0, q1, m[0][0].y);
            m[0][1].x = fma(p0, q2, m[0][1].x);
            m[0][1].y = fma(p0, q3, m[0][1].y);
            m[1][0].x = fma(p1, q0, m[1][0].x);
            m[1][0].y = fma(p1, q1, m[1][0].y);
            m[1][1].x = fma(p1, q2, m[1][1].x);
            m[1][1].y = fma(p1, q3, m[1][1].y);
          }
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) *reinterpret_cast<double2 *>(Cb + i * Lp + 2 * jj) = m[i][jj];
        }
      } else {
        constexpr int TR = SPFRG_TILE_R, TC = SPFRG_TILE_C;
        const int nr = Lp / TR, nc = Lp / TC;
        const int tiles = 2 * nr * nc;
        for (int jt = tid; jt < tiles; jt += NT) {
          const int cch = jt / (nr * nc);
          const int r = jt - cch * nr * nc;
          const int a = TR * (r / nc), b = TC * (r - (r / nc) * nc);
          const double *P = reinterpret_cast<const double *>(s_P) + cch * nb * Lp + a;
          const double *Q = reinterpret_cast<const double *>(s_Q) + cch * nb * Lp + b;
          double *Cb = s_C + cch * Lp * Lp + a * Lp + b;
          double2 m[TR][TC / 2];
#pragma unroll
          for (int i = 0; i < TR; ++i)
#pragma unroll
            for (int jj = 0; jj < TC / 2; ++jj) m[i][jj] = *reinterpret_cast<const double2 *>(Cb + i * Lp + 2 * jj);
          for (int e = 0; e < nvalid; ++e) {
            double2 pv[TR / 2], qv[TC / 2];
#pragma unroll
            for (int i = 0; i < TR / 2; ++i) pv[i] = *reinterpret_cast<const double2 *>(P + e * Lp + 2 * i);
#pragma unroll
            for (int jj = 0; jj < TC / 2; ++jj) qv[jj] = *reinterpret_cast<const double2 *>(Q + e * Lp + 2 * jj);
#pragma unroll
            for (int i = 0; i < TR; ++i) {
              const double pi_ = (i & 1) ? pv[i >> 1].y : pv[i >> 1].x;
#pragma unroll
              for (int jj = 0; jj < TC / 2; ++jj) {
                m[i][jj].x = fma(pi_, qv[jj].x, m[i][jj].x);
                m[i][jj].y = fma(pi_, qv[jj].y, m[i][jj].y);
              }
            }
          }
#pragma unroll
          for (int i = 0; i < TR; ++i)
#pragma unroll
            for (int jj = 0; jj < TC / 2; ++jj) *reinterpret_cast<double2 *>(Cb + i * Lp + 2 * jj) = m[i][jj];
        }
      }
      __syncthreads();
      if (jc == cS + cT - 1) {
        // RPA(rid) = sum over getOverlap(rid) of C[rid1][rid2] (src/SU2/SU2FrgCore.cpp:250-266)
        const int wave = tid >> 6;
        for (int r = wave; r < L; r += NT / 64) {
          double ps = 0.0, pd = 0.0;
          const int b = F.ovl_off[r], e = F.ovl_off[r + 1];
          for (int i = b + lane; i < e; i += 64) {
            const int idx = F.ovl_idx[i];
            ps += s_C[idx];
            pd += s_C[Lp * Lp + idx];
          }
          ps = wave_sum(ps);
          pd = wave_sum(pd);
          if (lane == 0) { s_rpa[2 * r] = 2.0 * F.spin * ps; s_rpa[2 * r + 1] = 8.0 * F.spin * pd; }
        }
        __syncthreads();
      }
    }
  }

  // reduce the slots; prefactor 1/(2 pi) (src/SU2/SU2FrgCore.cpp:426-430)
  if (MIXED) {
    acc_s = double(ks.s) + double(ks.c);
    acc_d = double(kd.s) + double(kd.c);
  }
  if (active) { s_red[(slot * L + rid) * 2] = acc_s; s_red[(slot * L + rid) * 2 + 1] = acc_d; }
  __syncthreads();
  if (tid < L) {
    double ts = s_rpa[2 * tid], td = s_rpa[2 * tid + 1];
    for (int k = 0; k < slots; ++k) { ts += s_red[(k * L + tid) * 2]; td += s_red[(k * L + tid) * 2 + 1]; }
    ts *= (1.0 / kTwoPi);
    td *= (1.0 / kTwoPi);
    if (MIXED) {
      // undo the shadow's power-of-two scale (quadratic in the vertex)
      const int S2 = 2 * F.flags[4];
      ts = ldexp(ts, S2);
      td = ldexp(td, S2);
    }
    F.dG[(size_t)item * L + tid] = make_double2(ts, td);
    if (isnan(ts) || isnan(td)) atomicOr(&F.flags[0], 1);
  }
}

// The mixed form's FP32 shadow holds G * 2^-S.  Every term of dGamma is a
// product of two vertex values with Gamma-independent coefficients
// (src/SU2/SU2FrgCore.cpp:199-431), so the mixed dGamma kernel multiplies its
// result by 2^(2S); the power-of-two scale is exact.  S is planned from the
// largest binary exponent of G so that shadow values stay below 2^kShadowExp:
// a flow whose vertex grows towards a divergence never overflows FP32.  S = 0
// (no scaling) while |G| < 2^kShadowExp.
constexpr int kShadowExp = 16;
__device__ __forceinline__ int shadow_plan(int max_exp) {
  return max_exp > kShadowExp ? (max_exp - kShadowExp < 1000 ? max_exp - kShadowExp : 1000) : 0;
}
// unbiased binary exponent of the larger of |a|, |b| (1024 for inf / NaN)
__device__ __forceinline__ int pair_exp(double a, double b) {
  const int ea = (int)((__double_as_longlong(a) >> 52) & 0x7ff);
  const int eb = (int)((__double_as_longlong(b) >> 52) & 0x7ff);
  return (ea > eb ? ea : eb) - 1023;
}
// wave maximum of v into *dst.  Exponents up to kShadowExp plan no scaling,
// so only larger ones are recorded, and only when they exceed the value
// already there: a flow of ordinary magnitude issues no atomic at all, where
// one atomic per wave on the one flag address serialized the Euler update
// (about 1 ms per step at 3 M vertex entries)
__device__ __forceinline__ void atomic_max_wave(int *dst, int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  if ((threadIdx.x & 63) == 0 && v > kShadowExp && v > __atomic_load_n(dst, __ATOMIC_RELAXED)) atomicMax(dst, v);
}

// in-place explicit Euler step (src/SU2/SU2FrgCore.cpp:111-137); a flow whose
// derivative holds a NaN is left untouched, as the reference stops before
// finalizeStep (src/SpinParser.cpp:151-159), and flags[2] records that it
// stopped: later steps enqueued behind it (sp_run_flow) leave it untouched
// too.  A flow in the mixed form also rewrites the FP32 shadow with the
// planned scale and records the largest exponent of the new G.
__global__ void k_euler(const FlowDev *__restrict__ fl) {
  const FlowDev &F = fl[blockIdx.y];
  if (!F.euler) return;
  if (F.flags[0] != 0 || F.flags[2] != 0) {
    if (F.flags[0] != 0 && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&F.flags[2], 1);
    return;
  }
  const double step = F.step;
  const size_t n = (size_t)F.nv;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  double2 *G = F.G;
  const double2 *dG = F.dG;
  float2 *G32 = F.precision == SP_PRECISION_MIXED ? F.G32 : nullptr;
  const int S = G32 ? F.flags[6] : 0;
  int emax = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double2 a = G[i];
    const double2 d = dG[i];
    a.x = fma(step, d.x, a.x);
    a.y = fma(step, d.y, a.y);
    G[i] = a;
    if (G32) {
      G32[i] = make_float2((float)ldexp(a.x, -S), (float)ldexp(a.y, -S));
      const int e = pair_exp(a.x, a.y);
      emax = e > emax ? e : emax;
    }
  }
  if (G32) {
    atomic_max_wave(&F.flags[5], emax);
    if (blockIdx.x == 0 && threadIdx.x == 0) F.flags[4] = S;
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < F.nw; i += blockDim.x) F.v2[i] = fma(step, F.dv2[i], F.v2[i]);
}

// scale plan for the next shadow write from the exponents recorded since the
// last plan (one thread per flow, after k_euler / k_maxexp)
__global__ void k_plan(const FlowDev *__restrict__ fl) {
  const FlowDev &F = fl[blockIdx.y];
  if (threadIdx.x != 0 || !F.G32) return;
  F.flags[6] = shadow_plan(F.flags[5]);
  F.flags[5] = 0;
}

// largest binary exponent of G (before the shadow of a state is rebuilt)
__global__ void k_maxexp(const FlowDev *__restrict__ fl) {
  const FlowDev &F = fl[blockIdx.y];
  if (!F.G32) return;
  const size_t n = (size_t)F.nv;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  int emax = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int e = pair_exp(F.G[i].x, F.G[i].y);
    emax = e > emax ? e : emax;
  }
  atomic_max_wave(&F.flags[5], emax);
}

// FP32 shadow of the current state with the planned scale (mixed precision)
__global__ void k_shadow(const FlowDev *__restrict__ fl) {
  const FlowDev &F = fl[blockIdx.y];
  if (!F.G32) return;
  const int S = F.flags[6];
  const size_t n = (size_t)F.nv;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    F.G32[i] = make_float2((float)ldexp(F.G[i].x, -S), (float)ldexp(F.G[i].y, -S));
  if (blockIdx.x == 0 && threadIdx.x == 0) F.flags[4] = S;
}

// ---------------------------------------------------------------------------
// static susceptibility (src/SU2/SU2MeasurementCorrelation.cpp:77-177), nu = 0
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_corr_partial(const FlowDev *__restrict__ fl) {
  __shared__ double red[kThreads * 2];
  extern __shared__ double s_cmesh[];  // the mesh (dynamic, nw doubles)
  const FlowDev &F = fl[blockIdx.y];
  const int o = blockIdx.x;
  const int n = *F.Cn;
  if (o >= n) return;
  const int nw = F.nw, L = F.L;
  for (int k = threadIdx.x; k < nw; k += kThreads) s_cmesh[k] = F.mesh[k];
  __syncthreads();
  const double *m = s_cmesh, *v2 = F.v2;
  const double2 *G = F.G;
  const double spin = F.spin;
  const int slots = kThreads / L;
  const int tid = threadIdx.x;
  const int slot = tid / L, rid = tid - (tid / L) * L;
  const bool active = slot < slots;
  const double nu = 0.0;
  const double w = F.Cwp[o];
  const double gw = (w + sigma_at(m, nw, v2, w)) * (w + nu + sigma_at(m, nw, v2, w + nu));
  double acc_s = 0.0, acc_d = 0.0;
  if (active) {
    for (int e = slot; e < n; e += slots) {
      const double wp = F.Cwp[e];
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
      interp_sel(m, nw, a, F.geo, F.ga, F.gb, lS, uS, bS);
      interp_sel(m, nw, b, F.geo, F.ga, F.gb, lT, uT, bT);
      interp_sel(m, nw, c, F.geo, F.ga, F.gb, lU, uU, bU);
      const int sr = exch ? F.inv_rid[rid] : F.site_rid[rid];
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
        const double2 vv = value_none(G, m, nw, L, 0, 0, w + wp + nu, w - wp, nu, F.geo, F.ga, F.gb);
        rs += spin * (-vv.x / 4.0 + vv.y);
        rd += spin * (3.0 * vv.x + 4.0 * vv.y);
      }
      const double norm = 1.0 / (gw * (wp + sigma_at(m, nw, v2, wp)) * (wp + nu + sigma_at(m, nw, v2, wp + nu)) *
                                 (4.0 * kPi * kPi));
      const double cf = F.Cw[e] * norm;
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
    F.corr_partial[((size_t)o * L + tid) * 2] = F.Cw[o] * ts;
    F.corr_partial[((size_t)o * L + tid) * 2 + 1] = F.Cw[o] * td;
  }
}

__global__ void k_corr_final(const FlowDev *__restrict__ fl) {
  __shared__ double susc[2 * 256];
  const FlowDev &F = fl[blockIdx.y];
  const int n = *F.Cn;
  const int L = F.L;
  for (int r = threadIdx.x; r < L; r += blockDim.x) {
    double s = 0.0, d = 0.0;
    for (int o = 0; o < n; ++o) {
      s += F.corr_partial[((size_t)o * L + r) * 2];
      d += F.corr_partial[((size_t)o * L + r) * 2 + 1];
    }
    susc[2 * r] = s;
    susc[2 * r + 1] = d;
  }
  __syncthreads();
  const int nout = F.B * F.Rb;
  for (int k = threadIdx.x; k < nout; k += blockDim.x) {
    F.corr_out[k] = susc[2 * F.corr_rid[k]];
    F.corr_out[nout + k] = susc[2 * F.corr_rid[k] + 1];
  }
  if (threadIdx.x == 0) {
    // the flags of the step after the record, so that one copy returns both
    const int2 fl = make_int2(F.flags[0], F.flags[1]);
    double packed;
    __builtin_memcpy(&packed, &fl, 8);
    F.corr_out[2 * nout] = packed;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
constexpr int kRing = 8;       // pinned staging slots of the descriptor uploads
constexpr int kRunBlock = 16;  // sp_run_flow: steps enqueued between two host checks

struct sp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  FlowDev fd{};             // host copy of the flow descriptor
  FlowDev *d_fd = nullptr;  // device copy used by single-flow launches
  // descriptor uploads go through pinned slots, so that they are truly
  // asynchronous; a slot is refilled only after the copy that read it ran
  FlowDev *h_ring = nullptr;
  hipEvent_t ring_ev[kRing] = {};
  int ring_pos = 0;
  // pinned copies of the last measurement and of the flags taken with it:
  // isDiverged right after a measurement needs no second round trip
  double *h_meas = nullptr;
  int *h_flags = nullptr;
  bool flags_cached = false;
  double list_cut = -1.0;   // cutoff of the correlation point list on the device
  double *h_run = nullptr;  // pinned records of one block of sp_run_flow steps
  double cutoff = 0.0;
  bool have_flow = false;
  size_t lds = 0;
  std::vector<void *> allocs;
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[4];
  std::vector<hipEvent_t> free_events;
  double prof_ms[4] = {0, 0, 0, 0};
  long prof_n[4] = {0, 0, 0, 0};
  std::string err;
};

struct sp_group {
  int capacity = 0;
  hipStream_t stream = nullptr;
  FlowDev *d_fd = nullptr;
  double *d_out = nullptr;  // per flow: 2*B*Rb correlation values, then 2 flag slots
  double *h_out = nullptr;  // pinned mirror of d_out
  size_t out_cap = 0;       // doubles
  std::vector<FlowDev> h_fd;
  std::string err;
};

namespace {

int fail(sp_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}
int gfail(sp_group *g, int code, const std::string &msg) {
  if (g) g->err = msg;
  return code;
}

#define SP_HIP(call)                                                                                    \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess) return fail(ctx, SP_EHIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

#define SP_GHIP(call)                                                                                    \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) return gfail(grp, SP_EHIP, std::string(#call ": ") + hipGetErrorString(e_)); \
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
    if (ctx && ctx->prof) {
      a = get_event(ctx);
      b = get_event(ctx);
      if (a && b) (void)hipEventRecord(a, ctx->stream);
    }
  }
  ~ProfScope() {
    if (ctx && ctx->prof && a && b) {
      (void)hipEventRecord(b, ctx->stream);
      ctx->pending[k].push_back({a, b});
    }
  }
};

int check_problem(sp_ctx *ctx, const sp_su2_problem *p) {
  if (!p || p->nw < 2 || !p->mesh) return fail(ctx, SP_EARG, "invalid frequency mesh");
  // the dGamma kernel stages a channel's 2 N_w + 8 evaluation points with two
  // loads per thread
  if (2 * p->nw + 8 > 2 * kThreads) return fail(ctx, SP_EARG, "at most 252 positive frequencies are supported");
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
  if (p->precision != SP_PRECISION_FP64 && p->precision != SP_PRECISION_MIXED)
    return fail(ctx, SP_EARG, "unsupported precision mode");
  return SP_OK;
}

// Workgroups per CU that the registers of a dGamma kernel allow: one wave per
// SIMD each for the 256-thread shape, so SPFRG_MIN_BLOCKS (FP64, 4) or
// SPFRG_MIN_BLOCKS_MIXED (mixed form, 5: it needs fewer registers; the mixed
// 256-thread kernel is also compiled for 4, which flows whose LDS plan needs
// more than 160 KB / 5 run, since its code is 1-3 % faster at L = 31).  The
// 128-thread shape of small lattices is compiled for 4 in both forms.
constexpr int resident_blocks(int vs, int nt) {
  return (vs == 4 && nt != kSmallThreads) ? SPFRG_MIN_BLOCKS_MIXED : SPFRG_MIN_BLOCKS;
}
// LDS budget of one dGamma workgroup: a quarter of the CU's 160 KB keeps four
// workgroups resident.  The 128-thread workgroups of small lattices keep it
// too: more points per chunk at six or seven workgroups per CU measured 2-6 %
// faster than the eighth of the LDS that eight resident workgroups would allow
// (L = 9, 11, 16).  Where the registers allow more workgroups per CU, the
// plan takes the smaller share 160 KB / resident_blocks if its chunks still
// hold at least three points per worker slot (mixed form, L = 23: 24 instead
// of 32 points, 3-14 % faster); with fewer (L = 31: 12 instead of 24) the
// larger chunks at four workgroups per CU were faster at small cutoffs.  A
// flow whose plan ends up within the smaller share runs the kernel compiled
// for more workgroups, any other the one compiled for four.
constexpr size_t kLdsBudget = 160 * 1024 / 4;
#ifndef SPFRG_MIN_PER_SLOT_HIGH
#define SPFRG_MIN_PER_SLOT_HIGH 3
#endif
constexpr int kMinPerSlotHigh = SPFRG_MIN_PER_SLOT_HIGH;

// worker slots (all waves but the last) and points per chunk of a flow whose
// dGamma workgroups have nt threads; the producer wave holds one lane per
// (point, pair), so nb <= 32 (points per slot and chunk shrink while the LDS
// plan exceeds the budget)
void launch_shape(int nw, int L, int vs, int nt, int &slots, int &nb, size_t &lds) {
  slots = (nt - 64) / L;
  const int maxb = vs == 4 ? kMaxBatchMixed : kMaxBatch;
  const int per0 = maxb / slots > 1 ? maxb / slots : 1;
  const int rb = resident_blocks(vs, nt);
  if (rb > 4) {
    for (int per = per0; per >= kMinPerSlotHigh; --per) {
      nb = slots * per < 32 ? slots * per : 32;
      lds = (size_t)lds_layout(nw, L, nb, slots, vs).total;
      if (lds <= (size_t)160 * 1024 / rb) return;
    }
  }
  for (int per = per0;; --per) {
    nb = slots * per < 32 ? slots * per : 32;
    lds = (size_t)lds_layout(nw, L, nb, slots, vs).total;
    if (lds <= kLdsBudget || per == 1) break;
  }
}

size_t lds_of(const FlowDev &f) {
  return (size_t)lds_layout(f.nw, f.L, f.nb, f.slots, f.precision == SP_PRECISION_MIXED ? 4 : 8).total;
}

// Flows of one dGamma launch share its workgroup size and its dynamic LDS
// size, the largest of them; a flow whose plan fits a budget must not run at
// the occupancy of a larger one.  Launch class = precision x workgroup size x
// LDS level (within 160 KB / resident_blocks, within kLdsBudget, above); a
// batch issues one launch per class.
constexpr int kLaunchClasses = 12;
int lds_level(const FlowDev &f) {
  const int vs = f.precision == SP_PRECISION_MIXED ? 4 : 8;
  const size_t lds = lds_of(f);
  return lds <= (size_t)160 * 1024 / resident_blocks(vs, f.nt) ? 0 : (lds <= kLdsBudget ? 1 : 2);
}
int launch_class(const FlowDev &f) {
  return ((f.precision == SP_PRECISION_MIXED ? 2 : 0) + (f.nt == kSmallThreads ? 1 : 0)) * 3 + lds_level(f);
}

// Launch geometry of a set of flows held in a host array (h) whose device copy is d.
struct Geometry {
  int maxnw = 0, maxemax = 0;
  long long maxnv = 0;
};

Geometry geometry(const FlowDev *h, int n) {
  Geometry g;
  for (int i = 0; i < n; ++i) {
    g.maxnw = std::max(g.maxnw, h[i].nw);
    g.maxemax = std::max(g.maxemax, h[i].emax);
    g.maxnv = std::max(g.maxnv, h[i].nv);
  }
  return g;
}

// dSigma, weight tables and dGamma for n flows; the flows come ordered by
// launch_class and h[i].item_base is relative to the first flow of its class.
hipError_t launch_compute(hipStream_t st, const FlowDev *d, const FlowDev *h, int n, sp_ctx *prof) {
  const Geometry geo = geometry(h, n);
  {
    ProfScope ps(prof, 1);
    hipLaunchKernelGGL(k_dsigma, dim3(geo.maxnw, n), dim3(256), geo.maxnw * sizeof(double), st, d);
  }
  {
    ProfScope ps(prof, 2);
    hipLaunchKernelGGL(k_weights, dim3(geo.maxnw + 1, n), dim3(64), geo.maxnw * sizeof(double), st, d, 0);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  ProfScope ps(prof, 0);
  int i0 = 0;
  while (i0 < n) {
    const int cls = launch_class(h[i0]);
    const bool mixed = h[i0].precision == SP_PRECISION_MIXED;
    int i1 = i0;
    long long total = 0;
    size_t lds = 0;
    while (i1 < n && launch_class(h[i1]) == cls) {
      total += h[i1].nf;
      lds = std::max(lds, lds_of(h[i1]));
      ++i1;
    }
    {
      const long long per_xcd = 0;  // unused by the chunked mapping
      const long long rounds = (total + 8 * kChunk - 1) / (8 * kChunk);
      const unsigned grid = (unsigned)(rounds * 8 * kChunk);
      const int nt = h[i0].nt;
      // plans within 160 KB / resident_blocks run the kernel compiled for
      // that many resident workgroups
      const bool high = lds_level(h[i0]) == 0 && resident_blocks(mixed ? 4 : 8, nt) > SPFRG_MIN_BLOCKS;
      if (mixed && nt == kSmallThreads)
        hipLaunchKernelGGL((k_dgamma<true, kSmallThreads, SPFRG_MIN_BLOCKS>), dim3(grid), dim3(nt), lds, st, d + i0,
                           i1 - i0, total, per_xcd);
      else if (mixed && high)
        hipLaunchKernelGGL((k_dgamma<true, kThreads, SPFRG_MIN_BLOCKS_MIXED>), dim3(grid), dim3(nt), lds, st, d + i0,
                           i1 - i0, total, per_xcd);
      else if (mixed)
        hipLaunchKernelGGL((k_dgamma<true, kThreads, SPFRG_MIN_BLOCKS>), dim3(grid), dim3(nt), lds, st, d + i0,
                           i1 - i0, total, per_xcd);
      else if (nt == kSmallThreads)
        hipLaunchKernelGGL((k_dgamma<false, kSmallThreads, SPFRG_MIN_BLOCKS>), dim3(grid), dim3(nt), lds, st, d + i0,
                           i1 - i0, total, per_xcd);
      else
        hipLaunchKernelGGL((k_dgamma<false, kThreads, SPFRG_MIN_BLOCKS>), dim3(grid), dim3(nt), lds, st, d + i0,
                           i1 - i0, total, per_xcd);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    i0 = i1;
  }
  return hipSuccess;
}

// build_list: rebuild the correlation point list; a dGamma launch at the same
// cutoff (k_weights over x <= nw) has built it already otherwise
hipError_t launch_measure(hipStream_t st, const FlowDev *d, const FlowDev *h, int n, bool build_list) {
  const Geometry geo = geometry(h, n);
  if (build_list) hipLaunchKernelGGL(k_weights, dim3(1, n), dim3(64), geo.maxnw * sizeof(double), st, d, 1);
  hipLaunchKernelGGL(k_corr_partial, dim3(geo.maxemax, n), dim3(kThreads), geo.maxnw * sizeof(double), st, d);
  hipLaunchKernelGGL(k_corr_final, dim3(1, n), dim3(256), 0, st, d);
  return hipGetLastError();
}

hipError_t launch_euler(hipStream_t st, const FlowDev *d, const FlowDev *h, int n, sp_ctx *prof) {
  const Geometry geo = geometry(h, n);
  long long blocks = (geo.maxnv + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  ProfScope ps(prof, 3);
  hipLaunchKernelGGL(k_euler, dim3((unsigned)blocks, n), dim3(256), 0, st, d);
  hipLaunchKernelGGL(k_plan, dim3(1, n), dim3(64), 0, st, d);
  return hipGetLastError();
}

int upload_fd(sp_ctx *ctx) {
  const int s = ctx->ring_pos;
  ctx->ring_pos = (s + 1) % kRing;
  SP_HIP(hipEventSynchronize(ctx->ring_ev[s]));
  ctx->h_ring[s] = ctx->fd;
  SP_HIP(hipMemcpyAsync(ctx->d_fd, ctx->h_ring + s, sizeof(FlowDev), hipMemcpyHostToDevice, ctx->stream));
  SP_HIP(hipEventRecord(ctx->ring_ev[s], ctx->stream));
  return SP_OK;
}

// FP32 shadow of the current state (mixed form): largest exponent of G ->
// scale plan -> shadow; a context in the FP64 form keeps its shadow stale
// until it returns to the mixed form
int rebuild_shadow(sp_ctx *ctx) {
  const FlowDev &f = ctx->fd;
  if (!f.G32 || f.precision != SP_PRECISION_MIXED) return SP_OK;
  SP_HIP(hipMemsetAsync(f.flags + 5, 0, sizeof(int), ctx->stream));
  hipLaunchKernelGGL(k_maxexp, dim3(1024, 1), dim3(256), 0, ctx->stream, ctx->d_fd);
  hipLaunchKernelGGL(k_plan, dim3(1, 1), dim3(64), 0, ctx->stream, ctx->d_fd);
  hipLaunchKernelGGL(k_shadow, dim3(2048, 1), dim3(256), 0, ctx->stream, ctx->d_fd);
  SP_HIP(hipGetLastError());
  return SP_OK;
}

}  // namespace

extern "C" {

int sp_ctx_footprint(const sp_su2_problem *p, long long *bytes) {
  if (!p || !bytes || p->nw < 2 || p->L < 1) return SP_EARG;
  const long long nf = (long long)p->nw * p->nw * (p->nw + 1) / 2;
  const long long emax = 2LL * p->nw + 8;
  long long b = 2 * nf * p->L * 16;  // Gamma, dGamma
  if (p->precision == SP_PRECISION_MIXED) b += nf * p->L * 8;  // FP32 shadow
  b += 3LL * p->nw * 8;                                          // Sigma, dSigma, mesh
  b += 2 * (long long)p->nw * emax * 8 + 2 * emax * 8 + (p->nw + 1) * 4;  // weight tables
  b += emax * p->L * 16 + 2LL * p->B * p->Rb * 8 + 32;                   // correlation scratch / output, flags
  b += 4LL * (3 * p->L + 1 + 2 * (p->ovl_off ? p->ovl_off[p->L] : 0) + 2 * p->R + p->B * p->Rb);
  b += 2LL * (p->ovl_off ? p->ovl_off[p->L] : 0);                         // packed overlap indices
  b += (long long)sizeof(FlowDev);
  *bytes = b;
  return SP_OK;
}

int sp_ctx_create(const sp_su2_problem *p, sp_ctx **out) {
  if (!out) return SP_EARG;
  *out = nullptr;
  sp_ctx *ctx = new (std::nothrow) sp_ctx();
  if (!ctx) return SP_ENOMEM;
  *out = ctx;
  int rc = check_problem(ctx, p);
  if (rc != SP_OK) return rc;
  SP_HIP(hipGetDevice(&ctx->device));
  SP_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  FlowDev &f = ctx->fd;
  f.nw = p->nw;
  f.L = p->L;
  f.R = p->R;
  f.B = p->B;
  f.Rb = p->Rb;
  f.nf = (long long)p->nw * p->nw * (p->nw + 1) / 2;
  f.nv = f.nf * p->L;
  f.emax = 2 * p->nw + 8;
  f.spin = p->spin;
  f.precision = p->precision;
  f.item_base = 0;
  f.nt = dgamma_threads(f.L);
  launch_shape(f.nw, f.L, f.precision == SP_PRECISION_MIXED ? 4 : 8, f.nt, f.slots, f.nb, ctx->lds);
  int max_lds = 0;
  SP_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
  if (ctx->lds > (size_t)max_lds) return fail(ctx, SP_EARG, "problem too large for the LDS plan of the dGamma kernel");
  if (f.precision == SP_PRECISION_MIXED) {
    // a mixed context may switch to the FP64 form (sp_set_precision)
    int s64, nb64;
    size_t lds64;
    launch_shape(f.nw, f.L, 8, f.nt, s64, nb64, lds64);
    if (lds64 > (size_t)max_lds) return fail(ctx, SP_EARG, "problem too large for the LDS plan of the dGamma kernel");
  }
  if (f.nv * 16 >= (1LL << 31)) return fail(ctx, SP_EARG, "vertex too large for 32-bit byte offsets");
  f.g_bytes = (unsigned)(f.nv * 16);
  f.g32_bytes = (unsigned)(f.nv * 8);
  const int O = p->ovl_off[p->L];
  const size_t nv = (size_t)f.nv;
  const int emax = f.emax;
  double *mesh = nullptr;
  int *site = nullptr, *inv = nullptr, *ooff = nullptr, *o1 = nullptr, *o2 = nullptr, *rf = nullptr, *ri = nullptr,
      *corr = nullptr;
  unsigned short *oidx = nullptr;
  if ((rc = dalloc(ctx, &oidx, O))) return rc;
  if ((rc = dalloc(ctx, &mesh, f.nw)) || (rc = dalloc(ctx, &site, f.L)) || (rc = dalloc(ctx, &inv, f.L)) ||
      (rc = dalloc(ctx, &ooff, f.L + 1)) || (rc = dalloc(ctx, &o1, O)) || (rc = dalloc(ctx, &o2, O)) ||
      (rc = dalloc(ctx, &rf, f.R)) || (rc = dalloc(ctx, &ri, f.R)) || (rc = dalloc(ctx, &corr, (size_t)f.B * f.Rb)) ||
      (rc = dalloc(ctx, &f.G, nv)) || (rc = dalloc(ctx, &f.dG, nv)) || (rc = dalloc(ctx, &f.v2, f.nw)) ||
      (rc = dalloc(ctx, &f.dv2, f.nw)) || (rc = dalloc(ctx, &f.Wwp, (size_t)f.nw * emax)) ||
      (rc = dalloc(ctx, &f.Wc, (size_t)f.nw * emax)) || (rc = dalloc(ctx, &f.Wn, f.nw)) ||
      (rc = dalloc(ctx, &f.Cwp, emax)) || (rc = dalloc(ctx, &f.Cw, emax)) || (rc = dalloc(ctx, &f.Cn, 1)) ||
      (rc = dalloc(ctx, &f.corr_partial, (size_t)emax * f.L * 2)) ||
      (rc = dalloc(ctx, &f.corr_out, (size_t)2 * f.B * f.Rb + 1)) || (rc = dalloc(ctx, &f.flags, 8)) ||
      (rc = dalloc(ctx, &ctx->d_fd, 1)))
    return rc;
  {
    void *q = nullptr;
    SP_HIP(hipHostMalloc(&q, kRing * sizeof(FlowDev), 0));
    ctx->h_ring = static_cast<FlowDev *>(q);
    SP_HIP(hipHostMalloc(&q, ((size_t)2 * f.B * f.Rb + 1) * 8, 0));
    ctx->h_meas = static_cast<double *>(q);
    SP_HIP(hipHostMalloc(&q, 16, 0));
    ctx->h_flags = static_cast<int *>(q);
    SP_HIP(hipHostMalloc(&q, (size_t)kRunBlock * (2 * (size_t)f.B * f.Rb + 2) * 8, 0));
    ctx->h_run = static_cast<double *>(q);
    for (int k = 0; k < kRing; ++k) SP_HIP(hipEventCreateWithFlags(&ctx->ring_ev[k], hipEventDisableTiming));
  }
  if (f.precision == SP_PRECISION_MIXED && (rc = dalloc(ctx, &f.G32, nv))) return rc;
  f.mesh = mesh;
  f.site_rid = site;
  f.inv_rid = inv;
  f.ovl_off = ooff;
  f.ovl_rid1 = o1;
  f.ovl_rid2 = o2;
  f.ovl_idx = oidx;
  f.range_fwd = rf;
  f.range_inv = ri;
  f.corr_rid = corr;

  std::vector<double> m(f.nw);
  for (int i = 0; i < f.nw; ++i) m[i] = (double)p->mesh[i];
  geo_mesh_params(m.data(), f.nw, f.geo, f.ga, f.gb);
  SP_HIP(hipMemcpy(mesh, m.data(), f.nw * 8, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(site, p->site_rid, f.L * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(inv, p->inv_rid, f.L * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(ooff, p->ovl_off, (f.L + 1) * 4, hipMemcpyHostToDevice));
  if (O > 0) {
    SP_HIP(hipMemcpy(o1, p->ovl_rid1, (size_t)O * 4, hipMemcpyHostToDevice));
    SP_HIP(hipMemcpy(o2, p->ovl_rid2, (size_t)O * 4, hipMemcpyHostToDevice));
    // packed C-matrix index of every overlap entry (rid < L <= 64, so it fits 16 bits)
    const int Lp = (f.L + 3) & ~3;
    std::vector<unsigned short> idx(O);
    for (int i = 0; i < O; ++i) idx[i] = (unsigned short)(p->ovl_rid1[i] * Lp + p->ovl_rid2[i]);
    SP_HIP(hipMemcpy(oidx, idx.data(), (size_t)O * 2, hipMemcpyHostToDevice));
  }
  SP_HIP(hipMemcpy(rf, p->range_fwd, (size_t)f.R * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(ri, p->range_inv, (size_t)f.R * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemcpy(corr, p->corr_rid, (size_t)f.B * f.Rb * 4, hipMemcpyHostToDevice));
  SP_HIP(hipMemset(f.G, 0, nv * sizeof(double2)));
  SP_HIP(hipMemset(f.dG, 0, nv * sizeof(double2)));
  if (f.G32) SP_HIP(hipMemset(f.G32, 0, nv * sizeof(float2)));
  SP_HIP(hipMemset(f.v2, 0, f.nw * 8));
  SP_HIP(hipMemset(f.dv2, 0, f.nw * 8));
  SP_HIP(hipMemset(f.flags, 0, 32));
  SP_HIP(hipMemset(f.Wn, 0, f.nw * 4));
  SP_HIP(hipMemset(f.Cn, 0, 4));
  SP_HIP(hipFuncSetAttribute((const void *)k_dgamma<false, kThreads, SPFRG_MIN_BLOCKS>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  SP_HIP(hipFuncSetAttribute((const void *)k_dgamma<true, kThreads, SPFRG_MIN_BLOCKS>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  SP_HIP(hipFuncSetAttribute((const void *)k_dgamma<true, kThreads, SPFRG_MIN_BLOCKS_MIXED>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  SP_HIP(hipFuncSetAttribute((const void *)k_dgamma<false, kSmallThreads, SPFRG_MIN_BLOCKS>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  SP_HIP(hipFuncSetAttribute((const void *)k_dgamma<true, kSmallThreads, SPFRG_MIN_BLOCKS>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  SP_HIP(hipMemcpy(ctx->d_fd, &f, sizeof(FlowDev), hipMemcpyHostToDevice));
  return SP_OK;
}

int sp_ctx_destroy(sp_ctx *ctx) {
  if (!ctx) return SP_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (int k = 0; k < 4; ++k)
    for (auto &pr : ctx->pending[k]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
  for (auto e : ctx->free_events) (void)hipEventDestroy(e);
  for (int k = 0; k < kRing; ++k)
    if (ctx->ring_ev[k]) (void)hipEventDestroy(ctx->ring_ev[k]);
  if (ctx->h_ring) (void)hipHostFree(ctx->h_ring);
  if (ctx->h_meas) (void)hipHostFree(ctx->h_meas);
  if (ctx->h_flags) (void)hipHostFree(ctx->h_flags);
  if (ctx->h_run) (void)hipHostFree(ctx->h_run);
  for (void *q : ctx->allocs) (void)hipFree(q);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SP_OK;
}

int sp_set_state(sp_ctx *ctx, double cutoff, const double *v2, const double *ss, const double *dd) {
  if (!ctx || !v2 || !ss || !dd) return fail(ctx, SP_EARG, "null argument");
  if (!std::isfinite(cutoff) || cutoff <= 0.0) return fail(ctx, SP_EARG, "cutoff must be finite and positive");
  FlowDev &f = ctx->fd;
  const size_t nv = (size_t)f.nv;
  std::vector<double2> buf;
  try { buf.resize(nv); } catch (...) { return fail(ctx, SP_ENOMEM, "host buffer"); }
  for (size_t i = 0; i < nv; ++i) buf[i] = make_double2(ss[i], dd[i]);
  SP_HIP(hipMemcpyAsync(f.G, buf.data(), nv * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
  SP_HIP(hipMemcpyAsync(f.v2, v2, f.nw * 8, hipMemcpyHostToDevice, ctx->stream));
  ctx->cutoff = cutoff;
  f.cut = cutoff;
  int rc = upload_fd(ctx);
  if (rc) return rc;
  if ((rc = rebuild_shadow(ctx))) return rc;
  SP_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_flow = false;
  ctx->flags_cached = false;
  return SP_OK;
}

int sp_set_precision(sp_ctx *ctx, int precision) {
  if (!ctx) return SP_EARG;
  if (precision != SP_PRECISION_FP64 && precision != SP_PRECISION_MIXED)
    return fail(ctx, SP_EARG, "unsupported precision mode");
  FlowDev &f = ctx->fd;
  if (precision == SP_PRECISION_MIXED && !f.G32)
    return fail(ctx, SP_EARG, "the context holds no FP32 shadow (create it with SP_PRECISION_MIXED)");
  if (precision == f.precision) return SP_OK;
  f.precision = precision;
  launch_shape(f.nw, f.L, precision == SP_PRECISION_MIXED ? 4 : 8, f.nt, f.slots, f.nb, ctx->lds);
  int rc = upload_fd(ctx);
  if (rc) return rc;
  return rebuild_shadow(ctx);
}

static int download_pair(sp_ctx *ctx, const double2 *src, double *a, double *b) {
  const size_t nv = (size_t)ctx->fd.nv;
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
    SP_HIP(hipMemcpyAsync(v2, ctx->fd.v2, ctx->fd.nw * 8, hipMemcpyDeviceToHost, ctx->stream));
    SP_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (ss || dd) return download_pair(ctx, ctx->fd.G, ss, dd);
  return SP_OK;
}

int sp_get_flow(sp_ctx *ctx, double *dv2, double *dss, double *ddd) {
  if (!ctx) return SP_EARG;
  if (!ctx->have_flow) return fail(ctx, SP_ESTATE, "no flow computed since the last state change");
  if (dv2) {
    SP_HIP(hipMemcpyAsync(dv2, ctx->fd.dv2, ctx->fd.nw * 8, hipMemcpyDeviceToHost, ctx->stream));
    SP_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (dss || ddd) return download_pair(ctx, ctx->fd.dG, dss, ddd);
  return SP_OK;
}

int sp_compute_step(sp_ctx *ctx) {
  if (!ctx) return SP_EARG;
  FlowDev &f = ctx->fd;
  f.cut = ctx->cutoff;
  int rc = upload_fd(ctx);
  if (rc) return rc;
  SP_HIP(hipMemsetAsync(f.flags, 0, 16, ctx->stream));
  SP_HIP(launch_compute(ctx->stream, ctx->d_fd, &f, 1, ctx));
  ctx->have_flow = true;
  ctx->flags_cached = false;
  ctx->list_cut = f.cut;
  return SP_OK;
}

int sp_finalize_step(sp_ctx *ctx, double new_cutoff) {
  if (!ctx) return SP_EARG;
  if (!ctx->have_flow) return fail(ctx, SP_ESTATE, "finalize_step without a computed flow");
  if (!std::isfinite(new_cutoff) || new_cutoff <= 0.0) return fail(ctx, SP_EARG, "cutoff must be finite and positive");
  FlowDev &f = ctx->fd;
  f.step = new_cutoff - ctx->cutoff;
  f.euler = 1;
  f.cut = new_cutoff;
  int rc = upload_fd(ctx);
  if (rc) return rc;
  // the reference updates even a flow the caller did not check for NaN; the
  // kernel only skips on the flag, so clear it for this explicit call
  SP_HIP(hipMemsetAsync(f.flags, 0, 4, ctx->stream));
  SP_HIP(launch_euler(ctx->stream, ctx->d_fd, &f, 1, ctx));
  f.euler = 0;
  ctx->cutoff = new_cutoff;
  ctx->have_flow = false;
  ctx->flags_cached = false;
  return SP_OK;
}

int sp_is_diverged(sp_ctx *ctx, int *diverged) {
  if (!ctx || !diverged) return SP_EARG;
  if (!ctx->have_flow) return fail(ctx, SP_ESTATE, "no flow computed");
  if (!ctx->flags_cached) {
    SP_HIP(hipMemcpyAsync(ctx->h_flags, ctx->fd.flags, 16, hipMemcpyDeviceToHost, ctx->stream));
    SP_HIP(hipStreamSynchronize(ctx->stream));
    ctx->flags_cached = true;
  }
  if (ctx->h_flags[1]) return fail(ctx, SP_EARG, "evaluation list overflow");
  *diverged = ctx->h_flags[0] ? 1 : 0;
  return SP_OK;
}

int sp_measure_correlation(sp_ctx *ctx, double *zz, double *dd) {
  if (!ctx || !zz || !dd) return SP_EARG;
  FlowDev &f = ctx->fd;
  f.cut = ctx->cutoff;
  int rc = upload_fd(ctx);
  if (rc) return rc;
  SP_HIP(launch_measure(ctx->stream, ctx->d_fd, &f, 1, !(ctx->list_cut == f.cut)));
  ctx->list_cut = f.cut;
  const int nout = f.B * f.Rb;
  SP_HIP(hipMemcpyAsync(ctx->h_meas, f.corr_out, (2 * (size_t)nout + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
  SP_HIP(hipStreamSynchronize(ctx->stream));
  // the flags that k_corr_final stored after the record
  std::memcpy(ctx->h_flags, ctx->h_meas + 2 * nout, 8);
  ctx->flags_cached = ctx->have_flow;
  std::memcpy(zz, ctx->h_meas, (size_t)nout * 8);
  std::memcpy(dd, ctx->h_meas + nout, (size_t)nout * 8);
  return SP_OK;
}

int sp_run_flow(sp_ctx *ctx, const double *cutoffs, int ncut, double *zz, double *dd, int *nrec, int *diverged) {
  if (!ctx || !cutoffs || !zz || !dd || !nrec || !diverged || ncut < 1) return fail(ctx, SP_EARG, "null argument");
  if (!(cutoffs[0] == ctx->cutoff)) return fail(ctx, SP_EARG, "the state is not at the first cutoff");
  for (int k = 0; k < ncut; ++k)
    if (!std::isfinite(cutoffs[k]) || cutoffs[k] <= 0.0) return fail(ctx, SP_EARG, "cutoff must be finite and positive");
  FlowDev &f = ctx->fd;
  const size_t nout = (size_t)f.B * f.Rb, slot = 2 * nout + 2;
  const int nsteps = ncut - 1;
  *nrec = 0;
  *diverged = 0;
  ctx->flags_cached = false;
  SP_HIP(hipMemsetAsync(f.flags, 0, 16, ctx->stream));  // clears the stop flag
  for (int kb = 0; kb < nsteps; kb += kRunBlock) {
    const int ke = nsteps - kb < kRunBlock ? nsteps : kb + kRunBlock;
    for (int j = kb; j < ke; ++j) {
      // computeStep and takeMeasurements at cutoffs[j]
      f.cut = cutoffs[j];
      f.euler = 0;
      int rc = upload_fd(ctx);
      if (rc) return rc;
      SP_HIP(hipMemsetAsync(f.flags, 0, 8, ctx->stream));
      SP_HIP(launch_compute(ctx->stream, ctx->d_fd, &f, 1, ctx));
      SP_HIP(launch_measure(ctx->stream, ctx->d_fd, &f, 1, false));
      double *rec = ctx->h_run + (size_t)(j - kb) * slot;  // record, then the step's flags
      SP_HIP(hipMemcpyAsync(rec, f.corr_out, (2 * nout + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
      // finalizeStep(cutoffs[j + 1]): skipped, and the flow stopped, on a NaN
      f.step = cutoffs[j + 1] - cutoffs[j];
      f.euler = 1;
      f.cut = cutoffs[j + 1];
      rc = upload_fd(ctx);
      if (rc) return rc;
      SP_HIP(launch_euler(ctx->stream, ctx->d_fd, &f, 1, ctx));
      f.euler = 0;
    }
    SP_HIP(hipStreamSynchronize(ctx->stream));
    for (int j = kb; j < ke; ++j) {
      const double *rec = ctx->h_run + (size_t)(j - kb) * slot;
      int fl[2];
      std::memcpy(fl, rec + 2 * nout, 8);
      if (fl[1]) return fail(ctx, SP_EARG, "evaluation list overflow");
      std::memcpy(zz + (size_t)j * nout, rec, nout * 8);
      std::memcpy(dd + (size_t)j * nout, rec + nout, nout * 8);
      *nrec = j + 1;
      if (fl[0]) {
        // isDiverged after step j: the state stayed at cutoffs[j]; compute its
        // flow again so that the context is the one the step-by-step loop
        // leaves (flow of step j computed, not finalized)
        *diverged = 1;
        ctx->cutoff = cutoffs[j];
        f.cut = cutoffs[j];
        int rc = upload_fd(ctx);
        if (rc) return rc;
        SP_HIP(hipMemsetAsync(f.flags, 0, 16, ctx->stream));
        SP_HIP(launch_compute(ctx->stream, ctx->d_fd, &f, 1, ctx));
        SP_HIP(hipStreamSynchronize(ctx->stream));
        ctx->have_flow = true;
        ctx->list_cut = cutoffs[j];
        return SP_OK;
      }
    }
  }
  ctx->cutoff = cutoffs[nsteps];
  f.cut = ctx->cutoff;
  ctx->have_flow = false;
  ctx->list_cut = nsteps > 0 ? cutoffs[nsteps - 1] : ctx->list_cut;
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

// ---------------------------------------------------------------------------
// flow groups: one launch per kernel for many resident flows
// ---------------------------------------------------------------------------
int sp_group_create(int capacity, sp_group **out) {
  if (!out || capacity < 1 || capacity > 65535) return SP_EARG;
  *out = nullptr;
  sp_group *grp = new (std::nothrow) sp_group();
  if (!grp) return SP_ENOMEM;
  *out = grp;
  grp->capacity = capacity;
  grp->h_fd.resize(capacity);
  SP_GHIP(hipStreamCreateWithFlags(&grp->stream, hipStreamNonBlocking));
  void *q = nullptr;
  if (hipMalloc(&q, sizeof(FlowDev) * capacity) != hipSuccess) return gfail(grp, SP_ENOMEM, "hipMalloc");
  grp->d_fd = static_cast<FlowDev *>(q);
  return SP_OK;
}

int sp_group_destroy(sp_group *grp) {
  if (!grp) return SP_OK;
  if (grp->stream) (void)hipStreamSynchronize(grp->stream);
  if (grp->d_fd) (void)hipFree(grp->d_fd);
  if (grp->d_out) (void)hipFree(grp->d_out);
  if (grp->h_out) (void)hipHostFree(grp->h_out);
  if (grp->stream) (void)hipStreamDestroy(grp->stream);
  delete grp;
  return SP_OK;
}

const char *sp_group_last_error(const sp_group *grp) { return grp ? grp->err.c_str() : "null group"; }

static int group_prepare(sp_group *grp, sp_ctx *const *ctxs, int n, size_t *out_off) {
  if (n < 1 || n > grp->capacity || !ctxs) return gfail(grp, SP_EARG, "invalid flow count");
  // FP64 flows first, then mixed; item bases restart per precision class
  std::vector<int> order;
  for (int cls = 0; cls < kLaunchClasses; ++cls)
    for (int i = 0; i < n; ++i)
      if (ctxs[i] && launch_class(ctxs[i]->fd) == cls) order.push_back(i);
  if ((int)order.size() != n) return gfail(grp, SP_EARG, "null context in group");
  size_t need = 0;
  for (int i = 0; i < n; ++i) {
    out_off[i] = need;
    need += 2 * (size_t)ctxs[i]->fd.B * ctxs[i]->fd.Rb + 2;
  }
  if (need > grp->out_cap) {
    if (grp->d_out) (void)hipFree(grp->d_out);
    if (grp->h_out) (void)hipHostFree(grp->h_out);
    grp->d_out = nullptr;
    grp->h_out = nullptr;
    void *q = nullptr;
    if (hipMalloc(&q, need * 8) != hipSuccess) return gfail(grp, SP_ENOMEM, "hipMalloc");
    grp->d_out = static_cast<double *>(q);
    if (hipHostMalloc(&q, need * 8, 0) != hipSuccess) return gfail(grp, SP_ENOMEM, "hipHostMalloc");
    grp->h_out = static_cast<double *>(q);
    grp->out_cap = need;
  }
  long long base = 0;
  int prev = -1;
  for (int j = 0; j < n; ++j) {
    sp_ctx *c = ctxs[order[j]];
    SP_GHIP(hipStreamSynchronize(c->stream));
    c->list_cut = -1.0;  // the group's launches rebuild the flow's point lists
    FlowDev f = c->fd;
    if (launch_class(f) != prev) { base = 0; prev = launch_class(f); }
    f.item_base = base;
    base += f.nf;
    f.cut = c->cutoff;
    f.euler = 0;
    f.corr_out = grp->d_out + out_off[order[j]];
    grp->h_fd[j] = f;
  }
  // remember the order for the caller
  for (int j = 0; j < n; ++j) grp->h_fd[j].pad_index = order[j];
  return SP_OK;
}

int sp_group_step(sp_group *grp, sp_ctx *const *ctxs, int n, const double *new_cutoffs, double *const *zz,
                  double *const *dd, int *diverged) {
  if (!grp) return SP_EARG;
  std::vector<size_t> off(n > 0 ? n : 1);
  int rc = group_prepare(grp, ctxs, n, off.data());
  if (rc) return rc;
  for (int j = 0; j < n; ++j) {
    sp_ctx *c = ctxs[grp->h_fd[j].pad_index];
    SP_GHIP(hipMemsetAsync(c->fd.flags, 0, 16, grp->stream));
  }
  SP_GHIP(hipMemcpyAsync(grp->d_fd, grp->h_fd.data(), sizeof(FlowDev) * n, hipMemcpyHostToDevice, grp->stream));
  SP_GHIP(launch_compute(grp->stream, grp->d_fd, grp->h_fd.data(), n, nullptr));
  // the divergence flags land next to each flow's record (k_corr_final)
  SP_GHIP(launch_measure(grp->stream, grp->d_fd, grp->h_fd.data(), n, false));
  if (new_cutoffs) {
    for (int j = 0; j < n; ++j) {
      const int i = grp->h_fd[j].pad_index;
      const double nc = new_cutoffs[i];
      FlowDev &f = grp->h_fd[j];
      if (std::isfinite(nc) && nc > 0.0) {
        f.euler = 1;
        f.step = nc - ctxs[i]->cutoff;
      }
    }
    SP_GHIP(hipMemcpyAsync(grp->d_fd, grp->h_fd.data(), sizeof(FlowDev) * n, hipMemcpyHostToDevice, grp->stream));
    SP_GHIP(launch_euler(grp->stream, grp->d_fd, grp->h_fd.data(), n, nullptr));
  }
  size_t total = 0;
  for (int i = 0; i < n; ++i) total = std::max(total, off[i] + 2 * (size_t)ctxs[i]->fd.B * ctxs[i]->fd.Rb + 2);
  SP_GHIP(hipMemcpyAsync(grp->h_out, grp->d_out, total * 8, hipMemcpyDeviceToHost, grp->stream));
  SP_GHIP(hipStreamSynchronize(grp->stream));
  for (int j = 0; j < n; ++j) {
    const int i = grp->h_fd[j].pad_index;
    sp_ctx *c = ctxs[i];
    const size_t nout = (size_t)c->fd.B * c->fd.Rb;
    const double *rec = grp->h_out + off[i];
    if (zz && zz[i]) std::memcpy(zz[i], rec, nout * 8);
    if (dd && dd[i]) std::memcpy(dd[i], rec + nout, nout * 8);
    int fl[2];
    std::memcpy(fl, rec + 2 * nout, 8);
    if (fl[1]) return gfail(grp, SP_EARG, "evaluation list overflow");
    if (diverged) diverged[i] = fl[0] ? 1 : 0;
    c->have_flow = true;
    if (grp->h_fd[j].euler && !fl[0]) {
      c->cutoff = new_cutoffs[i];
      c->have_flow = false;
    }
  }
  return SP_OK;
}

int sp_group_measure(sp_group *grp, sp_ctx *const *ctxs, int n, double *const *zz, double *const *dd) {
  if (!grp) return SP_EARG;
  std::vector<size_t> off(n > 0 ? n : 1);
  int rc = group_prepare(grp, ctxs, n, off.data());
  if (rc) return rc;
  SP_GHIP(hipMemcpyAsync(grp->d_fd, grp->h_fd.data(), sizeof(FlowDev) * n, hipMemcpyHostToDevice, grp->stream));
  SP_GHIP(launch_measure(grp->stream, grp->d_fd, grp->h_fd.data(), n, true));
  size_t total = 0;
  for (int i = 0; i < n; ++i) total = std::max(total, off[i] + 2 * (size_t)ctxs[i]->fd.B * ctxs[i]->fd.Rb);
  SP_GHIP(hipMemcpyAsync(grp->h_out, grp->d_out, total * 8, hipMemcpyDeviceToHost, grp->stream));
  SP_GHIP(hipStreamSynchronize(grp->stream));
  for (int i = 0; i < n; ++i) {
    const size_t nout = (size_t)ctxs[i]->fd.B * ctxs[i]->fd.Rb;
    if (zz && zz[i]) std::memcpy(zz[i], grp->h_out + off[i], nout * 8);
    if (dd && dd[i]) std::memcpy(dd[i], grp->h_out + off[i] + nout, nout * 8);
  }
  return SP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// sp_compute_step_noalloc: the FP64 computeStep on caller-owned device buffers.
// Everything else the kernels need -- descriptor, flags, mesh, lattice tables
// and the per-step weight tables -- lives in one file-scope device array of
// SP_NOALLOC_SCRATCH_BYTES, so the call allocates nothing.  Each device holds
// its own copy of the array; its address is looked up once per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[SP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
unsigned char *g_noalloc_ptr[kNoAllocDevices] = {};  // device address of scratch, per device
int g_noalloc_lds[kNoAllocDevices] = {};            // LDS limit per workgroup, per device
unsigned char g_noalloc_image[SP_NOALLOC_SCRATCH_BYTES];  // host image of the scratch contents

struct NoAllocPlan {
  size_t fd, flags, mesh, site, inv, ooff, o1, o2, oidx, rf, ri, wwp, wc, wn, cwp, cw, cn, total;
};

NoAllocPlan noalloc_plan(const sp_su2_problem *p) {
  NoAllocPlan q{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = (o + bytes + 255) & ~size_t(255);
    return at;
  };
  const size_t O = (size_t)p->ovl_off[p->L];
  const size_t emax = 2 * (size_t)p->nw + 8;
  q.fd = take(sizeof(FlowDev));
  q.flags = take(8 * sizeof(int));
  q.mesh = take((size_t)p->nw * 8);
  q.site = take((size_t)p->L * 4);
  q.inv = take((size_t)p->L * 4);
  q.ooff = take(((size_t)p->L + 1) * 4);
  q.o1 = take(O * 4);
  q.o2 = take(O * 4);
  q.oidx = take(O * 2);
  q.rf = take((size_t)p->R * 4);
  q.ri = take((size_t)p->R * 4);
  q.wwp = take((size_t)p->nw * emax * 8);
  q.wc = take((size_t)p->nw * emax * 8);
  q.wn = take((size_t)p->nw * 4);
  q.cwp = take(emax * 8);
  q.cw = take(emax * 8);
  q.cn = take(4);
  q.total = o;
  return q;
}

// device address of this device's scratch array (looked up once per device
// with the device current); callers hold g_noalloc_mu
int noalloc_scratch(int dev, unsigned char **base, int *max_lds) {
  if (dev < 0 || dev >= kNoAllocDevices) return SP_EARG;
  if (!g_noalloc_ptr[dev]) {
    void *q = nullptr;
    if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return SP_EHIP;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return SP_EHIP;
    if (hipFuncSetAttribute((const void *)k_dgamma<false, kThreads, SPFRG_MIN_BLOCKS>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_dgamma<false, kSmallThreads, SPFRG_MIN_BLOCKS>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
      return SP_EHIP;
    g_noalloc_lds[dev] = lds;
    g_noalloc_ptr[dev] = static_cast<unsigned char *>(q);
  }
  *base = g_noalloc_ptr[dev];
  *max_lds = g_noalloc_lds[dev];
  return SP_OK;
}

}  // namespace

extern "C" int sp_compute_step_noalloc(const sp_su2_problem *p, double cutoff, const double *d_v2,
                                       const double *d_v4, double *d_dv2, double *d_dv4, void *stream,
                                       int *diverged) {
  if (check_problem(nullptr, p) != SP_OK) return SP_EARG;
  if (p->precision != SP_PRECISION_FP64) return SP_EARG;
  if (!d_v2 || !d_v4 || !d_dv2 || !d_dv4 || !std::isfinite(cutoff) || cutoff <= 0.0) return SP_EARG;
  const NoAllocPlan q = noalloc_plan(p);
  if (q.total > (size_t)SP_NOALLOC_SCRATCH_BYTES) return SP_EARG;
  FlowDev f{};
  f.nw = p->nw;
  f.L = p->L;
  f.R = p->R;
  f.B = p->B;
  f.Rb = p->Rb;
  f.nf = (long long)p->nw * p->nw * (p->nw + 1) / 2;
  f.nv = f.nf * p->L;
  if (f.nv * 16 >= (1LL << 31)) return SP_EARG;
  f.emax = 2 * p->nw + 8;
  f.spin = p->spin;
  f.precision = SP_PRECISION_FP64;
  f.g_bytes = (unsigned)(f.nv * 16);
  f.cut = cutoff;
  size_t lds = 0;
  f.nt = dgamma_threads(f.L);
  launch_shape(f.nw, f.L, 8, f.nt, f.slots, f.nb, lds);
  double mesh_d[256];
  for (int i = 0; i < f.nw; ++i) mesh_d[i] = (double)p->mesh[i];
  geo_mesh_params(mesh_d, f.nw, f.geo, f.ga, f.gb);

  hipStream_t st = static_cast<hipStream_t>(stream);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return SP_EHIP;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  unsigned char *base = nullptr;
  int max_lds = 0;
  int rc = noalloc_scratch(dev, &base, &max_lds);
  if (rc != SP_OK) return rc;
  if (lds > (size_t)max_lds) return SP_EARG;
  // an earlier call on this stream may still read the scratch
  if (hipStreamSynchronize(st) != hipSuccess) return SP_EHIP;

  unsigned char *h = g_noalloc_image;
  std::memset(h, 0, q.total);  // zero flags and counters
  const int O = p->ovl_off[p->L];
  const int Lp = (f.L + 3) & ~3;
  std::memcpy(h + q.mesh, mesh_d, (size_t)f.nw * 8);
  std::memcpy(h + q.site, p->site_rid, (size_t)f.L * 4);
  std::memcpy(h + q.inv, p->inv_rid, (size_t)f.L * 4);
  std::memcpy(h + q.ooff, p->ovl_off, ((size_t)f.L + 1) * 4);
  if (O > 0) {
    std::memcpy(h + q.o1, p->ovl_rid1, (size_t)O * 4);
    std::memcpy(h + q.o2, p->ovl_rid2, (size_t)O * 4);
    unsigned short *idx = reinterpret_cast<unsigned short *>(h + q.oidx);
    for (int i = 0; i < O; ++i) idx[i] = (unsigned short)(p->ovl_rid1[i] * Lp + p->ovl_rid2[i]);
  }
  std::memcpy(h + q.rf, p->range_fwd, (size_t)f.R * 4);
  std::memcpy(h + q.ri, p->range_inv, (size_t)f.R * 4);
  f.G = reinterpret_cast<double2 *>(const_cast<double *>(d_v4));
  f.dG = reinterpret_cast<double2 *>(d_dv4);
  f.G32 = nullptr;
  f.v2 = const_cast<double *>(d_v2);
  f.dv2 = d_dv2;
  f.mesh = reinterpret_cast<const double *>(base + q.mesh);
  f.site_rid = reinterpret_cast<const int *>(base + q.site);
  f.inv_rid = reinterpret_cast<const int *>(base + q.inv);
  f.ovl_off = reinterpret_cast<const int *>(base + q.ooff);
  f.ovl_rid1 = reinterpret_cast<const int *>(base + q.o1);
  f.ovl_rid2 = reinterpret_cast<const int *>(base + q.o2);
  f.ovl_idx = reinterpret_cast<const unsigned short *>(base + q.oidx);
  f.range_fwd = reinterpret_cast<const int *>(base + q.rf);
  f.range_inv = reinterpret_cast<const int *>(base + q.ri);
  f.corr_rid = nullptr;
  f.Wwp = reinterpret_cast<double *>(base + q.wwp);
  f.Wc = reinterpret_cast<double *>(base + q.wc);
  f.Wn = reinterpret_cast<int *>(base + q.wn);
  f.Cwp = reinterpret_cast<double *>(base + q.cwp);
  f.Cw = reinterpret_cast<double *>(base + q.cw);
  f.Cn = reinterpret_cast<int *>(base + q.cn);
  f.corr_partial = nullptr;
  f.corr_out = nullptr;
  f.flags = reinterpret_cast<int *>(base + q.flags);
  std::memcpy(h + q.fd, &f, sizeof(FlowDev));
  if (hipMemcpy(base, h, q.total, hipMemcpyHostToDevice) != hipSuccess) return SP_EHIP;
  if (launch_compute(st, reinterpret_cast<const FlowDev *>(base + q.fd), &f, 1, nullptr) != hipSuccess)
    return SP_EHIP;
  if (diverged) {
    int fl[2] = {0, 0};
    if (hipStreamSynchronize(st) != hipSuccess) return SP_EHIP;
    if (hipMemcpy(fl, base + q.flags, sizeof(fl), hipMemcpyDeviceToHost) != hipSuccess) return SP_EHIP;
    if (fl[1]) return SP_EARG;  // evaluation list overflow
    *diverged = fl[0] ? 1 : 0;
  }
  return SP_OK;
}
