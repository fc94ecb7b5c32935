/*
 * spfrg.h -- C ABI of the MI355X pf-FRG flow engine (SU(2) core).
 *
 * One context holds one flow: the vertex state (self-energy Sigma and the
 * two-particle vertex Gamma in its spin and density channels), the flow
 * (dSigma/dLambda, dGamma/dLambda), the lattice tables and all scratch, resident
 * in device memory.  The entry points are the operations the reference's FrgCore
 * plugin boundary performs on its cores; each cites the reference interface it
 * replaces (paths relative to the SpinParser repository).
 *
 * Conventions
 *   - Every call returns an int status: SP_OK (0) or one of the SP_E* codes;
 *     sp_last_error() returns a message for the last failure on a context.
 *     Nothing throws across the ABI.
 *   - Host pointers are borrowed for the duration of the call only.
 *   - Vertex arrays use the reference layout: Gamma[(s(s+1)/2 + u) * Nw * L +
 *     t * L + rid] with s >= u (src/SU2/SU2VertexTwoParticle.hpp:595-605), one
 *     array per channel (spin "ss", density "dd"); Sigma[Nw].
 *   - Cutoff values are the reference's float grid values passed as double.
 *   - A context is used from one host thread at a time; its work runs on the
 *     context's own HIP stream (default priority) on the current device.
 */
#ifndef SPFRG_H
#define SPFRG_H

#ifdef __cplusplus
extern "C" {
#endif

#define SP_OK 0
#define SP_EARG 1      /* invalid argument or unsupported problem size        */
#define SP_ENOMEM 2    /* device or host allocation failed                     */
#define SP_EHIP 3      /* HIP runtime error                                    */
#define SP_ESTATE 4    /* call out of order (e.g. flow read before a step)     */

#define SP_PRECISION_FP64 0   /* plain FP64 form                              */
#define SP_PRECISION_MIXED 1  /* FP32 frequency integrals with compensated sums;
                                 FP64 state, Euler step, self-energy, measurement */

/* Problem descriptor: the flattened lattice / mesh tables of one SU(2) flow.
 * Built host-side by the lattice factory; replaces what the reference reads
 * through FrgCommon::frequency() / FrgCommon::lattice()
 * (src/FrequencyDiscretization.hpp:164-361, src/Lattice.hpp:297-552). */
typedef struct sp_su2_problem {
  int nw;                 /* positive Matsubara frequencies N_w               */
  const float *mesh;      /* N_w ascending positive mesh points (float grid)  */
  int L;                  /* representative lattice sites                      */
  const int *site_rid;    /* L: getSites()[j].rid       (Lattice.hpp:491)     */
  const int *inv_rid;     /* L: getInvertedSites()[j].rid (Lattice.hpp:481)   */
  const int *ovl_off;     /* L+1: CSR offsets of getOverlap(rid) (Lattice.hpp:470) */
  const int *ovl_rid1;    /* ovl_off[L]: overlap rid1                          */
  const int *ovl_rid2;    /* ovl_off[L]: overlap rid2                          */
  int R;                  /* sites in range of the origin                      */
  const int *range_fwd;   /* R: symmetryTransform(0, j), j in getRange(0)      */
  const int *range_inv;   /* R: symmetryTransform(j, 0)                        */
  int B;                  /* basis sites                                       */
  int Rb;                 /* sites in range of each basis site                 */
  const int *corr_rid;    /* B*Rb: symmetryTransform(b, j) for the .obs output */
  double spin;            /* spin length S (core option "spin")                */
  int precision;          /* SP_PRECISION_FP64 or SP_PRECISION_MIXED           */
} sp_su2_problem;

typedef struct sp_ctx sp_ctx;

/* Create a flow context and upload the tables; replaces the construction of
 * SU2FrgCore + SU2EffectiveAction (src/SU2/SU2FrgCore.cpp:17-81).  The state is
 * zero until sp_set_state(). */
int sp_ctx_create(const sp_su2_problem *problem, sp_ctx **out);

/* Release every device buffer; replaces SU2FrgCore::~SU2FrgCore
 * (src/SU2/SU2FrgCore.cpp:83-87). */
int sp_ctx_destroy(sp_ctx *ctx);

/* Upload the state at a cutoff; replaces writing EffectiveAction::cutoff and the
 * vertex arrays, as SU2EffectiveAction's initial condition and
 * readCheckpoint do (src/SU2/SU2EffectiveAction.hpp:38-60,159-205). */
int sp_set_state(sp_ctx *ctx, double cutoff, const double *v2, const double *v4ss, const double *v4dd);

/* Switch the precision form of a context between SP_PRECISION_FP64 and
 * SP_PRECISION_MIXED; no reference counterpart (the reference computes in
 * float throughout).  The FP64 state is the master copy in both forms, so a
 * switch changes how later derivatives are evaluated, never the state.  The
 * mixed form needs the FP32 shadow, which a context created with
 * SP_PRECISION_MIXED holds; switching to it rebuilds the shadow from the
 * state.  A served flow switches to the FP64 form over the cutoff window in
 * which its correlations are most sensitive to the FP32 rounding.
 * Asynchronous on the context stream. */
int sp_set_precision(sp_ctx *ctx, int precision);

/* Download the state; the read side of SU2EffectiveAction::writeCheckpoint
 * (src/SU2/SU2EffectiveAction.hpp:79-150).  Null pointers skip that array. */
int sp_get_state(sp_ctx *ctx, double *cutoff, double *v2, double *v4ss, double *v4dd);

/* Download the flow of the last sp_compute_step (the reference's FrgCore::_flow,
 * src/FrgCore.hpp:140).  Null pointers skip that array. */
int sp_get_flow(sp_ctx *ctx, double *dv2, double *dv4ss, double *dv4dd);

/* Evaluate the right-hand side at the state's cutoff: dSigma first, then dGamma
 * with its Katanin term; replaces SU2FrgCore::computeStep
 * (src/SU2/SU2FrgCore.cpp:89-109, 139-431).  Asynchronous on the context
 * stream. */
int sp_compute_step(sp_ctx *ctx);

/* Explicit Euler update in place, Gamma += (new_cutoff - cutoff) * dGamma,
 * Sigma likewise, then cutoff = new_cutoff; replaces SU2FrgCore::finalizeStep
 * (src/SU2/SU2FrgCore.cpp:111-137).  Asynchronous. */
int sp_finalize_step(sp_ctx *ctx, double new_cutoff);

/* 1 if the flow of the last sp_compute_step holds a NaN, else 0; replaces
 * SU2EffectiveAction::isDiverged on FrgCore::_flow
 * (src/SU2/SU2EffectiveAction.hpp:212-230, src/SpinParser.cpp:151). */
int sp_is_diverged(sp_ctx *ctx, int *diverged);

/* Static (nu = 0) spin and density susceptibilities at the state's cutoff,
 * B*Rb values each in the .obs site order (SU2CorZZ, SU2CorDD); replaces
 * SU2MeasurementCorrelation::_calculateCorrelation
 * (src/SU2/SU2MeasurementCorrelation.cpp:77-177).  Synchronous. */
int sp_measure_correlation(sp_ctx *ctx, double *zz, double *dd);

/* The cutoff loop of SpinParser::runCore (src/SpinParser.cpp:141-172) over
 * cutoffs[0..ncut-1], the state being at cutoffs[0]: for k = 0, 1, ...
 * computeStep at cutoffs[k], takeMeasurements (record k: B*Rb values at
 * zz + k*B*Rb and dd + k*B*Rb), stop if isDiverged, finalizeStep(cutoffs[k+1]).
 * The steps are queued without a host round trip each: records come back
 * through pinned memory, a flow whose derivative holds a NaN is stopped on the
 * device, and the host checks the records every 16 steps.  Records and final
 * state equal those of the step-by-step calls.  *nrec = records written
 * (ncut - 1 unless diverged), *diverged = 1 if the loop stopped on a NaN;
 * the context is then left as the step-by-step loop leaves it (flow of the
 * last step computed, not finalized).  zz and dd hold (ncut - 1)*B*Rb values;
 * the final takeMeasurements of runCore is a separate sp_measure_correlation. */
int sp_run_flow(sp_ctx *ctx, const double *cutoffs, int ncut, double *zz, double *dd, int *nrec, int *diverged);

/* Wait for all work queued on the context stream. */
int sp_synchronize(sp_ctx *ctx);

/* Message of the last failed call on ctx (static storage per context). */
const char *sp_last_error(const sp_ctx *ctx);

/* Kernel timing with HIP events on the context stream (no reference
 * counterpart; replaces LoadManager's per-stack compute-time accounting,
 * src/lib/LoadManager.hpp:583-648).  kernel ids: 0 dGamma, 1 dSigma,
 * 2 propagator weight tables, 3 Euler update. */
int sp_profile_enable(sp_ctx *ctx, int enable);
int sp_profile_get(sp_ctx *ctx, int kernel, double *total_ms, long *launches);

/* Device bytes held by the context (admission by footprint). */
int sp_ctx_footprint(const sp_su2_problem *problem, long long *bytes);

/* The FP64 computeStep without allocating: for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  No counterpart in the reference.  The caller owns the
 * state and the flow on the device:
 *   d_v2[Nw] (Sigma) and d_v4[2 * Nf * L] (Gamma) in, d_dv2[Nw] and
 *   d_dv4[2 * Nf * L] out, Gamma and dGamma interleaved as {spin, density}
 *   pairs at the reference index (Gamma[2 i] = spin, Gamma[2 i + 1] = density
 *   of entry i, i = (s(s+1)/2 + u) * Nw * L + t * L + rid), Nf = Nw^2 (Nw+1)/2.
 * problem->precision must be SP_PRECISION_FP64.  Descriptor, flags, lattice
 * tables and weight tables live in one library-owned device array of
 * SP_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends); problems
 * whose tables exceed it are rejected with SP_EARG (N_w = 64 with 8,000
 * overlap entries needs about 0.2 MB).  The call waits for earlier work on
 * `stream` (a hipStream_t; null = the null stream), uploads the tables
 * synchronously and queues dSigma and dGamma on `stream`.  diverged non-null:
 * the call also waits for the result and reports a NaN in the flow
 * (SpinParser::runCore's isDiverged check); null: fully asynchronous after
 * the upload.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other threads, other streams) race.  Results equal
 * sp_compute_step on a context holding the same state. */
#define SP_NOALLOC_SCRATCH_BYTES (4 << 20)
int sp_compute_step_noalloc(const sp_su2_problem *problem, double cutoff, const double *d_v2, const double *d_v4,
                            double *d_dv2, double *d_dv4, void *stream, int *diverged);

/* Flow groups (no reference counterpart; replaces running several SpinParser
 * processes side by side).  A group batches resident flows of any sizes into
 * one launch per kernel and step.  sp_group_step performs, for every flow i,
 * one iteration of the reference's cutoff loop (src/SpinParser.cpp:141-160):
 * computeStep at the flow's current cutoff, the correlation measurement
 * (zz[i], dd[i], B*Rb values each; null skips), the divergence check
 * (diverged[i]) and finalizeStep to new_cutoffs[i]; a non-finite or
 * non-positive new_cutoffs[i], or a diverged flow, keeps its state.  The call
 * is synchronous; contexts must not have work queued on their own streams.
 * sp_group_measure performs the measurement only (the final one of a flow). */
typedef struct sp_group sp_group;
int sp_group_create(int capacity, sp_group **out);
int sp_group_destroy(sp_group *group);
int sp_group_step(sp_group *group, sp_ctx *const *ctxs, int n, const double *new_cutoffs, double *const *zz,
                  double *const *dd, int *diverged);
int sp_group_measure(sp_group *group, sp_ctx *const *ctxs, int n, double *const *zz, double *const *dd);
const char *sp_group_last_error(const sp_group *group);

/* ------------------------------------------------------------------------
 * TRI core: general time-reversal-invariant two-spin couplings, 16 spin
 * components per vertex entry (src/TRI/TRIFrgCore.cpp).  FP64.
 *
 * Vertex layout (src/TRI/TRIVertexTwoParticle.hpp:62-76,657-666):
 *   Gamma[((s(s+1)/2 + u) * Nw + t) * 16 L + (4 s1 + s2) L + rid], s >= u,
 *   spin components s1, s2 in (x, y, z, d) = (0, 1, 2, 3); Sigma[Nw].
 * Spin permutations are 3 entries per site transformation, the image of the
 * components x, y, z (LatticeSiteDescriptor::spinPermutation,
 * src/Lattice.hpp:150-161; the overlap transforms transformedX/Y/Z 1/2,
 * src/Lattice.hpp:38-75).
 * ---------------------------------------------------------------------- */
typedef struct sp_tri_problem {
  int nw;                 /* positive Matsubara frequencies N_w               */
  const float *mesh;      /* N_w ascending positive mesh points               */
  int L;                  /* representative lattice sites                      */
  const int *site_rid;    /* L: getSites()[j].rid                              */
  const int *site_perm;   /* 3L: getSites()[j].spinPermutation                 */
  const int *inv_rid;     /* L: getInvertedSites()[j].rid                      */
  const int *inv_perm;    /* 3L: getInvertedSites()[j].spinPermutation         */
  const int *ovl_off;     /* L+1: CSR offsets of getOverlap(rid)               */
  const int *ovl_rid1;    /* O: overlap rid1                                   */
  const int *ovl_rid2;    /* O: overlap rid2                                   */
  const int *ovl_perm1;   /* 3O: transformedX1, Y1, Z1                         */
  const int *ovl_perm2;   /* 3O: transformedX2, Y2, Z2                         */
  int R;                  /* sites in range of the origin                      */
  const int *range_fwd;   /* R: symmetryTransform(0, j)                        */
  const int *range_inv;   /* R: symmetryTransform(j, 0)                        */
  int B;                  /* basis sites                                       */
  int Rb;                 /* sites in range of each basis site                 */
  const int *corr_rid;    /* B*Rb: symmetryTransform(b, j) rid                 */
  const int *corr_perm;   /* 3*B*Rb: its spin transformation of x, y, z        */
} sp_tri_problem;

typedef struct sp_tri_ctx sp_tri_ctx;

/* Construction of TRIFrgCore + TRIEffectiveAction (src/TRI/TRIFrgCore.cpp:17-71). */
int sp_tri_ctx_create(const sp_tri_problem *problem, sp_tri_ctx **out);
/* TRIFrgCore::~TRIFrgCore (src/TRI/TRIFrgCore.cpp:73-77). */
int sp_tri_ctx_destroy(sp_tri_ctx *ctx);
/* State upload: the initial condition / readCheckpoint of TRIEffectiveAction
 * (src/TRI/TRIEffectiveAction.hpp:38-64,159-205). */
int sp_tri_set_state(sp_tri_ctx *ctx, double cutoff, const double *v2, const double *v4);
/* State download (writeCheckpoint's read side, src/TRI/TRIEffectiveAction.hpp:79-150). */
int sp_tri_get_state(sp_tri_ctx *ctx, double *cutoff, double *v2, double *v4);
/* FrgCore::_flow of the last step (src/FrgCore.hpp:140). */
int sp_tri_get_flow(sp_tri_ctx *ctx, double *dv2, double *dv4);
/* TRIFrgCore::computeStep (src/TRI/TRIFrgCore.cpp:79-96,122-3001).  Asynchronous. */
int sp_tri_compute_step(sp_tri_ctx *ctx);
/* TRIFrgCore::finalizeStep (src/TRI/TRIFrgCore.cpp:98-120).  Asynchronous. */
int sp_tri_finalize_step(sp_tri_ctx *ctx, double new_cutoff);
/* TRIEffectiveAction::isDiverged on the flow (src/TRI/TRIEffectiveAction.hpp:212-230). */
int sp_tri_is_diverged(sp_tri_ctx *ctx, int *diverged);
/* TRIMeasurementCorrelation::_calculateCorrelation
 * (src/TRI/TRIMeasurementCorrelation.cpp:147-316): 10*B*Rb values, the groups
 * TRICorDD, XX, XY, XZ, YX, YY, YZ, ZX, ZY, ZZ in this order, each in the
 * .obs site order.  Synchronous. */
int sp_tri_measure_correlation(sp_tri_ctx *ctx, double *out);
int sp_tri_synchronize(sp_tri_ctx *ctx);
const char *sp_tri_last_error(const sp_tri_ctx *ctx);
/* Kernel timing (kernel ids as for the SU(2) core). */
int sp_tri_profile_enable(sp_tri_ctx *ctx, int enable);
int sp_tri_profile_get(sp_tri_ctx *ctx, int kernel, double *total_ms, long *launches);
/* Device bytes held by a TRI context. */
int sp_tri_ctx_footprint(const sp_tri_problem *problem, long long *bytes);

/* ------------------------------------------------------------------------
 * XYZ core: diagonal two-spin couplings J_xx, J_yy, J_zz, four vertex
 * components X, Y, Z, D (src/XYZ/XYZFrgCore.cpp).  FP64.  The lattice tables
 * are those of the TRI core (the spin permutations act on X, Y, Z).
 *
 * Vertex layout (src/XYZ/XYZVertexTwoParticle.hpp:63-77,606-640): four planes
 * X, Y, Z, D (_dataXX, _dataYY, _dataZZ, _dataDD) of N_f * L values, plane c at
 * c * N_f * L, entry ((s(s+1)/2 + u) * Nw + t) * L + rid, s >= u; Sigma[Nw].
 * ---------------------------------------------------------------------- */
typedef sp_tri_problem sp_xyz_problem;
typedef struct sp_xyz_ctx sp_xyz_ctx;

/* Construction of XYZFrgCore + XYZEffectiveAction (src/XYZ/XYZFrgCore.cpp:17-98). */
int sp_xyz_ctx_create(const sp_xyz_problem *problem, sp_xyz_ctx **out);
/* XYZFrgCore::~XYZFrgCore (src/XYZ/XYZFrgCore.cpp:100-104). */
int sp_xyz_ctx_destroy(sp_xyz_ctx *ctx);
/* State upload: the initial condition / readCheckpoint of XYZEffectiveAction
 * (src/XYZ/XYZEffectiveAction.hpp:38-66,164-217). */
int sp_xyz_set_state(sp_xyz_ctx *ctx, double cutoff, const double *v2, const double *v4);
/* State download (writeCheckpoint's read side, src/XYZ/XYZEffectiveAction.hpp:82-155). */
int sp_xyz_get_state(sp_xyz_ctx *ctx, double *cutoff, double *v2, double *v4);
/* FrgCore::_flow of the last step (src/FrgCore.hpp:140). */
int sp_xyz_get_flow(sp_xyz_ctx *ctx, double *dv2, double *dv4);
/* XYZFrgCore::computeStep (src/XYZ/XYZFrgCore.cpp:106-126,164-571).  Asynchronous. */
int sp_xyz_compute_step(sp_xyz_ctx *ctx);
/* XYZFrgCore::finalizeStep (src/XYZ/XYZFrgCore.cpp:128-162).  Asynchronous. */
int sp_xyz_finalize_step(sp_xyz_ctx *ctx, double new_cutoff);
/* XYZEffectiveAction::isDiverged on the flow (src/XYZ/XYZEffectiveAction.hpp:219-238). */
int sp_xyz_is_diverged(sp_xyz_ctx *ctx, int *diverged);
/* XYZMeasurementCorrelation::_calculateCorrelation
 * (src/XYZ/XYZMeasurementCorrelation.cpp:98-212): 4*B*Rb values, the groups
 * XYZCorXX, YY, ZZ, DD in this order, each in the .obs site order.  Synchronous. */
int sp_xyz_measure_correlation(sp_xyz_ctx *ctx, double *out);
int sp_xyz_synchronize(sp_xyz_ctx *ctx);
const char *sp_xyz_last_error(const sp_xyz_ctx *ctx);
/* Kernel timing (kernel ids as for the SU(2) core). */
int sp_xyz_profile_enable(sp_xyz_ctx *ctx, int enable);
int sp_xyz_profile_get(sp_xyz_ctx *ctx, int kernel, double *total_ms, long *launches);
/* Device bytes held by an XYZ context. */
int sp_xyz_ctx_footprint(const sp_xyz_problem *problem, long long *bytes);

#ifdef __cplusplus
}
#endif
#endif /* SPFRG_H */
