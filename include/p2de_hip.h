/*
 * p2de_hip.h -- C-ABI of the MI355X drop-in for P2DE.jl's limited ESDG right-hand side
 * and SSP-RK3 stage update (1D/2D compressible Euler, Lobatto or Gauss collocation).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never throws, and
 * returns a status (P2DE_OK = 0).  Host arrays are owned by the caller and only read
 * or written during the call.  Device memory, the HIP stream and all kernels belong to
 * the context.  A context is used by one host thread at a time.
 *
 * Memory layout of every array is the reference's Julia layout (column-major, isbits
 * SVector elements stored inline, component fastest):
 *   Uq, rhsU :: Array{SVector{Nc,Float64},2}(Nq, K)   -> double[K][Nq][Nc]
 *   operators are Matrix{Float64} column-major; index arrays hold 1-based Int64.
 */
#ifndef P2DE_HIP_H
#define P2DE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define P2DE_OK 0
#define P2DE_E_INVALID 1     /* null pointer, size mismatch, index out of range */
#define P2DE_E_UNSUPPORTED 2 /* option or operator structure not implemented */
#define P2DE_E_HIP 3         /* HIP runtime error (message in p2de_last_error) */
#define P2DE_E_ALLOC 4       /* device or host allocation failed */
#define P2DE_E_STATE 5       /* call out of order (e.g. no state uploaded) */
#define P2DE_E_POSITIVITY 6  /* a stage produced rho <= 0 or p <= 0 (or NaN) */

/* ------------------------------------------------ option enums (Solver.jl:1-121) */
enum { P2DE_RHS_LIMITED_DG = 0 };                 /* LimitedDG / ESLimitedLowOrderPos */
enum { P2DE_LF_ON_NODAL_VAL = 0,                  /* LaxFriedrichsOnNodalVal */
       P2DE_LF_ON_PROJECTED_VAL = 1 };            /* LaxFriedrichsOnProjectedVal */
enum { P2DE_CHANDRASHEKAR_FLUX = 0 };             /* ChandrashekarFlux */
enum { P2DE_BASIS_LOBATTO = 0, P2DE_BASIS_GAUSS = 1 };
enum { P2DE_NO_ENTROPY_PROJ_LIMITER = 0, P2DE_NODEWISE_SCALED_EXTRAPOLATION = 1 };
enum { P2DE_SUBCELL_LIMITER = 0, P2DE_ZHANG_SHU_LIMITER = 1 };
enum { P2DE_POSITIVITY_BOUND = 0,                          /* PositivityBound */
       P2DE_POSITIVITY_AND_MIN_ENTROPY_BOUND = 1,          /* PositivityAndMinEntropyBound */
       P2DE_POSITIVITY_AND_RELAXED_MIN_ENTROPY_BOUND = 2,  /* PositivityAndRelaxedMinEntropyBound */
       P2DE_POSITIVITY_AND_CELL_ENTROPY_BOUND = 3,         /* PositivityAndCellEntropyBound (2D Lobatto) */
       P2DE_POSITIVITY_AND_RELAXED_CELL_ENTROPY_BOUND = 4, /* PositivityAndRelaxedCellEntropyBound(beta) */
       /* TVD density bounds (subcell.jl:87-141) with the same entropy parts: code - 5 */
       P2DE_TVD_BOUND = 5,                                 /* TVDBound */
       P2DE_TVD_AND_MIN_ENTROPY_BOUND = 6,                 /* TVDAndMinEntropyBound */
       P2DE_TVD_AND_RELAXED_MIN_ENTROPY_BOUND = 7,         /* TVDAndRelaxedMinEntropyBound */
       P2DE_TVD_AND_CELL_ENTROPY_BOUND = 8,                /* TVDAndCellEntropyBound (2D Lobatto) */
       P2DE_TVD_AND_RELAXED_CELL_ENTROPY_BOUND = 9 };      /* TVDAndRelaxedCellEntropyBound(beta) */
enum { P2DE_NO_SHOCK_CAPTURE = 0, P2DE_HENNEMANN_SHOCK_CAPTURE = 1 };
/* FP64: the drop-in, matching the reference's FP64 arithmetic.  MIXED (2D only): FP32 flux
 * differencing, surface and low-order fluxes and element assembly; FP64 node pressures,
 * subcell flux potentials, limiter bounds and coefficients, RK stages and state.  A context
 * created with MIXED holds a one-instance p2de_batch and forwards every call to it. */
enum { P2DE_PRECISION_FP64 = 0, P2DE_PRECISION_MIXED = 1 };
/* Upper density bound of the subcell limiter.  REFERENCE: as limiter_utils.jl:11
 * destructures the 5-value bound tuple of subcell.jl:299 into 4 names, so
 * Urho = Lphi (0 for PositivityBound).  INTENDED: Urho = Urhoe = Inf. */
enum { P2DE_UPPER_BOUND_REFERENCE = 0, P2DE_UPPER_BOUND_INTENDED = 1 };

/* Param (Solver.jl:151-170) as scalars and enums. */
typedef struct {
    int Nd;            /* 1 or 2 */
    int N;             /* polynomial degree, 1..4 */
    int Kx, Ky;        /* elements per direction (Ky = 1 in 1D); K = Kx*Ky */
    double gamma, zeta, eta, POSTOL, ZEROTOL;
    double CFL, dt0, t0, T;
    int rhs;                      /* P2DE_RHS_LIMITED_DG */
    int low_order_surface_flux;   /* P2DE_LF_ON_NODAL_VAL or P2DE_LF_ON_PROJECTED_VAL */
    int high_order_surface_flux;  /* P2DE_LF_ON_PROJECTED_VAL */
    int high_order_volume_flux;   /* P2DE_CHANDRASHEKAR_FLUX */
    int basis;                    /* P2DE_BASIS_LOBATTO or P2DE_BASIS_GAUSS */
    int entropyproj_limiter;      /* P2DE_NO_ENTROPY_PROJ_LIMITER; with Gauss also
                                     P2DE_NODEWISE_SCALED_EXTRAPOLATION */
    int rhs_limiter;              /* P2DE_SUBCELL_LIMITER or P2DE_ZHANG_SHU_LIMITER */
    int bound;                    /* P2DE_POSITIVITY*_BOUND (SubcellLimiter; Zhang-Shu: POSITIVITY);
                                     the cell-entropy bounds in 2D with the Lobatto basis */
    int shockcapture;             /* P2DE_NO_SHOCK_CAPTURE or P2DE_HENNEMANN_SHOCK_CAPTURE */
    int upper_bound_mode;         /* P2DE_UPPER_BOUND_* */
    int precision;                /* P2DE_PRECISION_* */
    double shock_a, shock_c;      /* HennemannShockCapture(a, c) (Solver.jl:67-74) */
    double cell_beta;             /* (Positivity|TVD)AndRelaxedCellEntropyBound(beta) (Solver.jl:52-64) */
} p2de_config;

/* Operators{DIM} (Solver.jl:192-212), sizes from SizeData (Solver.jl:173-183). */
typedef struct {
    int Nq, Nfp, Nh;
    const double *Srsh_db;   /* Nd blocks of Nh x Nh */
    const double *Srs0;      /* Nd blocks of Nq x Nq (dense copy of the sparse matrices) */
    int nnz_h;
    const int64_t *Srsh_nnz; /* nnz_h pairs (i, j), i > j, 1-based, in the reference's order */
    int nnz_0;
    const int64_t *Srs0_nnz;
    const double *Brs;       /* Nd blocks of Nfp x Nfp (diagonal) */
    const double *MinvVhT;   /* Nq x Nh */
    const double *MinvVfT;   /* Nq x Nfp */
    const double *Vf;        /* Nfp x Nq */
    const double *Vf_low;    /* Nfp x Nq */
    const double *wq;        /* Nq */
    const int64_t *fq2q;     /* Nfp, 1-based */
    const int64_t *q2fq;     /* Nq x 2 column-major, 1-based, 0 = absent */
    const double *VDM_inv;   /* Nq x Nq (Operators.VDM_inv); read by HennemannShockCapture and the
                                relaxed bound only, may be NULL otherwise */
} p2de_operators;

/* GeomData (Solver.jl:185-190). */
typedef struct {
    const double *Jq;        /* Nq x K */
    const double *GJh;       /* NGEO blocks of Nh x K: rxJh (, sxJh, ryJh, syJh) */
} p2de_geometry;

/* BCData (StateParam.jl:1-7). */
typedef struct {
    const int64_t *mapP;     /* Nfp x K, 1-based linear index into (Nfp, K) */
    int nI;
    const int64_t *mapI;     /* inflow face nodes, 1-based */
    const double *Ival;      /* nI x Nc, component fastest */
    int nO;
    const int64_t *mapO;     /* outflow face nodes, 1-based */
} p2de_bcdata;

typedef struct p2de_ctx p2de_ctx;

/* Positivity report: nodes with !(rho > 0) or !(p > 0), counted over every stage
 * since the last reset (p2de_check_positivity resets it). */
typedef struct {
    int64_t violations;
    int64_t first_node;      /* i + Nq*k of the first offending node, -1 if none */
    int first_stage;         /* 1..3, 0 if none */
    double min_rho, min_p;   /* over the last state checked by p2de_check_positivity */
} p2de_pos_report;

typedef struct {
    int steps;               /* steps taken */
    double t;                /* final time */
    double last_dt;
    int64_t violations;      /* positivity violations seen (every stage) */
} p2de_run_stats;

/* Fields readable with p2de_get_field.  RHSH, RHSL and L_LOCAL need debug output on
 * (p2de_set_debug); the others are always kept. */
enum {
    P2DE_FIELD_RHSU = 0,     /* (Nq, K) x Nc */
    P2DE_FIELD_RHSH = 1,     /* (Nq, K) x Nc */
    P2DE_FIELD_RHSL = 2,     /* (Nq, K) x Nc */
    P2DE_FIELD_L_LOCAL = 3,  /* (Nq+N1D, Nd, K) for the stage of the last call; Zhang-Shu:
                                the first K values are L[:, nstage] */
    P2DE_FIELD_THETA = 4,    /* (Nfp, K) theta_local[:, :, nstage] of the last stage
                                (NodewiseScaledExtrapolation) */
    P2DE_FIELD_L = 5,        /* (K, 3) the reference's L (State.jl:17): the Zhang-Shu element
                                parameters of each stage (zhangshu.jl:44), zeros with the
                                subcell limiter, which does not write it */
    P2DE_FIELD_THETA_MEAN = 6/* (K, 3) the reference's theta (State.jl:19): the mean of
                                theta_local over the element's face nodes per stage (filter.jl:58),
                                summed in node order; zeros without NodewiseScaledExtrapolation */
};

/* Replaces initialize_DG's products as rhs!/SSP33! consume them (init.jl:44-60):
 * validates the options and operator structure, builds the per-element tables,
 * allocates device memory and creates the stream.  Returns P2DE_E_UNSUPPORTED for
 * option sets or operators outside the implemented path. */
int p2de_create(const p2de_config *cfg, const p2de_operators *ops, const p2de_geometry *geom,
                const p2de_bcdata *bc, p2de_ctx **out);
void p2de_destroy(p2de_ctx *ctx);

/* state.preallocation.Uq (State.jl:2): host -> device and device -> host copies,
 * complete on return. */
int p2de_set_state(p2de_ctx *ctx, const double *Uq);
int p2de_get_state(p2de_ctx *ctx, double *Uq);

/* rhs!(state, solver, state_param, TimeParam(t, dt, nstage)) (rhs.jl:5-13): evaluates the
 * limited RHS of the device state with the limiter using dt.  *dt_out receives the CFL
 * time step at nstage 1 (low_order_graph_viscosity.jl:222-243) and dt otherwise.  If
 * rhsU is not NULL the RHS is copied to it. */
int p2de_rhs(p2de_ctx *ctx, double t, double dt, int nstage, double *dt_out, double *rhsU);

/* Keep rhsH, rhsL and L_local of the next p2de_rhs calls for p2de_get_field (debug). */
int p2de_set_debug(p2de_ctx *ctx, int on);
int p2de_get_field(p2de_ctx *ctx, int field, double *out);

/* One iteration of SSP33!'s time loop (SSPRK33.jl:29-43) on the device state:
 * dt = min(CFL*dt0, T - t), three limited stages, t += dt.  *t_inout is updated,
 * *dt_out receives the step.  Does nothing (dt = 0) once t >= T. */
int p2de_ssprk3_step(p2de_ctx *ctx, double *t_inout, double *dt_out);

/* SSP33!'s loop (SSPRK33.jl:28-56) without history output: steps from the current
 * device time until t >= T or max_steps steps (max_steps <= 0: no limit).  Launches
 * run back to back; time and dt stay on the device. */
int p2de_run(p2de_ctx *ctx, int max_steps, p2de_run_stats *stats);

/* Device time control for p2de_run/p2de_ssprk3_step. */
int p2de_set_time(p2de_ctx *ctx, double t);
int p2de_get_time(p2de_ctx *ctx, double *t, double *last_dt);

/* One segment of SSP33!'s time loop with its history output (SSPRK33.jl:28-56): steps from
 * the device time, writing each step's dt to dthist, until after a step the reference would
 * record a history entry -- mod(i, output_interval) == 0 or |t - T| < 1e-10, where i is the
 * reference's step index (1 at t0, incremented by every step; *step_index in and out) -- or
 * t >= T, or max_steps steps (<= 0: no limit) or dthist_cap steps were taken.  *steps_taken
 * receives the count and *at_output 1 when the segment ended at a history entry; the caller
 * then reads the state (Uhist), p2de_get_time (thist), P2DE_FIELD_L (Lhist),
 * P2DE_FIELD_THETA_MEAN (thetahist) and p2de_check_conservation.  Launches stay queued
 * between history entries (no host round trip per step in FP64).  With precision MIXED
 * the steps are taken one per host round trip. */
int p2de_run_segment(p2de_ctx *ctx, int output_interval, int max_steps, int64_t *step_index,
                     double *dthist, int dthist_cap, int *steps_taken, int *at_output);

/* check_conservation (src/dg/utils.jl:1-12): sum over the components of
 * sum_k sum_i Jq[i,k] wq[i] Uq[i,k] for the device state, accumulated on the host in the
 * reference's order. */
int p2de_check_conservation(p2de_ctx *ctx, double *total);

/* Checkpoints.  The reference keeps its history in memory and saves it with JLD2 after the
 * run (postprocess.jl:131-142); it has no restart.  These three calls have no counterpart in
 * it: they snapshot a context (state, device time, last dt, step counter, the t0 entropy
 * minimum of the min-entropy bounds, and the caller's step index) as one self-describing
 * byte block -- a header with the problem's sizes, the state double[K][Nq][Nc], a 64-bit
 * FNV-1a checksum of the state -- that the caller stores where it likes.  Loading it into a
 * context of the same problem (Nd, N, basis, K) and stepping on gives the trajectory of the
 * run that was not stopped, bit for bit.  Load rejects a block of another problem, a
 * truncated block and a checksum mismatch with P2DE_E_INVALID and changes nothing then. */
int64_t p2de_checkpoint_size(const p2de_ctx *ctx);
int p2de_checkpoint_save(p2de_ctx *ctx, int64_t step_index, void *buf, int64_t size);
int p2de_checkpoint_load(p2de_ctx *ctx, const void *buf, int64_t size, int64_t *step_index);

/* Positivity of the device state plus the count of violations seen in every stage
 * since the last call; then resets the count. */
int p2de_check_positivity(p2de_ctx *ctx, p2de_pos_report *rep);

/* Launch only the kernels of `steps` SSP-RK3 steps, no host synchronisation. */
int p2de_enqueue_steps(p2de_ctx *ctx, int steps);
/* Wait for all queued device work of the context. */
int p2de_synchronize(p2de_ctx *ctx);
/* Per-kernel timing with HIP events recorded on the context's stream around every
 * launch while enabled.  kinds: 0 = element RHS (phase A; on the Gauss basis also the
 * entropy projection of the face states), 1 = symmetrise+update (phase B), 2 = step
 * bookkeeping.  ms[k] = total milliseconds, count[k] = launches. */
int p2de_set_kernel_timing(p2de_ctx *ctx, int on);
int p2de_get_kernel_timing(p2de_ctx *ctx, double ms[3], int64_t count[3]);

/* Launch shapes of the element kernel (elements per workgroup E, threads per workgroup NT).
 * All shapes give bit-identical results.  p2de_autotune times each shape on `steps`
 * SSP-RK3 steps of the current device state (state and time are restored afterwards),
 * selects the fastest and, if cache_path is not NULL, records the choice there keyed by
 * (Nd, N, K); a later p2de_autotune with the same key reads the cache instead of timing. */
int p2de_num_variants(const p2de_ctx *ctx);
int p2de_set_variant(p2de_ctx *ctx, int variant);
int p2de_get_variant(const p2de_ctx *ctx, int *variant, int *E, int *NT);
int p2de_autotune(p2de_ctx *ctx, int steps, const char *cache_path, int *chosen);

/* Number of degrees of freedom (nodes) K*Nq, and the device address of the state
 * (double[K][Nq][Nc]) for callers that keep data on the GPU. */
int64_t p2de_num_dofs(const p2de_ctx *ctx);
int p2de_device_state(p2de_ctx *ctx, void **dptr);

/* ------------------------------------------------------------------ batched solves
 * Many independent 2D solves with their own mesh sizes, degrees (N = 1..4), times and time
 * steps, advanced together in precision P2DE_PRECISION_MIXED.  This has no counterpart in
 * the reference, which runs one solve per Julia process (SSPRK33.jl:1-62 per instance); each
 * instance is described exactly as for p2de_create (arrays of n descriptors), and instances
 * of equal N must pass the same operators.  Per SSP-RK3 step one CFL launch and three fused
 * stage launches (RHS, limiter, stage update) cover every element of every instance; the
 * launches of a step are captured once in a hipGraph and replayed.  Instance i's state is
 * Array{SVector{4,Float64},2}(Nq_i, K_i) as for p2de_set_state. */
typedef struct p2de_batch p2de_batch;
int p2de_batch_create(int n, const p2de_config *cfg, const p2de_operators *ops, const p2de_geometry *geom,
                      const p2de_bcdata *bc, p2de_batch **out);
void p2de_batch_destroy(p2de_batch *b);
int p2de_batch_num_instances(const p2de_batch *b);
int64_t p2de_batch_num_dofs(const p2de_batch *b, int inst);          /* inst < 0: all instances */
int p2de_batch_set_state(p2de_batch *b, int inst, const double *Uq);
int p2de_batch_get_state(p2de_batch *b, int inst, double *Uq);
int p2de_batch_device_state(p2de_batch *b, int inst, void **dptr);
int p2de_batch_set_time(p2de_batch *b, int inst, double t);          /* inst < 0: all instances */
int p2de_batch_get_time(p2de_batch *b, int inst, double *t, double *last_dt, int *steps);
/* rhs! of one instance (as p2de_rhs); the other instances are untouched. */
int p2de_batch_rhs(p2de_batch *b, int inst, double t, double dt, int nstage, double *dt_out, double *rhsU);
int p2de_batch_set_debug(p2de_batch *b, int on);                     /* keep rhsH, rhsL of p2de_batch_rhs */
int p2de_batch_get_field(p2de_batch *b, int inst, int field, double *out);   /* RHSU, RHSH, RHSL */
/* Queue `steps` SSP-RK3 steps of every instance not yet at its T (no host synchronisation). */
int p2de_batch_enqueue_steps(p2de_batch *b, int steps);
int p2de_batch_synchronize(p2de_batch *b);
/* Step until every instance reaches its T or max_steps steps (max_steps <= 0: no limit). */
int p2de_batch_run(p2de_batch *b, int max_steps, int *steps_taken);
int p2de_batch_check_positivity(p2de_batch *b, int inst, p2de_pos_report *rep);
/* HIP-event timing per launch kind: 0 = fused stage, 1 = CFL pass, 2 = bookkeeping.  While
 * on, steps are launched directly instead of through the graph. */
int p2de_batch_set_kernel_timing(p2de_batch *b, int on);
int p2de_batch_get_kernel_timing(p2de_batch *b, double ms[3], int64_t count[3]);
int p2de_batch_set_graph(p2de_batch *b, int on);                     /* default on */
/* Launch shapes (threads NT, elements per tile E[N+1] for each degree) that cover the
 * batch's degrees; p2de_batch_autotune times each on `steps` steps (state and times are
 * restored) and keeps the fastest, caching the choice in cache_path when not NULL.  The
 * cache is a text file of "key value" lines shared with p2de_autotune; a batch line holds the
 * shape itself (NT:E2,E3,E4,E5, keyed by the elements per degree), so a build whose list of
 * shapes differs times again rather than taking another shape. */
int p2de_batch_num_variants(const p2de_batch *b);
int p2de_batch_set_variant(p2de_batch *b, int v);
int p2de_batch_get_variant(const p2de_batch *b, int *v, int *NT, int *E /* [6] */);
int p2de_batch_autotune(p2de_batch *b, int steps, const char *cache_path, int *chosen);
const char *p2de_batch_last_error(const p2de_batch *b);

const char *p2de_last_error(const p2de_ctx *ctx);
const char *p2de_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
