/*
 * p2de_oracle.h -- FP64 CPU oracle of P2DE.jl's limited ESDG right-hand side and
 * SSP-RK3 step (1D/2D compressible Euler).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared against and
 * the CPU baseline timed beside it; nothing in the product links, loads or calls it.
 *
 * It restates, loop for loop, the reference files
 *   src/dg/rhs/{rhs,rhs_utils,flux_differencing,low_order_graph_viscosity}.jl
 *   src/dg/limiter/{limiter,limiter_utils,subcell,shock_capture}.jl
 *   src/math/compressible_Navier_Stokes.jl, src/timestepping/SSPRK33.jl
 * and src/dg/filter.jl, src/math/nonlinear_solvers.jl
 * for the option set ESLimitedLowOrderPos(LaxFriedrichsOnNodalVal,
 * LaxFriedrichsOnProjectedVal) + SubcellLimiter(PositivityBound, PositivityAndMinEntropyBound or
 * PositivityAndRelaxedMinEntropyBound) or ZhangShuLimiter, each with NoShockCapture or
 * HennemannShockCapture, and
 * NoEntropyProjectionLimiter (Lobatto or Gauss operators) or NodewiseScaledExtrapolation
 * (Gauss operators), with the operators passed in (the reference reads them from
 * Solver.discrete_data).
 *
 * All arrays are column-major, exactly as the Julia arrays they mirror; index arrays
 * hold the reference's 1-based values.
 */
#ifndef P2DE_ORACLE_H
#define P2DE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* SizeData (Solver.jl:173-183) */
    int Nd, Nc, N1D, Nq, Nfp, Nh, K;
    /* Param scalars (Solver.jl:123-170) */
    double gamma, zeta, eta, POSTOL, ZEROTOL, CFL, dt0, t0, T;
    /* 0: bound tuple destructured as the reference does (limiter_utils.jl:11 takes 4 of
     *    the 5 values built at subcell.jl:299, so Urho = Lphi = 0 for PositivityBound);
     * 1: the evident intent, Urho = Urhoe = Inf. */
    int upper_bound_mode;
    /* Param.entropyproj_limiter: 0 NoEntropyProjectionLimiter, 1 NodewiseScaledExtrapolation
     * (filter.jl:6-130, rhs.jl:113-133, flux_differencing.jl:288-328; Gauss basis only).  The
     * basis itself enters through the operators: Vf, Vf_low, MinvVhT, MinvVfT, Srsh_db. */
    int entropyproj_limiter;
    /* Param.rhs_limiter: 0 SubcellLimiter, 1 ZhangShuLimiter (zhangshu.jl:4-56) */
    int rhs_limiter;
    /* SubcellLimiter bound: 0 PositivityBound, 1 PositivityAndMinEntropyBound,
     * 2 PositivityAndRelaxedMinEntropyBound (subcell.jl:4-75, 927-956, limiter_utils.jl:19-50),
     * 3 PositivityAndCellEntropyBound, 4 PositivityAndRelaxedCellEntropyBound(cell_beta)
     * (subcell.jl:458-823; 2D); 5..9 the same entropy parts with TVD density bounds instead of
     * positivity ones: TVDBound, TVDAndMinEntropyBound, TVDAndRelaxedMinEntropyBound,
     * TVDAndCellEntropyBound, TVDAndRelaxedCellEntropyBound (subcell.jl:87-141, 350-368) */
    int bound;
    /* shockcapture: 0 NoShockCapture, 1 HennemannShockCapture(shock_a, shock_c)
     * (shock_capture.jl:10-132) */
    int shockcapture;
    double shock_a, shock_c;
    double cell_beta;
    /* ESLimitedLowOrderPos low_order_surface_flux: 0 LaxFriedrichsOnNodalVal, 1
     * LaxFriedrichsOnProjectedVal (low_order_graph_viscosity.jl:63-71, 249-327) */
    int low_order_surface_flux;
    /* Operators (Solver.jl:192-212) */
    const double *Srsh_db;   /* Nd blocks of Nh x Nh */
    const double *Srs0;      /* Nd blocks of Nq x Nq */
    int nnz_h;
    const int64_t *Srsh_nnz; /* nnz_h pairs (i, j), 1-based */
    int nnz_0;
    const int64_t *Srs0_nnz;
    const double *Brs;       /* Nd blocks of Nfp x Nfp (diagonal) */
    const double *MinvVhT;   /* Nq x Nh */
    const double *MinvVfT;   /* Nq x Nfp */
    const double *Vf;        /* Nfp x Nq */
    const double *Vf_low;    /* Nfp x Nq */
    const double *wq;        /* Nq */
    const int64_t *fq2q;     /* Nfp, 1-based */
    const int64_t *q2fq;     /* Nq x 2, 1-based, 0 = absent (column-major: q2fq[i + Nq*m]) */
    const double *VDM_inv;   /* Nq x Nq, inverse of the orthonormal-Legendre Vandermonde (smoothness
                                indicator; may be NULL when no option reads it) */
    /* GeomData (Solver.jl:185-190) */
    const double *Jq;        /* Nq x K */
    const double *GJh;       /* NGEO blocks of Nh x K: rxJh (, sxJh, ryJh, syJh) */
    /* BCData (StateParam.jl:1-7) */
    const int64_t *mapP;     /* Nfp x K, 1-based linear index into (Nfp, K) */
    int nI;
    const int64_t *mapI;
    const double *Ival;      /* nI x Nc, component fastest */
    int nO;
    const int64_t *mapO;
} oracle_problem;

typedef struct oracle_ctx oracle_ctx;

/* Fields readable after oracle_rhs (State.jl:1-185 names). */
enum {
    ORACLE_RHSU = 0,   /* (Nq, K) x Nc */
    ORACLE_RHSH,       /* (Nq, K) x Nc */
    ORACLE_RHSL,       /* (Nq, K) x Nc */
    ORACLE_RHSXYH,     /* (Nq, K) x Nd x Nc */
    ORACLE_RHSXYL,     /* (Nq, K) x Nd x Nc */
    ORACLE_L_LOCAL,    /* (Nq+N1D, Nd, K, 3) */
    ORACLE_BF_H,       /* (Nfp, K) x Nd x Nc */
    ORACLE_BF_L,       /* (Nfp, K) x Nd x Nc */
    ORACLE_U_TILDE,    /* (Nh, K) x Nc */
    ORACLE_QF1,        /* (Nh, K) x Nd x Nc */
    ORACLE_LAMBDA,     /* (Nq, Nq, K) */
    ORACLE_LAMBDAB,    /* (Nfp, K) */
    ORACLE_FBAR_H,     /* Nd blocks of (Nq+N1D, K) x Nc */
    ORACLE_FBAR_L,
    ORACLE_VQ,         /* (Nq, K) x Nc */
    ORACLE_PSI_TILDE,  /* (Nh, K) x Nd */
    ORACLE_FSTAR_H,    /* (Nfp, K) x Nd x Nc */
    ORACLE_FSTAR_L,
    ORACLE_THETA_LOCAL, /* (Nfp, K, 3) entropy-projection limiting parameters */
    ORACLE_SMOOTH_INDICATOR, /* (K) */
    ORACLE_BLENDING_FACTOR,  /* (K, 3) */
    ORACLE_SMOOTH_FACTOR,    /* (K, 3) */
    ORACLE_LBOUND_S,         /* (Nq, K) lower bound of the modified specific entropy */
    ORACLE_L_ZHANGSHU,       /* (K, 3) Zhang-Shu limiting parameters */
    ORACLE_S_MODIFIED_MIN,   /* (1) */
    ORACLE_LBOUND_RHO,       /* (Nq, K) TVD density bounds */
    ORACLE_UBOUND_RHO,
    ORACLE_NFIELDS
};

int oracle_create(const oracle_problem *p, oracle_ctx **out);
void oracle_destroy(oracle_ctx *c);
/* Number of doubles in a field, and a read-only pointer to it. */
int64_t oracle_field_size(const oracle_ctx *c, int field);
const double *oracle_field(const oracle_ctx *c, int field);

/* rhs!(state, solver, state_param, TimeParam(t, dt, nstage)) (rhs.jl:5-13).
 * Uq: (Nq, K) x Nc.  Returns the CFL dt at nstage 1, dt otherwise, in *dt_out. */
int oracle_rhs(oracle_ctx *c, const double *Uq, double t, double dt, int nstage, double *dt_out);

/* One iteration of the SSP33! time loop (SSPRK33.jl:29-43), in place on Uq;
 * t is advanced by the step taken, whose size is written to *dt_out. */
int oracle_ssprk3_step(oracle_ctx *c, double *Uq, double *t, double *dt_out);

/* Positivity of a state: counts nodes with rho <= 0 or p <= 0 (or NaN);
 * first offending node index (0-based node*K order: i + Nq*k) in *first, -1 if none. */
int64_t oracle_count_nonpositive(const oracle_ctx *c, const double *Uq, int64_t *first);

/* check_conservation (utils.jl:1-12): sum_k sum_i J wq U summed over components. */
double oracle_check_conservation(const oracle_ctx *c, const double *Uq);

/* Pointwise evaluations of the restated physics, for the tests that pin them.
 * fS takes conservative states and forms U() (flux_differencing.jl:179-191) itself. */
void oracle_pt_v_of_u(int Nd, double gamma, const double *U, double *V);
void oracle_pt_u_of_v(int Nd, double gamma, const double *V, double *U);
void oracle_pt_flux(int Nd, double gamma, const double *U, double *F);
void oracle_pt_fS(int Nd, double gamma, const double *UL, const double *UR, double *F);
double oracle_pt_logmean(double aL, double aR);
double oracle_pt_wavespeed(int Nd, double gamma, const double *U, const double *n);
double oracle_pt_limiting_param(int Nd, double zeta, double ZEROTOL, int upper_bound_mode,
                                const double *uL, const double *P);
double oracle_pt_find_alpha(int Nd, double POSTOL, const double *ui, const double *ut);

/* OpenMP threads used by the loops (0 = runtime default). */
void oracle_set_threads(int n);
int oracle_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
