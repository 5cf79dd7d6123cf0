/*
 * p2de_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE; see p2de_oracle.h).
 *
 * Each function names the reference function it restates (file:line under src/).
 * The arithmetic keeps the reference's operation order and IEEE semantics: build with
 * -O2 -fno-fast-math -ffp-contract=off so no product/sum is fused or reassociated.
 * Julia's min/max (NaN-propagating, -0 < +0) and its generic 2-norm are reproduced.
 *
 * Parallelism mirrors Polyester @batch: OpenMP parallel-for over elements.  The one
 * cross-element write of the reference (2D symmetrize, subcell.jl:428-455, which writes
 * the neighbour's entry from inside @batch) is done as a two-pass gather here so the
 * C code has no data race; the values are the same because both sides store the min.
 */
#include "p2de_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NC_MAX 4
#define ND_MAX 2

/* ------------------------------------------------------------------------------- */
/* Julia Base semantics                                                             */
/* ------------------------------------------------------------------------------- */
/* Base.min / Base.max for IEEE floats (Julia 1.8 base/math.jl) */
static inline double jmin(double x, double y) {
    double diff = x - y;
    double am = signbit(diff) ? x : y;
    return (isnan(x) || isnan(y)) ? diff : am;
}
static inline double jmax(double x, double y) {
    double diff = x - y;
    double am = signbit(diff) ? y : x;
    return (isnan(x) || isnan(y)) ? diff : am;
}
/* LinearAlgebra.generic_norm2 on a 2-tuple */
static double jnorm2(double a, double b) {
    double maxabs = jmax(fabs(a), fabs(b));
    if (maxabs == 0.0 || isinf(maxabs) || isnan(maxabs)) return maxabs;
    if (isfinite(2.0 * maxabs * maxabs) && (maxabs * maxabs) != 0.0) {
        double s = a * a;
        s += b * b;
        return sqrt(s);
    }
    double s = (fabs(a) / maxabs) * (fabs(a) / maxabs);
    s += (fabs(b) / maxabs) * (fabs(b) / maxabs);
    return maxabs * sqrt(s);
}

/* ------------------------------------------------------------------------------- */
/* Pointwise Euler physics: src/math/compressible_Navier_Stokes.jl                  */
/* ------------------------------------------------------------------------------- */
/* pfun :18-28 */
static inline double pfun(int Nd, double g, const double *U) {
    if (Nd == 1) return (g - 1.0) * (U[2] - 0.5 * (U[1] * U[1]) / U[0]);
    return (g - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0]);
}
/* betafun :42-45 */
static inline double betafun(int Nd, double g, const double *U) {
    double p = pfun(Nd, g, U);
    return U[0] / (2.0 * p);
}
/* rhoe_ufun :70-78 */
static inline double rhoe_ufun(int Nd, const double *U) {
    if (Nd == 1) return U[2] - 0.5 * U[1] * U[1] / U[0];
    return U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0];
}
/* wavespeed_estimate :48-62 (2D rotates to a 1D state along n, then uses the 1D formula) */
static inline double wavespeed1(double g, double rho, double m, double E) {
    double U1[3] = {rho, m, E};
    double p = pfun(1, g, U1);
    return fabs(m / rho) + sqrt(g * p / rho);
}
static inline double wavespeed(int Nd, double g, const double *U, const double *n) {
    if (Nd == 1) return wavespeed1(g, U[0], U[1], U[2]);
    return wavespeed1(g, U[0], n[0] * U[1] + n[1] * U[2], U[3]);
}
/* sfun :64-68 */
static inline double sfun(int Nd, double g, const double *U) {
    double p = pfun(Nd, g, U);
    return log(p / pow(U[0], g));
}
/* v_ufun :113-122, :134-144 */
static void v_ufun(int Nd, double g, const double *U, double *V) {
    double s = sfun(Nd, g, U);
    double p = pfun(Nd, g, U);
    if (Nd == 1) {
        V[0] = (g + 1.0 - s) - (g - 1.0) * U[2] / p;
        V[1] = U[1] * (g - 1.0) / p;
        V[2] = -U[0] * (g - 1.0) / p;
    } else {
        V[0] = (g + 1.0 - s) - (g - 1.0) * U[3] / p;
        V[1] = U[1] * (g - 1.0) / p;
        V[2] = U[2] * (g - 1.0) / p;
        V[3] = -U[0] * (g - 1.0) / p;
    }
}
/* s_vfun :87-97, rhoe_vfun :99-104 */
static inline double rhoe_vfun(int Nd, double g, const double *V) {
    double s, vE;
    if (Nd == 1) {
        s = g - V[0] + V[1] * V[1] / (2.0 * V[2]);
        vE = V[2];
    } else {
        s = g - V[0] + (V[1] * V[1] + V[2] * V[2]) / (2.0 * V[3]);
        vE = V[3];
    }
    return pow((g - 1.0) / pow(-vE, g), 1.0 / (g - 1.0)) * exp(-s / (g - 1.0));
}
/* u_vfun :146-163 */
static void u_vfun(int Nd, double g, const double *V, double *U) {
    double r = rhoe_vfun(Nd, g, V);
    if (Nd == 1) {
        U[0] = -r * V[2];
        U[1] = r * V[1];
        U[2] = r * (1.0 - V[1] * V[1] / (2.0 * V[2]));
    } else {
        U[0] = -r * V[3];
        U[1] = r * V[1];
        U[2] = r * V[2];
        U[3] = r * (1.0 - (V[1] * V[1] + V[2] * V[2]) / (2.0 * V[3]));
    }
}
/* fluxes :165-194; F[d*Nc + c] */
static void fluxes(int Nd, double g, const double *U, double *F) {
    if (Nd == 1) {
        double p = pfun(1, g, U);
        double u = U[1] / U[0];
        F[0] = U[1];
        F[1] = U[1] * u + p;
        F[2] = u * (U[2] + p);
        return;
    }
    double p = pfun(2, g, U);
    double u = U[1] / U[0];
    double v = U[2] / U[0];
    double rhouv = U[0] * u * v;
    double Ep = U[3] + p;
    F[0] = U[1];
    F[1] = U[1] * u + p;
    F[2] = rhouv;
    F[3] = u * Ep;
    F[4] = U[2];
    F[5] = rhouv;
    F[6] = U[2] * v + p;
    F[7] = v * Ep;
}
/* logmean :307-321 (Winters et al. series with the reference's coefficients) */
static inline double logmean(double aL, double aR, double logL, double logR) {
    double da = aR - aL;
    double aavg = 0.5 * (aR + aL);
    double f = da / aavg;
    double v = f * f;
    if (fabs(f) < 1e-4) {
        return aavg * (1.0 + v * (-0.2 - v * (0.0512 - v * 0.026038857142857)));
    }
    return -da / (logL - logR);
}
/* fS :196-249.  UL/UR = (rho, u, [v,] beta, log rho, log beta); F[d*Nc + c] */
static void fS(int Nd, double g, const double *UL, const double *UR, double *F) {
    if (Nd == 1) {
        double rhoL = UL[0], uL = UL[1], betaL = UL[2], rhologL = UL[3], betalogL = UL[4];
        double rhoR = UR[0], uR = UR[1], betaR = UR[2], rhologR = UR[3], betalogR = UR[4];
        double rholog = logmean(rhoL, rhoR, rhologL, rhologR);
        double betalog = logmean(betaL, betaR, betalogL, betalogR);
        double rhoavg = 0.5 * (rhoL + rhoR);
        double uavg = 0.5 * (uL + uR);
        double unorm = uL * uR;
        double pa = rhoavg / (betaL + betaR);
        double f4aux = rholog / (2.0 * (g - 1.0) * betalog) + pa + 0.5 * rholog * unorm;
        F[0] = rholog * uavg;
        F[1] = F[0] * uavg + pa;
        F[2] = f4aux * uavg;
        return;
    }
    double rhoL = UL[0], uL = UL[1], vL = UL[2], betaL = UL[3], rhologL = UL[4], betalogL = UL[5];
    double rhoR = UR[0], uR = UR[1], vR = UR[2], betaR = UR[3], rhologR = UR[4], betalogR = UR[5];
    double rholog = logmean(rhoL, rhoR, rhologL, rhologR);
    double betalog = logmean(betaL, betaR, betalogL, betalogR);
    double rhoavg = 0.5 * (rhoL + rhoR);
    double uavg = 0.5 * (uL + uR);
    double vavg = 0.5 * (vL + vR);
    double unorm = uL * uR + vL * vR;
    double pa = rhoavg / (betaL + betaR);
    double f4aux = rholog / (2.0 * (g - 1.0) * betalog) + pa + 0.5 * rholog * unorm;
    double FxS1 = rholog * uavg;
    double FxS2 = FxS1 * uavg + pa;
    double FxS3 = FxS1 * vavg;
    double FxS4 = f4aux * uavg;
    double FyS1 = rholog * vavg;
    double FyS3 = FyS1 * vavg + pa;
    double FyS4 = f4aux * vavg;
    F[0] = FxS1; F[1] = FxS2; F[2] = FxS3; F[3] = FxS4;
    F[4] = FyS1; F[5] = FxS3; F[6] = FyS3; F[7] = FyS4;
}

/* ------------------------------------------------------------------------------- */
/* Context                                                                          */
/* ------------------------------------------------------------------------------- */
struct oracle_ctx {
    oracle_problem p;
    int NGEO, NLs; /* NLs = Nq + N1D */
    /* Preallocation (State.jl:1-26) */
    double *vq, *v_tilde, *u_tilde, *psi_tilde;
    double *rhsH, *rhsL, *rhsU, *rhsxyH, *rhsxyL, *rhsxyU;
    double *BF_H, *BF_L, *fstar_H, *fstar_L, *L_local;
    double *resW, *resZ;
    /* LowOrderPositivityCache (State.jl:29-50) */
    double *flux, *Q0F1, *wavespeed_f, *Uf, *uPL, *lambda, *lambdaB;
    /* FluxDiffCache (State.jl:52-85) */
    double *beta, *rholog, *betalog, *uPH, *betaP, *rhologP, *betalogP, *lam, *LFc;
    double *QF1, *MinvVhTQF1, *MinvVfTBF1;
    /* SubcellLimiterCache (State.jl:131-185) */
    double *f_bar_H, *f_bar_L, *f_bar_lim, *lbound_s_modified, *uL, *L_copy;
    double *rhoL, *lbound_rho, *ubound_rho;   /* TVD bounds (SubcellLimiterCache State.jl:132,146-147) */
    /* ShockCaptureCache */
    double *blending_factor; /* (K, 3) */
    double *smooth_factor;   /* (K, 3) */
    double *dtarr;
    /* EntropyProjectionLimiterCache (State.jl:87-129) and theta_local (State.jl:20) */
    double *epUf, *epVUf, *eprhoef, *theta_local;
    /* smoothness indicator (State.jl:23-25), entropy bounds (State.jl:148-153), Zhang-Shu L */
    double *indicator, *indicator_modal, *smooth_indicator, *s_modified, *L_zs;
    double *alpha;           /* (Nfp, K) LaxFriedrichsOnProjectedVal CFL factor */
    double s_modified_min;
    int nthreads_alloc;
    double *stage_buf;       /* rhsU copy for the RK driver */
};

static int g_threads = 0;
void oracle_set_threads(int n) { g_threads = n; }
int oracle_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(nthr)")

static double *zalloc(int64_t n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

int oracle_create(const oracle_problem *p, oracle_ctx **out) {
    if (!p || !out) return 1;
    if (p->Nd != 1 && p->Nd != 2) return 2;
    if (p->Nc != p->Nd + 2) return 2;
    if (p->entropyproj_limiter != 0 && p->entropyproj_limiter != 1) return 2;
    if (p->rhs_limiter < 0 || p->rhs_limiter > 1 || p->bound < 0 || p->bound > 9 || p->shockcapture < 0 ||
        p->shockcapture > 1 || p->low_order_surface_flux < 0 || p->low_order_surface_flux > 1)
        return 2;
    if (!p->VDM_inv && (p->shockcapture != 0 || p->bound % 5 == 2 || p->bound % 5 == 4)) return 2;
    /* the cell-entropy bounds size their per-element scratch (Nq - N1D) x K (State.jl:182-183),
     * which is empty in 1D, so the reference indexes out of bounds there (subcell.jl:503-505) */
    if (p->bound % 5 >= 3 && p->Nd != 2) return 2;
    oracle_ctx *c = (oracle_ctx *)calloc(1, sizeof(oracle_ctx));
    if (!c) return 3;
    c->p = *p;
    int64_t K = p->K, Nq = p->Nq, Nh = p->Nh, Nfp = p->Nfp, Nc = p->Nc, Nd = p->Nd;
    c->NGEO = Nd == 1 ? 1 : 4;
    c->NLs = p->Nq + p->N1D;
    int64_t NLs = c->NLs;
    c->vq = zalloc(Nq * K * Nc);
    c->v_tilde = zalloc(Nh * K * Nc);
    c->u_tilde = zalloc(Nh * K * Nc);
    c->psi_tilde = zalloc(Nh * K * Nd);
    c->rhsH = zalloc(Nq * K * Nc);
    c->rhsL = zalloc(Nq * K * Nc);
    c->rhsU = zalloc(Nq * K * Nc);
    c->rhsxyH = zalloc(Nq * K * Nd * Nc);
    c->rhsxyL = zalloc(Nq * K * Nd * Nc);
    c->rhsxyU = zalloc(Nq * K * Nd * Nc);
    c->BF_H = zalloc(Nfp * K * Nd * Nc);
    c->BF_L = zalloc(Nfp * K * Nd * Nc);
    c->fstar_H = zalloc(Nfp * K * Nd * Nc);
    c->fstar_L = zalloc(Nfp * K * Nd * Nc);
    c->L_local = zalloc(NLs * Nd * K * 3);
    c->L_copy = zalloc(NLs * Nd * K);
    c->resW = zalloc(Nq * K * Nc);
    c->resZ = zalloc(Nq * K * Nc);
    c->flux = zalloc(Nh * K * Nd * Nc);
    c->Q0F1 = zalloc(Nq * K * Nd * Nc);
    c->wavespeed_f = zalloc(Nfp * K);
    c->Uf = zalloc(Nfp * K * Nc);
    c->uPL = zalloc(Nfp * K * Nc);
    c->lambda = zalloc(Nq * Nq * K);
    c->lambdaB = zalloc(Nfp * K);
    c->beta = zalloc(Nh * K);
    c->rholog = zalloc(Nh * K);
    c->betalog = zalloc(Nh * K);
    c->uPH = zalloc(Nfp * K * Nc);
    c->betaP = zalloc(Nfp * K);
    c->rhologP = zalloc(Nfp * K);
    c->betalogP = zalloc(Nfp * K);
    c->lam = zalloc(Nfp * K);
    c->LFc = zalloc(Nfp * K);
    c->QF1 = zalloc(Nh * K * Nd * Nc);
    c->MinvVhTQF1 = zalloc(Nq * K * Nd * Nc);
    c->MinvVfTBF1 = zalloc(Nq * K * Nd * Nc);
    c->f_bar_H = zalloc(Nd * NLs * K * Nc);
    c->f_bar_L = zalloc(Nd * NLs * K * Nc);
    c->f_bar_lim = zalloc(Nd * NLs * K * Nc);
    c->lbound_s_modified = zalloc(Nq * K);
    c->rhoL = zalloc(Nq * K);
    c->lbound_rho = zalloc(Nq * K);
    c->ubound_rho = zalloc(Nq * K);
    c->uL = zalloc(Nq * K * Nc);
    c->blending_factor = zalloc(K * 3);
    c->smooth_factor = zalloc(K * 3);
    c->dtarr = zalloc(K);
    c->stage_buf = zalloc(Nq * K * Nc);
    c->epUf = zalloc(Nfp * K * Nc);
    c->epVUf = zalloc(Nfp * K * Nc);
    c->eprhoef = zalloc(Nfp * K);
    c->theta_local = zalloc(Nfp * K * 3);
    for (int64_t x = 0; x < Nfp * K * 3; x++) c->theta_local[x] = 1.0;
    c->indicator = zalloc(Nq * K);
    c->indicator_modal = zalloc(Nq * K);
    c->smooth_indicator = zalloc(K);
    c->s_modified = zalloc(Nq * K);
    c->L_zs = zalloc(K * 3);
    c->alpha = zalloc(Nfp * K);
    c->s_modified_min = 0.0;   /* SubcellLimiterCache: zeros(Float64, 1) (State.jl:180) */
    *out = c;
    return 0;
}

void oracle_destroy(oracle_ctx *c) {
    if (!c) return;
    double **f[] = {&c->vq, &c->v_tilde, &c->u_tilde, &c->psi_tilde, &c->rhsH, &c->rhsL, &c->rhsU,
                    &c->rhsxyH, &c->rhsxyL, &c->rhsxyU, &c->BF_H, &c->BF_L, &c->fstar_H, &c->fstar_L,
                    &c->L_local, &c->L_copy, &c->resW, &c->resZ, &c->flux, &c->Q0F1, &c->wavespeed_f,
                    &c->Uf, &c->uPL, &c->lambda, &c->lambdaB, &c->beta, &c->rholog, &c->betalog,
                    &c->uPH, &c->betaP, &c->rhologP, &c->betalogP, &c->lam, &c->LFc, &c->QF1,
                    &c->MinvVhTQF1, &c->MinvVfTBF1, &c->f_bar_H, &c->f_bar_L, &c->f_bar_lim,
                    &c->lbound_s_modified, &c->uL, &c->blending_factor, &c->smooth_factor,
                    &c->dtarr, &c->stage_buf, &c->epUf, &c->epVUf, &c->eprhoef, &c->theta_local,
                    &c->indicator, &c->indicator_modal, &c->smooth_indicator, &c->s_modified, &c->L_zs,
                    &c->alpha, &c->rhoL, &c->lbound_rho, &c->ubound_rho};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); i++) free(*f[i]);
    free(c);
}

int64_t oracle_field_size(const oracle_ctx *c, int field) {
    const oracle_problem *p = &c->p;
    int64_t K = p->K, Nq = p->Nq, Nh = p->Nh, Nfp = p->Nfp, Nc = p->Nc, Nd = p->Nd;
    switch (field) {
    case ORACLE_RHSU: case ORACLE_RHSH: case ORACLE_RHSL: case ORACLE_VQ: return Nq * K * Nc;
    case ORACLE_RHSXYH: case ORACLE_RHSXYL: return Nq * K * Nd * Nc;
    case ORACLE_L_LOCAL: return (int64_t)c->NLs * Nd * K * 3;
    case ORACLE_BF_H: case ORACLE_BF_L: case ORACLE_FSTAR_H: case ORACLE_FSTAR_L: return Nfp * K * Nd * Nc;
    case ORACLE_U_TILDE: return Nh * K * Nc;
    case ORACLE_QF1: return Nh * K * Nd * Nc;
    case ORACLE_LAMBDA: return Nq * Nq * K;
    case ORACLE_LAMBDAB: return Nfp * K;
    case ORACLE_FBAR_H: case ORACLE_FBAR_L: return Nd * (int64_t)c->NLs * K * Nc;
    case ORACLE_PSI_TILDE: return Nh * K * Nd;
    case ORACLE_THETA_LOCAL: return Nfp * K * 3;
    case ORACLE_SMOOTH_INDICATOR: return K;
    case ORACLE_BLENDING_FACTOR: case ORACLE_SMOOTH_FACTOR: case ORACLE_L_ZHANGSHU: return K * 3;
    case ORACLE_LBOUND_S: return Nq * K;
    case ORACLE_S_MODIFIED_MIN: return 1;
    case ORACLE_LBOUND_RHO: case ORACLE_UBOUND_RHO: return Nq * K;
    default: return -1;
    }
}

const double *oracle_field(const oracle_ctx *c, int field) {
    switch (field) {
    case ORACLE_RHSU: return c->rhsU;
    case ORACLE_RHSH: return c->rhsH;
    case ORACLE_RHSL: return c->rhsL;
    case ORACLE_RHSXYH: return c->rhsxyH;
    case ORACLE_RHSXYL: return c->rhsxyL;
    case ORACLE_L_LOCAL: return c->L_local;
    case ORACLE_BF_H: return c->BF_H;
    case ORACLE_BF_L: return c->BF_L;
    case ORACLE_U_TILDE: return c->u_tilde;
    case ORACLE_QF1: return c->QF1;
    case ORACLE_LAMBDA: return c->lambda;
    case ORACLE_LAMBDAB: return c->lambdaB;
    case ORACLE_FBAR_H: return c->f_bar_H;
    case ORACLE_FBAR_L: return c->f_bar_L;
    case ORACLE_VQ: return c->vq;
    case ORACLE_PSI_TILDE: return c->psi_tilde;
    case ORACLE_FSTAR_H: return c->fstar_H;
    case ORACLE_FSTAR_L: return c->fstar_L;
    case ORACLE_THETA_LOCAL: return c->theta_local;
    case ORACLE_SMOOTH_INDICATOR: return c->smooth_indicator;
    case ORACLE_BLENDING_FACTOR: return c->blending_factor;
    case ORACLE_SMOOTH_FACTOR: return c->smooth_factor;
    case ORACLE_LBOUND_S: return c->lbound_s_modified;
    case ORACLE_L_ZHANGSHU: return c->L_zs;
    case ORACLE_S_MODIFIED_MIN: return &c->s_modified_min;
    case ORACLE_LBOUND_RHO: return c->lbound_rho;
    case ORACLE_UBOUND_RHO: return c->ubound_rho;
    default: return NULL;
    }
}

/* ------------------------------------------------------------------------------- */
/* Geometry helpers: rhs_utils.jl                                                    */
/* ------------------------------------------------------------------------------- */
#define IDX_V(i, k, N, Nc) ((int64_t)(Nc) * ((int64_t)(i) + (int64_t)(N) * (int64_t)(k)))
#define IDX_XY(d, i, k, N, Nd, Nc) ((int64_t)(Nc) * ((int64_t)(d) + (int64_t)(Nd) * ((int64_t)(i) + (int64_t)(N) * (int64_t)(k))))

static inline double gjh(const oracle_ctx *c, int g, int i, int k) {
    return c->p.GJh[(int64_t)g * c->p.Nh * c->p.K + i + (int64_t)c->p.Nh * k];
}
/* reference_to_physical :1-12 */
static inline void ref2phys(const oracle_ctx *c, const double *Urs, int i, int k, double *out) {
    if (c->p.Nd == 1) {
        out[0] = gjh(c, 0, i, k) * Urs[0];
        return;
    }
    double rxJ = gjh(c, 0, i, k), sxJ = gjh(c, 1, i, k), ryJ = gjh(c, 2, i, k), syJ = gjh(c, 3, i, k);
    out[0] = rxJ * Urs[0] + sxJ * Urs[1];
    out[1] = ryJ * Urs[0] + syJ * Urs[1];
}
/* Bx :13-27 (face node i, 0-based; geometric factors at hybrid index Nq+i).  The
 * reference's 1D method passes its arguments in the wrong order (:18) and would throw;
 * this computes its evident intent rxJh * Br[i,i]. */
static inline void Bx(const oracle_ctx *c, int i, int k, double *B) {
    int Nfp = c->p.Nfp;
    double Urs[2];
    for (int d = 0; d < c->p.Nd; d++) Urs[d] = c->p.Brs[(int64_t)d * Nfp * Nfp + i + (int64_t)Nfp * i];
    ref2phys(c, Urs, c->p.Nq + i, k, B);
}
/* Bx_with_n :29-37 */
static inline double Bx_with_n(const oracle_ctx *c, int i, int k, double *B) {
    Bx(c, i, k, B);
    if (c->p.Nd == 1) return fabs(B[0]);
    return jnorm2(B[0], B[1]);
}
/* Sx :39-51 (geometric factors at node i, 0-based hybrid indices) */
static inline void Sx(const oracle_ctx *c, int i, int j, int k, double *S) {
    int Nh = c->p.Nh;
    double Urs[2];
    for (int d = 0; d < c->p.Nd; d++) Urs[d] = c->p.Srsh_db[(int64_t)d * Nh * Nh + i + (int64_t)Nh * j];
    ref2phys(c, Urs, i, k, S);
}
/* Sx0_with_n :53-75 */
static inline double Sx0_with_n(const oracle_ctx *c, int i, int j, int k, double *S) {
    int Nq = c->p.Nq;
    double Urs[2];
    for (int d = 0; d < c->p.Nd; d++) Urs[d] = c->p.Srs0[(int64_t)d * Nq * Nq + i + (int64_t)Nq * j];
    ref2phys(c, Urs, i, k, S);
    if (c->p.Nd == 1) return fabs(S[0]);
    return jnorm2(S[0], S[1]);
}

/* ------------------------------------------------------------------------------- */
/* Entropy-projection limiter: init_rhs!(::ScaledExtrapolation) rhs.jl:19-23,        */
/* compute_entropyproj_limiting_param!(::GaussCollocation) filter.jl:6-15            */
/* ------------------------------------------------------------------------------- */
#define THETA(i, k, s) (c->theta_local[(int64_t)(i) + (int64_t)c->p.Nfp * ((int64_t)(k) + (int64_t)c->p.K * (s))])

/* calc_face_values!(::CompressibleEulerIdealGas) filter.jl:26-42; the two mul! calls are the
 * generic matrix-vector product, which adds column by column from zero */
static void calc_face_values(oracle_ctx *c, const double *Uq, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nfp = p->Nfp, Nc = p->Nc, Nd = p->Nd;
    double g = p->gamma;
    OMP_FOR
    for (int k = 0; k < K; k++) {
        for (int i = 0; i < Nfp; i++) {
            double *uf = c->epUf + IDX_V(i, k, Nfp, Nc);
            for (int cc = 0; cc < Nc; cc++) uf[cc] = 0.0;
            for (int j = 0; j < Nq; j++) {
                double a = p->Vf[i + (int64_t)Nfp * j];
                for (int cc = 0; cc < Nc; cc++) uf[cc] += a * Uq[IDX_V(j, k, Nq, Nc) + cc];
            }
        }
        for (int i = 0; i < Nq; i++) v_ufun(Nd, g, Uq + IDX_V(i, k, Nq, Nc), c->vq + IDX_V(i, k, Nq, Nc));
        for (int i = 0; i < Nfp; i++) c->eprhoef[i + (int64_t)Nfp * k] = rhoe_ufun(Nd, c->epUf + IDX_V(i, k, Nfp, Nc));
        for (int i = 0; i < Nfp; i++) {
            double *vf = c->epVUf + IDX_V(i, k, Nfp, Nc);
            for (int cc = 0; cc < Nc; cc++) vf[cc] = 0.0;
            for (int j = 0; j < Nq; j++) {
                double a = p->Vf[i + (int64_t)Nfp * j];
                for (int cc = 0; cc < Nc; cc++) vf[cc] += a * c->vq[IDX_V(j, k, Nq, Nc) + cc];
            }
        }
    }
}

/* projected entropy variables of face node i with parameter l (rhs.jl:87-92, filter.jl:118-121):
 * v = sum_j (l*Vf[i,j] + (1-l)*Vf_low[i,j]) vq_j, from zero, j ascending */
static void project_face_v(const oracle_ctx *c, int i, int k, double l, double *vt) {
    const oracle_problem *p = &c->p;
    int Nq = p->Nq, Nfp = p->Nfp, Nc = p->Nc;
    for (int cc = 0; cc < Nc; cc++) vt[cc] = 0.0;
    for (int j = 0; j < Nq; j++) {
        double coef = l * p->Vf[i + (int64_t)Nfp * j] + (1.0 - l) * p->Vf_low[i + (int64_t)Nfp * j];
        for (int cc = 0; cc < Nc; cc++) vt[cc] += coef * c->vq[IDX_V(j, k, Nq, Nc) + cc];
    }
}

/* update_and_check_bound_limited_entropyproj_var_on_face_node! filter.jl:101-107 with
 * update_limited_entropyproj_vars_on_face_node! :109-130 and check_bound_on_face_node :87-98 */
static int theta_admissible(const oracle_ctx *c, int i, int k, double theta) {
    const oracle_problem *p = &c->p;
    int Nfp = p->Nfp, Nc = p->Nc, Nd = p->Nd;
    double eps = p->POSTOL, zeta = p->zeta, eta = p->eta;
    double vt[NC_MAX], ut[NC_MAX];
    project_face_v(c, i, k, theta, vt);
    if (!(vt[Nc - 1] < -eps)) return 0;                   /* well-definedness :123 */
    u_vfun(Nd, p->gamma, vt, ut);
    double v3tilde = vt[Nc - 1], rhotilde = ut[0], rhoetilde = rhoe_ufun(Nd, ut);
    const double *Uf = c->epUf + IDX_V(i, k, Nfp, Nc), *VUf = c->epVUf + IDX_V(i, k, Nfp, Nc);
    double rhoef = c->eprhoef[i + (int64_t)Nfp * k];
    return v3tilde < jmin(zeta * VUf[Nc - 1], -eps) && rhotilde > jmax((1.0 - eta) * Uf[0], eps) &&
           rhotilde < (1.0 + eta) * Uf[0] && rhoetilde > jmax((1.0 - eta) * rhoef, eps) &&
           rhoetilde < (1.0 + eta) * rhoef;
}

/* solve_theta!(::NodewiseScaledExtrapolation) filter.jl:49-59 with bisection
 * nonlinear_solvers.jl:3-20 (x_valid = 0, x_invalid = 1, iterations 0..20) */
static void compute_entropyproj_limiting_param(oracle_ctx *c, const double *Uq, int nstage, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nfp = p->Nfp;
    int s = nstage - 1;
    /* clear_entropyproj_limiting_parameter_cache! :18-20 */
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nfp; i++) THETA(i, k, s) = 1.0;
    calc_face_values(c, Uq, nthr);
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nfp; i++) {
            double x_valid = 0.0, x_invalid = 1.0, th;
            if (theta_admissible(c, i, k, x_invalid)) {
                th = x_invalid;
            } else {
                for (int iter = 0; iter <= 20; iter++) {
                    double x_new = 0.5 * (x_valid + x_invalid);
                    if (theta_admissible(c, i, k, x_new)) x_valid = x_new;
                    else x_invalid = x_new;
                }
                th = x_valid;
            }
            THETA(i, k, s) = th;
        }
}

/* ------------------------------------------------------------------------------- */
/* entropy_projection! rhs.jl:59-133 (NoEntropyProjectionLimiter: l = 1;             */
/* NodewiseScaledExtrapolation: l = theta_local[i, k, nstage])                        */
/* ------------------------------------------------------------------------------- */
static void entropy_projection(oracle_ctx *c, const double *Uq, int nstage, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nfp = p->Nfp, Nh = p->Nh, Nc = p->Nc, Nd = p->Nd;
    double g = p->gamma;
    OMP_FOR
    for (int k = 0; k < K; k++) {
        /* calculate_entropy_var! :72-77 */
        for (int i = 0; i < Nq; i++) v_ufun(Nd, g, Uq + IDX_V(i, k, Nq, Nc), c->vq + IDX_V(i, k, Nq, Nc));
        /* entropy_projection_volume_node! :79-82 */
        for (int i = 0; i < Nq; i++)
            for (int cc = 0; cc < Nc; cc++) {
                c->v_tilde[IDX_V(i, k, Nh, Nc) + cc] = c->vq[IDX_V(i, k, Nq, Nc) + cc];
                c->u_tilde[IDX_V(i, k, Nh, Nc) + cc] = Uq[IDX_V(i, k, Nq, Nc) + cc];
            }
        /* entropy_projection_face_node! :84-94 */
        for (int i = 0; i < Nfp; i++) {
            double l = p->entropyproj_limiter == 1 ? THETA(i, k, nstage - 1) : 1.0;
            double *vt = c->v_tilde + IDX_V(i + Nq, k, Nh, Nc);
            project_face_v(c, i, k, l, vt);
            u_vfun(Nd, g, vt, c->u_tilde + IDX_V(i + Nq, k, Nh, Nc));
        }
    }
}

/* find_alpha(::CompressibleIdealGas) low_order_graph_viscosity.jl:299-328: the smallest alpha
 * (50 bisection steps after doubling) with alpha*u_i - u~ of positive density and internal energy.
 * The doubling loop of the reference has no exit when no alpha works; here it stops once alpha
 * overflows to Inf (returned). */
static int alpha_ok(const oracle_ctx *c, const double *ui, const double *ut, double a) {
    double sub[NC_MAX];
    for (int cc = 0; cc < c->p.Nc; cc++) sub[cc] = a * ui[cc] - ut[cc];
    return sub[0] > c->p.POSTOL && rhoe_ufun(c->p.Nd, sub) > c->p.POSTOL;
}
static double find_alpha(const oracle_ctx *c, const double *ui, const double *ut) {
    double alphaL = 0.0, alphaR = 1.0;
    while (!alpha_ok(c, ui, ut, alphaR)) {
        alphaR = 2.0 * alphaR;
        if (isinf(alphaR)) return alphaR;
    }
    for (int iter = 0; iter < 50; iter++) {
        double alphaM = (alphaL + alphaR) / 2.0;
        if (alpha_ok(c, ui, ut, alphaM)) alphaR = alphaM;
        else alphaL = alphaM;
    }
    return alphaR;
}

/* ------------------------------------------------------------------------------- */
/* rhs_low_graph_visc! low_order_graph_viscosity.jl:4-41                            */
/* ------------------------------------------------------------------------------- */
static double rhs_low_graph_visc(oracle_ctx *c, const double *Uq, double t, double dt, int nstage, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nfp = p->Nfp, Nh = p->Nh, Nc = p->Nc, Nd = p->Nd;
    double g = p->gamma;
    /* calculate_wavespeed_and_inviscid_flux! :43-49 */
    OMP_FOR
    for (int k = 0; k < K; k++) {
        /* update_face_values!(::LaxFriedrichsOnNodalVal) :51-61, (::LaxFriedrichsOnProjectedVal) :63-71 */
        for (int i = 0; i < Nfp; i++) {
            int iq = (int)p->fq2q[i] - 1;
            const double *src = p->low_order_surface_flux == 1 ? c->u_tilde + IDX_V(Nq + i, k, Nh, Nc)
                                                               : Uq + IDX_V(iq, k, Nq, Nc);
            for (int cc = 0; cc < Nc; cc++) c->Uf[IDX_V(i, k, Nfp, Nc) + cc] = src[cc];
        }
        /* update_wavespeed_and_inviscid_flux! :73-92 */
        for (int i = 0; i < Nq; i++) fluxes(Nd, g, Uq + IDX_V(i, k, Nq, Nc), c->flux + IDX_XY(0, i, k, Nh, Nd, Nc));
        for (int i = 0; i < Nfp; i++) {
            const double *u_i = c->Uf + IDX_V(i, k, Nfp, Nc);
            double B[2] = {0.0, 0.0}, n[2] = {0.0, 0.0};
            double nrm = Bx_with_n(c, i, k, B);
            for (int d = 0; d < Nd; d++) n[d] = B[d] / nrm;
            c->wavespeed_f[i + (int64_t)Nfp * k] = wavespeed(Nd, g, u_i, n);
            fluxes(Nd, g, u_i, c->flux + IDX_XY(0, i + Nq, k, Nh, Nd, Nc));
        }
    }
    /* get_uP_and_enforce_BC! :94-124 */
    OMP_FOR
    for (int k = 0; k < K; k++) {
        for (int i = 0; i < Nfp; i++) {
            int64_t m = p->mapP[i + (int64_t)Nfp * k] - 1;
            int64_t iP = m % Nfp, kP = m / Nfp;
            for (int cc = 0; cc < Nc; cc++) c->uPL[IDX_V(i, k, Nfp, Nc) + cc] = c->Uf[IDX_V(iP, kP, Nfp, Nc) + cc];
        }
    }
    for (int i = 0; i < p->nI; i++) {
        int64_t ii = p->mapI[i] - 1;
        for (int cc = 0; cc < Nc; cc++) c->uPL[ii * Nc + cc] = p->Ival[(int64_t)i * Nc + cc];
    }
    for (int i = 0; i < p->nO; i++) {
        int64_t io = p->mapO[i] - 1;
        int64_t iP = io % Nfp, kP = io / Nfp;
        int64_t iq = p->fq2q[iP] - 1;
        for (int cc = 0; cc < Nc; cc++) c->uPL[io * Nc + cc] = Uq[IDX_V(iq, kP, Nq, Nc) + cc];
    }
    /* clear_low_order_rhs! :126-137 */
    memset(c->rhsxyL, 0, sizeof(double) * (size_t)Nq * K * Nd * Nc);
    memset(c->Q0F1, 0, sizeof(double) * (size_t)Nq * K * Nd * Nc);
    /* accumulate_low_order_rhs_volume! :139-173 */
    OMP_FOR
    for (int k = 0; k < K; k++) {
        for (int pidx = 0; pidx < p->nnz_0; pidx++) {
            int i = (int)p->Srs0_nnz[2 * pidx] - 1, j = (int)p->Srs0_nnz[2 * pidx + 1] - 1;
            const double *u_i = Uq + IDX_V(i, k, Nq, Nc), *u_j = Uq + IDX_V(j, k, Nq, Nc);
            double Fxy[ND_MAX * NC_MAX];
            for (int d = 0; d < Nd; d++)
                for (int cc = 0; cc < Nc; cc++)
                    Fxy[d * Nc + cc] = 0.5 * (c->flux[IDX_XY(d, i, k, Nh, Nd, Nc) + cc] + c->flux[IDX_XY(d, j, k, Nh, Nd, Nc) + cc]);
            double S0[2] = {0.0, 0.0}, n_ij[2] = {0.0, 0.0}, n_ji[2] = {0.0, 0.0};
            double nrm = Sx0_with_n(c, i, j, k, S0);
            for (int d = 0; d < Nd; d++) { n_ij[d] = S0[d] / nrm; n_ji[d] = -n_ij[d]; }
            double ws = jmax(wavespeed(Nd, g, u_i, n_ij), wavespeed(Nd, g, u_j, n_ji));
            double lam_ij = nrm * ws;
            c->lambda[i + (int64_t)Nq * (j + (int64_t)Nq * k)] = lam_ij;
            c->lambda[j + (int64_t)Nq * (i + (int64_t)Nq * k)] = lam_ij;
            /* graph_viscosity rhs_utils.jl:105-128 */
            double visc[NC_MAX];
            for (int cc = 0; cc < Nc; cc++) visc[cc] = lam_ij * (u_j[cc] - u_i[cc]);
            int vdir = 0;
            if (Nd == 2) vdir = fabs(S0[0]) > 1e-10 ? 0 : 1;
            for (int d = 0; d < Nd; d++)
                for (int cc = 0; cc < Nc; cc++) {
                    double lD = (d == vdir) ? visc[cc] : 0.0;
                    double v = 2.0 * S0[d] * Fxy[d * Nc + cc] - lD;
                    c->Q0F1[IDX_XY(d, i, k, Nq, Nd, Nc) + cc] += v;
                    c->Q0F1[IDX_XY(d, j, k, Nq, Nd, Nc) + cc] += -v;
                }
        }
    }
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nq; i++)
            for (int x = 0; x < Nd * Nc; x++) c->rhsxyL[IDX_XY(0, i, k, Nq, Nd, Nc) + x] -= c->Q0F1[IDX_XY(0, i, k, Nq, Nd, Nc) + x];
    /* accumulate_low_order_rhs_surface! :175-204 */
    int N1D = p->N1D;
    OMP_FOR
    for (int k = 0; k < K; k++) {
        for (int i = 0; i < Nfp; i++) {
            int64_t m = p->mapP[i + (int64_t)Nfp * k] - 1;
            int64_t iP = m % Nfp, kP = m / Nfp;
            double B[2] = {0.0, 0.0};
            double nrm = Bx_with_n(c, i, k, B);
            double lB = 0.5 * nrm * jmax(c->wavespeed_f[i + (int64_t)Nfp * k], c->wavespeed_f[iP + (int64_t)Nfp * kP]);
            c->lambdaB[i + (int64_t)Nfp * k] = lB;
            const double *uP = c->uPL + IDX_V(i, k, Nfp, Nc), *Uf = c->Uf + IDX_V(i, k, Nfp, Nc);
            double fP[ND_MAX * NC_MAX];
            fluxes(Nd, g, uP, fP);
            double *fs = c->fstar_L + IDX_XY(0, i, k, Nfp, Nd, Nc);
            double *BF = c->BF_L + IDX_XY(0, i, k, Nfp, Nd, Nc);
            for (int d = 0; d < Nd; d++)
                for (int cc = 0; cc < Nc; cc++) {
                    fs[d * Nc + cc] = 0.5 * (c->flux[IDX_XY(d, i + Nq, k, Nh, Nd, Nc) + cc] + fP[d * Nc + cc]);
                    BF[d * Nc + cc] = B[d] * fs[d * Nc + cc];
                }
            /* apply_LF_dissipation_to_BF / _to_fstar (rhs_utils.jl:78-102) */
            int ld = (Nd == 1 || i < 2 * N1D) ? 0 : 1;
            for (int cc = 0; cc < Nc; cc++) {
                double lf = lB * (uP[cc] - Uf[cc]);
                BF[ld * Nc + cc] = BF[ld * Nc + cc] - lf;
                fs[ld * Nc + cc] = fs[ld * Nc + cc] - lf / B[ld];
            }
            int iq = (int)p->fq2q[i] - 1;
            for (int x = 0; x < Nd * Nc; x++) c->rhsxyL[IDX_XY(0, iq, k, Nq, Nd, Nc) + x] -= BF[x];
        }
    }
    /* scale_low_order_rhs_by_mass! :206-220 */
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nq; i++) {
            double wJq = p->Jq[i + (int64_t)Nq * k] * p->wq[i];
            double *r = c->rhsxyL + IDX_XY(0, i, k, Nq, Nd, Nc);
            for (int x = 0; x < Nd * Nc; x++) r[x] = r[x] / wJq;
            for (int cc = 0; cc < Nc; cc++) {
                double s = r[cc];
                for (int d = 1; d < Nd; d++) s = s + r[d * Nc + cc];
                c->rhsL[IDX_V(i, k, Nq, Nc) + cc] = s;
            }
        }
    /* calculate_lambda_and_low_order_CFL! :222-243 (nstage 1 only) */
    if (nstage == 1) {
        double dt0 = jmin(p->CFL * p->dt0, p->T - t);
        OMP_FOR
        for (int k = 0; k < K; k++) {
            /* accumulate_alpha!(::LaxFriedrichsOnProjectedVal) :249-261 */
            if (p->low_order_surface_flux == 1)
                for (int i = 0; i < Nfp; i++) {
                    int iq = (int)p->fq2q[i] - 1;
                    c->alpha[i + (int64_t)Nfp * k] =
                        find_alpha(c, Uq + IDX_V(iq, k, Nq, Nc), c->u_tilde + IDX_V(Nq + i, k, Nh, Nc));
                }
            double dtk = dt0;
            for (int i = 0; i < Nq; i++) {
                double wq_i = p->wq[i];
                double wJq_i = p->Jq[i + (int64_t)Nq * k] * wq_i;
                /* lambda_i :263-281 */
                double li = 0.0;
                for (int j = 0; j < Nq; j++) li += c->lambda[i + (int64_t)Nq * (j + (int64_t)Nq * k)];
                for (int m = 0; m < 2; m++) {
                    int64_t f = p->q2fq[i + (int64_t)Nq * m];
                    if (f <= 0) continue;
                    /* lambda_B_CFL :283-291 */
                    if (p->low_order_surface_flux == 1) {
                        double B[2] = {0.0, 0.0};
                        double nrm = Bx_with_n(c, (int)f - 1, k, B);
                        int64_t fk = (f - 1) + (int64_t)Nfp * k;
                        li += c->alpha[fk] * c->lambdaB[fk] + 0.5 * nrm * c->wavespeed_f[fk];
                    } else {
                        li += c->lambdaB[(f - 1) + (int64_t)Nfp * k];
                    }
                }
                dtk = jmin(dtk, p->CFL * 0.5 * wJq_i / li);
            }
            c->dtarr[k] = dtk;
        }
        double dmin = dt0;
        for (int k = 0; k < K; k++) dmin = jmin(dmin, c->dtarr[k]);
        dt = dmin;
    }
    return dt;
}

/* ------------------------------------------------------------------------------- */
/* rhs_fluxdiff! flux_differencing.jl:4-37                                          */
/* ------------------------------------------------------------------------------- */
static inline void Uvol(const oracle_ctx *c, int idx, int k, double *out) {
    /* U(dim, ::CompressibleIdealGas, ::ChandrashekarFlux, ...) :179-191 */
    const double *u = c->u_tilde + IDX_V(idx, k, c->p.Nh, c->p.Nc);
    int64_t h = idx + (int64_t)c->p.Nh * k;
    out[0] = u[0];
    if (c->p.Nd == 1) {
        out[1] = u[1] / u[0];
        out[2] = c->beta[h]; out[3] = c->rholog[h]; out[4] = c->betalog[h];
    } else {
        out[1] = u[1] / u[0];
        out[2] = u[2] / u[0];
        out[3] = c->beta[h]; out[4] = c->rholog[h]; out[5] = c->betalog[h];
    }
}

static void rhs_fluxdiff(oracle_ctx *c, const double *Uq, int nstage, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nfp = p->Nfp, Nh = p->Nh, Nc = p->Nc, Nd = p->Nd, N1D = p->N1D;
    double g = p->gamma;
    /* calculate_primitive_variables! :39-70 */
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nh; i++) {
            const double *u = c->u_tilde + IDX_V(i, k, Nh, Nc);
            int64_t h = i + (int64_t)Nh * k;
            for (int d = 0; d < Nd; d++) c->psi_tilde[h * Nd + d] = (g - 1.0) * u[1 + d];
            c->beta[h] = betafun(Nd, g, u);
            c->rholog[h] = log(u[0]);
            c->betalog[h] = log(c->beta[h]);
        }
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nfp; i++) {
            int64_t m = p->mapP[i + (int64_t)Nfp * k] - 1;
            int64_t iP = m % Nfp, kP = m / Nfp;
            int64_t hP = Nq + iP + (int64_t)Nh * kP;
            for (int cc = 0; cc < Nc; cc++) c->uPH[IDX_V(i, k, Nfp, Nc) + cc] = c->u_tilde[hP * Nc + cc];
            c->betaP[i + (int64_t)Nfp * k] = c->beta[hP];
            c->rhologP[i + (int64_t)Nfp * k] = c->rholog[hP];
            c->betalogP[i + (int64_t)Nfp * k] = c->betalog[hP];
        }
    /* calculate_interface_dissipation_coeff! :90-114 */
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nfp; i++) {
            double B[2] = {0.0, 0.0}, n[2] = {0.0, 0.0};
            double nrm = Bx_with_n(c, i, k, B);
            for (int d = 0; d < Nd; d++) n[d] = B[d] / nrm;
            c->lam[i + (int64_t)Nfp * k] = wavespeed(Nd, g, c->u_tilde + IDX_V(Nq + i, k, Nh, Nc), n);
            c->LFc[i + (int64_t)Nfp * k] = 0.5 * nrm;
        }
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nfp; i++) {
            int64_t f = i + (int64_t)Nfp * k;
            c->LFc[f] = c->LFc[f] * jmax(c->lam[f], c->lam[p->mapP[f] - 1]);
        }
    /* enforce_BC! :116-151 */
    for (int i = 0; i < p->nI; i++) c->LFc[p->mapI[i] - 1] = 0.0;
    for (int i = 0; i < p->nO; i++) c->LFc[p->mapO[i] - 1] = 0.0;
    for (int i = 0; i < p->nI; i++) {
        int64_t ii = p->mapI[i] - 1;
        double *uP = c->uPH + ii * Nc;
        for (int cc = 0; cc < Nc; cc++) uP[cc] = p->Ival[(int64_t)i * Nc + cc];
        c->betaP[ii] = betafun(Nd, g, uP);
        c->rhologP[ii] = log(uP[0]);
        c->betalogP[ii] = log(c->betaP[ii]);
    }
    for (int i = 0; i < p->nO; i++) {
        int64_t io = p->mapO[i] - 1;
        int64_t iP = io % Nfp, kP = io / Nfp;
        int64_t iq = p->fq2q[iP] - 1;
        double *uP = c->uPH + io * Nc;
        for (int cc = 0; cc < Nc; cc++) uP[cc] = Uq[IDX_V(iq, kP, Nq, Nc) + cc];
        c->betaP[io] = betafun(Nd, g, uP);
        c->rhologP[io] = log(uP[0]);
        c->betalogP[io] = log(c->betaP[io]);
    }
    /* clear_flux_differencing_cache! :153-162 */
    memset(c->QF1, 0, sizeof(double) * (size_t)Nh * K * Nd * Nc);
    /* flux_differencing_volume! :164-176 with accumulate_QF1! :201-211 */
    OMP_FOR
    for (int k = 0; k < K; k++) {
        for (int pidx = 0; pidx < p->nnz_h; pidx++) {
            int i = (int)p->Srsh_nnz[2 * pidx] - 1, j = (int)p->Srsh_nnz[2 * pidx + 1] - 1;
            double Ui[6] = {0}, Uj[6] = {0}, S[2] = {0}, F[ND_MAX * NC_MAX];
            Uvol(c, i, k, Ui);
            Uvol(c, j, k, Uj);
            Sx(c, i, j, k, S);
            fS(Nd, g, Ui, Uj, F);
            for (int d = 0; d < Nd; d++)
                for (int cc = 0; cc < Nc; cc++) {
                    double v = S[d] * F[d * Nc + cc];
                    c->QF1[IDX_XY(d, i, k, Nh, Nd, Nc) + cc] += v;
                    c->QF1[IDX_XY(d, j, k, Nh, Nd, Nc) + cc] += -v;
                }
        }
    }
    /* flux_differencing_surface! :223-247 with eval_high_order_surface_flux(::LaxFriedrichsOnProjectedVal) :262-272 */
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nfp; i++) {
            const double *uf = c->u_tilde + IDX_V(Nq + i, k, Nh, Nc);
            const double *uP = c->uPH + IDX_V(i, k, Nfp, Nc);
            double ff[ND_MAX * NC_MAX], fP[ND_MAX * NC_MAX], B[2] = {0.0, 0.0};
            fluxes(Nd, g, uf, ff);
            fluxes(Nd, g, uP, fP);
            double *fs = c->fstar_H + IDX_XY(0, i, k, Nfp, Nd, Nc);
            double *BF = c->BF_H + IDX_XY(0, i, k, Nfp, Nd, Nc);
            for (int x = 0; x < Nd * Nc; x++) fs[x] = 0.5 * (ff[x] + fP[x]);
            Bx(c, i, k, B);
            for (int d = 0; d < Nd; d++)
                for (int cc = 0; cc < Nc; cc++) BF[d * Nc + cc] = B[d] * fs[d * Nc + cc];
            int ld = (Nd == 1 || i < 2 * N1D) ? 0 : 1;
            double LF = c->LFc[i + (int64_t)Nfp * k];
            for (int cc = 0; cc < Nc; cc++) {
                double lf = LF * (uP[cc] - uf[cc]);
                BF[ld * Nc + cc] = BF[ld * Nc + cc] - lf;
                fs[ld * Nc + cc] = fs[ld * Nc + cc] - lf / B[ld];
            }
        }
    /* assemble_rhs! :331-361: generic mul! with MinvVhT, MinvVfT
     * (project_flux_difference_to_quad_unlimited! :274-281), or, on an element whose
     * entropy-projection parameters are limited (is_Vf_limited :326-328), with
     * MinvVhT_new = (1 ./ wq) .* [I Vf_new'] and Vf_new = l Vf + (1-l) Vf_low row by row
     * (project_flux_difference_to_quad!(::ScaledExtrapolation) :288-308, update_limited_extrapolation! :310-319) */
    OMP_FOR
    for (int k = 0; k < K; k++) {
        int limited = 0;
        if (p->entropyproj_limiter == 1) {
            double m = THETA(0, k, nstage - 1);
            for (int f = 1; f < Nfp; f++) m = jmin(m, THETA(f, k, nstage - 1));
            limited = m < 1.0;
        }
        for (int i = 0; i < Nq; i++) {
            double A[ND_MAX * NC_MAX], Bv[ND_MAX * NC_MAX];
            for (int x = 0; x < Nd * Nc; x++) { A[x] = 0.0; Bv[x] = 0.0; }
            double rw = 1.0 / p->wq[i];
            for (int j = 0; j < Nh; j++) {
                double a;
                if (!limited) {
                    a = p->MinvVhT[i + (int64_t)Nq * j];
                } else if (j < Nq) {
                    a = rw * (j == i ? 1.0 : 0.0);
                } else {
                    double l = THETA(j - Nq, k, nstage - 1);
                    a = rw * (l * p->Vf[(j - Nq) + (int64_t)Nfp * i] + (1.0 - l) * p->Vf_low[(j - Nq) + (int64_t)Nfp * i]);
                }
                const double *q = c->QF1 + IDX_XY(0, j, k, Nh, Nd, Nc);
                for (int x = 0; x < Nd * Nc; x++) A[x] += a * q[x];
            }
            for (int j = 0; j < Nfp; j++) {
                double a;
                if (!limited) {
                    a = p->MinvVfT[i + (int64_t)Nq * j];
                } else {
                    double l = THETA(j, k, nstage - 1);
                    a = rw * (l * p->Vf[j + (int64_t)Nfp * i] + (1.0 - l) * p->Vf_low[j + (int64_t)Nfp * i]);
                }
                const double *q = c->BF_H + IDX_XY(0, j, k, Nfp, Nd, Nc);
                for (int x = 0; x < Nd * Nc; x++) Bv[x] += a * q[x];
            }
            double J = p->Jq[i + (int64_t)Nq * k];
            double *r = c->rhsxyH + IDX_XY(0, i, k, Nq, Nd, Nc);
            for (int x = 0; x < Nd * Nc; x++) {
                c->MinvVhTQF1[IDX_XY(0, i, k, Nq, Nd, Nc) + x] = A[x];
                c->MinvVfTBF1[IDX_XY(0, i, k, Nq, Nd, Nc) + x] = Bv[x];
                r[x] = -(A[x] + Bv[x]) / J;
            }
            for (int cc = 0; cc < Nc; cc++) {
                double s = r[cc];
                for (int d = 1; d < Nd; d++) s = s + r[d * Nc + cc];
                c->rhsH[IDX_V(i, k, Nq, Nc) + cc] = s;
            }
        }
    }
}

/* ------------------------------------------------------------------------------- */
/* Subcell limiter: limiter.jl:20-56, subcell.jl, limiter_utils.jl                  */
/* ------------------------------------------------------------------------------- */
/* rhoe_quadratic_coefficients limiter_utils.jl:78-90 and rhoe_quadratic_solve :52-76 */
static double rhoe_quadratic_solve(const oracle_ctx *c, const double *U, const double *P, double Lrhoe) {
    double ZEROTOL = c->p.ZEROTOL;
    if (Lrhoe == INFINITY) return 1.0;
    double a, b, cc;
    if (c->p.Nd == 1) {
        a = P[0] * P[2] - 1.0 / 2.0 * (P[1] * P[1]);
        b = U[2] * P[0] + U[0] * P[2] - U[1] * P[1] - P[0] * Lrhoe;
        cc = U[2] * U[0] - 1.0 / 2.0 * (U[1] * U[1]) - U[0] * Lrhoe;
    } else {
        a = P[0] * P[3] - 1.0 / 2.0 * (P[1] * P[1] + P[2] * P[2]);
        b = U[3] * P[0] + U[0] * P[3] - U[1] * P[1] - U[2] * P[2] - P[0] * Lrhoe;
        cc = U[3] * U[0] - 1.0 / 2.0 * (U[1] * U[1] + U[2] * U[2]) - U[0] * Lrhoe;
    }
    double l_eps_ij = 1.0;
    if (b * b - 4.0 * a * cc >= 0.0) {
        double r1 = (-b + sqrt(b * b - 4.0 * a * cc)) / (2.0 * a);
        double r2 = (-b - sqrt(b * b - 4.0 * a * cc)) / (2.0 * a);
        if (r1 > ZEROTOL && r2 > ZEROTOL) {
            l_eps_ij = jmin(r1, r2);
        } else if (r1 > ZEROTOL && r2 < -ZEROTOL) {
            l_eps_ij = r1;
        } else if (r2 > ZEROTOL && r1 < -ZEROTOL) {
            l_eps_ij = r2;
        }
    }
    return l_eps_ij;
}
/* limiting_param_bound_rho_rhoe limiter_utils.jl:26-40 */
static double limiting_param_bound_rho_rhoe(const oracle_ctx *c, const double *U, const double *P,
                                            double Lrho, double Lrhoe, double Urho, double Urhoe) {
    double l = 1.0;
    if (U[0] + P[0] < Lrho) l = jmax((Lrho - U[0]) / P[0], 0.0);
    if (U[0] + P[0] > Urho) l = jmin(l, jmax((Urho - U[0]) / P[0], 0.0));
    l = jmin(jmin(l, rhoe_quadratic_solve(c, U, P, Lrhoe)), rhoe_quadratic_solve(c, U, P, Urhoe));
    return l;
}
/* s_modified_ufun compressible_Navier_Stokes.jl:80-85 */
static inline double s_modified_ufun(int Nd, double g, const double *U) {
    double rho = U[0];
    double rhoe = rhoe_ufun(Nd, U);
    return rhoe * pow(rho, -g);
}
/* limiting_param_bound_phi limiter_utils.jl:42-50 with bisection nonlinear_solvers.jl:3-20
 * (x_valid = 0, x_invalid = lpos) */
static int phi_admissible(const oracle_ctx *c, const double *U, const double *P, double Lphi, double l) {
    double Ul[NC_MAX];
    for (int cc = 0; cc < c->p.Nc; cc++) Ul[cc] = U[cc] + l * P[cc];
    return s_modified_ufun(c->p.Nd, c->p.gamma, Ul) >= Lphi - c->p.POSTOL;
}
static double limiting_param_bound_phi(const oracle_ctx *c, const double *U, const double *P, double Lphi, double lpos) {
    if (phi_admissible(c, U, P, Lphi, lpos)) return lpos;
    double x_valid = 0.0, x_invalid = lpos;
    for (int iter = 0; iter <= 20; iter++) {
        double x_new = 0.5 * (x_valid + x_invalid);
        if (phi_admissible(c, U, P, Lphi, x_new)) x_valid = x_new;
        else x_invalid = x_new;
    }
    return x_valid;
}
/* limiting_param(::SubcellLimiter, bound) limiter_utils.jl:10-24 with the bound tuple of
 * subcell.jl:231,299 = (Lrho, Lrhoe, Lphi, Urho, Urhoe) and rho_bound(::Positivity...)
 * subcell.jl:370-387 (Lrho = zeta rho_L, Urho = Inf) */
/* limiting_param(::SubcellLimiter, bound) limiter_utils.jl:10-24 with the bound tuple of
 * subcell.jl:299 (Lrho, zeta*rhoe(u_L), Lphi, Urho, Inf), where (Lrho, Urho) is rho_bound
 * subcell.jl:350-387: (zeta*rho_L, Inf) for the positivity bounds, the TVD bounds of the node for
 * the TVD ones (lrho_tvd, urho_tvd) */
static double limiting_param(const oracle_ctx *c, const double *uL, const double *P, double Lphi, double lrho_tvd,
                             double urho_tvd) {
    const int tvd = c->p.bound >= 5, ent = c->p.bound % 5;
    double Lrho = tvd ? lrho_tvd : c->p.zeta * uL[0];
    double Urho_i = tvd ? urho_tvd : INFINITY, Urhoe = INFINITY;
    double Lrhoe = c->p.zeta * rhoe_ufun(c->p.Nd, uL);
    if (ent == 1 || ent == 2) {
        /* min-entropy bounds destructure all five values (limiter_utils.jl:20-22); the
         * cell-entropy bounds take the positivity form (:10-14) */
        double lpos = limiting_param_bound_rho_rhoe(c, uL, P, Lrho, Lrhoe, Urho_i, Urhoe);
        return limiting_param_bound_phi(c, uL, P, Lphi, lpos);
    }
    if (c->p.upper_bound_mode == 0) /* `Lrho, Lrhoe, Urho, Urhoe = bound_val` takes the first four */
        return limiting_param_bound_rho_rhoe(c, uL, P, Lrho, Lrhoe, Lphi, Urho_i);
    return limiting_param_bound_rho_rhoe(c, uL, P, Lrho, Lrhoe, Urho_i, Urhoe);
}

/* ------------------------------------------------------------------------------- */
/* Pointwise hooks for the tests                                                    */
/* ------------------------------------------------------------------------------- */
void oracle_pt_v_of_u(int Nd, double gamma, const double *U, double *V) { v_ufun(Nd, gamma, U, V); }
double oracle_pt_find_alpha(int Nd, double POSTOL, const double *ui, const double *ut) {
    oracle_ctx c;
    memset(&c, 0, sizeof(c));
    c.p.Nd = Nd;
    c.p.Nc = Nd + 2;
    c.p.POSTOL = POSTOL;
    return find_alpha(&c, ui, ut);
}
void oracle_pt_u_of_v(int Nd, double gamma, const double *V, double *U) { u_vfun(Nd, gamma, V, U); }
void oracle_pt_flux(int Nd, double gamma, const double *U, double *F) { fluxes(Nd, gamma, U, F); }
double oracle_pt_logmean(double aL, double aR) { return logmean(aL, aR, log(aL), log(aR)); }
double oracle_pt_wavespeed(int Nd, double gamma, const double *U, const double *n) { return wavespeed(Nd, gamma, U, n); }
void oracle_pt_fS(int Nd, double gamma, const double *UL, const double *UR, double *F) {
    double L[6], R[6];
    const double *Us[2] = {UL, UR};
    double *Ps[2] = {L, R};
    for (int s = 0; s < 2; s++) {
        const double *u = Us[s];
        double *P = Ps[s];
        double beta = betafun(Nd, gamma, u);
        P[0] = u[0];
        P[1] = u[1] / u[0];
        if (Nd == 2) P[2] = u[2] / u[0];
        P[Nd + 1] = beta;
        P[Nd + 2] = log(u[0]);
        P[Nd + 3] = log(beta);
    }
    fS(Nd, gamma, L, R, F);
}
double oracle_pt_limiting_param(int Nd, double zeta, double ZEROTOL, int upper_bound_mode, const double *uL,
                                const double *P) {
    oracle_ctx c;
    memset(&c, 0, sizeof(c));
    c.p.Nd = Nd;
    c.p.zeta = zeta;
    c.p.ZEROTOL = ZEROTOL;
    c.p.upper_bound_mode = upper_bound_mode;
    return limiting_param(&c, uL, P, 0.0, 0.0, INFINITY);
}

#define LLOC(i, d, k, s) (c->L_local[(int64_t)(i) + (int64_t)c->NLs * ((d) + (int64_t)Nd * ((int64_t)(k) + (int64_t)K * (s)))])
#define FB(arr, d, i, k) ((arr) + (int64_t)(d) * c->NLs * K * Nc + (int64_t)Nc * ((int64_t)(i) + (int64_t)c->NLs * (k)))

/* initialize_indicator!(::CompressibleIdealGas) shock_capture.jl:82-94 and
 * initialize_smoothness_indicator!(::Dim1 / ::Dim2) :14-80: rho*p at the nodes, its modal
 * coefficients (generic mul! with VDM_inv), the energy share of the top two mode shells */
static void smoothness_indicator(oracle_ctx *c, const double *Uq, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nc = p->Nc, Nd = p->Nd, N = p->N1D - 1;
    double g = p->gamma;
    OMP_FOR
    for (int k = 0; k < K; k++) {
        double *ind = c->indicator + (int64_t)Nq * k, *mod = c->indicator_modal + (int64_t)Nq * k;
        for (int i = 0; i < Nq; i++) {
            const double *Ui = Uq + IDX_V(i, k, Nq, Nc);
            ind[i] = Ui[0] * pfun(Nd, g, Ui);
        }
        for (int i = 0; i < Nq; i++) mod[i] = 0.0;
        for (int j = 0; j < Nq; j++)
            for (int i = 0; i < Nq; i++) mod[i] += p->VDM_inv[i + (int64_t)Nq * j] * ind[j];
        int count = 0;
        double modeN = 0.0, modeNm1 = 0.0, total = 0.0;
        if (Nd == 1) {
            for (int i = 0; i <= N; i++) {
                double energy = mod[count] * mod[count];
                if (i == N) modeN += energy;
                if (i == N - 1) modeNm1 += energy;
                total += energy;
                count++;
            }
        } else {
            for (int j = 0; j <= N; j++)
                for (int i = 0; i <= N; i++) {
                    double energy = mod[count] * mod[count];
                    if (i == N || j == N) modeN += energy;
                    if (i == N - 1 || j == N - 1) modeNm1 += energy;
                    total += energy;
                    count++;
                }
        }
        c->smooth_indicator[k] = jmax(modeN / total, modeNm1 / total);
    }
}

/* update_blending_factor! shock_capture.jl:111-132 (Hennemann et al., eq. (46)) */
static void update_blending_factor(oracle_ctx *c, int s) {
    const oracle_problem *p = &c->p;
    int K = p->K, N = p->N1D - 1;
    if (p->shockcapture == 0) {
        for (int k = 0; k < K; k++) c->blending_factor[k + (int64_t)K * s] = 1.0;
        return;
    }
    double TN = p->shock_a * pow(10.0, -p->shock_c * pow((double)(N + 1), 0.25));
    double alphamax = 0.5, alphaE0 = 0.0001;
    double s_factor = log((1.0 - alphaE0) / alphaE0);
    for (int k = 0; k < K; k++) {
        double alpha = 1.0 / (1.0 + exp(-s_factor / TN * (c->smooth_indicator[k] - TN)));
        c->blending_factor[k + (int64_t)K * s] = jmax(jmin(1.0 - alpha, 1.0), alphamax);
    }
}

/* update_smoothness_factor! subcell.jl:927-956 */
static void update_smoothness_factor(oracle_ctx *c, int s) {
    const oracle_problem *p = &c->p;
    int K = p->K, N = p->N1D - 1;
    for (int k = 0; k < K; k++) {
        double f;
        int ent = p->bound % 5;
        if (ent == 0 || ent == 3) {
            f = 0.0;
        } else if (ent == 1) {
            f = 1.0;
        } else {
            double kappa = 1.0;
            double s0 = log(pow((double)N, -4.0)) / log(10.0);   /* log(10, N^-4) */
            double sk = log(c->smooth_indicator[k]) / log(10.0);
            if (sk < s0 - kappa) f = 0.0;
            else if (sk > s0 + kappa) f = 1.0;
            else f = 0.5 - 0.5 * sin(3.141592653589793 * (sk - s0) / (2.0 * kappa));   /* Float64(pi) */
        }
        c->smooth_factor[k + (int64_t)K * s] = f;
    }
}

/* the stencil node across face q2fq[i][m] (quad_index_to_quad_index_P limiter_utils.jl:184-212):
 * element kP and node iP (0-based) */
static void node_across(const oracle_ctx *c, int i, int k, int m, int *iP, int *kP) {
    const oracle_problem *p = &c->p;
    int Nq = p->Nq, Nfp = p->Nfp;
    int iface = (int)p->q2fq[i + (int64_t)Nq * m] - 1;
    int64_t mp = p->mapP[iface + (int64_t)Nfp * k] - 1;
    *iP = (int)p->fq2q[mp % Nfp] - 1;
    *kP = (int)(mp / Nfp);
}

/* low_order_stencil limiter_utils.jl:214-231: the node's neighbours (sl, sr[, sb, st]) along its
 * lines, across the element boundary by face q2fq[q][direction] (a node on one face only takes its
 * one face); returns their count */
static int low_order_stencil(const oracle_ctx *c, int q, int k, int *nodes, int *elems) {
    const oracle_problem *p = &c->p;
    int Nq = p->Nq, Nd = p->Nd, N1D = p->N1D;
    int si = Nd == 1 ? q : q % N1D, sj = Nd == 1 ? 0 : q / N1D;
    int n = 0;
    int nfq = p->q2fq[q + (int64_t)Nq * 1] > 0 ? 2 : 1;
    for (int d = 0; d < Nd; d++) {
        int pos = d == 0 ? si : sj;
        for (int side = -1; side <= 1; side += 2) {
            int np = pos + side;
            if (np >= 0 && np < N1D) {
                nodes[n] = d == 0 ? (Nd == 1 ? np : np + N1D * sj) : si + N1D * np;
                elems[n] = k;
            } else {
                node_across(c, q, k, nfq == 1 ? 0 : d, &nodes[n], &elems[n]);
            }
            n++;
        }
    }
    return n;
}

/* initialize_TVD_bounds! subcell.jl:87-141: rho_L = rho + dt*rhsL_rho at every node, then its
 * minimum and maximum over the node and its low-order stencil */
static void initialize_tvd_bounds(oracle_ctx *c, const double *Uq, double dt, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nc = p->Nc;
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nq; i++)
            c->rhoL[i + (int64_t)Nq * k] = Uq[IDX_V(i, k, Nq, Nc)] + dt * c->rhsL[IDX_V(i, k, Nq, Nc)];
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int q = 0; q < Nq; q++) {
            int nodes[4], elems[4];
            int n = low_order_stencil(c, q, k, nodes, elems);
            double lb = c->rhoL[q + (int64_t)Nq * k], ub = lb;
            for (int x = 0; x < n; x++) {
                double r = c->rhoL[nodes[x] + (int64_t)Nq * elems[x]];
                lb = jmin(lb, r);
                ub = jmax(ub, r);
            }
            c->lbound_rho[q + (int64_t)Nq * k] = lb;
            c->ubound_rho[q + (int64_t)Nq * k] = ub;
        }
}

/* initialize_entropy_bounds! subcell.jl:4-75: s_modified at the nodes, its global minimum at the
 * first stage of t0, and the lower bound epsk*min(stencil) + (1-epsk)*min(s_modified(U0)) with
 * the low-order stencil of low_order_stencil limiter_utils.jl:214-231 */
static void initialize_entropy_bounds(oracle_ctx *c, const double *Uq, double t, int nstage, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nc = p->Nc, Nd = p->Nd;
    int s = nstage - 1;
    if (p->bound % 5 == 0 || p->bound % 5 >= 3) {      /* subcell.jl:4-6 */
        memset(c->lbound_s_modified, 0, sizeof(double) * (size_t)Nq * K);
        return;
    }
    double g = p->gamma;
    OMP_FOR
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nq; i++) c->s_modified[i + (int64_t)Nq * k] = s_modified_ufun(Nd, g, Uq + IDX_V(i, k, Nq, Nc));
    if (t == p->t0 && nstage == 1) {
        double m = c->s_modified[0];
        for (int64_t x = 1; x < (int64_t)Nq * K; x++) m = jmin(m, c->s_modified[x]);
        c->s_modified_min = m;
    }
    OMP_FOR
    for (int k = 0; k < K; k++) {
        double epsk = c->smooth_factor[k + (int64_t)K * s];
        for (int q = 0; q < Nq; q++) {
            int nodes[4], elems[4];
            int n = low_order_stencil(c, q, k, nodes, elems);
            double lb = c->s_modified[q + (int64_t)Nq * k];
            for (int x = 0; x < n; x++) lb = jmin(lb, c->s_modified[nodes[x] + (int64_t)Nq * elems[x]]);
            c->lbound_s_modified[q + (int64_t)Nq * k] = epsk * lb + (1.0 - epsk) * c->s_modified_min;
        }
    }
}

/* apply_zhang_shu_limiter! zhangshu.jl:4-24 with zhang_shu_bound_limiter! :35-45 and
 * limiting_param(::ZhangShuLimiter) limiter_utils.jl:4-8 */
static void apply_zhang_shu_limiter(oracle_ctx *c, const double *Uq, double dt, int nstage, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nc = p->Nc, Nd = p->Nd;
    int s = nstage - 1;
    double zeta = p->zeta;
    OMP_FOR
    for (int k = 0; k < K; k++) {
        double l = 1.0;
        for (int i = 0; i < Nq; i++) {
            double uL[NC_MAX], P[NC_MAX];
            for (int cc = 0; cc < Nc; cc++) {
                uL[cc] = Uq[IDX_V(i, k, Nq, Nc) + cc] + dt * c->rhsL[IDX_V(i, k, Nq, Nc) + cc];
                P[cc] = dt * (c->rhsH[IDX_V(i, k, Nq, Nc) + cc] - c->rhsL[IDX_V(i, k, Nq, Nc) + cc]);
            }
            double li = jmin(1.0, limiting_param_bound_rho_rhoe(c, uL, P, zeta * uL[0], zeta * rhoe_ufun(Nd, uL),
                                                                 INFINITY, INFINITY));
            l = jmin(l, li);
        }
        c->L_zs[k + (int64_t)K * s] = l;
        l = jmin(c->L_zs[k + (int64_t)K * s], c->blending_factor[k + (int64_t)K * s]);
        for (int i = 0; i < Nq; i++)
            for (int cc = 0; cc < Nc; cc++)
                c->rhsU[IDX_V(i, k, Nq, Nc) + cc] =
                    (1.0 - l) * c->rhsL[IDX_V(i, k, Nq, Nc) + cc] + l * c->rhsH[IDX_V(i, k, Nq, Nc) + cc];
    }
}

/* Julia's isless on Float64 (NaN after every number, -0.0 before +0.0) and isequal */
static int jisless(double a, double b) {
    if (isnan(a)) return 0;
    if (isnan(b)) return 1;
    if (a < b) return 1;
    return a == b && signbit(a) && !signbit(b);
}
static int jisequal(double a, double b) {
    if (isnan(a) || isnan(b)) return isnan(a) && isnan(b);
    return a == b && signbit(a) == signbit(b);
}

/* the positions 0..n-1 in the order of sort!(dvdf_order_k, rev=true) over (dvdf[i], i): descending
 * by Julia's tuple isless, i.e. by value, then by index */
static void es_order(const double *v, int n, int *order) {
    int used[64] = {0};
    for (int t = 0; t < n; t++) {
        int best = -1;
        for (int j = 0; j < n; j++) {
            if (used[j]) continue;
            if (best < 0 || jisless(v[best], v[j]) || (jisequal(v[best], v[j]) && best < j)) best = j;
        }
        used[best] = 1;
        order[t] = best;
    }
}

/* enforce_ES_subcell_volume!(::Dim2) subcell.jl:628-734 with initialize_ES_subcell_limiting!(::Dim2)
 * :511-562 and rhs_es :736-743: per element and direction, dvdf = (v_{i-1} - v_i).(fbar_H - fbar_L)
 * at the interior subcell faces; if sum(l dvdf) - rhs exceeds tol = max(0, sum dv.fbar_L - 1'B psi),
 * zero the parameters of the faces with the largest dvdf, greedily, and scale the last one so that
 * the sum reaches rhs + tol */
static void enforce_es_subcell_volume(oracle_ctx *c, const double *Uq, int nstage, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nfp = p->Nfp, Nc = p->Nc, N1D = p->N1D, Nd = 2;
    int s = nstage - 1;
    int N1Dp1 = N1D + 1, N1Dm1 = N1D - 1, n = Nq - N1D;
    double g = p->gamma;
    OMP_FOR
    for (int k = 0; k < K; k++) {
        double V[25][NC_MAX];
        for (int i = 0; i < Nq; i++) v_ufun(2, g, Uq + IDX_V(i, k, Nq, Nc), V[i]);
        double sBpsi[2] = {0.0, 0.0};
        for (int i = 0; i < Nfp; i++) {
            const double *uf = Uq + IDX_V((int)p->fq2q[i] - 1, k, Nq, Nc);
            double psi[2] = {(g - 1.0) * uf[1], (g - 1.0) * uf[2]};   /* psi_ufun :129-132 */
            double B[2] = {0.0, 0.0};
            Bx(c, i, k, B);
            sBpsi[0] = sBpsi[0] + B[0] * psi[0];
            sBpsi[1] = sBpsi[1] + B[1] * psi[1];
        }
        double dvdf[2][20], sdvfL[2] = {0.0, 0.0};
        for (int sj = 0; sj < N1D; sj++)
            for (int si = 1; si < N1D; si++) {
                const double *fL = FB(c->f_bar_L, 0, si + N1Dp1 * sj, k), *fH = FB(c->f_bar_H, 0, si + N1Dp1 * sj, k);
                const double *va = V[(si - 1) + N1D * sj], *vb = V[si + N1D * sj];
                double a = 0.0, b = 0.0;
                for (int cc = 0; cc < Nc; cc++) {
                    double dv = va[cc] - vb[cc];
                    double ta = dv * (fH[cc] - fL[cc]), tb = dv * fL[cc];
                    a = cc == 0 ? ta : a + ta;
                    b = cc == 0 ? tb : b + tb;
                }
                dvdf[0][(si - 1) + N1Dm1 * sj] = a;
                sdvfL[0] += b;
            }
        for (int si = 0; si < N1D; si++)
            for (int sj = 1; sj < N1D; sj++) {
                const double *fL = FB(c->f_bar_L, 1, si + N1D * sj, k), *fH = FB(c->f_bar_H, 1, si + N1D * sj, k);
                const double *va = V[si + N1D * (sj - 1)], *vb = V[si + N1D * sj];
                double a = 0.0, b = 0.0;
                for (int cc = 0; cc < Nc; cc++) {
                    double dv = va[cc] - vb[cc];
                    double ta = dv * (fH[cc] - fL[cc]), tb = dv * fL[cc];
                    a = cc == 0 ? ta : a + ta;
                    b = cc == 0 ? tb : b + tb;
                }
                dvdf[1][si + N1D * (sj - 1)] = a;
                sdvfL[1] += b;
            }
        double epsk = c->smooth_factor[k + (int64_t)K * s];
        for (int d = 0; d < 2; d++) {
            /* the subcell face of dvdf[d][idx] in L_local (reshape (N1Dp1, N1D) / (N1D, N1Dp1)) */
#define ES_LI(idx) (d == 0 ? ((idx) % N1Dm1 + 1) + N1Dp1 * ((idx) / N1Dm1) : ((idx) % N1D) + N1D * ((idx) / N1D + 1))
            double poslim = 0.0;
            if (d == 0) {
                for (int sj = 0; sj < N1D; sj++)
                    for (int si = 1; si < N1D; si++) {
                        int idx = (si - 1) + N1Dm1 * sj;
                        poslim += LLOC(ES_LI(idx), 0, k, s) * dvdf[0][idx];
                    }
            } else {
                for (int si = 0; si < N1D; si++)
                    for (int sj = 1; sj < N1D; sj++) {
                        int idx = si + N1D * (sj - 1);
                        poslim += LLOC(ES_LI(idx), 1, k, s) * dvdf[1][idx];
                    }
            }
            double diff = sBpsi[d] - sdvfL[d];
            double rhs = p->bound % 5 == 3 ? diff : (1.0 - p->cell_beta * epsk) * diff;
            double est = poslim - rhs;
            double tol = jmax(0.0, sdvfL[d] - sBpsi[d]);
            if (!(est > tol)) continue;
            int order[20];
            es_order(dvdf[d], n, order);
            int curr = 0;
            double lhs = poslim;
            while (lhs > rhs + tol && curr < n) {
                int idx = order[curr];
                if (dvdf[d][idx] < p->ZEROTOL) break;
                lhs = lhs - LLOC(ES_LI(idx), d, k, s) * dvdf[d][idx];
                curr++;
            }
            for (int i = 0; i < curr; i++) {
                int idx = order[i];
                double l_new = i == curr - 1 ? jmax((rhs + tol - lhs) / dvdf[d][idx], 0.0) : 0.0;
                LLOC(ES_LI(idx), d, k, s) = jmin(LLOC(ES_LI(idx), d, k, s), l_new);
            }
#undef ES_LI
        }
    }
}

static void apply_rhs_limiter(oracle_ctx *c, const double *Uq, double t, double dt, int nstage, int nthr) {
    const oracle_problem *p = &c->p;
    int K = p->K, Nq = p->Nq, Nfp = p->Nfp, Nc = p->Nc, Nd = p->Nd, N1D = p->N1D;
    int s = nstage - 1;
    int NLs = c->NLs;
    /* initialize_smoothness_indicator! shock_capture.jl:4-12: nothing for (NoShockCapture,
     * PositivityBound), which is also the bound the Zhang-Shu limiter reports (traits.jl:61-63) */
    if ((p->shockcapture != 0 || (p->rhs_limiter == 0 && p->bound != 0)) && p->VDM_inv) smoothness_indicator(c, Uq, nthr);
    update_blending_factor(c, s);
    if (p->rhs_limiter == 1) {
        apply_zhang_shu_limiter(c, Uq, dt, nstage, nthr);
        return;
    }
    update_smoothness_factor(c, s);
    initialize_entropy_bounds(c, Uq, t, nstage, nthr);
    /* initialize_TVD_bounds!: nothing for the positivity-type bounds (subcell.jl:82-84) */
    if (p->bound >= 5) initialize_tvd_bounds(c, Uq, dt, nthr);

    /* accumulate_f_bar! subcell.jl:144-206 */
    OMP_FOR
    for (int k = 0; k < K; k++) {
        if (Nd == 1) {
            for (int cc = 0; cc < Nc; cc++) {
                FB(c->f_bar_H, 0, 0, k)[cc] = c->BF_H[IDX_XY(0, 0, k, Nfp, Nd, Nc) + cc];
                FB(c->f_bar_L, 0, 0, k)[cc] = c->BF_L[IDX_XY(0, 0, k, Nfp, Nd, Nc) + cc];
            }
            for (int i = 1; i < Nq + 1; i++) {
                double wJ = p->Jq[(i - 1) + (int64_t)Nq * k] * p->wq[i - 1];
                for (int cc = 0; cc < Nc; cc++) {
                    FB(c->f_bar_H, 0, i, k)[cc] = FB(c->f_bar_H, 0, i - 1, k)[cc] + wJ * c->rhsH[IDX_V(i - 1, k, Nq, Nc) + cc];
                    FB(c->f_bar_L, 0, i, k)[cc] = FB(c->f_bar_L, 0, i - 1, k)[cc] + wJ * c->rhsL[IDX_V(i - 1, k, Nq, Nc) + cc];
                }
            }
            continue;
        }
        int N1Dp1 = N1D + 1;
        /* x strides: fx_bar[(si, sj)] at linear si + N1Dp1*sj */
        for (int sj = 0; sj < N1D; sj++) {
            int iface = sj;
            for (int cc = 0; cc < Nc; cc++) {
                FB(c->f_bar_H, 0, 0 + N1Dp1 * sj, k)[cc] = c->BF_H[IDX_XY(0, iface, k, Nfp, Nd, Nc) + cc];
                FB(c->f_bar_L, 0, 0 + N1Dp1 * sj, k)[cc] = c->BF_L[IDX_XY(0, iface, k, Nfp, Nd, Nc) + cc];
            }
            for (int si = 1; si < N1Dp1; si++) {
                int q = (si - 1) + N1D * sj;
                double wJ = p->wq[q] * p->Jq[q + (int64_t)Nq * k];
                for (int cc = 0; cc < Nc; cc++) {
                    FB(c->f_bar_H, 0, si + N1Dp1 * sj, k)[cc] = FB(c->f_bar_H, 0, (si - 1) + N1Dp1 * sj, k)[cc] + wJ * c->rhsxyH[IDX_XY(0, q, k, Nq, Nd, Nc) + cc];
                    FB(c->f_bar_L, 0, si + N1Dp1 * sj, k)[cc] = FB(c->f_bar_L, 0, (si - 1) + N1Dp1 * sj, k)[cc] + wJ * c->rhsxyL[IDX_XY(0, q, k, Nq, Nd, Nc) + cc];
                }
            }
        }
        /* y strides: fy_bar[(si, sj)] at linear si + N1D*sj */
        for (int si = 0; si < N1D; si++) {
            int iface = si + 2 * N1D;
            for (int cc = 0; cc < Nc; cc++) {
                FB(c->f_bar_H, 1, si, k)[cc] = c->BF_H[IDX_XY(1, iface, k, Nfp, Nd, Nc) + cc];
                FB(c->f_bar_L, 1, si, k)[cc] = c->BF_L[IDX_XY(1, iface, k, Nfp, Nd, Nc) + cc];
            }
            for (int sj = 1; sj < N1Dp1; sj++) {
                int q = si + N1D * (sj - 1);
                double wJ = p->wq[q] * p->Jq[q + (int64_t)Nq * k];
                for (int cc = 0; cc < Nc; cc++) {
                    FB(c->f_bar_H, 1, si + N1D * sj, k)[cc] = FB(c->f_bar_H, 1, si + N1D * (sj - 1), k)[cc] + wJ * c->rhsxyH[IDX_XY(1, q, k, Nq, Nd, Nc) + cc];
                    FB(c->f_bar_L, 1, si + N1D * sj, k)[cc] = FB(c->f_bar_L, 1, si + N1D * (sj - 1), k)[cc] + wJ * c->rhsxyL[IDX_XY(1, q, k, Nq, Nd, Nc) + cc];
                }
            }
        }
    }

    /* subcell_bound_limiter! subcell.jl:208-246 (1D), :248-349 (2D) */
#define TVD_LB(q) (c->lbound_rho[(q) + (int64_t)Nq * k])
#define TVD_UB(q) (c->ubound_rho[(q) + (int64_t)Nq * k])
    for (int64_t x = 0; x < (int64_t)NLs * Nd * K; x++) c->L_local[x + (int64_t)NLs * Nd * K * s] = 1.0;
    OMP_FOR
    for (int k = 0; k < K; k++) {
        double *uLk = c->uL + IDX_V(0, k, Nq, Nc);
        for (int i = 0; i < Nq; i++)
            for (int cc = 0; cc < Nc; cc++) uLk[i * Nc + cc] = Uq[IDX_V(i, k, Nq, Nc) + cc] + dt * c->rhsL[IDX_V(i, k, Nq, Nc) + cc];
        double P[NC_MAX];
        if (Nd == 1) {
            for (int i = 0; i < Nq; i++) {
                double wJq_i = p->wq[i] * p->Jq[i + (int64_t)Nq * k];
                double Lphi = c->lbound_s_modified[i + (int64_t)Nq * k];
                for (int cc = 0; cc < Nc; cc++) P[cc] = -2.0 * dt * (FB(c->f_bar_H, 0, i, k)[cc] - FB(c->f_bar_L, 0, i, k)[cc]) / wJq_i;
                LLOC(i, 0, k, s) = jmin(LLOC(i, 0, k, s), limiting_param(c, uLk + i * Nc, P, Lphi, TVD_LB(i), TVD_UB(i)));
            }
            for (int i = 1; i < Nq + 1; i++) {
                double wJq_im1 = p->wq[i - 1] * p->Jq[(i - 1) + (int64_t)Nq * k];
                double Lphi = c->lbound_s_modified[(i - 1) + (int64_t)Nq * k];
                for (int cc = 0; cc < Nc; cc++) P[cc] = 2.0 * dt * (FB(c->f_bar_H, 0, i, k)[cc] - FB(c->f_bar_L, 0, i, k)[cc]) / wJq_im1;
                LLOC(i, 0, k, s) = jmin(LLOC(i, 0, k, s), limiting_param(c, uLk + (i - 1) * Nc, P, Lphi, TVD_LB((i - 1)), TVD_UB((i - 1))));
            }
            double l_shock = c->blending_factor[k + (int64_t)K * s];
            for (int i = 0; i < NLs; i++) LLOC(i, 0, k, s) = jmin(LLOC(i, 0, k, s), l_shock);
            continue;
        }
        int N1Dp1 = N1D + 1;
        for (int sj = 0; sj < N1D; sj++) {
            for (int si = 0; si < N1D; si++) {           /* left subcell faces */
                int q = si + N1D * sj;
                double wJq_i = p->wq[q] * p->Jq[q + (int64_t)Nq * k];
                double Lphi = c->lbound_s_modified[q + (int64_t)Nq * k];
                for (int cc = 0; cc < Nc; cc++) P[cc] = -4.0 * dt * (FB(c->f_bar_H, 0, si + N1Dp1 * sj, k)[cc] - FB(c->f_bar_L, 0, si + N1Dp1 * sj, k)[cc]) / wJq_i;
                int li = si + N1Dp1 * sj;
                LLOC(li, 0, k, s) = jmin(LLOC(li, 0, k, s), limiting_param(c, uLk + q * Nc, P, Lphi, TVD_LB(q), TVD_UB(q)));
            }
            for (int si = 1; si < N1Dp1; si++) {         /* right subcell faces */
                int q = (si - 1) + N1D * sj;
                double wJq_i = p->wq[q] * p->Jq[q + (int64_t)Nq * k];
                double Lphi = c->lbound_s_modified[q + (int64_t)Nq * k];
                for (int cc = 0; cc < Nc; cc++) P[cc] = 4.0 * dt * (FB(c->f_bar_H, 0, si + N1Dp1 * sj, k)[cc] - FB(c->f_bar_L, 0, si + N1Dp1 * sj, k)[cc]) / wJq_i;
                int li = si + N1Dp1 * sj;
                LLOC(li, 0, k, s) = jmin(LLOC(li, 0, k, s), limiting_param(c, uLk + q * Nc, P, Lphi, TVD_LB(q), TVD_UB(q)));
            }
        }
        for (int si = 0; si < N1D; si++) {
            for (int sj = 0; sj < N1D; sj++) {           /* bottom subcell faces */
                int q = si + N1D * sj;
                double wJq_i = p->wq[q] * p->Jq[q + (int64_t)Nq * k];
                double Lphi = c->lbound_s_modified[q + (int64_t)Nq * k];
                for (int cc = 0; cc < Nc; cc++) P[cc] = -4.0 * dt * (FB(c->f_bar_H, 1, si + N1D * sj, k)[cc] - FB(c->f_bar_L, 1, si + N1D * sj, k)[cc]) / wJq_i;
                int li = si + N1D * sj;
                LLOC(li, 1, k, s) = jmin(LLOC(li, 1, k, s), limiting_param(c, uLk + q * Nc, P, Lphi, TVD_LB(q), TVD_UB(q)));
            }
            for (int sj = 1; sj < N1Dp1; sj++) {         /* top subcell faces */
                int q = si + N1D * (sj - 1);
                double wJq_i = p->wq[q] * p->Jq[q + (int64_t)Nq * k];
                double Lphi = c->lbound_s_modified[q + (int64_t)Nq * k];
                for (int cc = 0; cc < Nc; cc++) P[cc] = 4.0 * dt * (FB(c->f_bar_H, 1, si + N1D * sj, k)[cc] - FB(c->f_bar_L, 1, si + N1D * sj, k)[cc]) / wJq_i;
                int li = si + N1D * sj;
                LLOC(li, 1, k, s) = jmin(LLOC(li, 1, k, s), limiting_param(c, uLk + q * Nc, P, Lphi, TVD_LB(q), TVD_UB(q)));
            }
        }
        double l_shock = c->blending_factor[k + (int64_t)K * s];
        for (int x = 0; x < N1Dp1 * N1D; x++) {
            LLOC(x, 0, k, s) = jmin(LLOC(x, 0, k, s), l_shock);
            LLOC(x, 1, k, s) = jmin(LLOC(x, 1, k, s), l_shock);
        }
    }

#undef TVD_LB
#undef TVD_UB
    /* enforce_ES_subcell! subcell.jl:458-466: nothing for the positivity and min-entropy bounds;
     * the cell-entropy bounds limit the interior subcell faces further (the interface step,
     * :745-789, is the Gauss basis's; for Lobatto the interface fluxes coincide) */
    if (p->bound % 5 >= 3) enforce_es_subcell_volume(c, Uq, nstage, nthr);

    /* symmetrize_limiting_parameters! */
    if (Nd == 1) {
        /* subcell.jl:405-416: `end` of the first dimension is Nq+N1D, not Nq+1, and the
         * element index wraps periodically whatever the boundary conditions */
        for (int k = 0; k < K; k++) {
            int km1 = (k - 1 + K) % K;
            double l = jmin(LLOC(0, 0, k, s), LLOC(NLs - 1, 0, km1, s));
            LLOC(0, 0, k, s) = l;
            LLOC(NLs - 1, 0, km1, s) = l;
        }
    } else {
        /* subcell.jl:418-456 via subcell_index_P_x/_y (limiter_utils.jl:123-181) */
        int N1Dp1 = N1D + 1;
        for (int k = 0; k < K; k++)
            for (int d = 0; d < 2; d++)
                for (int i = 0; i < NLs; i++) c->L_copy[i + (int64_t)NLs * (d + 2 * (int64_t)k)] = LLOC(i, d, k, s);
#define LCOPY(i, d, k) (c->L_copy[(int64_t)(i) + (int64_t)NLs * ((d) + 2 * (int64_t)(k))])
        OMP_FOR
        for (int k = 0; k < K; k++) {
            for (int sj = 0; sj < N1D; sj++) {
                for (int si = 0; si < N1Dp1; si += N1D) {
                    int iface = (si == 0) ? sj : sj + N1D;
                    int64_t m = p->mapP[iface + (int64_t)Nfp * k] - 1;
                    int64_t iP = m % Nfp, kP = m / Nfp;
                    int sjP = (int)(iP % N1D);
                    int siP = (iP / N1D == 0) ? 0 : N1D;
                    int idx = si + sj * N1Dp1, idxP = siP + sjP * N1Dp1;
                    LLOC(idx, 0, k, s) = jmin(LCOPY(idx, 0, k), LCOPY(idxP, 0, kP));
                }
            }
            for (int si = 0; si < N1D; si++) {
                for (int sj = 0; sj < N1Dp1; sj += N1D) {
                    int iface = (sj == 0) ? si + 2 * N1D : si + 3 * N1D;
                    int64_t m = p->mapP[iface + (int64_t)Nfp * k] - 1;
                    int64_t iP = m % Nfp, kP = m / Nfp;
                    int siP = (int)(iP % N1D);
                    int sjP = (iP / N1D == 2) ? 0 : N1D;
                    int idx = si + sj * N1D, idxP = siP + sjP * N1D;
                    LLOC(idx, 1, k, s) = jmin(LCOPY(idx, 1, k), LCOPY(idxP, 1, kP));
                }
            }
        }
#undef LCOPY
    }

    /* accumulate_f_bar_limited! subcell.jl:826-876 */
    OMP_FOR
    for (int k = 0; k < K; k++) {
        if (Nd == 1) {
            for (int i = 0; i < Nq + 1; i++) {
                double l = LLOC(i, 0, k, s);
                for (int cc = 0; cc < Nc; cc++)
                    FB(c->f_bar_lim, 0, i, k)[cc] = l * FB(c->f_bar_H, 0, i, k)[cc] + (1.0 - l) * FB(c->f_bar_L, 0, i, k)[cc];
            }
            continue;
        }
        int N1Dp1 = N1D + 1;
        for (int d = 0; d < 2; d++)
            for (int x = 0; x < N1Dp1 * N1D; x++) {
                double l = LLOC(x, d, k, s);
                for (int cc = 0; cc < Nc; cc++)
                    FB(c->f_bar_lim, d, x, k)[cc] = l * FB(c->f_bar_H, d, x, k)[cc] + (1.0 - l) * FB(c->f_bar_L, d, x, k)[cc];
            }
    }
    /* apply_subcell_limiter! subcell.jl:878-924 */
    OMP_FOR
    for (int k = 0; k < K; k++) {
        if (Nd == 1) {
            for (int i = 0; i < Nq; i++) {
                double wJq_i = p->wq[i] * p->Jq[i + (int64_t)Nq * k];
                for (int cc = 0; cc < Nc; cc++)
                    c->rhsU[IDX_V(i, k, Nq, Nc) + cc] = (FB(c->f_bar_lim, 0, i + 1, k)[cc] - FB(c->f_bar_lim, 0, i, k)[cc]) / wJq_i;
            }
            continue;
        }
        int N1Dp1 = N1D + 1;
        for (int j = 0; j < N1D; j++)
            for (int i = 0; i < N1D; i++) {
                int q = i + N1D * j;
                double wJq_ij = p->wq[q] * p->Jq[q + (int64_t)Nq * k];
                double *rxy = c->rhsxyU + IDX_XY(0, q, k, Nq, Nd, Nc);
                for (int cc = 0; cc < Nc; cc++) {
                    rxy[cc] = (FB(c->f_bar_lim, 0, (i + 1) + N1Dp1 * j, k)[cc] - FB(c->f_bar_lim, 0, i + N1Dp1 * j, k)[cc]) / wJq_ij;
                    rxy[Nc + cc] = (FB(c->f_bar_lim, 1, i + N1D * (j + 1), k)[cc] - FB(c->f_bar_lim, 1, i + N1D * j, k)[cc]) / wJq_ij;
                }
                for (int cc = 0; cc < Nc; cc++) c->rhsU[IDX_V(q, k, Nq, Nc) + cc] = rxy[cc] + rxy[Nc + cc];
            }
    }
}

/* ------------------------------------------------------------------------------- */
/* rhs! rhs.jl:5-13, 41-55                                                          */
/* ------------------------------------------------------------------------------- */
int oracle_rhs(oracle_ctx *c, const double *Uq, double t, double dt, int nstage, double *dt_out) {
    if (!c || !Uq || nstage < 1 || nstage > 3) return 1;
    int nthr = oracle_get_threads();
    /* init_rhs! rhs.jl:15-23 */
    if (c->p.entropyproj_limiter == 1) compute_entropyproj_limiting_param(c, Uq, nstage, nthr);
    entropy_projection(c, Uq, nstage, nthr);
    double dt_ret = rhs_low_graph_visc(c, Uq, t, dt, nstage, nthr);
    rhs_fluxdiff(c, Uq, nstage, nthr);
    apply_rhs_limiter(c, Uq, t, dt, nstage, nthr);
    if (dt_out) *dt_out = dt_ret;
    return 0;
}

/* SSP33! loop body SSPRK33.jl:29-43 */
int oracle_ssprk3_step(oracle_ctx *c, double *Uq, double *t, double *dt_out) {
    const oracle_problem *p = &c->p;
    int64_t n = (int64_t)p->Nq * p->K * p->Nc;
    double dt = jmin(p->CFL * p->dt0, p->T - *t);
    memcpy(c->resW, Uq, sizeof(double) * (size_t)n);
    int err = oracle_rhs(c, Uq, *t, dt, 1, &dt);
    if (err) return err;
    for (int64_t x = 0; x < n; x++) Uq[x] = c->resW[x] + dt * c->rhsU[x];
    if ((err = oracle_rhs(c, Uq, *t, dt, 2, NULL))) return err;
    for (int64_t x = 0; x < n; x++) c->resZ[x] = Uq[x] + dt * c->rhsU[x];
    for (int64_t x = 0; x < n; x++) Uq[x] = 3.0 / 4.0 * c->resW[x] + 1.0 / 4.0 * c->resZ[x];
    if ((err = oracle_rhs(c, Uq, *t, dt, 3, NULL))) return err;
    for (int64_t x = 0; x < n; x++) c->resZ[x] = Uq[x] + dt * c->rhsU[x];
    for (int64_t x = 0; x < n; x++) Uq[x] = 1.0 / 3.0 * c->resW[x] + 2.0 / 3.0 * c->resZ[x];
    *t = *t + dt;
    if (dt_out) *dt_out = dt;
    return 0;
}

int64_t oracle_count_nonpositive(const oracle_ctx *c, const double *Uq, int64_t *first) {
    const oracle_problem *p = &c->p;
    int64_t n = (int64_t)p->Nq * p->K, bad = 0, f = -1;
    for (int64_t x = 0; x < n; x++) {
        const double *U = Uq + x * p->Nc;
        double pr = pfun(p->Nd, p->gamma, U);
        if (!(U[0] > 0.0) || !(pr > 0.0)) {
            if (f < 0) f = x;
            bad++;
        }
    }
    if (first) *first = f;
    return bad;
}

double oracle_check_conservation(const oracle_ctx *c, const double *Uq) {
    const oracle_problem *p = &c->p;
    double tot[NC_MAX] = {0, 0, 0, 0};
    for (int k = 0; k < p->K; k++)
        for (int i = 0; i < p->Nq; i++) {
            double a = p->Jq[i + (int64_t)p->Nq * k] * p->wq[i];
            for (int cc = 0; cc < p->Nc; cc++) tot[cc] += a * Uq[IDX_V(i, k, p->Nq, p->Nc) + cc];
        }
    double s = tot[0];
    for (int cc = 1; cc < p->Nc; cc++) s += tot[cc];
    return s;
}
