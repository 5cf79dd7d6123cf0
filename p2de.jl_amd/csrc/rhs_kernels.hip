// rhs_kernels.hip -- FP64 kernels of the limited ESDG right-hand side and the SSP-RK3
// stage update for gfx950.
//
// One stage = two launches:
//   phase A  (element-local; a workgroup owns E elements, node data staged in LDS)
//     low-order graph-viscosity RHS, high-order flux-differencing RHS, subcell flux
//     potentials and the limiter parameters of every subcell face with the element's own
//     data; writes the element RHS plus, for each element-boundary subcell face, its
//     parameter and f_bar_H - f_bar_L; at stage 1 it also writes its workgroup's minimum of
//     the CFL time step, which dt_reduce_kernel folds into the global minimum.
//   phase B  (node-parallel) symmetrises the element-boundary parameters with the
//     neighbour (min), corrects the adjacent nodes' RHS, applies the RK stage, checks
//     positivity of the new state and forms its projected face states u(v(u)) once per
//     boundary node for the next phase A (face_state_kernel does the same for a state
//     set from the host).
// On the Gauss basis a third launch per stage comes first: entropy_projection_kernel forms the
// projected face states (and the NodewiseScaledExtrapolation parameters) from the whole line.
// The reference runs the same work as ~27 threaded loops per stage over HBM-sized
// intermediates (src/dg/rhs/*.jl, src/dg/limiter/subcell.jl); the formulas and their
// operation order are kept (built with -ffp-contract=off) so the FP64 result matches the
// reference's arithmetic.
#include <hip/hip_runtime.h>

#include "euler_device.h"
#include "p2de_internal.h"

namespace p2de {

// BASIS 0: Lobatto (one face-volume pair per face node); 1: Gauss (a face node pairs with
// every node of its line)
template <int DIM, int N1D, int BASIS = 0>
struct Sz {
    static constexpr int NQ = DIM == 1 ? N1D : N1D * N1D;
    static constexpr int NFP = DIM == 1 ? 2 : 4 * N1D;
    static constexpr int NH = NQ + NFP;
    static constexpr int NC = DIM + 2;
    static constexpr int NFV = BASIS == 0 ? NFP : NFP * N1D;   // face-volume pairs
    static constexpr int NPH = DIM == 1 ? N1D * (N1D - 1) / 2 + NFV : N1D * N1D * (N1D - 1) + NFV;
    static constexpr int NPL = DIM == 1 ? N1D - 1 : 2 * N1D * (N1D - 1);
    static constexpr int NLINE = DIM == 1 ? 1 : 2 * N1D;
    static constexpr int NLS = NQ + N1D;  // first dimension of L_local (State.jl:21)
};

__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long double_to_bits(double d) { return (unsigned long long)__double_as_longlong(d); }

// Bx / Bx_with_n (rhs_utils.jl:13-37): physical boundary normal at a face node, its norm and
// the unit normal.  The reference's 1D method passes its arguments in the wrong order
// (rhs_utils.jl:18); this is its evident intent rxJh * Br[i,i].
template <int DIM>
__device__ __forceinline__ void face_normal(const Geo& G, double Br, double Bs, double* B, double& nrm, double* n) {
    if (DIM == 1) {
        B[0] = G.rxJ * Br;
        B[1] = 0.0;
        nrm = fabs(B[0]);
        n[0] = B[0] / nrm;
        n[1] = 0.0;
    } else {
        B[0] = G.rxJ * Br + G.sxJ * Bs;
        B[1] = G.ryJ * Br + G.syJ * Bs;
        nrm = jnorm2(B[0], B[1]);
        n[0] = B[0] / nrm;
        n[1] = B[1] / nrm;
    }
}

template <int DIM>
__device__ __forceinline__ void make_prim(double g, const double* u, Prim<DIM>& P) {
    // U() of flux_differencing.jl:179-191 on top of calculate_primitive_variables! (:39-52)
    P.rho = u[0];
    P.u = u[1] / u[0];
    P.v = DIM == 2 ? u[2] / u[0] : 0.0;
    double pp = Eq<DIM>::p(g, u);
    P.beta = u[0] / (2.0 * pp);
    P.lrho = log(u[0]);
    P.lbeta = log(P.beta);
}

// rhoe_quadratic_solve / _coefficients (limiter_utils.jl:52-90)
template <int DIM>
__device__ __forceinline__ double rhoe_quadratic_solve(const double* U, const double* P, double Lrhoe, double ZEROTOL) {
    if (Lrhoe == INFINITY) return 1.0;
    double a, b, c;
    if (DIM == 1) {
        a = P[0] * P[2] - 1.0 / 2.0 * (P[1] * P[1]);
        b = U[2] * P[0] + U[0] * P[2] - U[1] * P[1] - P[0] * Lrhoe;
        c = U[2] * U[0] - 1.0 / 2.0 * (U[1] * U[1]) - U[0] * Lrhoe;
    } else {
        a = P[0] * P[3] - 1.0 / 2.0 * (P[1] * P[1] + P[2] * P[2]);
        b = U[3] * P[0] + U[0] * P[3] - U[1] * P[1] - U[2] * P[2] - P[0] * Lrhoe;
        c = U[3] * U[0] - 1.0 / 2.0 * (U[1] * U[1] + U[2] * U[2]) - U[0] * Lrhoe;
    }
    double l = 1.0;
    if (b * b - 4.0 * a * c >= 0.0) {
        double r1 = (-b + sqrt(b * b - 4.0 * a * c)) / (2.0 * a);
        double r2 = (-b - sqrt(b * b - 4.0 * a * c)) / (2.0 * a);
        if (r1 > ZEROTOL && r2 > ZEROTOL) {
            l = jmin(r1, r2);
        } else if (r1 > ZEROTOL && r2 < -ZEROTOL) {
            l = r1;
        } else if (r2 > ZEROTOL && r1 < -ZEROTOL) {
            l = r2;
        }
    }
    return l;
}

// limiting_param(::SubcellLimiter, ::PositivityBound) (limiter_utils.jl:10-14, 26-40) with the
// bound tuple of subcell.jl:299 = (zeta*rho_L, zeta*rhoe(u_L), Lphi = 0, Inf, Inf).
// upper_mode 0 reproduces the reference's 4-of-5 destructuring (Urho = Lphi = 0).
template <int DIM>
__device__ __forceinline__ double limiting_param(const double* uL, const double* P, double zeta, double ZEROTOL,
                                                 int upper_mode) {
    double Lrho = zeta * uL[0];
    double Lrhoe = zeta * Eq<DIM>::rhoe(uL);
    double Urho = upper_mode == 0 ? 0.0 : INFINITY;
    double Urhoe = INFINITY;
    double l = 1.0;
    if (uL[0] + P[0] < Lrho) l = jmax((Lrho - uL[0]) / P[0], 0.0);
    if (uL[0] + P[0] > Urho) l = jmin(l, jmax((Urho - uL[0]) / P[0], 0.0));
    l = jmin(jmin(l, rhoe_quadratic_solve<DIM>(uL, P, Lrhoe, ZEROTOL)), rhoe_quadratic_solve<DIM>(uL, P, Urhoe, ZEROTOL));
    return l;
}

// limiting_param_bound_rho_rhoe (limiter_utils.jl:26-40)
template <int DIM>
__device__ __forceinline__ double bound_rho_rhoe(const double* U, const double* P, double Lrho, double Lrhoe, double Urho,
                                                 double Urhoe, double ZEROTOL) {
    double l = 1.0;
    if (U[0] + P[0] < Lrho) l = jmax((Lrho - U[0]) / P[0], 0.0);
    if (U[0] + P[0] > Urho) l = jmin(l, jmax((Urho - U[0]) / P[0], 0.0));
    l = jmin(jmin(l, rhoe_quadratic_solve<DIM>(U, P, Lrhoe, ZEROTOL)), rhoe_quadratic_solve<DIM>(U, P, Urhoe, ZEROTOL));
    return l;
}

// s_modified_ufun (compressible_Navier_Stokes.jl:80-85): rhoe * rho^-gamma
template <int DIM>
__device__ __forceinline__ double s_modified(double g, const double* U) {
    const double rho = U[0];
    const double rhoe = Eq<DIM>::rhoe(U);
    return rhoe * pow(rho, -g);
}

// limiting_param(::SubcellLimiter, <min-entropy bounds>) (limiter_utils.jl:19-24) with the bound
// tuple (Lrho, zeta*rhoe(u_L), Lphi, Urho, Inf) -- (Lrho, Urho) from rho_bound (subcell.jl:350-387):
// (zeta*rho_L, Inf), or the node's TVD bounds -- the positivity parameter lpos, then
// limiting_param_bound_phi (:42-50), the bisection (nonlinear_solvers.jl:3-20) between 0 and lpos
// for s_modified(u_L + l P) >= Lphi - POSTOL
template <int DIM>
__device__ __noinline__ double limiting_param_phi(const double* uL, const double* P, double Lrho, double Urho,
                                                  double zeta, double ZEROTOL, double Lphi, double POSTOL, double g) {
    constexpr int NC = DIM + 2;
    const double Lrhoe = zeta * Eq<DIM>::rhoe(uL);
    const double lpos = bound_rho_rhoe<DIM>(uL, P, Lrho, Lrhoe, Urho, INFINITY, ZEROTOL);
    auto ok = [&](double l) {
        double Ul[NC];
        for (int c = 0; c < NC; c++) Ul[c] = uL[c] + l * P[c];
        return s_modified<DIM>(g, Ul) >= Lphi - POSTOL;
    };
    if (ok(lpos)) return lpos;
    double x_valid = 0.0, x_invalid = lpos;
    for (int it = 0; it <= 20; it++) {
        const double x_new = 0.5 * (x_valid + x_invalid);
        if (ok(x_new)) x_valid = x_new;
        else x_invalid = x_new;
    }
    return x_valid;
}

// Julia's isless on Float64 (NaN after every number, -0.0 before +0.0) and isequal
__device__ __forceinline__ bool jisless(double a, double b) {
    if (isnan(a)) return false;
    if (isnan(b)) return true;
    if (a < b) return true;
    return a == b && signbit(a) && !signbit(b);
}
__device__ __forceinline__ bool jisequal(double a, double b) {
    if (isnan(a) || isnan(b)) return isnan(a) && isnan(b);
    return a == b && signbit(a) == signbit(b);
}

// enforce_ES_subcell_volume!(::Dim2) (subcell.jl:628-734) for one element and direction d, with
// initialize_ES_subcell_limiting!(::Dim2) (:511-562) and rhs_es (:736-743): dvdf = (v_a - v_b).(fbar_H
// - fbar_L) at the interior subcell faces of direction d (a before b along the line); when
// sum(l dvdf) - rhs exceeds tol = max(0, sum dv.fbar_L - 1'B psi), the faces are taken in the order of
// sort!((dvdf, idx), rev=true) and their parameters zeroed until the sum reaches rhs + tol, the last one
// scaled to meet it.  L (the element's subcell-face parameters, LDS) is updated in place; the function
// returns the number of faces it lowered and writes their positions to `changed`.
template <int N1D>
__device__ __noinline__ int es_subcell_volume(int d, double* L, const double* fH, const double* fL, const double* V,
                                              const double* U, const ElemTables& T, const Geo& G, double g,
                                              int bound, double beta, double epsk, double ZEROTOL, int* changed) {
    constexpr int NC = 4, NQ = N1D * N1D, NFP = 4 * N1D, NS = NQ - N1D;
    // 1'B psi over the face nodes, psi = (gamma - 1) (rho u, rho v) at the node fq2q
    double sB = 0.0;
    for (int f = 0; f < NFP; f++) {
        const double B = d == 0 ? G.rxJ * T.Br[f] + G.sxJ * T.Bs[f] : G.ryJ * T.Br[f] + G.syJ * T.Bs[f];
        const double psi = (g - 1.0) * U[T.fq2q[f] * NC + 1 + d];
        sB = sB + B * psi;
    }
    double dd[NS];
    int pos[NS];
    double sL = 0.0;
    for (int o = 0; o < N1D; o++)          // x: sj = o; y: si = o
        for (int t = 1; t < N1D; t++) {    // x: si = t; y: sj = t
            const int idx = d == 0 ? (t - 1) + (N1D - 1) * o : o + N1D * (t - 1);
            const int r = d == 0 ? o * (N1D + 1) + t : (N1D + o) * (N1D + 1) + t;
            const int qa = d == 0 ? (t - 1) + N1D * o : o + N1D * (t - 1);
            const int qb = d == 0 ? t + N1D * o : o + N1D * t;
            double a = 0.0, b = 0.0;
            for (int c = 0; c < NC; c++) {
                const double dv = V[qa * NC + c] - V[qb * NC + c];
                const double ta = dv * (fH[r * NC + c] - fL[r * NC + c]), tb = dv * fL[r * NC + c];
                a = c == 0 ? ta : a + ta;
                b = c == 0 ? tb : b + tb;
            }
            dd[idx] = a;
            pos[idx] = r;
            sL += b;
        }
    double poslim = 0.0;
    for (int o = 0; o < N1D; o++)
        for (int t = 1; t < N1D; t++) {
            const int idx = d == 0 ? (t - 1) + (N1D - 1) * o : o + N1D * (t - 1);
            poslim += L[pos[idx]] * dd[idx];
        }
    const double diff = sB - sL;
    const double rhs = bound % 5 == 3 ? diff : (1.0 - beta * epsk) * diff;
    const double est = poslim - rhs;
    const double tol = jmax(0.0, sL - sB);
    if (!(est > tol)) return 0;
    unsigned used = 0;
    int order[NS];
    int curr = 0;
    double lhs = poslim;
    while (lhs > rhs + tol && curr < NS) {
        int best = -1;   // the next entry of the descending (dvdf, idx) order
        for (int j = 0; j < NS; j++) {
            if (used & (1u << j)) continue;
            if (best < 0 || jisless(dd[best], dd[j]) || (jisequal(dd[best], dd[j]) && best < j)) best = j;
        }
        if (dd[best] < ZEROTOL) break;
        used |= 1u << best;
        order[curr] = best;
        lhs = lhs - L[pos[best]] * dd[best];
        curr++;
    }
    int n = 0;
    for (int i = 0; i < curr; i++) {
        const int idx = order[i];
        const double l_new = i == curr - 1 ? jmax((rhs + tol - lhs) / dd[idx], 0.0) : 0.0;
        const double l = jmin(L[pos[idx]], l_new);
        if (!(l == L[pos[idx]])) changed[n++] = pos[idx];
        L[pos[idx]] = l;
    }
    return n;
}

// find_alpha (low_order_graph_viscosity.jl:299-328): the smallest alpha, by doubling and then 50
// bisection steps, with alpha*u_i - u~ of positive density and internal energy.  The reference's
// doubling has no exit when no alpha works; here it ends when alpha overflows (alpha = Inf).
template <int DIM>
__device__ __noinline__ double find_alpha(const double* ui, const double* ut, double POSTOL) {
    constexpr int NC = DIM + 2;
    auto ok = [&](double al) {
        double sub[NC];
        for (int c = 0; c < NC; c++) sub[c] = al * ui[c] - ut[c];
        return sub[0] > POSTOL && Eq<DIM>::rhoe(sub) > POSTOL;
    };
    double alphaL = 0.0, alphaR = 1.0;
    while (!ok(alphaR)) {
        alphaR = 2.0 * alphaR;
        if (isinf(alphaR)) return alphaR;
    }
    for (int it = 0; it < 50; it++) {
        const double alphaM = (alphaL + alphaR) / 2.0;
        if (ok(alphaM)) alphaR = alphaM;
        else alphaL = alphaM;
    }
    return alphaR;
}

// LDS of one workgroup.  Arrays whose lifetimes do not overlap share storage:
// node/face primitives and fluxes (phases 1-3) with the directional RHS and u_L (phases 4-5),
// pair results (phases 3-4) with the subcell flux potentials and parameters (phase 5).
template <int DIM, int N1D, int E, int BASIS>
struct SmemA {
    using S = Sz<DIM, N1D, BASIS>;
    static constexpr int NFB = S::NLINE * (N1D + 1);  // subcell faces per element
    double U[E][S::NQ][S::NC];
    double BFH[E][S::NFP][S::NC];
    double BFL[E][S::NFP][S::NC];
    double lamB[E][S::NFP];
    double th[BASIS ? E : 1][S::NFP];                  // entropy-projection parameters (Gauss)
    union {
        struct {
            Prim<DIM> np[E][S::NQ];
            double nF[E][S::NQ][DIM * S::NC];
            Prim<DIM> fp[E][S::NFP];
        } g1;
        struct {
            double RH[E][S::NQ][DIM][S::NC];
            double RL[E][S::NQ][DIM][S::NC];
            double UL[E][S::NQ][S::NC];
        } g2;
        struct {
            double V[E][S::NQ][S::NC];                   // entropy variables (cell-entropy bounds)
        } g3;
    } a;
    union {
        struct {
            double PH[E][S::NPH][S::NC];
            double PL[E][S::NPL][S::NC + 1];
        } p;
        struct {
            double fH[E][NFB][S::NC];
            double fL[E][NFB][S::NC];
            double l[E][NFB];
        } fb;
    } x;
    double red[64];
};

// node of subcell line `line` at position s (0..N1D-1); lines are x lines (sj = line) then
// y lines (si = line - N1D) in 2D, a single line in 1D
template <int DIM, int N1D>
__device__ __forceinline__ int line_node(int line, int s) {
    if (DIM == 1) return s;
    return line < N1D ? s + N1D * line : (line - N1D) + N1D * s;
}

// face node at the element-boundary end of a subcell line (s = 0 or s = N1D)
template <int DIM, int N1D>
__device__ __forceinline__ int line_face(int line, int s) {
    if (DIM == 1) return s == 0 ? 0 : 1;
    if (line < N1D) return s == 0 ? line : N1D + line;
    return s == 0 ? 2 * N1D + (line - N1D) : 3 * N1D + (line - N1D);
}

// entropy_projection_face_node! (rhs.jl:84-94) for an LGL face node:
// v_tilde = sum_j (1*Vf + 0*Vf_low)[f,j] vq_j with the row's single nonzero, u_tilde = u(v_tilde)
template <int DIM>
__device__ __forceinline__ void project_face(double g, double coef, const double* V, double* ut) {
    constexpr int NC = DIM + 2;
    double Vt[NC];
    for (int c = 0; c < NC; c++) Vt[c] = 0.0 + coef * V[c];
    Eq<DIM>::u_of_v(g, Vt, ut);
}

template <int DIM, int N1D, int E, int NT, int BASIS, int OPT>
__global__ __launch_bounds__(NT) void phaseA_kernel(PhaseAArgs a) {
    using S = Sz<DIM, N1D, BASIS>;
    constexpr int NQ = S::NQ, NFP = S::NFP, NC = S::NC, NPH = S::NPH, NPL = S::NPL, NLINE = S::NLINE;
    constexpr int NFB = NLINE * (N1D + 1);
    __shared__ SmemA<DIM, N1D, E, BASIS> sm;
    if (a.sc->done) return;
    const int tid = threadIdx.x;
    const int k0 = blockIdx.x * E;
    const int K = a.K;
    const double g = a.gamma;
    const ElemTables& T = *a.T;

    // ---- phase 0: element states -> LDS (coalesced)
    {
        const int nel = min(E, K - k0);
        const double* src = a.U + (size_t)k0 * NQ * NC;
        double* dst = &sm.U[0][0][0];
        for (int w = tid; w < nel * NQ * NC; w += NT) dst[w] = src[w];
    }
    __syncthreads();

    // ---- phase 1: node primitives and physical fluxes
    for (int w = tid; w < E * NQ; w += NT) {
        const int e = w / NQ, i = w - e * NQ;
        if (k0 + e >= K) continue;
        const double* u = sm.U[e][i];
        make_prim<DIM>(g, u, sm.a.g1.np[e][i]);          // u_tilde = Uq at volume nodes (rhs.jl:79-82)
        Eq<DIM>::flux(g, u, sm.a.g1.nF[e][i]);           // update_wavespeed_and_inviscid_flux! :79-82
    }

    // ---- phase 2: face nodes (LF coefficients, both surface fluxes).  The projected face
    //      states u_tilde were formed once per node by the kernel that wrote the state.
    for (int w = tid; w < E * NFP; w += NT) {
        const int e = w / NFP, f = w - e * NFP;
        const int k = k0 + e;
        if (k >= K) continue;
        const int q = T.fq2q[f];
        const double* Uq = sm.U[e][q];
        const Geo G = a.geo[k];
        double B[2], nrm, n[2];
        face_normal<DIM>(G, T.Br[f], T.Bs[f], B, nrm, n);
        const int fi = k * NFP + f;
        double ut[NC];
        for (int c = 0; c < NC; c++) ut[c] = a.ut[(size_t)fi * NC + c];
        make_prim<DIM>(g, ut, sm.a.g1.fp[e][f]);
        if (BASIS) sm.th[BASIS ? e : 0][f] = a.theta ? a.theta[fi] : 1.0;

        // mapP target (calculate_primitive_variables! :54-69; get_uP_and_enforce_BC! :101-108)
        const int nbn = a.nb_node[fi], nbf = a.nb_face[fi];
        const int kP = nbf / NFP, iP = nbf - kP * NFP;
        double Ut[NC], utP[NC];
        for (int c = 0; c < NC; c++) {
            Ut[c] = a.U[(size_t)nbn * NC + c];
            utP[c] = a.ut[(size_t)nbf * NC + c];
        }
        const Geo GP = a.geo[kP];
        double BP[2], nrmP, nP[2];
        face_normal<DIM>(GP, T.Br[iP], T.Bs[iP], BP, nrmP, nP);

        // calculate_interface_dissipation_coeff! (flux_differencing.jl:90-114) + enforce_BC! (:116-151)
        const double lam = Eq<DIM>::wavespeed(g, ut, n);
        const double lamP = Eq<DIM>::wavespeed(g, utP, nP);
        double LFc = 0.5 * nrm;
        LFc = LFc * jmax(lam, lamP);
        const int bcc = a.bc[fi];
        double uPH[NC], uPL[NC];
        for (int c = 0; c < NC; c++) {
            if (bcc == 0) {
                uPH[c] = utP[c];
                uPL[c] = Ut[c];
            } else if (bcc < 0) {
                uPH[c] = Uq[c];
                uPL[c] = Uq[c];
            } else {
                uPH[c] = a.Ival[(bcc - 1) * NC + c];
                uPL[c] = a.Ival[(bcc - 1) * NC + c];
            }
        }
        if (bcc != 0) LFc = 0.0;

        // accumulate_numerical_flux! with LaxFriedrichsOnProjectedVal (:231-247, :262-272)
        const int ld = T.ld[f];
        double Ff[DIM * NC], FP[DIM * NC];
        Eq<DIM>::flux(g, ut, Ff);
        Eq<DIM>::flux(g, uPH, FP);
        for (int c = 0; c < NC; c++) {
            double fs = 0.5 * (Ff[ld * NC + c] + FP[ld * NC + c]);
            double BF = B[ld] * fs;
            double lf = LFc * (uPH[c] - ut[c]);
            sm.BFH[e][f][c] = BF - lf;
        }

        // low-order surface term (low_order_graph_viscosity.jl:51-71, 85-91, 94-124, 184-198): the face
        // states are the nodal values (LaxFriedrichsOnNodalVal) or the projected ones
        // (LaxFriedrichsOnProjectedVal); outflow takes the node's state, inflow Ival
        const bool lproj = a.lo_proj != 0;
        const double* Uf = lproj ? ut : Uq;
        if (lproj && bcc == 0)
            for (int c = 0; c < NC; c++) uPL[c] = utP[c];
        const double wsf = Eq<DIM>::wavespeed(g, Uf, n);
        const double wsP = Eq<DIM>::wavespeed(g, lproj ? utP : Ut, nP);
        const double lB = 0.5 * nrm * jmax(wsf, wsP);
        Eq<DIM>::flux(g, uPL, FP);
        for (int c = 0; c < NC; c++) {
            double fs = 0.5 * ((lproj ? Ff[ld * NC + c] : sm.a.g1.nF[e][q][ld * NC + c]) + FP[ld * NC + c]);
            double BF = B[ld] * fs;
            double lf = lB * (uPL[c] - Uf[c]);
            sm.BFL[e][f][c] = BF - lf;
        }
        // the face term of the CFL sum (lambda_B_CFL :283-291)
        sm.lamB[e][f] = (lproj && a.stage1) ? find_alpha<DIM>(Uq, ut, a.POSTOL) * lB + 0.5 * nrm * wsf : lB;
    }
    __syncthreads();

    // ---- phase 3a: high-order volume pairs (flux_differencing_volume! :164-176, accumulate_QF1! :201-211)
    for (int w = tid; w < E * NPH; w += NT) {
        const int e = w / NPH, p = w - e * NPH;
        const int k = k0 + e;
        if (k >= K) continue;
        const int i = T.hp_i[p], j = T.hp_j[p], dir = T.hp_dir[p];
        const Prim<DIM>& Pi = i < NQ ? sm.a.g1.np[e][i] : sm.a.g1.fp[e][i - NQ];
        const Prim<DIM>& Pj = j < NQ ? sm.a.g1.np[e][j] : sm.a.g1.fp[e][j - NQ];
        const Geo G = a.geo[k];
        double Sd;
        if (DIM == 1) Sd = G.rxJ * T.hp_cr[p];
        else Sd = dir == 0 ? G.rxJ * T.hp_cr[p] + G.sxJ * T.hp_cs[p] : G.ryJ * T.hp_cr[p] + G.syJ * T.hp_cs[p];
        double F[NC];
        fS_dir<DIM>(g, Pi, Pj, dir, F);
        for (int c = 0; c < NC; c++) sm.x.p.PH[e][p][c] = Sd * F[c];
    }
    // ---- phase 3b: low-order volume pairs (accumulate_low_order_rhs_volume! :139-166, graph_viscosity rhs_utils.jl:105-128)
    for (int w = tid; w < E * NPL; w += NT) {
        const int e = w / NPL, p = w - e * NPL;
        const int k = k0 + e;
        if (k >= K) continue;
        const int i = T.lp_i[p], j = T.lp_j[p], dir = T.lp_dir[p];
        const Geo G = a.geo[k];
        double S0[2], nrm, nij[2], nji[2];
        if (DIM == 1) {
            S0[0] = G.rxJ * T.lp_cr[p];
            S0[1] = 0.0;
            nrm = fabs(S0[0]);
        } else {
            S0[0] = G.rxJ * T.lp_cr[p] + G.sxJ * T.lp_cs[p];
            S0[1] = G.ryJ * T.lp_cr[p] + G.syJ * T.lp_cs[p];
            nrm = jnorm2(S0[0], S0[1]);
        }
        nij[0] = S0[0] / nrm;
        nij[1] = S0[1] / nrm;
        nji[0] = -nij[0];
        nji[1] = -nij[1];
        const double* ui = sm.U[e][i];
        const double* uj = sm.U[e][j];
        const double ws = jmax(Eq<DIM>::wavespeed(g, ui, nij), Eq<DIM>::wavespeed(g, uj, nji));
        const double lam = nrm * ws;
        for (int c = 0; c < NC; c++) {
            double Favg = 0.5 * (sm.a.g1.nF[e][i][dir * NC + c] + sm.a.g1.nF[e][j][dir * NC + c]);
            double visc = lam * (uj[c] - ui[c]);
            sm.x.p.PL[e][p][c] = 2.0 * S0[dir] * Favg - visc;
        }
        sm.x.p.PL[e][p][NC] = lam;
    }
    __syncthreads();

    // ---- phase 4: node assembly of rhsxyH, rhsxyL, CFL and u_L
    double dtmin = INFINITY;
    const double dt_lim = a.sc->dt_lim;
    for (int w = tid; w < E * NQ; w += NT) {
        const int e = w / NQ, i = w - e * NQ;
        const int k = k0 + e;
        if (k >= K) continue;
        const Geo G = a.geo[k];
        // QF1 of hybrid node i and of the face hybrid nodes at i, in pair order
        double Q[DIM][NC], A[DIM][NC], Bv[DIM][NC];
        for (int d = 0; d < DIM; d++)
            for (int c = 0; c < NC; c++) { Q[d][c] = 0.0; Bv[d][c] = 0.0; }
        for (int m = T.hl_start[i]; m < T.hl_start[i + 1]; m++) {
            const int p = T.hl_pair[m], d = T.hp_dir[p];
            if (T.hl_sign[m] > 0)
                for (int c = 0; c < NC; c++) Q[d][c] += sm.x.p.PH[e][p][c];
            else
                for (int c = 0; c < NC; c++) Q[d][c] += -sm.x.p.PH[e][p][c];
        }
        // assemble_rhs! (:331-361): generic mul! adds column by column from zero.  On a Gauss
        // element whose entropy-projection parameters are limited (is_Vf_limited :326-328) the
        // matrices are MinvVhT_new = (1 ./ wq) .* [I Vf_new'] with Vf_new = l Vf + (1-l) Vf_low
        // (project_flux_difference_to_quad!(::ScaledExtrapolation) :288-319)
        bool limited = false;
        if (BASIS)
            for (int f = 0; f < NFP; f++) limited = limited || sm.th[BASIS ? e : 0][f] < 1.0;
        const double av = limited ? T.rwq[i] * 1.0 : T.a_v[i];
        for (int d = 0; d < DIM; d++)
            for (int c = 0; c < NC; c++) A[d][c] = 0.0 + av * Q[d][c];
        for (int m = 0; m < T.nh[i]; m++) {
            const int f = T.hface[i][m], h = NQ + f;
            double Qf[DIM][NC];
            for (int d = 0; d < DIM; d++)
                for (int c = 0; c < NC; c++) Qf[d][c] = 0.0;
            for (int mm = T.hl_start[h]; mm < T.hl_start[h + 1]; mm++) {
                const int p = T.hl_pair[mm], d = T.hp_dir[p];
                if (T.hl_sign[mm] > 0)
                    for (int c = 0; c < NC; c++) Qf[d][c] += sm.x.p.PH[e][p][c];
                else
                    for (int c = 0; c < NC; c++) Qf[d][c] += -sm.x.p.PH[e][p][c];
            }
            double ah = T.a_h[i][m], af = T.a_f[i][m];
            if (limited) {
                const double l = sm.th[BASIS ? e : 0][f];
                ah = T.rwq[i] * (l * T.vf_h[i][m] + (1.0 - l) * T.vfl_h[i][m]);
                af = ah;
            }
            for (int d = 0; d < DIM; d++)
                for (int c = 0; c < NC; c++) A[d][c] += ah * Qf[d][c];
            const int ld = T.ld[f];
            for (int c = 0; c < NC; c++) Bv[ld][c] += af * sm.BFH[e][f][c];
        }
        const double J = G.J;
        double rH[DIM][NC];
        for (int d = 0; d < DIM; d++)
            for (int c = 0; c < NC; c++) rH[d][c] = -(A[d][c] + Bv[d][c]) / J;

        // low order: rhsxyL = 0 - Q0F1, then -= BF_L at the face nodes, then / (Jq*wq)
        double rL[DIM][NC];
        for (int d = 0; d < DIM; d++)
            for (int c = 0; c < NC; c++) Q[d][c] = 0.0;
        for (int m = T.ll_start[i]; m < T.ll_start[i + 1]; m++) {
            const int p = T.ll_pair[m], d = T.lp_dir[p];
            if (T.ll_sign[m] > 0)
                for (int c = 0; c < NC; c++) Q[d][c] += sm.x.p.PL[e][p][c];
            else
                for (int c = 0; c < NC; c++) Q[d][c] += -sm.x.p.PL[e][p][c];
        }
        for (int d = 0; d < DIM; d++)
            for (int c = 0; c < NC; c++) rL[d][c] = 0.0 - Q[d][c];
        for (int m = 0; m < T.nf[i]; m++) {
            const int f = T.nface[i][m], ld = T.ld[f];
            for (int c = 0; c < NC; c++) rL[ld][c] = rL[ld][c] - sm.BFL[e][f][c];
        }
        const double wJ = J * T.wq[i];
        for (int d = 0; d < DIM; d++)
            for (int c = 0; c < NC; c++) rL[d][c] = rL[d][c] / wJ;

        // CFL dt (calculate_lambda_and_low_order_CFL! :222-243, lambda_i :263-281)
        if (a.stage1) {
            double li = 0.0;
            for (int m = T.cl_start[i]; m < T.cl_start[i + 1]; m++) li += sm.x.p.PL[e][T.cl_pair[m]][NC];
            for (int m = 0; m < T.nq2fq[i]; m++) li += sm.lamB[e][T.q2fq[i][m]];
            const double dti = a.CFL * 0.5 * wJ / li;
            dtmin = dti < dtmin ? dti : dtmin;
        }
        double rhsL[NC];
        for (int c = 0; c < NC; c++) {
            double s = rL[0][c];
            if (DIM == 2) s = s + rL[DIM - 1][c];
            rhsL[c] = s;
        }
        // u_L = Uq + dt*rhsL (subcell.jl:269)
        for (int c = 0; c < NC; c++) sm.a.g2.UL[e][i][c] = sm.U[e][i][c] + dt_lim * rhsL[c];
        // TVD bounds' first pass: rho_L = rho + dt*rhsL_rho of every node (initialize_TVD_bounds!
        // subcell.jl:112-118), the same bits the limiter's u_L holds in the second pass
        if (OPT == OPT_SUBCELL_GENERAL && a.tvd_pre) a.rhoL_out[(size_t)k * NQ + i] = sm.a.g2.UL[e][i][0];
        for (int d = 0; d < DIM; d++)
            for (int c = 0; c < NC; c++) {
                sm.a.g2.RH[e][i][d][c] = rH[d][c];
                sm.a.g2.RL[e][i][d][c] = rL[d][c];
            }
        if (a.dbg_rhsH) {
            for (int c = 0; c < NC; c++) {
                double s = rH[0][c];
                if (DIM == 2) s = s + rH[DIM - 1][c];
                a.dbg_rhsH[((size_t)k * NQ + i) * NC + c] = s;
                a.dbg_rhsL[((size_t)k * NQ + i) * NC + c] = rhsL[c];
            }
        }
    }
    if (OPT == OPT_SUBCELL_GENERAL && a.tvd_pre) return;   // uniform over the workgroup
    if (a.stage1) {
        // block min, then one global atomic min on the (positive) bit pattern
        for (int off = 32; off > 0; off >>= 1) {
            double o = __shfl_xor(dtmin, off, 64);
            dtmin = o < dtmin ? o : dtmin;
        }
        if ((tid & 63) == 0) sm.red[tid >> 6] = dtmin;
    }
    __syncthreads();
    if (a.stage1 && tid == 0) {
        double m = sm.red[0];
        for (int x = 1; x < NT / 64; x++) m = sm.red[x] < m ? sm.red[x] : m;
        a.dtp[blockIdx.x] = m;   // reduced by dt_reduce_kernel (one-address atomics serialise across XCDs)
    }

    if (OPT == OPT_ZHANG_SHU) {
        // apply_zhang_shu_limiter! (zhangshu.jl:4-24): one parameter per element, the largest l
        // keeping u_L + l dt (rhsH - rhsL) within (zeta rho_L, zeta rhoe_L) at every node
        // (zhang_shu_bound_limiter! :35-45, limiting_param(::ZhangShuLimiter) limiter_utils.jl:4-8)
        for (int w = tid; w < E * NQ; w += NT) {
            const int e = w / NQ, i = w - e * NQ;
            if (k0 + e >= K) continue;
            const double* uL = sm.a.g2.UL[e][i];
            double P[NC];
            for (int c = 0; c < NC; c++) {
                double h = sm.a.g2.RH[e][i][0][c], lo = sm.a.g2.RL[e][i][0][c];
                if (DIM == 2) {
                    h = h + sm.a.g2.RH[e][i][DIM - 1][c];
                    lo = lo + sm.a.g2.RL[e][i][DIM - 1][c];
                }
                P[c] = dt_lim * (h - lo);
            }
            sm.x.fb.l[e][i] = jmin(1.0, bound_rho_rhoe<DIM>(uL, P, a.zeta * uL[0], a.zeta * Eq<DIM>::rhoe(uL), INFINITY,
                                                            INFINITY, a.ZEROTOL));
        }
        __syncthreads();
        for (int w = tid; w < E * NQ; w += NT) {
            const int e = w / NQ, i = w - e * NQ;
            const int k = k0 + e;
            if (k >= K) continue;
            double Lk = 1.0;
            for (int q = 0; q < NQ; q++) Lk = jmin(Lk, sm.x.fb.l[e][q]);
            const double l = jmin(Lk, a.blend ? a.blend[k] : 1.0);
            for (int c = 0; c < NC; c++) {
                double h = sm.a.g2.RH[e][i][0][c], lo = sm.a.g2.RL[e][i][0][c];
                if (DIM == 2) {
                    h = h + sm.a.g2.RH[e][i][DIM - 1][c];
                    lo = lo + sm.a.g2.RL[e][i][DIM - 1][c];
                }
                a.rhsUA[((size_t)k * NQ + i) * NC + c] = (1.0 - l) * lo + l * h;
            }
            if (a.dbg_L && i == 0) a.dbg_L[k] = Lk;
            if (a.zsL && i == 0) a.zsL[k] = Lk;
        }
        // no subcell faces: the update kernel's symmetrisation finds nothing to correct
        for (int w = tid; w < E * NFP; w += NT) {
            const int e = w / NFP, f = w - e * NFP;
            const int k = k0 + e;
            if (k >= K) continue;
            a.lA[(size_t)k * NFP + f] = 1.0;
            for (int c = 0; c < NC; c++) a.dfA[((size_t)k * NFP + f) * NC + c] = 0.0;
        }
        return;
    }

    // ---- phase 5a: subcell flux potentials along every line (accumulate_f_bar! subcell.jl:163-206)
    for (int w = tid; w < E * NLINE; w += NT) {
        const int e = w / NLINE, line = w - e * NLINE;
        const int k = k0 + e;
        if (k >= K) continue;
        const double J = a.geo[k].J;
        const int d = (DIM == 1 || line < N1D) ? 0 : 1;
        const int f0 = line_face<DIM, N1D>(line, 0);
        const int base = line * (N1D + 1);
        double fH[NC], fL[NC];
        for (int c = 0; c < NC; c++) {
            fH[c] = sm.BFH[e][f0][c];
            fL[c] = sm.BFL[e][f0][c];
            sm.x.fb.fH[e][base][c] = fH[c];
            sm.x.fb.fL[e][base][c] = fL[c];
        }
        for (int s = 1; s <= N1D; s++) {
            const int q = line_node<DIM, N1D>(line, s - 1);
            const double wJ = T.wq[q] * J;
            for (int c = 0; c < NC; c++) {
                fH[c] = fH[c] + wJ * sm.a.g2.RH[e][q][d][c];
                fL[c] = fL[c] + wJ * sm.a.g2.RL[e][q][d][c];
                sm.x.fb.fH[e][base + s][c] = fH[c];
                sm.x.fb.fL[e][base + s][c] = fL[c];
            }
        }
    }
    __syncthreads();

    // ---- phase 5b: limiter parameter of every subcell face (subcell_bound_limiter! :248-349)
    const double zeta = a.zeta, ZEROTOL = a.ZEROTOL;
    constexpr double fac = 2.0 * DIM;                      // 2 in 1D, 4 in 2D
    for (int w = tid; w < E * NFB; w += NT) {
        const int e = w / NFB, r = w - e * NFB;
        const int line = r / (N1D + 1), s = r - line * (N1D + 1);
        const int k = k0 + e;
        if (k >= K) continue;
        const double J = a.geo[k].J;
        double dfb[NC], P[NC];
        for (int c = 0; c < NC; c++) dfb[c] = sm.x.fb.fH[e][r][c] - sm.x.fb.fL[e][r][c];
        // the bound's limiting_param (limiter_utils.jl:10-24); Lphi is the node's lower entropy bound
        auto lim = [&](int q) {
            if (OPT == OPT_SUBCELL_GENERAL && (a.bound != 0 && a.bound != 3 && a.bound != 4)) {
                const double* uL = sm.a.g2.UL[e][q];
                const int ent = a.bound % 5;
                double Lrho = zeta * uL[0], Urho = INFINITY;
                if (a.bound >= 5) {   // rho_bound(::TVD...) (subcell.jl:350-368): the stencil's extrema of rho_L
                    Lrho = Urho = a.rhoL[(size_t)k * NQ + q];
                    for (int x = 0; x < 2 * DIM; x++) {
                        const int st = T.stn[q][x];
                        const double r = st >= 0 ? a.rhoL[(size_t)k * NQ + st] : a.rhoL[a.nb_node[(size_t)k * NFP + (-st - 1)]];
                        Lrho = jmin(Lrho, r);
                        Urho = jmax(Urho, r);
                    }
                }
                if (ent == 1 || ent == 2)
                    return limiting_param_phi<DIM>(uL, P, Lrho, Urho, zeta, ZEROTOL, a.lphi[(size_t)k * NQ + q], a.POSTOL, g);
                const double Lrhoe = zeta * Eq<DIM>::rhoe(uL);
                // the 4-of-5 destructuring (limiter_utils.jl:11) takes (Lrho, Lrhoe, Lphi = 0, Urho)
                return a.upper_mode == 0 ? bound_rho_rhoe<DIM>(uL, P, Lrho, Lrhoe, 0.0, Urho, ZEROTOL)
                                         : bound_rho_rhoe<DIM>(uL, P, Lrho, Lrhoe, Urho, INFINITY, ZEROTOL);
            }
            return limiting_param<DIM>(sm.a.g2.UL[e][q], P, zeta, ZEROTOL, a.upper_mode);
        };
        double l = 1.0;
        if (s < N1D) {   // node on the right of (above) the subcell face
            const int q = line_node<DIM, N1D>(line, s);
            const double wJ = T.wq[q] * J;
            for (int c = 0; c < NC; c++) P[c] = -fac * dt_lim * dfb[c] / wJ;
            l = jmin(l, lim(q));
        }
        if (s > 0) {     // node on the left of (below) the subcell face
            const int q = line_node<DIM, N1D>(line, s - 1);
            const double wJ = T.wq[q] * J;
            for (int c = 0; c < NC; c++) P[c] = fac * dt_lim * dfb[c] / wJ;
            l = jmin(l, lim(q));
        }
        // blending factor of the shock capture (shock_capture.jl:111-132; 1 without)
        l = jmin(l, (OPT == OPT_SUBCELL_GENERAL && a.blend) ? a.blend[k] : 1.0);
        sm.x.fb.l[e][r] = l;
        if (s == 0 || s == N1D) {
            const int f = line_face<DIM, N1D>(line, s);
            a.lA[(size_t)k * NFP + f] = l;
            for (int c = 0; c < NC; c++) a.dfA[((size_t)k * NFP + f) * NC + c] = dfb[c];
        }
        if (a.dbg_L) {
            const int d = (DIM == 1 || line < N1D) ? 0 : 1;
            const int o = d == 0 ? line : line - N1D;
            const int idx = DIM == 1 ? s : (d == 0 ? s + (N1D + 1) * o : o + N1D * s);
            a.dbg_L[idx + (size_t)S::NLS * (d + DIM * (size_t)k)] = l;
            if (DIM == 1 && s == N1D)
                for (int x = N1D + 1; x < S::NLS; x++) a.dbg_L[x + (size_t)S::NLS * k] = 1.0;
        }
    }
    __syncthreads();

    // ---- phase 5b': cell-entropy bounds (enforce_ES_subcell! subcell.jl:462-466; the interface step
    //      :745-789 is the Gauss basis's, for Lobatto the interface fluxes coincide)
    if constexpr (DIM == 2 && OPT == OPT_SUBCELL_GENERAL) {
        if (a.bound % 5 >= 3) {
            for (int w = tid; w < E * NQ; w += NT) {
                const int e = w / NQ, i = w - e * NQ;
                if (k0 + e >= K) continue;
                Eq<DIM>::v_of_u(g, sm.U[e][i], sm.a.g3.V[e][i]);
            }
            __syncthreads();
            for (int w = tid; w < E * DIM; w += NT) {
                const int e = w / DIM, d = w - e * DIM;
                const int k = k0 + e;
                if (k >= K) continue;
                int changed[N1D * N1D];
                const int n = es_subcell_volume<N1D>(d, sm.x.fb.l[e], &sm.x.fb.fH[e][0][0], &sm.x.fb.fL[e][0][0],
                                                     &sm.a.g3.V[e][0][0], &sm.U[e][0][0], T, a.geo[k], g, a.bound,
                                                     a.cell_beta, a.sfac ? a.sfac[k] : 0.0, a.ZEROTOL, changed);
                if (a.dbg_L)
                    for (int m = 0; m < n; m++) {
                        const int r = changed[m];
                        const int line = r / (N1D + 1), sp = r - line * (N1D + 1);
                        const int o = d == 0 ? line : line - N1D;
                        const int idx = d == 0 ? sp + (N1D + 1) * o : o + N1D * sp;
                        a.dbg_L[idx + (size_t)S::NLS * (d + DIM * (size_t)k)] = sm.x.fb.l[e][r];
                    }
            }
            __syncthreads();
        }
    }

    // ---- phase 5c: rhsU with the element's own parameters (accumulate_f_bar_limited! :841-876,
    //      apply_subcell_limiter! :894-924)
    for (int w = tid; w < E * NQ; w += NT) {
        const int e = w / NQ, i = w - e * NQ;
        const int k = k0 + e;
        if (k >= K) continue;
        const double wJ = T.wq[i] * a.geo[k].J;
        const int si = DIM == 1 ? i : i % N1D, sj = DIM == 1 ? 0 : i / N1D;
        double rx[NC], ry[NC];
        {
            const int rl = (DIM == 1 ? 0 : sj) * (N1D + 1) + si;   // left subcell face of the node
            const double lL = sm.x.fb.l[e][rl], lR = sm.x.fb.l[e][rl + 1];
            for (int c = 0; c < NC; c++) {
                const double fl = lL * sm.x.fb.fH[e][rl][c] + (1.0 - lL) * sm.x.fb.fL[e][rl][c];
                const double fr = lR * sm.x.fb.fH[e][rl + 1][c] + (1.0 - lR) * sm.x.fb.fL[e][rl + 1][c];
                rx[c] = (fr - fl) / wJ;
            }
        }
        if (DIM == 2) {
            const int rb = (N1D + si) * (N1D + 1) + sj;            // bottom subcell face of the node
            const double lB = sm.x.fb.l[e][rb], lT = sm.x.fb.l[e][rb + 1];
            for (int c = 0; c < NC; c++) {
                const double fb = lB * sm.x.fb.fH[e][rb][c] + (1.0 - lB) * sm.x.fb.fL[e][rb][c];
                const double ft = lT * sm.x.fb.fH[e][rb + 1][c] + (1.0 - lT) * sm.x.fb.fL[e][rb + 1][c];
                ry[c] = (ft - fb) / wJ;
            }
        }
        for (int c = 0; c < NC; c++) a.rhsUA[((size_t)k * NQ + i) * NC + c] = DIM == 2 ? rx[c] + ry[c] : rx[c];
    }
}

// Limiter options, per stage before the element kernel (apply_rhs_limiter! limiter.jl:7-44):
// initialize_smoothness_indicator! (shock_capture.jl:47-94: rho*p at the nodes, its modal
// coefficients VDM_inv*(rho*p), the energy share of the top two mode shells), the Hennemann
// blending factor (shock_capture.jl:117-132), the smoothness factor (subcell.jl:927-956) and the
// lower bound of the modified specific entropy over each node's low-order stencil
// (initialize_lower_bound! subcell.jl:37-75), which reaches into the face neighbour.
template <int DIM, int N1D, int E, int NT>
__global__ __launch_bounds__(NT) void limiter_prep_kernel(PrepArgs a) {
    using S = Sz<DIM, N1D>;
    constexpr int NQ = S::NQ, NFP = S::NFP, NC = S::NC;
    __shared__ double sS[E][NQ];
    __shared__ double sI[E][NQ];
    __shared__ double sF[E];
    if (a.sc->done) return;
    const int tid = threadIdx.x;
    const int k0 = blockIdx.x * E;
    const int nel = min(E, a.K - k0);
    const ElemTables& T = *a.T;
    const double g = a.gamma;
    const int ent = a.bound % 5;   // the entropy part of the bound (TVD bounds add 5)
    const bool need_ind = a.shock != 0 || ent == 2 || ent == 4;
    for (int w = tid; w < nel * NQ; w += NT) {
        const int e = w / NQ, i = w - e * NQ;
        const double* src = a.U + ((size_t)(k0 + e) * NQ + i) * NC;
        double u[NC];
        for (int c = 0; c < NC; c++) u[c] = src[c];
        if (ent == 1 || ent == 2) sS[e][i] = s_modified<DIM>(g, u);
        if (need_ind) sI[e][i] = u[0] * Eq<DIM>::p(g, u);   // initialize_indicator! :82-94
    }
    __syncthreads();
    for (int e = tid; e < nel; e += NT) {
        const int k = k0 + e;
        double f = ent == 1 ? 1.0 : 0.0;
        if (need_ind) {
            double mod[NQ];
            for (int i = 0; i < NQ; i++) mod[i] = 0.0;
            for (int j = 0; j < NQ; j++)
                for (int i = 0; i < NQ; i++) mod[i] += T.vdm_inv[i][j] * sI[e][j];
            const int N = N1D - 1;
            double modeN = 0.0, modeNm1 = 0.0, total = 0.0;
            int count = 0;
            for (int j = 0; j <= (DIM == 2 ? N : 0); j++)
                for (int i = 0; i <= N; i++) {
                    const double energy = mod[count] * mod[count];
                    if (i == N || (DIM == 2 && j == N)) modeN += energy;
                    if (i == N - 1 || (DIM == 2 && j == N - 1)) modeNm1 += energy;
                    total += energy;
                    count++;
                }
            const double ind = jmax(modeN / total, modeNm1 / total);
            if (a.shock) {
                const double TN = a.shock_a * pow(10.0, -a.shock_c * pow((double)(N + 1), 0.25));
                const double alphamax = 0.5, alphaE0 = 0.0001;
                const double s_factor = log((1.0 - alphaE0) / alphaE0);
                const double alpha = 1.0 / (1.0 + exp(-s_factor / TN * (ind - TN)));
                a.blend[k] = jmax(jmin(1.0 - alpha, 1.0), alphamax);
            }
            if (ent == 2 || ent == 4) {
                const double kappa = 1.0;
                const double s0 = log(pow((double)N, -4.0)) / log(10.0);
                const double sk = log(ind) / log(10.0);
                if (sk < s0 - kappa) f = 0.0;
                else if (sk > s0 + kappa) f = 1.0;
                else f = 0.5 - 0.5 * sin(3.141592653589793 * (sk - s0) / (2.0 * kappa));
            }
        }
        sF[e] = f;
        if (a.sfac) a.sfac[k] = f;
    }
    __syncthreads();
    if (ent == 0 || ent >= 3) return;
    for (int w = tid; w < nel * NQ; w += NT) {
        const int e = w / NQ, q = w - e * NQ;
        const int k = k0 + e;
        double lb = sS[e][q];
        for (int x = 0; x < 2 * DIM; x++) {
            const int st = T.stn[q][x];
            if (st >= 0) {
                lb = jmin(lb, sS[e][st]);
            } else {
                const double* src = a.U + (size_t)a.nb_node[(size_t)k * NFP + (-st - 1)] * NC;
                double u[NC];
                for (int c = 0; c < NC; c++) u[c] = src[c];
                lb = jmin(lb, s_modified<DIM>(g, u));
            }
        }
        a.lphi[(size_t)k * NQ + q] = sF[e] * lb + (1.0 - sF[e]) * a.sc->s_min;
    }
}

// initialize_s_modified! (subcell.jl:19-35): at the first stage of t0, the minimum of
// rhoe*rho^-gamma over every node, per workgroup then in one workgroup
template <int DIM>
__global__ __launch_bounds__(256) void s_min_partial_kernel(const double* U, int64_t nodes, double g, const double* tp,
                                                            double tv, double t0, double* part) {
    __shared__ double red[4];
    const double t = tp ? *tp : tv;
    if (!(t == t0)) return;
    const size_t gidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double m = INFINITY;
    if (gidx < (size_t)nodes) m = s_modified<DIM>(g, U + gidx * (DIM + 2));
    for (int off = 32; off > 0; off >>= 1) m = jmin(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = jmin(jmin(red[0], red[1]), jmin(red[2], red[3]));
}

__global__ __launch_bounds__(1024) void s_min_reduce_kernel(const double* part, int n, const double* tp, double tv,
                                                            double t0, StepScalars* sc) {
    __shared__ double red[16];
    const double t = tp ? *tp : tv;
    if (!(t == t0)) return;
    double m = INFINITY;
    for (int b = (int)threadIdx.x; b < n; b += 1024) m = jmin(m, part[b]);
    for (int off = 32; off > 0; off >>= 1) m = jmin(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int x = 1; x < 16; x++) m = jmin(m, red[x]);
        m = jmin(m, red[0]);
        sc->s_min = m;
    }
}

// Entropy projection of the face states on the Gauss basis (init_rhs! and entropy_projection!,
// rhs.jl:19-23, 113-133).  Per face node, the NodewiseScaledExtrapolation parameter theta by
// bisection (solve_theta! filter.jl:49-59, bisection nonlinear_solvers.jl:3-20) against the
// bounds of check_bound_on_face_node (filter.jl:87-98) taken from calc_face_values!
// (filter.jl:26-42), then u_tilde = u(sum_j (theta Vf + (1-theta) Vf_low)[f, j] v_j).  Rows of
// Vf and Vf_low vanish off the face's line, so every sum runs over the line's nodes in ascending
// order: the generic product's order without its zero terms.  Without the limiter theta = 1.
template <int DIM, int N1D, int E, int NT>
__global__ __launch_bounds__(NT) void entropy_projection_kernel(EPArgs a) {
    using S = Sz<DIM, N1D, 1>;
    constexpr int NQ = S::NQ, NFP = S::NFP, NC = S::NC;
    __shared__ double sU[E * NQ * NC];
    __shared__ double sV[E * NQ * NC];
    if (a.sc->done) return;
    const int tid = threadIdx.x;
    const int k0 = blockIdx.x * E;
    const int nel = min(E, a.K - k0);
    const ElemTables& T = *a.T;
    const double g = a.gamma;
    {
        const double* src = a.U + (size_t)k0 * NQ * NC;
        for (int w = tid; w < nel * NQ * NC; w += NT) sU[w] = src[w];
    }
    __syncthreads();
    // calculate_entropy_var! (rhs.jl:72-77)
    for (int w = tid; w < nel * NQ; w += NT) Eq<DIM>::v_of_u(g, &sU[w * NC], &sV[w * NC]);
    __syncthreads();
    for (int w = tid; w < nel * NFP; w += NT) {
        const int e = w / NFP, f = w - e * NFP;
        const size_t fi = (size_t)(k0 + e) * NFP + f;
        const double* Ue = &sU[e * NQ * NC];
        const double* Ve = &sV[e * NQ * NC];
        int nd[N1D];
        double vf[N1D], vl[N1D];
        for (int s = 0; s < N1D; s++) {
            nd[s] = T.fl_node[f][s] * NC;
            vf[s] = T.fl_vf[f][s];
            vl[s] = T.fl_vfl[f][s];
        }
        double vt[NC], ut[NC];
        auto project = [&](double l) {   // rhs.jl:89-92, filter.jl:118-121
            for (int c = 0; c < NC; c++) vt[c] = 0.0;
            for (int s = 0; s < N1D; s++) {
                const double coef = l * vf[s] + (1.0 - l) * vl[s];
                for (int c = 0; c < NC; c++) vt[c] += coef * Ve[nd[s] + c];
            }
        };
        double theta = 1.0;
        bool have_ut = false;
        if (a.limit) {
            double Uf[NC], VUf[NC];
            for (int c = 0; c < NC; c++) {
                Uf[c] = 0.0;
                VUf[c] = 0.0;
            }
            for (int s = 0; s < N1D; s++)
                for (int c = 0; c < NC; c++) Uf[c] += vf[s] * Ue[nd[s] + c];
            for (int s = 0; s < N1D; s++)
                for (int c = 0; c < NC; c++) VUf[c] += vf[s] * Ve[nd[s] + c];
            const double rhoef = Eq<DIM>::rhoe(Uf);
            const double eps = a.POSTOL;
            const double v_hi = jmin(a.zeta * VUf[NC - 1], -eps);
            const double r_lo = jmax((1.0 - a.eta) * Uf[0], eps), r_hi = (1.0 + a.eta) * Uf[0];
            const double e_lo = jmax((1.0 - a.eta) * rhoef, eps), e_hi = (1.0 + a.eta) * rhoef;
            auto admissible = [&](double l) -> bool {   // filter.jl:101-130, 87-98
                project(l);
                if (!(vt[NC - 1] < -eps)) return false;
                Eq<DIM>::u_of_v(g, vt, ut);
                const double re = Eq<DIM>::rhoe(ut);
                return vt[NC - 1] < v_hi && ut[0] > r_lo && ut[0] < r_hi && re > e_lo && re < e_hi;
            };
            if (admissible(1.0)) {
                have_ut = true;                       // u_tilde of theta = 1 is already formed
            } else {
                double x_valid = 0.0, x_invalid = 1.0;
                for (int it = 0; it <= 20; it++) {
                    const double x_new = 0.5 * (x_valid + x_invalid);
                    if (admissible(x_new)) x_valid = x_new;
                    else x_invalid = x_new;
                }
                theta = x_valid;
            }
        }
        if (!have_ut) {
            project(theta);
            Eq<DIM>::u_of_v(g, vt, ut);
        }
        for (int c = 0; c < NC; c++) a.ut[fi * NC + c] = ut[c];
        a.theta[fi] = theta;
    }
}

// Projected face states u_tilde of every face node from the state (entropy projection of
// the face nodes, rhs.jl:84-94), for states that did not come out of phase B.
template <int DIM, int N1D>
__global__ __launch_bounds__(256) void face_state_kernel(const double* U, const ElemTables* T, int K, double gamma,
                                                         double* ut) {
    using S = Sz<DIM, N1D>;
    constexpr int NQ = S::NQ, NFP = S::NFP, NC = S::NC;
    const size_t fi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= (size_t)K * NFP) return;
    const int k = (int)(fi / NFP), f = (int)(fi - (size_t)k * NFP);
    const double* u = U + ((size_t)k * NQ + T->fq2q[f]) * NC;
    double V[NC], uu[NC], o[NC];
    for (int c = 0; c < NC; c++) uu[c] = u[c];
    Eq<DIM>::v_of_u(gamma, uu, V);
    project_face<DIM>(gamma, T->vf_coef[f], V, o);
    for (int c = 0; c < NC; c++) ut[fi * NC + c] = o[c];
}

// Symmetrise the element-boundary parameters (subcell.jl:405-456), correct rhsU at the
// adjacent nodes, then the RK stage (SSPRK33.jl:33-39), the positivity check and the
// projected face states of the new state.
template <int DIM, int N1D>
__global__ __launch_bounds__(256) void phaseB_kernel(PhaseBArgs a) {
    using S = Sz<DIM, N1D>;
    constexpr int NQ = S::NQ, NFP = S::NFP, NC = S::NC;
    if (a.sc->done) return;
    const size_t gidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double dt = bits_to_double(a.sc->dt_bits);
    if (a.finalize && gidx == 0) {
        const double t_new = a.sc->t + dt;
        const int steps = a.sc->steps + 1;
        a.sc->t = t_new;
        a.sc->steps = steps;
        if (a.out_interval > 0) {
            // SSPRK33.jl:42-46: i = i + 1, then a history point when mod(i, output_interval) == 0
            // or |t - T| < 1e-10
            const int n = steps - a.steps_base;
            if (n >= 1 && n <= a.dt_cap) a.dt_log[n - 1] = dt;
            const long long i = a.i0 + n;
            if (i % a.out_interval == 0 || fabs(t_new - a.T_end) < 1e-10 || n >= a.dt_cap) a.sc->hold = 1;
        }
    }
    if (a.mode == MODE_STAGE1 && gidx == 0) a.sc->dt_lim = dt;  // stages 2 and 3 limit with the CFL dt
    if (gidx >= (size_t)a.K * NQ) return;
    const int k = (int)(gidx / NQ), i = (int)(gidx - (size_t)k * NQ);
    const ElemTables& T = *a.T;
    double r[NC];
    for (int c = 0; c < NC; c++) r[c] = a.rhsUA[gidx * NC + c];
    const double wJ = T.wq[i] * a.geo[k].J;
    for (int m = 0; m < T.nf[i]; m++) {
        const int f = T.nface[i][m];
        const size_t fi = (size_t)k * NFP + f;
        const double lown = a.lA[fi];
        double lsym;
        if (DIM == 1) {
            // the reference min's face 1 with slot `end` = Nq+N1D of element k-1, which the
            // limiter never writes (1.0): the element-boundary faces are not symmetrised
            lsym = f == 0 ? jmin(lown, 1.0) : lown;
        } else {
            lsym = jmin(lown, a.lA[a.nb_face[fi]]);
        }
        if (!(lsym == lown)) {
            for (int c = 0; c < NC; c++) {
                const double delta = (lsym - lown) * a.dfA[fi * NC + c];
                r[c] = T.bsign[f] > 0 ? r[c] + delta / wJ : r[c] - delta / wJ;
            }
        }
        if (a.dbg_L) {
            int idx, d;
            if (DIM == 1) {
                idx = f == 0 ? 0 : N1D;
                d = 0;
            } else if (f < 2 * N1D) {
                d = 0;
                idx = f < N1D ? (N1D + 1) * f : N1D + (N1D + 1) * (f - N1D);
            } else {
                d = 1;
                idx = f < 3 * N1D ? f - 2 * N1D : (f - 3 * N1D) + N1D * N1D;
            }
            a.dbg_L[idx + (size_t)S::NLS * (d + DIM * (size_t)k)] = lsym;
            if (DIM == 1 && f == 0) {
                const int km1 = (k - 1 + a.K) % a.K;
                a.dbg_L[(S::NLS - 1) + (size_t)S::NLS * (size_t)km1] = lsym;
            }
        }
    }
    if (a.mode == MODE_RHS) {
        for (int c = 0; c < NC; c++) a.rhsU[gidx * NC + c] = r[c];
        return;
    }
    double u[NC], w[NC], un[NC];
    for (int c = 0; c < NC; c++) {
        u[c] = a.Uin[gidx * NC + c];
        w[c] = a.W[gidx * NC + c];
    }
    if (a.mode == MODE_STAGE1) {
        for (int c = 0; c < NC; c++) un[c] = w[c] + dt * r[c];
    } else if (a.mode == MODE_STAGE2) {
        for (int c = 0; c < NC; c++) {
            double z = u[c] + dt * r[c];
            un[c] = 3.0 / 4.0 * w[c] + 1.0 / 4.0 * z;
        }
    } else {
        for (int c = 0; c < NC; c++) {
            double z = u[c] + dt * r[c];
            un[c] = 1.0 / 3.0 * w[c] + 2.0 / 3.0 * z;
        }
    }
    for (int c = 0; c < NC; c++) a.Uout[gidx * NC + c] = un[c];
    const double pr = Eq<DIM>::p(a.gamma, un);
    if (!(un[0] > 0.0) || !(pr > 0.0)) {
        atomicAdd(&a.sc->viol, 1ull);
        atomicMin(&a.sc->first_bad, ((unsigned long long)a.mode << 40) | (unsigned long long)gidx);
    }
    // projected face states of the new state, once per node (corner nodes feed two faces);
    // the Gauss basis forms them in entropy_projection_kernel from the whole line instead
    if (a.ut && T.nf[i] > 0) {
        double V[NC], o[NC];
        Eq<DIM>::v_of_u(a.gamma, un, V);
        double coef_prev = 0.0;
        for (int m = 0; m < T.nf[i]; m++) {
            const int f = T.nface[i][m];
            const double coef = T.vf_coef[f];
            if (m == 0 || coef != coef_prev) project_face<DIM>(a.gamma, coef, V, o);
            coef_prev = coef;
            for (int c = 0; c < NC; c++) a.ut[((size_t)k * NFP + f) * NC + c] = o[c];
        }
    }
}

__global__ void step_init_kernel(StepScalars* sc, double CFL, double dt0, double T) {
    if (threadIdx.x != 0) return;
    const double t = sc->t;
    if (!(t < T) || sc->hold) {
        sc->done = 1;
        return;
    }
    sc->done = 0;
    const double dt = jmin(CFL * dt0, T - t);   // SSPRK33.jl:30
    sc->dt_lim = dt;
    sc->dt_bits = double_to_bits(dt);
}

__global__ void rhs_init_kernel(StepScalars* sc, double dt_lim, double dt_cap) {
    if (threadIdx.x != 0) return;
    sc->done = 0;
    sc->dt_lim = dt_lim;
    sc->dt_bits = double_to_bits(dt_cap);
}

// dt = min(cap, min over the phase-A workgroups' minima); a NaN minimum is ignored, as the
// bit-pattern minimum over positive doubles ignores it
__global__ __launch_bounds__(1024) void dt_reduce_kernel(StepScalars* sc, const double* dtp, int n) {
    __shared__ double red[16];
    if (sc->done) return;
    double m = INFINITY;
    for (int b = (int)threadIdx.x; b < n; b += 1024) m = dtp[b] < m ? dtp[b] : m;
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o < m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int x = 0; x < 16; x++) m = red[x] < m ? red[x] : m;
        const double cap = bits_to_double(sc->dt_bits);
        sc->dt_bits = double_to_bits(m < cap ? m : cap);
    }
}

// order-preserving map of a double to an unsigned key (for atomicMin)
__device__ __forceinline__ unsigned long long order_key(double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

template <int DIM>
__global__ void positivity_kernel(const double* U, int64_t nodes, double gamma, unsigned long long* keys,
                                  unsigned long long* cnt) {
    const size_t gidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double mr = INFINITY, mp = INFINITY;
    unsigned long long bad = 0;
    if (gidx < (size_t)nodes) {
        const double* u = U + gidx * (DIM + 2);
        const double pr = Eq<DIM>::p(gamma, u);
        mr = u[0];
        mp = pr;
        bad = (!(u[0] > 0.0) || !(pr > 0.0)) ? 1ull : 0ull;
    }
    for (int off = 32; off > 0; off >>= 1) {
        double o1 = __shfl_xor(mr, off, 64), o2 = __shfl_xor(mp, off, 64);
        mr = o1 < mr ? o1 : mr;
        mp = o2 < mp ? o2 : mp;
        bad += __shfl_xor(bad, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&keys[0], order_key(mr));
        atomicMin(&keys[1], order_key(mp));
        if (bad) atomicAdd(cnt, bad);
    }
}

// --------------------------------------------------------------------------- launchers
template <int DIM, int N1D, int E, int NT, int BASIS, int OPT>
static int launchA(const PhaseAArgs& a, hipStream_t s) {
    const int blocks = (a.K + E - 1) / E;
    hipLaunchKernelGGL((phaseA_kernel<DIM, N1D, E, NT, BASIS, OPT>), dim3(blocks), dim3(NT), 0, s, a);
    return (int)hipGetLastError();
}
template <int DIM, int N1D>
static int launchB(const PhaseBArgs& a, hipStream_t s) {
    const size_t nodes = (size_t)a.K * Sz<DIM, N1D>::NQ;
    const int blocks = (int)((nodes + 255) / 256);
    hipLaunchKernelGGL((phaseB_kernel<DIM, N1D>), dim3(blocks), dim3(256), 0, s, a);
    return (int)hipGetLastError();
}

// Launch shapes of phase A (elements per workgroup E, threads NT).  All variants compute
// bit-identical results (the gather order does not depend on the shape); the autotuner
// picks one per (Nd, N).  The first entry is the default.
struct VariantA {
    int E, NT;
    int (*launch[3])(const PhaseAArgs&, hipStream_t);   // by OPT_*
};
#define VA(D, N, E, T) {E, T, {&launchA<D, N, E, T, 0, 0>, &launchA<D, N, E, T, 0, 1>, &launchA<D, N, E, T, 0, 2>}}
static const VariantA kV1_2[] = {VA(1, 2, 32, 64), VA(1, 2, 32, 128)};
static const VariantA kV1_3[] = {VA(1, 3, 21, 64), VA(1, 3, 21, 128)};
static const VariantA kV1_4[] = {VA(1, 4, 16, 64), VA(1, 4, 16, 128)};
static const VariantA kV1_5[] = {VA(1, 5, 12, 64), VA(1, 5, 12, 128)};
static const VariantA kV2_2[] = {VA(2, 2, 16, 64), VA(2, 2, 16, 128), VA(2, 2, 32, 128), VA(2, 2, 32, 256)};
static const VariantA kV2_3[] = {VA(2, 3, 7, 64), VA(2, 3, 7, 128), VA(2, 3, 14, 128), VA(2, 3, 14, 256), VA(2, 3, 4, 64)};
static const VariantA kV2_4[] = {VA(2, 4, 4, 64), VA(2, 4, 4, 128), VA(2, 4, 4, 256), VA(2, 4, 8, 128),
                                 VA(2, 4, 8, 256), VA(2, 4, 2, 64), VA(2, 4, 2, 128)};
static const VariantA kV2_5[] = {VA(2, 5, 2, 64), VA(2, 5, 2, 128), VA(2, 5, 4, 128), VA(2, 5, 4, 256), VA(2, 5, 1, 64)};
#undef VA
// Gauss: more face-volume pairs per element, so fewer elements per workgroup
#define VG(D, N, E, T) {E, T, {&launchA<D, N, E, T, 1, 0>, &launchA<D, N, E, T, 1, 1>, &launchA<D, N, E, T, 1, 2>}}
static const VariantA kG1_2[] = {VG(1, 2, 32, 64), VG(1, 2, 32, 128)};
static const VariantA kG1_3[] = {VG(1, 3, 16, 64), VG(1, 3, 16, 128)};
static const VariantA kG1_4[] = {VG(1, 4, 12, 64), VG(1, 4, 12, 128)};
static const VariantA kG1_5[] = {VG(1, 5, 8, 64), VG(1, 5, 8, 128)};
static const VariantA kG2_2[] = {VG(2, 2, 16, 64), VG(2, 2, 16, 128), VG(2, 2, 32, 256)};
static const VariantA kG2_3[] = {VG(2, 3, 4, 64), VG(2, 3, 7, 128), VG(2, 3, 7, 64), VG(2, 3, 14, 256)};
static const VariantA kG2_4[] = {VG(2, 4, 4, 128), VG(2, 4, 4, 64), VG(2, 4, 2, 64), VG(2, 4, 8, 256)};
static const VariantA kG2_5[] = {VG(2, 5, 2, 128), VG(2, 5, 2, 64), VG(2, 5, 1, 64), VG(2, 5, 4, 256)};
#undef VG

static const VariantA* variants(int Nd, int N1D, int basis, int* n) {
#define CASE(D, N, B, arr)                                 \
    if (Nd == D && N1D == N && basis == B) {               \
        *n = (int)(sizeof(arr) / sizeof(arr[0]));          \
        return arr;                                        \
    }
    CASE(1, 2, 0, kV1_2) CASE(1, 3, 0, kV1_3) CASE(1, 4, 0, kV1_4) CASE(1, 5, 0, kV1_5)
    CASE(2, 2, 0, kV2_2) CASE(2, 3, 0, kV2_3) CASE(2, 4, 0, kV2_4) CASE(2, 5, 0, kV2_5)
    CASE(1, 2, 1, kG1_2) CASE(1, 3, 1, kG1_3) CASE(1, 4, 1, kG1_4) CASE(1, 5, 1, kG1_5)
    CASE(2, 2, 1, kG2_2) CASE(2, 3, 1, kG2_3) CASE(2, 4, 1, kG2_4) CASE(2, 5, 1, kG2_5)
#undef CASE
    *n = 0;
    return nullptr;
}

int num_variants_A(int Nd, int N1D, int basis) {
    int n = 0;
    variants(Nd, N1D, basis, &n);
    return n;
}

int variant_shape_A(int Nd, int N1D, int basis, int v, int* E, int* NT) {
    int n = 0;
    const VariantA* t = variants(Nd, N1D, basis, &n);
    if (!t || v < 0 || v >= n) return -1;
    *E = t[v].E;
    *NT = t[v].NT;
    return 0;
}

int launch_phaseA(int Nd, int N1D, int basis, int variant, int opt, const PhaseAArgs& a, void* stream) {
    int n = 0;
    const VariantA* t = variants(Nd, N1D, basis, &n);
    if (!t || variant < 0 || variant >= n || opt < 0 || opt > 2) return (int)hipErrorInvalidValue;
    return t[variant].launch[opt](a, (hipStream_t)stream);
}

template <int DIM, int N1D>
static int launchP(const PrepArgs& a, hipStream_t s) {
    constexpr int NQ = Sz<DIM, N1D>::NQ;
    constexpr int E = 256 / NQ;
    const int blocks = (a.K + E - 1) / E;
    hipLaunchKernelGGL((limiter_prep_kernel<DIM, N1D, E, 256>), dim3(blocks), dim3(256), 0, s, a);
    return (int)hipGetLastError();
}

int launch_limiter_prep(int Nd, int N1D, const PrepArgs& a, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (Nd == 1) {
        switch (N1D) {
            case 2: return launchP<1, 2>(a, s);
            case 3: return launchP<1, 3>(a, s);
            case 4: return launchP<1, 4>(a, s);
            case 5: return launchP<1, 5>(a, s);
        }
    } else {
        switch (N1D) {
            case 2: return launchP<2, 2>(a, s);
            case 3: return launchP<2, 3>(a, s);
            case 4: return launchP<2, 4>(a, s);
            case 5: return launchP<2, 5>(a, s);
        }
    }
    return (int)hipErrorInvalidValue;
}

int launch_s_min(const double* U, int64_t nodes, int Nd, double gamma, const double* tp, double t_value, double t0,
                 double* part, StepScalars* sc, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const int blocks = (int)((nodes + 255) / 256);
    if (Nd == 1)
        hipLaunchKernelGGL(s_min_partial_kernel<1>, dim3(blocks), dim3(256), 0, s, U, nodes, gamma, tp, t_value, t0, part);
    else
        hipLaunchKernelGGL(s_min_partial_kernel<2>, dim3(blocks), dim3(256), 0, s, U, nodes, gamma, tp, t_value, t0, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(s_min_reduce_kernel, dim3(1), dim3(1024), 0, s, part, blocks, tp, t_value, t0, sc);
    return (int)hipGetLastError();
}

// entropy projection: about 128 face nodes per workgroup
template <int DIM, int N1D>
static int launchEP(const EPArgs& a, hipStream_t s) {
    constexpr int NFP = Sz<DIM, N1D, 1>::NFP;
    constexpr int E = 128 / NFP;
    const int blocks = (a.K + E - 1) / E;
    hipLaunchKernelGGL((entropy_projection_kernel<DIM, N1D, E, 128>), dim3(blocks), dim3(128), 0, s, a);
    return (int)hipGetLastError();
}

int launch_entropy_projection(int Nd, int N1D, const EPArgs& a, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (Nd == 1) {
        switch (N1D) {
            case 2: return launchEP<1, 2>(a, s);
            case 3: return launchEP<1, 3>(a, s);
            case 4: return launchEP<1, 4>(a, s);
            case 5: return launchEP<1, 5>(a, s);
        }
    } else {
        switch (N1D) {
            case 2: return launchEP<2, 2>(a, s);
            case 3: return launchEP<2, 3>(a, s);
            case 4: return launchEP<2, 4>(a, s);
            case 5: return launchEP<2, 5>(a, s);
        }
    }
    return (int)hipErrorInvalidValue;
}

int launch_phaseB(int Nd, int N1D, const PhaseBArgs& a, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (Nd == 1) {
        switch (N1D) {
            case 2: return launchB<1, 2>(a, s);
            case 3: return launchB<1, 3>(a, s);
            case 4: return launchB<1, 4>(a, s);
            case 5: return launchB<1, 5>(a, s);
        }
    } else {
        switch (N1D) {
            case 2: return launchB<2, 2>(a, s);
            case 3: return launchB<2, 3>(a, s);
            case 4: return launchB<2, 4>(a, s);
            case 5: return launchB<2, 5>(a, s);
        }
    }
    return (int)hipErrorInvalidValue;
}

template <int DIM, int N1D>
static int launchF(const double* U, const ElemTables* T, int K, double gamma, double* ut, hipStream_t s) {
    const size_t n = (size_t)K * Sz<DIM, N1D>::NFP;
    hipLaunchKernelGGL((face_state_kernel<DIM, N1D>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, U, T, K, gamma, ut);
    return (int)hipGetLastError();
}

int launch_face_state(int Nd, int N1D, const double* U, const ElemTables* T, int K, double gamma, double* ut,
                      void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (Nd == 1) {
        switch (N1D) {
            case 2: return launchF<1, 2>(U, T, K, gamma, ut, s);
            case 3: return launchF<1, 3>(U, T, K, gamma, ut, s);
            case 4: return launchF<1, 4>(U, T, K, gamma, ut, s);
            case 5: return launchF<1, 5>(U, T, K, gamma, ut, s);
        }
    } else {
        switch (N1D) {
            case 2: return launchF<2, 2>(U, T, K, gamma, ut, s);
            case 3: return launchF<2, 3>(U, T, K, gamma, ut, s);
            case 4: return launchF<2, 4>(U, T, K, gamma, ut, s);
            case 5: return launchF<2, 5>(U, T, K, gamma, ut, s);
        }
    }
    return (int)hipErrorInvalidValue;
}

int launch_step_init(StepScalars* sc, double CFL, double dt0, double T, void* stream) {
    hipLaunchKernelGGL(step_init_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, sc, CFL, dt0, T);
    return (int)hipGetLastError();
}

int launch_rhs_init(StepScalars* sc, double dt_lim, double dt_cap, void* stream) {
    hipLaunchKernelGGL(rhs_init_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, sc, dt_lim, dt_cap);
    return (int)hipGetLastError();
}

int launch_dt_reduce(StepScalars* sc, const double* dtp, int n, void* stream) {
    hipLaunchKernelGGL(dt_reduce_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, sc, dtp, n);
    return (int)hipGetLastError();
}

int phaseA_blocks(int Nd, int N1D, int basis, int variant, int K) {
    int n = 0;
    const VariantA* t = variants(Nd, N1D, basis, &n);
    if (!t || variant < 0 || variant >= n) return -1;
    return (K + t[variant].E - 1) / t[variant].E;
}

int launch_positivity(const double* U, int64_t nodes, int Nd, double gamma, unsigned long long* keys,
                      unsigned long long* cnt, void* stream) {
    const int blocks = (int)((nodes + 255) / 256);
    if (Nd == 1)
        hipLaunchKernelGGL(positivity_kernel<1>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, U, nodes, gamma,
                           keys, cnt);
    else
        hipLaunchKernelGGL(positivity_kernel<2>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, U, nodes, gamma,
                           keys, cnt);
    return (int)hipGetLastError();
}

}  // namespace p2de
