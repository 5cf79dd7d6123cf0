// mixed_kernels.hip -- mixed-precision (FP32 bulk, FP64 limiter and RK) fused stage kernels of
// the limited ESDG scheme on 2D LGL quads, for gfx950.
//
// Per SSP-RK3 step: the CFL pass and its reduction (global min of the low-order time step per
// instance, low_order_graph_viscosity.jl:222-243), three fused stage launches (low- and
// high-order RHS, subcell limiter, RK stage, positivity check) and a post kernel (t += dt, then
// the next step's done flag and dt cap; one init launch starts the first step of a sequence).
// Each launch is a persistent workgroup pool: the tile
// list (E_N elements of one instance per tile) is cut into 8 contiguous chunks, one per XCD,
// and the workgroups of an XCD (blockIdx % 8) stride through its chunk, so neighbouring
// elements share an L2.  No global atomics sit on the tile path: device-scope atomics to one
// address serialise across the XCDs (tens of ns each), so the CFL minimum is written per
// tile and reduced per instance by a small kernel.
//
// Algebra that the LGL collocation makes exact, used here instead of computing roundoff:
//  * the entropy projection of a face node is u(v(u_q)) = u_q (rhs.jl:84-94 with Vf a
//    selection), so the high-order face state is the node state;
//  * then the high- and low-order surface fluxes are the same expression at interior faces
//    (flux_differencing.jl:262-272 and low_order_graph_viscosity.jl:184-198 with equal states
//    and wave speeds), so f_bar_H - f_bar_L = 0 on element-boundary subcell faces and the
//    limiter symmetrisation across elements (subcell.jl:418-456) changes nothing;
//  * a face-volume pair (Nq+f, fq2q[f]) enters MinvVhT*QF1 as (a_h - a_v)*S*f(u) = 0 when
//    MinvVhT[i,Nq+f] == MinvVhT[i,i] (checked on the host), so those pairs are not formed;
//  * f_bar on the far end of a line is -BF of the far face (telescoping), which is stored
//    exactly so that the element-interface fluxes are conservative to FP64 roundoff.
// The limited update is assembled from the same FP64 flux potentials the limiter saw (the
// limited flux of each subcell face is formed once, in FP64, from them and the parameter), so
// rho >= zeta*rho_L and rhoe >= zeta*rhoe_L hold to FP64 roundoff exactly as in the FP64 path.
#include <hip/hip_runtime.h>

#include "euler_device.h"
#include "p2de_internal.h"
#include "p2de_mixed.h"

namespace p2de {
namespace {

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrtf_fast(float x) { return __builtin_amdgcn_sqrtf(x); }

// FP64 reciprocal: hardware estimate plus two Newton steps (about 1 ulp)
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double bits_to_d(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long d_to_bits(double d) { return (unsigned long long)__double_as_longlong(d); }

// FP32 node state for the pair fluxes.  wsx/wsy are the wave speeds of wavespeed_estimate
// (compressible_Navier_Stokes.jl:48-62) along x and y: the rotated 1D state (rho, m.n, E)
// keeps the tangential kinetic energy inside its "internal" energy.
struct NodeP {
    float rho, u, v, p, beta, E, wsx, wsy;
};

// pfun / betafun / wavespeed with the cancellation-prone energy differences in FP64
__device__ __forceinline__ NodeP node_prims(double g, const double* U) {
    const double ri = rcp64(U[0]);
    const double u = U[1] * ri, v = U[2] * ri;
    const double kx = 0.5 * U[1] * u, ky = 0.5 * U[2] * v;
    const double gm1 = g - 1.0;
    const double p = gm1 * (U[3] - (kx + ky));
    const double px = gm1 * (U[3] - kx), py = gm1 * (U[3] - ky);
    NodeP P;
    P.rho = (float)U[0];
    P.u = (float)u;
    P.v = (float)v;
    P.p = (float)p;
    P.E = (float)U[3];
    P.beta = 0.5f * P.rho * rcpf(P.p);
    const float gri = (float)(g * ri);
    P.wsx = fabsf(P.u) + sqrtf_fast(gri * (float)px);
    P.wsy = fabsf(P.v) + sqrtf_fast(gri * (float)py);
    return P;
}

// Euler flux in direction d (fluxes, compressible_Navier_Stokes.jl:175-194)
__device__ __forceinline__ void flux_dir(const NodeP& P, int d, float F[4]) {
    const float un = d == 0 ? P.u : P.v;
    const float m = P.rho * un;
    F[0] = m;
    F[1] = fmaf(m, P.u, d == 0 ? P.p : 0.f);
    F[2] = fmaf(m, P.v, d == 0 ? 0.f : P.p);
    F[3] = un * (P.E + P.p);
}

// Logarithmic mean in ratio form: ln(b/a) = 2 atanh(f), f = (b-a)/(b+a), so the mean is
// (a+b) / (2 (1 + f^2/3 + f^4/5 + f^6/7 + ...)) (Ismail & Roe 2009).  Its relative accuracy
// does not depend on the magnitude of a and b (near-vacuum densities and betas included);
// the series (truncation < 1.2e-9) is used for f^2 < 1e-2, the logarithm of the ratio above.
// The reference's series (compressible_Navier_Stokes.jl:307-321) differs from the exact
// mean by < 1.2e-9 relative, below FP32 resolution.
__device__ __forceinline__ float logmean_f(float a, float b) {
    const float s = a + b, d = b - a;
    const float f = d * rcpf(s);
    const float u = f * f;
    if (u < 1e-2f) return 0.5f * s * rcpf(fmaf(u, fmaf(u, fmaf(u, 1.f / 7.f, 0.2f), 1.f / 3.f), 1.f));
    return d * rcpf(__logf(b * rcpf(a)));
}

// 1 / logmean(a, b), the same two forms: 2 (1 + f^2/3 + f^4/5 + f^6/7) / (a+b), ln(b/a) / (b-a)
__device__ __forceinline__ float inv_logmean_f(float a, float b) {
    const float rs = rcpf(a + b), d = b - a;
    const float f = d * rs;
    const float u = f * f;
    if (u < 1e-2f) return 2.f * rs * fmaf(u, fmaf(u, fmaf(u, 1.f / 7.f, 0.2f), 1.f / 3.f), 1.f);
    return __logf(b * rcpf(a)) * rcpf(d);
}

// Chandrashekar flux fS (compressible_Navier_Stokes.jl:196-249), components of direction d;
// ig2m1 = 1 / (2 (gamma - 1)).  The beta log mean enters only as its reciprocal.
__device__ __forceinline__ void fS_f(float ig2m1, const NodeP& L, const NodeP& R, int d, float F[4]) {
    const float rl = logmean_f(L.rho, R.rho), ibl = inv_logmean_f(L.beta, R.beta);
    const float ra = 0.5f * (L.rho + R.rho), ua = 0.5f * (L.u + R.u), va = 0.5f * (L.v + R.v);
    const float un = fmaf(L.u, R.u, L.v * R.v);
    const float pa = ra * rcpf(L.beta + R.beta);
    const float f4 = fmaf(0.5f * rl, un, fmaf(rl, ibl * ig2m1, pa));
    const float vn = d == 0 ? ua : va;
    const float m = rl * vn;
    F[0] = m;
    F[1] = fmaf(m, ua, d == 0 ? pa : 0.f);
    F[2] = fmaf(m, va, d == 0 ? 0.f : pa);
    F[3] = f4 * vn;
}

// limiting_param(::SubcellLimiter, ::PositivityBound) (limiter_utils.jl:10-14, 26-40, 52-90) in
// FP64 with precomputed bounds Lrho = zeta*rho(uL), Lrhoe = zeta*rhoe(uL).  upper_mode 0
// keeps the reference's bound tuple destructuring (Urho = 0, limiter_utils.jl:11).
__device__ __forceinline__ double lim_param(const double* uL, double Lrho, double Lrhoe, const double* P,
                                            double ZEROTOL, int upper_mode) {
    double l = 1.0;
    const double r0 = uL[0] + P[0];
    const bool lo = r0 < Lrho, up = upper_mode == 0 && r0 > 0.0;
    if (lo || up) {
        const double iP = rcp64(P[0]);
        if (lo) l = fmax((Lrho - uL[0]) * iP, 0.0);
        if (up) l = fmin(l, fmax(-uL[0] * iP, 0.0));
    }
    // the quadratic only returns values in (ZEROTOL, 1]: it cannot lower l = 0
    if (!(l > 0.0)) return l;
    const double a = P[0] * P[3] - 0.5 * (P[1] * P[1] + P[2] * P[2]);
    const double b = uL[3] * P[0] + uL[0] * P[3] - uL[1] * P[1] - uL[2] * P[2] - P[0] * Lrhoe;
    const double c = uL[3] * uL[0] - 0.5 * (uL[1] * uL[1] + uL[2] * uL[2]) - uL[0] * Lrhoe;
    // g(x) = a x^2 + b x + c is rho * (rhoe - Lrhoe) along uL + x P.  With g(0) > 0 and g(l) > 0,
    // no root lies in (0, l] when g is concave (its one positive root is beyond l) or its vertex
    // -b/(2a) is outside (0, l): the root selection below then leaves l unchanged, so the
    // square root and the division are skipped.
    if (c > 0.0 && fma(fma(a, l, b), l, c) > 0.0 && (a <= 0.0 || b >= 0.0 || -b >= 2.0 * a * l)) return l;
    const double disc = b * b - 4.0 * a * c;
    // a == 0 gives non-finite roots that fail every test of rhoe_quadratic_solve: l stays 1
    if (disc >= 0.0 && a != 0.0) {
        const double sq = sqrt(disc), i2a = rcp64(2.0 * a);
        const double x1 = (-b + sq) * i2a, x2 = (-b - sq) * i2a;
        double q = 1.0;
        if (x1 > ZEROTOL && x2 > ZEROTOL) q = fmin(x1, x2);
        else if (x1 > ZEROTOL && x2 < -ZEROTOL) q = x1;
        else if (x2 > ZEROTOL && x1 < -ZEROTOL) q = x2;
        l = fmin(l, q);
    }
    return l;
}

// LDS of one tile, component-major (structure of arrays) so that consecutive lanes touch
// consecutive words: node n = e*NQ + i, face node e*NFP + f, pairs e*NPV + p and e*NPL + p,
// subcell faces e*NFB + r.  Arrays whose lifetimes do not overlap share storage.
template <int N1D, int E>
struct SmemM {
    static constexpr int NQ = N1D * N1D, NFP = 4 * N1D;
    static constexpr int NPV = MixTab<N1D>::NPV, NPL = MixTab<N1D>::NPL, NFB = 2 * N1D * (N1D + 1);
    static constexpr int M = E * NQ, MF = E * NFP, MV = E * NPV, ML = E * NPL, MB = E * NFB;
    double U[4][M];
    double geo[E][4];            // J, rxJ, syJ, 1/J
    float BF[4][MF];             // B*f* - LF (high order == low order at interior faces)
    float BFL[4][MF];            // low-order surface term (differs at inflow faces)
    // subcell flux potentials, accumulated in FP64 and stored once rounded (fd: f_bar_L and
    // f_bar_H - f_bar_L per face and component); the limiter (5b) replaces each face's pair by
    // the limited flux f_bar_L + l (f_bar_H - f_bar_L) in FP64, in place (g), from the same
    // values it bounded, so the update (5c) is the convex combination the limiter bounded
    union {
        float2 fd[4][MB];
        double g[4][MB];
    } p;
    union {
        struct {                 // phases 1-4
            float np[8][M];      // NodeP fields
            float PH[4][MV];
            float PL[4][ML];
        } a;
        struct {                 // phases 5u-5c
            double UL[4][M];
            double Lrhoe[M];
            double iwJ[M];       // 1 / (wq J)
        } b;
    } x;
};

template <class S>
__device__ __forceinline__ NodeP ld_np(const S& sm, int n) {
    NodeP P;
    P.rho = sm.x.a.np[0][n];
    P.u = sm.x.a.np[1][n];
    P.v = sm.x.a.np[2][n];
    P.p = sm.x.a.np[3][n];
    P.beta = sm.x.a.np[4][n];
    P.E = sm.x.a.np[5][n];
    P.wsx = sm.x.a.np[6][n];
    P.wsy = sm.x.a.np[7][n];
    return P;
}
template <class S>
__device__ __forceinline__ void st_np(S& sm, int n, const NodeP& P) {
    sm.x.a.np[0][n] = P.rho;
    sm.x.a.np[1][n] = P.u;
    sm.x.a.np[2][n] = P.v;
    sm.x.a.np[3][n] = P.p;
    sm.x.a.np[4][n] = P.beta;
    sm.x.a.np[5][n] = P.E;
    sm.x.a.np[6][n] = P.wsx;
    sm.x.a.np[7][n] = P.wsy;
}
template <class S>
__device__ __forceinline__ float ws_dir(const S& sm, int n, int d) {
    return sm.x.a.np[6 + d][n];
}

template <int N1D, int E>
struct SmemDt {
    static constexpr int NQ = N1D * N1D, M = E * NQ;
    float ws[2][M];
    double geo[E][3];
    double red[16];
};

template <int N1D, int E, int NT>
__device__ __forceinline__ void stage_tile(const MixArgs& a, const MixInst& I, const MixTab<N1D>& T, int k0,
                                           int nel, SmemM<N1D, E>& sm) {
    using Sm = SmemM<N1D, E>;
    constexpr int NQ = Sm::NQ, NFP = Sm::NFP, NPV = Sm::NPV, NPL = Sm::NPL, NFB = Sm::NFB;
    constexpr int LW = N1D - 1;
    const int tid = threadIdx.x;
    const double gd = I.gamma;
    const float ig2m1 = (float)(0.5 / (gd - 1.0));
    const double* Uin = a.Uin + I.node_off * 4;
    StepScalars* sc = I.sc;
    const int mN = nel * NQ;      // valid nodes of the tile
    // the step's dt and the limiter's dt (stage 1 limits with the input dt min(CFL*dt0, T - t),
    // not the CFL one) are fixed during the launch: read them before the first barrier
    const double dt = bits_to_d(sc->dt_bits);
    const double dt_lim = (a.mode == MODE_STAGE2 || a.mode == MODE_STAGE3) ? dt : sc->dt_lim;
    const double zeta = I.zeta;

    // ---- 0: element states (one node per lane, 32 contiguous bytes) and metric -> LDS.  The
    //      face table entries of the lane's first face node are loaded here too: phase 2 then
    //      waits for one load (the neighbour's state) instead of two dependent ones.
    int f_bc = 0, f_nb = 0;
    if (tid < nel * NFP) {
        const size_t fi = (size_t)k0 * NFP + tid;
        f_bc = I.bc[fi];
        f_nb = I.nb_node[fi];
    }
    {
        const double4* src = (const double4*)(Uin + (size_t)k0 * NQ * 4);
#pragma unroll 1
        for (int n = tid; n < mN; n += NT) {
            const double4 v = src[n];
            sm.U[0][n] = v.x;
            sm.U[1][n] = v.y;
            sm.U[2][n] = v.z;
            sm.U[3][n] = v.w;
            // ---- 1: node primitives (FP64 energy differences -> FP32), from the lane's own
            //      node: no barrier between the two phases
            const double U[4] = {v.x, v.y, v.z, v.w};
            st_np(sm, n, node_prims(gd, U));
        }
        for (int w = tid; w < nel; w += NT) {
            const Geo G = I.geo[k0 + w];
            sm.geo[w][0] = G.J;
            sm.geo[w][1] = G.rxJ;
            sm.geo[w][2] = G.syJ;
            sm.geo[w][3] = 1.0 / G.J;
        }
    }
    __syncthreads();

    // ---- 2: surface fluxes at the face nodes (flux_differencing.jl:90-151, 223-272;
    //      low_order_graph_viscosity.jl:51-124, 175-204)
#pragma unroll 1
    for (int w = tid; w < nel * NFP; w += NT) {
        const int e = w / NFP, f = w - e * NFP;
        const int k = k0 + e;
        const int ld = f < 2 * N1D ? 0 : 1;
        const int nq = e * NQ + T.fq[f];
        const NodeP P = ld_np(sm, nq);
        const size_t fi = (size_t)k * NFP + f;
        const int bcc = w == tid ? f_bc : I.bc[fi];
        const int nb = w == tid ? f_nb : I.nb_node[fi];
        const double4 up = ((const double4*)Uin)[nb];
        const double UP[4] = {up.x, up.y, up.z, up.w};
        const NodeP PP = node_prims(gd, UP);
        // |B| from both elements' metric, so that the two sides of an interface compute exactly
        // negated fluxes (the per-element metrics agree to FP64 roundoff, which FP32 rounding
        // could otherwise turn into a one-ulp mismatch and a conservation defect)
        const Geo& GP = I.geo[nb / NQ];
        const double gsym = 0.5 * (sm.geo[e][1 + ld] + (ld == 0 ? GP.rxJ : GP.syJ));
        const float Bmag = (float)gsym * fabsf(T.Bf[f]);
        const float Bf = T.Bf[f] < 0.f ? -Bmag : Bmag;
        const float LF = 0.5f * Bmag * fmaxf(ld == 0 ? P.wsx : P.wsy, ld == 0 ? PP.wsx : PP.wsy);
        float F[4], FP[4];
        flux_dir(P, ld, F);
        double Uq[4];
        for (int c = 0; c < 4; c++) Uq[c] = sm.U[c][nq];
        if (bcc == 0) {
            flux_dir(PP, ld, FP);
            for (int c = 0; c < 4; c++) {
                const float v = fmaf(Bf, 0.5f * (F[c] + FP[c]), -LF * (float)(UP[c] - Uq[c]));
                sm.BF[c][w] = v;
                sm.BFL[c][w] = v;
            }
        } else if (bcc < 0) {   // outflow: uP = own node, no dissipation
            for (int c = 0; c < 4; c++) {
                const float v = Bf * F[c];
                sm.BF[c][w] = v;
                sm.BFL[c][w] = v;
            }
        } else {                // inflow: uP = Ival; the low-order flux keeps its dissipation
            double UI[4];
            for (int c = 0; c < 4; c++) UI[c] = I.Ival[(size_t)(bcc - 1) * 4 + c];
            const NodeP PI = node_prims(gd, UI);
            flux_dir(PI, ld, FP);
            for (int c = 0; c < 4; c++) {
                const float v = Bf * (0.5f * (F[c] + FP[c]));
                sm.BF[c][w] = v;
                sm.BFL[c][w] = v - LF * (float)(UI[c] - Uq[c]);
            }
        }
    }

    // ---- 3a: high-order volume pairs (flux_differencing_volume!, flux_differencing.jl:164-211)
#pragma unroll 1
    for (int w = tid; w < nel * NPV; w += NT) {
        const int e = w / NPV, p = w - e * NPV;
        const uint32_t pk = T.pH[p];
        const int d = (int)(pk >> 16);
        const int i = e * NQ + (int)(pk & 255u), j = e * NQ + (int)((pk >> 8) & 255u);
        const float Sd = (float)sm.geo[e][1 + d] * T.cH[p];
        float F[4];
        fS_f(ig2m1, ld_np(sm, i), ld_np(sm, j), d, F);
        for (int c = 0; c < 4; c++) sm.x.a.PH[c][w] = Sd * F[c];
    }
    // ---- 3b: low-order line pairs (accumulate_low_order_rhs_volume!, :139-173)
#pragma unroll 1
    for (int w = tid; w < nel * NPL; w += NT) {
        const int e = w / NPL, p = w - e * NPL;
        const uint32_t pk = T.pL[p];
        const int d = (int)(pk >> 16);
        const int i = e * NQ + (int)(pk & 255u), j = e * NQ + (int)((pk >> 8) & 255u);
        const float S0 = (float)sm.geo[e][1 + d] * T.cL[p];
        const NodeP Pi = ld_np(sm, i), Pj = ld_np(sm, j);
        const float lam = fabsf(S0) * fmaxf(d == 0 ? Pi.wsx : Pi.wsy, d == 0 ? Pj.wsx : Pj.wsy);
        float Fi[4], Fj[4];
        flux_dir(Pi, d, Fi);
        flux_dir(Pj, d, Fj);
        for (int c = 0; c < 4; c++)
            sm.x.a.PL[c][w] = fmaf(S0, Fi[c] + Fj[c], -lam * (float)(sm.U[c][j] - sm.U[c][i]));
    }
    __syncthreads();

    // ---- 4/5a: element assembly along lines (assemble_rhs!, flux_differencing.jl:331-361; the
    //      low-order line and surface terms, low_order_graph_viscosity.jl:175-220) fused with the
    //      subcell flux potentials (accumulate_f_bar!, subcell.jl:163-206).  One work item per
    //      (line, component): node q of a line in direction d contributes
    //      wq J rhsxyH[d] = -wq hx and wq J rhsxyL[d] = lx, where hx, lx are its assembled sums
    //      in that direction (rhsxyH = -hx / J, rhsxyL = lx / (wq J)); the increments are
    //      summed in FP64 and the potentials stored rounded once.  Items are numbered over the
    //      tile's capacity E, so that the component index is a compile-time division; the lines
    //      of absent elements are skipped.
    constexpr int NLN = E * 2 * N1D;
#pragma unroll 1
    for (int w = tid; w < 4 * NLN; w += NT) {
        const int c = w / NLN, wl = w - c * NLN;
        const int e = wl / (2 * N1D), L = wl - e * (2 * N1D);
        if (e >= nel) continue;
        const int d = L / N1D, line = L - d * N1D;
        const int fe = e * NFP;
        const int f0 = fe + (d == 0 ? line : 2 * N1D + line);
        const int f1 = fe + (d == 0 ? N1D + line : 3 * N1D + line);
        const int base = e * NFB + L * (N1D + 1);
        // the line's volume pairs: pair (a > b) of positions a, b has index a(a-1)/2 + b
        const float* PHl = sm.x.a.PH[c] + e * NPV + (d * N1D + line) * MixTab<N1D>::NP1;
        const float* PLc = sm.x.a.PL[c] + e * NPL + (d * N1D + line) * LW;
        double fh = (double)sm.BF[c][f0], fl = (double)sm.BFL[c][f0];
        sm.p.fd[c][base] = make_float2((float)fl, (float)(fh - fl));
#pragma unroll
        for (int s = 0; s < N1D - 1; s++) {   // the far node's increment is -BF of the far face
            const int q = T.lq[L][s];
            // node s adds the pairs it is the upper node of and subtracts the others (S is
            // skew-symmetric, fS symmetric), over the other positions in increasing order
            float Q = 0.f;
#pragma unroll
            for (int o = 0; o < N1D; o++) {
                if (o < s) Q = Q + PHl[s * (s - 1) / 2 + o];
                if (o > s) Q = Q - PHl[o * (o - 1) / 2 + s];
            }
            float h = T.a_v[q] * Q;
            float l = 0.f;
            if (s > 0) l -= PLc[s - 1];
            l += PLc[s];
            if (s == 0) {   // the near face
                h = fmaf(d == 0 ? T.a_fx[q] : T.a_fy[q], sm.BF[c][f0], h);
                l -= sm.BFL[c][f0];
            }
            fh = fma(-T.wq[q], (double)h, fh);
            fl += (double)l;
            sm.p.fd[c][base + s + 1] = make_float2((float)fl, (float)(fh - fl));
        }
        // far end: -BF exactly (FP32 values), so interfaces conserve
        sm.p.fd[c][base + N1D] = make_float2(-sm.BFL[c][f1], (float)((double)sm.BFL[c][f1] - (double)sm.BF[c][f1]));
    }
    // debug outputs rhsH, rhsL per node (the sums of both directions, as rhs! stores them)
#pragma unroll 1
    for (int n = a.dbg_rhsH ? tid : mN; n < mN; n += NT) {
        const int e = n / NQ, i = n - e * NQ;
        const int ia = i % N1D, ib = i / N1D;
        const int pv = e * NPV, pl = e * NPL, fe = e * NFP;
        float Qx[4] = {0.f, 0.f, 0.f, 0.f}, Qy[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int m = 0; m < N1D - 1; m++) {
            const uint32_t gxm = T.gx[i][m], gym = T.gy[i][m];
            const float sx = (gxm & 0x8000u) ? -1.f : 1.f, sy = (gym & 0x8000u) ? -1.f : 1.f;
            const int px = pv + (int)(gxm & 0x7fffu), py = pv + (int)(gym & 0x7fffu);
            for (int c = 0; c < 4; c++) {
                Qx[c] = fmaf(sx, sm.x.a.PH[c][px], Qx[c]);
                Qy[c] = fmaf(sy, sm.x.a.PH[c][py], Qy[c]);
            }
        }
        const float av = T.a_v[i];
        float hx[4], hy[4];
        for (int c = 0; c < 4; c++) {
            hx[c] = av * Qx[c];
            hy[c] = av * Qy[c];
        }
        // low order: -Q0F1 of the line neighbours
        float lx[4] = {0.f, 0.f, 0.f, 0.f}, ly[4] = {0.f, 0.f, 0.f, 0.f};
        if (ia > 0)
            for (int c = 0; c < 4; c++) lx[c] -= sm.x.a.PL[c][pl + ib * LW + ia - 1];
        if (ia < N1D - 1)
            for (int c = 0; c < 4; c++) lx[c] += sm.x.a.PL[c][pl + ib * LW + ia];
        if (ib > 0)
            for (int c = 0; c < 4; c++) ly[c] -= sm.x.a.PL[c][pl + N1D * LW + ia * LW + ib - 1];
        if (ib < N1D - 1)
            for (int c = 0; c < 4; c++) ly[c] += sm.x.a.PL[c][pl + N1D * LW + ia * LW + ib];
        // face terms (at most one r-face and one s-face per node)
        const int fx = ia == 0 ? ib : ia == N1D - 1 ? N1D + ib : -1;
        const int fy = ib == 0 ? 2 * N1D + ia : ib == N1D - 1 ? 3 * N1D + ia : -1;
        if (fx >= 0) {
            const float af = T.a_fx[i];
            for (int c = 0; c < 4; c++) {
                hx[c] = fmaf(af, sm.BF[c][fe + fx], hx[c]);
                lx[c] -= sm.BFL[c][fe + fx];
            }
        }
        if (fy >= 0) {
            const float af = T.a_fy[i];
            for (int c = 0; c < 4; c++) {
                hy[c] = fmaf(af, sm.BF[c][fe + fy], hy[c]);
                ly[c] -= sm.BFL[c][fe + fy];
            }
        }
        const float nJ = -(float)sm.geo[e][3];
        const float iwJ = (float)(sm.geo[e][3] * rcp64(T.wq[i]));
        const size_t o = ((size_t)I.node_off + (size_t)k0 * NQ + n) * 4;
        for (int c = 0; c < 4; c++) {
            a.dbg_rhsH[o + c] = (double)(hx[c] * nJ) + (double)(hy[c] * nJ);
            a.dbg_rhsL[o + c] = (double)(lx[c] * iwJ) + (double)(ly[c] * iwJ);
        }
    }

    __syncthreads();

    // ---- 5u: u_L from the low-order potentials and the bounds (subcell.jl:269, 370-387)
#pragma unroll 1
    for (int n = tid; n < mN; n += NT) {
        const int e = n / NQ, i = n - e * NQ;
        const int rx = e * NFB + (T.nfc[i] & 255), ry = e * NFB + (T.nfc[i] >> 8);
        const double iwJ = sm.geo[e][3] * rcp64(T.wq[i]);
        sm.x.b.iwJ[n] = iwJ;
        double uL[4];
        for (int c = 0; c < 4; c++) {
            const double r = ((double)sm.p.fd[c][rx + 1].x - (double)sm.p.fd[c][rx].x) +
                             ((double)sm.p.fd[c][ry + 1].x - (double)sm.p.fd[c][ry].x);
            uL[c] = sm.U[c][n] + dt_lim * (r * iwJ);
            sm.x.b.UL[c][n] = uL[c];
        }
        sm.x.b.Lrhoe[n] = zeta * (uL[3] - 0.5 * (uL[1] * uL[1] + uL[2] * uL[2]) * rcp64(uL[0]));
    }
    __syncthreads();

    // ---- 5b: limiter parameter of every subcell face (subcell_bound_limiter!, subcell.jl:248-349):
    //      one work item per face, the minimum of the parameters of its two adjacent nodes (side 0:
    //      the node at position s of the line, above/right of face s, P = -4 dt df / (wq J); side
    //      1: the node at position s - 1, P = +4 dt df / (wq J)).  The faces inside elements
    //      come first, then the element-boundary faces, which have one node in the element and
    //      whose f_bar_H - f_bar_L is zero except at inflow boundaries (a zero difference makes
    //      the parameter irrelevant: 1).  The side whose density increment is not negative is
    //      evaluated first, and the other is skipped once the minimum is 0.
    const double fac_dt = 4.0 * dt_lim;
    constexpr int NFI = 2 * N1D * (N1D - 1);   // inner subcell faces per element
    constexpr int NFO = 4 * N1D;               // element-boundary subcell faces per element
#pragma unroll 1
    for (int w = tid; w < nel * (NFI + NFO); w += NT) {
        int e, L, f;
        if (w < nel * NFI) {
            e = w / NFI;
            const int q = w - e * NFI;
            L = q / (N1D - 1);
            f = q - L * (N1D - 1) + 1;
        } else {
            const int jb = w - nel * NFI;
            e = jb / NFO;
            const int q = jb - e * NFO;
            L = q >> 1;
            f = (q & 1) ? N1D : 0;
        }
        const int r = e * NFB + L * (N1D + 1) + f;
        double fLr[4], dfr[4];
        for (int c = 0; c < 4; c++) {
            const float2 v = sm.p.fd[c][r];
            fLr[c] = (double)v.x;
            dfr[c] = (double)v.y;
        }
        double l = 1.0;
        if (dfr[0] != 0.0 || dfr[1] != 0.0 || dfr[2] != 0.0 || dfr[3] != 0.0) {
            // P0 of side 0 is -df0 times a positive factor
            const int first = dfr[0] > 0.0 ? 1 : 0;
#pragma unroll 1
            for (int k = 0; k < 2; k++) {
                const int side = k == 0 ? first : 1 - first;
                const int pos = f - side;            // the node's position on the line
                if (pos < 0 || pos >= N1D) continue;  // boundary face: no node on that side
                const int n = e * NQ + T.lq[L][pos];
                const double sc_ = (side ? fac_dt : -fac_dt) * sm.x.b.iwJ[n];
                double P[4], uL[4];
                for (int c = 0; c < 4; c++) {
                    P[c] = sc_ * dfr[c];
                    uL[c] = sm.x.b.UL[c][n];
                }
                l = fmin(l, lim_param(uL, zeta * uL[0], sm.x.b.Lrhoe[n], P, I.ZEROTOL, I.upper_mode));
                if (!(l > 0.0)) break;
            }
        }
        // rounded toward zero (a smaller parameter stays admissible); the limited flux of the
        // face replaces its potentials
        const double lr = (double)__double2float_rz(l);
        for (int c = 0; c < 4; c++) sm.p.g[c][r] = fma(lr, dfr[c], fLr[c]);
    }
    __syncthreads();

    // ---- 5c: limited RHS (accumulate_f_bar_limited!, apply_subcell_limiter!, subcell.jl:826-924),
    //      RK stage (SSPRK33.jl:30-39) and positivity of the new state
    const size_t ob = ((size_t)I.node_off + (size_t)k0 * NQ) * 4;
#pragma unroll 1
    for (int n = tid; n < mN; n += NT) {
        const int e = n / NQ, i = n - e * NQ;
        const int rx = e * NFB + (T.nfc[i] & 255), ry = e * NFB + (T.nfc[i] >> 8);
        const double iwJ = sm.x.b.iwJ[n];
        double r[4];
        for (int c = 0; c < 4; c++) {
            const double gx0 = sm.p.g[c][rx], gx1 = sm.p.g[c][rx + 1];
            const double gy0 = sm.p.g[c][ry], gy1 = sm.p.g[c][ry + 1];
            r[c] = ((gx1 - gx0) + (gy1 - gy0)) * iwJ;
        }
        const size_t o = ob + (size_t)n * 4;
        if (a.mode == MODE_RHS) {
            *(double4*)(a.rhsU + o) = make_double4(r[0], r[1], r[2], r[3]);
            continue;
        }
        double un[4];
        if (a.mode == MODE_STAGE1) {
            for (int c = 0; c < 4; c++) un[c] = fma(dt, r[c], sm.U[c][n]);
        } else {
            const double wa = a.mode == MODE_STAGE2 ? 3.0 / 4.0 : 1.0 / 3.0;
            const double wb = a.mode == MODE_STAGE2 ? 1.0 / 4.0 : 2.0 / 3.0;
            const double4 W = *(const double4*)(a.W + o);
            const double Wc[4] = {W.x, W.y, W.z, W.w};
            for (int c = 0; c < 4; c++) un[c] = fma(wa, Wc[c], wb * fma(dt, r[c], sm.U[c][n]));
        }
        *(double4*)(a.Uout + o) = make_double4(un[0], un[1], un[2], un[3]);
        // the sign of the pressure: a reciprocal within an ulp is enough
        const double pr = (gd - 1.0) * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]) * rcp64(un[0]));
        if (!(un[0] > 0.0) || !(pr > 0.0)) {
            atomicAdd(&sc->viol, 1ull);
            atomicMin(&sc->first_bad, ((unsigned long long)a.mode << 40) | (unsigned long long)((size_t)k0 * NQ + n));
        }
    }
}

// CFL time step of the low-order scheme (calculate_lambda_and_low_order_CFL!,
// low_order_graph_viscosity.jl:222-285): per node CFL*wq*J/2 / (sum_j lambda_ij + sum_f lambda_B).
template <int N1D, int E, int NT>
__device__ __forceinline__ void dt_tile(const MixArgs& a, const MixInst& I, const MixTab<N1D>& T, int tile, int k0,
                                        int nel, SmemDt<N1D, E>& sm) {
    constexpr int NQ = N1D * N1D, NFP = 4 * N1D, LW = N1D - 1;
    const int tid = threadIdx.x;
    const double gd = I.gamma;
    const double* Uin = a.Uin + I.node_off * 4;
#pragma unroll 1
    for (int w = tid; w < E * NQ; w += NT) {
        const int e = w / NQ, i = w - e * NQ;
        if (e >= nel) continue;
        const double4 u = ((const double4*)Uin)[(size_t)(k0 + e) * NQ + i];
        const double U[4] = {u.x, u.y, u.z, u.w};
        const NodeP P = node_prims(gd, U);
        sm.ws[0][w] = P.wsx;
        sm.ws[1][w] = P.wsy;
    }
    for (int w = tid; w < nel; w += NT) {
        const Geo G = I.geo[k0 + w];
        sm.geo[w][0] = G.J;
        sm.geo[w][1] = G.rxJ;
        sm.geo[w][2] = G.syJ;
    }
    __syncthreads();
    double dtmin = INFINITY;
#pragma unroll 1
    for (int w = tid; w < E * NQ; w += NT) {
        const int e = w / NQ, i = w - e * NQ;
        if (e >= nel) continue;
        const int k = k0 + e;
        const int ia = i % N1D, ib = i / N1D;
        const float rx = (float)sm.geo[e][1], sy = (float)sm.geo[e][2];
        const float wx = sm.ws[0][w], wy = sm.ws[1][w];
        float lam = 0.f;
        if (ia > 0) lam += fabsf(rx * T.cL[ib * LW + ia - 1]) * fmaxf(wx, sm.ws[0][w - 1]);
        if (ia < N1D - 1) lam += fabsf(rx * T.cL[ib * LW + ia]) * fmaxf(wx, sm.ws[0][w + 1]);
        if (ib > 0) lam += fabsf(sy * T.cL[N1D * LW + ia * LW + ib - 1]) * fmaxf(wy, sm.ws[1][w - N1D]);
        if (ib < N1D - 1) lam += fabsf(sy * T.cL[N1D * LW + ia * LW + ib]) * fmaxf(wy, sm.ws[1][w + N1D]);
        const int fx = ia == 0 ? ib : ia == N1D - 1 ? N1D + ib : -1;
        const int fy = ib == 0 ? 2 * N1D + ia : ib == N1D - 1 ? 3 * N1D + ia : -1;
        for (int m = 0; m < 2; m++) {
            const int f = m == 0 ? fx : fy;
            if (f < 0) continue;
            const int nb = I.nb_node[(size_t)k * NFP + f];
            double UP[4];
            for (int c = 0; c < 4; c++) UP[c] = Uin[(size_t)nb * 4 + c];
            const NodeP PP = node_prims(gd, UP);
            const float Bf = (m == 0 ? rx : sy) * T.Bf[f];
            lam += 0.5f * fabsf(Bf) * fmaxf(m == 0 ? wx : wy, m == 0 ? PP.wsx : PP.wsy);
        }
        const double dti = I.CFL * 0.5 * T.wq[i] * sm.geo[e][0] * rcp64((double)lam);
        dtmin = dti < dtmin ? dti : dtmin;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(dtmin, off, 64);
        dtmin = o < dtmin ? o : dtmin;
    }
    if ((tid & 63) == 0) sm.red[tid >> 6] = dtmin;
    __syncthreads();
    if (tid == 0) {
        double m = sm.red[0];
        for (int x = 1; x < NT / 64; x++) m = sm.red[x] < m ? sm.red[x] : m;
        a.dtp[tile] = m;
    }
}

template <int A, int B>
struct cmax {
    static constexpr int v = A > B ? A : B;
};

template <int N1D, int E>
constexpr int stage_bytes() {
    return E > 0 ? (int)sizeof(SmemM<N1D, (E > 0 ? E : 1)>) : 0;
}
template <int N1D, int E>
constexpr int dt_bytes() {
    return E > 0 ? (int)sizeof(SmemDt<N1D, (E > 0 ? E : 1)>) : 0;
}
template <int N1D, int E>
constexpr int tab_bytes() {
    return E > 0 ? (int)sizeof(MixTab<N1D>) : 0;
}

// LDS offsets of the tables of the compiled degrees and the size of the tile scratch.
template <int E2, int E3, int E4, int E5>
struct Layout {
    static constexpr int T2 = 0;
    static constexpr int T3 = T2 + (tab_bytes<2, E2>() + 15) / 16 * 16;
    static constexpr int T4 = T3 + (tab_bytes<3, E3>() + 15) / 16 * 16;
    static constexpr int T5 = T4 + (tab_bytes<4, E4>() + 15) / 16 * 16;
    static constexpr int TEND = T5 + (tab_bytes<5, E5>() + 15) / 16 * 16;
    static constexpr int SMAX = cmax<cmax<stage_bytes<2, E2>(), stage_bytes<3, E3>()>::v,
                                     cmax<stage_bytes<4, E4>(), stage_bytes<5, E5>()>::v>::v;
    static constexpr int DMAX = cmax<cmax<dt_bytes<2, E2>(), dt_bytes<3, E3>()>::v,
                                     cmax<dt_bytes<4, E4>(), dt_bytes<5, E5>()>::v>::v;
};

template <int N1D, int EN, int NT>
__device__ __forceinline__ void load_tab(const MixArgs& a, unsigned char* dst) {
    // degrees compiled into the variant but absent from the batch have no table
    if (EN == 0 || a.tab[N1D] == nullptr) return;
    const int n = (int)sizeof(MixTab<N1D>) / 4;
    const unsigned* src = (const unsigned*)a.tab[N1D];
    unsigned* d = (unsigned*)dst;
    for (int w = threadIdx.x; w < n; w += NT) d[w] = src[w];
}

// This workgroup's tiles: t0, t0 + step, ... < hi within the chunk of its XCD.
__device__ __forceinline__ void tile_range(const MixArgs& a, int& t0, int& hi, int& step) {
    const int G = gridDim.x, b = blockIdx.x;
    if (G < MIX_CHUNKS) {
        t0 = b;
        hi = a.ntiles;
        step = G;
        return;
    }
    const int c = b % MIX_CHUNKS, r = b / MIX_CHUNKS;
    t0 = a.chunk[c] + r;
    hi = a.chunk[c + 1];
    step = G / MIX_CHUNKS + (c < G % MIX_CHUNKS ? 1 : 0);
}

// Kernels compiled for several degrees ask for 4 waves per SIMD: the union of the inlined
// tile functions would otherwise take more registers than any one of them (137-169 VGPRs
// against at most 112), and calling them out of line passes the uniform arguments in VGPRs.
template <int E2, int E3, int E4, int E5>
constexpr int min_waves() {
    return (E2 > 0) + (E3 > 0) + (E4 > 0) + (E5 > 0) > 1 ? 4 : 1;
}

template <int NT, int E2, int E3, int E4, int E5>
__global__ __launch_bounds__(NT, (min_waves<E2, E3, E4, E5>())) void mix_stage_kernel(MixArgs a) {
    using LY = Layout<E2, E3, E4, E5>;
    __shared__ __align__(16) unsigned char tabs[LY::TEND > 0 ? LY::TEND : 16];
    __shared__ __align__(16) unsigned char raw[LY::SMAX];
    load_tab<2, E2, NT>(a, tabs + LY::T2);
    load_tab<3, E3, NT>(a, tabs + LY::T3);
    load_tab<4, E4, NT>(a, tabs + LY::T4);
    load_tab<5, E5, NT>(a, tabs + LY::T5);
    __syncthreads();
    int t0, hi, step;
    tile_range(a, t0, hi, step);
    for (int t = t0; t < hi; t += step) {
        const MixTile tl = a.tiles[t];
        if (a.only >= 0 && tl.inst != a.only) continue;
        const MixInst& I = a.inst[tl.inst];
        if (I.sc->done) continue;
        switch (tl.n1d) {
            case 2:
                if constexpr (E2 > 0)
                {
                    auto& sm = *(SmemM<2, (E2 > 0 ? E2 : 1)>*)raw;
                    const auto& T = *(const MixTab<2>*)(tabs + LY::T2);
                    stage_tile<2, E2, NT>(a, I, T, tl.k0, tl.nel, sm);
                }
                break;
            case 3:
                if constexpr (E3 > 0)
                {
                    auto& sm = *(SmemM<3, (E3 > 0 ? E3 : 1)>*)raw;
                    const auto& T = *(const MixTab<3>*)(tabs + LY::T3);
                    stage_tile<3, E3, NT>(a, I, T, tl.k0, tl.nel, sm);
                }
                break;
            case 4:
                if constexpr (E4 > 0)
                {
                    auto& sm = *(SmemM<4, (E4 > 0 ? E4 : 1)>*)raw;
                    const auto& T = *(const MixTab<4>*)(tabs + LY::T4);
                    stage_tile<4, E4, NT>(a, I, T, tl.k0, tl.nel, sm);
                }
                break;
            case 5:
                if constexpr (E5 > 0)
                {
                    auto& sm = *(SmemM<5, (E5 > 0 ? E5 : 1)>*)raw;
                    const auto& T = *(const MixTab<5>*)(tabs + LY::T5);
                    stage_tile<5, E5, NT>(a, I, T, tl.k0, tl.nel, sm);
                }
                break;
        }
        __syncthreads();   // the next tile reuses the LDS
    }
}

template <int NT, int E2, int E3, int E4, int E5>
__global__ __launch_bounds__(NT) void mix_dt_kernel(MixArgs a) {
    using LY = Layout<E2, E3, E4, E5>;
    __shared__ __align__(16) unsigned char tabs[LY::TEND > 0 ? LY::TEND : 16];
    __shared__ __align__(16) unsigned char raw[LY::DMAX];
    load_tab<2, E2, NT>(a, tabs + LY::T2);
    load_tab<3, E3, NT>(a, tabs + LY::T3);
    load_tab<4, E4, NT>(a, tabs + LY::T4);
    load_tab<5, E5, NT>(a, tabs + LY::T5);
    __syncthreads();
    int t0, hi, step;
    tile_range(a, t0, hi, step);
    for (int t = t0; t < hi; t += step) {
        const MixTile tl = a.tiles[t];
        if (a.only >= 0 && tl.inst != a.only) continue;
        const MixInst& I = a.inst[tl.inst];
        if (I.sc->done) continue;
        switch (tl.n1d) {
            case 2:
                if constexpr (E2 > 0)
                    dt_tile<2, E2, NT>(a, I, *(const MixTab<2>*)(tabs + LY::T2), t, tl.k0, tl.nel,
                                       *(SmemDt<2, (E2 > 0 ? E2 : 1)>*)raw);
                break;
            case 3:
                if constexpr (E3 > 0)
                    dt_tile<3, E3, NT>(a, I, *(const MixTab<3>*)(tabs + LY::T3), t, tl.k0, tl.nel,
                                       *(SmemDt<3, (E3 > 0 ? E3 : 1)>*)raw);
                break;
            case 4:
                if constexpr (E4 > 0)
                    dt_tile<4, E4, NT>(a, I, *(const MixTab<4>*)(tabs + LY::T4), t, tl.k0, tl.nel,
                                       *(SmemDt<4, (E4 > 0 ? E4 : 1)>*)raw);
                break;
            case 5:
                if constexpr (E5 > 0)
                    dt_tile<5, E5, NT>(a, I, *(const MixTab<5>*)(tabs + LY::T5), t, tl.k0, tl.nel,
                                       *(SmemDt<5, (E5 > 0 ? E5 : 1)>*)raw);
                break;
        }
        __syncthreads();   // the reduction slots are reused by the next tile
    }
}

// per instance: step start (SSPRK33.jl:30) or RHS-evaluation set-up
// SSPRK33.jl:28-30: done once t >= T, else the step's dt cap min(CFL*dt0, T - t)
__device__ __forceinline__ void step_start(const MixInst& I, StepScalars* sc) {
    const double t = sc->t;
    if (!(t < I.T)) {
        sc->done = 1;
        return;
    }
    sc->done = 0;
    const double dt = jmin(I.CFL * I.dt0, I.T - t);
    sc->dt_lim = dt;
    sc->dt_bits = d_to_bits(dt);
}

__global__ void mix_init_kernel(const MixInst* inst, int n, int mode, int only, double dt_lim, double dt_cap) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    StepScalars* sc = inst[g].sc;
    if (mode == 0) {
        step_start(inst[g], sc);
    } else if (g == only) {
        sc->done = 0;
        sc->dt_lim = dt_lim;
        sc->dt_bits = d_to_bits(dt_cap);
    }
}

// dt = min(cap, min over the instance's tiles) (the global minimum of
// calculate_lambda_and_low_order_CFL!, low_order_graph_viscosity.jl:230-243); one workgroup
// per instance
__global__ __launch_bounds__(256) void mix_dtred_kernel(const MixInst* inst, int n, const double* dtp, int only) {
    __shared__ double red[4];
    const int i = blockIdx.x;
    if (i >= n || (only >= 0 && i != only)) return;
    const MixInst& I = inst[i];
    if (I.sc->done) return;
    // four independent chains, so that the loads of a thread are in flight together (a NaN
    // minimum is skipped in every chain, as the single chain skipped it)
    double m0 = INFINITY, m1 = INFINITY, m2 = INFINITY, m3 = INFINITY;
    int t = I.tile_lo + (int)threadIdx.x;
    for (; t + 768 < I.tile_hi; t += 1024) {
        const double a = dtp[t], b = dtp[t + 256], c = dtp[t + 512], d = dtp[t + 768];
        m0 = a < m0 ? a : m0;
        m1 = b < m1 ? b : m1;
        m2 = c < m2 ? c : m2;
        m3 = d < m3 ? d : m3;
    }
    for (; t < I.tile_hi; t += 256) m0 = dtp[t] < m0 ? dtp[t] : m0;
    m1 = m1 < m0 ? m1 : m0;
    m3 = m3 < m2 ? m3 : m2;
    double m = m3 < m1 ? m3 : m1;
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o < m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int x = 1; x < 4; x++) m = red[x] < m ? red[x] : m;
        m = red[0] < m ? red[0] : m;
        const double cap = bits_to_d(I.sc->dt_bits);
        I.sc->dt_bits = d_to_bits(m < cap ? m : cap);
    }
}

// the end of a step (t += dt, SSPRK33.jl:42) and the start of the next one, in one launch
__global__ void mix_post_kernel(const MixInst* inst, int n) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    StepScalars* sc = inst[g].sc;
    if (!sc->done) {
        const double dt = bits_to_d(sc->dt_bits);
        sc->t = sc->t + dt;
        sc->steps = sc->steps + 1;
        sc->last_dt = dt;
    }
    step_start(inst[g], sc);
}

__device__ __forceinline__ unsigned long long order_key_m(double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__global__ void mix_positivity_kernel(const double* U, int64_t nodes, double gamma, unsigned long long* keys,
                                      unsigned long long* cnt) {
    const size_t gidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double mr = INFINITY, mp = INFINITY;
    unsigned long long bad = 0;
    if (gidx < (size_t)nodes) {
        const double* u = U + gidx * 4;
        const double pr = (gamma - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
        mr = u[0];
        mp = pr;
        bad = (!(u[0] > 0.0) || !(pr > 0.0)) ? 1ull : 0ull;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o1 = __shfl_xor(mr, off, 64), o2 = __shfl_xor(mp, off, 64);
        mr = o1 < mr ? o1 : mr;
        mp = o2 < mp ? o2 : mp;
        bad += __shfl_xor(bad, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&keys[0], order_key_m(mr));
        atomicMin(&keys[1], order_key_m(mp));
        if (bad) atomicAdd(cnt, bad);
    }
}

// ------------------------------------------------------------------------------ variants
struct MixVariant {
    int NT, E[MIX_MAX_N1D + 1];
    const void* stage;
    const void* dt;
    int NTdt;            // the CFL pass has one work item per node: at most 128 threads
};

#define MV(NT, E2, E3, E4, E5)                                                                    \
    {NT, {0, 0, E2, E3, E4, E5}, (const void*)&mix_stage_kernel<NT, E2, E3, E4, E5>,             \
     (const void*)&mix_dt_kernel<(NT > 128 ? 128 : NT), E2, E3, E4, E5>, (NT > 128 ? 128 : NT)}

// single-degree shapes first (one degree per problem), then all-degree shapes for batches
const MixVariant kMix[] = {
    MV(128, 0, 0, 4, 0), MV(64, 0, 0, 4, 0), MV(128, 0, 0, 8, 0), MV(256, 0, 0, 8, 0), MV(64, 0, 0, 2, 0),
    MV(128, 0, 0, 0, 2), MV(64, 0, 0, 0, 2), MV(128, 0, 0, 0, 4), MV(256, 0, 0, 0, 4),
    MV(128, 0, 8, 0, 0), MV(64, 0, 4, 0, 0), MV(256, 0, 16, 0, 0), MV(256, 0, 12, 0, 0),
    MV(128, 16, 0, 0, 0), MV(64, 8, 0, 0, 0),
    MV(256, 0, 12, 8, 5), MV(256, 0, 12, 8, 4), MV(256, 12, 12, 8, 5), MV(256, 12, 12, 8, 4),
    MV(256, 0, 16, 8, 5), MV(256, 0, 10, 8, 5), MV(256, 0, 16, 8, 4), MV(256, 0, 12, 9, 5),
    // 512 threads, two workgroups per CU (up to 76 KB of LDS each): twice the elements per tile
    MV(512, 0, 32, 16, 10), MV(512, 0, 32, 16, 11),
    MV(512, 0, 0, 16, 0), MV(512, 0, 32, 0, 0), MV(512, 0, 0, 0, 10),
};
#undef MV
constexpr int kNumMix = (int)(sizeof(kMix) / sizeof(kMix[0]));

}  // namespace

int mix_num_variants() { return kNumMix; }

int mix_variant_shape(int v, int* NT, int E[MIX_MAX_N1D + 1]) {
    if (v < 0 || v >= kNumMix) return -1;
    *NT = kMix[v].NT;
    for (int n = 0; n <= MIX_MAX_N1D; n++) E[n] = kMix[v].E[n];
    return 0;
}

int mix_blocks_per_cu(int v, int* stage_blocks, int* dt_blocks) {
    if (v < 0 || v >= kNumMix) return (int)hipErrorInvalidValue;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(stage_blocks, kMix[v].stage, kMix[v].NT, 0);
    if (e != hipSuccess) return (int)e;
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(dt_blocks, kMix[v].dt, kMix[v].NTdt, 0);
}

int mix_launch_stage(int v, const MixArgs& a, int grid, void* stream) {
    if (v < 0 || v >= kNumMix || grid < 1) return (int)hipErrorInvalidValue;
    MixArgs arg = a;
    void* params[] = {&arg};
    return (int)hipLaunchKernel(kMix[v].stage, dim3(grid), dim3(kMix[v].NT), params, 0, (hipStream_t)stream);
}

int mix_launch_dt(int v, const MixArgs& a, int grid, void* stream) {
    if (v < 0 || v >= kNumMix || grid < 1) return (int)hipErrorInvalidValue;
    MixArgs arg = a;
    void* params[] = {&arg};
    return (int)hipLaunchKernel(kMix[v].dt, dim3(grid), dim3(kMix[v].NTdt), params, 0, (hipStream_t)stream);
}

int mix_launch_init(const MixInst* inst, int n, int mode, int only, double dt_lim, double dt_cap, void* stream) {
    hipLaunchKernelGGL(mix_init_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, inst, n, mode, only,
                       dt_lim, dt_cap);
    return (int)hipGetLastError();
}

int mix_launch_dt_reduce(const MixInst* inst, int n, const double* dtp, int only, void* stream) {
    hipLaunchKernelGGL(mix_dtred_kernel, dim3(n), dim3(256), 0, (hipStream_t)stream, inst, n, dtp, only);
    return (int)hipGetLastError();
}

int mix_launch_post(const MixInst* inst, int n, void* stream) {
    hipLaunchKernelGGL(mix_post_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, inst, n);
    return (int)hipGetLastError();
}

int mix_launch_positivity(const double* U, int64_t nodes, double gamma, unsigned long long* keys,
                          unsigned long long* cnt, void* stream) {
    const int blocks = (int)((nodes + 255) / 256);
    hipLaunchKernelGGL(mix_positivity_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, U, nodes, gamma,
                       keys, cnt);
    return (int)hipGetLastError();
}

}  // namespace p2de
