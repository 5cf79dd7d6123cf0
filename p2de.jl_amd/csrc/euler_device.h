// euler_device.h -- pointwise compressible-Euler physics for the device kernels.
//
// Each function follows the reference formula of src/math/compressible_Navier_Stokes.jl
// operation by operation (the library is built with -ffp-contract=off, so the FP64
// results round exactly like the CPU oracle's).  Min/max follow Julia's Base.min/max
// on IEEE floats: NaN-propagating and ordering -0 below +0.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace p2de {

__device__ __forceinline__ double jmin(double x, double y) {
    double diff = x - y;
    double am = signbit(diff) ? x : y;
    return (isnan(x) || isnan(y)) ? diff : am;
}
__device__ __forceinline__ double jmax(double x, double y) {
    double diff = x - y;
    double am = signbit(diff) ? y : x;
    return (isnan(x) || isnan(y)) ? diff : am;
}
// LinearAlgebra.generic_norm2 on a 2-tuple (used for |B| and |S0|, rhs_utils.jl:29-75)
__device__ __forceinline__ double jnorm2(double a, double b) {
    double maxabs = jmax(fabs(a), fabs(b));
    if (maxabs == 0.0 || isinf(maxabs) || isnan(maxabs)) return maxabs;
    if (isfinite(2.0 * maxabs * maxabs) && (maxabs * maxabs) != 0.0) {
        double s = a * a;
        s += b * b;
        return sqrt(s);
    }
    double ra = fabs(a) / maxabs, rb = fabs(b) / maxabs;
    double s = ra * ra;
    s += rb * rb;
    return maxabs * sqrt(s);
}

template <int DIM> struct Eq;

// ---------------------------------------------------------------------------- 1D
template <> struct Eq<1> {
    static constexpr int NC = 3;
    // pfun :18-22
    __device__ static __forceinline__ double p(double g, const double* U) {
        return (g - 1.0) * (U[2] - 0.5 * (U[1] * U[1]) / U[0]);
    }
    // rhoe_ufun :70-73
    __device__ static __forceinline__ double rhoe(const double* U) {
        return U[2] - 0.5 * U[1] * U[1] / U[0];
    }
    // v_ufun :113-122
    __device__ static __forceinline__ void v_of_u(double g, const double* U, double* V) {
        double pp = p(g, U);
        double s = log(pp / pow(U[0], g));
        V[0] = (g + 1.0 - s) - (g - 1.0) * U[2] / pp;
        V[1] = U[1] * (g - 1.0) / pp;
        V[2] = -U[0] * (g - 1.0) / pp;
    }
    // u_vfun :146-153 with rhoe_vfun :99-104 and s_vfun :87-91
    __device__ static __forceinline__ void u_of_v(double g, const double* V, double* U) {
        double s = g - V[0] + V[1] * V[1] / (2.0 * V[2]);
        double r = pow((g - 1.0) / pow(-V[2], g), 1.0 / (g - 1.0)) * exp(-s / (g - 1.0));
        U[0] = -r * V[2];
        U[1] = r * V[1];
        U[2] = r * (1.0 - V[1] * V[1] / (2.0 * V[2]));
    }
    // fluxes :165-173, F[d*NC + c]
    __device__ static __forceinline__ void flux(double g, const double* U, double* F) {
        double pp = p(g, U);
        double u = U[1] / U[0];
        F[0] = U[1];
        F[1] = U[1] * u + pp;
        F[2] = u * (U[2] + pp);
    }
    // wavespeed_estimate :48-56 (1D ignores the normal)
    __device__ static __forceinline__ double wavespeed(double g, const double* U, const double*) {
        double pp = p(g, U);
        return fabs(U[1] / U[0]) + sqrt(g * pp / U[0]);
    }
};

// ---------------------------------------------------------------------------- 2D
template <> struct Eq<2> {
    static constexpr int NC = 4;
    // pfun :24-28
    __device__ static __forceinline__ double p(double g, const double* U) {
        return (g - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0]);
    }
    // rhoe_ufun :75-78
    __device__ static __forceinline__ double rhoe(const double* U) {
        return U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0];
    }
    // v_ufun :134-144
    __device__ static __forceinline__ void v_of_u(double g, const double* U, double* V) {
        double pp = p(g, U);
        double s = log(pp / pow(U[0], g));
        V[0] = (g + 1.0 - s) - (g - 1.0) * U[3] / pp;
        V[1] = U[1] * (g - 1.0) / pp;
        V[2] = U[2] * (g - 1.0) / pp;
        V[3] = -U[0] * (g - 1.0) / pp;
    }
    // u_vfun :155-163 with rhoe_vfun :99-104 and s_vfun :93-97
    __device__ static __forceinline__ void u_of_v(double g, const double* V, double* U) {
        double s = g - V[0] + (V[1] * V[1] + V[2] * V[2]) / (2.0 * V[3]);
        double r = pow((g - 1.0) / pow(-V[3], g), 1.0 / (g - 1.0)) * exp(-s / (g - 1.0));
        U[0] = -r * V[3];
        U[1] = r * V[1];
        U[2] = r * V[2];
        U[3] = r * (1.0 - (V[1] * V[1] + V[2] * V[2]) / (2.0 * V[3]));
    }
    // fluxes :175-194, F[d*NC + c]
    __device__ static __forceinline__ void flux(double g, const double* U, double* F) {
        double pp = p(g, U);
        double u = U[1] / U[0];
        double v = U[2] / U[0];
        double rhouv = U[0] * u * v;
        double Ep = U[3] + pp;
        F[0] = U[1];
        F[1] = U[1] * u + pp;
        F[2] = rhouv;
        F[3] = u * Ep;
        F[4] = U[2];
        F[5] = rhouv;
        F[6] = U[2] * v + pp;
        F[7] = v * Ep;
    }
    // wavespeed_estimate :58-62: rotate to a 1D state along n, then the 1D formula
    __device__ static __forceinline__ double wavespeed(double g, const double* U, const double* n) {
        double m = n[0] * U[1] + n[1] * U[2];
        double U1[3] = {U[0], m, U[3]};
        return Eq<1>::wavespeed(g, U1, nullptr);
    }
};

// logmean :307-321 (series branch with the reference's coefficients)
__device__ __forceinline__ double logmean(double aL, double aR, double logL, double logR) {
    double da = aR - aL;
    double aavg = 0.5 * (aR + aL);
    double f = da / aavg;
    double v = f * f;
    if (fabs(f) < 1e-4) return aavg * (1.0 + v * (-0.2 - v * (0.0512 - v * 0.026038857142857)));
    return -da / (logL - logR);
}

// Primitive tuple of a hybrid node for the Chandrashekar flux: U() of
// flux_differencing.jl:179-191 = (rho, u, [v,] beta, log rho, log beta).
template <int DIM> struct Prim {
    double rho, u, v, beta, lrho, lbeta;
};

// fS :196-249, only the components of direction `dir` (0 = x, 1 = y); the other
// direction is multiplied by an exactly-zero metric term in the callers.
template <int DIM>
__device__ __forceinline__ void fS_dir(double g, const Prim<DIM>& L, const Prim<DIM>& R, int dir, double* F) {
    double rholog = logmean(L.rho, R.rho, L.lrho, R.lrho);
    double betalog = logmean(L.beta, R.beta, L.lbeta, R.lbeta);
    double rhoavg = 0.5 * (L.rho + R.rho);
    double uavg = 0.5 * (L.u + R.u);
    if (DIM == 1) {
        double unorm = L.u * R.u;
        double pa = rhoavg / (L.beta + R.beta);
        double f4aux = rholog / (2.0 * (g - 1.0) * betalog) + pa + 0.5 * rholog * unorm;
        F[0] = rholog * uavg;
        F[1] = F[0] * uavg + pa;
        F[2] = f4aux * uavg;
        return;
    }
    double vavg = 0.5 * (L.v + R.v);
    double unorm = L.u * R.u + L.v * R.v;
    double pa = rhoavg / (L.beta + R.beta);
    double f4aux = rholog / (2.0 * (g - 1.0) * betalog) + pa + 0.5 * rholog * unorm;
    if (dir == 0) {
        double FxS1 = rholog * uavg;
        F[0] = FxS1;
        F[1] = FxS1 * uavg + pa;
        F[2] = FxS1 * vavg;
        F[3] = f4aux * uavg;
    } else {
        double FxS1 = rholog * uavg;
        double FyS1 = rholog * vavg;
        F[0] = FyS1;
        F[1] = FxS1 * vavg;      // FyS2 = FxS3
        F[2] = FyS1 * vavg + pa;
        F[3] = f4aux * vavg;
    }
}

}  // namespace p2de
