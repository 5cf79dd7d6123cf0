// p2de_capi.cpp -- the C-ABI of include/p2de_hip.h: validation of options and operator
// structure, operator tables, device memory, the stream, and the stage/step sequencing.
// No C++ exception crosses the ABI; every failure becomes a status and a message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "p2de_hip.h"
#include "p2de_internal.h"

using namespace p2de;

namespace {

thread_local std::string g_last_error;

struct TimedLaunch {
    int kind;
    hipEvent_t a, b;
};

}  // namespace

struct p2de_ctx {
    p2de_config cfg;
    int Nd, Nc, N1D, Nq, Nfp, Nh, K;
    ElemTables tab;
    // device buffers
    ElemTables* d_tab = nullptr;
    Geo* d_geo = nullptr;
    int *d_nb_node = nullptr, *d_nb_face = nullptr, *d_bc = nullptr;
    double* d_Ival = nullptr;
    double* d_U[2] = {nullptr, nullptr};
    double *d_rhsUA = nullptr, *d_lA = nullptr, *d_dfA = nullptr, *d_rhsU = nullptr;
    double* d_ut = nullptr;                // projected face states of d_U[0]
    bool ut_valid = false;                 // d_ut matches d_U[0]
    double *d_dbg_rhsH = nullptr, *d_dbg_rhsL = nullptr, *d_dbg_L = nullptr;
    StepScalars* d_sc = nullptr;
    unsigned long long* d_keys = nullptr;  // [min rho key, min p key, count]
    StepScalars* h_sc = nullptr;           // pinned
    unsigned long long* h_keys = nullptr;  // pinned
    hipStream_t stream = nullptr;
    int variantA = 0;                      // phase-A launch shape (autotuned)
    bool state_set = false;
    bool debug = false;
    bool timing = false;
    std::vector<TimedLaunch> launches;
    std::vector<hipEvent_t> event_pool;
    double time_ms[3] = {0, 0, 0};
    int64_t time_count[3] = {0, 0, 0};
    std::string err;
};

namespace {

int fail(p2de_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    g_last_error = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(ctx, P2DE_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define LAUNCHCHK(ctx, expr)                                                                      \
    do {                                                                                          \
        int e_ = (expr);                                                                          \
        if (e_ != 0)                                                                              \
            return fail(ctx, P2DE_E_HIP, std::string(#expr) + ": " + hipGetErrorString((hipError_t)e_)); \
    } while (0)

template <class T>
int dalloc(p2de_ctx* c, T** p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(c, P2DE_E_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    return P2DE_OK;
}

void free_ctx(p2de_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& l : c->launches) {
        (void)hipEventDestroy(l.a);
        (void)hipEventDestroy(l.b);
    }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    void* dptrs[] = {c->d_tab, c->d_geo, c->d_nb_node, c->d_nb_face, c->d_bc, c->d_Ival, c->d_U[0], c->d_U[1],
                     c->d_rhsUA, c->d_lA, c->d_dfA, c->d_rhsU, c->d_ut, c->d_dbg_rhsH, c->d_dbg_rhsL, c->d_dbg_L,
                     c->d_sc, c->d_keys};
    for (void* p : dptrs)
        if (p) (void)hipFree(p);
    if (c->h_sc) (void)hipHostFree(c->h_sc);
    if (c->h_keys) (void)hipHostFree(c->h_keys);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int64_t ipow(int a, int b) {
    int64_t r = 1;
    for (int i = 0; i < b; i++) r *= a;
    return r;
}

// Column-major access helpers.
inline double cm(const double* A, int64_t ld, int64_t i, int64_t j) { return A[i + ld * j]; }

hipEvent_t take_event(p2de_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Record timing events around a launch when kernel timing is on.
template <class F>
int timed(p2de_ctx* c, int kind, F&& launch) {
    if (!c->timing) return launch();
    TimedLaunch t{kind, take_event(c), take_event(c)};
    (void)hipEventRecord(t.a, c->stream);
    int r = launch();
    (void)hipEventRecord(t.b, c->stream);
    c->launches.push_back(t);
    return r;
}

void harvest_timing(p2de_ctx* c) {
    for (auto& l : c->launches) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, l.a, l.b) == hipSuccess) {
            c->time_ms[l.kind] += ms;
            c->time_count[l.kind] += 1;
        }
        c->event_pool.push_back(l.a);
        c->event_pool.push_back(l.b);
    }
    c->launches.clear();
}

int build_tables(p2de_ctx* c, const p2de_operators* ops) {
    ElemTables& T = c->tab;
    std::memset(&T, 0, sizeof(T));
    const int Nd = c->Nd, N1D = c->N1D, Nq = c->Nq, Nfp = c->Nfp, Nh = c->Nh;
    T.Nd = Nd;
    T.Nc = c->Nc;
    T.N1D = N1D;
    T.Nq = Nq;
    T.Nfp = Nfp;
    T.Nh = Nh;
    const int64_t nph_expect = Nd == 1 ? N1D * (N1D - 1) / 2 + 2 : (int64_t)N1D * N1D * (N1D - 1) + Nfp;
    const int64_t npl_expect = Nd == 1 ? N1D - 1 : 2 * N1D * (N1D - 1);
    if (ops->nnz_h != nph_expect)
        return fail(c, P2DE_E_UNSUPPORTED, "Srsh_nnz has " + std::to_string(ops->nnz_h) + " pairs, the Lobatto operator has " +
                                               std::to_string(nph_expect));
    if (ops->nnz_0 != npl_expect)
        return fail(c, P2DE_E_UNSUPPORTED, "Srs0_nnz has " + std::to_string(ops->nnz_0) + " pairs, expected " +
                                               std::to_string(npl_expect));
    T.nph = (int)nph_expect;
    T.npl = (int)npl_expect;

    // fq2q and the tensor-product face layout the subcell limiter assumes (subcell.jl:173-206)
    for (int f = 0; f < Nfp; f++) {
        const int64_t q1 = ops->fq2q[f];
        if (q1 < 1 || q1 > Nq) return fail(c, P2DE_E_INVALID, "fq2q out of range");
        const int q = (int)q1 - 1;
        int expect;
        if (Nd == 1) {
            expect = f == 0 ? 0 : N1D - 1;
        } else {
            const int face = f / N1D, m = f % N1D;
            expect = face == 0 ? m * N1D : face == 1 ? m * N1D + N1D - 1 : face == 2 ? m : (N1D - 1) * N1D + m;
        }
        if (q != expect) return fail(c, P2DE_E_UNSUPPORTED, "fq2q does not follow the tensor-product face layout");
        T.fq2q[f] = (uint8_t)q;
        T.ld[f] = (int8_t)((Nd == 1 || f < 2 * N1D) ? 0 : 1);   // rhs_utils.jl:84-88
        if (Nd == 1) T.bsign[f] = f == 0 ? -1 : 1;
        else T.bsign[f] = (f < N1D || (f >= 2 * N1D && f < 3 * N1D)) ? -1 : 1;
        T.Br[f] = cm(ops->Brs, Nfp, f, f);
        T.Bs[f] = Nd == 2 ? cm(ops->Brs + (int64_t)Nfp * Nfp, Nfp, f, f) : 0.0;
        if (Nd == 2) {
            const bool xface = f < 2 * N1D;
            if (xface ? (T.Br[f] == 0.0 || T.Bs[f] != 0.0) : (T.Bs[f] == 0.0 || T.Br[f] != 0.0))
                return fail(c, P2DE_E_UNSUPPORTED, "Brs does not have one nonzero direction per face node");
        } else if (T.Br[f] == 0.0) {
            return fail(c, P2DE_E_UNSUPPORTED, "Br has a zero face entry");
        }
        // entropy projection row (rhs.jl:90-92 with l = 1): a single nonzero at fq2q[f]
        for (int j = 0; j < Nq; j++) {
            const double vf = cm(ops->Vf, Nfp, f, j);
            const double coef = 1.0 * vf + (1.0 - 1.0) * cm(ops->Vf_low, Nfp, f, j);
            if (j == q) {
                if (coef == 0.0) return fail(c, P2DE_E_UNSUPPORTED, "Vf is zero at fq2q");
                T.vf_coef[f] = coef;
            } else if (vf != 0.0) {
                return fail(c, P2DE_E_UNSUPPORTED, "Vf is not a face-node selection (Gauss basis is not implemented)");
            }
        }
    }
    // node data and the assembly structure of MinvVhT / MinvVfT (flux_differencing.jl:331-361)
    for (int i = 0; i < Nq; i++) {
        T.wq[i] = ops->wq[i];
        if (!(T.wq[i] > 0.0)) return fail(c, P2DE_E_INVALID, "wq must be positive");
        for (int j = 0; j < Nq; j++)
            if (j != i && cm(ops->MinvVhT, Nq, i, j) != 0.0)
                return fail(c, P2DE_E_UNSUPPORTED, "MinvVhT volume block is not diagonal");
        T.a_v[i] = cm(ops->MinvVhT, Nq, i, i);
        int n = 0;
        for (int f = 0; f < Nfp; f++) {
            const double ah = cm(ops->MinvVhT, Nq, i, Nq + f), af = cm(ops->MinvVfT, Nq, i, f);
            if (T.fq2q[f] == i) {
                if (n >= 2) return fail(c, P2DE_E_UNSUPPORTED, "more than two faces at a node");
                T.nface[i][n] = (int8_t)f;
                T.a_h[i][n] = ah;
                T.a_f[i][n] = af;
                n++;
            } else if (ah != 0.0 || af != 0.0) {
                return fail(c, P2DE_E_UNSUPPORTED, "MinvVhT/MinvVfT couple a node to a face not at that node");
            }
        }
        T.nf[i] = (int8_t)n;
        int m = 0;
        for (int s = 0; s < 2; s++) {
            const int64_t f1 = ops->q2fq[i + (int64_t)Nq * s];
            if (f1 == 0) continue;
            if (f1 < 1 || f1 > Nfp) return fail(c, P2DE_E_INVALID, "q2fq out of range");
            T.q2fq[i][m++] = (int8_t)(f1 - 1);
        }
        T.nq2fq[i] = (int8_t)m;
    }
    // high-order pairs
    std::vector<std::vector<std::pair<int, int>>> hl(Nh);
    for (int p = 0; p < T.nph; p++) {
        const int64_t i1 = ops->Srsh_nnz[2 * p], j1 = ops->Srsh_nnz[2 * p + 1];
        if (i1 < 1 || i1 > Nh || j1 < 1 || j1 > Nh || i1 <= j1) return fail(c, P2DE_E_INVALID, "Srsh_nnz entry out of range");
        const int i = (int)i1 - 1, j = (int)j1 - 1;
        const double cr = cm(ops->Srsh_db, Nh, i, j);
        const double cs = Nd == 2 ? cm(ops->Srsh_db + (int64_t)Nh * Nh, Nh, i, j) : 0.0;
        if ((cr != 0.0) == (cs != 0.0) && Nd == 2)
            return fail(c, P2DE_E_UNSUPPORTED, "a Srsh pair acts in both or no directions");
        if (Nd == 1 && cr == 0.0) return fail(c, P2DE_E_UNSUPPORTED, "zero Srh pair");
        T.hp_i[p] = (uint8_t)i;
        T.hp_j[p] = (uint8_t)j;
        T.hp_dir[p] = (uint8_t)(cr != 0.0 ? 0 : 1);
        T.hp_cr[p] = cr;
        T.hp_cs[p] = cs;
        hl[i].push_back({p, +1});
        hl[j].push_back({p, -1});
    }
    int pos = 0;
    for (int h = 0; h < Nh; h++) {
        T.hl_start[h] = (int16_t)pos;
        for (auto& e : hl[h]) {
            T.hl_pair[pos] = (int16_t)e.first;
            T.hl_sign[pos] = (int8_t)e.second;
            pos++;
        }
    }
    T.hl_start[Nh] = (int16_t)pos;
    // low-order pairs
    std::vector<std::vector<std::pair<int, int>>> ll(Nq), cl(Nq);
    for (int p = 0; p < T.npl; p++) {
        const int64_t i1 = ops->Srs0_nnz[2 * p], j1 = ops->Srs0_nnz[2 * p + 1];
        if (i1 < 1 || i1 > Nq || j1 < 1 || j1 > Nq || i1 <= j1) return fail(c, P2DE_E_INVALID, "Srs0_nnz entry out of range");
        const int i = (int)i1 - 1, j = (int)j1 - 1;
        const double cr = cm(ops->Srs0, Nq, i, j);
        const double cs = Nd == 2 ? cm(ops->Srs0 + (int64_t)Nq * Nq, Nq, i, j) : 0.0;
        if ((cr != 0.0) == (cs != 0.0) && Nd == 2)
            return fail(c, P2DE_E_UNSUPPORTED, "a Srs0 pair acts in both or no directions");
        if (Nd == 1 && cr == 0.0) return fail(c, P2DE_E_UNSUPPORTED, "zero Sr0 pair");
        T.lp_i[p] = (uint8_t)i;
        T.lp_j[p] = (uint8_t)j;
        T.lp_dir[p] = (uint8_t)(cr != 0.0 ? 0 : 1);
        T.lp_cr[p] = cr;
        T.lp_cs[p] = cs;
        ll[i].push_back({p, +1});
        ll[j].push_back({p, -1});
        cl[i].push_back({j, p});
        cl[j].push_back({i, p});
    }
    pos = 0;
    int cpos = 0;
    for (int i = 0; i < Nq; i++) {
        T.ll_start[i] = (int16_t)pos;
        for (auto& e : ll[i]) {
            T.ll_pair[pos] = (int16_t)e.first;
            T.ll_sign[pos] = (int8_t)e.second;
            pos++;
        }
        std::sort(cl[i].begin(), cl[i].end());
        T.cl_start[i] = (int16_t)cpos;
        for (auto& e : cl[i]) T.cl_pair[cpos++] = (int16_t)e.second;
    }
    T.ll_start[Nq] = (int16_t)pos;
    T.cl_start[Nq] = (int16_t)cpos;
    return P2DE_OK;
}

}  // namespace

extern "C" {

const char* p2de_status_string(int s) {
    switch (s) {
        case P2DE_OK: return "ok";
        case P2DE_E_INVALID: return "invalid argument";
        case P2DE_E_UNSUPPORTED: return "unsupported option or operator structure";
        case P2DE_E_HIP: return "HIP runtime error";
        case P2DE_E_ALLOC: return "allocation failed";
        case P2DE_E_STATE: return "call out of order";
        case P2DE_E_POSITIVITY: return "non-positive density or pressure";
        default: return "unknown status";
    }
}

const char* p2de_last_error(const p2de_ctx* ctx) { return ctx ? ctx->err.c_str() : g_last_error.c_str(); }

int p2de_create(const p2de_config* cfg, const p2de_operators* ops, const p2de_geometry* geom, const p2de_bcdata* bc,
                p2de_ctx** out) {
    if (!out) return fail(nullptr, P2DE_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!cfg || !ops || !geom || !bc) return fail(nullptr, P2DE_E_INVALID, "NULL config, operators, geometry or bcdata");
    p2de_ctx* c = nullptr;
    try {
        c = new p2de_ctx();
    } catch (...) {
        return fail(nullptr, P2DE_E_ALLOC, "host allocation failed");
    }
    auto bail = [&](int code) {
        g_last_error = c->err;
        free_ctx(c);
        return code;
    };
    try {
        c->cfg = *cfg;
        // option set (Solver.jl:1-121); anything else is not implemented here
        if (cfg->rhs != P2DE_RHS_LIMITED_DG || cfg->low_order_surface_flux != P2DE_LF_ON_NODAL_VAL ||
            cfg->high_order_surface_flux != P2DE_LF_ON_PROJECTED_VAL ||
            cfg->high_order_volume_flux != P2DE_CHANDRASHEKAR_FLUX || cfg->basis != P2DE_BASIS_LOBATTO ||
            cfg->entropyproj_limiter != P2DE_NO_ENTROPY_PROJ_LIMITER || cfg->rhs_limiter != P2DE_SUBCELL_LIMITER ||
            cfg->bound != P2DE_POSITIVITY_BOUND || cfg->shockcapture != P2DE_NO_SHOCK_CAPTURE ||
            cfg->precision != P2DE_PRECISION_FP64)
            return fail(c, P2DE_E_UNSUPPORTED, "option set not implemented"), bail(P2DE_E_UNSUPPORTED);
        if (cfg->upper_bound_mode != P2DE_UPPER_BOUND_REFERENCE && cfg->upper_bound_mode != P2DE_UPPER_BOUND_INTENDED)
            return fail(c, P2DE_E_INVALID, "upper_bound_mode"), bail(P2DE_E_INVALID);
        if (cfg->Nd != 1 && cfg->Nd != 2) return fail(c, P2DE_E_INVALID, "Nd must be 1 or 2"), bail(P2DE_E_INVALID);
        if (cfg->N < 1 || cfg->N > 4) return fail(c, P2DE_E_UNSUPPORTED, "N must be 1..4"), bail(P2DE_E_UNSUPPORTED);
        if (cfg->Kx < 1 || cfg->Ky < 1 || (cfg->Nd == 1 && cfg->Ky != 1))
            return fail(c, P2DE_E_INVALID, "Kx, Ky"), bail(P2DE_E_INVALID);
        if (!(cfg->gamma > 1.0)) return fail(c, P2DE_E_INVALID, "gamma must exceed 1"), bail(P2DE_E_INVALID);
        c->Nd = cfg->Nd;
        c->Nc = cfg->Nd + 2;
        c->N1D = cfg->N + 1;
        c->Nq = (int)ipow(c->N1D, c->Nd);
        c->Nfp = c->Nd == 1 ? 2 : 4 * c->N1D;
        c->Nh = c->Nq + c->Nfp;
        const int64_t K64 = (int64_t)cfg->Kx * cfg->Ky;
        if (K64 * c->Nq * c->Nc >= (int64_t)1 << 31 || K64 * c->Nfp >= (int64_t)1 << 31)
            return fail(c, P2DE_E_UNSUPPORTED, "problem too large for 32-bit node indices"), bail(P2DE_E_UNSUPPORTED);
        c->K = (int)K64;
        if (ops->Nq != c->Nq || ops->Nfp != c->Nfp || ops->Nh != c->Nh)
            return fail(c, P2DE_E_INVALID, "operator sizes do not match Nd and N"), bail(P2DE_E_INVALID);
        const void* req[] = {ops->Srsh_db, ops->Srs0, ops->Srsh_nnz, ops->Srs0_nnz, ops->Brs, ops->MinvVhT,
                             ops->MinvVfT, ops->Vf, ops->Vf_low, ops->wq, ops->fq2q, ops->q2fq, geom->Jq, geom->GJh,
                             bc->mapP};
        for (const void* p : req)
            if (!p) return fail(c, P2DE_E_INVALID, "NULL operator, geometry or mapP array"), bail(P2DE_E_INVALID);
        int st = build_tables(c, ops);
        if (st) return bail(st);

        const int K = c->K, Nq = c->Nq, Nfp = c->Nfp, Nh = c->Nh, Nc = c->Nc, N1D = c->N1D;
        // per-element metric, constant inside the element
        std::vector<Geo> geo(K);
        const int NGEO = c->Nd == 1 ? 1 : 4;
        for (int k = 0; k < K; k++) {
            const double J = geom->Jq[(int64_t)Nq * k];
            for (int i = 1; i < Nq; i++)
                if (geom->Jq[i + (int64_t)Nq * k] != J)
                    return fail(c, P2DE_E_UNSUPPORTED, "Jq varies inside an element"), bail(P2DE_E_UNSUPPORTED);
            if (!(J > 0.0)) return fail(c, P2DE_E_INVALID, "Jq must be positive"), bail(P2DE_E_INVALID);
            double gv[4] = {0, 0, 0, 0};
            for (int gi = 0; gi < NGEO; gi++) {
                const double* G = geom->GJh + (int64_t)gi * Nh * K;
                gv[gi] = G[(int64_t)Nh * k];
                for (int h = 1; h < Nh; h++)
                    if (G[h + (int64_t)Nh * k] != gv[gi])
                        return fail(c, P2DE_E_UNSUPPORTED, "metric terms vary inside an element"), bail(P2DE_E_UNSUPPORTED);
            }
            Geo g{J, gv[0], 0.0, 0.0, 0.0};
            if (c->Nd == 2) {
                g.sxJ = gv[1];
                g.ryJ = gv[2];
                g.syJ = gv[3];
                if (g.sxJ != 0.0 || g.ryJ != 0.0)
                    return fail(c, P2DE_E_UNSUPPORTED, "elements must be axis aligned (sxJ = ryJ = 0)"), bail(P2DE_E_UNSUPPORTED);
            }
            // the graph-viscosity direction test |Sx0| > 1e-10 (rhs_utils.jl:256) must pick the
            // direction the pair acts in
            for (int p = 0; p < c->tab.npl && c->Nd == 2; p++) {
                const double s0x = g.rxJ * c->tab.lp_cr[p] + g.sxJ * c->tab.lp_cs[p];
                const int vdir = std::fabs(s0x) > 1e-10 ? 0 : 1;
                if (vdir != c->tab.lp_dir[p])
                    return fail(c, P2DE_E_UNSUPPORTED, "graph-viscosity direction test disagrees with the pair direction"),
                           bail(P2DE_E_UNSUPPORTED);
            }
            geo[k] = g;
        }
        // face connectivity and boundary codes (StateParam.jl:1-7)
        std::vector<int> nb_node((size_t)K * Nfp), nb_face((size_t)K * Nfp), bcode((size_t)K * Nfp, 0);
        for (int k = 0; k < K; k++)
            for (int f = 0; f < Nfp; f++) {
                const int64_t m = bc->mapP[f + (int64_t)Nfp * k] - 1;
                if (m < 0 || m >= (int64_t)Nfp * K) return fail(c, P2DE_E_INVALID, "mapP out of range"), bail(P2DE_E_INVALID);
                const int kP = (int)(m / Nfp), iP = (int)(m % Nfp);
                if (c->Nd == 2 && ((f < 2 * N1D) != (iP < 2 * N1D)))
                    return fail(c, P2DE_E_UNSUPPORTED, "mapP joins an x face to a y face"), bail(P2DE_E_UNSUPPORTED);
                nb_face[(size_t)k * Nfp + f] = (int)m;
                nb_node[(size_t)k * Nfp + f] = kP * Nq + c->tab.fq2q[iP];
            }
        if (bc->nI > 0 && (!bc->mapI || !bc->Ival)) return fail(c, P2DE_E_INVALID, "mapI/Ival"), bail(P2DE_E_INVALID);
        if (bc->nO > 0 && !bc->mapO) return fail(c, P2DE_E_INVALID, "mapO"), bail(P2DE_E_INVALID);
        for (int i = 0; i < bc->nI; i++) {
            const int64_t m = bc->mapI[i] - 1;
            if (m < 0 || m >= (int64_t)Nfp * K) return fail(c, P2DE_E_INVALID, "mapI out of range"), bail(P2DE_E_INVALID);
            bcode[m] = i + 1;
        }
        for (int i = 0; i < bc->nO; i++) {   // enforce_BC! applies outflow after inflow
            const int64_t m = bc->mapO[i] - 1;
            if (m < 0 || m >= (int64_t)Nfp * K) return fail(c, P2DE_E_INVALID, "mapO out of range"), bail(P2DE_E_INVALID);
            bcode[m] = -1;
        }

        // device resources (default-priority stream)
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
            return fail(c, P2DE_E_HIP, "hipStreamCreate failed"), bail(P2DE_E_HIP);
        const size_t nU = (size_t)K * Nq * Nc;
        if ((st = dalloc(c, &c->d_tab, 1)) || (st = dalloc(c, &c->d_geo, (size_t)K)) ||
            (st = dalloc(c, &c->d_nb_node, (size_t)K * Nfp)) || (st = dalloc(c, &c->d_nb_face, (size_t)K * Nfp)) ||
            (st = dalloc(c, &c->d_bc, (size_t)K * Nfp)) || (st = dalloc(c, &c->d_Ival, (size_t)std::max(bc->nI, 1) * Nc)) ||
            (st = dalloc(c, &c->d_U[0], nU)) || (st = dalloc(c, &c->d_U[1], nU)) || (st = dalloc(c, &c->d_rhsUA, nU)) ||
            (st = dalloc(c, &c->d_lA, (size_t)K * Nfp)) || (st = dalloc(c, &c->d_dfA, (size_t)K * Nfp * Nc)) ||
            (st = dalloc(c, &c->d_rhsU, nU)) || (st = dalloc(c, &c->d_ut, (size_t)K * Nfp * Nc)) ||
            (st = dalloc(c, &c->d_sc, 1)) || (st = dalloc(c, &c->d_keys, 3)))
            return bail(st);
        if (hipHostMalloc((void**)&c->h_sc, sizeof(StepScalars)) != hipSuccess ||
            hipHostMalloc((void**)&c->h_keys, 3 * sizeof(unsigned long long)) != hipSuccess)
            return fail(c, P2DE_E_ALLOC, "hipHostMalloc failed"), bail(P2DE_E_ALLOC);
        std::memset(c->h_sc, 0, sizeof(StepScalars));
        c->h_sc->t = cfg->t0;
        c->h_sc->first_bad = ~0ull;
        hipError_t e;
        if ((e = hipMemcpy(c->d_tab, &c->tab, sizeof(ElemTables), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(c->d_geo, geo.data(), sizeof(Geo) * K, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(c->d_nb_node, nb_node.data(), sizeof(int) * nb_node.size(), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(c->d_nb_face, nb_face.data(), sizeof(int) * nb_face.size(), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(c->d_bc, bcode.data(), sizeof(int) * bcode.size(), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(c->d_sc, c->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(c, P2DE_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)), bail(P2DE_E_HIP);
        if (bc->nI > 0 &&
            (e = hipMemcpy(c->d_Ival, bc->Ival, sizeof(double) * bc->nI * Nc, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(c, P2DE_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)), bail(P2DE_E_HIP);
    } catch (const std::bad_alloc&) {
        return fail(c, P2DE_E_ALLOC, "host allocation failed"), bail(P2DE_E_ALLOC);
    } catch (...) {
        return fail(c, P2DE_E_INVALID, "unexpected exception"), bail(P2DE_E_INVALID);
    }
    *out = c;
    return P2DE_OK;
}

void p2de_destroy(p2de_ctx* ctx) { free_ctx(ctx); }

int64_t p2de_num_dofs(const p2de_ctx* ctx) { return ctx ? (int64_t)ctx->K * ctx->Nq : 0; }

int p2de_device_state(p2de_ctx* ctx, void** dptr) {
    if (!ctx || !dptr) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    *dptr = ctx->d_U[0];
    return P2DE_OK;
}

int p2de_set_state(p2de_ctx* ctx, const double* Uq) {
    if (!ctx || !Uq) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_U[0], Uq, sizeof(double) * (size_t)ctx->K * ctx->Nq * ctx->Nc,
                               hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->state_set = true;
    ctx->ut_valid = false;
    return P2DE_OK;
}

int p2de_get_state(p2de_ctx* ctx, double* Uq) {
    if (!ctx || !Uq) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    HIPCHK(ctx, hipMemcpyAsync(Uq, ctx->d_U[0], sizeof(double) * (size_t)ctx->K * ctx->Nq * ctx->Nc,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return P2DE_OK;
}

int p2de_set_debug(p2de_ctx* ctx, int on) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (on && !ctx->d_dbg_rhsH) {
        const size_t nU = (size_t)ctx->K * ctx->Nq * ctx->Nc;
        const size_t nL = (size_t)(ctx->Nq + ctx->N1D) * ctx->Nd * ctx->K;
        int st;
        if ((st = dalloc(ctx, &ctx->d_dbg_rhsH, nU)) || (st = dalloc(ctx, &ctx->d_dbg_rhsL, nU)) ||
            (st = dalloc(ctx, &ctx->d_dbg_L, nL)))
            return st;
    }
    ctx->debug = on != 0;
    return P2DE_OK;
}

}  // extern "C"

namespace {

PhaseAArgs argsA(p2de_ctx* c, const double* U, int stage1) {
    PhaseAArgs a{};
    a.U = U;
    a.ut = c->d_ut;
    a.T = c->d_tab;
    a.geo = c->d_geo;
    a.nb_node = c->d_nb_node;
    a.nb_face = c->d_nb_face;
    a.bc = c->d_bc;
    a.Ival = c->d_Ival;
    a.sc = c->d_sc;
    a.K = c->K;
    a.gamma = c->cfg.gamma;
    a.zeta = c->cfg.zeta;
    a.ZEROTOL = c->cfg.ZEROTOL;
    a.CFL = c->cfg.CFL;
    a.upper_mode = c->cfg.upper_bound_mode;
    a.stage1 = stage1;
    a.rhsUA = c->d_rhsUA;
    a.lA = c->d_lA;
    a.dfA = c->d_dfA;
    if (c->debug) {
        a.dbg_rhsH = c->d_dbg_rhsH;
        a.dbg_rhsL = c->d_dbg_rhsL;
        a.dbg_L = c->d_dbg_L;
    }
    return a;
}

PhaseBArgs argsB(p2de_ctx* c, int mode, const double* W, const double* Uin, double* Uout, int finalize) {
    PhaseBArgs b{};
    b.rhsUA = c->d_rhsUA;
    b.lA = c->d_lA;
    b.dfA = c->d_dfA;
    b.nb_face = c->d_nb_face;
    b.T = c->d_tab;
    b.geo = c->d_geo;
    b.sc = c->d_sc;
    b.K = c->K;
    b.gamma = c->cfg.gamma;
    b.mode = mode;
    b.W = W;
    b.Uin = Uin;
    b.Uout = Uout;
    b.rhsU = c->d_rhsU;
    b.dbg_L = c->debug ? c->d_dbg_L : nullptr;
    b.ut = c->d_ut;
    b.finalize = finalize;
    return b;
}

// Projected face states of d_U[0] when they are not already current.
int ensure_face_states(p2de_ctx* c) {
    if (c->ut_valid) return P2DE_OK;
    LAUNCHCHK(c, timed(c, 2, [&] {
                  return launch_face_state(c->Nd, c->N1D, c->d_U[0], c->d_tab, c->K, c->cfg.gamma, c->d_ut, c->stream);
              }));
    c->ut_valid = true;
    return P2DE_OK;
}

int enqueue_step(p2de_ctx* c) {
    void* s = c->stream;
    const p2de_config& g = c->cfg;
    double* A = c->d_U[0];
    double* B = c->d_U[1];
    int st = ensure_face_states(c);
    if (st) return st;
    LAUNCHCHK(c, timed(c, 2, [&] { return launch_step_init(c->d_sc, g.CFL, g.dt0, g.T, s); }));
    // stage 1: U1 = W + dt*rhsU (SSPRK33.jl:32-33)
    LAUNCHCHK(c, timed(c, 0, [&] { return launch_phaseA(c->Nd, c->N1D, c->variantA, argsA(c, A, 1), s); }));
    LAUNCHCHK(c, timed(c, 1, [&] { return launch_phaseB(c->Nd, c->N1D, argsB(c, MODE_STAGE1, A, A, B, 0), s); }));
    // stage 2: U2 = 3/4 W + 1/4 (U1 + dt*rhsU) (SSPRK33.jl:34-36)
    LAUNCHCHK(c, timed(c, 0, [&] { return launch_phaseA(c->Nd, c->N1D, c->variantA, argsA(c, B, 0), s); }));
    LAUNCHCHK(c, timed(c, 1, [&] { return launch_phaseB(c->Nd, c->N1D, argsB(c, MODE_STAGE2, A, B, B, 0), s); }));
    // stage 3: U^{n+1} = 1/3 W + 2/3 (U2 + dt*rhsU) (SSPRK33.jl:37-39), t += dt
    LAUNCHCHK(c, timed(c, 0, [&] { return launch_phaseA(c->Nd, c->N1D, c->variantA, argsA(c, B, 0), s); }));
    LAUNCHCHK(c, timed(c, 1, [&] { return launch_phaseB(c->Nd, c->N1D, argsB(c, MODE_STAGE3, A, B, A, 1), s); }));
    return P2DE_OK;
}

int read_scalars(p2de_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->h_sc, c->d_sc, sizeof(StepScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->timing) harvest_timing(c);
    return P2DE_OK;
}

double bits_to_double_h(unsigned long long b) {
    double d;
    std::memcpy(&d, &b, sizeof(d));
    return d;
}

double jmin_h(double x, double y) {
    double diff = x - y;
    double am = std::signbit(diff) ? x : y;
    return (std::isnan(x) || std::isnan(y)) ? diff : am;
}

}  // namespace

extern "C" {

int p2de_rhs(p2de_ctx* ctx, double t, double dt, int nstage, double* dt_out, double* rhsU) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (nstage < 1 || nstage > 3) return fail(ctx, P2DE_E_INVALID, "nstage must be 1, 2 or 3");
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    void* s = ctx->stream;
    const p2de_config& g = ctx->cfg;
    // the stage-1 CFL reduction starts from min(CFL*dt0, T - t) (low_order_graph_viscosity.jl:230)
    const double cap = nstage == 1 ? jmin_h(g.CFL * g.dt0, g.T - t) : dt;
    LAUNCHCHK(ctx, timed(ctx, 2, [&] { return launch_rhs_init(ctx->d_sc, dt, cap, s); }));
    int st0 = ensure_face_states(ctx);
    if (st0) return st0;
    LAUNCHCHK(ctx, timed(ctx, 0, [&] { return launch_phaseA(ctx->Nd, ctx->N1D, ctx->variantA, argsA(ctx, ctx->d_U[0], nstage == 1), s); }));
    LAUNCHCHK(ctx, timed(ctx, 1, [&] {
                  return launch_phaseB(ctx->Nd, ctx->N1D, argsB(ctx, MODE_RHS, ctx->d_U[0], ctx->d_U[0], nullptr, 0), s);
              }));
    if (rhsU)
        HIPCHK(ctx, hipMemcpyAsync(rhsU, ctx->d_rhsU, sizeof(double) * (size_t)ctx->K * ctx->Nq * ctx->Nc,
                                   hipMemcpyDeviceToHost, ctx->stream));
    int st = read_scalars(ctx);
    if (st) return st;
    if (dt_out) *dt_out = bits_to_double_h(ctx->h_sc->dt_bits);
    return P2DE_OK;
}

int p2de_get_field(p2de_ctx* ctx, int field, double* out) {
    if (!ctx || !out) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    const size_t nU = (size_t)ctx->K * ctx->Nq * ctx->Nc;
    const double* src = nullptr;
    size_t n = nU;
    switch (field) {
        case P2DE_FIELD_RHSU: src = ctx->d_rhsU; break;
        case P2DE_FIELD_RHSH: src = ctx->d_dbg_rhsH; break;
        case P2DE_FIELD_RHSL: src = ctx->d_dbg_rhsL; break;
        case P2DE_FIELD_L_LOCAL:
            src = ctx->d_dbg_L;
            n = (size_t)(ctx->Nq + ctx->N1D) * ctx->Nd * ctx->K;
            break;
        default: return fail(ctx, P2DE_E_INVALID, "unknown field");
    }
    if (!src) return fail(ctx, P2DE_E_STATE, "debug output is off (p2de_set_debug)");
    HIPCHK(ctx, hipMemcpyAsync(out, src, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return P2DE_OK;
}

int p2de_set_time(p2de_ctx* ctx, double t) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    int st = read_scalars(ctx);
    if (st) return st;
    ctx->h_sc->t = t;
    ctx->h_sc->done = 0;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_sc, ctx->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return P2DE_OK;
}

int p2de_get_time(p2de_ctx* ctx, double* t, double* last_dt) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    int st = read_scalars(ctx);
    if (st) return st;
    if (t) *t = ctx->h_sc->t;
    if (last_dt) *last_dt = bits_to_double_h(ctx->h_sc->dt_bits);
    return P2DE_OK;
}

int p2de_ssprk3_step(p2de_ctx* ctx, double* t_inout, double* dt_out) {
    if (!ctx || !t_inout) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    int st = p2de_set_time(ctx, *t_inout);
    if (st) return st;
    if ((st = enqueue_step(ctx))) return st;
    if ((st = read_scalars(ctx))) return st;
    const bool done = ctx->h_sc->done != 0;
    *t_inout = ctx->h_sc->t;
    if (dt_out) *dt_out = done ? 0.0 : bits_to_double_h(ctx->h_sc->dt_bits);
    return P2DE_OK;
}

int p2de_enqueue_steps(p2de_ctx* ctx, int steps) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    for (int i = 0; i < steps; i++) {
        int st = enqueue_step(ctx);
        if (st) return st;
    }
    return P2DE_OK;
}

int p2de_synchronize(p2de_ctx* ctx) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->timing) harvest_timing(ctx);
    return P2DE_OK;
}

int p2de_run(p2de_ctx* ctx, int max_steps, p2de_run_stats* stats) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    int st = read_scalars(ctx);
    if (st) return st;
    const int steps0 = ctx->h_sc->steps;
    const int chunk = 32;
    int launched = 0;
    while (true) {
        int n = chunk;
        if (max_steps > 0) n = std::min(n, max_steps - launched);
        if (n <= 0) break;
        if ((st = p2de_enqueue_steps(ctx, n))) return st;
        launched += n;
        if ((st = read_scalars(ctx))) return st;
        if (ctx->h_sc->done || !(ctx->h_sc->t < ctx->cfg.T)) break;
    }
    if (stats) {
        stats->steps = ctx->h_sc->steps - steps0;
        stats->t = ctx->h_sc->t;
        stats->last_dt = bits_to_double_h(ctx->h_sc->dt_bits);
        stats->violations = (int64_t)ctx->h_sc->viol;
    }
    return P2DE_OK;
}

int p2de_check_positivity(p2de_ctx* ctx, p2de_pos_report* rep) {
    if (!ctx || !rep) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    ctx->h_keys[0] = ~0ull;
    ctx->h_keys[1] = ~0ull;
    ctx->h_keys[2] = 0;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_keys, ctx->h_keys, 3 * sizeof(unsigned long long), hipMemcpyHostToDevice, ctx->stream));
    LAUNCHCHK(ctx, launch_positivity(ctx->d_U[0], (int64_t)ctx->K * ctx->Nq, ctx->Nd, ctx->cfg.gamma, ctx->d_keys,
                                     ctx->d_keys + 2, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_keys, ctx->d_keys, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    int st = read_scalars(ctx);
    if (st) return st;
    auto unkey = [](unsigned long long k) {
        unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
        return bits_to_double_h(b);
    };
    rep->min_rho = unkey(ctx->h_keys[0]);
    rep->min_p = unkey(ctx->h_keys[1]);
    rep->violations = (int64_t)ctx->h_sc->viol + (int64_t)ctx->h_keys[2];
    const unsigned long long fb = ctx->h_sc->first_bad;
    rep->first_node = fb == ~0ull ? -1 : (int64_t)(fb & ((1ull << 40) - 1));
    rep->first_stage = fb == ~0ull ? 0 : (int)(fb >> 40);
    ctx->h_sc->viol = 0;
    ctx->h_sc->first_bad = ~0ull;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_sc, ctx->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return P2DE_OK;
}

int p2de_set_kernel_timing(p2de_ctx* ctx, int on) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    harvest_timing(ctx);
    for (int k = 0; k < 3; k++) {
        ctx->time_ms[k] = 0;
        ctx->time_count[k] = 0;
    }
    ctx->timing = on != 0;
    return P2DE_OK;
}

int p2de_get_kernel_timing(p2de_ctx* ctx, double ms[3], int64_t count[3]) {
    if (!ctx || !ms || !count) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    harvest_timing(ctx);
    for (int k = 0; k < 3; k++) {
        ms[k] = ctx->time_ms[k];
        count[k] = ctx->time_count[k];
    }
    return P2DE_OK;
}

}  // extern "C"

namespace {

std::string tune_key(const p2de_ctx* c) {
    std::ostringstream o;
    o << "Nd" << c->Nd << "_N" << (c->N1D - 1) << "_K" << c->K;
    return o.str();
}

// Tuning cache: one "key variant" pair per line.
std::map<std::string, int> read_cache(const char* path) {
    std::map<std::string, int> m;
    if (!path) return m;
    std::ifstream f(path);
    std::string k;
    int v;
    while (f >> k >> v) m[k] = v;
    return m;
}

void write_cache(const char* path, const std::map<std::string, int>& m) {
    if (!path) return;
    std::ofstream f(path, std::ios::trunc);
    for (auto& kv : m) f << kv.first << " " << kv.second << "\n";
}

}  // namespace

extern "C" {

int p2de_num_variants(const p2de_ctx* ctx) { return ctx ? num_variants_A(ctx->Nd, ctx->N1D) : 0; }

int p2de_set_variant(p2de_ctx* ctx, int variant) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (variant < 0 || variant >= num_variants_A(ctx->Nd, ctx->N1D)) return fail(ctx, P2DE_E_INVALID, "variant out of range");
    ctx->variantA = variant;
    return P2DE_OK;
}

int p2de_get_variant(const p2de_ctx* ctx, int* variant, int* E, int* NT) {
    if (!ctx) return P2DE_E_INVALID;
    int e = 0, nt = 0;
    variant_shape_A(ctx->Nd, ctx->N1D, ctx->variantA, &e, &nt);
    if (variant) *variant = ctx->variantA;
    if (E) *E = e;
    if (NT) *NT = nt;
    return P2DE_OK;
}

int p2de_autotune(p2de_ctx* ctx, int steps, const char* cache_path, int* chosen) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    const std::string key = tune_key(ctx);
    auto cache = read_cache(cache_path);
    const int nv = num_variants_A(ctx->Nd, ctx->N1D);
    auto it = cache.find(key);
    if (it != cache.end() && it->second >= 0 && it->second < nv) {
        ctx->variantA = it->second;
        if (chosen) *chosen = ctx->variantA;
        return P2DE_OK;
    }
    // save state and time, time each variant, restore
    const size_t nU = (size_t)ctx->K * ctx->Nq * ctx->Nc;
    double* saved = nullptr;
    int st = dalloc(ctx, &saved, nU);
    if (st) return st;
    HIPCHK(ctx, hipMemcpyAsync(saved, ctx->d_U[0], nU * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if ((st = read_scalars(ctx))) return st;
    const StepScalars keep = *ctx->h_sc;
    const bool timing_was = ctx->timing;
    int best = 0;
    double best_ms = 1e300;
    for (int v = 0; v < nv; v++) {
        ctx->variantA = v;
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_U[0], saved, nU * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        ctx->ut_valid = false;
        *ctx->h_sc = keep;
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_sc, ctx->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice, ctx->stream));
        if ((st = enqueue_step(ctx))) return st;   // warm-up
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        harvest_timing(ctx);
        ctx->timing = true;
        for (int k = 0; k < 3; k++) { ctx->time_ms[k] = 0; ctx->time_count[k] = 0; }
        for (int i = 0; i < std::max(steps, 1); i++)
            if ((st = enqueue_step(ctx))) return st;
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        harvest_timing(ctx);
        const double ms = ctx->time_ms[0] / std::max<int64_t>(ctx->time_count[0], 1);
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
        ctx->timing = false;
    }
    for (int k = 0; k < 3; k++) { ctx->time_ms[k] = 0; ctx->time_count[k] = 0; }
    ctx->timing = timing_was;
    ctx->variantA = best;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_U[0], saved, nU * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    ctx->ut_valid = false;
    *ctx->h_sc = keep;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_sc, ctx->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(saved);
    cache[key] = best;
    write_cache(cache_path, cache);
    if (chosen) *chosen = best;
    return P2DE_OK;
}

}  // extern "C"
