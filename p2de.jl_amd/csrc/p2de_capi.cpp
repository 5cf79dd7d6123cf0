// p2de_capi.cpp -- the C-ABI of include/p2de_hip.h: validation of options and operator
// structure, operator tables, device memory, the stream, and the stage/step sequencing.
// No C++ exception crosses the ABI; every failure becomes a status and a message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "p2de_hip.h"
#include "p2de_internal.h"
#include "p2de_mixed.h"
#include "p2de_setup.h"

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
    double* d_ut = nullptr;                // projected face states of d_U[0] (Lobatto) or of the stage input (Gauss)
    double* d_theta = nullptr;             // entropy-projection parameters per face node and stage (Gauss), [3][K][Nfp]
    double* d_Lz = nullptr;                // Zhang-Shu element parameters per stage, [3][K] (the reference's L)
    int last_stage = 0;                    // stage of the last element-kernel launch
    int basis = 0;                         // P2DE_BASIS_*
    bool ep_limit = false;                 // NodewiseScaledExtrapolation
    int optA = OPT_SUBCELL;                // element-kernel form for the limiter options
    double *d_blend = nullptr, *d_lphi = nullptr, *d_spart = nullptr;   // limiter option buffers
    double* d_sfac = nullptr;              // smoothness factor per element (relaxed cell-entropy bounds)
    double* d_rhoL = nullptr;              // rho + dt*rhsL_rho per node (TVD bounds)
    double* d_dtp = nullptr;               // per-workgroup CFL minima of phase A (stage 1)
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
    p2de_batch* mix = nullptr;             // precision MIXED: a one-instance batch does the work
    std::vector<double> h_Jq, h_wq;        // check_conservation's weights (Nq x K, Nq)
    // history segment of p2de_run_segment while it runs (out_interval 0: none)
    double* d_dtlog = nullptr;
    int dtlog_cap = 0;
    int seg_interval = 0, seg_cap = 0, seg_base = 0;
    long long seg_i0 = 0;
};

namespace {

int fail(p2de_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    g_last_error = msg;
    return code;
}

// status of a call forwarded to the context's one-instance batch, with its message
int fwd(p2de_ctx* c, int st) {
    if (st) {
        c->err = p2de_batch_last_error(c->mix);
        g_last_error = c->err;
    }
    return st;
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
    if (c->mix) p2de_batch_destroy(c->mix);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& l : c->launches) {
        (void)hipEventDestroy(l.a);
        (void)hipEventDestroy(l.b);
    }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    void* dptrs[] = {c->d_tab, c->d_geo, c->d_nb_node, c->d_nb_face, c->d_bc, c->d_Ival, c->d_U[0], c->d_U[1],
                     c->d_rhsUA, c->d_lA, c->d_dfA, c->d_rhsU, c->d_ut, c->d_theta, c->d_dtp, c->d_dbg_rhsH,
                     c->d_dbg_rhsL, c->d_dbg_L, c->d_blend, c->d_lphi, c->d_spart,
                     c->d_sc, c->d_keys, c->d_Lz, c->d_dtlog, c->d_sfac, c->d_rhoL};
    for (void* p : dptrs)
        if (p) (void)hipFree(p);
    if (c->h_sc) (void)hipHostFree(c->h_sc);
    if (c->h_keys) (void)hipHostFree(c->h_keys);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}


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
        int st = setup_config(cfg, &c->Nq, &c->Nfp, &c->Nh, &c->K, c->err);
        if (st) return bail(st);
        if (!ops->wq || !geom->Jq) return fail(c, P2DE_E_INVALID, "NULL wq or Jq"), bail(P2DE_E_INVALID);
        c->h_wq.assign(ops->wq, ops->wq + c->Nq);
        c->h_Jq.assign(geom->Jq, geom->Jq + (size_t)c->Nq * c->K);
        if (cfg->precision == P2DE_PRECISION_MIXED) {
            c->Nd = cfg->Nd;
            c->Nc = cfg->Nd + 2;
            c->N1D = cfg->N + 1;
            if ((st = p2de_batch_create(1, cfg, ops, geom, bc, &c->mix))) {
                c->err = p2de_batch_last_error(nullptr);
                return bail(st);
            }
            *out = c;
            return P2DE_OK;
        }
        c->Nd = cfg->Nd;
        c->Nc = cfg->Nd + 2;
        c->N1D = cfg->N + 1;
        if (ops->Nq != c->Nq || ops->Nfp != c->Nfp || ops->Nh != c->Nh)
            return fail(c, P2DE_E_INVALID, "operator sizes do not match Nd and N"), bail(P2DE_E_INVALID);
        const void* req[] = {ops->Srsh_db, ops->Srs0, ops->Srsh_nnz, ops->Srs0_nnz, ops->Brs, ops->MinvVhT,
                             ops->MinvVfT, ops->Vf, ops->Vf_low, ops->wq, ops->fq2q, ops->q2fq, geom->Jq, geom->GJh,
                             bc->mapP};
        for (const void* p : req)
            if (!p) return fail(c, P2DE_E_INVALID, "NULL operator, geometry or mapP array"), bail(P2DE_E_INVALID);
        c->basis = cfg->basis;
        c->ep_limit = cfg->entropyproj_limiter == P2DE_NODEWISE_SCALED_EXTRAPOLATION;
        const bool shock = cfg->shockcapture == P2DE_HENNEMANN_SHOCK_CAPTURE;
        const int ent = cfg->bound % 5;
        if ((shock || ent == 2 || ent == 4) && !ops->VDM_inv)
            return fail(c, P2DE_E_INVALID, "the smoothness indicator needs VDM_inv"), bail(P2DE_E_INVALID);
        c->optA = cfg->rhs_limiter == P2DE_ZHANG_SHU_LIMITER                        ? OPT_ZHANG_SHU
                  : (shock || cfg->bound != P2DE_POSITIVITY_BOUND) ? OPT_SUBCELL_GENERAL
                                                                    : OPT_SUBCELL;
        if ((st = setup_tables(c->tab, c->Nd, c->N1D, cfg->basis, ops, c->err))) return bail(st);
        MeshArrays mesh;
        if ((st = setup_mesh(c->Nd, c->N1D, c->K, c->tab, geom, bc, mesh, c->err))) return bail(st);
        const std::vector<Geo>& geo = mesh.geo;
        const std::vector<int>& nb_node = mesh.nb_node;
        const std::vector<int>& nb_face = mesh.nb_face;
        const std::vector<int>& bcode = mesh.bcode;
        const int K = c->K, Nq = c->Nq, Nfp = c->Nfp, Nc = c->Nc;

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
            (st = dalloc(c, &c->d_theta, (size_t)3 * K * Nfp)) || (st = dalloc(c, &c->d_blend, (size_t)K)) ||
            (st = dalloc(c, &c->d_lphi, (size_t)K * Nq)) || (st = dalloc(c, &c->d_spart, ((size_t)K * Nq + 255) / 256)) ||
            (st = dalloc(c, &c->d_dtp, (size_t)K)) ||
            (st = dalloc(c, &c->d_sc, 1)) || (st = dalloc(c, &c->d_keys, 3)))
            return bail(st);
        if (c->optA == OPT_ZHANG_SHU && (st = dalloc(c, &c->d_Lz, (size_t)3 * K))) return bail(st);
        if (cfg->bound % 5 == 4 && (st = dalloc(c, &c->d_sfac, (size_t)K))) return bail(st);
        if (cfg->bound >= P2DE_TVD_BOUND && (st = dalloc(c, &c->d_rhoL, (size_t)K * Nq))) return bail(st);
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
            (e = hipMemcpy(c->d_sc, c->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemset(c->d_theta, 0, sizeof(double) * 3 * K * Nfp)) != hipSuccess ||
            (c->d_Lz && (e = hipMemset(c->d_Lz, 0, sizeof(double) * 3 * K)) != hipSuccess))
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

int64_t p2de_num_dofs(const p2de_ctx* ctx) {
    if (ctx && ctx->mix) return p2de_batch_num_dofs(ctx->mix, 0);
    return ctx ? (int64_t)ctx->K * ctx->Nq : 0;
}

int p2de_device_state(p2de_ctx* ctx, void** dptr) {
    if (!ctx || !dptr) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (ctx->mix) return fwd(ctx, p2de_batch_device_state(ctx->mix, 0, dptr));
    *dptr = ctx->d_U[0];
    return P2DE_OK;
}

int p2de_set_state(p2de_ctx* ctx, const double* Uq) {
    if (!ctx || !Uq) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (ctx->mix) return fwd(ctx, p2de_batch_set_state(ctx->mix, 0, Uq));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_U[0], Uq, sizeof(double) * (size_t)ctx->K * ctx->Nq * ctx->Nc,
                               hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->state_set = true;
    ctx->ut_valid = false;
    return P2DE_OK;
}

int p2de_get_state(p2de_ctx* ctx, double* Uq) {
    if (!ctx || !Uq) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (ctx->mix) return fwd(ctx, p2de_batch_get_state(ctx->mix, 0, Uq));
    HIPCHK(ctx, hipMemcpyAsync(Uq, ctx->d_U[0], sizeof(double) * (size_t)ctx->K * ctx->Nq * ctx->Nc,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return P2DE_OK;
}

int p2de_set_debug(p2de_ctx* ctx, int on) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (ctx->mix) return fwd(ctx, p2de_batch_set_debug(ctx->mix, on));
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

double* theta_slot(p2de_ctx* c, int nstage) { return c->d_theta + (size_t)(nstage - 1) * c->K * c->Nfp; }

PhaseAArgs argsA(p2de_ctx* c, const double* U, int nstage) {
    const int stage1 = nstage == 1;
    PhaseAArgs a{};
    a.U = U;
    a.ut = c->d_ut;
    a.theta = c->ep_limit ? theta_slot(c, nstage) : nullptr;
    a.zsL = c->d_Lz ? c->d_Lz + (size_t)(nstage - 1) * c->K : nullptr;
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
    a.dtp = c->d_dtp;
    a.bound = c->cfg.bound;
    a.POSTOL = c->cfg.POSTOL;
    a.blend = c->cfg.shockcapture == P2DE_HENNEMANN_SHOCK_CAPTURE ? c->d_blend : nullptr;
    const bool min_entropy = c->cfg.bound % 5 == 1 || c->cfg.bound % 5 == 2;
    a.lphi = min_entropy ? c->d_lphi : nullptr;
    a.sfac = c->d_sfac;
    a.cell_beta = c->cfg.cell_beta;
    a.rhoL = c->d_rhoL;
    a.lo_proj = c->cfg.low_order_surface_flux == P2DE_LF_ON_PROJECTED_VAL;
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
    b.dbg_L = (c->debug && c->optA != OPT_ZHANG_SHU) ? c->d_dbg_L : nullptr;   // Zhang-Shu: L[k] stays
    b.ut = c->basis == P2DE_BASIS_GAUSS ? nullptr : c->d_ut;
    b.finalize = finalize;
    if (finalize && c->seg_interval > 0) {
        b.dt_log = c->d_dtlog;
        b.dt_cap = c->seg_cap;
        b.steps_base = c->seg_base;
        b.i0 = c->seg_i0;
        b.out_interval = c->seg_interval;
        b.T_end = c->cfg.T;
    }
    return b;
}

// Lobatto: projected face states of d_U[0] when they are not already current (phase B forms
// them for the states it writes).  Gauss: nothing, each stage projects its own input.
int ensure_face_states(p2de_ctx* c) {
    if (c->ut_valid || c->basis == P2DE_BASIS_GAUSS) return P2DE_OK;
    LAUNCHCHK(c, timed(c, 2, [&] {
                  return launch_face_state(c->Nd, c->N1D, c->d_U[0], c->d_tab, c->K, c->cfg.gamma, c->d_ut, c->stream);
              }));
    c->ut_valid = true;
    return P2DE_OK;
}

// Gauss: entropy projection of the stage input (init_rhs! and entropy_projection!, rhs.jl:19-23,
// 41-44): theta and u_tilde of every face node, before phase A reads them.
int project_faces(p2de_ctx* c, const double* U, int nstage) {
    if (c->basis != P2DE_BASIS_GAUSS) return P2DE_OK;
    EPArgs e{};
    e.U = U;
    e.T = c->d_tab;
    e.sc = c->d_sc;
    e.K = c->K;
    e.limit = c->ep_limit ? 1 : 0;
    e.gamma = c->cfg.gamma;
    e.zeta = c->cfg.zeta;
    e.eta = c->cfg.eta;
    e.POSTOL = c->cfg.POSTOL;
    e.ut = c->d_ut;
    e.theta = theta_slot(c, nstage);
    LAUNCHCHK(c, timed(c, 0, [&] { return launch_entropy_projection(c->Nd, c->N1D, e, c->stream); }));
    return P2DE_OK;
}

// Limiter options: the entropy minimum of the state at t0 (first stage), then the smoothness
// indicator, blending factor and entropy lower bounds of the stage input.  t is read from the
// device scalars when tp is set (steps), else it is t_value (p2de_rhs).
int limiter_prep(p2de_ctx* c, const double* U, int nstage, const double* tp, double t_value) {
    const p2de_config& g = c->cfg;
    const bool shock = g.shockcapture == P2DE_HENNEMANN_SHOCK_CAPTURE;
    const int ent = g.bound % 5;
    const bool min_entropy = ent == 1 || ent == 2;
    // no entropy bound and the plain cell-entropy bound read neither the indicator nor entropy
    // lower bounds
    if (!shock && (ent == 0 || ent == 3)) return P2DE_OK;
    if (min_entropy && nstage == 1)
        LAUNCHCHK(c, timed(c, 2, [&] {
                      return launch_s_min(U, (int64_t)c->K * c->Nq, c->Nd, g.gamma, tp, t_value, g.t0, c->d_spart, c->d_sc,
                                          c->stream);
                  }));
    PrepArgs p{};
    p.U = U;
    p.T = c->d_tab;
    p.nb_node = c->d_nb_node;
    p.sc = c->d_sc;
    p.K = c->K;
    p.N = c->N1D - 1;
    p.stage = nstage;
    p.shock = g.shockcapture;
    p.bound = g.bound;
    p.gamma = g.gamma;
    p.shock_a = g.shock_a;
    p.shock_c = g.shock_c;
    p.blend = c->d_blend;
    p.lphi = c->d_lphi;
    p.sfac = c->d_sfac;
    LAUNCHCHK(c, timed(c, 0, [&] { return launch_limiter_prep(c->Nd, c->N1D, p, c->stream); }));
    return P2DE_OK;
}

int stage_rhs(p2de_ctx* c, const double* U, int nstage, const double* tp, double t_value) {
    const int stage1 = nstage == 1;
    int st = project_faces(c, U, nstage);
    if (st) return st;
    if ((st = limiter_prep(c, U, nstage, tp, t_value))) return st;
    c->last_stage = nstage;
    if (c->cfg.bound >= P2DE_TVD_BOUND) {
        // TVD bounds: rho + dt*rhsL of every node first, since a node's bounds take the neighbour
        // element's nodes across the face (initialize_TVD_bounds! subcell.jl:112-141)
        PhaseAArgs pre = argsA(c, U, nstage);
        pre.tvd_pre = 1;
        pre.rhoL_out = c->d_rhoL;
        LAUNCHCHK(c, timed(c, 0, [&] { return launch_phaseA(c->Nd, c->N1D, c->basis, c->variantA, c->optA, pre, c->stream); }));
    }
    LAUNCHCHK(c, timed(c, 0, [&] {
                  return launch_phaseA(c->Nd, c->N1D, c->basis, c->variantA, c->optA, argsA(c, U, nstage), c->stream);
              }));
    if (stage1)
        LAUNCHCHK(c, timed(c, 2, [&] {
                      return launch_dt_reduce(c->d_sc, c->d_dtp, phaseA_blocks(c->Nd, c->N1D, c->basis, c->variantA, c->K),
                                              c->stream);
                  }));
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
    const double* tp = (const double*)((const char*)c->d_sc + offsetof(StepScalars, t));
    // stage 1: U1 = W + dt*rhsU (SSPRK33.jl:32-33)
    if ((st = stage_rhs(c, A, 1, tp, 0.0))) return st;
    LAUNCHCHK(c, timed(c, 1, [&] { return launch_phaseB(c->Nd, c->N1D, argsB(c, MODE_STAGE1, A, A, B, 0), s); }));
    // stage 2: U2 = 3/4 W + 1/4 (U1 + dt*rhsU) (SSPRK33.jl:34-36)
    if ((st = stage_rhs(c, B, 2, tp, 0.0))) return st;
    LAUNCHCHK(c, timed(c, 1, [&] { return launch_phaseB(c->Nd, c->N1D, argsB(c, MODE_STAGE2, A, B, B, 0), s); }));
    // stage 3: U^{n+1} = 1/3 W + 2/3 (U2 + dt*rhsU) (SSPRK33.jl:37-39), t += dt
    if ((st = stage_rhs(c, B, 3, tp, 0.0))) return st;
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
    if (ctx->mix) return fwd(ctx, p2de_batch_rhs(ctx->mix, 0, t, dt, nstage, dt_out, rhsU));
    if (nstage < 1 || nstage > 3) return fail(ctx, P2DE_E_INVALID, "nstage must be 1, 2 or 3");
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    void* s = ctx->stream;
    const p2de_config& g = ctx->cfg;
    // the stage-1 CFL reduction starts from min(CFL*dt0, T - t) (low_order_graph_viscosity.jl:230)
    const double cap = nstage == 1 ? jmin_h(g.CFL * g.dt0, g.T - t) : dt;
    LAUNCHCHK(ctx, timed(ctx, 2, [&] { return launch_rhs_init(ctx->d_sc, dt, cap, s); }));
    int st0 = ensure_face_states(ctx);
    if (st0) return st0;
    if ((st0 = stage_rhs(ctx, ctx->d_U[0], nstage, nullptr, t))) return st0;
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
    const size_t K = (size_t)ctx->K;
    if (field == P2DE_FIELD_L || field == P2DE_FIELD_THETA_MEAN) {
        // the reference writes L with the Zhang-Shu limiter only (zhangshu.jl:44) and theta with
        // NodewiseScaledExtrapolation only (filter.jl:58); both start at zero (init.jl:20-22)
        const bool have = field == P2DE_FIELD_L ? ctx->d_Lz != nullptr : (ctx->ep_limit && !ctx->mix);
        if (!have) {
            std::fill(out, out + 3 * K, 0.0);
            return P2DE_OK;
        }
        if (field == P2DE_FIELD_L) {
            HIPCHK(ctx, hipMemcpyAsync(out, ctx->d_Lz, sizeof(double) * 3 * K, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            return P2DE_OK;
        }
        const int Nfp = ctx->Nfp;
        std::vector<double> th((size_t)3 * K * Nfp);
        HIPCHK(ctx, hipMemcpyAsync(th.data(), ctx->d_theta, sizeof(double) * th.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t sk = 0; sk < 3 * K; sk++) {          // theta[k, nstage] = sum(theta_local[:, k, nstage]) / Nfp
            double sum = 0.0;
            for (int i = 0; i < Nfp; i++) sum += th[sk * Nfp + i];
            out[sk] = sum / Nfp;
        }
        return P2DE_OK;
    }
    if (ctx->mix) return fwd(ctx, p2de_batch_get_field(ctx->mix, 0, field, out));
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
        case P2DE_FIELD_THETA:
            if (!ctx->ep_limit)
                return fail(ctx, P2DE_E_INVALID, "theta exists with NodewiseScaledExtrapolation only");
            src = theta_slot(ctx, std::max(ctx->last_stage, 1));
            n = (size_t)ctx->Nfp * ctx->K;
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
    if (ctx->mix) return fwd(ctx, p2de_batch_set_time(ctx->mix, 0, t));
    int st = read_scalars(ctx);
    if (st) return st;
    ctx->h_sc->t = t;
    ctx->h_sc->done = 0;
    ctx->h_sc->hold = 0;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_sc, ctx->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return P2DE_OK;
}

int p2de_get_time(p2de_ctx* ctx, double* t, double* last_dt) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (ctx->mix) return fwd(ctx, p2de_batch_get_time(ctx->mix, 0, t, last_dt, nullptr));
    int st = read_scalars(ctx);
    if (st) return st;
    if (t) *t = ctx->h_sc->t;
    if (last_dt) *last_dt = bits_to_double_h(ctx->h_sc->dt_bits);
    return P2DE_OK;
}

int p2de_ssprk3_step(p2de_ctx* ctx, double* t_inout, double* dt_out) {
    if (!ctx || !t_inout) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (ctx->mix) {
        int n0 = 0, n1 = 0, st;
        double dt = 0.0;
        if ((st = fwd(ctx, p2de_batch_set_time(ctx->mix, 0, *t_inout))) ||
            (st = fwd(ctx, p2de_batch_get_time(ctx->mix, 0, nullptr, nullptr, &n0))) ||
            (st = fwd(ctx, p2de_batch_enqueue_steps(ctx->mix, 1))) ||
            (st = fwd(ctx, p2de_batch_get_time(ctx->mix, 0, t_inout, &dt, &n1))))
            return st;
        if (dt_out) *dt_out = n1 > n0 ? dt : 0.0;
        return P2DE_OK;
    }
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
    if (ctx->mix) return fwd(ctx, p2de_batch_enqueue_steps(ctx->mix, steps));
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    for (int i = 0; i < steps; i++) {
        int st = enqueue_step(ctx);
        if (st) return st;
    }
    return P2DE_OK;
}

int p2de_synchronize(p2de_ctx* ctx) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (ctx->mix) return fwd(ctx, p2de_batch_synchronize(ctx->mix));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->timing) harvest_timing(ctx);
    return P2DE_OK;
}

int p2de_run(p2de_ctx* ctx, int max_steps, p2de_run_stats* stats) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (ctx->mix) {
        int n0 = 0, n1 = 0, taken = 0, st;
        if ((st = fwd(ctx, p2de_batch_get_time(ctx->mix, 0, nullptr, nullptr, &n0))) ||
            (st = fwd(ctx, p2de_batch_run(ctx->mix, max_steps, &taken))))
            return st;
        if (stats) {
            int64_t viol = 0;
            if ((st = fwd(ctx, batch_instance_stats(ctx->mix, 0, &stats->t, &stats->last_dt, &n1, &viol)))) return st;
            stats->steps = n1 - n0;
            stats->violations = viol;
        }
        return P2DE_OK;
    }
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

int p2de_run_segment(p2de_ctx* ctx, int output_interval, int max_steps, int64_t* step_index, double* dthist,
                     int dthist_cap, int* steps_taken, int* at_output) {
    if (!ctx || !step_index || !dthist || !steps_taken || !at_output)
        return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (output_interval < 1) return fail(ctx, P2DE_E_INVALID, "output_interval must be >= 1");
    if (dthist_cap < 1) return fail(ctx, P2DE_E_INVALID, "dthist_cap must be >= 1");
    const int limit = max_steps > 0 ? std::min(max_steps, dthist_cap) : dthist_cap;
    const double T = ctx->cfg.T;
    // SSPRK33.jl:42-46: after a step, i = i + 1; a history point when mod(i, output_interval) == 0
    // or |t - T| < 1e-10
    auto history_point = [&](long long i, double t) { return i % output_interval == 0 || std::fabs(t - T) < 1e-10; };
    *steps_taken = 0;
    *at_output = 0;
    if (ctx->mix) {
        // the batched kernels keep no history: one step per call
        long long i = *step_index;
        for (int n = 0; n < limit; n++) {
            int n0 = 0, n1 = 0, st;
            double t = 0.0, dt = 0.0;
            if ((st = fwd(ctx, p2de_batch_get_time(ctx->mix, 0, nullptr, nullptr, &n0))) ||
                (st = fwd(ctx, p2de_batch_enqueue_steps(ctx->mix, 1))) ||
                (st = fwd(ctx, p2de_batch_get_time(ctx->mix, 0, &t, &dt, &n1))))
                return st;
            if (n1 == n0) break;
            dthist[n] = dt;
            *steps_taken = n + 1;
            *step_index = ++i;
            if (history_point(i, t)) {
                *at_output = 1;
                break;
            }
        }
        return P2DE_OK;
    }
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    if (limit > ctx->dtlog_cap) {
        if (ctx->d_dtlog) {
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->d_dtlog);
            ctx->d_dtlog = nullptr;
            ctx->dtlog_cap = 0;
        }
        int st = dalloc(ctx, &ctx->d_dtlog, (size_t)limit);
        if (st) return st;
        ctx->dtlog_cap = limit;
    }
    int st = read_scalars(ctx);
    if (st) return st;
    ctx->h_sc->hold = 0;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_sc, ctx->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice, ctx->stream));
    ctx->seg_interval = output_interval;
    ctx->seg_cap = limit;
    ctx->seg_base = ctx->h_sc->steps;
    ctx->seg_i0 = *step_index;
    // steps after the history point (or after the limit) find hold set and do nothing
    const int chunk = 32;
    int launched = 0;
    while (launched < limit) {
        const int n = std::min(chunk, limit - launched);
        for (int j = 0; j < n && !st; j++) st = enqueue_step(ctx);
        if (!st) st = read_scalars(ctx);
        if (st) break;
        launched += n;
        if (ctx->h_sc->hold || ctx->h_sc->done || !(ctx->h_sc->t < T)) break;
    }
    ctx->seg_interval = 0;
    if (st) return st;
    const int taken = ctx->h_sc->steps - ctx->seg_base;
    if (taken > 0)
        HIPCHK(ctx, hipMemcpyAsync(dthist, ctx->d_dtlog, sizeof(double) * taken, hipMemcpyDeviceToHost, ctx->stream));
    ctx->h_sc->hold = 0;
    ctx->h_sc->done = 0;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_sc, ctx->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *steps_taken = taken;
    *step_index = ctx->seg_i0 + taken;
    *at_output = taken > 0 && history_point(*step_index, ctx->h_sc->t);
    return P2DE_OK;
}

int p2de_check_conservation(p2de_ctx* ctx, double* total) {
    if (!ctx || !total) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    const int Nq = ctx->Nq, Nc = ctx->Nc, K = ctx->K;
    std::vector<double> U((size_t)K * Nq * Nc);
    int st = p2de_get_state(ctx, U.data());
    if (st) return st;
    // utils.jl:5-11: total += J[i, k] * wq[i] * Uq[i, k] over k, then i; then the sum of the components
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < K; k++)
        for (int i = 0; i < Nq; i++) {
            const double w = ctx->h_Jq[(size_t)k * Nq + i] * ctx->h_wq[i];
            const double* u = &U[((size_t)k * Nq + i) * Nc];
            for (int c = 0; c < Nc; c++) acc[c] = acc[c] + w * u[c];
        }
    double sum = acc[0];
    for (int c = 1; c < Nc; c++) sum = sum + acc[c];
    *total = sum;
    return P2DE_OK;
}

namespace {

// Checkpoint layout: header, then the state double[K][Nq][Nc]; integers and doubles little-endian.
struct CheckpointHeader {
    char magic[8];            // "P2DECKPT"
    uint32_t version;         // 1
    uint32_t header_bytes;
    int32_t Nd, N, Nc, basis, precision, reserved;
    int64_t K, Nq;
    double t, last_dt, s_min;
    int64_t steps, step_index;
    uint64_t payload_bytes, checksum;
};
constexpr char kCkptMagic[8] = {'P', '2', 'D', 'E', 'C', 'K', 'P', 'T'};

// FNV-1a over the payload bytes
uint64_t fnv1a64(const unsigned char* p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

}  // namespace

int64_t p2de_checkpoint_size(const p2de_ctx* ctx) {
    if (!ctx) return 0;
    return (int64_t)sizeof(CheckpointHeader) + (int64_t)sizeof(double) * ctx->K * ctx->Nq * ctx->Nc;
}

int p2de_checkpoint_save(p2de_ctx* ctx, int64_t step_index, void* buf, int64_t size) {
    if (!ctx || !buf) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (size < p2de_checkpoint_size(ctx)) return fail(ctx, P2DE_E_INVALID, "buffer smaller than p2de_checkpoint_size");
    CheckpointHeader h{};
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = 1;
    h.header_bytes = sizeof(CheckpointHeader);
    h.Nd = ctx->Nd;
    h.N = ctx->N1D - 1;
    h.Nc = ctx->Nc;
    h.basis = ctx->cfg.basis;
    h.precision = ctx->cfg.precision;
    h.K = ctx->K;
    h.Nq = ctx->Nq;
    h.step_index = step_index;
    double* U = (double*)((char*)buf + sizeof(CheckpointHeader));
    int st = p2de_get_state(ctx, U);
    if (st) return st;
    if (ctx->mix) {
        int steps = 0;
        if ((st = fwd(ctx, batch_instance_stats(ctx->mix, 0, &h.t, &h.last_dt, &steps, nullptr)))) return st;
        h.steps = steps;
        h.s_min = 0.0;
    } else {
        if ((st = read_scalars(ctx))) return st;
        h.t = ctx->h_sc->t;
        h.last_dt = bits_to_double_h(ctx->h_sc->dt_bits);
        h.s_min = ctx->h_sc->s_min;
        h.steps = ctx->h_sc->steps;
    }
    h.payload_bytes = sizeof(double) * (uint64_t)ctx->K * ctx->Nq * ctx->Nc;
    h.checksum = fnv1a64((const unsigned char*)U, h.payload_bytes);
    std::memcpy(buf, &h, sizeof(h));
    return P2DE_OK;
}

int p2de_checkpoint_load(p2de_ctx* ctx, const void* buf, int64_t size, int64_t* step_index) {
    if (!ctx || !buf) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (size < (int64_t)sizeof(CheckpointHeader)) return fail(ctx, P2DE_E_INVALID, "checkpoint truncated (header)");
    CheckpointHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, kCkptMagic, 8) != 0) return fail(ctx, P2DE_E_INVALID, "not a checkpoint (magic)");
    if (h.version != 1 || h.header_bytes != sizeof(CheckpointHeader))
        return fail(ctx, P2DE_E_INVALID, "unsupported checkpoint version");
    if (h.Nd != ctx->Nd || h.N != ctx->N1D - 1 || h.Nc != ctx->Nc || h.basis != ctx->cfg.basis || h.K != ctx->K ||
        h.Nq != ctx->Nq)
        return fail(ctx, P2DE_E_INVALID, "checkpoint of another problem (dimension, degree, basis or mesh size)");
    const uint64_t payload = sizeof(double) * (uint64_t)ctx->K * ctx->Nq * ctx->Nc;
    if (h.payload_bytes != payload || (uint64_t)size < sizeof(CheckpointHeader) + payload)
        return fail(ctx, P2DE_E_INVALID, "checkpoint truncated (state)");
    const double* U = (const double*)((const char*)buf + sizeof(CheckpointHeader));
    if (fnv1a64((const unsigned char*)U, payload) != h.checksum)
        return fail(ctx, P2DE_E_INVALID, "checkpoint checksum mismatch");
    int st = p2de_set_state(ctx, U);
    if (st) return st;
    if (ctx->mix) {
        if ((st = fwd(ctx, batch_instance_restore(ctx->mix, 0, h.t, h.last_dt, (int)h.steps)))) return st;
    } else {
        if ((st = read_scalars(ctx))) return st;
        ctx->h_sc->t = h.t;
        std::memcpy(&ctx->h_sc->dt_bits, &h.last_dt, sizeof(double));
        ctx->h_sc->s_min = h.s_min;
        ctx->h_sc->steps = (int)h.steps;
        ctx->h_sc->done = 0;
        ctx->h_sc->hold = 0;
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_sc, ctx->h_sc, sizeof(StepScalars), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (step_index) *step_index = h.step_index;
    return P2DE_OK;
}

int p2de_check_positivity(p2de_ctx* ctx, p2de_pos_report* rep) {
    if (!ctx || !rep) return fail(ctx, P2DE_E_INVALID, "NULL argument");
    if (ctx->mix) return fwd(ctx, p2de_batch_check_positivity(ctx->mix, 0, rep));
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
    if (ctx->mix) return fwd(ctx, p2de_batch_set_kernel_timing(ctx->mix, on));
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
    if (ctx->mix) return fwd(ctx, p2de_batch_get_kernel_timing(ctx->mix, ms, count));
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
    o << "Nd" << c->Nd << "_N" << (c->N1D - 1) << "_K" << c->K << (c->basis == P2DE_BASIS_GAUSS ? "_gauss" : "");
    return o.str();
}

// Tuning cache: one "key variant" pair per line.
// Autotuner cache: one "key value" line per problem (the element-kernel shape index here; the
// batched path stores its launch shape), read line by line so that either kind of line is kept.
std::map<std::string, std::string> read_cache(const char* path) {
    std::map<std::string, std::string> m;
    if (!path) return m;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        std::string k, v;
        if (ls >> k >> v) m[k] = v;
    }
    return m;
}

void write_cache(const char* path, const std::map<std::string, std::string>& m) {
    if (!path) return;
    std::ofstream f(path, std::ios::trunc);
    for (auto& kv : m) f << kv.first << " " << kv.second << "\n";
}

// the cached shape index, or -1 when the entry is absent or not an index
int cached_index(const std::map<std::string, std::string>& m, const std::string& key) {
    const auto it = m.find(key);
    if (it == m.end() || it->second.empty()) return -1;
    char* end = nullptr;
    const long v = std::strtol(it->second.c_str(), &end, 10);
    return (*end == '\0' && v >= 0 && v < 1000000) ? (int)v : -1;
}

}  // namespace

extern "C" {

int p2de_num_variants(const p2de_ctx* ctx) {
    if (ctx && ctx->mix) return p2de_batch_num_variants(ctx->mix);
    return ctx ? num_variants_A(ctx->Nd, ctx->N1D, ctx->basis) : 0;
}

int p2de_set_variant(p2de_ctx* ctx, int variant) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (ctx->mix) return fwd(ctx, p2de_batch_set_variant(ctx->mix, variant));
    if (variant < 0 || variant >= num_variants_A(ctx->Nd, ctx->N1D, ctx->basis)) return fail(ctx, P2DE_E_INVALID, "variant out of range");
    ctx->variantA = variant;
    return P2DE_OK;
}

int p2de_get_variant(const p2de_ctx* ctx, int* variant, int* E, int* NT) {
    if (!ctx) return P2DE_E_INVALID;
    if (ctx->mix) {
        int Ev[6] = {0, 0, 0, 0, 0, 0};
        const int st = p2de_batch_get_variant(ctx->mix, variant, NT, Ev);
        if (E) *E = Ev[ctx->N1D];
        return st;
    }
    int e = 0, nt = 0;
    variant_shape_A(ctx->Nd, ctx->N1D, ctx->basis, ctx->variantA, &e, &nt);
    if (variant) *variant = ctx->variantA;
    if (E) *E = e;
    if (NT) *NT = nt;
    return P2DE_OK;
}

int p2de_autotune(p2de_ctx* ctx, int steps, const char* cache_path, int* chosen) {
    if (!ctx) return fail(ctx, P2DE_E_INVALID, "NULL context");
    if (ctx->mix) return fwd(ctx, p2de_batch_autotune(ctx->mix, steps, cache_path, chosen));
    if (!ctx->state_set) return fail(ctx, P2DE_E_STATE, "no state uploaded (p2de_set_state)");
    const std::string key = tune_key(ctx);
    auto cache = read_cache(cache_path);
    const int nv = num_variants_A(ctx->Nd, ctx->N1D, ctx->basis);
    const int cv = cached_index(cache, key);
    if (cv >= 0 && cv < nv) {
        ctx->variantA = cv;
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
    cache[key] = std::to_string(best);
    write_cache(cache_path, cache);
    if (chosen) *chosen = best;
    return P2DE_OK;
}

}  // extern "C"
