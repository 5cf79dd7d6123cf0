// p2de_batch.cpp -- C-ABI of the batched mixed-precision path (include/p2de_hip.h, p2de_batch_*):
// instance set-up and validation, the tile list, device memory, the per-step launch sequence
// (captured once in a hipGraph and replayed), kernel timing and the launch-shape autotuner.
// No C++ exception crosses the ABI; every failure becomes a status and a message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "p2de_hip.h"
#include "p2de_internal.h"
#include "p2de_mixed.h"
#include "p2de_setup.h"

using namespace p2de;

namespace {

thread_local std::string g_batch_error;

struct BInst {
    p2de_config cfg;
    int N1D, Nq, Nfp, K;
    int64_t node_off;      // first node in the concatenated state
    int64_t elem_off;      // first element in the concatenated per-element arrays
    int64_t face_off;      // first face node in the concatenated per-face arrays
    int64_t ival_off;      // first Ival row
};

struct BTimed {
    int kind;
    hipEvent_t a, b;
};

}  // namespace

struct p2de_batch {
    int n = 0;
    std::vector<BInst> inst;
    int64_t total_nodes = 0, total_elems = 0, total_faces = 0;
    bool present[MIX_MAX_N1D + 1] = {false, false, false, false, false, false};
    // device
    MixInst* d_inst = nullptr;
    MixTile* d_tiles = nullptr;
    int ntiles = 0;
    int chunk[MIX_CHUNKS + 1] = {};            // XCD chunk boundaries in the tile list
    double* d_dtp = nullptr;                   // per-tile CFL minima
    std::vector<MixInst> hinst;                // host copy of d_inst
    void* d_tab[MIX_MAX_N1D + 1] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    Geo* d_geo = nullptr;
    int *d_nb_node = nullptr, *d_bc = nullptr;
    double* d_Ival = nullptr;
    double* d_U[3] = {nullptr, nullptr, nullptr};
    double *d_rhsU = nullptr, *d_dbg_rhsH = nullptr, *d_dbg_rhsL = nullptr;
    StepScalars* d_sc = nullptr;
    StepScalars* h_sc = nullptr;               // pinned, n entries
    unsigned long long* d_keys = nullptr;
    unsigned long long* h_keys = nullptr;      // pinned
    hipStream_t stream = nullptr;
    int cus = 256;
    std::vector<int> variants;                 // applicable global variant indices
    int vsel = 0;                              // index into variants
    int grid_stage = 1, grid_dt = 1;
    bool state_set = false;
    bool debug = false;
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool timing = false;
    std::vector<BTimed> launches;
    std::vector<hipEvent_t> event_pool;
    double time_ms[3] = {0, 0, 0};
    int64_t time_count[3] = {0, 0, 0};
    std::string err;
};

namespace {

int bfail(p2de_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    g_batch_error = msg;
    return code;
}

#define BHIP(b, expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return bfail(b, P2DE_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define BLAUNCH(b, expr)                                                                              \
    do {                                                                                              \
        int e_ = (expr);                                                                              \
        if (e_ != 0)                                                                                  \
            return bfail(b, P2DE_E_HIP, std::string(#expr) + ": " + hipGetErrorString((hipError_t)e_)); \
    } while (0)

template <class T>
int balloc(p2de_batch* b, T** p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return bfail(b, P2DE_E_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    return P2DE_OK;
}

void drop_graph(p2de_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

void free_batch(p2de_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (auto& l : b->launches) {
        (void)hipEventDestroy(l.a);
        (void)hipEventDestroy(l.b);
    }
    for (auto e : b->event_pool) (void)hipEventDestroy(e);
    void* dptrs[] = {b->d_inst, b->d_tiles, b->d_dtp, b->d_geo, b->d_nb_node, b->d_bc, b->d_Ival,
                     b->d_U[0], b->d_U[1], b->d_U[2], b->d_rhsU, b->d_dbg_rhsH, b->d_dbg_rhsL, b->d_sc, b->d_keys};
    for (void* p : dptrs)
        if (p) (void)hipFree(p);
    for (void* p : b->d_tab)
        if (p) (void)hipFree(p);
    if (b->h_sc) (void)hipHostFree(b->h_sc);
    if (b->h_keys) (void)hipHostFree(b->h_keys);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

inline double cmat(const double* A, int64_t ld, int64_t i, int64_t j) { return A[i + ld * j]; }

// MixTab<N1D> from the validated tables and the caller's dense operators (the pair layout of
// the LGL tensor-product operators, init.jl:133-228, is checked entry by entry).
template <int N1D>
int build_mixtab(const ElemTables& T, const p2de_operators* ops, MixTab<N1D>& M, std::string& err) {
    using MT = MixTab<N1D>;
    constexpr int NQ = MT::NQ, NFP = MT::NFP, NP1 = MT::NP1, NH = NQ + NFP;
    std::memset(&M, 0, sizeof(M));
    auto node = [](int d, int line, int s) { return d == 0 ? s + N1D * line : line + N1D * s; };
    // volume pairs and the dense Srsh_db blocks
    std::vector<char> covered((size_t)NH * NH, 0);
    for (int d = 0; d < 2; d++)
        for (int line = 0; line < N1D; line++)
            for (int a = 1; a < N1D; a++)
                for (int bb = 0; bb < a; bb++) {
                    const int i = node(d, line, a), j = node(d, line, bb);
                    const int p = d * N1D * NP1 + line * NP1 + a * (a - 1) / 2 + bb;
                    const double c = cmat(ops->Srsh_db + (int64_t)d * NH * NH, NH, i, j);
                    const double o = cmat(ops->Srsh_db + (int64_t)(1 - d) * NH * NH, NH, i, j);
                    if (o != 0.0) {
                        err = "a Srsh line pair acts in both directions";
                        return P2DE_E_UNSUPPORTED;
                    }
                    M.cH[p] = (float)c;
                    covered[(size_t)i * NH + j] = 1;
                }
    // every other Srsh pair must be a face-volume pair whose contribution cancels:
    // MinvVhT[i, Nq+f] == MinvVhT[i, i] at i = fq2q[f]
    for (int p = 0; p < T.nph; p++) {
        const int i = T.hp_i[p], j = T.hp_j[p];
        if (covered[(size_t)i * NH + j]) continue;
        if (i < NQ || j != T.fq2q[i - NQ]) {
            err = "Srsh_nnz holds a pair outside the tensor-product line structure";
            return P2DE_E_UNSUPPORTED;
        }
        const int f = i - NQ, q = j;
        bool ok = false;
        for (int m = 0; m < T.nf[q]; m++)
            if (T.nface[q][m] == f && T.a_h[q][m] == T.a_v[q]) ok = true;
        if (!ok) {
            err = "face-volume pairs do not cancel (MinvVhT[i, Nq+f] != MinvVhT[i, i])";
            return P2DE_E_UNSUPPORTED;
        }
    }
    // low-order pairs: line neighbours only
    for (int p = 0; p < T.npl; p++) {
        const int i = T.lp_i[p], j = T.lp_j[p];
        const int d = T.lp_dir[p];
        const bool nbr = d == 0 ? (i == j + 1 && i % N1D != 0) : (i == j + N1D);
        if (!nbr) {
            err = "Srs0_nnz holds a pair that is not a line neighbour";
            return P2DE_E_UNSUPPORTED;
        }
    }
    for (int d = 0; d < 2; d++)
        for (int line = 0; line < N1D; line++)
            for (int s = 0; s + 1 < N1D; s++) {
                const int i = node(d, line, s + 1), j = node(d, line, s);
                M.cL[d * N1D * (N1D - 1) + line * (N1D - 1) + s] = (float)cmat(ops->Srs0 + (int64_t)d * NQ * NQ, NQ, i, j);
            }
    for (int i = 0; i < NQ; i++) {
        M.wq[i] = T.wq[i];
        M.a_v[i] = (float)T.a_v[i];
        for (int m = 0; m < T.nf[i]; m++) {
            const int f = T.nface[i][m];
            if (f < 2 * N1D) M.a_fx[i] = (float)T.a_f[i][m];
            else M.a_fy[i] = (float)T.a_f[i][m];
        }
    }
    for (int f = 0; f < NFP; f++) {
        M.Bf[f] = (float)(f < 2 * N1D ? T.Br[f] : T.Bs[f]);
        M.fq[f] = (uint8_t)T.fq2q[f];
        if (T.vf_coef[f] != 1.0) {
            err = "Vf is not a unit face-node selection";
            return P2DE_E_UNSUPPORTED;
        }
    }
    // index tables (layout in p2de_mixed.h)
    auto pack = [](int a, int b, int c) { return (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16); };
    for (int d = 0; d < 2; d++)
        for (int line = 0; line < N1D; line++) {
            for (int a = 1; a < N1D; a++)
                for (int bb = 0; bb < a; bb++)
                    M.pH[d * N1D * NP1 + line * NP1 + a * (a - 1) / 2 + bb] = pack(node(d, line, a), node(d, line, bb), d);
            for (int s = 0; s + 1 < N1D; s++)
                M.pL[d * N1D * (N1D - 1) + line * (N1D - 1) + s] = pack(node(d, line, s + 1), node(d, line, s), d);
            for (int s = 0; s < N1D; s++) M.lq[d * N1D + line][s] = (uint16_t)node(d, line, s);
        }
    for (int i = 0; i < NQ; i++) {
        const int ia = i % N1D, ib = i / N1D;
        int m = 0;
        for (int o = 0; o < N1D; o++) {
            if (o == ia) continue;
            const int hi = std::max(ia, o), lo = std::min(ia, o);
            M.gx[i][m++] = (uint16_t)(ib * NP1 + hi * (hi - 1) / 2 + lo) | (ia > o ? 0 : 0x8000);
        }
        m = 0;
        for (int o = 0; o < N1D; o++) {
            if (o == ib) continue;
            const int hi = std::max(ib, o), lo = std::min(ib, o);
            M.gy[i][m++] = (uint16_t)(N1D * NP1 + ia * NP1 + hi * (hi - 1) / 2 + lo) | (ib > o ? 0 : 0x8000);
        }
        M.nfc[i] = (uint16_t)((ib * (N1D + 1) + ia) | (((N1D + ia) * (N1D + 1) + ib) << 8));
    }
    return P2DE_OK;
}

int upload_tab(p2de_batch* b, int N1D, const ElemTables& T, const p2de_operators* ops) {
    std::string e;
    int st = P2DE_OK;
    const void* host = nullptr;
    size_t bytes = 0;
    MixTab<2> t2;
    MixTab<3> t3;
    MixTab<4> t4;
    MixTab<5> t5;
    switch (N1D) {
        case 2: st = build_mixtab<2>(T, ops, t2, e), host = &t2, bytes = sizeof(t2); break;
        case 3: st = build_mixtab<3>(T, ops, t3, e), host = &t3, bytes = sizeof(t3); break;
        case 4: st = build_mixtab<4>(T, ops, t4, e), host = &t4, bytes = sizeof(t4); break;
        case 5: st = build_mixtab<5>(T, ops, t5, e), host = &t5, bytes = sizeof(t5); break;
        default: return bfail(b, P2DE_E_UNSUPPORTED, "degree");
    }
    if (st) return bfail(b, st, e);
    if (!b->d_tab[N1D]) {
        BHIP(b, hipMalloc(&b->d_tab[N1D], bytes));
        BHIP(b, hipMemcpy(b->d_tab[N1D], host, bytes, hipMemcpyHostToDevice));
        return P2DE_OK;
    }
    // instances of one degree share the table: it must be the same
    std::vector<unsigned char> cur(bytes);
    BHIP(b, hipMemcpy(cur.data(), b->d_tab[N1D], bytes, hipMemcpyDeviceToHost));
    if (std::memcmp(cur.data(), host, bytes) != 0)
        return bfail(b, P2DE_E_UNSUPPORTED, "instances of one degree pass different operators");
    return P2DE_OK;
}

// tiles of E_N consecutive elements, instance by instance
int build_tiles(p2de_batch* b) {
    int NT = 0, E[MIX_MAX_N1D + 1];
    mix_variant_shape(b->variants[b->vsel], &NT, E);
    std::vector<MixTile> tiles;
    for (int i = 0; i < b->n; i++) {
        const BInst& I = b->inst[i];
        const int e = E[I.N1D];
        b->hinst[i].tile_lo = (int)tiles.size();
        for (int k0 = 0; k0 < I.K; k0 += e) tiles.push_back(MixTile{i, k0, std::min(e, I.K - k0), I.N1D});
        b->hinst[i].tile_hi = (int)tiles.size();
    }
    b->ntiles = (int)tiles.size();
    // XCD chunks of about equal cost: DOFs (the fused stage kernel takes about the same time per
    // DOF at degrees 2, 3 and 4; on the 48-instance batch, weighting the degree-4 tiles by 1.3
    // made a step 4% slower, and any weight other than 1 was slower)
    {
        std::vector<double> cost(tiles.size() + 1, 0.0);
        for (size_t t = 0; t < tiles.size(); t++) {
            const int n1d = tiles[t].n1d;
            cost[t + 1] = cost[t] + (double)tiles[t].nel * n1d * n1d;
        }
        b->chunk[0] = 0;
        size_t t = 0;
        for (int c = 1; c < MIX_CHUNKS; c++) {
            const double target = cost.back() * c / MIX_CHUNKS;
            while (t < tiles.size() && cost[t + 1] <= target) t++;
            b->chunk[c] = (int)t;
        }
        b->chunk[MIX_CHUNKS] = b->ntiles;
    }
    BHIP(b, hipMemcpyAsync(b->d_tiles, tiles.data(), sizeof(MixTile) * tiles.size(), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_inst, b->hinst.data(), sizeof(MixInst) * b->n, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    int sb = 0, db = 0;
    BLAUNCH(b, mix_blocks_per_cu(b->variants[b->vsel], &sb, &db));
    sb = std::max(sb, 1);
    db = std::max(db, 1);
    b->grid_stage = std::max(1, std::min(b->ntiles, sb * b->cus));
    b->grid_dt = std::max(1, std::min(b->ntiles, db * b->cus));
    drop_graph(b);
    return P2DE_OK;
}

hipEvent_t take_event(p2de_batch* b) {
    if (!b->event_pool.empty()) {
        hipEvent_t e = b->event_pool.back();
        b->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

template <class F>
int timed(p2de_batch* b, int kind, F&& launch) {
    if (!b->timing) return launch();
    BTimed t{kind, take_event(b), take_event(b)};
    (void)hipEventRecord(t.a, b->stream);
    int r = launch();
    (void)hipEventRecord(t.b, b->stream);
    b->launches.push_back(t);
    return r;
}

void harvest(p2de_batch* b) {
    for (auto& l : b->launches) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, l.a, l.b) == hipSuccess) {
            b->time_ms[l.kind] += ms;
            b->time_count[l.kind] += 1;
        }
        b->event_pool.push_back(l.a);
        b->event_pool.push_back(l.b);
    }
    b->launches.clear();
}

MixArgs margs(p2de_batch* b, int mode, int only, const double* Uin, const double* W, double* Uout) {
    MixArgs a{};
    a.inst = b->d_inst;
    a.tiles = b->d_tiles;
    a.ntiles = b->ntiles;
    for (int c = 0; c <= MIX_CHUNKS; c++) a.chunk[c] = b->chunk[c];
    a.dtp = b->d_dtp;
    for (int n = 0; n <= MIX_MAX_N1D; n++) a.tab[n] = b->d_tab[n];
    a.mode = mode;
    a.only = only;
    a.Uin = Uin;
    a.W = W;
    a.Uout = Uout;
    a.rhsU = b->d_rhsU;
    if (b->debug && mode == MODE_RHS) {
        a.dbg_rhsH = b->d_dbg_rhsH;
        a.dbg_rhsL = b->d_dbg_rhsL;
    }
    return a;
}

// One SSP-RK3 step of every instance (SSPRK33.jl:29-43): A holds U^n, B and C the stages.  The
// last launch also starts the next step (done flag, dt cap), so the first step of a sequence
// needs `with_init`; repeating the start is harmless.
int launch_step(p2de_batch* b, bool with_init) {
    void* s = b->stream;
    const int v = b->variants[b->vsel];
    double *A = b->d_U[0], *B = b->d_U[1], *C = b->d_U[2];
    if (with_init) BLAUNCH(b, timed(b, 2, [&] { return mix_launch_init(b->d_inst, b->n, 0, -1, 0.0, 0.0, s); }));
    BLAUNCH(b, timed(b, 1, [&] { return mix_launch_dt(v, margs(b, MODE_STAGE1, -1, A, A, nullptr), b->grid_dt, s); }));
    BLAUNCH(b, timed(b, 2, [&] { return mix_launch_dt_reduce(b->d_inst, b->n, b->d_dtp, -1, s); }));
    BLAUNCH(b, timed(b, 0, [&] { return mix_launch_stage(v, margs(b, MODE_STAGE1, -1, A, A, B), b->grid_stage, s); }));
    BLAUNCH(b, timed(b, 0, [&] { return mix_launch_stage(v, margs(b, MODE_STAGE2, -1, B, A, C), b->grid_stage, s); }));
    BLAUNCH(b, timed(b, 0, [&] { return mix_launch_stage(v, margs(b, MODE_STAGE3, -1, C, A, A), b->grid_stage, s); }));
    BLAUNCH(b, timed(b, 2, [&] { return mix_launch_post(b->d_inst, b->n, s); }));
    return P2DE_OK;
}

int ensure_graph(p2de_batch* b) {
    if (b->gexec) return P2DE_OK;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const bool timing = b->timing;
    b->timing = false;
    int st = launch_step(b, false);
    b->timing = timing;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return bfail(b, P2DE_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    BHIP(b, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    return P2DE_OK;
}

int enqueue(p2de_batch* b, int steps) {
    if (b->use_graph && !b->timing) {
        int st = ensure_graph(b);
        if (st) return st;
        if (steps > 0) BLAUNCH(b, mix_launch_init(b->d_inst, b->n, 0, -1, 0.0, 0.0, b->stream));
        for (int i = 0; i < steps; i++) BHIP(b, hipGraphLaunch(b->gexec, b->stream));
        return P2DE_OK;
    }
    for (int i = 0; i < steps; i++) {
        int st = launch_step(b, true);
        if (st) return st;
    }
    return P2DE_OK;
}

int read_sc(p2de_batch* b) {
    BHIP(b, hipMemcpyAsync(b->h_sc, b->d_sc, sizeof(StepScalars) * b->n, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    if (b->timing) harvest(b);
    return P2DE_OK;
}

int write_sc(p2de_batch* b) {
    BHIP(b, hipMemcpyAsync(b->d_sc, b->h_sc, sizeof(StepScalars) * b->n, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return P2DE_OK;
}

double h_bits_to_double(unsigned long long u) {
    double d;
    std::memcpy(&d, &u, sizeof(d));
    return d;
}

// Autotuner cache: one "key value" line per problem; the batch stores the chosen launch shape
// (NT:E2,E3,E4,E5), not its index, so a build with another list of shapes re-tunes instead of
// taking a different shape.  Lines written by the FP64 contexts are kept as they are.
std::string shape_str(int v) {
    int NT = 0, E[MIX_MAX_N1D + 1] = {0};
    mix_variant_shape(v, &NT, E);
    std::ostringstream o;
    o << NT << ':' << E[2] << ',' << E[3] << ',' << E[4] << ',' << E[5];
    return o.str();
}

std::map<std::string, std::string> read_tune_cache(const char* path) {
    std::map<std::string, std::string> m;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        std::string k, v;
        if (ls >> k >> v) m[k] = v;
    }
    return m;
}

std::string tune_key(const p2de_batch* b) {
    std::ostringstream k;
    k << "mixed";
    for (int n = MIX_MIN_N1D; n <= MIX_MAX_N1D; n++) {
        int64_t K = 0;
        for (auto& I : b->inst)
            if (I.N1D == n) K += I.K;
        if (K) k << ":N" << (n - 1) << "K" << K;
    }
    return k.str();
}

}  // namespace

int batch_instance_stats(p2de_batch* b, int inst, double* t, double* last_dt, int* steps, int64_t* viol) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (inst < 0 || inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    int st = read_sc(b);
    if (st) return st;
    const StepScalars& S = b->h_sc[inst];
    if (t) *t = S.t;
    if (last_dt) *last_dt = S.last_dt;
    if (steps) *steps = S.steps;
    if (viol) *viol = (int64_t)S.viol;
    return P2DE_OK;
}

// Device time, last step and step counter of one instance as a checkpoint recorded them.
int batch_instance_restore(p2de_batch* b, int inst, double t, double last_dt, int steps) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (inst < 0 || inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    int st = read_sc(b);
    if (st) return st;
    StepScalars& S = b->h_sc[inst];
    S.t = t;
    S.done = 0;
    S.hold = 0;
    S.steps = steps;
    S.last_dt = last_dt;
    return write_sc(b);
}

extern "C" {

int p2de_batch_create(int n, const p2de_config* cfg, const p2de_operators* ops, const p2de_geometry* geom,
                      const p2de_bcdata* bc, p2de_batch** out) {
    if (!out) return bfail(nullptr, P2DE_E_INVALID, "out is NULL");
    *out = nullptr;
    if (n < 1 || !cfg || !ops || !geom || !bc) return bfail(nullptr, P2DE_E_INVALID, "n < 1 or NULL descriptor array");
    p2de_batch* b = nullptr;
    try {
        b = new p2de_batch();
    } catch (...) {
        return bfail(nullptr, P2DE_E_ALLOC, "host allocation failed");
    }
    auto bail = [&](int code) {
        g_batch_error = b->err;
        free_batch(b);
        return code;
    };
    try {
        b->n = n;
        b->inst.resize(n);
        std::vector<MeshArrays> meshes(n);
        std::vector<ElemTables> tabs(n);
        int64_t ival_rows = 0;
        for (int i = 0; i < n; i++) {
            BInst& I = b->inst[i];
            I.cfg = cfg[i];
            int Nh = 0;
            std::string e;
            int st = setup_config(&cfg[i], &I.Nq, &I.Nfp, &Nh, &I.K, e);
            if (st) return bfail(b, st, "instance " + std::to_string(i) + ": " + e), bail(st);
            if (cfg[i].Nd != 2 || cfg[i].precision != P2DE_PRECISION_MIXED)
                return bfail(b, P2DE_E_UNSUPPORTED, "instance " + std::to_string(i) + ": batches hold 2D mixed-precision solves"),
                       bail(P2DE_E_UNSUPPORTED);
            I.N1D = cfg[i].N + 1;
            if (ops[i].Nq != I.Nq || ops[i].Nfp != I.Nfp || ops[i].Nh != Nh)
                return bfail(b, P2DE_E_INVALID, "instance " + std::to_string(i) + ": operator sizes"), bail(P2DE_E_INVALID);
            const void* req[] = {ops[i].Srsh_db, ops[i].Srs0, ops[i].Srsh_nnz, ops[i].Srs0_nnz, ops[i].Brs,
                                 ops[i].MinvVhT, ops[i].MinvVfT, ops[i].Vf, ops[i].Vf_low, ops[i].wq, ops[i].fq2q,
                                 ops[i].q2fq, geom[i].Jq, geom[i].GJh, bc[i].mapP};
            for (const void* p : req)
                if (!p) return bfail(b, P2DE_E_INVALID, "instance " + std::to_string(i) + ": NULL array"), bail(P2DE_E_INVALID);
            if ((st = setup_tables(tabs[i], 2, I.N1D, P2DE_BASIS_LOBATTO, &ops[i], e)))
                return bfail(b, st, "instance " + std::to_string(i) + ": " + e), bail(st);
            if ((st = setup_mesh(2, I.N1D, I.K, tabs[i], &geom[i], &bc[i], meshes[i], e)))
                return bfail(b, st, "instance " + std::to_string(i) + ": " + e), bail(st);
            for (const Geo& G : meshes[i].geo)
                if (!(G.rxJ > 0.0) || !(G.syJ > 0.0))
                    return bfail(b, P2DE_E_UNSUPPORTED, "instance " + std::to_string(i) + ": rxJ and syJ must be positive"),
                           bail(P2DE_E_UNSUPPORTED);
            I.node_off = b->total_nodes;
            I.elem_off = b->total_elems;
            I.face_off = b->total_faces;
            I.ival_off = ival_rows;
            b->total_nodes += (int64_t)I.K * I.Nq;
            b->total_elems += I.K;
            b->total_faces += (int64_t)I.K * I.Nfp;
            ival_rows += bc[i].nI;
            b->present[I.N1D] = true;
        }
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
            return bfail(b, P2DE_E_HIP, "hipStreamCreate failed"), bail(P2DE_E_HIP);
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            b->cus = prop.multiProcessorCount;
        // applicable launch shapes: every present degree compiled in
        for (int v = 0; v < mix_num_variants(); v++) {
            int NT = 0, E[MIX_MAX_N1D + 1];
            mix_variant_shape(v, &NT, E);
            bool ok = true;
            for (int d = MIX_MIN_N1D; d <= MIX_MAX_N1D; d++)
                if (b->present[d] && E[d] == 0) ok = false;
            if (ok) b->variants.push_back(v);
        }
        if (b->variants.empty()) return bfail(b, P2DE_E_UNSUPPORTED, "no launch shape covers these degrees"), bail(P2DE_E_UNSUPPORTED);
        int st;
        const size_t nU = (size_t)b->total_nodes * 4;
        if ((st = balloc(b, &b->d_inst, (size_t)n)) || (st = balloc(b, &b->d_tiles, (size_t)b->total_elems)) ||
            (st = balloc(b, &b->d_dtp, (size_t)b->total_elems)) || (st = balloc(b, &b->d_geo, (size_t)b->total_elems)) ||
            (st = balloc(b, &b->d_nb_node, (size_t)b->total_faces)) || (st = balloc(b, &b->d_bc, (size_t)b->total_faces)) ||
            (st = balloc(b, &b->d_Ival, (size_t)std::max<int64_t>(ival_rows, 1) * 4)) || (st = balloc(b, &b->d_U[0], nU)) ||
            (st = balloc(b, &b->d_U[1], nU)) || (st = balloc(b, &b->d_U[2], nU)) || (st = balloc(b, &b->d_rhsU, nU)) ||
            (st = balloc(b, &b->d_sc, (size_t)n)) || (st = balloc(b, &b->d_keys, 3)))
            return bail(st);
        if (hipHostMalloc((void**)&b->h_sc, sizeof(StepScalars) * n) != hipSuccess ||
            hipHostMalloc((void**)&b->h_keys, 3 * sizeof(unsigned long long)) != hipSuccess)
            return bfail(b, P2DE_E_ALLOC, "hipHostMalloc failed"), bail(P2DE_E_ALLOC);
        std::vector<MixInst>& hinst = b->hinst;
        hinst.assign(n, MixInst{});
        for (int i = 0; i < n; i++) {
            BInst& I = b->inst[i];
            if ((st = upload_tab(b, I.N1D, tabs[i], &ops[i]))) return bail(st);
            const MeshArrays& M = meshes[i];
            hipError_t e;
            if ((e = hipMemcpy(b->d_geo + I.elem_off, M.geo.data(), sizeof(Geo) * I.K, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(b->d_nb_node + I.face_off, M.nb_node.data(), sizeof(int) * M.nb_node.size(),
                               hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(b->d_bc + I.face_off, M.bcode.data(), sizeof(int) * M.bcode.size(), hipMemcpyHostToDevice)) !=
                    hipSuccess)
                return bfail(b, P2DE_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)), bail(P2DE_E_HIP);
            if (bc[i].nI > 0) {
                if (!bc[i].Ival) return bfail(b, P2DE_E_INVALID, "Ival is NULL"), bail(P2DE_E_INVALID);
                if ((e = hipMemcpy(b->d_Ival + I.ival_off * 4, bc[i].Ival, sizeof(double) * bc[i].nI * 4,
                                   hipMemcpyHostToDevice)) != hipSuccess)
                    return bfail(b, P2DE_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)), bail(P2DE_E_HIP);
            }
            MixInst& H = hinst[i];
            std::memset(&H, 0, sizeof(H));
            H.node_off = I.node_off;
            H.geo = b->d_geo + I.elem_off;
            H.nb_node = b->d_nb_node + I.face_off;
            H.bc = b->d_bc + I.face_off;
            H.Ival = b->d_Ival + I.ival_off * 4;
            H.sc = b->d_sc + i;
            H.gamma = I.cfg.gamma;
            H.zeta = I.cfg.zeta;
            H.ZEROTOL = I.cfg.ZEROTOL;
            H.CFL = I.cfg.CFL;
            H.dt0 = I.cfg.dt0;
            H.T = I.cfg.T;
            H.K = I.K;
            H.N1D = I.N1D;
            H.upper_mode = I.cfg.upper_bound_mode;
            StepScalars& S = b->h_sc[i];
            std::memset(&S, 0, sizeof(S));
            S.t = I.cfg.t0;
            S.first_bad = ~0ull;
        }
        hipError_t e;
        if ((e = hipMemcpy(b->d_inst, hinst.data(), sizeof(MixInst) * n, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(b->d_sc, b->h_sc, sizeof(StepScalars) * n, hipMemcpyHostToDevice)) != hipSuccess)
            return bfail(b, P2DE_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)), bail(P2DE_E_HIP);
        b->vsel = 0;
        if ((st = build_tiles(b))) return bail(st);
    } catch (const std::bad_alloc&) {
        return bfail(b, P2DE_E_ALLOC, "host allocation failed"), bail(P2DE_E_ALLOC);
    } catch (...) {
        return bfail(b, P2DE_E_INVALID, "unexpected exception"), bail(P2DE_E_INVALID);
    }
    *out = b;
    return P2DE_OK;
}

void p2de_batch_destroy(p2de_batch* b) { free_batch(b); }

int p2de_batch_num_instances(const p2de_batch* b) { return b ? b->n : 0; }

int64_t p2de_batch_num_dofs(const p2de_batch* b, int inst) {
    if (!b) return 0;
    if (inst < 0) return b->total_nodes;
    if (inst >= b->n) return 0;
    return (int64_t)b->inst[inst].K * b->inst[inst].Nq;
}

int p2de_batch_set_state(p2de_batch* b, int inst, const double* Uq) {
    if (!b || !Uq) return bfail(b, P2DE_E_INVALID, "NULL argument");
    if (inst < 0 || inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    const BInst& I = b->inst[inst];
    BHIP(b, hipMemcpyAsync(b->d_U[0] + I.node_off * 4, Uq, sizeof(double) * (size_t)I.K * I.Nq * 4,
                           hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    b->state_set = true;
    return P2DE_OK;
}

int p2de_batch_get_state(p2de_batch* b, int inst, double* Uq) {
    if (!b || !Uq) return bfail(b, P2DE_E_INVALID, "NULL argument");
    if (inst < 0 || inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    const BInst& I = b->inst[inst];
    BHIP(b, hipMemcpyAsync(Uq, b->d_U[0] + I.node_off * 4, sizeof(double) * (size_t)I.K * I.Nq * 4,
                           hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return P2DE_OK;
}

int p2de_batch_device_state(p2de_batch* b, int inst, void** dptr) {
    if (!b || !dptr) return bfail(b, P2DE_E_INVALID, "NULL argument");
    if (inst < 0 || inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    *dptr = b->d_U[0] + b->inst[inst].node_off * 4;
    return P2DE_OK;
}

int p2de_batch_set_time(p2de_batch* b, int inst, double t) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    int st = read_sc(b);
    if (st) return st;
    for (int i = 0; i < b->n; i++)
        if (inst < 0 || i == inst) {
            b->h_sc[i].t = t;
            b->h_sc[i].done = 0;
        }
    return write_sc(b);
}

int p2de_batch_get_time(p2de_batch* b, int inst, double* t, double* last_dt, int* steps) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (inst < 0 || inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    int st = read_sc(b);
    if (st) return st;
    if (t) *t = b->h_sc[inst].t;
    if (last_dt) *last_dt = b->h_sc[inst].last_dt;
    if (steps) *steps = b->h_sc[inst].steps;
    return P2DE_OK;
}

int p2de_batch_set_debug(p2de_batch* b, int on) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (on && !b->d_dbg_rhsH) {
        int st;
        if ((st = balloc(b, &b->d_dbg_rhsH, (size_t)b->total_nodes * 4)) ||
            (st = balloc(b, &b->d_dbg_rhsL, (size_t)b->total_nodes * 4)))
            return st;
    }
    b->debug = on != 0;
    return P2DE_OK;
}

int p2de_batch_rhs(p2de_batch* b, int inst, double t, double dt, int nstage, double* dt_out, double* rhsU) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (inst < 0 || inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    if (nstage < 1 || nstage > 3) return bfail(b, P2DE_E_INVALID, "nstage must be 1, 2 or 3");
    if (!b->state_set) return bfail(b, P2DE_E_STATE, "no state uploaded (p2de_batch_set_state)");
    const BInst& I = b->inst[inst];
    const p2de_config& g = I.cfg;
    double cap = dt;
    if (nstage == 1) {   // low_order_graph_viscosity.jl:230
        const double a = g.CFL * g.dt0, c = g.T - t;
        cap = std::isnan(a) || std::isnan(c) ? a - c : (std::signbit(a - c) ? a : c);
    }
    void* s = b->stream;
    const int v = b->variants[b->vsel];
    double* A = b->d_U[0];
    BLAUNCH(b, timed(b, 2, [&] { return mix_launch_init(b->d_inst, b->n, 1, inst, dt, cap, s); }));
    if (nstage == 1) {
        BLAUNCH(b, timed(b, 1, [&] { return mix_launch_dt(v, margs(b, MODE_RHS, inst, A, A, nullptr), b->grid_dt, s); }));
        BLAUNCH(b, timed(b, 2, [&] { return mix_launch_dt_reduce(b->d_inst, b->n, b->d_dtp, inst, s); }));
    }
    BLAUNCH(b, timed(b, 0, [&] { return mix_launch_stage(v, margs(b, MODE_RHS, inst, A, A, nullptr), b->grid_stage, s); }));
    if (rhsU)
        BHIP(b, hipMemcpyAsync(rhsU, b->d_rhsU + I.node_off * 4, sizeof(double) * (size_t)I.K * I.Nq * 4,
                               hipMemcpyDeviceToHost, b->stream));
    int st = read_sc(b);
    if (st) return st;
    if (dt_out) *dt_out = h_bits_to_double(b->h_sc[inst].dt_bits);
    return P2DE_OK;
}

int p2de_batch_get_field(p2de_batch* b, int inst, int field, double* out) {
    if (!b || !out) return bfail(b, P2DE_E_INVALID, "NULL argument");
    if (inst < 0 || inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    const BInst& I = b->inst[inst];
    const double* src = nullptr;
    switch (field) {
        case P2DE_FIELD_RHSU: src = b->d_rhsU; break;
        case P2DE_FIELD_RHSH: src = b->d_dbg_rhsH; break;
        case P2DE_FIELD_RHSL: src = b->d_dbg_rhsL; break;
        default: return bfail(b, P2DE_E_UNSUPPORTED, "field not kept by the mixed-precision path");
    }
    if (!src) return bfail(b, P2DE_E_STATE, "debug output is off (p2de_batch_set_debug)");
    BHIP(b, hipMemcpyAsync(out, src + I.node_off * 4, sizeof(double) * (size_t)I.K * I.Nq * 4, hipMemcpyDeviceToHost,
                           b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return P2DE_OK;
}

int p2de_batch_enqueue_steps(p2de_batch* b, int steps) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (!b->state_set) return bfail(b, P2DE_E_STATE, "no state uploaded (p2de_batch_set_state)");
    if (steps < 0) return bfail(b, P2DE_E_INVALID, "steps < 0");
    return enqueue(b, steps);
}

int p2de_batch_synchronize(p2de_batch* b) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    BHIP(b, hipStreamSynchronize(b->stream));
    if (b->timing) harvest(b);
    return P2DE_OK;
}

int p2de_batch_run(p2de_batch* b, int max_steps, int* steps_taken) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (!b->state_set) return bfail(b, P2DE_E_STATE, "no state uploaded (p2de_batch_set_state)");
    int done_steps = 0;
    for (;;) {
        int st = read_sc(b);
        if (st) return st;
        bool all_done = true;
        for (int i = 0; i < b->n; i++)
            if (b->h_sc[i].t < b->inst[i].cfg.T) all_done = false;
        if (all_done || (max_steps > 0 && done_steps >= max_steps)) break;
        int chunk = 32;
        if (max_steps > 0) chunk = std::min(chunk, max_steps - done_steps);
        if ((st = enqueue(b, chunk))) return st;
        done_steps += chunk;
    }
    if (steps_taken) *steps_taken = done_steps;
    return P2DE_OK;
}

int p2de_batch_check_positivity(p2de_batch* b, int inst, p2de_pos_report* rep) {
    if (!b || !rep) return bfail(b, P2DE_E_INVALID, "NULL argument");
    if (inst < 0 || inst >= b->n) return bfail(b, P2DE_E_INVALID, "instance out of range");
    const BInst& I = b->inst[inst];
    b->h_keys[0] = ~0ull;
    b->h_keys[1] = ~0ull;
    b->h_keys[2] = 0;
    BHIP(b, hipMemcpyAsync(b->d_keys, b->h_keys, 3 * sizeof(unsigned long long), hipMemcpyHostToDevice, b->stream));
    BLAUNCH(b, mix_launch_positivity(b->d_U[0] + I.node_off * 4, (int64_t)I.K * I.Nq, I.cfg.gamma, b->d_keys,
                                     b->d_keys + 2, b->stream));
    BHIP(b, hipMemcpyAsync(b->h_keys, b->d_keys, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, b->stream));
    int st = read_sc(b);
    if (st) return st;
    auto key_to_double = [](unsigned long long k) {
        const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
        return h_bits_to_double(u);
    };
    StepScalars& S = b->h_sc[inst];
    rep->violations = (int64_t)S.viol + (int64_t)b->h_keys[2];
    rep->first_node = S.first_bad == ~0ull ? -1 : (int64_t)(S.first_bad & ((1ull << 40) - 1));
    rep->first_stage = S.first_bad == ~0ull ? 0 : (int)(S.first_bad >> 40);
    rep->min_rho = key_to_double(b->h_keys[0]);
    rep->min_p = key_to_double(b->h_keys[1]);
    S.viol = 0;
    S.first_bad = ~0ull;
    return write_sc(b);
}

int p2de_batch_set_kernel_timing(p2de_batch* b, int on) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    BHIP(b, hipStreamSynchronize(b->stream));
    harvest(b);
    b->timing = on != 0;
    for (int k = 0; k < 3; k++) {
        b->time_ms[k] = 0;
        b->time_count[k] = 0;
    }
    return P2DE_OK;
}

int p2de_batch_get_kernel_timing(p2de_batch* b, double ms[3], int64_t count[3]) {
    if (!b || !ms || !count) return bfail(b, P2DE_E_INVALID, "NULL argument");
    BHIP(b, hipStreamSynchronize(b->stream));
    harvest(b);
    for (int k = 0; k < 3; k++) {
        ms[k] = b->time_ms[k];
        count[k] = b->time_count[k];
    }
    return P2DE_OK;
}

int p2de_batch_set_graph(p2de_batch* b, int on) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    b->use_graph = on != 0;
    return P2DE_OK;
}

int p2de_batch_num_variants(const p2de_batch* b) { return b ? (int)b->variants.size() : 0; }

int p2de_batch_set_variant(p2de_batch* b, int v) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (v < 0 || v >= (int)b->variants.size()) return bfail(b, P2DE_E_INVALID, "variant out of range");
    BHIP(b, hipStreamSynchronize(b->stream));
    b->vsel = v;
    return build_tiles(b);
}

int p2de_batch_get_variant(const p2de_batch* b, int* v, int* NT, int* E) {
    if (!b) return bfail(nullptr, P2DE_E_INVALID, "NULL batch");
    int nt = 0, e[MIX_MAX_N1D + 1];
    mix_variant_shape(b->variants[b->vsel], &nt, e);
    if (v) *v = b->vsel;
    if (NT) *NT = nt;
    if (E)
        for (int n = 0; n <= MIX_MAX_N1D; n++) E[n] = e[n];
    return P2DE_OK;
}

int p2de_batch_autotune(p2de_batch* b, int steps, const char* cache_path, int* chosen) {
    if (!b) return bfail(b, P2DE_E_INVALID, "NULL batch");
    if (!b->state_set) return bfail(b, P2DE_E_STATE, "no state uploaded (p2de_batch_set_state)");
    const std::string key = tune_key(b);
    const int nv = (int)b->variants.size();
    if (cache_path) {
        const auto cache = read_tune_cache(cache_path);
        const auto it = cache.find(key);
        if (it != cache.end())
            for (int v = 0; v < nv; v++)
                if (shape_str(b->variants[v]) == it->second) {
                    if (chosen) *chosen = v;
                    return p2de_batch_set_variant(b, v);
                }
    }
    const size_t nU = (size_t)b->total_nodes * 4;
    double* saved = nullptr;
    int st = balloc(b, &saved, nU);
    if (st) return st;
    BHIP(b, hipMemcpyAsync(saved, b->d_U[0], nU * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    if ((st = read_sc(b))) return st;
    std::vector<StepScalars> keep(b->h_sc, b->h_sc + b->n);
    const bool timing_was = b->timing;
    int best = 0;
    double best_ms = 1e300;
    for (int v = 0; v < nv; v++) {
        b->vsel = v;
        if ((st = build_tiles(b))) return st;
        BHIP(b, hipMemcpyAsync(b->d_U[0], saved, nU * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
        std::copy(keep.begin(), keep.end(), b->h_sc);
        if ((st = write_sc(b))) return st;
        b->timing = false;
        if ((st = enqueue(b, 1))) return st;   // warm-up
        BHIP(b, hipStreamSynchronize(b->stream));
        b->timing = true;
        for (int k = 0; k < 3; k++) {
            b->time_ms[k] = 0;
            b->time_count[k] = 0;
        }
        if ((st = enqueue(b, std::max(steps, 1)))) return st;
        BHIP(b, hipStreamSynchronize(b->stream));
        harvest(b);
        const double ms = b->time_ms[0] + b->time_ms[1];
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    b->timing = timing_was;
    for (int k = 0; k < 3; k++) {
        b->time_ms[k] = 0;
        b->time_count[k] = 0;
    }
    b->vsel = best;
    if ((st = build_tiles(b))) return st;
    BHIP(b, hipMemcpyAsync(b->d_U[0], saved, nU * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    std::copy(keep.begin(), keep.end(), b->h_sc);
    if ((st = write_sc(b))) return st;
    (void)hipFree(saved);
    if (cache_path) {
        auto cache = read_tune_cache(cache_path);
        cache[key] = shape_str(b->variants[best]);
        std::ofstream f(cache_path, std::ios::trunc);
        for (auto& kv : cache) f << kv.first << " " << kv.second << "\n";
    }
    if (chosen) *chosen = best;
    return P2DE_OK;
}

const char* p2de_batch_last_error(const p2de_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

}  // extern "C"
