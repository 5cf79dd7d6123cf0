// p2de_setup.cpp -- validation of the caller's operators and mesh and the tables built from
// them (see p2de_setup.h).
#include "p2de_setup.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace p2de {
namespace {

int fail_s(std::string& err, int code, const std::string& msg) {
    err = msg;
    return code;
}

inline double cm(const double* A, int64_t ld, int64_t i, int64_t j) { return A[i + ld * j]; }

int64_t ipow(int a, int b) {
    int64_t r = 1;
    for (int i = 0; i < b; i++) r *= a;
    return r;
}

}  // namespace

int setup_config(const p2de_config* cfg, int* Nq, int* Nfp, int* Nh, int* K, std::string& err) {
    // option set (Solver.jl:1-121); anything else is not implemented here
    if (cfg->rhs != P2DE_RHS_LIMITED_DG ||
        (cfg->low_order_surface_flux != P2DE_LF_ON_NODAL_VAL && cfg->low_order_surface_flux != P2DE_LF_ON_PROJECTED_VAL) ||
        cfg->high_order_surface_flux != P2DE_LF_ON_PROJECTED_VAL ||
        cfg->high_order_volume_flux != P2DE_CHANDRASHEKAR_FLUX)
        return fail_s(err, P2DE_E_UNSUPPORTED, "option set not implemented");
    if (cfg->rhs_limiter != P2DE_SUBCELL_LIMITER && cfg->rhs_limiter != P2DE_ZHANG_SHU_LIMITER)
        return fail_s(err, P2DE_E_UNSUPPORTED, "rhs limiter not implemented");
    if (cfg->bound < P2DE_POSITIVITY_BOUND || cfg->bound > P2DE_TVD_AND_RELAXED_CELL_ENTROPY_BOUND)
        return fail_s(err, P2DE_E_UNSUPPORTED, "limiter bound not implemented");
    const int ent = cfg->bound % 5;   // entropy part: none, min, relaxed min, cell, relaxed cell
    // the cell-entropy bounds: in 1D the reference sizes their per-element scratch (Nq - N1D) x K,
    // empty, and indexes past it (State.jl:182-183, subcell.jl:503-505); on the Gauss basis their
    // interface step (subcell.jl:745-789) reads neighbour parameters that the neighbour's own step
    // may already have lowered, so its result depends on the thread schedule
    if (ent >= 3 && (cfg->Nd != 2 || cfg->basis != P2DE_BASIS_LOBATTO))
        return fail_s(err, P2DE_E_UNSUPPORTED, "the cell-entropy bounds are implemented for 2D Lobatto");
    if (ent == 4 && !(cfg->cell_beta == cfg->cell_beta))
        return fail_s(err, P2DE_E_INVALID, "the relaxed cell-entropy bound needs a beta");
    // bound(::ZhangShuLimiter) is PositivityBound (traits.jl:61-63)
    if (cfg->rhs_limiter == P2DE_ZHANG_SHU_LIMITER && cfg->bound != P2DE_POSITIVITY_BOUND)
        return fail_s(err, P2DE_E_UNSUPPORTED, "the Zhang-Shu limiter takes no subcell bound");
    if (cfg->shockcapture != P2DE_NO_SHOCK_CAPTURE && cfg->shockcapture != P2DE_HENNEMANN_SHOCK_CAPTURE)
        return fail_s(err, P2DE_E_UNSUPPORTED, "shock capture not implemented");
    if (cfg->shockcapture == P2DE_HENNEMANN_SHOCK_CAPTURE && !(cfg->shock_a > 0.0))
        return fail_s(err, P2DE_E_INVALID, "HennemannShockCapture needs a > 0");
    if (cfg->basis != P2DE_BASIS_LOBATTO && cfg->basis != P2DE_BASIS_GAUSS)
        return fail_s(err, P2DE_E_UNSUPPORTED, "approximation basis not implemented");
    if (cfg->entropyproj_limiter != P2DE_NO_ENTROPY_PROJ_LIMITER &&
        cfg->entropyproj_limiter != P2DE_NODEWISE_SCALED_EXTRAPOLATION)
        return fail_s(err, P2DE_E_UNSUPPORTED, "entropy-projection limiter not implemented");
    // the Lobatto method of compute_entropyproj_limiting_param! (filter.jl:1-3) does not run
    if (cfg->basis == P2DE_BASIS_LOBATTO && cfg->entropyproj_limiter != P2DE_NO_ENTROPY_PROJ_LIMITER)
        return fail_s(err, P2DE_E_UNSUPPORTED, "the Lobatto basis takes NoEntropyProjectionLimiter");
    if (cfg->precision != P2DE_PRECISION_FP64 && cfg->precision != P2DE_PRECISION_MIXED)
        return fail_s(err, P2DE_E_INVALID, "precision");
    if (cfg->upper_bound_mode != P2DE_UPPER_BOUND_REFERENCE && cfg->upper_bound_mode != P2DE_UPPER_BOUND_INTENDED)
        return fail_s(err, P2DE_E_INVALID, "upper_bound_mode");
    if (cfg->Nd != 1 && cfg->Nd != 2) return fail_s(err, P2DE_E_INVALID, "Nd must be 1 or 2");
    if (cfg->N < 1 || cfg->N > 4) return fail_s(err, P2DE_E_UNSUPPORTED, "N must be 1..4");
    if (cfg->Kx < 1 || cfg->Ky < 1 || (cfg->Nd == 1 && cfg->Ky != 1)) return fail_s(err, P2DE_E_INVALID, "Kx, Ky");
    if (!(cfg->gamma > 1.0)) return fail_s(err, P2DE_E_INVALID, "gamma must exceed 1");
    if (cfg->precision == P2DE_PRECISION_MIXED && cfg->Nd != 2)
        return fail_s(err, P2DE_E_UNSUPPORTED, "the mixed-precision path is 2D only");
    if (cfg->precision == P2DE_PRECISION_MIXED && cfg->basis != P2DE_BASIS_LOBATTO)
        return fail_s(err, P2DE_E_UNSUPPORTED, "the mixed-precision path runs the Lobatto basis");
    if (cfg->precision == P2DE_PRECISION_MIXED &&
        (cfg->rhs_limiter != P2DE_SUBCELL_LIMITER || cfg->bound != P2DE_POSITIVITY_BOUND ||
         cfg->shockcapture != P2DE_NO_SHOCK_CAPTURE || cfg->low_order_surface_flux != P2DE_LF_ON_NODAL_VAL))
        return fail_s(err, P2DE_E_UNSUPPORTED,
                      "the mixed-precision path runs SubcellLimiter(PositivityBound, NoShockCapture)");
    const int N1D = cfg->N + 1;
    *Nq = (int)ipow(N1D, cfg->Nd);
    *Nfp = cfg->Nd == 1 ? 2 : 4 * N1D;
    *Nh = *Nq + *Nfp;
    const int64_t K64 = (int64_t)cfg->Kx * cfg->Ky;
    if (K64 * *Nq * (cfg->Nd + 2) >= (int64_t)1 << 31 || K64 * *Nfp >= (int64_t)1 << 31)
        return fail_s(err, P2DE_E_UNSUPPORTED, "problem too large for 32-bit node indices");
    *K = (int)K64;
    return P2DE_OK;
}

int setup_tables(ElemTables& T, int Nd, int N1D, int basis, const p2de_operators* ops, std::string& err) {
    std::memset(&T, 0, sizeof(T));
    const int Nq = Nd == 1 ? N1D : N1D * N1D, Nfp = Nd == 1 ? 2 : 4 * N1D, Nh = Nq + Nfp;
    const bool gauss = basis == P2DE_BASIS_GAUSS;
    T.basis = gauss ? 1 : 0;
    T.Nd = Nd;
    T.Nc = Nd + 2;
    T.N1D = N1D;
    T.Nq = Nq;
    T.Nfp = Nfp;
    T.Nh = Nh;
    // face-volume pairs: one per face node (Lobatto), a whole line per face node (Gauss)
    const int64_t nfv = gauss ? (int64_t)Nfp * N1D : Nfp;
    const int64_t nph_expect = Nd == 1 ? N1D * (N1D - 1) / 2 + nfv : (int64_t)N1D * N1D * (N1D - 1) + nfv;
    const int64_t npl_expect = Nd == 1 ? N1D - 1 : 2 * N1D * (N1D - 1);
    if (ops->nnz_h != nph_expect)
        return fail_s(err, P2DE_E_UNSUPPORTED, "Srsh_nnz has " + std::to_string(ops->nnz_h) + " pairs, the " +
                                               (gauss ? "Gauss" : "Lobatto") + " operator has " + std::to_string(nph_expect));
    if (ops->nnz_0 != npl_expect)
        return fail_s(err, P2DE_E_UNSUPPORTED, "Srs0_nnz has " + std::to_string(ops->nnz_0) + " pairs, expected " +
                                               std::to_string(npl_expect));
    T.nph = (int)nph_expect;
    T.npl = (int)npl_expect;

    // fq2q and the tensor-product face layout the subcell limiter assumes (subcell.jl:173-206)
    for (int f = 0; f < Nfp; f++) {
        const int64_t q1 = ops->fq2q[f];
        if (q1 < 1 || q1 > Nq) return fail_s(err, P2DE_E_INVALID, "fq2q out of range");
        const int q = (int)q1 - 1;
        int expect;
        if (Nd == 1) {
            expect = f == 0 ? 0 : N1D - 1;
        } else {
            const int face = f / N1D, m = f % N1D;
            expect = face == 0 ? m * N1D : face == 1 ? m * N1D + N1D - 1 : face == 2 ? m : (N1D - 1) * N1D + m;
        }
        if (q != expect) return fail_s(err, P2DE_E_UNSUPPORTED, "fq2q does not follow the tensor-product face layout");
        T.fq2q[f] = (uint8_t)q;
        T.ld[f] = (int8_t)((Nd == 1 || f < 2 * N1D) ? 0 : 1);   // rhs_utils.jl:84-88
        if (Nd == 1) T.bsign[f] = f == 0 ? -1 : 1;
        else T.bsign[f] = (f < N1D || (f >= 2 * N1D && f < 3 * N1D)) ? -1 : 1;
        T.Br[f] = cm(ops->Brs, Nfp, f, f);
        T.Bs[f] = Nd == 2 ? cm(ops->Brs + (int64_t)Nfp * Nfp, Nfp, f, f) : 0.0;
        if (Nd == 2) {
            const bool xface = f < 2 * N1D;
            if (xface ? (T.Br[f] == 0.0 || T.Bs[f] != 0.0) : (T.Bs[f] == 0.0 || T.Br[f] != 0.0))
                return fail_s(err, P2DE_E_UNSUPPORTED, "Brs does not have one nonzero direction per face node");
        } else if (T.Br[f] == 0.0) {
            return fail_s(err, P2DE_E_UNSUPPORTED, "Br has a zero face entry");
        }
        // the line of face node f: the nodes the rows of Vf and Vf_low may reach
        bool online[MAX_NQ] = {};
        for (int s = 0; s < N1D; s++) {
            int node;
            if (Nd == 1) {
                node = s;
            } else {
                const int face = f / N1D, m = f % N1D;
                node = face < 2 ? m * N1D + s : s * N1D + m;
            }
            online[node] = true;
            T.fl_node[f][s] = (uint8_t)node;
            T.fl_vf[f][s] = cm(ops->Vf, Nfp, f, node);
            T.fl_vfl[f][s] = cm(ops->Vf_low, Nfp, f, node);
        }
        for (int j = 0; j < Nq; j++) {
            const double vf = cm(ops->Vf, Nfp, f, j), vfl = cm(ops->Vf_low, Nfp, f, j);
            if (!online[j] && (vf != 0.0 || vfl != 0.0))
                return fail_s(err, P2DE_E_UNSUPPORTED, "a row of Vf or Vf_low reaches off the face node's line");
            if (vfl != (j == q ? 1.0 : 0.0))
                return fail_s(err, P2DE_E_UNSUPPORTED, "Vf_low is not the selection of the fq2q node");
            if (!gauss) {
                // entropy projection row (rhs.jl:90-92 with l = 1): a single nonzero at fq2q[f]
                const double coef = 1.0 * vf + (1.0 - 1.0) * vfl;
                if (j == q) {
                    if (coef == 0.0) return fail_s(err, P2DE_E_UNSUPPORTED, "Vf is zero at fq2q");
                    T.vf_coef[f] = coef;
                } else if (vf != 0.0) {
                    return fail_s(err, P2DE_E_UNSUPPORTED, "Vf is not a face-node selection: pass basis = Gauss");
                }
            }
        }
    }
    // node data and the assembly structure of MinvVhT / MinvVfT (flux_differencing.jl:331-361)
    for (int i = 0; i < Nq; i++) {
        T.wq[i] = ops->wq[i];
        if (!(T.wq[i] > 0.0)) return fail_s(err, P2DE_E_INVALID, "wq must be positive");
        T.rwq[i] = 1.0 / T.wq[i];
        for (int j = 0; j < Nq; j++)
            if (j != i && cm(ops->MinvVhT, Nq, i, j) != 0.0)
                return fail_s(err, P2DE_E_UNSUPPORTED, "MinvVhT volume block is not diagonal");
        T.a_v[i] = cm(ops->MinvVhT, Nq, i, i);
        int n = 0, nh = 0;
        for (int f = 0; f < Nfp; f++) {
            const double ah = cm(ops->MinvVhT, Nq, i, Nq + f), af = cm(ops->MinvVfT, Nq, i, f);
            const double vf = cm(ops->Vf, Nfp, f, i), vfl = cm(ops->Vf_low, Nfp, f, i);
            if (T.fq2q[f] == i) {
                if (n >= 2) return fail_s(err, P2DE_E_UNSUPPORTED, "more than two faces at a node");
                T.nface[i][n++] = (int8_t)f;
            }
            // assembly list: faces whose Vf or Vf_low row reaches the node (the only nonzeros of
            // MinvVhT_new's face block); the caller's MinvVhT/MinvVfT must vanish elsewhere
            const bool reach = gauss ? (vf != 0.0 || vfl != 0.0) : T.fq2q[f] == i;
            if (reach) {
                if (nh >= MAX_HF) return fail_s(err, P2DE_E_UNSUPPORTED, "more than four faces reach a node");
                T.hface[i][nh] = (int8_t)f;
                T.a_h[i][nh] = ah;
                T.a_f[i][nh] = af;
                T.vf_h[i][nh] = vf;
                T.vfl_h[i][nh] = vfl;
                nh++;
            } else if (ah != 0.0 || af != 0.0) {
                return fail_s(err, P2DE_E_UNSUPPORTED, "MinvVhT/MinvVfT couple a node to a face whose row does not reach it");
            }
        }
        T.nf[i] = (int8_t)n;
        T.nh[i] = (int8_t)nh;
        int m = 0;
        for (int s = 0; s < 2; s++) {
            const int64_t f1 = ops->q2fq[i + (int64_t)Nq * s];
            if (f1 == 0) continue;
            if (f1 < 1 || f1 > Nfp) return fail_s(err, P2DE_E_INVALID, "q2fq out of range");
            T.q2fq[i][m++] = (int8_t)(f1 - 1);
        }
        T.nq2fq[i] = (int8_t)m;
    }
    // high-order pairs
    std::vector<std::vector<std::pair<int, int>>> hl(Nh);
    for (int p = 0; p < T.nph; p++) {
        const int64_t i1 = ops->Srsh_nnz[2 * p], j1 = ops->Srsh_nnz[2 * p + 1];
        if (i1 < 1 || i1 > Nh || j1 < 1 || j1 > Nh || i1 <= j1) return fail_s(err, P2DE_E_INVALID, "Srsh_nnz entry out of range");
        const int i = (int)i1 - 1, j = (int)j1 - 1;
        const double cr = cm(ops->Srsh_db, Nh, i, j);
        const double cs = Nd == 2 ? cm(ops->Srsh_db + (int64_t)Nh * Nh, Nh, i, j) : 0.0;
        if ((cr != 0.0) == (cs != 0.0) && Nd == 2)
            return fail_s(err, P2DE_E_UNSUPPORTED, "a Srsh pair acts in both or no directions");
        if (Nd == 1 && cr == 0.0) return fail_s(err, P2DE_E_UNSUPPORTED, "zero Srh pair");
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
        if (i1 < 1 || i1 > Nq || j1 < 1 || j1 > Nq || i1 <= j1) return fail_s(err, P2DE_E_INVALID, "Srs0_nnz entry out of range");
        const int i = (int)i1 - 1, j = (int)j1 - 1;
        const double cr = cm(ops->Srs0, Nq, i, j);
        const double cs = Nd == 2 ? cm(ops->Srs0 + (int64_t)Nq * Nq, Nq, i, j) : 0.0;
        if ((cr != 0.0) == (cs != 0.0) && Nd == 2)
            return fail_s(err, P2DE_E_UNSUPPORTED, "a Srs0 pair acts in both or no directions");
        if (Nd == 1 && cr == 0.0) return fail_s(err, P2DE_E_UNSUPPORTED, "zero Sr0 pair");
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
    // smoothness indicator operator (Operators.VDM_inv)
    if (ops->VDM_inv)
        for (int i = 0; i < Nq; i++)
            for (int j = 0; j < Nq; j++) T.vdm_inv[i][j] = cm(ops->VDM_inv, Nq, i, j);
    // low-order stencil (low_order_stencil limiter_utils.jl:214-231): the line neighbours, and at
    // the element boundary the node across face q2fq[q][direction] (or the node's only face)
    for (int q = 0; q < Nq; q++) {
        const int si = Nd == 1 ? q : q % N1D, sj = Nd == 1 ? 0 : q / N1D;
        const int nq = T.nq2fq[q];
        int x = 0;
        for (int d = 0; d < Nd; d++) {
            const int posd = d == 0 ? si : sj;
            for (int side = -1; side <= 1; side += 2) {
                const int np = posd + side;
                if (np >= 0 && np < N1D) {
                    T.stn[q][x] = (int8_t)(d == 0 ? (Nd == 1 ? np : np + N1D * sj) : si + N1D * np);
                } else {
                    if (nq < 1) return fail_s(err, P2DE_E_UNSUPPORTED, "an element-boundary node without a face in q2fq");
                    const int f = T.q2fq[q][nq == 1 ? 0 : d];
                    T.stn[q][x] = (int8_t)(-(f + 1));
                }
                x++;
            }
        }
    }
    return P2DE_OK;
}

int setup_mesh(int Nd, int N1D, int K, const ElemTables& T, const p2de_geometry* geom, const p2de_bcdata* bc,
               MeshArrays& out, std::string& err) {
    const int Nq = Nd == 1 ? N1D : N1D * N1D, Nfp = Nd == 1 ? 2 : 4 * N1D, Nh = Nq + Nfp;
    // per-element metric, constant inside the element
    std::vector<Geo>& geo = out.geo;
    geo.assign(K, Geo{});
    const int NGEO = Nd == 1 ? 1 : 4;
    for (int k = 0; k < K; k++) {
        const double J = geom->Jq[(int64_t)Nq * k];
        for (int i = 1; i < Nq; i++)
            if (geom->Jq[i + (int64_t)Nq * k] != J)
                return fail_s(err, P2DE_E_UNSUPPORTED, "Jq varies inside an element");
        if (!(J > 0.0)) return fail_s(err, P2DE_E_INVALID, "Jq must be positive");
        double gv[4] = {0, 0, 0, 0};
        for (int gi = 0; gi < NGEO; gi++) {
            const double* G = geom->GJh + (int64_t)gi * Nh * K;
            gv[gi] = G[(int64_t)Nh * k];
            for (int h = 1; h < Nh; h++)
                if (G[h + (int64_t)Nh * k] != gv[gi])
                    return fail_s(err, P2DE_E_UNSUPPORTED, "metric terms vary inside an element");
        }
        Geo g{J, gv[0], 0.0, 0.0, 0.0};
        if (Nd == 2) {
            g.sxJ = gv[1];
            g.ryJ = gv[2];
            g.syJ = gv[3];
            if (g.sxJ != 0.0 || g.ryJ != 0.0)
                return fail_s(err, P2DE_E_UNSUPPORTED, "elements must be axis aligned (sxJ = ryJ = 0)");
        }
        // the graph-viscosity direction test |Sx0| > 1e-10 (rhs_utils.jl:122) must pick the
        // direction the pair acts in
        for (int p = 0; p < T.npl && Nd == 2; p++) {
            const double s0x = g.rxJ * T.lp_cr[p] + g.sxJ * T.lp_cs[p];
            const int vdir = std::fabs(s0x) > 1e-10 ? 0 : 1;
            if (vdir != T.lp_dir[p])
                return fail_s(err, P2DE_E_UNSUPPORTED, "graph-viscosity direction test disagrees with the pair direction");
        }
        geo[k] = g;
    }
    // face connectivity and boundary codes (StateParam.jl:1-7)
    std::vector<int>& nb_node = out.nb_node;
    std::vector<int>& nb_face = out.nb_face;
    std::vector<int>& bcode = out.bcode;
    nb_node.assign((size_t)K * Nfp, 0);
    nb_face.assign((size_t)K * Nfp, 0);
    bcode.assign((size_t)K * Nfp, 0);
    for (int k = 0; k < K; k++)
        for (int f = 0; f < Nfp; f++) {
            const int64_t m = bc->mapP[f + (int64_t)Nfp * k] - 1;
            if (m < 0 || m >= (int64_t)Nfp * K) return fail_s(err, P2DE_E_INVALID, "mapP out of range");
            const int kP = (int)(m / Nfp), iP = (int)(m % Nfp);
            if (Nd == 2 && ((f < 2 * N1D) != (iP < 2 * N1D)))
                return fail_s(err, P2DE_E_UNSUPPORTED, "mapP joins an x face to a y face");
            nb_face[(size_t)k * Nfp + f] = (int)m;
            nb_node[(size_t)k * Nfp + f] = kP * Nq + T.fq2q[iP];
        }
    if (bc->nI > 0 && (!bc->mapI || !bc->Ival)) return fail_s(err, P2DE_E_INVALID, "mapI/Ival");
    if (bc->nO > 0 && !bc->mapO) return fail_s(err, P2DE_E_INVALID, "mapO");
    for (int i = 0; i < bc->nI; i++) {
        const int64_t m = bc->mapI[i] - 1;
        if (m < 0 || m >= (int64_t)Nfp * K) return fail_s(err, P2DE_E_INVALID, "mapI out of range");
        bcode[m] = i + 1;
    }
    for (int i = 0; i < bc->nO; i++) {   // enforce_BC! applies outflow after inflow
        const int64_t m = bc->mapO[i] - 1;
        if (m < 0 || m >= (int64_t)Nfp * K) return fail_s(err, P2DE_E_INVALID, "mapO out of range");
        bcode[m] = -1;
    }

    return P2DE_OK;
}

}  // namespace p2de
