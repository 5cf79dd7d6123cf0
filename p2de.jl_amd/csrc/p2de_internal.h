// p2de_internal.h -- data shared between the C-ABI layer (p2de_capi.cpp) and the kernels
// (rhs_kernels.hip).
//
// Device data layout (HBM):
//   state      double[K][Nq][Nc]           Julia's Array{SVector{Nc},2}(Nq, K)
//   rhsUA      double[K][Nq][Nc]           element RHS with pre-symmetrisation limiter
//   lA         double[K][Nfp]              limiter parameter of each element-boundary
//                                          subcell face (one per face node)
//   dfA        double[K][Nfp][Nc]          f_bar_H - f_bar_L at those faces
//   ut         double[K][Nfp][Nc]          projected face states u(v(U)) of the current state
//   theta      double[K][Nfp]              entropy-projection limiting parameters (Gauss basis)
//   geo        Geo[K]                      per-element metric (constant inside an element)
//   nb_node    int32[K][Nfp]               kP*Nq + fq2q[iP] of the mapP target
//   nb_face    int32[K][Nfp]               kP*Nfp + iP of the mapP target
//   bc         int32[K][Nfp]               0 interior, -1 outflow, m+1 inflow (Ival row m)
//   tables     ElemTables                  operator tables shared by all elements
#pragma once
#include <stdint.h>

namespace p2de {

constexpr int MAX_N1D = 5;
constexpr int MAX_NQ = 25;
constexpr int MAX_NFP = 20;
constexpr int MAX_NH = 45;
constexpr int MAX_NPH = 200;      // Gauss 2D N=4: 100 line pairs + 100 face-volume pairs
constexpr int MAX_HF = 4;         // faces whose interpolation row reaches a node (Gauss 2D)
constexpr int MAX_NPL = 40;

// Per-element metric terms.  The reference stores them per node (GeomData, Solver.jl:185-190);
// on its affine uniform meshes they are constant inside an element, which p2de_create checks.
struct Geo {
    double J, rxJ, sxJ, ryJ, syJ;
};

// Operator tables of the reference element, built once from the operators the caller
// passes (Operators, Solver.jl:192-212).  Indices are 0-based.
struct ElemTables {
    int Nd, Nc, N1D, Nq, Nfp, Nh, nph, npl;
    int basis;                    // 0 Lobatto (Vf selects fq2q), 1 Gauss (Vf interpolates along the line)
    // high-order pairs, Srsh_nnz order (init.jl:188-195): hybrid nodes i > j, the metric
    // direction the pair acts in, and Srh_db[i,j], Ssh_db[i,j]
    uint8_t hp_i[MAX_NPH], hp_j[MAX_NPH], hp_dir[MAX_NPH];
    double hp_cr[MAX_NPH], hp_cs[MAX_NPH];
    // low-order pairs, Srs0_nnz order (init.jl:197-204): Sr0[i,j], Ss0[i,j]
    uint8_t lp_i[MAX_NPL], lp_j[MAX_NPL], lp_dir[MAX_NPL];
    double lp_cr[MAX_NPL], lp_cs[MAX_NPL];
    // QF1 gather lists per hybrid node, in pair order: pair index and sign (+1 if the
    // node is the pair's i, -1 if it is j)
    int16_t hl_start[MAX_NH + 1];
    int16_t hl_pair[2 * MAX_NPH];
    int8_t hl_sign[2 * MAX_NPH];
    // Q0F1 gather lists per volume node, in pair order
    int16_t ll_start[MAX_NQ + 1];
    int16_t ll_pair[2 * MAX_NPL];
    int8_t ll_sign[2 * MAX_NPL];
    // CFL lists: the low-order pairs of node i ordered by partner index j, so
    // sum_j lambda[i,j] adds in the reference's j order (low_order_graph_viscosity.jl:271-273)
    int16_t cl_start[MAX_NQ + 1];
    int16_t cl_pair[2 * MAX_NPL];
    // node data
    double wq[MAX_NQ];
    double rwq[MAX_NQ];           // 1 / wq[i] (the scaling of MinvVhT_new, flux_differencing.jl:304)
    double a_v[MAX_NQ];           // MinvVhT[i, i]
    int8_t nf[MAX_NQ];            // faces f with fq2q[f] == i (ascending): the node's element-boundary
    int8_t nface[MAX_NQ][2];      // subcell faces and low-order surface terms
    // faces f whose row of Vf or Vf_low reaches node i (ascending), for the assembly
    // MinvVhT*QF1 + MinvVfT*BF_H: Lobatto the nface list, Gauss the four line ends of its lines
    int8_t nh[MAX_NQ];
    int8_t hface[MAX_NQ][MAX_HF];
    double a_h[MAX_NQ][MAX_HF];   // MinvVhT[i, Nq+f]
    double a_f[MAX_NQ][MAX_HF];   // MinvVfT[i, f]
    double vf_h[MAX_NQ][MAX_HF];  // Vf[f, i]       (limited assembly, flux_differencing.jl:310-319)
    double vfl_h[MAX_NQ][MAX_HF]; // Vf_low[f, i]
    int8_t nq2fq[MAX_NQ];         // q2fq[i] (CFL face order)
    int8_t q2fq[MAX_NQ][2];
    // face data
    uint8_t fq2q[MAX_NFP];
    double vf_coef[MAX_NFP];      // Lobatto: (1*Vf + 0*Vf_low)[f, fq2q[f]], the face row's only nonzero
    // the line of face node f (Gauss): its nodes in ascending order, Vf and Vf_low on them
    uint8_t fl_node[MAX_NFP][MAX_N1D];
    double fl_vf[MAX_NFP][MAX_N1D];
    double fl_vfl[MAX_NFP][MAX_N1D];
    double Br[MAX_NFP], Bs[MAX_NFP];
    int8_t ld[MAX_NFP];           // direction the LF term is applied to (rhs_utils.jl:78-102)
    int8_t bsign[MAX_NFP];        // -1: left/bottom element-boundary subcell face, +1 right/top
    // limiter options
    double vdm_inv[MAX_NQ][MAX_NQ];   // VDM_inv[i][j] (smoothness indicator, shock_capture.jl:55-57)
    // low-order stencil of each node (low_order_stencil limiter_utils.jl:214-231), in the order
    // (left, right[, bottom, top]): >= 0 a node of the element, < 0 -(f+1) for the node across face f
    int8_t stn[MAX_NQ][4];
};

// Device-side time-stepping scalars.
struct StepScalars {
    double t;
    double dt_lim;                // dt the limiter uses (time_param.dt)
    unsigned long long dt_bits;   // CFL dt accumulator (positive double bit pattern), then the step dt
    int done;                     // t >= T: kernels of the step do nothing
    int steps;
    int hold;                     // a history point was reached: later steps do nothing until cleared
    unsigned long long viol;      // positivity violations
    unsigned long long first_bad; // min over (stage<<40 | node) of offending nodes
    double s_min;                 // minimum modified specific entropy of the state at t0 (min-entropy bounds)
    double last_dt;               // the last step's dt (batched path: dt_bits already holds the next step's cap)
};

struct PhaseAArgs {
    const double* U;
    const double* ut;             // projected face states u_tilde, double[K][Nfp][Nc]
    const double* theta;          // Gauss with the entropy-projection limiter: double[K][Nfp] (else nullptr)
    const ElemTables* T;
    const Geo* geo;
    const int* nb_node;
    const int* nb_face;
    const int* bc;
    const double* Ival;
    StepScalars* sc;
    int K;
    double gamma, zeta, ZEROTOL, CFL;
    int upper_mode;
    int stage1;                   // accumulate the CFL dt (nstage == 1)
    double* rhsUA;
    double* lA;
    double* dfA;
    double* dbg_rhsH;             // optional debug outputs (nullptr: off)
    double* dbg_rhsL;
    double* dbg_L;                // L_local[:, :, :] of this stage, (Nq+N1D, Nd, K)
    double* zsL;                  // Zhang-Shu: L[:, nstage] of this stage, [K] (else nullptr)
    const double* sfac;           // smoothness factor per element (relaxed cell-entropy bound), [K]
    double cell_beta;             // PositivityAndRelaxedCellEntropyBound(beta)
    // TVD bounds: the first launch (tvd_pre) writes rho_L = rho + dt*rhsL_rho of every node to
    // rhoL_out and stops; the second reads the node stencils' extrema from rhoL
    const double* rhoL;
    double* rhoL_out;
    int tvd_pre;
    double* dtp;                  // stage 1: per-workgroup minimum of the CFL time step
    // limiter options (element-kernel variants OPT 1 and 2)
    int bound;                    // P2DE_*_BOUND: entropy part bound % 5, TVD density bounds when >= 5
    double POSTOL;
    const double* blend;          // per element blending factor of the shock capture (nullptr: 1)
    const double* lphi;           // per node lower bound of the modified specific entropy
    int lo_proj;                  // low-order surface flux LaxFriedrichsOnProjectedVal (else OnNodalVal)
};

// element-kernel forms: 0 SubcellLimiter(PositivityBound, NoShockCapture); 1 SubcellLimiter with
// the min-entropy bounds and/or shock capture; 2 ZhangShuLimiter
enum { OPT_SUBCELL = 0, OPT_SUBCELL_GENERAL = 1, OPT_ZHANG_SHU = 2 };

enum { MODE_RHS = 0, MODE_STAGE1 = 1, MODE_STAGE2 = 2, MODE_STAGE3 = 3 };

struct PhaseBArgs {
    const double* rhsUA;
    const double* lA;
    const double* dfA;
    const int* nb_face;
    const ElemTables* T;
    const Geo* geo;
    StepScalars* sc;
    int K;
    double gamma;
    int mode;
    const double* W;              // state at the start of the step (stage 2/3), or the input (stage 1)
    const double* Uin;            // stage input
    double* Uout;                 // stage output (may alias Uin or W elementwise)
    double* rhsU;                 // MODE_RHS output
    double* dbg_L;
    double* ut;                   // projected face states of Uout (stage modes; nullptr: not formed here)
    int finalize;                 // stage 3: advance t
    // history segments (finalize only): each step's dt goes to dt_log[steps - steps_base - 1]
    // while that index is below dt_cap; hold is set after the step whose index
    // i = i0 + steps - steps_base is a history point of the time loop, or that fills dt_log
    double* dt_log;
    int dt_cap;
    int steps_base;
    long long i0;
    int out_interval;             // 0: no history segment
    double T_end;                 // final time T
};

// Per stage before the element kernel, for the limiter options: smoothness indicator, blending
// factor, smoothness factor and the lower entropy bound of every node.
struct PrepArgs {
    const double* U;
    const ElemTables* T;
    const int* nb_node;
    StepScalars* sc;
    int K, N, stage;
    int shock, bound;             // P2DE_*_SHOCK_CAPTURE, P2DE_POSITIVITY*_BOUND
    double gamma, shock_a, shock_c;
    double* blend;                // [K] (shock capture)
    double* lphi;                 // [K][Nq] (min-entropy bounds)
    double* sfac;                 // [K] smoothness factor (relaxed cell-entropy bound), may be nullptr
};

// Entropy projection of the face states (Gauss basis): theta per face node and u_tilde.
struct EPArgs {
    const double* U;
    const ElemTables* T;
    StepScalars* sc;
    int K;
    int limit;                    // NodewiseScaledExtrapolation (else theta = 1)
    double gamma, zeta, eta, POSTOL;
    double* ut;                   // double[K][Nfp][Nc]
    double* theta;                // double[K][Nfp]
};

// launchers (rhs_kernels.hip); return hipError_t as int
// basis: 0 Lobatto, 1 Gauss (phase-A variants are per basis)
int launch_phaseA(int Nd, int N1D, int basis, int variant, int opt, const PhaseAArgs& a, void* stream);
int launch_limiter_prep(int Nd, int N1D, const PrepArgs& a, void* stream);
// s_min = min over all nodes of rhoe*rho^-gamma, when t == t0 (t from *tp or t_value) at stage 1
int launch_s_min(const double* U, int64_t nodes, int Nd, double gamma, const double* tp, double t_value, double t0,
                 double* part, StepScalars* sc, void* stream);
int num_variants_A(int Nd, int N1D, int basis);
int variant_shape_A(int Nd, int N1D, int basis, int variant, int* E, int* NT);
int launch_entropy_projection(int Nd, int N1D, const EPArgs& a, void* stream);
int launch_phaseB(int Nd, int N1D, const PhaseBArgs& a, void* stream);
int launch_face_state(int Nd, int N1D, const double* U, const ElemTables* T, int K, double gamma, double* ut,
                      void* stream);
int launch_step_init(StepScalars* sc, double CFL, double dt0, double T, void* stream);
int launch_rhs_init(StepScalars* sc, double dt_lim, double dt_cap, void* stream);
// dt = min(cap, min_b dtp[b]) over the phase-A workgroups (no global atomics)
int launch_dt_reduce(StepScalars* sc, const double* dtp, int n, void* stream);
int phaseA_blocks(int Nd, int N1D, int basis, int variant, int K);
int launch_positivity(const double* U, int64_t nodes, int Nd, double gamma, unsigned long long* keys,
                      unsigned long long* cnt, void* stream);

}  // namespace p2de
