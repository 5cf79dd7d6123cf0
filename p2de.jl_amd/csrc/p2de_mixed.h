// p2de_mixed.h -- data of the mixed-precision batched path, shared by its C-ABI layer
// (p2de_batch.cpp) and its kernels (mixed_kernels.hip).
//
// The path advances many independent 2D Lobatto (LGL) quad solves together.  Flux
// differencing, surface fluxes, the low-order graph-viscosity fluxes and the element
// assembly run in FP32; node pressures (the E - |m|^2/(2 rho) cancellation), the subcell
// flux potentials, the limiter's bounds and coefficients and the RK stages run in FP64, and
// the state stays FP64 in HBM.  One launch per RK stage covers every element of every
// instance (a persistent workgroup pool striding over the tile list), plus one CFL-time-step
// launch and its per-instance reduction per step.
//
// Device layout (HBM), all instances concatenated:
//   state      double[sum_i K_i*Nq_i][4]    instance i starts at node node_off[i]
//   geo        Geo[sum_i K_i]               per element
//   nb_node    int32[...][Nfp_i]            instance-local node index of the mapP target
//   bc         int32[...][Nfp_i]            0 interior, -1 outflow, m+1 inflow (Ival row m)
//   tiles      MixTile[ntiles]              E_N consecutive elements of one instance
#pragma once
#include <stdint.h>

#include "p2de_hip.h"
#include "p2de_internal.h"

namespace p2de {

constexpr int MIX_MIN_N1D = 2;
constexpr int MIX_MAX_N1D = 5;
constexpr int MIX_CHUNKS = 8;        // contiguous tile chunks, one per XCD of the MI355X

// Compact operator tables of one degree for LGL tensor-product quads.
//
// Volume pairs are enumerated line by line: pair (a > b) of line `line` in direction d has
// index d*N1D*NP1 + line*NP1 + a*(a-1)/2 + b (NP1 = N1D*(N1D-1)/2) and joins node
// i = node(d, line, a) to j = node(d, line, b), where node(0, line, s) = s + N1D*line (an
// r-line) and node(1, line, s) = line + N1D*s (an s-line).  Low-order pairs are the line
// neighbours (s+1, s), index d*N1D*(N1D-1) + line*(N1D-1) + s.  Faces follow the reference
// layout [r=-1, r=+1, s=-1, s=+1] with N1D nodes each.
template <int N1D>
struct MixTab {
    static constexpr int NQ = N1D * N1D, NFP = 4 * N1D, NP1 = N1D * (N1D - 1) / 2;
    static constexpr int NPV = 2 * N1D * NP1, NPL = 2 * N1D * (N1D - 1);
    static constexpr int NFB = 2 * N1D * (N1D + 1);          // subcell faces per element
    double wq[NQ];
    float cH[NPV];        // Srh_db[i, j] (d = 0) or Ssh_db[i, j] (d = 1)
    float cL[NPL];        // Sr0[i, j] or Ss0[i, j]
    float a_v[NQ];        // MinvVhT[i, i]
    float a_fx[NQ];       // MinvVfT[i, f] of the r-face at node i (0 if none)
    float a_fy[NQ];       // MinvVfT[i, f] of the s-face at node i (0 if none)
    float Bf[NFP];        // Br[f, f] on r-faces, Bs[f, f] on s-faces
    // index tables of the element (the kernels read them instead of recomputing the layout):
    uint32_t pH[NPV];     // volume pair p: i | j << 8 | d << 16
    uint32_t pL[NPL];     // low-order pair p: i | j << 8 | d << 16
    uint16_t gx[NQ][N1D - 1];   // node i's volume pairs along its r-line: p | 0x8000 if i is the pair's j
    uint16_t gy[NQ][N1D - 1];   // ... along its s-line
    uint16_t nfc[NQ];     // node i's left r-line subcell face | bottom s-line subcell face << 8
    uint16_t lq[2 * N1D][N1D];  // node at position s of line L (r-lines, then s-lines)
    uint8_t fq[NFP];      // volume node of face node f
    uint8_t pad[(4 - (NFP % 4)) % 4 + 4];
};

// One instance as the kernels see it.
struct MixInst {
    int64_t node_off;     // first node of the instance in the concatenated state
    const Geo* geo;       // per element (already offset to the instance)
    const int* nb_node;
    const int* bc;
    const double* Ival;
    StepScalars* sc;
    double gamma, zeta, ZEROTOL, CFL, dt0, T;
    int K, N1D, upper_mode;
    int tile_lo, tile_hi;     // the instance's tiles in the tile list
    int pad;
};

struct MixTile {
    int inst, k0, nel, n1d;
};

struct MixArgs {
    const MixInst* inst;
    const MixTile* tiles;
    int ntiles;
    int chunk[MIX_CHUNKS + 1];  // tile ranges of the XCD chunks (cut by estimated cost)
    double* dtp;          // per-tile minimum of the CFL time step (CFL pass)
    const void* tab[MIX_MAX_N1D + 1];   // MixTab<N1D> by N1D
    int mode;             // MODE_RHS, MODE_STAGE1..3
    int only;             // evaluate only this instance (-1: all)
    const double* Uin;    // stage input
    const double* W;      // state at the start of the step
    double* Uout;         // stage output
    double* rhsU;         // MODE_RHS output
    double* dbg_rhsH;     // optional (MODE_RHS)
    double* dbg_rhsL;
};

// launchers (mixed_kernels.hip); return hipError_t as int
int mix_num_variants();
// NT, E[N1D] for N1D = 2..5 (0: the degree is not compiled into the variant)
int mix_variant_shape(int v, int* NT, int E[MIX_MAX_N1D + 1]);
int mix_blocks_per_cu(int v, int* stage_blocks, int* dt_blocks);
int mix_launch_stage(int v, const MixArgs& a, int grid, void* stream);
int mix_launch_dt(int v, const MixArgs& a, int grid, void* stream);
int mix_launch_init(const MixInst* inst, int n, int mode, int only, double dt_lim, double dt_cap, void* stream);
int mix_launch_dt_reduce(const MixInst* inst, int n, const double* dtp, int only, void* stream);
int mix_launch_post(const MixInst* inst, int n, void* stream);
int mix_launch_positivity(const double* U, int64_t nodes, double gamma, unsigned long long* keys,
                          unsigned long long* cnt, void* stream);

}  // namespace p2de

// time, last dt, steps and the positivity violations counted since the last check of one
// instance, without resetting the count (for p2de_run on a mixed-precision context)
int batch_instance_stats(p2de_batch* b, int inst, double* t, double* last_dt, int* steps, int64_t* viol);
int batch_instance_restore(p2de_batch* b, int inst, double t, double last_dt, int steps);
