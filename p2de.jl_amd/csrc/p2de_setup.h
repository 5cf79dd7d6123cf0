// p2de_setup.h -- host-side validation of the operators and mesh the caller passes, and the
// operator tables and connectivity arrays the kernels read.  Shared by the single-solve C-ABI
// (p2de_capi.cpp) and the batched one (p2de_batch.cpp).
#pragma once
#include <string>
#include <vector>

#include "p2de_hip.h"
#include "p2de_internal.h"

namespace p2de {

struct MeshArrays {
    std::vector<Geo> geo;       // per element
    std::vector<int> nb_node;   // [K][Nfp] kP*Nq + fq2q[iP] of the mapP target
    std::vector<int> nb_face;   // [K][Nfp] kP*Nfp + iP
    std::vector<int> bcode;     // [K][Nfp] 0 interior, -1 outflow, m+1 inflow
};

// Checks the option set of a configuration (Solver.jl:1-121) against what the build
// implements and the basic sizes; fills Nq/Nfp/Nh/K.
int setup_config(const p2de_config* cfg, int* Nq, int* Nfp, int* Nh, int* K, std::string& err);

// Operator tables of the reference element from Operators{DIM} (Solver.jl:192-212) for the
// basis the configuration names (P2DE_BASIS_*); returns P2DE_E_UNSUPPORTED for operator
// structures outside the tensor-product Lobatto / Gauss layouts.
int setup_tables(ElemTables& T, int Nd, int N1D, int basis, const p2de_operators* ops, std::string& err);

// Per-element metric, face connectivity and boundary codes from GeomData and BCData
// (Solver.jl:185-190, StateParam.jl:1-7).
int setup_mesh(int Nd, int N1D, int K, const ElemTables& T, const p2de_geometry* geom, const p2de_bcdata* bc,
               MeshArrays& out, std::string& err);

}  // namespace p2de
