// mechgen.h -- mechanism compiler: turns a mass-action mechanism into
// straight-line, sparsity-specialised code for one Rosenbrock23 attempt on one
// grid cell (RHS, analytic Jacobian, W = I - h*gamma*J, static sparse LU
// without pivoting, triangular solves) and for its hand-derived adjoint.
//
// The reference builds the per-cell block Jacobian numerically or symbolically
// and solves it with a dense pivot-free elimination (rx_traced_integrator.jl:
// 110-135, 151-253).  Here the structure is compiled once per mechanism: the
// emitted code touches only structural non-zeros of W and of its LU factors,
// in a fill-reducing (Markowitz) elimination order, so the per-cell state stays
// in registers on the GPU.  The same text also compiles as plain C++ for the
// CPU-side check of the generator against the oracle.
#pragma once

#include <string>
#include <vector>

namespace rs {

struct MechSpec {
    int ns = 0, nr = 0;
    std::vector<int> ra, rb, pa, pb, photo;    // reactants a,b (-1), products a,b (-1)
    // General products (empty: pa/pb with coefficient 1): per reaction a list
    // of (species, coefficient), coefficients real (e.g. 0.87 CH2O), a
    // species may appear on both sides (a catalyst).
    std::vector<std::vector<std::pair<int, double>>> prod;
};

// Operation counts of the emitted code (FP64 flops; an FMA counts 2).
struct MechWork {
    long long nnz_J = 0, nnz_LU = 0;
    long long f_rhs = 0, f_jac = 0, f_w = 0, f_lu = 0, f_solve = 0, f_solveT = 0;
    long long f_jtv = 0, f_hess = 0, f_vec_attempt = 0, f_vec_vjp = 0;
    long long attempt() const { return 3 * f_rhs + f_jac + f_w + f_lu + 3 * f_solve + f_vec_attempt; }
    long long vjp() const {
        return 2 * f_rhs + f_jac + f_w + f_lu + 2 * f_solve + 2 * f_solveT + 2 * f_jtv + f_hess +
               f_vec_vjp;
    }
};

struct MechPlan {
    int ns = 0, nr = 0;
    int ordering = 1;                           // 0 natural, 1 Markowitz
    std::vector<std::vector<std::pair<int, double>>> net;   // per reaction: (species, nu != 0)
    std::vector<int> perm;                      // elimination position -> species
    std::vector<int> pos;                       // species -> position
    std::vector<std::vector<int>> idx;          // [row pos][col pos] -> W slot, -1 = zero
    std::vector<std::vector<char>> jpat;        // [species i][species a] J structurally != 0
    int nnz = 0;
    MechWork work;
};

// Validates the spec; returns an empty string on success, else a message.
std::string validate_mech(const MechSpec& m);
MechPlan plan_mechanism(const MechSpec& m, int ordering);
// Device/host functions of the plan (no kernels).  The spec is needed as well as
// the plan: catalysts (nu == 0) are absent from the net stoichiometry but set
// the reaction order.
std::string generate_mech_source(const MechSpec& m, const MechPlan& p);

}  // namespace rs
