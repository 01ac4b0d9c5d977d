// The reference's ia/ibk series approximations of the LOD history sum
// (TG_solver dt_alg = 4: calc_va, fr_d_t_g_2d.cpp:246-741, used by Om1 at
// 3438-3469), as a plan of per-step operations on per-node coefficient
// slots (planner in frm_iaibk.cpp, kernel in frm_lod2d.hip).
//
// For every node and direction the reference keeps lists of "long" (ia)
// and "short" (ibk) approximations of g(x, r) = (x - r + a)^A - (x - r + b)^A,
// A = 1 - alpha, whose coefficients collect the increments F[r] step by
// step.  Which approximations exist, their ranges, centres and term counts,
// and when they are created, extended, split or dropped depend only on A,
// (a, b) and the step -- never on the data -- so the structure is planned
// once on the host, and the GPU applies the same plan to every node's
// coefficients: each step is a list of operations on slots of MAXAK
// coefficients (evaluate, create from the newest increment, add the newest
// increment, copy).
#pragma once
#include <cstdint>
#include <vector>

namespace frm {

constexpr int kIaAk = 10;     // terms of a new approximation (AK)
constexpr int kIaMaxAk = 20;  // coefficients per slot (MAXAK)

enum IaOpKind : int32_t {
    kIaEvalIa = 0,   // V += sum_k<nt c[slot][k] * vals[off + k]   (x^(A-k))
    kIaEvalIbk = 1,  // V += sum_k<nt c[slot][k] * vals[off + k]   ((x - centre)^k)
    kIaExact = 2,    // V += vals[off] * F                          (the newest block, exact)
    kIaCreate = 3,   // c[slot][k] = ak[k] * (F * vals[off + k]), k < nt; 0 above
    kIaAdd = 4,      // c[slot][k] += ak[k] * (F * vals[off + k]), k < nt
    kIaCopy = 5      // c[slot][:] = c[src][:]
};

struct IaOp {
    int32_t kind, slot, src, nt, off;
};

struct IaPlan {
    int nslots = 0;              // slots alive at once, at most
    std::vector<int> step_off;   // ops of step n (n >= 1): [step_off[n-1], step_off[n])
    std::vector<IaOp> ops;
    std::vector<double> vals;
    double ak[kIaMaxAk] = {};    // ak[k] = A (A-1) ... (A-k+1) / k!
};

// The plan of one direction (a, b) = (3/2, 1/2) for rows, (2, 1) for
// columns, for the steps n = 1 .. steps (step n folds block n - 1 and
// evaluates at x = n).
IaPlan iaibk_plan(double A, double a, double b, int steps);

}  // namespace frm
