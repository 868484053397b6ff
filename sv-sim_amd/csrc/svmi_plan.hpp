// svmi_plan.hpp -- scheduling of lowered ops into device sweeps.
#pragma once

#include <string>
#include <vector>

#include "svmi_internal.hpp"
#include "svmi_lower.hpp"

namespace svmi {

struct PlanOptions {
    bool fusion = true;        // tiled multi-op sweeps; false = one pass per op
    int tile_qubits = 12;      // k: qubits per LDS tile (13 allowed: one 128 KiB tile per CU)
    int low_qubits = 4;        // c: physical qubits 0..c-1 in every tile
    long sweep_budget = 0;     // per-amplitude FMA budget of one sweep (0 = none)
    long op_overhead = 2;      // fusion cost model: per-op cost on top of its FMAs (per amplitude)
    long coef_budget = 16000;  // split a sweep whose op records and coefficients exceed this many bytes
                               // (0 = never)
    bool rotations = true;     // one-state phases: one-qubit unitaries as lifting rotations between
                               // merged register tables where the cost model prefers it
    int merge_qubits = 2;      // widest dense block the register-run merging may form (1..4)
    bool tile_map = true;      // tile-local CX and SWAP ops outside a phase's register bits become
                               // a relabelling of the tile's LDS indices, not a phase
};

// One step of a program: a tiled sweep or a standalone full-state op.
struct Step {
    enum Kind { SWEEP = 0, STANDALONE = 1, MEASURE = 2 } kind;
    uint32_t index;
};

// A mid-circuit measurement: P(1) of `qubit` over the state, the outcome
// draw < P(1), and the collapse factors written into coefficient slots
// coef, coef + 1 (the DIAG op with LOp::dyn == this measurement's index,
// which runs first in the next segment).
struct MeasureStep {
    uint32_t qubit = 0;
    uint32_t coef = UINT32_MAX;
    uint32_t n = 0;        // P(1) sums over the first 2^n amplitudes
    uint32_t dm_n = 0;     // > 0: density matrix of dm_n qubits (2^n = 4^dm_n entries, index
                           // row | col << dm_n): P(1) = sum of |Re rho[i, i]| over rows i with
                           // bit `qubit` set, collapse factors 1/P(1) or 1/(1 - P(1))
    double draw = 0.0;
};

// Host copy of a device program.
struct Program {
    int n = 0;
    int n_states = 1;                  // 2 for adjoint sweeps (phi, lambda)
    std::vector<Step> steps;
    std::vector<SweepInfo> sweeps;
    std::vector<PhaseInfo> phases;
    std::vector<DOp> ops;
    std::vector<c64> coef;
    std::vector<uint64_t> hitab;
    std::vector<StandaloneOp> standalone;
    std::vector<MeasureStep> measures;
    // adjoint: accumulator g of sweep s -> gradient parameter index
    std::vector<int32_t> grad_param;
    uint32_t n_grad = 0;
    uint64_t n_lowered = 0;            // ops after lowering + peephole fusion
};

// Merge runs of single-qubit ops on one qubit (no other op on that qubit in
// between) into one op.  Ops on other qubits commute past them.
void fuse_single_qubit_runs(const std::vector<LOp>& in, std::vector<LOp>& out);

// Forward program for n qubits.
void build_forward_program(const std::vector<LOp>& ops, int n, const PlanOptions& opt, Program& prog);

// Program of segments separated by measurements: segments[i] is scheduled
// on its own (no op moves across a measurement), then measurement i runs if
// i < measures.size().  segments[i + 1] must contain the collapse op of
// measurement i (a DIAG op with dyn == i).
void build_segmented_program(const std::vector<std::vector<LOp>>& segments, const std::vector<MeasureStep>& measures,
                             int n, const PlanOptions& opt, Program& prog);

// An entry of an adjoint (backward) schedule: the inverse of one gate applied
// to phi and lambda, or a gradient accumulation <lam| dU |phi> between them.
struct AdjEntry {
    LOp op;             // U^dagger (targets phi and lambda) or dU (grad)
    bool grad = false;
    int param = -1;     // gradient parameter (grad entries)
};

// Backward program over (phi, lambda); entries in execution order.
void build_adjoint_program(const std::vector<AdjEntry>& entries, int n, const PlanOptions& opt,
                           Program& prog);

// FP64 multiply-adds per amplitude of the program's state ops.
double program_fma_per_amp(const Program& prog);

// Human-readable summary (sweeps, phases, ops per sweep).
std::string describe_program(const Program& prog);

}  // namespace svmi
