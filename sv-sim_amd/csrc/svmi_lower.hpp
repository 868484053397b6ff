// svmi_lower.hpp -- lowering of reference gates to small matrices.
#pragma once

#include <string>
#include <vector>

#include "../../include/svmi.h"
#include "svmi_internal.hpp"

namespace svmi {

// A gate (or a fusion of gates) as one operator on k <= 5 qubits.
// Local index bit i <-> qubit q[i]; DENSE: m[row * 2^k + col];
// DIAG: m[idx] (2^k entries); SRN: the reference's real-linear quirk on q[0].
struct LOp {
    enum Kind { DENSE = 0, DIAG = 1, SRN = 2 } kind = DENSE;
    int k = 0;
    int q[5] = {0, 0, 0, 0, 0};
    std::vector<c64> m;
    int gate = -1;        // source gate index
    // Control qubits (bit mask, disjoint from q): the op acts only on the
    // amplitudes whose index has every bit of ctrl set (the identity
    // elsewhere).  Controlled DENSE ops have k == 1.
    uint64_t ctrl = 0;
    // >= 0: a DIAG op (k == 1) whose two entries are written on the device
    // by measurement `dyn` of the program (the collapse factors); m holds
    // placeholders and the op is never merged, folded or dropped.
    int dyn = -1;
};

// Number of qubit slots used by a reference op (1..5), -1 if unknown.
int gate_arity(int op);

// Check a gate against an n-qubit register; returns SVMI_OK or an error and
// fills err.  Mirrors Simulation::append's asserts (svsim_cpu_omp.hpp:242-255)
// and the non-cloning asserts of the two-qubit kernels.
svmi_status validate_gate(const svmi_gate& g, int n, std::string& err);

// Lower one gate to its composite matrix on its own qubits, computed by
// running the reference decomposition (svsim_cpu_omp.hpp:2016-2374) with the
// reference primitive formulas on every basis state.  Entries within 1e-14
// of 0, +-1 are snapped to those values (the reference's multi-pass rounding
// leaves residues such as S2I*S2I*2 != 1); the effect on amplitudes is at the
// 1e-16 level.
void lower_gate(const svmi_gate& g, LOp& out);

// d(matrix)/d(slot) of a gate, slot 0 = theta, 1 = phi, 2 = lambda, by
// forward-mode differentiation through the same decomposition.  Returns
// false if the gate's matrix does not depend on that slot.
bool gate_derivative(const svmi_gate& g, int slot, LOp& out);

// True if the gate's action is unitary (everything except R and SRN).
bool gate_is_unitary(int op);

// Matrix helpers.
void lop_to_dense(const LOp& a, std::vector<c64>& M);         // DIAG -> dense
void lop_adjoint(const LOp& a, LOp& out);                      // conjugate transpose
// Product b*a (apply a first) of two ops on the same ordered qubit list.
void lop_compose_same(const LOp& first, const LOp& second, LOp& out);

}  // namespace svmi
