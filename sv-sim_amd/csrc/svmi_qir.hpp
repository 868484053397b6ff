// svmi_qir.hpp -- the QIR gate set (multi-controlled single-qubit gates,
// Swap, Pauli measurement with collapse) lowered to the engine's ops.
#pragma once

#include <string>
#include <vector>

#include "../../include/svmi.h"
#include "svmi_lower.hpp"
#include "svmi_plan.hpp"

namespace svmi {

// Check one record against n allocated qubits, as Circuit::append
// (svsim_qsharp/svsim_cpu_omp.hpp:153-166: target >= n is rejected) and
// SWAP_GATE's non-cloning assert (1312-1320) do; a Swap's second qubit must
// also be < n.  Returns SVMI_OK or an error and fills err.
svmi_status validate_qir_gate(const svmi_qir_gate& g, int n, std::string& err);

// Lower a validated QIR gate list on n qubits into segments separated by
// measurements (for build_segmented_program).  Measurement m in basis
// X / Y ends segment m with the basis change (H; AdjointS then H) and starts
// segment m + 1 with its collapse op (LOp::dyn == m) and the inverse change
// (H; H then S), as Measure_GATE does (1406-1500).  A controlled op whose
// mask names its target or a qubit >= n never acts (OP_HEAD_MASK, 606-615)
// and is dropped.  A controlled diagonal diag(d0, d1) becomes the phase d0
// where all controls are 1 and the phase d1/d0 where the target is 1 too.
void lower_qir(const svmi_qir_gate* gates, size_t n_gates, int n, std::vector<std::vector<LOp>>& segments,
               std::vector<MeasureStep>& measures);

// The same gate list for the density-matrix engine
// (svsim_qsharp_noise/svsim_cpu_omp.hpp): rho of n qubits as a 2n-qubit
// state, index row | (col << n).  Every gate is lowered as U on (q, mask)
// followed by conj(U) on (q + n, mask << n) (the *_OP table, 2785-2990);
// Swap exchanges row qubits only (SWAP_GATE, 1560-1606); a measurement
// changes basis on both halves (X: H, H; Y: AdjointS, S, H, H), takes P(1)
// from the diagonal (MeasureStep::dm_n) and collapses the row qubit only
// with 1/P(1) (Measure_GATE, 1710-1830).
void lower_dm_qir(const svmi_qir_gate* gates, size_t n_gates, int n, std::vector<std::vector<LOp>>& segments,
                  std::vector<MeasureStep>& measures);

}  // namespace svmi
