/*
 * qir_oracle.h -- CPU oracle for the QIR gate set: multi-controlled
 * single-qubit gates and mid-circuit Pauli measurement with collapse.
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C restatement of the reference's
 * QIR-facing engine, svsim_qsharp/svsim_cpu_omp.hpp, used to check the HIP
 * engine.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Conventions follow that engine:
 *   - state is structure-of-arrays: re[2^n], im[2^n] (cpu_omp:215-232);
 *     only the first 2^n amplitudes are touched (Simulation::update,
 *     cpu_omp:452-460, sizes the run by the allocated qubits);
 *   - gate record {op, qubit, mask, theta} (cpu_omp:113-147): for the
 *     Controlled* ops `mask` holds the control qubits as bits, for Swap it
 *     holds the second qubit, for Measure it holds the Pauli basis (0 Z,
 *     1 X, 2 Y) and theta holds the uniform draw;
 *   - a controlled op acts on the pairs whose pos0 has every mask bit set
 *     (OP_HEAD_MASK, cpu_omp:606-615), so a mask naming the target or a
 *     qubit >= n makes it the identity;
 *   - every primitive keeps the reference's expression order; build with
 *     -ffp-contract=off so results are bit-identical to the reference.
 */
#ifndef QIR_ORACLE_H
#define QIR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same numbering as the reference enum OP (svsim_qsharp/svsim_cpu_omp.hpp:45-59). */
enum orc_qir_op {
    ORQ_X, ORQ_Y, ORQ_Z, ORQ_H, ORQ_S, ORQ_T,
    ORQ_RI, ORQ_RX, ORQ_RY, ORQ_RZ, ORQ_EI, ORQ_EX, ORQ_EY, ORQ_EZ,
    ORQ_CX, ORQ_CY, ORQ_CZ, ORQ_CH, ORQ_CS, ORQ_CT,
    ORQ_CRI, ORQ_CRX, ORQ_CRY, ORQ_CRZ, ORQ_CEI, ORQ_CEX, ORQ_CEY, ORQ_CEZ,
    ORQ_SDG, ORQ_TDG, ORQ_CSDG, ORQ_CTDG,
    ORQ_SWAP, ORQ_MEASURE,
    ORQ_NUM_OPS
};

/* Gate record: identical layout to svmi_qir_gate in include/svmi.h (24 B). */
typedef struct orc_qir_gate {
    int32_t op;
    int32_t qubit;
    uint64_t mask;
    double theta;
} orc_qir_gate;

/* Apply gates[0..ng) to the first 2^n amplitudes, as Simulation::sim does
 * (cpu_omp:467-505, 562-567).  For the m-th Measure record, probs[m]
 * receives its probability of |1> and outcomes[m] = (theta < probs[m]), the
 * result the QIR wrapper reports (svsim_qsharp_wrapper.cpp:520-555); either
 * may be NULL.  *last_prob (may be NULL) plays the role of m_real[0]: it is
 * overwritten by every Measure and left alone otherwise.  nthreads <= 0
 * means the OpenMP default.  Returns 0, or -1 (nothing applied) on an
 * unknown op, a target >= n, or a Swap whose second qubit is >= n or equal
 * to the first. */
int orc_qir_apply(int n, double* re, double* im, const orc_qir_gate* gates, size_t ng,
                  double* last_prob, double* probs, int32_t* outcomes, int nthreads);

/* The density-matrix engine (svsim_qsharp_noise/svsim_cpu_omp.hpp) on n
 * qubits: re/im hold rho as 2^(2n) entries, index row | (col << n); same
 * records, probs/outcomes as above.  n <= 20.  Returns 0 or -1 as above. */
int orc_dm_qir_apply(int n, double* re, double* im, const orc_qir_gate* gates, size_t ng, double* probs,
                     int32_t* outcomes, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
