/*
 * svsim_oracle.h -- CPU oracle for the state-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU-OpenMP engine (svsim/src/svsim_cpu_omp.hpp) used to check the HIP
 * engine.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Conventions follow the reference exactly:
 *   - state is structure-of-arrays: re[2^n], im[2^n] (cpu_omp:199-220)
 *   - qubit q is bit q of the amplitude index (cpu_omp:620-628)
 *   - gate records carry the reference enum value and the reference slot
 *     layout {qb0..qb4, theta, phi, lambda} (cpu_omp:103-194, 377-564)
 *   - every primitive keeps the reference's expression order; build with
 *     -ffp-contract=off so results are bit-identical to the reference.
 * Indices are 64-bit (the reference's 32-bit IdxType overflows at n >= 29).
 */
#ifndef SVSIM_ORACLE_H
#define SVSIM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same numbering as the reference enum OP (svsim_cpu_omp.hpp:46-52). */
enum orc_op {
    ORC_U3, ORC_U2, ORC_U1, ORC_CX, ORC_ID, ORC_X, ORC_Y, ORC_Z, ORC_H, ORC_S,
    ORC_SDG, ORC_T, ORC_TDG, ORC_RX, ORC_RY, ORC_RZ, ORC_CZ, ORC_CY, ORC_SWAP, ORC_CH,
    ORC_CCX, ORC_CSWAP, ORC_CRX, ORC_CRY, ORC_CRZ, ORC_CU1, ORC_CU3, ORC_RXX, ORC_RZZ, ORC_RCCX,
    ORC_RC3X, ORC_C3X, ORC_C3SQRTX, ORC_C4X, ORC_R, ORC_SRN, ORC_W, ORC_RYY,
    ORC_NUM_OPS
};

/* Gate record: identical layout to svmi_gate in include/svmi.h. */
typedef struct orc_gate {
    int32_t op;
    int32_t qb[5];
    double theta, phi, lambda;
} orc_gate;

/* Pauli term: coeff * i^{|x&z|} X^x Z^z (Y = iXZ on a qubit with both bits). */
typedef struct orc_pauli_term {
    double coeff;
    uint64_t xmask;
    uint64_t zmask;
} orc_pauli_term;

/* Number of qubit slots a gate uses (1..5) or -1 for an unknown op. */
int orc_gate_arity(int op);

/* Apply gates[0..ng) in order to the state, as Simulation::sim does
 * (cpu_omp:288-319, 590-596).  nthreads <= 0 means omp default.
 * Returns 0, or -1 on an invalid gate (bad op, qubit >= n, repeated qubit). */
int orc_apply_gates(int n, double* re, double* im,
                    const orc_gate* gates, size_t ng, int nthreads);

/* Sample shots with fed uniform draws r in [0,1] following
 * Simulation::measure (cpu_omp:339-361): sequential inclusive scan, result j
 * with scan[j] <= r < scan[j+1], 0 if none.  Uses a binary search over the
 * same scan (equivalent result; see orc_measure_linear).  Returns sum of
 * probabilities (scan[dim]). */
double orc_measure(int n, const double* re, const double* im,
                   const double* uniforms, size_t shots, uint64_t* out);

/* The reference's O(shots * 2^n) linear search, kept for small-n tests. */
double orc_measure_linear(int n, const double* re, const double* im,
                          const double* uniforms, size_t shots, uint64_t* out);

/* Write the sequential scan scan[0..dim] (dim+1 doubles). */
void orc_prob_scan(int n, const double* re, const double* im, double* scan);

/* <psi| sum_t coeff_t P_t |psi> (real part). No reference counterpart:
 * the textbook definition, evaluated term by term. */
double orc_expval(int n, const double* re, const double* im,
                  const orc_pauli_term* terms, size_t nterms, int nthreads);

/* out = H|psi> for the same Pauli sum. */
void orc_apply_pauli_sum(int n, const double* re, const double* im,
                         const orc_pauli_term* terms, size_t nterms,
                         double* out_re, double* out_im);

#ifdef __cplusplus
}
#endif

#endif
