/*
 * svmi.h -- C ABI of the MI355X state-vector engine (libsvmi.so).
 *
 * The engine replaces the hot path of the reference SV-Sim engines
 * (svsim/src/svsim_cpu_omp.hpp, svsim/src/svsim_amdgpu_sin.hpp): applying
 * the reference gate set to a 2^n complex FP64 amplitude array, then
 * measuring and sampling.  The reference's C++ Simulation class and its
 * pybind11 module (svsim/src/py_wrapper.cpp) sit on top of these entry points
 * (sv-sim_amd/csrc/svsim_amd.hpp, sv-sim_amd/csrc/py_wrapper.cpp).
 *
 * Conventions (identical to the reference):
 *   - qubit q is bit q of the amplitude index;
 *   - gate records use the reference enum OP numbering and slot layout
 *     (svsim_cpu_omp.hpp:46-52, 103-194): U3(theta,phi,lambda,q),
 *     U2(phi,lambda,q), U1(lambda,q), RX/RY(theta,q), RZ(phi,q),
 *     CRX/CRY/CRZ/CU1(lambda,c,t), CU3(theta,phi,lambda,c,t),
 *     RXX/RYY/RZZ/R(theta,...); controls first;
 *   - every gate has the reference's exact action, including global phases
 *     (RZ == U1, RZZ == CX.U1.CX, ...) and the non-unitary R and SRN.
 * Every call returns an svmi_status; svmi_last_error() describes the last
 * failure on the calling thread.  A handle owns its device state and one HIP
 * stream and must not be used from two threads at once.
 */
#ifndef SVMI_C_ABI_H_
#define SVMI_C_ABI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVMI_ABI_VERSION 2

typedef enum svmi_status {
    SVMI_OK = 0,
    SVMI_ERR_ARG = 1,        /* null pointer, bad size or option           */
    SVMI_ERR_QUBIT = 2,      /* qubit index >= n or repeated in one gate   */
    SVMI_ERR_OP = 3,         /* unknown gate opcode                        */
    SVMI_ERR_STATE = 4,      /* call out of order (run before upload, ...) */
    SVMI_ERR_DEVICE = 5,     /* no such device / HIP failure               */
    SVMI_ERR_OOM = 6,        /* device allocation failed                   */
    SVMI_ERR_UNSUPPORTED = 7 /* e.g. adjoint through a non-unitary gate    */
} svmi_status;

/* Reference enum OP (svsim_cpu_omp.hpp:46-52), same values. */
typedef enum svmi_op {
    SVMI_U3, SVMI_U2, SVMI_U1, SVMI_CX, SVMI_ID, SVMI_X, SVMI_Y, SVMI_Z, SVMI_H, SVMI_S,
    SVMI_SDG, SVMI_T, SVMI_TDG, SVMI_RX, SVMI_RY, SVMI_RZ, SVMI_CZ, SVMI_CY, SVMI_SWAP, SVMI_CH,
    SVMI_CCX, SVMI_CSWAP, SVMI_CRX, SVMI_CRY, SVMI_CRZ, SVMI_CU1, SVMI_CU3, SVMI_RXX, SVMI_RZZ, SVMI_RCCX,
    SVMI_RC3X, SVMI_C3X, SVMI_C3SQRTX, SVMI_C4X, SVMI_R, SVMI_SRN, SVMI_W, SVMI_RYY,
    SVMI_NUM_OPS
} svmi_op;

/* One gate: the reference Gate record (svsim_cpu_omp.hpp:103-194) without
 * its function pointer.  Unused qubit slots must be valid qubits (the
 * reference factories set them to 0; append() asserts all five < n). */
typedef struct svmi_gate {
    int32_t op;
    int32_t qb[5];
    double theta;
    double phi;
    double lambda;
} svmi_gate;

/* Pauli-string term coeff * P with P = i^{|x&z|} X^x Z^z: bit q of xmask
 * and zmask selects X (x only), Z (z only) or Y (both) on qubit q. */
typedef struct svmi_pauli_term {
    double coeff;
    uint64_t xmask;
    uint64_t zmask;
} svmi_pauli_term;

typedef struct svmi_sim* svmi_handle;

/* Execution statistics of the last run (for benchmarks and tests). */
typedef struct svmi_stats {
    uint64_t sweeps;         /* full state passes executed on the device      */
    uint64_t kernel_launches;
    uint64_t ops;            /* device ops after lowering and fusion          */
    uint64_t gates;          /* reference gates in the uploaded circuit       */
    double device_ms;        /* HIP-event time of the last svmi_run           */
    /* with SVMI_OPT_PROFILE: per-launch HIP-event times of the last run      */
    uint64_t sweep_launches; /* tiled sweep kernels launched                  */
    double sweep_ms;         /* summed duration of those launches             */
    uint64_t other_launches; /* standalone-op kernels                         */
    double other_ms;
    /* with SVMI_OPT_SPECIALIZE: sweeps of the uploaded circuit that run as
     * compiled straight-line kernels (0 if compilation failed: the sweeps are
     * then interpreted, and svmi_last_error names the failure) and the hipRTC
     * compile time of the last upload                                         */
    uint64_t specialized_sweeps;
    double compile_ms;
} svmi_stats;

/* Engine options (svmi_set_option). */
typedef enum svmi_option {
    SVMI_OPT_FUSION = 0,      /* 1 = tiled multi-gate sweeps (default), 0 = one pass per op */
    SVMI_OPT_TILE_QUBITS = 1, /* qubits per LDS tile (4..13, default 12)                    */
    SVMI_OPT_LOW_QUBITS = 2,  /* low qubits every tile keeps for coalescing (2..6, default 4) */
    SVMI_OPT_SWEEP_FLOPS = 3, /* per-amplitude FMA budget of one sweep (0 = unlimited)      */
    SVMI_OPT_PROFILE = 4,     /* 1 = time every launch with HIP events (svmi_stats)         */
    SVMI_OPT_PIPELINE = 5,    /* 1 = persistent sweeps prefetching the next tile,
                                 2 = the same with one 512-VGPR workgroup per CU, 0 = off (default) */
    SVMI_OPT_OP_COST = 6,     /* fusion cost model: per-op cost in FMA per amplitude (default 2;
                                 0 for SVMI_OPT_SPECIALIZE unless set)                      */
    SVMI_OPT_COEF_BUDGET = 7, /* split a forward sweep whose op records and coefficients exceed
                                 this many bytes into consecutive sweeps (default 16000, 0 = never) */
    SVMI_OPT_SPECIALIZE = 8,  /* 1 = compile each uploaded circuit's sweeps into straight-line
                                 kernels with hipRTC at upload (run-time compilation, seconds
                                 per circuit); 0 = interpret the sweep programs (default)     */
    SVMI_OPT_MERGE_QUBITS = 9, /* widest dense block (1..4 register qubits) that fusion may
                                 form from a run of gates inside one phase (default 2)       */
    SVMI_OPT_PREFETCH = 10,   /* 1 = one-state sweeps of 12-qubit tiles over large states run
                                 persistent with two LDS tile buffers, the next tile streaming
                                 in by LDS DMA during the current one's phases (one workgroup
                                 per CU); 0 = one workgroup per tile, two per CU (default)   */
    SVMI_OPT_TILE_MAP = 11    /* 1 = CX and SWAP gates on two tile qubits that a phase does not
                                 hold in registers relabel the tile's LDS indices instead of
                                 needing a phase (default; not with SVMI_OPT_SPECIALIZE);
                                 0 = off                                                     */
} svmi_option;

/* ---- lifecycle: Simulation::Simulation / ~Simulation (svsim_amdgpu_sin.hpp:207-254,
 *      svsim_cpu_omp.hpp:199-228).  `device` is the HIP ordinal, the second
 *      constructor argument of the AMD single-GPU engine (i_gpu). */
svmi_status svmi_create(int n_qubits, int device, svmi_handle* out);
svmi_status svmi_destroy(svmi_handle h);
int svmi_num_qubits(svmi_handle h);

/* ---- state: reset_sv (svsim_cpu_omp.hpp:235-240) and the public
 *      sv_real/sv_imag members (svsim_cpu_omp.hpp:576-577). Host SoA arrays of
 *      2^n doubles each. */
svmi_status svmi_reset_state(svmi_handle h);
/* Seeded random state (no reference counterpart): amplitude x is
 * (g0 + i g1) / norm with (g0, g1) = Box-Muller of the two uniforms
 * u = ((splitmix64(seed + 2x + k) >> 11) + 1) * 2^-53, k = 1, 2 (u0 -> radius,
 * u1 -> angle); norm makes the state unit length.  Generated on the device,
 * so 2^33-amplitude states need no host copy. */
svmi_status svmi_set_random_state(svmi_handle h, uint64_t seed);
svmi_status svmi_set_state(svmi_handle h, const double* re, const double* im);
svmi_status svmi_get_state(svmi_handle h, double* re, double* im);
/* Amplitudes at the given basis indices (no reference counterpart; reads a
 * few amplitudes of a large state without copying all of it). */
svmi_status svmi_get_amplitudes(svmi_handle h, const uint64_t* idx, size_t count,
                                double* re, double* im);

/* ---- circuit: append+upload / clear_circuit (svsim_cpu_omp.hpp:242-275,
 *      321-338).  upload validates (qubit bounds, distinct qubits, opcode),
 *      lowers and schedules the gate list; the engine keeps its own copy. */
svmi_status svmi_upload(svmi_handle h, const svmi_gate* gates, size_t n_gates);
svmi_status svmi_clear_circuit(svmi_handle h);

/* ---- run: Simulation::sim (svsim_cpu_omp.hpp:288-319,
 *      svsim_amdgpu_sin.hpp:320-353).  Applies the uploaded circuit to the
 *      current device state; blocks until done.  No host copy is made. */
svmi_status svmi_run(svmi_handle h);

/* ---- measure: Simulation::measure (svsim_cpu_omp.hpp:339-361).  For each
 *      draw r = uniforms[i] in [0,1] returns the basis index j with
 *      scan[j] <= r < scan[j+1] (scan = inclusive prefix sum of |amp|^2),
 *      or 0 if no interval matches.  The state is not collapsed.  If
 *      total_prob is non-null it receives scan[2^n]. */
svmi_status svmi_measure(svmi_handle h, const double* uniforms, size_t shots,
                         uint64_t* out, double* total_prob);

/* ---- expectation of a Pauli sum (no reference counterpart). */
svmi_status svmi_expval(svmi_handle h, const svmi_pauli_term* terms, size_t n_terms,
                        double* out);

/* ---- the same expectation without any device allocation (no reference
 *      counterpart): for callers that cannot allocate device memory (code
 *      under HIP graph capture, tight serving loops, harnesses that forbid
 *      hipMalloc).  Same result as svmi_expval (same kernels, same summation
 *      order).  The Pauli tables and block partials live in a 1 MiB device
 *      array that the library holds on every GPU it runs on until the process
 *      ends; at most 16384 terms (SVMI_ERR_ARG beyond).  Calls on one GPU
 *      share that array: calls that overlap in time on the same GPU (from
 *      several threads or handles) race and must be serialised by the
 *      caller. */
svmi_status svmi_expval_static(svmi_handle h, const svmi_pauli_term* terms, size_t n_terms,
                               double* out);

/* ---- adjoint differentiation (no reference counterpart).  Applies the
 *      uploaded circuit to the current state psi0 (the state afterwards is
 *      U psi0, as after svmi_run), returns E = <H> and
 *      grad[p] = dE/dtheta_p.  param_slots has 3 entries per uploaded gate
 *      (theta, phi, lambda slot) holding a parameter index in [0, n_params)
 *      or -1; a parameter may be shared by several slots.  Fails with
 *      SVMI_ERR_UNSUPPORTED if the circuit contains R or SRN (not unitary). */
svmi_status svmi_adjoint(svmi_handle h, const svmi_pauli_term* terms, size_t n_terms,
                         const int32_t* param_slots, size_t n_params,
                         double* expval_out, double* grad_out);

/* ---- batched expectation + adjoint gradient of many independent circuits,
 *      each started from |0..0>.  Step l of every circuit (forward sweeps,
 *      <H>, adjoint sweeps, reduction) is dependency level l; one device
 *      launch per level runs the tasks of all circuits at that level. */
typedef struct svmi_batch_item {
    int32_t n_qubits;
    /* planning of this circuit; 0 = the autotuned table for n_qubits
       (sv-sim_amd/csrc/svmi_tuning.hpp, written by svmi/autotune.py)      */
    uint8_t fwd_tile;               /* tile qubits of the forward sweeps, 4..13      */
    uint8_t adj_tile;               /* tile qubits of the backward sweeps, 4..11     */
    uint8_t op_cost;                /* fusion op cost + 1 (1..33)                    */
    uint8_t reserved;
    const svmi_gate* gates;
    size_t n_gates;
    const int32_t* param_slots;     /* 3 per gate, -1 = not trainable */
    size_t n_params;
    const svmi_pauli_term* terms;
    size_t n_terms;
} svmi_batch_item;

svmi_status svmi_batch_adjoint(int device, const svmi_batch_item* items, size_t n_items,
                               double* expvals, double* grads /* concatenated n_params */);

/* ---- QIR gate set: the reference's QIR-facing engine
 *      (svsim_qsharp/svsim_cpu_omp.hpp), which its Q#/QIR runtime driver
 *      (svsim_qsharp/svsim_qsharp_wrapper.cpp) feeds.  Same numbering as its
 *      enum OP (svsim_cpu_omp.hpp:45-59). */
typedef enum svmi_qir_op {
    SVMI_QIR_X, SVMI_QIR_Y, SVMI_QIR_Z, SVMI_QIR_H, SVMI_QIR_S, SVMI_QIR_T,
    SVMI_QIR_RI, SVMI_QIR_RX, SVMI_QIR_RY, SVMI_QIR_RZ, SVMI_QIR_EI, SVMI_QIR_EX, SVMI_QIR_EY, SVMI_QIR_EZ,
    SVMI_QIR_CONTROLLED_X, SVMI_QIR_CONTROLLED_Y, SVMI_QIR_CONTROLLED_Z,
    SVMI_QIR_CONTROLLED_H, SVMI_QIR_CONTROLLED_S, SVMI_QIR_CONTROLLED_T,
    SVMI_QIR_CONTROLLED_RI, SVMI_QIR_CONTROLLED_RX, SVMI_QIR_CONTROLLED_RY, SVMI_QIR_CONTROLLED_RZ,
    SVMI_QIR_CONTROLLED_EI, SVMI_QIR_CONTROLLED_EX, SVMI_QIR_CONTROLLED_EY, SVMI_QIR_CONTROLLED_EZ,
    SVMI_QIR_ADJOINT_S, SVMI_QIR_ADJOINT_T, SVMI_QIR_CONTROLLED_ADJOINT_S, SVMI_QIR_CONTROLLED_ADJOINT_T,
    SVMI_QIR_SWAP, SVMI_QIR_MEASURE,
    SVMI_QIR_NUM_OPS
} svmi_qir_op;

/* The reference's QIR Gate record (svsim_cpu_omp.hpp:113-147) without its
 * function pointer.  Controlled ops: `mask` holds the control qubits as bits
 * (the op acts where all of them are 1; a mask naming the target or a qubit
 * >= n_qubits makes it the identity).  Swap: `mask` is the second qubit.
 * Measure: `theta` is the uniform draw and `mask` the basis (0 Z, 1 X, 2 Y). */
typedef struct svmi_qir_gate {
    int32_t op;
    int32_t qubit;
    uint64_t mask;
    double theta;
} svmi_qir_gate;

/* ---- Simulation::sim of the QIR engine (svsim_cpu_omp.hpp:467-505) with
 *      mid-circuit measurement (Measure_GATE, 1406-1500).  Applies the gates
 *      in order to the first 2^n_qubits amplitudes of the device state
 *      (n_qubits = the allocated qubits, 1 <= n_qubits <= the handle's
 *      qubits; with n_qubits < 4 the amplitudes at index >= 2^n_qubits must
 *      be zero, as they are after qubits are released in |0>).  A Measure
 *      record computes P(1) of its qubit in its basis, then collapses: if
 *      draw < P(1) the bit-0 amplitudes are zeroed and the bit-1 ones scaled
 *      by 1/sqrt(P(1)), otherwise the converse with 1/sqrt(1 - P(1)).
 *      probs[m] / outcomes[m] (optional, one per Measure record) receive
 *      P(1) and (draw < P(1)); *last_prob (optional) receives the P(1) of
 *      the latest Measure run on this handle, 0 before any (the value
 *      Simulation::sim returns).  Every decision stays on the device: the
 *      call synchronises once, at its end. */
svmi_status svmi_qir_run(svmi_handle h, int n_qubits, const svmi_qir_gate* gates, size_t n_gates,
                         double* probs, int32_t* outcomes, double* last_prob);

/* ---- Simulation::sim of the density-matrix engine
 *      (svsim_qsharp_noise/svsim_cpu_omp.hpp:480-520): the same gate records
 *      on rho of n_qubits qubits, held as the first 2^(2 n_qubits)
 *      amplitudes of the device state (index row | col << n_qubits; the
 *      handle needs 2 n_qubits qubits; after svmi_reset_state it holds
 *      rho = |0..0><0..0|).  Each gate acts as U on (qubit, mask) and as
 *      conj(U) on (qubit + n_qubits, mask << n_qubits) (2785-2990); Swap
 *      exchanges the row qubits only (1560-1606); Measure changes basis on
 *      both halves, takes P(1) = sum of |Re rho[i, i]| over the rows with the
 *      bit set and divides the rows with the outcome's bit by P(1) (or
 *      1 - P(1)) and zeroes the others, rows only (1710-1830).  The engine's
 *      noise superoperators are not applied there (their calls are commented
 *      out, 2792-2840), nor here.  probs / outcomes as above.  Control masks
 *      follow that engine's OP_HEAD_MASK (619-628) over the 2 n_qubits-qubit
 *      state: a control bit c in [n_qubits, 2 n_qubits) makes the row half
 *      act where column qubit c - n_qubits is 1, and the column half (mask
 *      shifted past the state) the identity -- rho then need not stay
 *      Hermitian, exactly as in the reference; a bit >= 2 n_qubits or one
 *      naming the target makes the whole op the identity. */
svmi_status svmi_dm_qir_run(svmi_handle h, int n_qubits, const svmi_qir_gate* gates, size_t n_gates,
                            double* probs, int32_t* outcomes);

/* ---- OpenQASM 2.0 front end (replaces the translator
 *      svsim/QASMBench/svsim_qasm.py, which writes a Python script of
 *      sim.append(sim.OP(...)) calls).  Host only: parses `text` (len bytes)
 *      into the same gate records, quantum registers laid out in declaration
 *      order.  measure / reset / barrier / if statements make no gates.
 *      gates may be NULL to query the size; otherwise it must hold
 *      info->n_gates records (SVMI_ERR_ARG if capacity is smaller).  Parse
 *      errors return SVMI_ERR_ARG with "line L: ..." in svmi_last_error. */
typedef struct svmi_qasm_info {
    int32_t n_qubits;        /* sum of the qreg sizes                                */
    int32_t reserved;
    uint64_t n_gates;        /* gate records produced                                 */
    uint64_t basic_gates;    /* gate count by the translator's table (svsim_qasm.py:30-74) */
    uint64_t cx_gates;       /* CX count by the translator's table (svsim_qasm.py:78-88)   */
    uint64_t skipped;        /* measure / reset / barrier / if statements             */
} svmi_qasm_info;

svmi_status svmi_qasm_parse(const char* text, size_t len, svmi_gate* gates, size_t capacity,
                            svmi_qasm_info* info);

/* ---- options, stats, errors */
svmi_status svmi_set_option(svmi_handle h, int option, int64_t value);
svmi_status svmi_get_stats(svmi_handle h, svmi_stats* out);
const char* svmi_last_error(void);
int svmi_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SVMI_C_ABI_H_ */
