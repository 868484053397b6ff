/*
 * svmi_debug.h -- host-logic checks of libsvmi.so (no reference counterpart).
 *
 * These entry points run the engine's lowering and scheduling and then
 * execute the resulting device program on the host with the same op code the
 * HIP kernels use (sv-sim_amd/csrc/svmi_exec.hpp).  They exist so the CPU
 * test suite can check the planner without a GPU; they never touch a device
 * and are not reachable from svmi_run or the Simulation API, which always
 * execute on the GPU.
 */
#ifndef SVMI_DEBUG_H
#define SVMI_DEBUG_H

#include "svmi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Plan `gates` for n qubits with the given options and execute the program
 * on the host state re/im (2^n doubles each, updated in place). */
svmi_status svmi_debug_emulate(int n, const svmi_gate* gates, size_t n_gates,
                               int fusion, int tile_qubits, int low_qubits, int64_t sweep_budget,
                               double* re, double* im);

/* Host execution of the adjoint program: forward from (re0, im0), then
 * <H> and the gradient exactly as svmi_adjoint computes them. */
svmi_status svmi_debug_emulate_adjoint(int n, const svmi_gate* gates, size_t n_gates,
                                       const svmi_pauli_term* terms, size_t n_terms,
                                       const int32_t* param_slots, size_t n_params,
                                       int tile_qubits, int low_qubits,
                                       const double* re0, const double* im0,
                                       double* expval_out, double* grad_out);

/* Host execution of svmi_batch_adjoint's task queue, level by level, the
 * tasks of each level in reverse order (circuits of up to 20 qubits). */
svmi_status svmi_debug_emulate_batch(const svmi_batch_item* items, size_t n_items,
                                     double* expvals, double* grads);

/* Plan a batch exactly as svmi_batch_adjoint does, without executing it:
 * number of dependency levels, device launches, tasks and state bytes. */
svmi_status svmi_debug_plan_batch(const svmi_batch_item* items, size_t n_items, uint64_t* n_levels,
                                  uint64_t* n_launches, uint64_t* n_tasks, uint64_t* state_bytes);

/* The same for the backward (adjoint) program svmi_adjoint builds. */
svmi_status svmi_debug_plan_adjoint(int n, const svmi_gate* gates, size_t n_gates, const int32_t* param_slots,
                                    size_t n_params, int tile_qubits, int low_qubits, uint64_t* n_passes,
                                    char* buf, size_t buflen);

/* Host execution of svmi_qir_run's program (n <= 26 qubits, state of 2^n
 * amplitudes in and out; probs / outcomes optional, one per Measure). */
svmi_status svmi_debug_emulate_qir(int n, const svmi_qir_gate* gates, size_t n_gates, int tile_qubits,
                                   int low_qubits, double* re, double* im, double* probs, int32_t* outcomes);

/* Plan a QIR gate list as svmi_qir_run does: sweeps, measurement steps and
 * a text description (truncated to buflen-1 bytes; buf may be NULL). */
/* svmi_dm_qir_run's program on the host for a density matrix of n <= 13
 * qubits: re/im hold 2^(2n) entries in and out. */
svmi_status svmi_debug_emulate_dm_qir(int n, const svmi_qir_gate* gates, size_t n_gates, int tile_qubits,
                                      int low_qubits, double* re, double* im, double* probs, int32_t* outcomes);

svmi_status svmi_debug_plan_qir(int n, const svmi_qir_gate* gates, size_t n_gates, int tile_qubits, int low_qubits,
                                uint64_t* n_sweeps, uint64_t* n_measures, char* buf, size_t buflen);

/* Source of the specialised sweep kernels (SVMI_OPT_SPECIALIZE) for a
 * circuit (buf, truncated to buflen-1 bytes; may be NULL) and, if compile
 * != 0, its hipRTC compilation for gfx950 (no device needed): time and
 * code-object size.  compile = 2: with SVMI_OPT_PREFETCH's persistent
 * double-buffered kernels, sized for a 256-CU device. */
svmi_status svmi_debug_jit(int n, const svmi_gate* gates, size_t n_gates, int tile_qubits, int low_qubits,
                           int compile, double* compile_ms, uint64_t* code_bytes, char* buf, size_t buflen);

/* Per-step device times (ms, HIP events) of the last program the handle ran
 * with SVMI_OPT_PROFILE set, in program order (sweeps, standalone ops,
 * measurements); *n_steps receives their number, ms[0..capacity) the times. */
svmi_status svmi_debug_step_times(svmi_handle h, double* ms, size_t capacity, uint64_t* n_steps);

/* FP64 issue rates of the device (calibration): TFLOP/s of plain vector
 * FMAs (v_fma_f64) and of FP64 matrix-core products (v_mfma_f64_16x16x4f64),
 * each with eight independent accumulator chains, 256-thread workgroups and
 * waves_per_simd (1..8) workgroups per CU. */
svmi_status svmi_debug_fp64_rates(int device, int waves_per_simd, double* vfma_tflops, double* mfma_tflops);

/* Planner summary: number of full-state passes and a text description
 * (truncated to buflen-1 bytes; buf may be NULL). */
svmi_status svmi_debug_plan(int n, const svmi_gate* gates, size_t n_gates,
                            int fusion, int tile_qubits, int low_qubits, int64_t sweep_budget,
                            uint64_t* n_passes, char* buf, size_t buflen);

#ifdef __cplusplus
}
#endif

#endif /* SVMI_DEBUG_H */
