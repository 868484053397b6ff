// svmi_kernels.hpp -- host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "svmi_internal.hpp"

namespace svmi {

struct PauliGroup {
    uint64_t xmask;
    uint32_t begin, end;
};
struct PauliTerm {
    uint64_t zmask;
    double cre, cim;   // coeff * i^{|xmask & zmask|}
};

// ---- batched execution (one launch per dependency level) ----
struct BatchStep {
    uint32_t kind;        // BSTEP_*
    uint32_t circuit;
    int32_t dep;          // previous step of the same circuit (an earlier level), -1 = none
    uint32_t sweep;       // FWD/ADJ: index into the sweep array
    uint64_t task_begin;  // first queue task of this step
    uint64_t n_tasks;
    uint64_t partial;     // PAULI: first E partial; ADJ: first gradient partial
    uint32_t init_zero;   // FWD: generate |0..0> instead of loading
    uint32_t pad;
};
enum : uint32_t { BSTEP_FWD = 0, BSTEP_PAULI = 1, BSTEP_ADJ = 2 };

struct BatchCircuit {
    c64* phi;             // state (forward result, then adjoint phi)
    c64* lam;             // H psi, then adjoint lambda
    uint32_t n;           // qubits of the device state
    uint32_t group_begin, group_end;   // Pauli groups
    uint32_t e_partial, n_e_partials;  // PAULI partials
    uint32_t slot_begin, n_slots;      // gradient accumulators (global numbering)
};

struct BatchArgs {
    const BatchStep* steps;
    const BatchCircuit* circuits;
    const SweepInfo* sweeps;
    const PhaseInfo* phases;
    const DOp* ops;
    const c64* coef;
    const uint64_t* hitab;
    const PauliGroup* groups;
    const PauliTerm* terms;
    double* e_partials;
    double* grad_partials;
};

constexpr uint32_t kPauliChunk = 4096;     // amplitudes per Pauli task
static_assert(kPauliChunk % kThreads == 0, "a Pauli task is whole registers of its 256 threads");
constexpr int kPauliPerThread = 4;         // amplitudes per thread per k_pauli block


// One launch of a batch: the steps [step_begin, step_end) of one level,
// all of one kind and (sweeps) one tile size k; their tasks are
// [task_begin, task_begin + n_tasks).
struct BatchLaunch {
    uint32_t step_begin, step_end;
    uint32_t kind, k;
    uint32_t max_grad;      // ADJ: largest n_grad of the group's sweeps
    uint32_t init;          // FWD: first sweeps (tiles start as |0..0>)
    uint32_t max_xf;        // largest n_xf of the group's sweeps
    uint32_t max_dense;     // FWD: widest dense block of the group's sweeps
    uint32_t mapped;        // FWD: some of the group's sweeps have tile maps
    uint64_t task_begin, n_tasks;
};
size_t batch_launch_smem(uint32_t kind, uint32_t k, uint32_t max_grad, uint32_t max_xf);
hipError_t launch_batch(hipStream_t st, const BatchArgs& a, const BatchLaunch& L);

hipError_t init_sweep_kernels();
hipError_t init_sample_kernels();   // after init_sweep_kernels
size_t sweep_smem_limit();
size_t sweep_smem_bytes(const SweepInfo& sw, int ns);
// pipeline != 0: one-state sweeps larger than one resident grid run
// persistent and prefetch the next tile (k_sweep1p); 2 = 512-VGPR variant.
// prefetch (with pipeline 0): one-state sweeps of 12-qubit tiles over at
// least two resident grids run k_sweep1d (double-buffered LDS DMA).
hipError_t launch_sweep(hipStream_t st, int ns, c64* s0, c64* s1, const SweepInfo& sw,
                        const PhaseInfo* phases, const DOp* ops, const c64* coef,
                        const uint64_t* hitab, double* grad_partials, int pipeline, bool prefetch);
hipError_t launch_standalone(hipStream_t st, c64* s, int n, const StandaloneOp& so, const c64* coef);
hipError_t launch_fill_basis(hipStream_t st, c64* s, uint64_t dim);
hipError_t launch_gather(hipStream_t st, const c64* s, uint64_t dim, const uint64_t* idx, uint64_t cnt, c64* out);
uint32_t random_state_grid(uint64_t dim);
hipError_t launch_random_state(hipStream_t st, c64* s, uint64_t dim, uint64_t seed, double* partials);
hipError_t launch_scale(hipStream_t st, c64* s, uint64_t dim, double f);
hipError_t launch_copy(hipStream_t st, c64* d, const c64* s, uint64_t dim);
hipError_t launch_block_totals(hipStream_t st, const c64* s, uint64_t nblocks, uint32_t nthr,
                               uint32_t chunk, double* totals);
hipError_t launch_sample(hipStream_t st, const c64* s, uint32_t nthr, uint32_t chunk,
                         const uint64_t* jobs, uint32_t njobs, const double* offsets,
                         const double* draws, uint64_t* out);
uint32_t pauli_grid(uint64_t dim);
// copy (optional): also copy[y] = psi[y]
hipError_t launch_pauli(hipStream_t st, const c64* psi, c64* lam, uint64_t dim, const PauliGroup* groups,
                        uint32_t ngroups, const PauliTerm* terms, double* partials, c64* copy = nullptr);
// dst[seg.dst] = sum of src[seg.src, seg.src + seg.len), per segment, in a
// fixed order (one workgroup per segment).
struct SumSegment {
    uint64_t src, len, dst;
};
hipError_t launch_sum_segments(hipStream_t st, const double* src, const SumSegment* segs, uint64_t nseg,
                               double* dst);
// dst[r] = sum of src[r * len, (r + 1) * len), r < rows, same order.
hipError_t launch_sum_rows(hipStream_t st, const double* src, uint32_t rows, uint64_t len, double* dst);

// One mid-circuit measurement on the device.
struct MeasureResult {
    double prob;         // P(1)
    int32_t outcome;     // draw < P(1)
    int32_t pad;
};
// P(1) of qubit q over the 2^n-amplitude state s, the outcome for `draw`,
// and the collapse factors written into collapse[0] (bit 0) and collapse[1]
// (bit 1): (0, 1/sqrt(P)) for outcome 1, (1/sqrt(1-P), 0) for outcome 0.
// dm_n > 0: s is a density matrix of dm_n qubits (n = 2 dm_n, index row |
// col << dm_n), P(1) sums |Re rho[i, i]| over the rows with bit q set and the
// factors are 1/P and 1/(1-P).  partials: measure_partials(n) doubles of
// scratch.
uint32_t measure_partials(int n);
hipError_t launch_measure(hipStream_t st, const c64* s, int n, uint32_t q, double draw, double* partials,
                          c64* collapse, MeasureResult* res, int dm_n = 0);

}  // namespace svmi
