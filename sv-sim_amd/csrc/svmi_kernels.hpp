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

hipError_t init_sweep_kernels();
size_t sweep_smem_limit();
size_t sweep_smem_bytes(const SweepInfo& sw, int ns);
hipError_t launch_sweep(hipStream_t st, int ns, c64* s0, c64* s1, const SweepInfo& sw,
                        const PhaseInfo* phases, const DOp* ops, const c64* coef,
                        const uint64_t* hitab, double* grad_partials);
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
hipError_t launch_pauli(hipStream_t st, const c64* psi, c64* lam, uint64_t dim, const PauliGroup* groups,
                        uint32_t ngroups, const PauliTerm* terms, double* partials);
hipError_t launch_reduce_grads(hipStream_t st, const double* partials, uint64_t ntiles, uint32_t ngrad,
                               double* out);

}  // namespace svmi
