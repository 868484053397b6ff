// svmi_jit.hpp -- specialised sweep kernels compiled at upload time.
//
// k_sweep1 interprets a sweep's op list: per op it decodes a header and
// dispatches on the op type and register slots, and the compiler must keep
// the 16 amplitudes of a register group in the same registers at every join
// of that dispatch, which costs register copies per op.  For a fixed circuit
// the op list is known at upload, so each sweep can instead be emitted as
// straight-line HIP code -- the same tile staging (svmi_exec.hpp) and, per
// op, a direct call of its template (svmi_internal.hpp) with register slots,
// condition masks and coefficient offsets as literals -- and compiled with
// hipRTC for the device.  The coefficients stay in the program's coefficient
// array, so circuits that differ only in angles give the same source (and
// hit hipRTC's cache).  Same arithmetic in the same order: the results equal
// the interpreted sweeps'.
//
// Cost: compilation is serialised inside the process and takes from about
// a second (QFT, 33 qubits, 6 sweeps) to tens of seconds (20 layers of a
// dense u3 brickwall, 24 sweeps of 16x16 ops), so this pays for circuits
// that run many times.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "svmi_plan.hpp"

namespace svmi {

struct JitSweeps {
    hipModule_t module = nullptr;
    std::vector<hipFunction_t> fn;   // per sweep of the program; nullptr: run by k_sweep1
    std::vector<char> dma;           // per sweep: 1 = persistent double-buffered kernel
    uint64_t dma_grid = 0;           // its workgroups (one per CU)
    double compile_ms = 0.0;
    void release();
};

// HIP source of the specialised kernels (`svmi_jit_<i>` for sweep i) of a
// forward program; sweeps the generator does not cover (two states, tiles
// above 12 qubits) get no kernel.  covered[i] = 1 for the sweeps it emits,
// 2 for those emitted as persistent double-buffered kernels (dma_grid != 0:
// sweeps of 12-qubit tiles over at least two grids of dma_grid workgroups).
std::string jit_source(const Program& prog, std::vector<char>& covered, uint64_t dma_grid);

// Compile with hipRTC for `arch` and load the module on the current device.
// Returns false (and fills err) if anything fails; `out` is then empty.
bool jit_build(const Program& prog, const char* arch, JitSweeps& out, std::string& err, uint64_t dma_grid);

// Compile only (no device needed); returns the code object size or 0.
size_t jit_compile_only(const Program& prog, const char* arch, double* ms, std::string& err, uint64_t dma_grid);

// Launch sweep `i` (out.fn[i] != nullptr) over all its tiles.
hipError_t jit_launch(hipStream_t st, const JitSweeps& j, uint32_t i, const SweepInfo& sw, c64* s, const c64* coef,
                      const uint64_t* hitab);

}  // namespace svmi
