// svmi_kernels.hip -- CDNA4 (gfx950) kernels of the state-vector engine.
//
// k_sweep<NS>: one tiled sweep (the hot kernel).  One workgroup per tile of
//   2^k amplitudes, 2^(k-4) threads; the tile is staged in LDS (swizzled),
//   each phase gathers one 16-amplitude register group per thread, runs the
//   phase's ops in registers and writes the group back; after the last phase
//   the tile goes back to HBM.  Loads/stores are 16-byte per lane and the
//   low c physical qubits of every tile are contiguous, so each group of
//   2^c lanes reads one 2^c*16-byte segment.  NS = 2 runs the adjoint sweep
//   over (phi, lambda) and accumulates gradient partial sums per tile.
// Standalone ops (5-qubit dense, per-op mode), sampling, Pauli sums and
//   state fills are simple grid-stride kernels.
#include <hip/hip_runtime.h>

#include "svmi_exec.hpp"
#include "svmi_kernels.hpp"

namespace svmi {

// ------------------------------------------------------------------ sweeps

// Tile staging.  A tile is 2^k amplitudes; thread tid owns tile-local
// indices l = tid + nt*j (j < J: 16 for one state, 8 per state for two);
// global index = base | row(l >> c) | (l & cmask).  LDS holds the tile
// XOR-swizzled (lds_slot) at LDS index lo + l: lo = 0, or 2^k for lambda's
// half of a two-state tile.

// HBM (or |0..0> when init_zero) -> registers r[j] for the tile at `base`.
template <int J>
__device__ __attribute__((always_inline)) inline void
tile_fetch(c64 (&r)[J], const c64* __restrict__ s, const SweepInfo& sw, const uint64_t* __restrict__ ht,
           uint64_t base, uint32_t tid, uint32_t nt, bool init_zero)
{
    const uint32_t cmask = (1u << sw.c) - 1u;
#pragma unroll
    for (int j = 0; j < J; j++) {
        const uint32_t l = tid + nt * j;
        const uint64_t g = base | ht[l >> sw.c] | (l & cmask);
        r[j] = init_zero ? c64{g == 0 ? 1.0 : 0.0, 0.0} : s[g];
    }
}

template <int J>
__device__ __attribute__((always_inline)) inline void
tile_put(const c64 (&r)[J], c64* t, uint32_t lo, uint32_t tid, uint32_t nt)
{
#pragma unroll
    for (int j = 0; j < J; j++) t[lds_slot(lo + tid + nt * j)] = r[j];
}

// LDS -> HBM for the tile at `base` when the sweep ends with a tile map:
// tile-local index l = tid + nt j sits at LDS index M(tid) ^ M(nt j), the
// second part precomputed per sweep (SweepInfo::fj).
template <int J>
__device__ __attribute__((always_inline)) inline void
tile_store_mapped(const c64* t, uint32_t lo, c64* __restrict__ s, const SweepInfo& sw,
                  const uint64_t* __restrict__ ht, uint64_t base, uint32_t tid, uint32_t nt)
{
    const uint32_t cmask = (1u << sw.c) - 1u;
    const uint32_t mt = sweep_map_thread(sw, tid);
#pragma unroll
    for (int j = 0; j < J; j++) {
        const uint32_t l = tid + nt * j;
        s[base | ht[l >> sw.c] | (l & cmask)] = t[lds_slot(lo + (mt ^ sw.fj[j]))];
    }
}

// LDS -> HBM for the tile at `base`.
template <int J>
__device__ __attribute__((always_inline)) inline void
tile_store(const c64* t, uint32_t lo, c64* __restrict__ s, const SweepInfo& sw, const uint64_t* __restrict__ ht,
           uint64_t base, uint32_t tid, uint32_t nt)
{
    const uint32_t cmask = (1u << sw.c) - 1u;
#pragma unroll
    for (int j = 0; j < J; j++) {
        const uint32_t l = tid + nt * j;
        s[base | ht[l >> sw.c] | (l & cmask)] = t[lds_slot(lo + l)];
    }
}

// v of a lane moved by one DPP pattern (both halves of the double from the
// same source lane); lanes without a source, or outside row_mask, get 0.
template <int kCtrl, int kRowMask>
__device__ __attribute__((always_inline)) inline double dpp_move(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, kRowMask, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, kRowMask, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum of v over the 64 lanes of a wave, left in lane 63: an inclusive scan
// inside each 16-lane row (row_shr 1, 2, 4, 8), then row 0's total into row 1
// and row 2's into row 3 (row_bcast:15), then rows 0+1 into rows 2-3
// (row_bcast:31).  A fixed order, so the sum is the same on every run.
__device__ __attribute__((always_inline)) inline double wave_sum_lane63(double v)
{
    v += dpp_move<0x111, 0xf>(v);
    v += dpp_move<0x112, 0xf>(v);
    v += dpp_move<0x114, 0xf>(v);
    v += dpp_move<0x118, 0xf>(v);
    v += dpp_move<0x142, 0xa>(v);
    v += dpp_move<0x143, 0xc>(v);
    return v;
}

// The phases of one sweep on the tile in LDS (t0; for NS = 2 phi then lambda).
// Threads tid < nt own one register group per phase; with kIdle (batched
// kernel: blockDim >= nt) the others only join the barriers.  Without kIdle
// every thread is active, which keeps the phase body free of divergence
// (and its register groups out of scratch).  Ends with a barrier.
// kMap: the phases address LDS through their tile maps (every two-state
// sweep; one-state sweeps whose SweepInfo::mapped is set).
template <int NS, bool kIdle, int kMaxQ, bool kMap = (NS == 2)>
__device__ __attribute__((always_inline)) inline void
tile_phases(c64* t0, double* gacc, const c64* xf, const SweepInfo& sw,
            const PhaseInfo* __restrict__ phases, const DOp* __restrict__ ops, const c64* __restrict__ coef,
            uint64_t base, uint32_t tid, uint32_t nt)
{
    const bool active = !kIdle || tid < nt;
    const uint32_t nwaves = (nt + 63) / 64;
    const uint32_t vhi = uint32_t(extract_bits(base, sw.exthi));
    for (uint32_t p = sw.phase_begin; p < sw.phase_end; p++) {
        if (active) {
            const PhaseInfo ph = phases[p];
            uint32_t off[kRegs], goffs[kRegs];
            phase_offsets(ph, off);
            phase_gather_offsets(ph, goffs);
            const uint32_t goff = group_offset(ph, tid);
            // NS = 2: the phase's last register bit is k, so a[0..7] are
            // phi's amplitudes and a[8..15] lambda's (t1 = t0 + 2^k) and every
            // op, dispatched once, acts on both.  The group's amplitudes sit
            // at LDS indices under the tile map (gphys, gather and scatter
            // offsets); the ops see the plain tile-local indices (goff, off).
            uint32_t gphys = goff;
            if constexpr (kMap) gphys = mapped_group_offset(ph, tid);
            c64 a[kRegs];
#pragma unroll
            for (int r = 0; r < kRegs; r++) a[r] = t0[lds_slot(kMap ? (gphys ^ goffs[r]) : (goff | goffs[r]))];
            // op headers are fetched one op ahead so the scalar-load latency
            // of the next op overlaps the arithmetic of the current one (every
            // op array ends with a NOP, so the read past a phase's last op
            // stays inside it)
            const DOp* __restrict__ op = ops + ph.op_begin;
            DOpHdr nxt = op->h;
            for (uint32_t o = ph.op_begin; o < ph.op_end; o++, op++) {
                const DOpHdr h = nxt;
                nxt = op[1].h;
                if constexpr (NS == 2) {
                    if (h.type == DOP_GRAD_DENSE || h.type == DOP_GRAD_DIAG) {
                        // every gradient op of a sweep has its own slot and each
                        // wave runs it once: a plain store, no read
                        double v = exec_grad_op<kMaxDenseQubitsAdj, kRegsAdj>(h, op, a, coef, goff, off, base);
                        if (nt >= 64) {
                            v = wave_sum_lane63(v);
                            if ((tid & 63u) == 63u) gacc[h.grad * nwaves + (tid >> 6)] = v;
                        } else {
                            const int width = int(nt);
                            for (int d = width >> 1; d > 0; d >>= 1) v += __shfl_xor(v, d, width);
                            if (tid == 0) gacc[h.grad * nwaves] = v;
                        }
                        continue;
                    }
                }
                exec_state_op<kMaxQ>(h, op, a, coef, goff, tid, off, base, xf, sw.exthi, vhi);
            }
            if constexpr (kMap) {
                uint32_t soffs[kRegs];
                phase_scatter_offsets(ph, soffs);
#pragma unroll
                for (int r = 0; r < kRegs; r++) t0[lds_slot(gphys ^ soffs[r])] = a[r];
            } else {
#pragma unroll
                for (int r = 0; r < kRegs; r++) t0[lds_slot((goff | off[r]) ^ ph.sxor)] = a[r];
            }
        }
        __syncthreads();
    }
}

// The tile's phase factors (one per factor slot of the sweep) into LDS; the
// caller's next barrier publishes them.
__device__ __attribute__((always_inline)) inline void
tile_factors(c64* xf, const SweepInfo& sw, const DOp* __restrict__ ops, const c64* __restrict__ coef,
             uint64_t base, uint32_t tid)
{
    if (sw.n_xf == 0) return;
    // the sweep's lane tables, then the tile's factors
    const uint32_t nlt = 16u * sw.n_lt, lt0 = sw.n_xf - nlt;
    for (uint32_t i = tid; i < nlt; i += blockDim.x) xf[lt0 + i] = coef[sw.lt_coef + i];
    const uint64_t x = ext_index(sw, base);
    for (uint32_t o = sw.op_begin + tid; o < sw.op_end; o += blockDim.x) {
        const DOpHdr h = ops[o].h;
        if (has_tile_factor(h)) xf[h.grad] = tile_factor(h, coef, base, x);
    }
}

// LDS of a tile: t0 (NS = 2: phi's 2^k then lambda's) | phase factors |
// gradient accumulators.
// One tile of one sweep: load, phases, store (and for NS = 2 the tile's
// gradient partials).  init_zero: the tile is generated as |0..0>.
template <int NS, bool kIdle, int kMaxQ = (NS == 2 ? kMaxDenseQubitsAdj : kRegBits), bool kMap = (NS == 2)>
__device__ __attribute__((always_inline)) inline void
sweep_tile(c64* __restrict__ s0, c64* __restrict__ s1, const SweepInfo& sw,
           const PhaseInfo* __restrict__ phases, const DOp* __restrict__ ops,
           const c64* __restrict__ coef, const uint64_t* __restrict__ hitab,
           double* __restrict__ grad_partials, uint64_t tile, uint32_t nt, unsigned char* smem,
           bool init_zero)
{
    const uint32_t tsize = 1u << sw.k;
    c64* t0 = reinterpret_cast<c64*>(smem);
    c64* xf = t0 + tsize * NS;
    double* gacc = reinterpret_cast<double*>(xf + sw.n_xf);
    const uint32_t tid = threadIdx.x;
    const bool active = !kIdle || tid < nt;
    const uint32_t nwaves = (nt + 63) / 64;
    const uint64_t* ht = hitab + sw.hitab;
    const uint64_t base = tile_base(sw, tile);

    constexpr int J = NS == 2 ? kRegsAdj : kRegs;   // amplitudes per thread per state
    if (active) {
        // both states' loads are issued before either waits
        c64 r[J], r1[J];
        tile_fetch(r, s0, sw, ht, base, tid, nt, init_zero);
        if constexpr (NS == 2) tile_fetch(r1, s1, sw, ht, base, tid, nt, false);
        tile_put(r, t0, 0u, tid, nt);
        if constexpr (NS == 2) tile_put(r1, t0, tsize, tid, nt);
    }
    if (NS == 2)
        for (uint32_t i = tid; i < sw.n_grad * nwaves; i += blockDim.x) gacc[i] = 0.0;
    tile_factors(xf, sw, ops, coef, base, tid);
    __syncthreads();
    tile_phases<NS, kIdle, kMaxQ, kMap>(t0, gacc, xf, sw, phases, ops, coef, base, tid, nt);
    if (active) {
        if (kMap && sw.mapped) {
            tile_store_mapped<J>(t0, 0u, s0, sw, ht, base, tid, nt);
            if constexpr (NS == 2) tile_store_mapped<J>(t0, tsize, s1, sw, ht, base, tid, nt);
        } else {
            tile_store<J>(t0, 0u, s0, sw, ht, base, tid, nt);
            if constexpr (NS == 2) tile_store<J>(t0, tsize, s1, sw, ht, base, tid, nt);
        }
    }
    // gradient partials are stored grad-major ([i][tile]) so the sum over
    // tiles of one gradient reads one contiguous segment
    if (NS == 2) {
        for (uint32_t i = tid; i < sw.n_grad; i += blockDim.x) {
            double v = 0.0;
            for (uint32_t w = 0; w < nwaves; w++) v += gacc[i * nwaves + w];
            grad_partials[i * sw.n_tiles + tile] = v;
        }
    }
}

// One state, one tile per workgroup: up to 8 waves per CU (one 2^13 tile of
// 512 threads, or two 2^12 tiles of 256), so 256 VGPRs per thread.
// stagger_lo..stagger_hi: workgroups of the first resident wave that start
// later (by `stagger` sleeps of 127*64 cycles), so that the co-resident
// workgroups of a CU alternate loading and computing.
// kMaxQ: widest dense block compiled in (sweeps with blocks of at most two
// register qubits run the kMaxQ = 2 build, which holds fewer registers).
// kMap: the build for sweeps with tile maps (SweepInfo::mapped).
template <int kMaxQ, bool kMap>
__global__ void __launch_bounds__(kMaxTileThreads, 1)
k_sweep1(c64* __restrict__ s0, const SweepInfo sw, const PhaseInfo* __restrict__ phases,
         const DOp* __restrict__ ops, const c64* __restrict__ coef, const uint64_t* __restrict__ hitab,
         uint64_t tile_offset, uint32_t stagger_lo, uint32_t stagger_hi, uint32_t stagger)
{
    extern __shared__ __align__(16) unsigned char smem[];
    if (blockIdx.x >= stagger_lo && blockIdx.x < stagger_hi)
        for (uint32_t i = 0; i < stagger; i++) __builtin_amdgcn_s_sleep(127);
    sweep_tile<1, false, kMaxQ, kMap>(s0, nullptr, sw, phases, ops, coef, hitab, nullptr, tile_offset + blockIdx.x, blockDim.x,
                         smem, false);
}

// One state, persistent and software-pipelined: workgroup w handles tiles
// w, w + gridDim.x, ...  While the phases of tile t run out of LDS, the
// loads of the next tile are already in flight into registers, so HBM
// traffic overlaps the arithmetic inside each workgroup (co-resident
// workgroups start in step and would otherwise load and compute together).
// kWide: 256-thread workgroups with the full 512-VGPR budget (one per CU);
// otherwise up to 512 threads and 256 VGPRs.
template <bool kWide>
__global__ void __launch_bounds__(kWide ? 256 : kMaxTileThreads, 1)
k_sweep1p(c64* __restrict__ s0, const SweepInfo sw, const PhaseInfo* __restrict__ phases,
          const DOp* __restrict__ ops, const c64* __restrict__ coef, const uint64_t* __restrict__ hitab,
          uint64_t tile_offset, uint64_t n_tiles)
{
    extern __shared__ __align__(16) unsigned char smem[];
    c64* t0 = reinterpret_cast<c64*>(smem);
    c64* xf = t0 + (size_t(1) << sw.k);
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint64_t* ht = hitab + sw.hitab;
    uint64_t t = blockIdx.x;
    c64 nx[kRegs];
    tile_fetch(nx, s0, sw, ht, tile_base(sw, tile_offset + t), tid, nt, false);
    for (;;) {
        const uint64_t base = tile_base(sw, tile_offset + t);
        tile_put(nx, t0, 0u, tid, nt);
        tile_factors(xf, sw, ops, coef, base, tid);
        __syncthreads();
        const uint64_t tn = t + gridDim.x;
        const bool more = tn < n_tiles;
        if (more) tile_fetch(nx, s0, sw, ht, tile_base(sw, tile_offset + tn), tid, nt, false);
        tile_phases<1, false, kRegBits>(t0, nullptr, xf, sw, phases, ops, coef, base, tid, nt);
        tile_store<kRegs>(t0, 0u, s0, sw, ht, base, tid, nt);
        if (!more) break;
        __syncthreads();  // the store's LDS reads finish before the next put
        t = tn;
    }
}

// ---------------------------------------------- double-buffered sweeps
//
// LDS DMA helpers (lds_dma16, tile_dma): svmi_exec.hpp.

size_t dma_sweep_smem_bytes(const SweepInfo& sw)
{
    return (size_t(32) << kDmaTileQubits) + (size_t(8) << (kDmaTileQubits - sw.c)) + size_t(16) * sw.n_xf;
}

// One state, persistent, two tile buffers of 2^12 amplitudes: while the
// phases of tile t run in one buffer, tile t + gridDim.x streams into the
// other, so HBM traffic overlaps the arithmetic inside every workgroup (one
// per CU).  LDS: two tiles | the sweep's row table | phase factors.  Per
// tile: wait for its DMA (all but the previous tile's kRegs stores, the
// youngest memory operations), barrier, the tile's factors, the next tile's
// DMA, barrier, phases, stores.
template <int kMaxQ>
__global__ void __launch_bounds__(kDmaThreads, 1)
k_sweep1d(c64* __restrict__ s0, const SweepInfo sw, const PhaseInfo* __restrict__ phases,
          const DOp* __restrict__ ops, const c64* __restrict__ coef, const uint64_t* __restrict__ hitab)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr uint32_t kT = 1u << kDmaTileQubits;
    c64* buf = reinterpret_cast<c64*>(smem);
    uint64_t* ht = reinterpret_cast<uint64_t*>(buf + 2 * kT);
    c64* xf = reinterpret_cast<c64*>(ht + (kT >> sw.c));
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    for (uint32_t i = tid; i < (kT >> sw.c); i += kDmaThreads) ht[i] = hitab[sw.hitab + i];
    __syncthreads();
    const uint32_t lds0 = lds_address(buf);
    uint64_t t = blockIdx.x;
    tile_dma(s0, ht, sw.c, tile_base(sw, t), lds0, wave, lane);
    uint32_t cur = 0;
    bool first = true;
    for (;;) {
        // tile_store<kRegs> below issues kRegs = 16 stores per thread, so
        // vmcnt(16) leaves only those outstanding and this tile's DMA done
        if (first) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        else asm volatile("s_waitcnt vmcnt(16)" ::: "memory");
        __syncthreads();
        const uint64_t base = tile_base(sw, t);
        const uint64_t tn = t + gridDim.x;
        const bool more = tn < sw.n_tiles;
        // factors first: their global loads would otherwise wait for the
        // DMA; the explicit vmcnt(0) tells the compiler's wait analysis that
        // no load of theirs is still in flight (a pending one whose register
        // a later LDS read reuses would make it wait for everything)
        tile_factors(xf, sw, ops, coef, base, tid);
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0), expcnt/lgkmcnt unconstrained
        if (more) tile_dma(s0, ht, sw.c, tile_base(sw, tn), lds0 + (cur ^ 1u) * kT * 16u, wave, lane);
        if (sw.n_xf) __syncthreads();
        c64* t0 = buf + cur * kT;
        tile_phases<1, false, kMaxQ>(t0, nullptr, xf, sw, phases, ops, coef, base, tid, kDmaThreads);
        tile_store<kRegs>(t0, 0u, s0, sw, ht, base, tid, kDmaThreads);
        if (!more) break;
        t = tn;
        cur ^= 1u;
        first = false;
    }
}

// Two states (adjoint): two register groups per thread (one wave per SIMD),
// dense ops of at most 2 qubits.
__global__ void __launch_bounds__(kThreads, 1)
k_sweep2(c64* __restrict__ s0, c64* __restrict__ s1, const SweepInfo sw, const PhaseInfo* __restrict__ phases,
         const DOp* __restrict__ ops, const c64* __restrict__ coef, const uint64_t* __restrict__ hitab,
         double* __restrict__ grad_partials, uint64_t tile_offset)
{
    extern __shared__ __align__(16) unsigned char smem[];
    sweep_tile<2, false>(s0, s1, sw, phases, ops, coef, hitab, grad_partials, tile_offset + blockIdx.x, blockDim.x,
                         smem, false);
}

static size_t g_max_smem = 64 * 1024;  // dynamic LDS the sweep kernels may use

// --------------------------------------------------------------- batching
//
// A batch is a list of steps (forward sweep, Pauli step, adjoint sweep,
// reduction) of many circuits; a task is one tile (or chunk) of one step.
// Level l holds step l of every circuit that has one, so the tasks of a
// level are independent; inside a level the steps are grouped by kind and
// tile size, and each group is one launch (blockIdx.x -> task, blockDim =
// the group's tile threads).  Stream order between the launches supplies
// every dependency, so no workgroup ever waits for another one.

// The step of task t among steps [s0, s1): the last one whose task_begin is
// not above t.
__device__ inline uint32_t batch_step_of(const BatchStep* __restrict__ steps, uint32_t s0, uint32_t s1, uint64_t t)
{
    uint32_t lo = s0, hi = s1;
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (steps[mid].task_begin <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Read-only program tables of a batch, as restrict kernel arguments so the
// tile code reads op headers through the scalar path as k_sweep does.
#define SVMI_BATCH_TABLES                                                                            \
    const BatchStep* __restrict__ steps, const BatchCircuit* __restrict__ circuits,                  \
        const SweepInfo* __restrict__ sweeps, const PhaseInfo* __restrict__ phases,                 \
        const DOp* __restrict__ ops, const c64* __restrict__ coef, const uint64_t* __restrict__ hitab

// kInit: the steps are first forward sweeps (tiles generated as |0..0>).
// kMap: some of the launch's sweeps have tile maps.
template <bool kInit, int kMaxQ, bool kMap>
__global__ void __launch_bounds__(kMaxTileThreads, 1)
k_batch_fwd(SVMI_BATCH_TABLES, uint32_t s0, uint32_t s1, uint64_t task0)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t t = task0 + blockIdx.x;
    const BatchStep st = steps[batch_step_of(steps, s0, s1, t)];
    const SweepInfo sw = sweeps[st.sweep];
    sweep_tile<1, false, kMaxQ, kMap>(circuits[st.circuit].phi, nullptr, sw, phases, ops, coef, hitab, nullptr,
                         t - st.task_begin, blockDim.x, smem, kInit);
}

__global__ void __launch_bounds__(kThreads, 1)
k_batch_adj(SVMI_BATCH_TABLES, double* __restrict__ grad_partials, uint32_t s0, uint32_t s1, uint64_t task0)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t t = task0 + blockIdx.x;
    const BatchStep st = steps[batch_step_of(steps, s0, s1, t)];
    const SweepInfo sw = sweeps[st.sweep];
    const BatchCircuit& c = circuits[st.circuit];
    sweep_tile<2, false>(c.phi, c.lam, sw, phases, ops, coef, hitab, grad_partials + st.partial, t - st.task_begin,
                         blockDim.x, smem, false);
}

// <H> step: lam = H phi on a chunk of kPauliChunk amplitudes and the chunk's
// partial sum of Re <phi|lam>.
__global__ void __launch_bounds__(kThreads)
k_batch_pauli(BatchArgs a, uint32_t s0, uint32_t s1, uint64_t task0)
{
    __shared__ double red[kThreads];
    const uint64_t t = task0 + blockIdx.x;
    const BatchStep st = a.steps[batch_step_of(a.steps, s0, s1, t)];
    const BatchCircuit& c = a.circuits[st.circuit];
    const uint64_t task = t - st.task_begin;
    const uint64_t dim = 1ull << c.n;
    const uint64_t y0 = task * kPauliChunk;
    const uint64_t y1 = y0 + kPauliChunk < dim ? y0 + kPauliChunk : dim;
    // the thread's kPauliPerThread amplitudes y0 + tid + 256 j in registers,
    // groups and terms outermost: each term is read once per thread, not
    // once per amplitude (same sums in the same order as one amplitude at a
    // time)
    constexpr int M = kPauliChunk / kThreads;
    c64 l[M];
#pragma unroll
    for (int j = 0; j < M; j++) l[j] = c64{0.0, 0.0};
    for (uint32_t g = c.group_begin; g < c.group_end; g++) {
        const PauliGroup G = a.groups[g];
        double fr[M], fi[M];
#pragma unroll
        for (int j = 0; j < M; j++) fr[j] = fi[j] = 0.0;
        for (uint32_t e = G.begin; e < G.end; e++) {
            const PauliTerm T = a.terms[e];
#pragma unroll
            for (int j = 0; j < M; j++) {
                const uint64_t x = (y0 + threadIdx.x + uint64_t(kThreads) * j) ^ G.xmask;
                const bool neg = __popcll(x & T.zmask) & 1;
                fr[j] += neg ? -T.cre : T.cre;
                fi[j] += neg ? -T.cim : T.cim;
            }
        }
#pragma unroll
        for (int j = 0; j < M; j++) {
            const uint64_t y = y0 + threadIdx.x + uint64_t(kThreads) * j;
            if (y < y1) l[j] = cmac(l[j], c64{fr[j], fi[j]}, c.phi[y ^ G.xmask]);
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < M; j++) {
        const uint64_t y = y0 + threadIdx.x + uint64_t(kThreads) * j;
        if (y < y1) {
            c.lam[y] = l[j];
            const c64 p = c.phi[y];
            acc += p.re * l[j].re + p.im * l[j].im;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.e_partials[st.partial + task] = red[0];
}

size_t batch_launch_smem(uint32_t kind, uint32_t k, uint32_t max_grad, uint32_t max_xf)
{
    if (kind == BSTEP_FWD) return (size_t(16) << k) + size_t(16) * max_xf;
    if (kind == BSTEP_ADJ) {
        const size_t nwaves = (size_t(sweep_threads(k, 2)) + 63) / 64;
        return (size_t(32) << k) + size_t(16) * max_xf + size_t(8) * max_grad * nwaves;
    }
    return 0;
}

// The k_batch_fwd build of a forward launch.
using BatchFwdFn = void (*)(const BatchStep*, const BatchCircuit*, const SweepInfo*, const PhaseInfo*, const DOp*,
                            const c64*, const uint64_t*, uint32_t, uint32_t, uint64_t);
template <bool kInit, bool kMap>
BatchFwdFn batch_fwd_kernel_of(const BatchLaunch& L)
{
    return L.max_dense <= 2 ? &k_batch_fwd<kInit, 2, kMap> : &k_batch_fwd<kInit, kRegBits, kMap>;
}
BatchFwdFn batch_fwd_kernel(const BatchLaunch& L)
{
    if (L.init) return L.mapped ? batch_fwd_kernel_of<true, true>(L) : batch_fwd_kernel_of<true, false>(L);
    return L.mapped ? batch_fwd_kernel_of<false, true>(L) : batch_fwd_kernel_of<false, false>(L);
}

hipError_t launch_batch(hipStream_t st, const BatchArgs& a, const BatchLaunch& L)
{
    const size_t smem = batch_launch_smem(L.kind, L.k, L.max_grad, L.max_xf);
    if (smem > g_max_smem) return hipErrorInvalidConfiguration;
    const uint32_t threads = L.kind == BSTEP_FWD ? sweep_threads(L.k, 1)
                             : L.kind == BSTEP_ADJ ? sweep_threads(L.k, 2)
                                                   : kThreads;
    const uint64_t max_grid = 1ull << 30;
    for (uint64_t t0 = 0; t0 < L.n_tasks; t0 += max_grid) {
        const uint32_t cnt = uint32_t(L.n_tasks - t0 < max_grid ? L.n_tasks - t0 : max_grid);
        const uint64_t first = L.task_begin + t0;
        switch (L.kind) {
        case BSTEP_FWD: {
            auto* fn = batch_fwd_kernel(L);
            hipLaunchKernelGGL(fn, dim3(cnt), dim3(threads), smem, st, a.steps, a.circuits, a.sweeps, a.phases,
                               a.ops, a.coef, a.hitab, L.step_begin, L.step_end, first);
            break;
        }
        case BSTEP_ADJ:
            hipLaunchKernelGGL(k_batch_adj, dim3(cnt), dim3(threads), smem, st, a.steps, a.circuits, a.sweeps,
                               a.phases, a.ops, a.coef, a.hitab, a.grad_partials, L.step_begin, L.step_end, first);
            break;
        case BSTEP_PAULI:
            hipLaunchKernelGGL(k_batch_pauli, dim3(cnt), dim3(threads), 0, st, a, L.step_begin, L.step_end, first);
            break;
        default:
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

size_t sweep_smem_bytes(const SweepInfo& sw, int ns)
{
    const size_t nt = sweep_threads(sw.k, ns);
    const size_t nwaves = (nt + 63) / 64;
    return (size_t(16) << sw.k) * ns + size_t(16) * sw.n_xf + size_t(8) * sw.n_grad * nwaves;
}

// Workgroups of k_sweep1p resident at once for tile size k (occupancy x CUs).
template <bool kWide>
static uint64_t pipeline_grid(uint32_t k)
{
    static uint64_t cache[kMaxTileQubits + 1] = {};
    if (cache[k]) return cache[k];
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(&k_sweep1p<kWide>),
                                                     1 << (k - kRegBits), size_t(16) << k) != hipSuccess ||
        per < 1)
        per = 1;
    (void)hipGetLastError();
    cache[k] = uint64_t(cus) * uint64_t(per);
    return cache[k];
}

// Workgroups of k_sweep1d resident at once (CUs x occupancy at the largest
// LDS it may use; 0 if it cannot run).
static uint64_t dma_grid()
{
    static uint64_t cache = ~0ull;
    if (cache != ~0ull) return cache;
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1 ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(&k_sweep1d<kRegBits>),
                                                     int(kDmaThreads), g_max_smem) != hipSuccess)
        per = 0;
    (void)hipGetLastError();
    cache = per > 0 ? uint64_t(cus) * uint64_t(per) : 0;
    return cache;
}

hipError_t launch_sweep(hipStream_t st, int ns, c64* s0, c64* s1, const SweepInfo& sw,
                        const PhaseInfo* phases, const DOp* ops, const c64* coef,
                        const uint64_t* hitab, double* grad_partials, int pipeline, bool prefetch)
{
    const uint32_t nt = sweep_threads(sw.k, ns);
    const size_t smem = sweep_smem_bytes(sw, ns);
    if (smem > g_max_smem) return hipErrorInvalidConfiguration;
    // sweeps with tile maps run on k_sweep1's mapped build only
    if (ns == 1 && sw.mapped) pipeline = 0, prefetch = false;
    if (ns == 1 && prefetch && pipeline == 0 && sw.k == kDmaTileQubits && nt == kDmaThreads &&
        dma_sweep_smem_bytes(sw) <= g_max_smem) {
        const uint64_t grid = dma_grid();
        if (grid && sw.n_tiles >= 2 * grid) {
            hipLaunchKernelGGL(sw.max_dense <= 2 ? k_sweep1d<2> : k_sweep1d<kRegBits>, dim3(uint32_t(grid)),
                               dim3(kDmaThreads), dma_sweep_smem_bytes(sw), st, s0, sw, phases, ops, coef, hitab);
            return hipGetLastError();
        }
    }
    if (ns == 1 && (pipeline == 1 || pipeline == 2)) {
        const bool wide = pipeline == 2 && nt <= 256;
        const uint64_t resident = wide ? pipeline_grid<true>(sw.k) : pipeline_grid<false>(sw.k);
        if (sw.n_tiles > resident) {
            if (wide)
                hipLaunchKernelGGL(k_sweep1p<true>, dim3(uint32_t(resident)), dim3(nt), smem, st, s0, sw, phases, ops,
                                   coef, hitab, uint64_t(0), sw.n_tiles);
            else
                hipLaunchKernelGGL(k_sweep1p<false>, dim3(uint32_t(resident)), dim3(nt), smem, st, s0, sw, phases,
                                   ops, coef, hitab, uint64_t(0), sw.n_tiles);
            return hipGetLastError();
        }
    }
    const uint64_t max_grid = 1ull << 30;
    for (uint64_t t0 = 0; t0 < sw.n_tiles; t0 += max_grid) {
        const uint64_t cnt = sw.n_tiles - t0 < max_grid ? sw.n_tiles - t0 : max_grid;
        if (ns == 1) {
            // pipeline 3: stagger the second workgroup slot of every CU
            uint32_t lo = 0, hi = 0, sl = 0;
            if (pipeline >= 3) {
                const uint64_t res = pipeline_grid<false>(sw.k);
                lo = uint32_t(res / 2);
                hi = uint32_t(res);
                sl = uint32_t(pipeline - 2);
            }
            auto* fn = sw.mapped ? (sw.max_dense <= 2 ? &k_sweep1<2, true> : &k_sweep1<kRegBits, true>)
                                 : (sw.max_dense <= 2 ? &k_sweep1<2, false> : &k_sweep1<kRegBits, false>);
            hipLaunchKernelGGL(fn, dim3(uint32_t(cnt)), dim3(nt), smem, st, s0, sw, phases, ops, coef, hitab, t0, lo,
                               hi, sl);
        }
        else
            hipLaunchKernelGGL(k_sweep2, dim3(uint32_t(cnt)), dim3(nt), smem, st, s0, s1, sw, phases, ops,
                               coef, hitab, grad_partials, t0);
    }
    return hipGetLastError();
}

size_t sweep_smem_limit() { return g_max_smem; }

// Raise the dynamic-LDS limit of the sweep kernels to what the device allows
// per workgroup (gfx950: 160 KiB per CU); 64 KiB always works.
hipError_t init_sweep_kernels()
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || optin <= 0)
        optin = 64 * 1024;
    size_t want = size_t(optin);
    if (want > 160 * 1024) want = 160 * 1024;
    if (want > 64 * 1024) {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep1<kRegBits, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(want));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep1<2, false>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(want));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep1<kRegBits, true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(want));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep1<2, true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(want));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep1p<false>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(want));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep1p<true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(want));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep1d<kRegBits>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(want));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep1d<2>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(want));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep2),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(want));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            want = 64 * 1024;
        }
    }
    auto raise = [&](const void* fn) {
        return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(want)) == hipSuccess;
    };
    if (want > 64 * 1024 &&
        !(raise(reinterpret_cast<const void*>(&k_batch_fwd<true, kRegBits, false>)) &&
          raise(reinterpret_cast<const void*>(&k_batch_fwd<false, kRegBits, false>)) &&
          raise(reinterpret_cast<const void*>(&k_batch_fwd<true, 2, false>)) &&
          raise(reinterpret_cast<const void*>(&k_batch_fwd<false, 2, false>)) &&
          raise(reinterpret_cast<const void*>(&k_batch_fwd<true, kRegBits, true>)) &&
          raise(reinterpret_cast<const void*>(&k_batch_fwd<false, kRegBits, true>)) &&
          raise(reinterpret_cast<const void*>(&k_batch_fwd<true, 2, true>)) &&
          raise(reinterpret_cast<const void*>(&k_batch_fwd<false, 2, true>)) &&
          raise(reinterpret_cast<const void*>(&k_batch_adj)))) {
        (void)hipGetLastError();
        want = 64 * 1024;
    }
    g_max_smem = want;
    return hipSuccess;
}

// -------------------------------------------------------- standalone ops

struct QArgs {
    int q;
    int qubits[5];      // op order
    int sorted[5];      // ascending
};

template <int K>
__global__ void __launch_bounds__(256)
k_dense_full(c64* __restrict__ s, uint64_t ngroups, QArgs qa, const c64* __restrict__ M)
{
    constexpr int D = 1 << K;
    for (uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < ngroups;
         g += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t b = g;
#pragma unroll
        for (int i = 0; i < K; i++) {
            const int p = qa.sorted[i];
            b = ((b >> p) << (p + 1)) | (b & ((1ull << p) - 1ull));
        }
        uint64_t idx[D];
        c64 v[D];
#pragma unroll
        for (int j = 0; j < D; j++) {
            uint64_t o = b;
#pragma unroll
            for (int i = 0; i < K; i++)
                if ((j >> i) & 1) o |= 1ull << qa.qubits[i];
            idx[j] = o;
            v[j] = s[o];
        }
#pragma unroll
        for (int r = 0; r < D; r++) {
            c64 acc{0.0, 0.0};
#pragma unroll
            for (int j = 0; j < D; j++) acc = cmac(acc, M[r * D + j], v[j]);
            s[idx[r]] = acc;
        }
    }
}

__global__ void __launch_bounds__(256)
k_diag_full(c64* __restrict__ s, uint64_t dim, QArgs qa, const c64* __restrict__ tab)
{
    for (uint64_t x = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; x < dim;
         x += uint64_t(gridDim.x) * blockDim.x) {
        uint32_t idx = 0;
        for (int i = 0; i < qa.q; i++) idx |= uint32_t((x >> qa.qubits[i]) & 1u) << i;
        s[x] = cmul(s[x], tab[idx]);
    }
}

__global__ void __launch_bounds__(256) k_srn_full(c64* __restrict__ s, uint64_t half, int q)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < half;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t p0 = ((i >> q) << (q + 1)) | (i & ((1ull << q) - 1ull));
        const uint64_t p1 = p0 | (1ull << q);
        const c64 a0 = s[p0], a1 = s[p1];
        s[p0] = c64{0.5 * (a0.re + a1.re), 0.5 * (a0.im - a1.im)};
        s[p1] = c64{0.5 * (a0.re + a1.re), 0.5 * (-a0.im + a1.im)};
    }
}

static unsigned grid_for(uint64_t work, unsigned block)
{
    const uint64_t g = (work + block - 1) / block;
    const uint64_t cap = 65536;
    return unsigned(g < 1 ? 1 : (g > cap ? cap : g));
}

hipError_t launch_standalone(hipStream_t st, c64* s, int n, const StandaloneOp& so, const c64* coef)
{
    QArgs qa;
    qa.q = so.q;
    for (int i = 0; i < 5; i++) qa.qubits[i] = i < so.q ? so.qubits[i] : 0;
    for (int i = 0; i < 5; i++) qa.sorted[i] = qa.qubits[i];
    for (int i = 0; i < so.q; i++)
        for (int j = i + 1; j < so.q; j++)
            if (qa.sorted[j] < qa.sorted[i]) {
                const int t = qa.sorted[i];
                qa.sorted[i] = qa.sorted[j];
                qa.sorted[j] = t;
            }
    const uint64_t dim = 1ull << n;
    if (so.kind == 2) {
        hipLaunchKernelGGL(k_srn_full, dim3(grid_for(dim / 2, 256)), dim3(256), 0, st, s, dim / 2, so.qubits[0]);
    } else if (so.kind == 1) {
        hipLaunchKernelGGL(k_diag_full, dim3(grid_for(dim, 256)), dim3(256), 0, st, s, dim, qa, coef + so.coef);
    } else {
        const uint64_t ng = dim >> so.q;
        const dim3 gr(grid_for(ng, 256)), bl(256);
        const c64* M = coef + so.coef;
        switch (so.q) {
        case 1: hipLaunchKernelGGL(k_dense_full<1>, gr, bl, 0, st, s, ng, qa, M); break;
        case 2: hipLaunchKernelGGL(k_dense_full<2>, gr, bl, 0, st, s, ng, qa, M); break;
        case 3: hipLaunchKernelGGL(k_dense_full<3>, gr, bl, 0, st, s, ng, qa, M); break;
        case 4: hipLaunchKernelGGL(k_dense_full<4>, gr, bl, 0, st, s, ng, qa, M); break;
        default: hipLaunchKernelGGL(k_dense_full<5>, gr, bl, 0, st, s, ng, qa, M); break;
        }
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ state utils

__global__ void __launch_bounds__(256) k_fill_basis(c64* __restrict__ s, uint64_t dim)
{
    for (uint64_t x = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; x < dim;
         x += uint64_t(gridDim.x) * blockDim.x)
        s[x] = c64{x == 0 ? 1.0 : 0.0, 0.0};
}

__global__ void __launch_bounds__(256) k_copy(c64* __restrict__ d, const c64* __restrict__ s, uint64_t dim)
{
    for (uint64_t x = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; x < dim;
         x += uint64_t(gridDim.x) * blockDim.x)
        d[x] = s[x];
}

// splitmix64 finaliser (Steele, Lea & Flood 2014)
__device__ inline uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256)
k_random_state(c64* __restrict__ s, uint64_t dim, uint64_t seed, double* __restrict__ partials)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (uint64_t x = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; x < dim;
         x += uint64_t(gridDim.x) * blockDim.x) {
        const double u0 = double((splitmix64(seed + 2 * x + 1) >> 11) + 1) * 0x1p-53;
        const double u1 = double((splitmix64(seed + 2 * x + 2) >> 11) + 1) * 0x1p-53;
        const double r = sqrt(-2.0 * log(u0));
        double sn, cs;
        sincospi(2.0 * u1, &sn, &cs);
        const c64 v{r * cs, r * sn};
        s[x] = v;
        acc += v.re * v.re + v.im * v.im;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_scale(c64* __restrict__ s, uint64_t dim, double f)
{
    for (uint64_t x = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; x < dim;
         x += uint64_t(gridDim.x) * blockDim.x) {
        s[x].re *= f;
        s[x].im *= f;
    }
}

uint32_t random_state_grid(uint64_t dim) { return grid_for(dim, 256) > 4096 ? 4096 : grid_for(dim, 256); }

hipError_t launch_random_state(hipStream_t st, c64* s, uint64_t dim, uint64_t seed, double* partials)
{
    hipLaunchKernelGGL(k_random_state, dim3(random_state_grid(dim)), dim3(256), 0, st, s, dim, seed, partials);
    return hipGetLastError();
}

hipError_t launch_scale(hipStream_t st, c64* s, uint64_t dim, double f)
{
    hipLaunchKernelGGL(k_scale, dim3(grid_for(dim, 256)), dim3(256), 0, st, s, dim, f);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256)
k_gather(const c64* __restrict__ s, uint64_t dim, const uint64_t* __restrict__ idx, uint64_t cnt, c64* __restrict__ out)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < cnt) out[i] = idx[i] < dim ? s[idx[i]] : c64{0.0, 0.0};
}

hipError_t launch_gather(hipStream_t st, const c64* s, uint64_t dim, const uint64_t* idx, uint64_t cnt, c64* out)
{
    if (cnt == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, st, s, dim, idx, cnt, out);
    return hipGetLastError();
}

hipError_t launch_fill_basis(hipStream_t st, c64* s, uint64_t dim)
{
    hipLaunchKernelGGL(k_fill_basis, dim3(grid_for(dim, 256)), dim3(256), 0, st, s, dim);
    return hipGetLastError();
}

hipError_t launch_copy(hipStream_t st, c64* d, const c64* s, uint64_t dim)
{
    hipLaunchKernelGGL(k_copy, dim3(grid_for(dim, 256)), dim3(256), 0, st, d, s, dim);
    return hipGetLastError();
}

// -------------------------------------------------------------- sampling
//
// Scan definition (identical in both kernels, fixed operation order):
//   p_x = (re*re) + (im*im) with unfused IEEE ops (the reference's
//   expression, svsim_cpu_omp.hpp:345);
//   block of B = nthr*chunk amplitudes; thread t owns a contiguous chunk;
//   T_t = sequential sum of its chunk; P_t = exclusive prefix of T over the
//   threads (wave64 shuffle scan, then waves in order);
//   L_j = P_t + p_(first) + ... + p_j accumulated sequentially;
//   block total = L of the block's last amplitude.
// Global scan S_j = O_b + L_j with O_b the host's sequential prefix of the
// block totals; a draw r selects the first j with S_j > r (0 if none), which
// is the reference's interval rule scan[j] <= r < scan[j+1].

__device__ inline double prob_rn(c64 a) { return __dadd_rn(__dmul_rn(a.re, a.re), __dmul_rn(a.im, a.im)); }

__device__ double thread_prefix(const c64* __restrict__ blk, uint32_t chunk, double* lds_w)
{
    const uint32_t t = threadIdx.x, nthr = blockDim.x;
    double T = 0.0;
    for (uint32_t i = 0; i < chunk; i++) T = __dadd_rn(T, prob_rn(blk[size_t(t) * chunk + i]));
    const int width = nthr < 64 ? int(nthr) : 64;
    const int lane = int(t & 63u);
    double inc = T;
    for (int d = 1; d < width; d <<= 1) {
        const double u = __shfl_up(inc, d, width);
        if (lane >= d) inc = __dadd_rn(inc, u);
    }
    double exc = __shfl_up(inc, 1, width);
    if (lane == 0) exc = 0.0;
    const uint32_t w = t >> 6, nw = (nthr + 63) / 64;
    if (lane == width - 1) lds_w[w] = inc;
    __syncthreads();
    double wpre = 0.0;
    for (uint32_t v = 0; v < w; v++) wpre = __dadd_rn(wpre, lds_w[v]);
    (void)nw;
    __syncthreads();
    return w == 0 ? exc : __dadd_rn(wpre, exc);
}

__global__ void __launch_bounds__(256)
k_block_totals(const c64* __restrict__ s, uint32_t chunk, double* __restrict__ totals)
{
    __shared__ double lds_w[4];
    const uint64_t b = blockIdx.x;
    const c64* blk = s + b * uint64_t(blockDim.x) * chunk;
    double acc = thread_prefix(blk, chunk, lds_w);
    for (uint32_t i = 0; i < chunk; i++) acc = __dadd_rn(acc, prob_rn(blk[size_t(threadIdx.x) * chunk + i]));
    if (threadIdx.x == blockDim.x - 1) totals[b] = acc;
}

// The same two passes with the block's probabilities staged in LDS first:
// loaded lane-consecutive (coalesced), stored so that thread t's chunk sits
// at t * (chunk + 1) (the pad spreads the per-thread runs over the banks).
// Every sum keeps the order of the direct kernels above, so the totals and
// samples are bit-identical to theirs.
__device__ inline uint32_t staged_index(uint32_t j, uint32_t lg) { return j + (j >> lg); }

__device__ void stage_probs(const c64* __restrict__ blk, uint32_t lg, double* lp)
{
    const uint32_t bs = blockDim.x << lg;
    for (uint32_t j = threadIdx.x; j < bs; j += blockDim.x) lp[staged_index(j, lg)] = prob_rn(blk[j]);
    __syncthreads();
}

__device__ double thread_prefix_staged(const double* lp, uint32_t lg, double* lds_w)
{
    const uint32_t t = threadIdx.x, nthr = blockDim.x, chunk = 1u << lg;
    const double* mine = lp + t * (chunk + 1);
    double T = 0.0;
    for (uint32_t i = 0; i < chunk; i++) T = __dadd_rn(T, mine[i]);
    const int width = nthr < 64 ? int(nthr) : 64;
    const int lane = int(t & 63u);
    double inc = T;
    for (int d = 1; d < width; d <<= 1) {
        const double u = __shfl_up(inc, d, width);
        if (lane >= d) inc = __dadd_rn(inc, u);
    }
    double exc = __shfl_up(inc, 1, width);
    if (lane == 0) exc = 0.0;
    const uint32_t w = t >> 6;
    if (lane == width - 1) lds_w[w] = inc;
    __syncthreads();
    double wpre = 0.0;
    for (uint32_t v = 0; v < w; v++) wpre = __dadd_rn(wpre, lds_w[v]);
    __syncthreads();
    return w == 0 ? exc : __dadd_rn(wpre, exc);
}

__global__ void __launch_bounds__(256)
k_block_totals_staged(const c64* __restrict__ s, uint32_t lg, double* __restrict__ totals)
{
    extern __shared__ double lp[];
    __shared__ double lds_w[4];
    const uint64_t b = blockIdx.x;
    const uint32_t chunk = 1u << lg;
    stage_probs(s + (b * blockDim.x << lg), lg, lp);
    double acc = thread_prefix_staged(lp, lg, lds_w);
    const double* mine = lp + threadIdx.x * (chunk + 1);
    for (uint32_t i = 0; i < chunk; i++) acc = __dadd_rn(acc, mine[i]);
    if (threadIdx.x == blockDim.x - 1) totals[b] = acc;
}

// One workgroup per (block, shot range).  job = {block, first shot, n shots}.
__global__ void __launch_bounds__(256)
k_sample(const c64* __restrict__ s, uint32_t chunk, const uint64_t* __restrict__ jobs,
         const double* __restrict__ offsets, const double* __restrict__ draws,
         uint64_t* __restrict__ out)
{
    __shared__ double lds_w[4];
    __shared__ unsigned long long best;
    const uint64_t b = jobs[3 * blockIdx.x + 0];
    const uint64_t first = jobs[3 * blockIdx.x + 1];
    const uint64_t cnt = jobs[3 * blockIdx.x + 2];
    const uint64_t bsize = uint64_t(blockDim.x) * chunk;
    const c64* blk = s + b * bsize;
    const double P = thread_prefix(blk, chunk, lds_w);
    const double O = offsets[blockIdx.x];
    for (uint64_t q = 0; q < cnt; q++) {
        const double r = draws[first + q];
        if (threadIdx.x == 0) best = ~0ull;
        __syncthreads();
        double acc = P;
        uint64_t found = ~0ull;
        for (uint32_t i = 0; i < chunk; i++) {
            acc = __dadd_rn(acc, prob_rn(blk[size_t(threadIdx.x) * chunk + i]));
            if (__dadd_rn(O, acc) > r) {
                found = uint64_t(threadIdx.x) * chunk + i;
                break;
            }
        }
        if (found != ~0ull) atomicMin(&best, (unsigned long long)found);
        __syncthreads();
        if (threadIdx.x == 0) out[first + q] = best == ~0ull ? 0 : b * bsize + best;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256)
k_sample_staged(const c64* __restrict__ s, uint32_t lg, const uint64_t* __restrict__ jobs,
                const double* __restrict__ offsets, const double* __restrict__ draws,
                uint64_t* __restrict__ out)
{
    extern __shared__ double lp[];
    __shared__ double lds_w[4];
    __shared__ unsigned long long best;
    const uint32_t chunk = 1u << lg;
    const uint64_t b = jobs[3 * blockIdx.x + 0];
    const uint64_t first = jobs[3 * blockIdx.x + 1];
    const uint64_t cnt = jobs[3 * blockIdx.x + 2];
    const uint64_t bsize = uint64_t(blockDim.x) << lg;
    stage_probs(s + b * bsize, lg, lp);
    const double P = thread_prefix_staged(lp, lg, lds_w);
    const double O = offsets[blockIdx.x];
    const double* mine = lp + threadIdx.x * (chunk + 1);
    for (uint64_t q = 0; q < cnt; q++) {
        const double r = draws[first + q];
        if (threadIdx.x == 0) best = ~0ull;
        __syncthreads();
        double acc = P;
        uint64_t found = ~0ull;
        for (uint32_t i = 0; i < chunk; i++) {
            acc = __dadd_rn(acc, mine[i]);
            if (__dadd_rn(O, acc) > r) {
                found = uint64_t(threadIdx.x) * chunk + i;
                break;
            }
        }
        if (found != ~0ull) atomicMin(&best, (unsigned long long)found);
        __syncthreads();
        if (threadIdx.x == 0) out[first + q] = best == ~0ull ? 0 : b * bsize + best;
        __syncthreads();
    }
}

static size_t staged_smem(uint32_t nthr, uint32_t chunk) { return (size_t(nthr) * chunk + nthr) * sizeof(double); }
static size_t g_staged_smem = 0;   // dynamic LDS the staged sampling kernels may use

hipError_t init_sample_kernels()
{
    // the largest staged block: 2^14 amplitudes over 256 threads (the kernels
    // also use a little static LDS, so ask for exactly this)
    const size_t want = staged_smem(256, 64);
    g_staged_smem = 0;
    if (want > sweep_smem_limit()) return hipSuccess;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_block_totals_staged),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(want)) == hipSuccess &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sample_staged),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(want)) == hipSuccess)
        g_staged_smem = want;
    (void)hipGetLastError();
    return hipSuccess;
}

static uint32_t log2_u32(uint32_t v)
{
    uint32_t l = 0;
    while ((1u << l) < v) l++;
    return l;
}

hipError_t launch_block_totals(hipStream_t st, const c64* s, uint64_t nblocks, uint32_t nthr,
                               uint32_t chunk, double* totals)
{
    const size_t smem = staged_smem(nthr, chunk);
    const bool staged = smem <= g_staged_smem;
    for (uint64_t b0 = 0; b0 < nblocks; b0 += (1ull << 30)) {
        const uint64_t cnt = nblocks - b0 < (1ull << 30) ? nblocks - b0 : (1ull << 30);
        if (staged)
            hipLaunchKernelGGL(k_block_totals_staged, dim3(uint32_t(cnt)), dim3(nthr), smem, st,
                               s + b0 * uint64_t(nthr) * chunk, log2_u32(chunk), totals + b0);
        else
            hipLaunchKernelGGL(k_block_totals, dim3(uint32_t(cnt)), dim3(nthr), 0, st,
                               s + b0 * uint64_t(nthr) * chunk, chunk, totals + b0);
    }
    return hipGetLastError();
}

hipError_t launch_sample(hipStream_t st, const c64* s, uint32_t nthr, uint32_t chunk,
                         const uint64_t* jobs, uint32_t njobs, const double* offsets,
                         const double* draws, uint64_t* out)
{
    if (njobs == 0) return hipSuccess;
    const size_t smem = staged_smem(nthr, chunk);
    if (smem <= g_staged_smem)
        hipLaunchKernelGGL(k_sample_staged, dim3(njobs), dim3(nthr), smem, st, s, log2_u32(chunk), jobs, offsets,
                           draws, out);
    else
        hipLaunchKernelGGL(k_sample, dim3(njobs), dim3(nthr), 0, st, s, chunk, jobs, offsets, draws, out);
    return hipGetLastError();
}

// ----------------------------------------------------------- Pauli sums
//
// lam[y] = sum_g sum_{t in g} c'_t (-1)^{|x & z_t|} psi[x], x = y ^ xmask_g,
// c'_t = coeff_t * i^{|xmask & zmask_t|}.  Block partials of
// Re(conj(psi[y]) lam[y]) give <psi|H|psi>.

__global__ void __launch_bounds__(256)
k_pauli(const c64* __restrict__ psi, c64* __restrict__ lam, uint64_t dim,
        const PauliGroup* __restrict__ groups, uint32_t ngroups, const PauliTerm* __restrict__ terms,
        double* __restrict__ partials, c64* __restrict__ copy)
{
    __shared__ double red[256];
    double acc = 0.0;
    // blocks of kPauliPerThread x 256 amplitudes, grid-strided; each thread
    // keeps its kPauliPerThread amplitudes in registers with the groups and
    // terms outermost, so each term is read once per block, not per amplitude
    constexpr int M = kPauliPerThread;
    for (uint64_t yb = uint64_t(blockIdx.x) * blockDim.x * M; yb < dim; yb += uint64_t(gridDim.x) * blockDim.x * M) {
        c64 l[M];
#pragma unroll
        for (int j = 0; j < M; j++) l[j] = c64{0.0, 0.0};
        for (uint32_t g = 0; g < ngroups; g++) {
            const PauliGroup G = groups[g];
            double fr[M], fi[M];
#pragma unroll
            for (int j = 0; j < M; j++) fr[j] = fi[j] = 0.0;
            for (uint32_t t = G.begin; t < G.end; t++) {
                const PauliTerm T = terms[t];
#pragma unroll
                for (int j = 0; j < M; j++) {
                    const uint64_t x = (yb + threadIdx.x + uint64_t(blockDim.x) * j) ^ G.xmask;
                    const bool neg = __popcll(x & T.zmask) & 1;
                    fr[j] += neg ? -T.cre : T.cre;
                    fi[j] += neg ? -T.cim : T.cim;
                }
            }
#pragma unroll
            for (int j = 0; j < M; j++) {
                const uint64_t y = yb + threadIdx.x + uint64_t(blockDim.x) * j;
                if (y < dim) l[j] = cmac(l[j], c64{fr[j], fi[j]}, psi[y ^ G.xmask]);
            }
        }
#pragma unroll
        for (int j = 0; j < M; j++) {
            const uint64_t y = yb + threadIdx.x + uint64_t(blockDim.x) * j;
            if (y < dim) {
                if (lam) lam[y] = l[j];
                const c64 p = psi[y];
                if (copy) copy[y] = p;
                acc += p.re * l[j].re + p.im * l[j].im;
            }
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

uint32_t pauli_grid(uint64_t dim)
{
    const uint64_t g = grid_for(dim, 256 * kPauliPerThread);
    return g > 2048 ? 2048 : uint32_t(g);
}

hipError_t launch_pauli(hipStream_t st, const c64* psi, c64* lam, uint64_t dim, const PauliGroup* groups,
                        uint32_t ngroups, const PauliTerm* terms, double* partials, c64* copy)
{
    hipLaunchKernelGGL(k_pauli, dim3(pauli_grid(dim)), dim3(256), 0, st, psi, lam, dim, groups, ngroups,
                       terms, partials, copy);
    return hipGetLastError();
}

// Sum of per-tile gradient partials, one thread per accumulator (fixed order).
// ------------------------------------------------------- segment sums
//
// dst[seg.dst] = sum of src[seg.src .. seg.src + seg.len): one workgroup per
// segment; thread t adds elements t, t + 256, ... in order, then a fixed
// tree over the threads -- the same order on every run.

__global__ void __launch_bounds__(kThreads)
k_sum_segments(const double* __restrict__ src, const SumSegment* __restrict__ segs, double* __restrict__ dst)
{
    __shared__ double red[kThreads];
    const SumSegment g = segs[blockIdx.x];
    double acc = 0.0;
    for (uint64_t i = threadIdx.x; i < g.len; i += blockDim.x) acc += src[g.src + i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[g.dst] = red[0];
}

// Rows of a [rows][len] array: dst[r] = sum of row r (same order as above).
__global__ void __launch_bounds__(kThreads)
k_sum_rows(const double* __restrict__ src, uint64_t len, double* __restrict__ dst)
{
    __shared__ double red[kThreads];
    const double* row = src + uint64_t(blockIdx.x) * len;
    double acc = 0.0;
    for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) acc += row[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[blockIdx.x] = red[0];
}

// ------------------------------------------------------- measurement
//
// P(1) of qubit q: thread t of the grid reads the amplitudes with bit q set
// at pair indices i = t, t + G, ... (G = grid threads; pos1 = i with a 1
// inserted at bit q), four loads in flight, and sums |a|^2 in that order;
// then a fixed tree per workgroup, and one workgroup sums the partials in a
// fixed order.  The same order on every run.

__global__ void __launch_bounds__(kThreads)
k_prob_one(const c64* __restrict__ s, uint64_t half, uint32_t q, double* __restrict__ partials)
{
    __shared__ double red[kThreads];
    const uint64_t G = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t lo = (1ull << q) - 1;
    double acc = 0.0;
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i + 3 * G < half; i += 4 * G) {
        c64 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t j = i + u * G;
            v[u] = s[((j & ~lo) << 1) | (1ull << q) | (j & lo)];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) acc += v[u].re * v[u].re + v[u].im * v[u].im;
    }
    for (; i < half; i += G) {
        const c64 v = s[((i & ~lo) << 1) | (1ull << q) | (i & lo)];
        acc += v.re * v.re + v.im * v.im;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Density matrix of nh qubits (index row | col << nh): partial sums of
// |Re rho[i, i]| over the rows i with bit q set -- the diagonal entries only
// (2^(nh-1) strided reads), as Measure_GATE of the density-matrix engine
// sums them (svsim_qsharp_noise/svsim_cpu_omp.hpp:1737-1757).
__global__ void __launch_bounds__(kThreads)
k_prob_one_dm(const c64* __restrict__ s, uint32_t nh, uint32_t q, double* __restrict__ partials)
{
    __shared__ double red[kThreads];
    const uint64_t G = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t half = 1ull << (nh - 1);
    const uint64_t lo = (1ull << q) - 1;
    double acc = 0.0;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < half; j += G) {
        const uint64_t i = ((j & ~lo) << 1) | (1ull << q) | (j & lo);
        acc += fabs(s[(i << nh) | i].re);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// dm: collapse factors 1/P(1) or 1/(1 - P(1)) (density matrix) instead of
// their square roots (state vector).
__global__ void __launch_bounds__(kThreads)
k_measure_decide(const double* __restrict__ partials, uint32_t nparts, double draw, c64* __restrict__ collapse,
                 MeasureResult* __restrict__ res, int dm)
{
    __shared__ double red[kThreads];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) acc += partials[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double p = red[0];
        const bool one = draw < p;
        const double w = one ? p : 1.0 - p;
        const double f = dm ? 1.0 / w : 1.0 / sqrt(w);
        collapse[0] = c64{one ? 0.0 : f, 0.0};
        collapse[1] = c64{one ? f : 0.0, 0.0};
        MeasureResult r;
        r.prob = p;
        r.outcome = one ? 1 : 0;
        r.pad = 0;
        *res = r;
    }
}

uint32_t measure_partials(int n)
{
    const uint64_t half = 1ull << (n - 1);
    const uint64_t g = (half + kThreads * 4 - 1) / (kThreads * 4);
    return uint32_t(g > 2048 ? 2048 : (g < 1 ? 1 : g));
}

hipError_t launch_measure(hipStream_t st, const c64* s, int n, uint32_t q, double draw, double* partials,
                          c64* collapse, MeasureResult* res, int dm_n)
{
    if (dm_n > 0) {
        const uint32_t g = measure_partials(dm_n);
        hipLaunchKernelGGL(k_prob_one_dm, dim3(g), dim3(kThreads), 0, st, s, uint32_t(dm_n), q, partials);
        hipLaunchKernelGGL(k_measure_decide, dim3(1), dim3(kThreads), 0, st, partials, g, draw, collapse, res, 1);
        return hipGetLastError();
    }
    const uint32_t g = measure_partials(n);
    hipLaunchKernelGGL(k_prob_one, dim3(g), dim3(kThreads), 0, st, s, 1ull << (n - 1), q, partials);
    hipLaunchKernelGGL(k_measure_decide, dim3(1), dim3(kThreads), 0, st, partials, g, draw, collapse, res, 0);
    return hipGetLastError();
}

hipError_t launch_sum_segments(hipStream_t st, const double* src, const SumSegment* segs, uint64_t nseg,
                               double* dst)
{
    const uint64_t max_grid = 1ull << 30;
    for (uint64_t s0 = 0; s0 < nseg; s0 += max_grid) {
        const uint32_t cnt = uint32_t(nseg - s0 < max_grid ? nseg - s0 : max_grid);
        hipLaunchKernelGGL(k_sum_segments, dim3(cnt), dim3(kThreads), 0, st, src, segs + s0, dst);
    }
    return hipGetLastError();
}

hipError_t launch_sum_rows(hipStream_t st, const double* src, uint32_t rows, uint64_t len, double* dst)
{
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_rows, dim3(rows), dim3(kThreads), 0, st, src, len, dst);
    return hipGetLastError();
}

}  // namespace svmi
