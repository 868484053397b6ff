// svmi_api.hip -- C ABI of the engine (include/svmi.h, include/svmi_debug.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/svmi.h"
#include "../../include/svmi_debug.h"
#include "svmi_exec.hpp"
#include "svmi_kernels.hpp"
#include "svmi_lower.hpp"
#include "svmi_plan.hpp"
#include "svmi_jit.hpp"
#include "svmi_qasm.hpp"
#include "svmi_qir.hpp"
#include "svmi_tuning.hpp"

using namespace svmi;

namespace {

thread_local std::string g_err;

svmi_status fail(svmi_status st, const std::string& msg)
{
    g_err = msg;
    return st;
}

// Adds v to a GF(2) basis kept by leading bit (basis[b] has top bit b);
// false when v is already in its span.
bool gf2_insert(uint32_t (&basis)[32], uint32_t v)
{
    for (int b = 31; b >= 0 && v; b--)
        if ((v >> b) & 1u) {
            if (!basis[b]) {
                basis[b] = v;
                return true;
            }
            v ^= basis[b];
        }
    return false;
}

bool gf2_in_span(const uint32_t (&basis)[32], uint32_t v)
{
    for (int b = 31; b >= 0 && v; b--)
        if ((v >> b) & 1u) {
            if (!basis[b]) return false;
            v ^= basis[b];
        }
    return v == 0;
}

// A phase's LDS addressing stays inside its tile and is a bijection there:
// the tile map's columns (scol at the tile register bits, gm at the group
// bits) are independent and below 2^k, the gather columns span the same
// register columns, gxor and sxor lie in that span, and a two-state phase's
// state column is 2^k.
bool phase_map_valid(const PhaseInfo& ph, uint32_t k, bool two)
{
    const int rb = two ? kRegBitsAdj : kRegBits;
    if (k < uint32_t(rb) || k - rb > 10) return false;
    const uint32_t lim = 1u << k;
    uint32_t all[32] = {}, reg[32] = {}, gat[32] = {};
    for (int s = 0; s < rb; s++) {
        if (ph.scol[s] >= lim || !gf2_insert(all, ph.scol[s])) return false;
        gf2_insert(reg, ph.scol[s]);
    }
    for (int i = 0; i < 10; i++) {
        if (uint32_t(i) >= k - rb) {
            if (ph.gm[i]) return false;
        } else if (ph.gm[i] >= lim || !gf2_insert(all, ph.gm[i])) {
            return false;
        }
    }
    for (int s = 0; s < rb; s++)
        if (!gf2_in_span(reg, ph.gcol[s]) || !gf2_insert(gat, ph.gcol[s])) return false;
    if (two && (ph.scol[kRegBits - 1] != lim || ph.gcol[kRegBits - 1] != lim)) return false;
    return gf2_in_span(reg, ph.gxor) && gf2_in_span(reg, ph.sxor);
}

#define SVMI_HIP(x)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? SVMI_ERR_OOM : SVMI_ERR_DEVICE,          \
                        std::string(#x) + ": " + hipGetErrorString(e_));                     \
    } while (0)

#define SVMI_TRY(x)                        \
    do {                                   \
        svmi_status s_ = (x);              \
        if (s_ != SVMI_OK) return s_;      \
    } while (0)

constexpr int kMaxQubits = 40;

// Device scratch of svmi_expval_static: a file-scope array that every GPU
// the library runs on holds until the process ends.  Host code never passes
// `scratch` itself to a kernel or copy (on the host it is only a stand-in for
// the symbol); it uses the device address hipGetSymbolAddress returns for
// the current device, looked up once per device by svmi_create.
constexpr size_t kStaticScratchBytes = size_t(1) << 20;
constexpr size_t kStaticMaxTerms = 16384;   // Pauli terms of one svmi_expval_static call
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
void* g_static_scratch[kMaxDevices] = {};   // per-device address of `scratch`
std::mutex g_static_scratch_mu;

// Device address of `scratch` on `device` (which must be current).
svmi_status static_scratch_ptr(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return fail(SVMI_ERR_DEVICE, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_static_scratch_mu);
    if (!g_static_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(SVMI_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_static_scratch[device] = ptr;
    }
    *out = g_static_scratch[device];
    return SVMI_OK;
}

// Growable device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    svmi_status ensure(size_t want)
    {
        if (want <= bytes) return SVMI_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        SVMI_HIP(hipMalloc(&p, want));
        bytes = want;
        return SVMI_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

struct DeviceProgram {
    DevBuf phases, ops, coef, hitab;
    template <class T>
    static svmi_status put(DevBuf& b, const std::vector<T>& v, hipStream_t st)
    {
        const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
        SVMI_TRY(b.ensure(bytes));
        if (!v.empty()) SVMI_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
        return SVMI_OK;
    }
    // op arrays end with one NOP record: the sweep kernels read the header
    // after each phase's last op
    static svmi_status put_ops(DevBuf& b, const std::vector<DOp>& v, hipStream_t st)
    {
        static const DOp nop = {};
        SVMI_TRY(b.ensure((v.size() + 1) * sizeof(DOp)));
        if (!v.empty()) SVMI_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(DOp), hipMemcpyHostToDevice, st));
        SVMI_HIP(hipMemcpyAsync(static_cast<DOp*>(b.p) + v.size(), &nop, sizeof(DOp), hipMemcpyHostToDevice, st));
        return SVMI_OK;
    }
    svmi_status upload(const Program& p, hipStream_t st)
    {
        SVMI_TRY(put(phases, p.phases, st));
        SVMI_TRY(put_ops(ops, p.ops, st));
        SVMI_TRY(put(coef, p.coef, st));
        SVMI_TRY(put(hitab, p.hitab, st));
        SVMI_HIP(hipStreamSynchronize(st));
        return SVMI_OK;
    }
    void release()
    {
        phases.release();
        ops.release();
        coef.release();
        hitab.release();
    }
};

svmi_status check_gates(const svmi_gate* gates, size_t n_gates, int n)
{
    if (n_gates && !gates) return fail(SVMI_ERR_ARG, "null gate array");
    for (size_t i = 0; i < n_gates; i++) {
        std::string err;
        const svmi_status st = validate_gate(gates[i], n, err);
        if (st != SVMI_OK) return fail(st, "gate " + std::to_string(i) + ": " + err);
    }
    return SVMI_OK;
}

void lower_all(const svmi_gate* gates, size_t n_gates, std::vector<LOp>& ops)
{
    ops.resize(n_gates);
    for (size_t i = 0; i < n_gates; i++) {
        lower_gate(gates[i], ops[i]);
        ops[i].gate = int(i);
    }
}

void plan_forward(const svmi_gate* gates, size_t n_gates, int n, const PlanOptions& opt, Program& prog)
{
    std::vector<LOp> low, fused;
    lower_all(gates, n_gates, low);
    if (opt.fusion) {
        fuse_single_qubit_runs(low, fused);
        build_forward_program(fused, n, opt, prog);
    } else {
        build_forward_program(low, n, opt, prog);
    }
}

// QIR gate list on n allocated qubits -> segmented program on n_prog >= n
// qubits (n_prog >= 4: every sweep needs 4 register bits).
svmi_status plan_qir(const svmi_qir_gate* gates, size_t n_gates, int n, int n_prog, const PlanOptions& opt,
                     Program& prog, bool dm = false)
{
    if (n_gates && !gates) return fail(SVMI_ERR_ARG, "null gate array");
    std::string err;
    for (size_t i = 0; i < n_gates; i++) {
        const svmi_status st = validate_qir_gate(gates[i], n, err);
        if (st != SVMI_OK) return fail(st, "QIR gate " + std::to_string(i) + ": " + err);
    }
    std::vector<std::vector<LOp>> segs;
    std::vector<MeasureStep> meas;
    if (dm)
        lower_dm_qir(gates, n_gates, n, segs, meas);
    else
        lower_qir(gates, n_gates, n, segs, meas);
    for (auto& seg : segs) {
        std::vector<LOp> f;
        fuse_single_qubit_runs(seg, f);
        seg.swap(f);
    }
    try {
        build_segmented_program(segs, meas, n_prog, opt, prog);
    } catch (const std::exception& e) {
        return fail(SVMI_ERR_ARG, std::string("QIR planning: ") + e.what());
    }
    return SVMI_OK;
}

// gfx target of the current device, without feature suffixes
std::string device_arch()
{
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return "gfx950";
    std::string a = p.gcnArchName;
    return a.substr(0, a.find(':'));
}

svmi_status check_terms(const svmi_pauli_term* terms, size_t n_terms, int n)
{
    if (n_terms && !terms) return fail(SVMI_ERR_ARG, "null term array");
    const uint64_t lim = n >= 64 ? ~0ull : ((1ull << n) - 1);
    for (size_t i = 0; i < n_terms; i++)
        if ((terms[i].xmask | terms[i].zmask) & ~lim)
            return fail(SVMI_ERR_QUBIT, "Pauli term " + std::to_string(i) + " acts on a qubit >= n");
    return SVMI_OK;
}

// Group terms by X mask; coefficients absorb i^{|x&z|}.
void group_terms(const svmi_pauli_term* terms, size_t n_terms, std::vector<PauliGroup>& groups,
                 std::vector<PauliTerm>& out)
{
    std::vector<size_t> idx(n_terms);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return terms[a].xmask < terms[b].xmask; });
    groups.clear();
    out.clear();
    for (size_t i : idx) {
        const svmi_pauli_term& t = terms[i];
        if (groups.empty() || groups.back().xmask != t.xmask)
            groups.push_back(PauliGroup{t.xmask, uint32_t(out.size()), uint32_t(out.size())});
        const int ny = __builtin_popcountll(t.xmask & t.zmask) & 3;
        const double re[4] = {1, 0, -1, 0}, im[4] = {0, 1, 0, -1};
        out.push_back(PauliTerm{t.zmask, t.coeff * re[ny], t.coeff * im[ny]});
        groups.back().end = uint32_t(out.size());
    }
}

svmi_status check_params(const svmi_gate* gates, size_t n_gates, const int32_t* param_slots, size_t n_params)
{
    if (n_gates && !param_slots) return fail(SVMI_ERR_ARG, "null param_slots");
    for (size_t i = 0; i < n_gates; i++) {
        if (!gate_is_unitary(gates[i].op))
            return fail(SVMI_ERR_UNSUPPORTED, "adjoint needs unitary gates; gate " + std::to_string(i) +
                                                  " (R or SRN) is not unitary");
        for (int s = 0; s < 3; s++) {
            const int32_t p = param_slots[3 * i + s];
            if (p < -1 || (p >= 0 && size_t(p) >= n_params))
                return fail(SVMI_ERR_ARG, "param_slots[" + std::to_string(3 * i + s) + "] out of range");
        }
    }
    return SVMI_OK;
}

// Backward entries: for gates in reverse, U^dagger on (phi, lambda), then
// K = U^dagger dU/dtheta_s for each trainable slot (read only).
void adjoint_entries(const svmi_gate* gates, size_t n_gates, const int32_t* param_slots,
                     std::vector<AdjEntry>& entries)
{
    entries.clear();
    for (size_t ii = n_gates; ii-- > 0;) {
        LOp U;
        lower_gate(gates[ii], U);
        U.gate = int(ii);
        AdjEntry inv;
        lop_adjoint(U, inv.op);
        entries.push_back(inv);
        for (int s = 0; s < 3; s++) {
            const int32_t p = param_slots[3 * ii + s];
            if (p < 0) continue;
            LOp dU;
            if (!gate_derivative(gates[ii], s, dU)) continue;
            dU.gate = int(ii);
            AdjEntry g;
            g.grad = true;
            g.param = p;
            // K = U^dagger * dU  (dU applied first, then U^dagger)
            if (inv.op.kind == LOp::DIAG && dU.kind == LOp::DIAG) {
                g.op = dU;
                for (size_t r = 0; r < dU.m.size(); r++) g.op.m[r] = cmul(inv.op.m[r], dU.m[r]);
            } else {
                lop_compose_same(dU, inv.op, g.op);
            }
            entries.push_back(g);
        }
    }
}

// ---- batch planning (shared by the device path and its host check) ----
struct BatchPlan {
    std::vector<Program> fwd, adj;
    std::vector<SweepInfo> sweeps;
    std::vector<PhaseInfo> phases;
    std::vector<DOp> ops;
    std::vector<c64> coef;
    std::vector<uint64_t> hitab;
    std::vector<PauliGroup> groups;
    std::vector<PauliTerm> terms;
    std::vector<BatchCircuit> circs;     // phi/lam unset; state_off gives the layout
    std::vector<uint64_t> state_off;
    uint64_t state_bytes = 0;
    std::vector<BatchStep> steps;        // level by level
    std::vector<uint32_t> level_begin;   // first step of each level, then steps.size()
    std::vector<BatchLaunch> launches;   // in stream order
    // final sums: partials = [gradient partials | <H> partials] ->
    // results = [gradient accumulators (n_slots) | <H> per circuit]
    std::vector<SumSegment> segs;
    uint64_t n_tasks = 0, n_e_partials = 0, n_grad_partials = 0;
    uint32_t n_slots = 0;
    std::vector<size_t> need_fill;       // circuits without forward sweeps
};

// Consistency of a batch plan: every index the kernel will follow is in range
// and every dependency points to a step of the same circuit in an earlier
// level (the level launches run in stream order).
svmi_status verify_batch(const BatchPlan& bp)
{
    if (bp.level_begin.empty() || bp.level_begin.front() != 0 || bp.level_begin.back() != bp.steps.size())
        return fail(SVMI_ERR_STATE, "batch plan: bad levels");
    std::vector<uint32_t> level_of(bp.steps.size());
    for (size_t l = 0; l + 1 < bp.level_begin.size(); l++) {
        if (bp.level_begin[l] >= bp.level_begin[l + 1]) return fail(SVMI_ERR_STATE, "batch plan: empty level");
        for (uint32_t s = bp.level_begin[l]; s < bp.level_begin[l + 1]; s++) level_of[s] = uint32_t(l);
    }
    // launches tile the step list in order; each holds steps of one level,
    // one kind and one tile size, within the LDS the kernels may use
    uint32_t next = 0;
    for (const BatchLaunch& L : bp.launches) {
        if (L.step_begin != next || L.step_end <= L.step_begin || L.step_end > bp.steps.size() ||
            level_of[L.step_begin] != level_of[L.step_end - 1] ||
            L.task_begin != bp.steps[L.step_begin].task_begin ||
            batch_launch_smem(L.kind, L.k, L.max_grad, L.max_xf) > size_t(160) * 1024)
            return fail(SVMI_ERR_STATE, "batch plan: bad launch");
        uint64_t nt = 0;
        for (uint32_t s = L.step_begin; s < L.step_end; s++) {
            const BatchStep& b = bp.steps[s];
            const bool sweep = b.kind == BSTEP_FWD || b.kind == BSTEP_ADJ;
            if (b.kind != L.kind || (sweep && bp.sweeps[b.sweep].k != L.k) || b.init_zero != L.init ||
                (b.kind == BSTEP_ADJ && bp.sweeps[b.sweep].n_grad > L.max_grad) ||
                (sweep && bp.sweeps[b.sweep].n_xf > L.max_xf) ||
                (sweep && bp.sweeps[b.sweep].max_dense > L.max_dense) ||
                (b.kind == BSTEP_FWD && bp.sweeps[b.sweep].mapped && !L.mapped))
                return fail(SVMI_ERR_STATE, "batch plan: mixed launch");
            nt += b.n_tasks;
        }
        if (nt != L.n_tasks) return fail(SVMI_ERR_STATE, "batch plan: launch task count");
        next = L.step_end;
    }
    if (next != bp.steps.size()) return fail(SVMI_ERR_STATE, "batch plan: steps outside launches");
    uint64_t t = 0;
    for (size_t s = 0; s < bp.steps.size(); s++) {
        const BatchStep& b = bp.steps[s];
        if (b.task_begin != t || b.n_tasks == 0 || b.circuit >= bp.circs.size())
            return fail(SVMI_ERR_STATE, "batch plan: bad task range");
        t += b.n_tasks;
        if (b.dep >= 0 && (size_t(b.dep) >= s || bp.steps[b.dep].circuit != b.circuit ||
                           level_of[size_t(b.dep)] >= level_of[s]))
            return fail(SVMI_ERR_STATE, "batch plan: bad dependency");
        if (b.kind == BSTEP_FWD || b.kind == BSTEP_ADJ) {
            if (b.sweep >= bp.sweeps.size()) return fail(SVMI_ERR_STATE, "batch plan: bad sweep index");
            const SweepInfo& sw = bp.sweeps[b.sweep];
            if (sw.n_tiles != b.n_tasks || sw.k > uint32_t(b.kind == BSTEP_ADJ ? kMaxTileQubitsAdj : kMaxTileQubits) ||
                sw.k < uint32_t(kRegBits) || sw.phase_end > bp.phases.size() || sw.phase_begin > sw.phase_end ||
                sw.hitab + (1ull << (sw.k - sw.c)) > bp.hitab.size() || sw.n_grad > uint32_t(kMaxGradsPerSweep))
                return fail(SVMI_ERR_STATE, "batch plan: bad sweep");
            if ((sw.k + uint32_t(__builtin_popcountll(sw.extmask))) != bp.circs[b.circuit].n)
                return fail(SVMI_ERR_STATE, "batch plan: sweep does not match the circuit size");
            if (b.kind == BSTEP_ADJ && b.partial + sw.n_tiles * sw.n_grad > bp.n_grad_partials)
                return fail(SVMI_ERR_STATE, "batch plan: bad gradient partial range");
            if (sw.op_begin > sw.op_end || sw.op_end > bp.ops.size() || sw.n_xf > uint32_t(kMaxXfSlots) ||
                16u * sw.n_lt > sw.n_xf || sw.lt_coef + 16ull * sw.n_lt > bp.coef.size())
                return fail(SVMI_ERR_STATE, "batch plan: bad sweep op range");
            if (sw.mapped) {   // the store's tile map: a bijection of the tile
                uint32_t basis[32] = {};
                bool ok = true;
                for (uint32_t i = 0; i < 14 && ok; i++)
                    ok = i < sw.k ? (sw.fcol[i] < (1u << sw.k) && gf2_insert(basis, sw.fcol[i])) : sw.fcol[i] == 0;
                const uint32_t rb = uint32_t(b.kind == BSTEP_ADJ ? kRegBitsAdj : kRegBits);
                for (uint32_t j = 0; j < 16 && ok; j++)
                    ok = j < (1u << rb) ? sw.fj[j] == sweep_map(sw, j << (sw.k - rb)) : sw.fj[j] == 0;
                if (!ok) return fail(SVMI_ERR_STATE, "batch plan: bad tile map");
            }
            for (uint32_t o = sw.op_begin; o < sw.op_end; o++)
                if ((has_tile_factor(bp.ops[o].h) && bp.ops[o].h.grad >= sw.n_xf) ||
                    (bp.ops[o].h.type == DOP_PHASETAB && bp.ops[o].h.gmask != kLaneTabGlobal &&
                     uint64_t(bp.ops[o].h.gmask) + 16ull * bp.ops[o].h.s0 > sw.n_xf))
                    return fail(SVMI_ERR_STATE, "batch plan: bad phase factor slot");
            for (uint32_t p = sw.phase_begin; p < sw.phase_end; p++) {
                const PhaseInfo& ph = bp.phases[p];
                if (ph.op_end > bp.ops.size() || ph.op_begin > ph.op_end || ph.op_begin < sw.op_begin ||
                    ph.op_end > sw.op_end)
                    return fail(SVMI_ERR_STATE, "batch plan: bad phase");
                // adjoint phases: 3 tile bits, then k (lambda's half of the LDS)
                const bool two = b.kind == BSTEP_ADJ;
                for (int r = 0; r < kRegBits; r++) {
                    const bool state_bit = two && r == kRegBits - 1;
                    if (state_bit ? ph.r[r] != sw.k : ph.r[r] >= sw.k)
                        return fail(SVMI_ERR_STATE, "batch plan: bad register bit");
                }
                if (!phase_map_valid(ph, sw.k, two) || (ph.mapped && !sw.mapped))
                    return fail(SVMI_ERR_STATE, "batch plan: bad gather map");
                for (uint32_t o = ph.op_begin; o < ph.op_end; o++)
                    if (bp.ops[o].h.coef > bp.coef.size()) return fail(SVMI_ERR_STATE, "batch plan: bad coef offset");
            }
        } else if (b.kind == BSTEP_PAULI) {
            if (b.partial + b.n_tasks > bp.n_e_partials) return fail(SVMI_ERR_STATE, "batch plan: bad E partials");
        }
    }
    if (t != bp.n_tasks) return fail(SVMI_ERR_STATE, "batch plan: task count");
    for (const SumSegment& g : bp.segs)
        if (g.len == 0 || g.src + g.len > bp.n_grad_partials + bp.n_e_partials ||
            g.dst >= uint64_t(bp.n_slots) + bp.circs.size())
            return fail(SVMI_ERR_STATE, "batch plan: bad sum segment");
    for (const BatchCircuit& c : bp.circs)
        if (c.group_end > bp.groups.size() || c.slot_begin + c.n_slots > bp.n_slots ||
            c.e_partial + c.n_e_partials > bp.n_e_partials)
            return fail(SVMI_ERR_STATE, "batch plan: bad circuit record");
    return SVMI_OK;
}

svmi_status build_batch(const svmi_batch_item* items, size_t n_items, const double* expvals, const double* grads,
                        BatchPlan& bp)
{
    if (!items || !expvals) return fail(SVMI_ERR_ARG, "null batch arrays");
    size_t total_params = 0;
    for (size_t i = 0; i < n_items; i++) {
        const svmi_batch_item& it = items[i];
        if (it.n_qubits < 1 || it.n_qubits > 30)
            return fail(SVMI_ERR_ARG, "batch item " + std::to_string(i) + ": n_qubits must be in [1, 30]");
        if ((it.fwd_tile && (it.fwd_tile < kRegBits || it.fwd_tile > kMaxTileQubits)) ||
            (it.adj_tile && (it.adj_tile < kRegBits || it.adj_tile > kMaxTileQubitsAdj)) || it.op_cost > 33)
            return fail(SVMI_ERR_ARG, "batch item " + std::to_string(i) + ": fwd_tile must be in [4, 13], adj_tile in "
                                                                             "[4, 11], op_cost + 1 in [1, 33]");
        SVMI_TRY(check_gates(it.gates, it.n_gates, it.n_qubits));
        SVMI_TRY(check_terms(it.terms, it.n_terms, it.n_qubits));
        SVMI_TRY(check_params(it.gates, it.n_gates, it.param_slots, it.n_params));
        total_params += it.n_params;
    }
    if (total_params && !grads) return fail(SVMI_ERR_ARG, "null gradient array");
    std::vector<int> ni(n_items);
    std::vector<std::vector<PauliGroup>> pg(n_items);
    std::vector<std::vector<PauliTerm>> pt(n_items);
    bp.fwd.resize(n_items);
    bp.adj.resize(n_items);
    {
        const unsigned nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<size_t> next{0};
        auto worker = [&]() {
            for (size_t i = next++; i < n_items; i = next++) {
                const svmi_batch_item& it = items[i];
                ni[i] = std::max(int(it.n_qubits), kRegBits);
                // tile sizes and fusion op cost: the item's, else the
                // autotuned table's for its qubit count
                const BatchTuning& tu = batch_tuning(ni[i]);
                PlanOptions of;
                of.tile_qubits = it.fwd_tile ? it.fwd_tile : tu.fwd_tile;
                of.op_overhead = it.op_cost ? it.op_cost - 1 : tu.op_cost;
                plan_forward(it.gates, it.n_gates, ni[i], of, bp.fwd[i]);
                std::vector<AdjEntry> entries;
                adjoint_entries(it.gates, it.n_gates, it.param_slots, entries);
                PlanOptions oa;
                oa.tile_qubits = it.adj_tile ? it.adj_tile : tu.adj_tile;
                oa.op_overhead = of.op_overhead;
                build_adjoint_program(entries, ni[i], oa, bp.adj[i]);
                group_terms(it.terms, it.n_terms, pg[i], pt[i]);
            }
        };
        std::vector<std::thread> pool;
        for (unsigned t = 0; t + 1 < nthr; t++) pool.emplace_back(worker);
        worker();
        for (auto& th : pool) th.join();
    }
    for (size_t i = 0; i < n_items; i++)
        for (const Program* pr : {&bp.fwd[i], &bp.adj[i]})
            for (const Step& stp : pr->steps)
                if (stp.kind != Step::SWEEP)
                    return fail(SVMI_ERR_UNSUPPORTED,
                                "batch item " + std::to_string(i) +
                                    ": an op does not fit a tile (a 5-qubit gate such as C4X, or a tile too small "
                                    "for the 4 low qubits plus the op's qubits)");
    auto append_program = [&](const Program& pr) {
        const uint32_t op_base = uint32_t(bp.ops.size()), coef_base = uint32_t(bp.coef.size());
        const uint32_t ph_base = uint32_t(bp.phases.size()), ht_base = uint32_t(bp.hitab.size());
        const uint32_t sw_base = uint32_t(bp.sweeps.size());
        for (DOp d : pr.ops) {
            d.h.coef += coef_base;
            if (d.h.type == DOP_ROT && d.h.s0) d.h.grad += coef_base;   // its table
            bp.ops.push_back(d);
        }
        for (PhaseInfo ph : pr.phases) {
            ph.op_begin += op_base;
            ph.op_end += op_base;
            bp.phases.push_back(ph);
        }
        for (SweepInfo sw : pr.sweeps) {
            sw.phase_begin += ph_base;
            sw.phase_end += ph_base;
            sw.op_begin += op_base;
            sw.op_end += op_base;
            sw.hitab += ht_base;
            sw.lt_coef += coef_base;
            bp.sweeps.push_back(sw);
        }
        bp.coef.insert(bp.coef.end(), pr.coef.begin(), pr.coef.end());
        bp.hitab.insert(bp.hitab.end(), pr.hitab.begin(), pr.hitab.end());
        return sw_base;
    };
    bp.circs.resize(n_items);
    bp.state_off.resize(n_items);
    std::vector<std::vector<BatchStep>> csteps(n_items);
    for (size_t i = 0; i < n_items; i++) {
        BatchCircuit& c = bp.circs[i];
        std::memset(&c, 0, sizeof(c));
        c.n = uint32_t(ni[i]);
        const uint64_t dim = 1ull << ni[i];
        bp.state_off[i] = bp.state_bytes;
        bp.state_bytes += ((2 * dim * sizeof(c64)) + 255) & ~uint64_t(255);
        c.group_begin = uint32_t(bp.groups.size());
        for (PauliGroup g : pg[i]) {
            g.begin += uint32_t(bp.terms.size());
            g.end += uint32_t(bp.terms.size());
            bp.groups.push_back(g);
        }
        bp.terms.insert(bp.terms.end(), pt[i].begin(), pt[i].end());
        c.group_end = uint32_t(bp.groups.size());
        auto& cs = csteps[i];
        const Program& fw = bp.fwd[i];
        const uint32_t fsw = append_program(fw);
        for (size_t s = 0; s < fw.sweeps.size(); s++) {
            BatchStep b;
            std::memset(&b, 0, sizeof(b));
            b.kind = BSTEP_FWD;
            b.circuit = uint32_t(i);
            b.sweep = fsw + uint32_t(s);
            b.n_tasks = fw.sweeps[s].n_tiles;
            b.init_zero = s == 0 ? 1u : 0u;
            cs.push_back(b);
        }
        if (fw.sweeps.empty()) bp.need_fill.push_back(i);
        BatchStep pb;
        std::memset(&pb, 0, sizeof(pb));
        pb.kind = BSTEP_PAULI;
        pb.circuit = uint32_t(i);
        pb.n_tasks = (dim + kPauliChunk - 1) / kPauliChunk;
        pb.partial = bp.n_e_partials;
        c.e_partial = uint32_t(bp.n_e_partials);
        c.n_e_partials = uint32_t(pb.n_tasks);
        bp.n_e_partials += pb.n_tasks;
        cs.push_back(pb);
        const Program& ad = bp.adj[i];
        const uint32_t asw = append_program(ad);
        c.slot_begin = bp.n_slots;
        c.n_slots = ad.n_grad;
        bp.n_slots += ad.n_grad;
        for (size_t s = 0; s < ad.sweeps.size(); s++) {
            BatchStep b;
            std::memset(&b, 0, sizeof(b));
            b.kind = BSTEP_ADJ;
            b.circuit = uint32_t(i);
            b.sweep = asw + uint32_t(s);
            b.n_tasks = ad.sweeps[s].n_tiles;
            b.partial = bp.n_grad_partials;
            bp.n_grad_partials += ad.sweeps[s].n_tiles * ad.sweeps[s].n_grad;
            cs.push_back(b);
        }
    }
    for (size_t i = 0; i < n_items; i++) {
        for (const BatchStep& b : csteps[i]) {
            if (b.kind != BSTEP_ADJ) continue;
            const SweepInfo& sw = bp.sweeps[b.sweep];
            for (uint32_t g = 0; g < sw.n_grad; g++)
                bp.segs.push_back({b.partial + uint64_t(g) * sw.n_tiles, sw.n_tiles,
                                   uint64_t(bp.circs[i].slot_begin) + sw.grad_base + g});
        }
        const BatchCircuit& c = bp.circs[i];
        bp.segs.push_back({bp.n_grad_partials + c.e_partial, c.n_e_partials, uint64_t(bp.n_slots) + i});
    }
    // level l = step l of every circuit that has one, largest circuits first
    std::vector<size_t> corder(n_items);
    std::iota(corder.begin(), corder.end(), 0);
    std::stable_sort(corder.begin(), corder.end(), [&](size_t a, size_t b) {
        return (csteps[a].size() << ni[a]) > (csteps[b].size() << ni[b]);
    });
    std::vector<std::vector<uint32_t>> idx_of(n_items);
    size_t max_levels = 0;
    for (auto& cs : csteps) max_levels = std::max(max_levels, cs.size());
    // inside a level, steps of one kind and tile size are consecutive and
    // form one launch
    auto tile_k = [&](const BatchStep& b) -> uint32_t {
        return (b.kind == BSTEP_FWD || b.kind == BSTEP_ADJ) ? bp.sweeps[b.sweep].k : 0u;
    };
    for (size_t lv = 0; lv < max_levels; lv++) {
        bp.level_begin.push_back(uint32_t(bp.steps.size()));
        std::vector<size_t> members;
        for (size_t i : corder)
            if (lv < csteps[i].size()) members.push_back(i);
        std::stable_sort(members.begin(), members.end(), [&](size_t x, size_t y) {
            const BatchStep &bx = csteps[x][lv], &by = csteps[y][lv];
            if (bx.kind != by.kind) return bx.kind < by.kind;
            return tile_k(bx) > tile_k(by);
        });
        for (size_t i : members) {
            BatchStep b = csteps[i][lv];
            b.dep = lv == 0 ? -1 : int32_t(idx_of[i][lv - 1]);
            b.task_begin = bp.n_tasks;
            bp.n_tasks += b.n_tasks;
            const uint32_t idx = uint32_t(bp.steps.size());
            idx_of[i].push_back(idx);
            bp.steps.push_back(b);
            const uint32_t k = tile_k(b);
            const uint32_t ng = b.kind == BSTEP_ADJ ? bp.sweeps[b.sweep].n_grad : 0u;
            const uint32_t nx = (b.kind == BSTEP_FWD || b.kind == BSTEP_ADJ) ? bp.sweeps[b.sweep].n_xf : 0u;
            const uint32_t md = (b.kind == BSTEP_FWD || b.kind == BSTEP_ADJ) ? bp.sweeps[b.sweep].max_dense : 0u;
            const uint32_t mp = b.kind == BSTEP_FWD ? bp.sweeps[b.sweep].mapped : 0u;
            if (!bp.launches.empty() && bp.launches.back().step_end == idx &&
                bp.launches.back().step_begin >= bp.level_begin.back() && bp.launches.back().kind == b.kind &&
                bp.launches.back().k == k && bp.launches.back().init == b.init_zero) {
                BatchLaunch& L = bp.launches.back();
                L.step_end = idx + 1;
                L.n_tasks += b.n_tasks;
                L.max_grad = std::max(L.max_grad, ng);
                L.max_xf = std::max(L.max_xf, nx);
                L.max_dense = std::max(L.max_dense, md);
                L.mapped |= mp;
            } else {
                BatchLaunch L;
                std::memset(&L, 0, sizeof(L));
                L.step_begin = idx;
                L.step_end = idx + 1;
                L.kind = b.kind;
                L.k = k;
                L.max_grad = ng;
                L.max_xf = nx;
                L.max_dense = md;
                L.mapped = mp;
                L.init = b.init_zero;
                L.task_begin = b.task_begin;
                L.n_tasks = b.n_tasks;
                bp.launches.push_back(L);
            }
        }
    }
    bp.level_begin.push_back(uint32_t(bp.steps.size()));
    return verify_batch(bp);
}

void finish_batch(const BatchPlan& bp, const svmi_batch_item* items, size_t n_items, const std::vector<double>& gs,
                  const std::vector<double>& eo, double* expvals, double* grads)
{
    size_t goff = 0;
    for (size_t i = 0; i < n_items; i++) {
        expvals[i] = eo[i];
        for (size_t p = 0; p < items[i].n_params; p++) grads[goff + p] = 0.0;
        for (uint32_t s = 0; s < bp.adj[i].n_grad; s++)
            grads[goff + bp.adj[i].grad_param[s]] += 2.0 * gs[bp.circs[i].slot_begin + s];
        goff += items[i].n_params;
    }
}

}  // namespace

struct svmi_sim {
    int n = 0;          // qubits seen by the caller
    int n_int = 0;      // qubits of the device state (>= 4)
    int device = 0;
    uint64_t dim = 0;   // 2^n_int
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    c64* sv = nullptr;
    DevBuf aux[2];
    DevBuf scratch[6];
    std::vector<svmi_gate> gates;
    bool uploaded = false;
    Program prog;
    DeviceProgram dprog;
    JitSweeps jit;             // specialised kernels of prog's sweeps (SVMI_OPT_SPECIALIZE)
    bool specialize = false;
    bool op_cost_set = false;  // SVMI_OPT_OP_COST given (else specialised sweeps plan with op cost 0)
    std::string jit_error;     // why the last specialisation failed (the sweeps then interpret)
    PlanOptions opt;
    bool profile = false;
    int pipeline = 0;
    bool prefetch = false;     // SVMI_OPT_PREFETCH
    std::vector<hipEvent_t> prof_events;
    std::vector<double> step_ms;   // with profiling: per-step times of the last program run
    svmi_stats stats{};
    // backward program of the uploaded circuit for the last parameter map
    // (reused while the circuit, the map and the options stay the same)
    bool adj_valid = false;
    std::vector<int32_t> adj_slots;
    size_t adj_n_params = 0;
    Program adj_prog;
    DeviceProgram adj_dprog;
    DevBuf adj_gsum;
    void* static_scratch = nullptr;   // this device's address of the file-scope scratch array
    // QIR runs: program, device copy, measurement results; the P(1) of the
    // latest measurement (0 after create / reset_state)
    Program qir_prog;
    DeviceProgram qir_dprog;
    DevBuf qir_meas;
    double qir_last_prob = 0.0;
};

namespace {

svmi_status ensure_kernels()
{
    // per-process: kernel attributes (dynamic LDS limit) for the current device
    static bool ready = false;
    if (!ready) {
        SVMI_HIP(init_sweep_kernels());
        SVMI_HIP(init_sample_kernels());
        ready = true;
    }
    return SVMI_OK;
}

svmi_status use_device(svmi_handle h)
{
    if (!h) return fail(SVMI_ERR_ARG, "null handle");
    SVMI_HIP(hipSetDevice(h->device));
    return SVMI_OK;
}

// largest gradient-partial matrix kept for one reduction at the end of a
// two-state program (beyond it each sweep reduces its own partials)
constexpr size_t kCombinedPartialsBytes = size_t(256) << 20;

svmi_status exec_program(svmi_handle h, const Program& prog, DeviceProgram& dp, c64* s0, c64* s1,
                         double* grad_sums /* device, prog.n_grad */,
                         MeasureResult* meas = nullptr /* device, prog.measures.size() */,
                         const JitSweeps* jit = nullptr /* specialised kernels of prog's sweeps */)
{
    const bool prof = h->profile;
    if (prof) {
        while (h->prof_events.size() < 2 * prog.steps.size()) {
            hipEvent_t e;
            SVMI_HIP(hipEventCreate(&e));
            h->prof_events.push_back(e);
        }
    }
    // Two-state programs: when every sweep with gradient ops has the same
    // tile count (one tile size per program), all their partials go into one
    // grad-major matrix [prog.n_grad][tiles] and one k_sum_rows launch after
    // the last sweep reduces it -- the same per-row sums as a launch per sweep.
    uint64_t common_tiles = 0;
    bool combined = prog.n_states == 2 && prog.n_grad && grad_sums;
    for (const SweepInfo& sw : prog.sweeps) {
        if (!combined || !sw.n_grad) continue;
        if (!common_tiles) common_tiles = sw.n_tiles;
        combined = sw.n_tiles == common_tiles;
    }
    combined = combined && common_tiles && prog.n_grad * common_tiles * sizeof(double) <= kCombinedPartialsBytes;
    if (combined) SVMI_TRY(h->scratch[5].ensure(prog.n_grad * common_tiles * sizeof(double)));
    size_t si = 0;
    for (const Step& st : prog.steps) {
        if (prof) SVMI_HIP(hipEventRecord(h->prof_events[2 * si], h->stream));
        if (st.kind == Step::SWEEP) {
            const SweepInfo& sw = prog.sweeps[st.index];
            double* partials = nullptr;
            if (combined && sw.n_grad) {
                partials = h->scratch[5].as<double>() + sw.grad_base * common_tiles;
            } else if (prog.n_states == 2 && sw.n_grad) {
                SVMI_TRY(h->scratch[5].ensure(sw.n_tiles * sw.n_grad * sizeof(double)));
                partials = h->scratch[5].as<double>();
            }
            if (jit && st.index < jit->fn.size() && jit->fn[st.index])
                SVMI_HIP(jit_launch(h->stream, *jit, st.index, sw, s0, dp.coef.as<c64>(), dp.hitab.as<uint64_t>()));
            else
                SVMI_HIP(launch_sweep(h->stream, prog.n_states, s0, s1, sw, dp.phases.as<PhaseInfo>(),
                                      dp.ops.as<DOp>(), dp.coef.as<c64>(), dp.hitab.as<uint64_t>(), partials,
                                      h->pipeline, h->prefetch));
            if (partials && !combined)
                SVMI_HIP(launch_sum_rows(h->stream, partials, sw.n_grad, sw.n_tiles, grad_sums + sw.grad_base));
        } else if (st.kind == Step::MEASURE) {
            const MeasureStep& m = prog.measures[st.index];
            SVMI_TRY(h->scratch[4].ensure(measure_partials(int(m.n)) * sizeof(double)));
            SVMI_HIP(launch_measure(h->stream, s0, int(m.n), m.qubit, m.draw, h->scratch[4].as<double>(),
                                    dp.coef.as<c64>() + m.coef, meas + st.index, int(m.dm_n)));
        } else {
            const StandaloneOp& so = prog.standalone[st.index];
            SVMI_HIP(launch_standalone(h->stream, s0, h->n_int, so, dp.coef.as<c64>()));
            if (prog.n_states == 2) SVMI_HIP(launch_standalone(h->stream, s1, h->n_int, so, dp.coef.as<c64>()));
        }
        if (prof) SVMI_HIP(hipEventRecord(h->prof_events[2 * si + 1], h->stream));
        h->stats.kernel_launches++;
        si++;
    }
    if (combined) SVMI_HIP(launch_sum_rows(h->stream, h->scratch[5].as<double>(), prog.n_grad, common_tiles, grad_sums));
    h->stats.sweeps += prog.steps.size();
    if (prof) {
        SVMI_HIP(hipStreamSynchronize(h->stream));
        h->step_ms.assign(prog.steps.size(), 0.0);
        for (size_t i = 0; i < prog.steps.size(); i++) {
            float ms = 0.f;
            SVMI_HIP(hipEventElapsedTime(&ms, h->prof_events[2 * i], h->prof_events[2 * i + 1]));
            h->step_ms[i] = ms;
            if (prog.steps[i].kind == Step::SWEEP) {
                h->stats.sweep_launches++;
                h->stats.sweep_ms += ms;
            } else {
                h->stats.other_launches++;
                h->stats.other_ms += ms;
            }
        }
    }
    return SVMI_OK;
}

// <psi|H|psi>, optionally writing lam = H psi.
// The Pauli tables of `terms` into scratch[0..1] (groups / pt hold them on
// the host and must outlive the stream's copies: the caller synchronises).
svmi_status pauli_upload(svmi_handle h, const svmi_pauli_term* terms, size_t n_terms, std::vector<PauliGroup>& groups,
                         std::vector<PauliTerm>& pt)
{
    group_terms(terms, n_terms, groups, pt);
    SVMI_TRY(h->scratch[0].ensure(std::max<size_t>(groups.size(), 1) * sizeof(PauliGroup)));
    SVMI_TRY(h->scratch[1].ensure(std::max<size_t>(pt.size(), 1) * sizeof(PauliTerm)));
    if (!groups.empty()) {
        SVMI_HIP(hipMemcpyAsync(h->scratch[0].p, groups.data(), groups.size() * sizeof(PauliGroup),
                                hipMemcpyHostToDevice, h->stream));
        SVMI_HIP(hipMemcpyAsync(h->scratch[1].p, pt.data(), pt.size() * sizeof(PauliTerm), hipMemcpyHostToDevice,
                                h->stream));
    }
    return SVMI_OK;
}

// lam = H psi (and copy = psi if given) with the uploaded tables of n_groups
// groups; the block partials of <psi|lam> in scratch[2] (grid of them,
// returned).
svmi_status pauli_launch(svmi_handle h, const c64* psi, c64* lam, c64* copy, size_t n_groups, uint32_t* grid_out)
{
    const uint32_t grid = pauli_grid(h->dim);
    SVMI_TRY(h->scratch[2].ensure(size_t(grid) * sizeof(double)));
    SVMI_HIP(launch_pauli(h->stream, psi, lam, h->dim, h->scratch[0].as<PauliGroup>(), uint32_t(n_groups),
                          h->scratch[1].as<PauliTerm>(), h->scratch[2].as<double>(), copy));
    *grid_out = grid;
    return SVMI_OK;
}

svmi_status pauli_device(svmi_handle h, const c64* psi, c64* lam, const svmi_pauli_term* terms, size_t n_terms,
                         double* out)
{
    std::vector<PauliGroup> groups;
    std::vector<PauliTerm> pt;
    uint32_t grid = 0;
    SVMI_TRY(pauli_upload(h, terms, n_terms, groups, pt));
    SVMI_TRY(pauli_launch(h, psi, lam, nullptr, groups.size(), &grid));
    std::vector<double> part(grid);
    SVMI_HIP(hipMemcpyAsync(part.data(), h->scratch[2].p, grid * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SVMI_HIP(hipStreamSynchronize(h->stream));
    double s = 0.0;
    for (double v : part) s += v;
    *out = s;
    return SVMI_OK;
}

// Host copies in chunks (AoS device <-> SoA host).
constexpr size_t kChunk = size_t(1) << 22;  // amplitudes per staging chunk (64 MiB)

}  // namespace

extern "C" {

int svmi_abi_version(void) { return SVMI_ABI_VERSION; }

const char* svmi_last_error(void) { return g_err.c_str(); }

svmi_status svmi_create(int n_qubits, int device, svmi_handle* out)
{
    if (!out) return fail(SVMI_ERR_ARG, "null output handle");
    *out = nullptr;
    if (n_qubits < 1 || n_qubits > kMaxQubits)
        return fail(SVMI_ERR_ARG, "n_qubits must be in [1, " + std::to_string(kMaxQubits) + "]");
    int ndev = 0;
    SVMI_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(SVMI_ERR_DEVICE, "device " + std::to_string(device) + " not present (" + std::to_string(ndev) +
                                         " visible): the second argument is the HIP ordinal of the one GPU the "
                                         "engine runs on (the single-GPU engine's i_gpu), not a device or core count");
    SVMI_HIP(hipSetDevice(device));
    SVMI_TRY(ensure_kernels());
    svmi_sim* h = new svmi_sim();
    h->n = n_qubits;
    h->n_int = std::max(n_qubits, kRegBits);
    h->device = device;
    h->dim = 1ull << h->n_int;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e == hipSuccess && static_scratch_ptr(device, &h->static_scratch) != SVMI_OK) {
        const std::string msg = g_err;
        svmi_destroy(h);
        return fail(SVMI_ERR_DEVICE, "create: " + msg);
    }
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&h->sv), h->dim * sizeof(c64));
    if (e == hipSuccess) e = launch_fill_basis(h->stream, h->sv, h->dim);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        const std::string msg = std::string("create: ") + hipGetErrorString(e);
        svmi_destroy(h);
        return fail(e == hipErrorOutOfMemory ? SVMI_ERR_OOM : SVMI_ERR_DEVICE, msg);
    }
    *out = h;
    return SVMI_OK;
}

svmi_status svmi_destroy(svmi_handle h)
{
    if (!h) return SVMI_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->sv) (void)hipFree(h->sv);
    for (auto& b : h->aux) b.release();
    for (auto& b : h->scratch) b.release();
    for (hipEvent_t e : h->prof_events) (void)hipEventDestroy(e);
    h->dprog.release();
    h->jit.release();
    h->adj_dprog.release();
    h->adj_gsum.release();
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SVMI_OK;
}

int svmi_num_qubits(svmi_handle h) { return h ? h->n : -1; }

svmi_status svmi_reset_state(svmi_handle h)
{
    SVMI_TRY(use_device(h));
    h->qir_last_prob = 0.0;   // reset() also clears m_real (svsim_qsharp/svsim_cpu_omp.hpp:430-438)
    SVMI_HIP(launch_fill_basis(h->stream, h->sv, h->dim));
    SVMI_HIP(hipStreamSynchronize(h->stream));
    return SVMI_OK;
}

svmi_status svmi_set_random_state(svmi_handle h, uint64_t seed)
{
    SVMI_TRY(use_device(h));
    const uint64_t dim_user = 1ull << h->n;
    if (h->n_int != h->n) SVMI_HIP(launch_fill_basis(h->stream, h->sv, h->dim));
    const uint32_t grid = random_state_grid(dim_user);
    SVMI_TRY(h->scratch[2].ensure(size_t(grid) * sizeof(double)));
    SVMI_HIP(launch_random_state(h->stream, h->sv, dim_user, seed, h->scratch[2].as<double>()));
    std::vector<double> part(grid);
    SVMI_HIP(hipMemcpyAsync(part.data(), h->scratch[2].p, grid * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SVMI_HIP(hipStreamSynchronize(h->stream));
    double norm2 = 0.0;
    for (double v : part) norm2 += v;
    SVMI_HIP(launch_scale(h->stream, h->sv, dim_user, 1.0 / std::sqrt(norm2)));
    SVMI_HIP(hipStreamSynchronize(h->stream));
    return SVMI_OK;
}

svmi_status svmi_set_state(svmi_handle h, const double* re, const double* im)
{
    SVMI_TRY(use_device(h));
    if (!re || !im) return fail(SVMI_ERR_ARG, "null state array");
    const uint64_t dim_user = 1ull << h->n;
    if (h->n_int != h->n) {
        SVMI_HIP(launch_fill_basis(h->stream, h->sv, h->dim));
    }
    c64* stage = nullptr;
    const size_t cs = std::min<uint64_t>(kChunk, dim_user);
    SVMI_HIP(hipHostMalloc(reinterpret_cast<void**>(&stage), cs * sizeof(c64), hipHostMallocDefault));
    svmi_status st = SVMI_OK;
    for (uint64_t x0 = 0; x0 < dim_user && st == SVMI_OK; x0 += cs) {
        const size_t cnt = std::min<uint64_t>(cs, dim_user - x0);
        hipError_t e = hipStreamSynchronize(h->stream);
        for (size_t i = 0; i < cnt; i++) stage[i] = c64{re[x0 + i], im[x0 + i]};
        if (e == hipSuccess)
            e = hipMemcpyAsync(h->sv + x0, stage, cnt * sizeof(c64), hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) st = fail(SVMI_ERR_DEVICE, std::string("set_state: ") + hipGetErrorString(e));
    }
    (void)hipHostFree(stage);
    return st;
}

svmi_status svmi_get_state(svmi_handle h, double* re, double* im)
{
    SVMI_TRY(use_device(h));
    if (!re || !im) return fail(SVMI_ERR_ARG, "null state array");
    const uint64_t dim_user = 1ull << h->n;
    c64* stage = nullptr;
    const size_t cs = std::min<uint64_t>(kChunk, dim_user);
    SVMI_HIP(hipHostMalloc(reinterpret_cast<void**>(&stage), cs * sizeof(c64), hipHostMallocDefault));
    svmi_status st = SVMI_OK;
    for (uint64_t x0 = 0; x0 < dim_user && st == SVMI_OK; x0 += cs) {
        const size_t cnt = std::min<uint64_t>(cs, dim_user - x0);
        hipError_t e = hipMemcpyAsync(stage, h->sv + x0, cnt * sizeof(c64), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) {
            st = fail(SVMI_ERR_DEVICE, std::string("get_state: ") + hipGetErrorString(e));
            break;
        }
        for (size_t i = 0; i < cnt; i++) {
            re[x0 + i] = stage[i].re;
            im[x0 + i] = stage[i].im;
        }
    }
    (void)hipHostFree(stage);
    return st;
}

svmi_status svmi_get_amplitudes(svmi_handle h, const uint64_t* idx, size_t count, double* re, double* im)
{
    SVMI_TRY(use_device(h));
    if (count && (!idx || !re || !im)) return fail(SVMI_ERR_ARG, "null argument");
    const uint64_t dim_user = 1ull << h->n;
    for (size_t i = 0; i < count; i++)
        if (idx[i] >= dim_user) return fail(SVMI_ERR_ARG, "amplitude index out of range");
    if (count == 0) return SVMI_OK;
    SVMI_TRY(h->scratch[0].ensure(count * sizeof(uint64_t)));
    SVMI_TRY(h->scratch[4].ensure(count * sizeof(c64)));
    SVMI_HIP(hipMemcpyAsync(h->scratch[0].p, idx, count * sizeof(uint64_t), hipMemcpyHostToDevice, h->stream));
    SVMI_HIP(launch_gather(h->stream, h->sv, h->dim, h->scratch[0].as<uint64_t>(), count, h->scratch[4].as<c64>()));
    std::vector<c64> out(count);
    SVMI_HIP(hipMemcpyAsync(out.data(), h->scratch[4].p, count * sizeof(c64), hipMemcpyDeviceToHost, h->stream));
    SVMI_HIP(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < count; i++) {
        re[i] = out[i].re;
        im[i] = out[i].im;
    }
    return SVMI_OK;
}

// Planner options of the forward program.  Specialised sweeps pay almost
// nothing per op (no dispatch), so unless the caller set an op cost their
// fusion model counts FMAs only: 20-layer u3 brick-wall at 30 qubits 300 ->
// 258 ms per circuit; the interpreted sweeps are fastest at the default 2.
PlanOptions forward_options(svmi_handle h)
{
    PlanOptions o = h->opt;
    if (h->specialize && !h->op_cost_set) o.op_overhead = 0;
    if (h->specialize) o.tile_map = false;   // the compiled sweeps address LDS without maps
    return o;
}

// Device copy of h->prog and, with SVMI_OPT_SPECIALIZE, its compiled sweeps.
svmi_status install_program(svmi_handle h)
{
    SVMI_TRY(h->dprog.upload(h->prog, h->stream));
    h->jit.release();
    h->jit_error.clear();
    h->stats.specialized_sweeps = 0;
    h->stats.compile_ms = 0.0;
    if (!h->specialize) return SVMI_OK;
    uint64_t dma_grid = 0;   // SVMI_OPT_PREFETCH: double-buffered kernels, one workgroup per CU
    if (h->prefetch) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess && cus > 0)
            dma_grid = uint64_t(cus);
    }
    if (!jit_build(h->prog, device_arch().c_str(), h->jit, h->jit_error, dma_grid)) {
        h->jit.release();
        // the sweeps are interpreted; the reason stays readable
        if (!h->jit_error.empty()) g_err = "specialisation: " + h->jit_error;
        return SVMI_OK;
    }
    for (hipFunction_t f : h->jit.fn) h->stats.specialized_sweeps += f != nullptr;
    h->stats.compile_ms = h->jit.compile_ms;
    return SVMI_OK;
}

svmi_status svmi_upload(svmi_handle h, const svmi_gate* gates, size_t n_gates)
{
    SVMI_TRY(use_device(h));
    SVMI_TRY(check_gates(gates, n_gates, h->n));
    h->gates.assign(gates, gates + n_gates);
    h->adj_valid = false;
    plan_forward(gates, n_gates, h->n_int, forward_options(h), h->prog);
    SVMI_TRY(install_program(h));
    h->uploaded = true;
    h->stats.gates = n_gates;
    h->stats.ops = h->prog.n_lowered;
    return SVMI_OK;
}

svmi_status svmi_clear_circuit(svmi_handle h)
{
    if (!h) return fail(SVMI_ERR_ARG, "null handle");
    h->gates.clear();
    h->prog = Program();
    h->uploaded = false;
    h->adj_valid = false;
    return SVMI_OK;
}

svmi_status svmi_run(svmi_handle h)
{
    SVMI_TRY(use_device(h));
    if (!h->uploaded) return fail(SVMI_ERR_STATE, "run() before upload()");
    h->stats.sweeps = 0;
    h->stats.kernel_launches = 0;
    h->stats.sweep_launches = h->stats.other_launches = 0;
    h->stats.sweep_ms = h->stats.other_ms = 0.0;
    SVMI_HIP(hipEventRecord(h->ev0, h->stream));
    SVMI_TRY(exec_program(h, h->prog, h->dprog, h->sv, nullptr, nullptr, nullptr, &h->jit));
    SVMI_HIP(hipEventRecord(h->ev1, h->stream));
    SVMI_HIP(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    SVMI_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->stats.device_ms = ms;
    return SVMI_OK;
}

svmi_status svmi_qasm_parse(const char* text, size_t len, svmi_gate* gates, size_t capacity, svmi_qasm_info* info)
{
    if (!text && len) return fail(SVMI_ERR_ARG, "null text");
    if (!info) return fail(SVMI_ERR_ARG, "null info");
    QasmProgram prog;
    std::string err;
    const svmi_status st = parse_qasm(text, len, prog, err);
    if (st != SVMI_OK) return fail(st, "qasm: " + err);
    info->n_qubits = prog.n_qubits;
    info->reserved = 0;
    info->n_gates = prog.gates.size();
    info->basic_gates = prog.basic_gates;
    info->cx_gates = prog.cx_gates;
    info->skipped = prog.skipped;
    if (gates) {
        if (capacity < prog.gates.size())
            return fail(SVMI_ERR_ARG, "qasm: " + std::to_string(prog.gates.size()) + " gates do not fit in " +
                                          std::to_string(capacity));
        std::copy(prog.gates.begin(), prog.gates.end(), gates);
    }
    return SVMI_OK;
}

svmi_status qir_run(svmi_handle h, int n_qubits, const svmi_qir_gate* gates, size_t n_gates, double* probs,
                    int32_t* outcomes, double* last_prob, bool dm)
{
    SVMI_TRY(use_device(h));
    const int n_state = dm ? 2 * n_qubits : n_qubits;
    if (n_qubits < 1 || n_state > h->n)
        return fail(SVMI_ERR_ARG, dm ? "density matrix of n_qubits needs a handle of 2 * n_qubits qubits (has " +
                                           std::to_string(h->n) + ")"
                                     : "n_qubits must be in [1, " + std::to_string(h->n) + "] (the handle's qubits)");
    SVMI_TRY(plan_qir(gates, n_gates, n_qubits, std::max(n_state, kRegBits), h->opt, h->qir_prog, dm));
    const Program& prog = h->qir_prog;
    SVMI_TRY(h->qir_dprog.upload(prog, h->stream));
    const size_t nm = prog.measures.size();
    SVMI_TRY(h->qir_meas.ensure(std::max<size_t>(nm, 1) * sizeof(MeasureResult)));
    h->stats.sweeps = 0;
    h->stats.kernel_launches = 0;
    h->stats.sweep_launches = h->stats.other_launches = 0;
    h->stats.sweep_ms = h->stats.other_ms = 0.0;
    h->stats.gates = n_gates;
    h->stats.ops = prog.n_lowered;
    SVMI_HIP(hipEventRecord(h->ev0, h->stream));
    SVMI_TRY(exec_program(h, prog, h->qir_dprog, h->sv, nullptr, nullptr, h->qir_meas.as<MeasureResult>()));
    SVMI_HIP(hipEventRecord(h->ev1, h->stream));
    std::vector<MeasureResult> res(nm);
    if (nm)
        SVMI_HIP(hipMemcpyAsync(res.data(), h->qir_meas.p, nm * sizeof(MeasureResult), hipMemcpyDeviceToHost,
                                h->stream));
    SVMI_HIP(hipStreamSynchronize(h->stream));
    float ms = 0.f;
    SVMI_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->stats.device_ms = ms;
    for (size_t m = 0; m < nm; m++) {
        if (probs) probs[m] = res[m].prob;
        if (outcomes) outcomes[m] = res[m].outcome;
    }
    if (nm) h->qir_last_prob = res[nm - 1].prob;
    if (last_prob) *last_prob = h->qir_last_prob;
    return SVMI_OK;
}

svmi_status svmi_qir_run(svmi_handle h, int n_qubits, const svmi_qir_gate* gates, size_t n_gates, double* probs,
                         int32_t* outcomes, double* last_prob)
{
    return qir_run(h, n_qubits, gates, n_gates, probs, outcomes, last_prob, false);
}

svmi_status svmi_dm_qir_run(svmi_handle h, int n_qubits, const svmi_qir_gate* gates, size_t n_gates, double* probs,
                            int32_t* outcomes)
{
    return qir_run(h, n_qubits, gates, n_gates, probs, outcomes, nullptr, true);
}

svmi_status svmi_measure(svmi_handle h, const double* uniforms, size_t shots, uint64_t* out, double* total_prob)
{
    SVMI_TRY(use_device(h));
    if (shots && (!uniforms || !out)) return fail(SVMI_ERR_ARG, "null draws/output");
    // blocks of 2^12 amplitudes: the staged probabilities of one (35 KiB of
    // LDS) leave room for four per CU, so loads overlap the scans
    // (2^14-amplitude blocks fit one per CU: 2.1 TB/s at 30 qubits)
    const int s = std::min(h->n_int, 12);
    const uint64_t bsize = 1ull << s;
    const uint32_t nthr = uint32_t(std::min<uint64_t>(256, bsize));
    const uint32_t chunk = uint32_t(bsize / nthr);
    const uint64_t nb = h->dim >> s;
    SVMI_TRY(h->scratch[3].ensure(nb * sizeof(double)));
    SVMI_HIP(launch_block_totals(h->stream, h->sv, nb, nthr, chunk, h->scratch[3].as<double>()));
    std::vector<double> tot(nb);
    SVMI_HIP(hipMemcpyAsync(tot.data(), h->scratch[3].p, nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SVMI_HIP(hipStreamSynchronize(h->stream));
    std::vector<double> O(nb + 1);
    O[0] = 0.0;
    for (uint64_t b = 0; b < nb; b++) O[b + 1] = O[b] + tot[b];
    if (total_prob) *total_prob = O[nb];
    if (shots == 0) return SVMI_OK;
    // block of each shot (-1: no interval matches -> result 0)
    std::vector<int64_t> blk(shots);
    for (size_t i = 0; i < shots; i++) {
        const double r = uniforms[i];
        if (!(r >= O[0]) || !(r < O[nb])) {
            blk[i] = -1;
            continue;
        }
        const uint64_t m = uint64_t(std::upper_bound(O.begin(), O.end(), r) - O.begin());  // first O[m] > r
        blk[i] = int64_t(m) - 1;
    }
    std::vector<size_t> order;
    order.reserve(shots);
    for (size_t i = 0; i < shots; i++)
        if (blk[i] >= 0) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return blk[a] < blk[b]; });
    std::vector<uint64_t> jobs;
    std::vector<double> offs, draws(order.size());
    for (size_t j = 0; j < order.size(); j++) {
        draws[j] = uniforms[order[j]];
        const uint64_t b = uint64_t(blk[order[j]]);
        if (jobs.empty() || jobs[jobs.size() - 3] != b) {
            jobs.push_back(b);
            jobs.push_back(j);
            jobs.push_back(0);
            offs.push_back(O[b]);
        }
        jobs.back()++;
    }
    const uint32_t njobs = uint32_t(offs.size());
    std::vector<uint64_t> res(order.size());
    if (njobs) {
        SVMI_TRY(h->scratch[0].ensure(jobs.size() * sizeof(uint64_t)));
        SVMI_TRY(h->scratch[1].ensure(offs.size() * sizeof(double)));
        SVMI_TRY(h->scratch[2].ensure(draws.size() * sizeof(double)));
        SVMI_TRY(h->scratch[4].ensure(res.size() * sizeof(uint64_t)));
        SVMI_HIP(hipMemcpyAsync(h->scratch[0].p, jobs.data(), jobs.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                                h->stream));
        SVMI_HIP(hipMemcpyAsync(h->scratch[1].p, offs.data(), offs.size() * sizeof(double), hipMemcpyHostToDevice,
                                h->stream));
        SVMI_HIP(hipMemcpyAsync(h->scratch[2].p, draws.data(), draws.size() * sizeof(double), hipMemcpyHostToDevice,
                                h->stream));
        SVMI_HIP(launch_sample(h->stream, h->sv, nthr, chunk, h->scratch[0].as<uint64_t>(), njobs,
                               h->scratch[1].as<double>(), h->scratch[2].as<double>(), h->scratch[4].as<uint64_t>()));
        SVMI_HIP(hipMemcpyAsync(res.data(), h->scratch[4].p, res.size() * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                h->stream));
        SVMI_HIP(hipStreamSynchronize(h->stream));
    }
    for (size_t i = 0; i < shots; i++) out[i] = 0;
    for (size_t j = 0; j < order.size(); j++) out[order[j]] = res[j];
    return SVMI_OK;
}

svmi_status svmi_expval(svmi_handle h, const svmi_pauli_term* terms, size_t n_terms, double* out)
{
    SVMI_TRY(use_device(h));
    if (!out) return fail(SVMI_ERR_ARG, "null output");
    SVMI_TRY(check_terms(terms, n_terms, h->n));
    return pauli_device(h, h->sv, nullptr, terms, n_terms, out);
}

svmi_status svmi_expval_static(svmi_handle h, const svmi_pauli_term* terms, size_t n_terms, double* out)
{
    SVMI_TRY(use_device(h));
    if (!out) return fail(SVMI_ERR_ARG, "null output");
    SVMI_TRY(check_terms(terms, n_terms, h->n));
    if (n_terms > kStaticMaxTerms)
        return fail(SVMI_ERR_ARG, "svmi_expval_static takes at most " + std::to_string(kStaticMaxTerms) +
                                      " Pauli terms (got " + std::to_string(n_terms) + "); use svmi_expval");
    std::vector<PauliGroup> groups;
    std::vector<PauliTerm> pt;
    group_terms(terms, n_terms, groups, pt);
    // scratch layout: groups | terms | block partials, each 256-B aligned
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const uint32_t grid = pauli_grid(h->dim);
    const size_t og = 0, ot = up(groups.size() * sizeof(PauliGroup)), op = ot + up(pt.size() * sizeof(PauliTerm));
    if (op + size_t(grid) * sizeof(double) > kStaticScratchBytes)
        return fail(SVMI_ERR_ARG, "svmi_expval_static: tables and partials exceed the static scratch");
    unsigned char* base = static_cast<unsigned char*>(h->static_scratch);
    PauliGroup* dg = reinterpret_cast<PauliGroup*>(base + og);
    PauliTerm* dt = reinterpret_cast<PauliTerm*>(base + ot);
    double* dp = reinterpret_cast<double*>(base + op);
    if (!groups.empty()) {
        SVMI_HIP(hipMemcpyAsync(dg, groups.data(), groups.size() * sizeof(PauliGroup), hipMemcpyHostToDevice,
                                h->stream));
        SVMI_HIP(hipMemcpyAsync(dt, pt.data(), pt.size() * sizeof(PauliTerm), hipMemcpyHostToDevice, h->stream));
    }
    SVMI_HIP(launch_pauli(h->stream, h->sv, nullptr, h->dim, dg, uint32_t(groups.size()), dt, dp));
    std::vector<double> part(grid);
    SVMI_HIP(hipMemcpyAsync(part.data(), dp, grid * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SVMI_HIP(hipStreamSynchronize(h->stream));
    double sum = 0.0;
    for (double v : part) sum += v;
    *out = sum;
    return SVMI_OK;
}

svmi_status svmi_adjoint(svmi_handle h, const svmi_pauli_term* terms, size_t n_terms, const int32_t* param_slots,
                         size_t n_params, double* expval_out, double* grad_out)
{
    SVMI_TRY(use_device(h));
    if (!h->uploaded) return fail(SVMI_ERR_STATE, "adjoint() before upload()");
    if (!expval_out || (n_params && !grad_out)) return fail(SVMI_ERR_ARG, "null output");
    SVMI_TRY(check_terms(terms, n_terms, h->n));
    SVMI_TRY(check_params(h->gates.data(), h->gates.size(), param_slots, n_params));
    // the device work of the whole gradient is queued without host
    // synchronisation; one copy-back and one sync at the end
    SVMI_TRY(h->aux[0].ensure(h->dim * sizeof(c64)));
    SVMI_TRY(h->aux[1].ensure(h->dim * sizeof(c64)));
    // the observable's tables go up while the stream is idle
    std::vector<PauliGroup> groups;
    std::vector<PauliTerm> pt;
    SVMI_TRY(pauli_upload(h, terms, n_terms, groups, pt));
    // forward
    SVMI_TRY(exec_program(h, h->prog, h->dprog, h->sv, nullptr, nullptr, nullptr, &h->jit));
    // phi = psi and lam = H psi in one pass, partials of E = <psi|lam>
    c64* phi = h->aux[0].as<c64>();
    c64* lam = h->aux[1].as<c64>();
    uint32_t grid = 0;
    SVMI_TRY(pauli_launch(h, h->sv, lam, phi, groups.size(), &grid));
    for (size_t p = 0; p < n_params; p++) grad_out[p] = 0.0;
    // backward (planned once per circuit and parameter map)
    const size_t n_slots = 3 * h->gates.size();
    if (!h->adj_valid || h->adj_n_params != n_params ||
        !std::equal(param_slots, param_slots + n_slots, h->adj_slots.begin(), h->adj_slots.end())) {
        h->adj_valid = false;
        std::vector<AdjEntry> entries;
        adjoint_entries(h->gates.data(), h->gates.size(), param_slots, entries);
        h->adj_prog = Program();
        build_adjoint_program(entries, h->n_int, h->opt, h->adj_prog);
        SVMI_TRY(h->adj_dprog.upload(h->adj_prog, h->stream));
        SVMI_TRY(h->adj_gsum.ensure(std::max<size_t>(h->adj_prog.n_grad, 1) * sizeof(double)));
        h->adj_slots.assign(param_slots, param_slots + n_slots);
        h->adj_n_params = n_params;
        h->adj_valid = true;
    }
    const Program& bp = h->adj_prog;
    SVMI_TRY(exec_program(h, bp, h->adj_dprog, phi, lam, h->adj_gsum.as<double>()));
    std::vector<double> g(bp.n_grad), part(grid);
    if (bp.n_grad)
        SVMI_HIP(hipMemcpyAsync(g.data(), h->adj_gsum.p, bp.n_grad * sizeof(double), hipMemcpyDeviceToHost,
                                h->stream));
    SVMI_HIP(hipMemcpyAsync(part.data(), h->scratch[2].p, grid * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SVMI_HIP(hipStreamSynchronize(h->stream));
    double E = 0.0;
    for (double v : part) E += v;
    *expval_out = E;
    for (uint32_t i = 0; i < bp.n_grad; i++) grad_out[bp.grad_param[i]] += 2.0 * g[i];
    return SVMI_OK;
}

// Batched expectation + adjoint gradients: every circuit is planned on the
// host (in parallel), the programs are concatenated into one device program
// and one launch per dependency level runs the tasks of every circuit at
// that level (no host synchronisation between the levels).  The circuits go
// in chunks, costliest first (gates x 2^n): while the GPU runs chunk c the
// host plans chunk c + 1, so planning hides behind the large circuits' device
// time instead of preceding it.
svmi_status svmi_batch_adjoint(int device, const svmi_batch_item* items, size_t n_items, double* expvals,
                               double* grads)
{
    if (n_items == 0) return SVMI_OK;
    if (!items || !expvals) return fail(SVMI_ERR_ARG, "null batch arrays");
    int ndev = 0;
    SVMI_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(SVMI_ERR_DEVICE, "device " + std::to_string(device) + " not present");
    SVMI_HIP(hipSetDevice(device));
    SVMI_TRY(ensure_kernels());

    // chunks by cumulative cost: the first quarter of the work, up to 90 %,
    // the rest (many small circuits: mostly planning, little device time)
    std::vector<size_t> order(n_items);
    std::iota(order.begin(), order.end(), 0);
    std::vector<double> cost(n_items);
    double total = 0.0;
    for (size_t i = 0; i < n_items; i++) {
        const int n = std::max(int(items[i].n_qubits), kRegBits);
        cost[i] = double(items[i].n_gates + 1) * std::ldexp(1.0, std::min(n, 60));
        total += cost[i];
    }
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<size_t> bounds{0};
    {
        const double cut[] = {0.25, 0.9};
        double acc = 0.0;
        int ci = 0;
        for (size_t j = 0; j < n_items; j++) {
            acc += cost[order[j]];
            if (ci < 2 && acc >= cut[ci] * total && j + 1 < n_items) {
                bounds.push_back(j + 1);
                while (ci < 2 && acc >= cut[ci] * total) ci++;
            }
        }
        bounds.push_back(n_items);
    }
    const size_t n_chunks = bounds.size() - 1;
    // parameter offset of each item in grads, state offset in the one state buffer
    std::vector<size_t> goff(n_items + 1, 0);
    std::vector<uint64_t> soff(n_items + 1, 0);
    for (size_t i = 0; i < n_items; i++) {
        goff[i + 1] = goff[i] + items[i].n_params;
        const uint64_t dim = 1ull << std::max(int(items[i].n_qubits), kRegBits);
        soff[i + 1] = soff[i] + (((2 * dim * sizeof(c64)) + 255) & ~uint64_t(255));
    }
    if (goff[n_items] && !grads) return fail(SVMI_ERR_ARG, "null gradient array");

    struct Chunk {
        std::vector<size_t> idx;                  // original item indices
        std::vector<svmi_batch_item> items;
        BatchPlan bp;
        DevBuf dsteps, dcircs, dsweeps, dphases, dops, dcoef, dhitab, dgroups, dterms, dpart, dsegs, dout;
        void release()
        {
            for (DevBuf* b : {&dsteps, &dcircs, &dsweeps, &dphases, &dops, &dcoef, &dhitab, &dgroups, &dterms, &dpart,
                              &dsegs, &dout})
                b->release();
        }
    };
    std::vector<Chunk> ch(n_chunks);
    auto plan_chunk = [&](size_t c) -> svmi_status {
        Chunk& k = ch[c];
        for (size_t j = bounds[c]; j < bounds[c + 1]; j++) {
            k.idx.push_back(order[j]);
            k.items.push_back(items[order[j]]);
        }
        std::vector<double> dummy(1);
        return build_batch(k.items.data(), k.items.size(), dummy.data(), dummy.data(), k.bp);
    };

    DevBuf dstate;
    hipStream_t st = nullptr;
    svmi_status rc = SVMI_OK;
    auto put = [&](DevBuf& b, const void* src, size_t bytes) -> svmi_status {
        SVMI_TRY(b.ensure(std::max<size_t>(bytes, 16)));
        if (bytes) SVMI_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
        return SVMI_OK;
    };
    auto launch_chunk = [&](size_t c) -> svmi_status {
        Chunk& k = ch[c];
        BatchPlan& bp = k.bp;
        std::vector<BatchCircuit> circs = bp.circs;
        for (size_t j = 0; j < circs.size(); j++) {
            circs[j].phi = reinterpret_cast<c64*>(static_cast<char*>(dstate.p) + soff[k.idx[j]]);
            circs[j].lam = circs[j].phi + (1ull << circs[j].n);
        }
        for (size_t j : bp.need_fill)
            SVMI_HIP(launch_fill_basis(st, circs[j].phi, 1ull << circs[j].n));
        SVMI_TRY(put(k.dsteps, bp.steps.data(), bp.steps.size() * sizeof(BatchStep)));
        SVMI_TRY(put(k.dcircs, circs.data(), circs.size() * sizeof(BatchCircuit)));
        SVMI_TRY(put(k.dsweeps, bp.sweeps.data(), bp.sweeps.size() * sizeof(SweepInfo)));
        SVMI_TRY(put(k.dphases, bp.phases.data(), bp.phases.size() * sizeof(PhaseInfo)));
        bp.ops.push_back(DOp{});   // a NOP past the end: the kernels read one header ahead
        SVMI_TRY(put(k.dops, bp.ops.data(), bp.ops.size() * sizeof(DOp)));
        SVMI_TRY(put(k.dcoef, bp.coef.data(), bp.coef.size() * sizeof(c64)));
        SVMI_TRY(put(k.dhitab, bp.hitab.data(), bp.hitab.size() * sizeof(uint64_t)));
        SVMI_TRY(put(k.dgroups, bp.groups.data(), bp.groups.size() * sizeof(PauliGroup)));
        SVMI_TRY(put(k.dterms, bp.terms.data(), bp.terms.size() * sizeof(PauliTerm)));
        SVMI_TRY(k.dpart.ensure(std::max<size_t>((bp.n_grad_partials + bp.n_e_partials) * sizeof(double), 16)));
        SVMI_TRY(put(k.dsegs, bp.segs.data(), bp.segs.size() * sizeof(SumSegment)));
        SVMI_TRY(k.dout.ensure((bp.n_slots + k.items.size()) * sizeof(double)));
        BatchArgs a;
        a.steps = k.dsteps.as<BatchStep>();
        a.circuits = k.dcircs.as<BatchCircuit>();
        a.sweeps = k.dsweeps.as<SweepInfo>();
        a.phases = k.dphases.as<PhaseInfo>();
        a.ops = k.dops.as<DOp>();
        a.coef = k.dcoef.as<c64>();
        a.hitab = k.dhitab.as<uint64_t>();
        a.groups = k.dgroups.as<PauliGroup>();
        a.terms = k.dterms.as<PauliTerm>();
        a.grad_partials = k.dpart.as<double>();
        a.e_partials = k.dpart.as<double>() + bp.n_grad_partials;
        for (const BatchLaunch& L : bp.launches) SVMI_HIP(launch_batch(st, a, L));
        SVMI_HIP(launch_sum_segments(st, k.dpart.as<double>(), k.dsegs.as<SumSegment>(), bp.segs.size(),
                                     k.dout.as<double>()));
        return SVMI_OK;
    };
    do {
        if ((rc = plan_chunk(0)) != SVMI_OK) break;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            rc = fail(SVMI_ERR_DEVICE, "batch: stream creation failed");
            break;
        }
        if ((rc = dstate.ensure(std::max<uint64_t>(soff[n_items], 16))) != SVMI_OK) break;
        for (size_t c = 0; c < n_chunks && rc == SVMI_OK; c++) {
            rc = launch_chunk(c);
            // the next chunk's plans while this one runs
            if (rc == SVMI_OK && c + 1 < n_chunks) rc = plan_chunk(c + 1);
        }
        if (rc != SVMI_OK) break;
        const hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            rc = fail(SVMI_ERR_DEVICE, std::string("batch launch: ") + hipGetErrorString(e));
            break;
        }
        for (size_t c = 0; c < n_chunks && rc == SVMI_OK; c++) {
            Chunk& k = ch[c];
            const size_t m = k.items.size();
            std::vector<double> res(k.bp.n_slots + m);
            const hipError_t e2 = hipMemcpy(res.data(), k.dout.p, res.size() * sizeof(double), hipMemcpyDeviceToHost);
            if (e2 != hipSuccess) {
                rc = fail(SVMI_ERR_DEVICE, std::string("batch results: ") + hipGetErrorString(e2));
                break;
            }
            for (size_t j = 0; j < m; j++) {
                const size_t i = k.idx[j];
                expvals[i] = res[k.bp.n_slots + j];
                double* g = grads ? grads + goff[i] : nullptr;
                for (size_t p = 0; p < items[i].n_params; p++) g[p] = 0.0;
                const Program& ad = k.bp.adj[j];
                for (uint32_t s2 = 0; s2 < ad.n_grad; s2++)
                    g[ad.grad_param[s2]] += 2.0 * res[k.bp.circs[j].slot_begin + s2];
            }
        }
    } while (false);
    if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (Chunk& k : ch) k.release();
    dstate.release();
    return rc;
}

svmi_status svmi_set_option(svmi_handle h, int option, int64_t value)
{
    if (!h) return fail(SVMI_ERR_ARG, "null handle");
    switch (option) {
    case SVMI_OPT_FUSION: h->opt.fusion = value != 0; break;
    case SVMI_OPT_TILE_QUBITS:
        if (value < kRegBits || value > kMaxTileQubits) return fail(SVMI_ERR_ARG, "tile qubits must be in [4, 13]");
        h->opt.tile_qubits = int(value);
        break;
    case SVMI_OPT_LOW_QUBITS:
        if (value < 0 || value > 8) return fail(SVMI_ERR_ARG, "low qubits must be in [0, 8]");
        h->opt.low_qubits = int(value);
        break;
    case SVMI_OPT_SWEEP_FLOPS:
        if (value < 0) return fail(SVMI_ERR_ARG, "sweep budget must be >= 0");
        h->opt.sweep_budget = long(value);
        break;
    case SVMI_OPT_OP_COST:
        if (value < 0 || value > 1024) return fail(SVMI_ERR_ARG, "op cost must be in [0, 1024]");
        h->opt.op_overhead = long(value);
        h->op_cost_set = true;
        break;
    case SVMI_OPT_COEF_BUDGET:
        if (value < 0) return fail(SVMI_ERR_ARG, "coefficient budget must be >= 0");
        h->opt.coef_budget = long(value);
        break;
    case SVMI_OPT_PROFILE: h->profile = value != 0; return SVMI_OK;
    case SVMI_OPT_PIPELINE:
        if (value < 0 || value > 64) return fail(SVMI_ERR_ARG, "pipeline must be in [0, 64]");
        h->pipeline = int(value);
        return SVMI_OK;
    case SVMI_OPT_PREFETCH: h->prefetch = value != 0; return SVMI_OK;
    case SVMI_OPT_SPECIALIZE: h->specialize = value != 0; break;
    case SVMI_OPT_MERGE_QUBITS:
        if (value < 1 || value > kRegBits) return fail(SVMI_ERR_ARG, "merge qubits must be in [1, 4]");
        h->opt.merge_qubits = int(value);
        break;
    case SVMI_OPT_TILE_MAP: h->opt.tile_map = value != 0; break;
    default: return fail(SVMI_ERR_ARG, "unknown option " + std::to_string(option));
    }
    h->adj_valid = false;
    if (h->uploaded) {
        plan_forward(h->gates.data(), h->gates.size(), h->n_int, forward_options(h), h->prog);
        SVMI_TRY(use_device(h));
        SVMI_TRY(install_program(h));
        h->stats.ops = h->prog.n_lowered;
    }
    return SVMI_OK;
}

svmi_status svmi_get_stats(svmi_handle h, svmi_stats* out)
{
    if (!h || !out) return fail(SVMI_ERR_ARG, "null argument");
    *out = h->stats;
    return SVMI_OK;
}

}  // extern "C"

// =====================================================================
// Host execution of device programs (include/svmi_debug.h).
// =====================================================================

namespace {

void host_standalone(std::vector<c64>& s, int n, const StandaloneOp& so, const std::vector<c64>& coef)
{
    const uint64_t dim = 1ull << n;
    if (so.kind == 2) {
        const int q = so.qubits[0];
        for (uint64_t i = 0; i < dim / 2; i++) {
            const uint64_t p0 = ((i >> q) << (q + 1)) | (i & ((1ull << q) - 1));
            const uint64_t p1 = p0 | (1ull << q);
            const c64 a0 = s[p0], a1 = s[p1];
            s[p0] = c64{0.5 * (a0.re + a1.re), 0.5 * (a0.im - a1.im)};
            s[p1] = c64{0.5 * (a0.re + a1.re), 0.5 * (-a0.im + a1.im)};
        }
        return;
    }
    if (so.kind == 1) {
        for (uint64_t x = 0; x < dim; x++) {
            uint32_t idx = 0;
            for (int i = 0; i < so.q; i++) idx |= uint32_t((x >> so.qubits[i]) & 1) << i;
            s[x] = cmul(s[x], coef[so.coef + idx]);
        }
        return;
    }
    const uint32_t D = 1u << so.q;
    uint64_t qm = 0;
    for (int i = 0; i < so.q; i++) qm |= 1ull << so.qubits[i];
    std::vector<c64> v(D);
    for (uint64_t b = 0; b < dim; b++) {
        if (b & qm) continue;
        for (uint32_t j = 0; j < D; j++) {
            uint64_t o = b;
            for (int i = 0; i < so.q; i++)
                if ((j >> i) & 1) o |= 1ull << so.qubits[i];
            v[j] = s[o];
        }
        for (uint32_t r = 0; r < D; r++) {
            c64 acc{0.0, 0.0};
            for (uint32_t j = 0; j < D; j++) acc = cmac(acc, coef[so.coef + r * D + j], v[j]);
            uint64_t o = b;
            for (int i = 0; i < so.q; i++)
                if ((r >> i) & 1) o |= 1ull << so.qubits[i];
            s[o] = acc;
        }
    }
}

// One tile of one sweep on the host (the kernels' sweep_tile without LDS).
// gsum_tile receives the tile's gradient accumulators (ns == 2).
void host_tile(const SweepInfo& sw, const PhaseInfo* phases, const DOp* ops, const c64* coef, const uint64_t* hitab,
               int ns, c64* s0, c64* s1, uint64_t tile, bool init_zero, double* gsum_tile)
{
    // as the kernels: one LDS-like array, lambda's half at [2^k, 2^(k+1));
    // a group is 16 amplitudes (two states: 8 of phi then 8 of lambda, the
    // phase's last register bit being k)
    const uint32_t tsize = 1u << sw.k, ngroups = sweep_threads(sw.k, ns);
    const uint32_t cmask = (1u << sw.c) - 1u;
    std::vector<c64> t(size_t(tsize) * ns);
    std::vector<uint64_t> gidx(tsize);
    for (uint32_t i = 0; i < sw.n_grad; i++) gsum_tile[i] = 0.0;
    const uint64_t base = tile_base(sw, tile);
    std::vector<c64> xf(std::max<uint32_t>(sw.n_xf, 1));
    const uint64_t xbits = ext_index(sw, base);
    for (uint32_t i = 0; i < 16u * sw.n_lt; i++) xf[sw.n_xf - 16u * sw.n_lt + i] = coef[sw.lt_coef + i];
    for (uint32_t o = sw.op_begin; o < sw.op_end; o++)
        if (has_tile_factor(ops[o].h)) xf[ops[o].h.grad] = tile_factor(ops[o].h, coef, base, xbits);
    for (uint32_t l = 0; l < tsize; l++) {
        gidx[l] = base | hitab[sw.hitab + (l >> sw.c)] | (l & cmask);
        t[l] = init_zero ? c64{gidx[l] == 0 ? 1.0 : 0.0, 0.0} : s0[gidx[l]];
        if (ns == 2) t[tsize + l] = s1[gidx[l]];
    }
    // two-state sweeps always address LDS through their tile maps (the
    // identity where nothing was folded), one-state sweeps when mapped
    const bool mp = ns == 2 || sw.mapped;
    for (uint32_t p = sw.phase_begin; p < sw.phase_end; p++) {
        const PhaseInfo& ph = phases[p];
        uint32_t off[kRegs], goffs[kRegs];
        phase_offsets(ph, off);
        phase_gather_offsets(ph, goffs);
        uint32_t soffs[kRegs];
        phase_scatter_offsets(ph, soffs);
        for (uint32_t g = 0; g < ngroups; g++) {
            const uint32_t goff = group_offset(ph, g);
            const uint32_t gphys = mp ? mapped_group_offset(ph, g) : goff;
            c64 a[kRegs];
            for (int r = 0; r < kRegs; r++) a[r] = t[gphys ^ goffs[r]];
            for (uint32_t o = ph.op_begin; o < ph.op_end; o++) {
                const DOp& op = ops[o];
                if (op.h.type == DOP_GRAD_DENSE || op.h.type == DOP_GRAD_DIAG)
                    gsum_tile[op.h.grad] += exec_grad_op<kRegBits, kRegsAdj>(op.h, &op, a, coef, goff, off, base);
                else
                    exec_state_op(op.h, &op, a, coef, goff, g, off, base, xf.data(), sw.exthi,
                                  uint32_t(extract_bits(base, sw.exthi)));
            }
            for (int r = 0; r < kRegs; r++) t[mp ? (gphys ^ soffs[r]) : ((goff | off[r]) ^ ph.sxor)] = a[r];
        }
    }
    const uint32_t tbits = sw.k - uint32_t(ns == 2 ? kRegBitsAdj : kRegBits);   // log2 of the tile's threads
    for (uint32_t l = 0; l < tsize; l++) {
        // as the kernels' store: M(tid) ^ M(nt j) with l = tid + nt j
        const uint32_t m = sw.mapped ? sweep_map_thread(sw, l & ((1u << tbits) - 1u)) ^ sw.fj[l >> tbits] : l;
        s0[gidx[l]] = t[m];
        if (ns == 2) s1[gidx[l]] = t[tsize + m];
    }
}

// A measurement step on the host: P(1) over the first 2^m.n amplitudes, the
// outcome and the collapse factors written into the program's coefficients
// (what k_prob_one / k_measure_decide do on the device).
MeasureResult host_measure(const MeasureStep& m, const std::vector<c64>& s, std::vector<c64>& coef)
{
    double p = 0.0;
    if (m.dm_n) {   // density matrix: the diagonal of the rows with bit q set
        for (uint64_t i = 0; i < (1ull << m.dm_n); i++)
            if ((i >> m.qubit) & 1) p += std::fabs(s[(i << m.dm_n) | i].re);
    } else {
        for (uint64_t x = 0; x < (1ull << m.n); x++)
            if ((x >> m.qubit) & 1) p += s[x].re * s[x].re + s[x].im * s[x].im;
    }
    const bool one = m.draw < p;
    const double w = one ? p : 1.0 - p;
    const double f = m.dm_n ? 1.0 / w : 1.0 / std::sqrt(w);
    coef[m.coef] = c64{one ? 0.0 : f, 0.0};
    coef[m.coef + 1] = c64{one ? f : 0.0, 0.0};
    MeasureResult r;
    r.prob = p;
    r.outcome = one ? 1 : 0;
    r.pad = 0;
    return r;
}

void host_program(Program& prog, std::vector<c64>& s0, std::vector<c64>& s1, std::vector<double>& gsum,
                  std::vector<MeasureResult>* meas = nullptr)
{
    const int ns = prog.n_states;
    gsum.assign(prog.n_grad, 0.0);
    for (const Step& st : prog.steps) {
        if (st.kind == Step::MEASURE) {
            const MeasureResult r = host_measure(prog.measures[st.index], s0, prog.coef);
            if (meas) meas->push_back(r);
            continue;
        }
        if (st.kind == Step::STANDALONE) {
            host_standalone(s0, prog.n, prog.standalone[st.index], prog.coef);
            if (ns == 2) host_standalone(s1, prog.n, prog.standalone[st.index], prog.coef);
            continue;
        }
        const SweepInfo& sw = prog.sweeps[st.index];
        std::vector<double> g(sw.n_grad);
        for (uint64_t tile = 0; tile < sw.n_tiles; tile++) {
            host_tile(sw, prog.phases.data(), prog.ops.data(), prog.coef.data(), prog.hitab.data(), ns, s0.data(),
                      ns == 2 ? s1.data() : nullptr, tile, false, g.data());
            for (uint32_t i = 0; i < sw.n_grad; i++) gsum[sw.grad_base + i] += g[i];
        }
    }
}

void host_pauli(int n, const std::vector<c64>& psi, std::vector<c64>* lam, const svmi_pauli_term* terms,
                size_t n_terms, double* E)
{
    std::vector<PauliGroup> groups;
    std::vector<PauliTerm> pt;
    group_terms(terms, n_terms, groups, pt);
    const uint64_t dim = 1ull << n;
    double acc = 0.0;
    for (uint64_t y = 0; y < dim; y++) {
        c64 l{0.0, 0.0};
        for (const PauliGroup& G : groups) {
            const uint64_t x = y ^ G.xmask;
            double fr = 0.0, fi = 0.0;
            for (uint32_t t = G.begin; t < G.end; t++) {
                const bool neg = __builtin_popcountll(x & pt[t].zmask) & 1;
                fr += neg ? -pt[t].cre : pt[t].cre;
                fi += neg ? -pt[t].cim : pt[t].cim;
            }
            l = cmac(l, c64{fr, fi}, psi[x]);
        }
        if (lam) (*lam)[y] = l;
        acc += psi[y].re * l.re + psi[y].im * l.im;
    }
    *E = acc;
}

PlanOptions debug_options(int fusion, int tile_qubits, int low_qubits, int64_t budget)
{
    PlanOptions o;
    o.fusion = fusion != 0;
    o.tile_qubits = tile_qubits;
    o.low_qubits = low_qubits;
    o.sweep_budget = long(budget);
    return o;
}

}  // namespace

extern "C" {

svmi_status svmi_debug_emulate(int n, const svmi_gate* gates, size_t n_gates, int fusion, int tile_qubits,
                               int low_qubits, int64_t sweep_budget, double* re, double* im)
{
    if (n < 1 || n > 26) return fail(SVMI_ERR_ARG, "emulation supports 1..26 qubits");
    if (!re || !im) return fail(SVMI_ERR_ARG, "null state");
    if (tile_qubits < kRegBits || tile_qubits > kMaxTileQubits) return fail(SVMI_ERR_ARG, "bad tile_qubits");
    SVMI_TRY(check_gates(gates, n_gates, n));
    const int ni = std::max(n, kRegBits);
    Program prog;
    plan_forward(gates, n_gates, ni, debug_options(fusion, tile_qubits, low_qubits, sweep_budget), prog);
    std::vector<c64> s(1ull << ni, c64{0.0, 0.0}), dummy;
    for (uint64_t x = 0; x < (1ull << n); x++) s[x] = c64{re[x], im[x]};
    std::vector<double> g;
    host_program(prog, s, dummy, g);
    for (uint64_t x = 0; x < (1ull << n); x++) {
        re[x] = s[x].re;
        im[x] = s[x].im;
    }
    return SVMI_OK;
}

svmi_status svmi_debug_emulate_adjoint(int n, const svmi_gate* gates, size_t n_gates, const svmi_pauli_term* terms,
                                       size_t n_terms, const int32_t* param_slots, size_t n_params, int tile_qubits,
                                       int low_qubits, const double* re0, const double* im0, double* expval_out,
                                       double* grad_out)
{
    if (n < 1 || n > 22) return fail(SVMI_ERR_ARG, "emulation supports 1..22 qubits");
    if (!re0 || !im0 || !expval_out || (n_params && !grad_out)) return fail(SVMI_ERR_ARG, "null argument");
    if (tile_qubits < kRegBits || tile_qubits > kMaxTileQubits) return fail(SVMI_ERR_ARG, "bad tile_qubits");
    SVMI_TRY(check_gates(gates, n_gates, n));
    SVMI_TRY(check_terms(terms, n_terms, n));
    SVMI_TRY(check_params(gates, n_gates, param_slots, n_params));
    const int ni = std::max(n, kRegBits);
    const PlanOptions opt = debug_options(1, tile_qubits, low_qubits, 0);
    Program fp;
    plan_forward(gates, n_gates, ni, opt, fp);
    std::vector<c64> psi(1ull << ni, c64{0.0, 0.0}), dummy;
    for (uint64_t x = 0; x < (1ull << n); x++) psi[x] = c64{re0[x], im0[x]};
    std::vector<double> g;
    host_program(fp, psi, dummy, g);
    std::vector<c64> lam(psi.size());
    double E = 0.0;
    host_pauli(ni, psi, &lam, terms, n_terms, &E);
    *expval_out = E;
    std::vector<AdjEntry> entries;
    adjoint_entries(gates, n_gates, param_slots, entries);
    Program bp;
    build_adjoint_program(entries, ni, opt, bp);
    host_program(bp, psi, lam, g);
    for (size_t p = 0; p < n_params; p++) grad_out[p] = 0.0;
    for (uint32_t i = 0; i < bp.n_grad; i++) grad_out[bp.grad_param[i]] += 2.0 * g[i];
    return SVMI_OK;
}

svmi_status svmi_debug_emulate_batch(const svmi_batch_item* items, size_t n_items, double* expvals, double* grads)
{
    if (n_items == 0) return SVMI_OK;
    for (size_t i = 0; i < n_items; i++)
        if (items && items[i].n_qubits > 20) return fail(SVMI_ERR_ARG, "batch emulation supports up to 20 qubits");
    BatchPlan bp;
    SVMI_TRY(build_batch(items, n_items, expvals, grads, bp));
    std::vector<c64> mem(bp.state_bytes / sizeof(c64) + 1);
    auto phi = [&](size_t c) { return mem.data() + bp.state_off[c] / sizeof(c64); };
    for (size_t i : bp.need_fill) {
        c64* p = phi(i);
        for (uint64_t x = 0; x < (1ull << bp.circs[i].n); x++) p[x] = c64{x == 0 ? 1.0 : 0.0, 0.0};
    }
    std::vector<double> part(bp.n_grad_partials + bp.n_e_partials);
    double* g_part = part.data();
    double* e_part = part.data() + bp.n_grad_partials;
    // launch by launch; inside a launch the tasks run last to first, since
    // the device runs them in no particular order
    for (const BatchLaunch& L : bp.launches)
    for (uint32_t si = L.step_end; si-- > L.step_begin;) {
        const BatchStep& st = bp.steps[si];
        const BatchCircuit& c = bp.circs[st.circuit];
        c64* ph = phi(st.circuit);
        c64* lm = ph + (1ull << c.n);
        for (uint64_t task = st.n_tasks; task-- > 0;) {
            if (st.kind == BSTEP_FWD || st.kind == BSTEP_ADJ) {
                const SweepInfo& sw = bp.sweeps[st.sweep];
                std::vector<double> g(sw.n_grad);
                host_tile(sw, bp.phases.data(), bp.ops.data(), bp.coef.data(), bp.hitab.data(),
                          st.kind == BSTEP_ADJ ? 2 : 1, ph, st.kind == BSTEP_ADJ ? lm : nullptr, task,
                          L.init != 0, g.data());
                for (uint32_t i = 0; i < sw.n_grad; i++) g_part[st.partial + i * sw.n_tiles + task] = g[i];
            } else {
                const uint64_t dim = 1ull << c.n;
                const uint64_t y0 = task * kPauliChunk, y1 = std::min<uint64_t>(y0 + kPauliChunk, dim);
                double acc = 0.0;
                for (uint64_t y = y0; y < y1; y++) {
                    c64 l{0.0, 0.0};
                    for (uint32_t gi = c.group_begin; gi < c.group_end; gi++) {
                        const PauliGroup& G = bp.groups[gi];
                        const uint64_t x = y ^ G.xmask;
                        double fr = 0.0, fi = 0.0;
                        for (uint32_t t = G.begin; t < G.end; t++) {
                            const bool neg = __builtin_popcountll(x & bp.terms[t].zmask) & 1;
                            fr += neg ? -bp.terms[t].cre : bp.terms[t].cre;
                            fi += neg ? -bp.terms[t].cim : bp.terms[t].cim;
                        }
                        l = cmac(l, c64{fr, fi}, ph[x]);
                    }
                    lm[y] = l;
                    acc += ph[y].re * l.re + ph[y].im * l.im;
                }
                e_part[st.partial + task] = acc;
            }
        }
    }
    // final segment sums, in the device kernel's order (256 strided
    // partial sums, then a halving tree)
    std::vector<double> res(bp.n_slots + n_items, 0.0);
    for (const SumSegment& g : bp.segs) {
        double red[256];
        for (uint32_t t = 0; t < 256; t++) {
            double acc = 0.0;
            for (uint64_t i = t; i < g.len; i += 256) acc += part[g.src + i];
            red[t] = acc;
        }
        for (uint32_t k = 128; k > 0; k >>= 1)
            for (uint32_t t = 0; t < k; t++) red[t] += red[t + k];
        res[g.dst] = red[0];
    }
    const std::vector<double> gs(res.begin(), res.begin() + bp.n_slots), eo(res.begin() + bp.n_slots, res.end());
    finish_batch(bp, items, n_items, gs, eo, expvals, grads);
    return SVMI_OK;
}

svmi_status svmi_debug_plan_batch(const svmi_batch_item* items, size_t n_items, uint64_t* n_levels,
                                  uint64_t* n_launches, uint64_t* n_tasks, uint64_t* state_bytes)
{
    if (n_items == 0) return fail(SVMI_ERR_ARG, "empty batch");
    std::vector<double> ev(n_items), gr(1);
    size_t total = 0;
    for (size_t i = 0; i < n_items; i++) total += items ? items[i].n_params : 0;
    gr.resize(std::max<size_t>(total, 1));
    BatchPlan bp;
    SVMI_TRY(build_batch(items, n_items, ev.data(), gr.data(), bp));
    if (n_levels) *n_levels = bp.level_begin.size() - 1;
    if (n_launches) *n_launches = bp.launches.size();
    if (n_tasks) *n_tasks = bp.n_tasks;
    if (state_bytes) *state_bytes = bp.state_bytes;
    return SVMI_OK;
}

svmi_status svmi_debug_jit(int n, const svmi_gate* gates, size_t n_gates, int tile_qubits, int low_qubits,
                           int compile, double* compile_ms, uint64_t* code_bytes, char* buf, size_t buflen)
{
    if (n < 1 || n > kMaxQubits) return fail(SVMI_ERR_ARG, "bad n");
    if (tile_qubits < kRegBits || tile_qubits > kMaxTileQubits) return fail(SVMI_ERR_ARG, "bad tile_qubits");
    SVMI_TRY(check_gates(gates, n_gates, n));
    Program prog;
    PlanOptions o = debug_options(1, tile_qubits, low_qubits, 0);
    o.tile_map = false;   // as SVMI_OPT_SPECIALIZE plans
    plan_forward(gates, n_gates, std::max(n, kRegBits), o, prog);
    std::vector<char> covered;
    const uint64_t dma_grid = compile == 2 ? 256 : 0;   // double-buffered kernels for a 256-CU device
    const std::string src = jit_source(prog, covered, dma_grid);
    if (buf && buflen) {
        const size_t m = std::min(buflen - 1, src.size());
        std::memcpy(buf, src.data(), m);
        buf[m] = 0;
    }
    if (compile) {
        std::string err;
        double ms = 0.0;
        const size_t bytes = jit_compile_only(prog, "gfx950", &ms, err, dma_grid);
        if (!bytes) return fail(SVMI_ERR_UNSUPPORTED, "jit: " + err);
        if (compile_ms) *compile_ms = ms;
        if (code_bytes) *code_bytes = bytes;
    }
    return SVMI_OK;
}

svmi_status svmi_debug_emulate_qir(int n, const svmi_qir_gate* gates, size_t n_gates, int tile_qubits,
                                   int low_qubits, double* re, double* im, double* probs, int32_t* outcomes)
{
    if (n < 1 || n > 26) return fail(SVMI_ERR_ARG, "emulation supports 1..26 qubits");
    if (!re || !im) return fail(SVMI_ERR_ARG, "null state");
    if (tile_qubits < kRegBits || tile_qubits > kMaxTileQubits) return fail(SVMI_ERR_ARG, "bad tile_qubits");
    const int ni = std::max(n, kRegBits);
    Program prog;
    SVMI_TRY(plan_qir(gates, n_gates, n, ni, debug_options(1, tile_qubits, low_qubits, 0), prog));
    std::vector<c64> s(1ull << ni, c64{0.0, 0.0}), dummy;
    for (uint64_t x = 0; x < (1ull << n); x++) s[x] = c64{re[x], im[x]};
    std::vector<double> g;
    std::vector<MeasureResult> res;
    host_program(prog, s, dummy, g, &res);
    for (uint64_t x = 0; x < (1ull << n); x++) {
        re[x] = s[x].re;
        im[x] = s[x].im;
    }
    for (size_t m = 0; m < res.size(); m++) {
        if (probs) probs[m] = res[m].prob;
        if (outcomes) outcomes[m] = res[m].outcome;
    }
    return SVMI_OK;
}

svmi_status svmi_debug_emulate_dm_qir(int n, const svmi_qir_gate* gates, size_t n_gates, int tile_qubits,
                                      int low_qubits, double* re, double* im, double* probs, int32_t* outcomes)
{
    if (n < 1 || n > 13) return fail(SVMI_ERR_ARG, "density-matrix emulation supports 1..13 qubits");
    if (!re || !im) return fail(SVMI_ERR_ARG, "null state");
    if (tile_qubits < kRegBits || tile_qubits > kMaxTileQubits) return fail(SVMI_ERR_ARG, "bad tile_qubits");
    const int n2 = 2 * n, ni = std::max(n2, kRegBits);
    Program prog;
    SVMI_TRY(plan_qir(gates, n_gates, n, ni, debug_options(1, tile_qubits, low_qubits, 0), prog, true));
    std::vector<c64> s(1ull << ni, c64{0.0, 0.0}), dummy;
    for (uint64_t x = 0; x < (1ull << n2); x++) s[x] = c64{re[x], im[x]};
    std::vector<double> g;
    std::vector<MeasureResult> res;
    host_program(prog, s, dummy, g, &res);
    for (uint64_t x = 0; x < (1ull << n2); x++) {
        re[x] = s[x].re;
        im[x] = s[x].im;
    }
    for (size_t m = 0; m < res.size(); m++) {
        if (probs) probs[m] = res[m].prob;
        if (outcomes) outcomes[m] = res[m].outcome;
    }
    return SVMI_OK;
}

svmi_status svmi_debug_plan_qir(int n, const svmi_qir_gate* gates, size_t n_gates, int tile_qubits, int low_qubits,
                                uint64_t* n_sweeps, uint64_t* n_measures, char* buf, size_t buflen)
{
    if (n < 1 || n > kMaxQubits) return fail(SVMI_ERR_ARG, "bad n");
    if (tile_qubits < kRegBits || tile_qubits > kMaxTileQubits) return fail(SVMI_ERR_ARG, "bad tile_qubits");
    Program prog;
    SVMI_TRY(plan_qir(gates, n_gates, n, std::max(n, kRegBits), debug_options(1, tile_qubits, low_qubits, 0), prog));
    if (n_sweeps) *n_sweeps = prog.sweeps.size();
    if (n_measures) *n_measures = prog.measures.size();
    if (buf && buflen) {
        const std::string d = describe_program(prog);
        const size_t m = std::min(buflen - 1, d.size());
        std::memcpy(buf, d.data(), m);
        buf[m] = 0;
    }
    return SVMI_OK;
}

svmi_status svmi_debug_step_times(svmi_handle h, double* ms, size_t capacity, uint64_t* n_steps)
{
    if (!h || !n_steps || (capacity && !ms)) return fail(SVMI_ERR_ARG, "null argument");
    *n_steps = h->step_ms.size();
    for (size_t i = 0; i < capacity && i < h->step_ms.size(); i++) ms[i] = h->step_ms[i];
    return SVMI_OK;
}

svmi_status svmi_debug_plan(int n, const svmi_gate* gates, size_t n_gates, int fusion, int tile_qubits,
                            int low_qubits, int64_t sweep_budget, uint64_t* n_passes, char* buf, size_t buflen)
{
    if (n < 1 || n > kMaxQubits) return fail(SVMI_ERR_ARG, "bad n");
    if (tile_qubits < kRegBits || tile_qubits > kMaxTileQubits) return fail(SVMI_ERR_ARG, "bad tile_qubits");
    SVMI_TRY(check_gates(gates, n_gates, n));
    Program prog;
    plan_forward(gates, n_gates, std::max(n, kRegBits), debug_options(fusion, tile_qubits, low_qubits, sweep_budget),
                 prog);
    if (n_passes) *n_passes = prog.steps.size();
    if (buf && buflen) {
        const std::string s = describe_program(prog);
        const size_t m = std::min(buflen - 1, s.size());
        std::memcpy(buf, s.data(), m);
        buf[m] = 0;
    }
    return SVMI_OK;
}

svmi_status svmi_debug_plan_adjoint(int n, const svmi_gate* gates, size_t n_gates, const int32_t* param_slots,
                                    size_t n_params, int tile_qubits, int low_qubits, uint64_t* n_passes, char* buf,
                                    size_t buflen)
{
    if (n < 1 || n > kMaxQubits) return fail(SVMI_ERR_ARG, "bad n");
    if (tile_qubits < kRegBits || tile_qubits > kMaxTileQubits) return fail(SVMI_ERR_ARG, "bad tile_qubits");
    SVMI_TRY(check_gates(gates, n_gates, n));
    SVMI_TRY(check_params(gates, n_gates, param_slots, n_params));
    std::vector<AdjEntry> entries;
    adjoint_entries(gates, n_gates, param_slots, entries);
    Program prog;
    build_adjoint_program(entries, std::max(n, kRegBits), debug_options(1, tile_qubits, low_qubits, 0), prog);
    if (n_passes) *n_passes = prog.steps.size();
    if (buf && buflen) {
        const std::string s = describe_program(prog);
        const size_t m = std::min(buflen - 1, s.size());
        std::memcpy(buf, s.data(), m);
        buf[m] = 0;
    }
    return SVMI_OK;
}

}  // extern "C"
