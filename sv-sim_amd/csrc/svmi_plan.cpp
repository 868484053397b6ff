// svmi_plan.cpp -- sweep/phase scheduling and device-program encoding.
//
// Scheduling (greedy, in circuit order):
//   sweep level: a sweep's tile set T starts as the low physical qubits
//   0..c-1 (coalescing) and grows by the qubits of the next non-diagonal ops
//   until it has k qubits.  An op is taken if it does not touch a qubit that
//   an earlier skipped op touches (which keeps every qubit's op order) and it
//   is diagonal or all its qubits fit in T.  Diagonal ops need no tile bits.
//   phase level: the same greedy over the sweep's ops with the 4 register
//   bits R instead of T.
// Every sweep is one read and one write of the whole state.
#include "svmi_plan.hpp"

#include <algorithm>
#include <cstring>
#include <sstream>

namespace svmi {

namespace {

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

uint64_t op_mask(const LOp& op)
{
    uint64_t m = 0;
    for (int i = 0; i < op.k; i++) m |= uint64_t(1) << op.q[i];
    return m;
}

// FMA cost per amplitude (for the sweep budget).
long op_cost(const LOp& op, bool grad)
{
    if (op.kind == LOp::DIAG) return grad ? 6 : 4;
    if (op.kind == LOp::SRN) return 4;
    return (4L << op.k) * (grad ? 2 : 1);
}

struct SOp {
    const LOp* op;
    bool grad;
    int param;
    uint64_t mask;       // qubits the op acts on
    uint64_t dep;        // qubits it must stay ordered with (SRN: all)
    bool diag;
    bool standalone;
    bool identity;
};

bool is_identity(const LOp& op)
{
    if (op.kind == LOp::SRN) return false;
    const size_t D = size_t(1) << op.k;
    for (size_t r = 0; r < D; r++) {
        if (op.kind == LOp::DIAG) {
            if (op.m[r].re != 1.0 || op.m[r].im != 0.0) return false;
        } else {
            for (size_t c = 0; c < D; c++) {
                const c64 v = op.m[r * D + c];
                const double want = r == c ? 1.0 : 0.0;
                if (v.re != want || v.im != 0.0) return false;
            }
        }
    }
    return true;
}

SOp make_sop(const LOp* op, bool grad, int param)
{
    SOp s;
    s.op = op;
    s.grad = grad;
    s.param = param;
    s.mask = op_mask(*op);
    // The reference SRN (svsim_cpu_omp.hpp:866-879) is only real-linear: it
    // does not commute with complex phases even on other qubits.
    s.dep = op->kind == LOp::SRN ? ~uint64_t(0) : s.mask;
    s.diag = op->kind == LOp::DIAG;
    s.standalone = op->kind == LOp::DENSE && op->k > kRegBits;
    s.identity = !grad && is_identity(*op);
    return s;
}

// Exact permutation patterns on the op's own qubit order.
bool is_perm(const LOp& op, const int* src /* out row r <- in col src[r] */)
{
    const size_t D = size_t(1) << op.k;
    for (size_t r = 0; r < D; r++)
        for (size_t c = 0; c < D; c++) {
            const c64 v = op.m[r * D + c];
            const double want = (int(c) == src[r]) ? 1.0 : 0.0;
            if (v.re != want || v.im != 0.0) return false;
        }
    return true;
}

struct Encoder {
    Program& prog;
    explicit Encoder(Program& p) : prog(p) {}

    uint32_t push_coef(const std::vector<c64>& v)
    {
        const uint32_t off = uint32_t(prog.coef.size());
        prog.coef.insert(prog.coef.end(), v.begin(), v.end());
        return off;
    }

    // Encode op for a phase: tloc[p] = tile-local bit of physical qubit p (or -1),
    // slot_of[l] = register slot of local bit l (or -1).
    void encode(const SOp& s, const int* tloc, const int* slot_of, uint32_t grad_slot)
    {
        const LOp& op = *s.op;
        DOp d;
        std::memset(&d, 0, sizeof(d));
        d.targets = prog.n_states == 2 ? DOP_TO_BOTH : DOP_TO_PSI;
        d.grad = grad_slot;
        if (op.kind == LOp::SRN) {
            d.type = DOP_SRN;
            d.s0 = uint8_t(slot_of[tloc[op.q[0]]]);
            prog.ops.push_back(d);
            return;
        }
        if (op.kind == LOp::DIAG) {
            const size_t D = size_t(1) << op.k;
            bool cphase = true;
            for (size_t r = 0; r + 1 < D; r++)
                if (op.m[r].re != 1.0 || op.m[r].im != 0.0) cphase = false;
            d.nq = uint8_t(op.k);
            for (int i = 0; i < op.k; i++) {
                const int l = tloc[op.q[i]];
                d.qloc[i] = int8_t(l);
                d.qphys[i] = uint8_t(op.q[i]);
                if (l >= 0) d.lmask |= 1u << l;
                else d.emask |= uint64_t(1) << op.q[i];
            }
            if (cphase && !s.grad) {
                d.type = DOP_CPHASE;
                d.coef = push_coef(std::vector<c64>{op.m[D - 1]});
            } else {
                d.type = s.grad ? DOP_GRAD_DIAG : DOP_DIAG;
                d.coef = push_coef(op.m);
            }
            prog.ops.push_back(d);
            return;
        }
        // dense on register slots
        int slots[5];
        for (int i = 0; i < op.k; i++) slots[i] = slot_of[tloc[op.q[i]]];
        if (!s.grad) {
            if (op.k == 1) {
                const int src[2] = {1, 0};
                if (is_perm(op, src)) {
                    d.type = DOP_X;
                    d.s0 = uint8_t(slots[0]);
                    prog.ops.push_back(d);
                    return;
                }
            } else if (op.k == 2) {
                // local index bit0 = q[0], bit1 = q[1]
                const int cx01[4] = {0, 3, 2, 1};   // control q0, target q1
                const int cx10[4] = {0, 1, 3, 2};   // control q1, target q0
                const int sw[4] = {0, 2, 1, 3};
                if (is_perm(op, cx01)) {
                    d.type = DOP_CX; d.s0 = uint8_t(slots[0]); d.s1 = uint8_t(slots[1]);
                    prog.ops.push_back(d);
                    return;
                }
                if (is_perm(op, cx10)) {
                    d.type = DOP_CX; d.s0 = uint8_t(slots[1]); d.s1 = uint8_t(slots[0]);
                    prog.ops.push_back(d);
                    return;
                }
                if (is_perm(op, sw)) {
                    d.type = DOP_SWAP;
                    d.s0 = uint8_t(std::min(slots[0], slots[1]));
                    d.s1 = uint8_t(std::max(slots[0], slots[1]));
                    prog.ops.push_back(d);
                    return;
                }
            }
        }
        // reorder the matrix so that matrix bit j <-> j-th lowest slot
        int order[5];
        for (int i = 0; i < op.k; i++) order[i] = i;
        std::sort(order, order + op.k, [&](int a, int b) { return slots[a] < slots[b]; });
        int newbit[5];  // op qubit i -> matrix bit newbit[i]
        for (int j = 0; j < op.k; j++) newbit[order[j]] = j;
        const size_t D = size_t(1) << op.k;
        std::vector<c64> M(D * D);
        auto remap = [&](size_t x) {
            size_t y = 0;
            for (int i = 0; i < op.k; i++)
                if ((x >> i) & 1) y |= size_t(1) << newbit[i];
            return y;
        };
        for (size_t r = 0; r < D; r++)
            for (size_t c = 0; c < D; c++) M[remap(r) * D + remap(c)] = op.m[r * D + c];
        d.type = s.grad ? DOP_GRAD_DENSE : DOP_DENSE;
        d.smask = 0;
        for (int i = 0; i < op.k; i++) d.smask |= uint16_t(1u << slots[i]);
        d.coef = push_coef(M);
        prog.ops.push_back(d);
    }
};

// Emit a standalone full-state op step.
void emit_standalone(Program& prog, const SOp& s)
{
    const LOp& op = *s.op;
    StandaloneOp so;
    std::memset(&so, 0, sizeof(so));
    so.kind = op.kind == LOp::DIAG ? 1 : (op.kind == LOp::SRN ? 2 : 0);
    so.q = op.k;
    for (int i = 0; i < op.k; i++) so.qubits[i] = op.q[i];
    so.coef = op.kind == LOp::SRN ? 0 : uint32_t(prog.coef.size());
    prog.coef.insert(prog.coef.end(), op.m.begin(), op.m.end());
    prog.steps.push_back(Step{Step::STANDALONE, uint32_t(prog.standalone.size())});
    prog.standalone.push_back(so);
}

void emit_sweep(Program& prog, const std::vector<const SOp*>& taken, uint64_t T, int n, int c)
{
    Encoder enc(prog);
    int tpos[64];
    int tloc[64];
    int k = 0;
    for (int p = 0; p < n; p++) {
        tloc[p] = -1;
        if ((T >> p) & 1) {
            tloc[p] = k;
            tpos[k++] = p;
        }
    }
    SweepInfo sw;
    std::memset(&sw, 0, sizeof(sw));
    sw.k = uint32_t(k);
    sw.c = uint32_t(c);
    sw.phase_begin = uint32_t(prog.phases.size());
    sw.hitab = uint32_t(prog.hitab.size());
    sw.n_ext = 0;
    for (int p = 0; p < n; p++)
        if (!((T >> p) & 1)) sw.ext[sw.n_ext++] = uint8_t(p);
    sw.n_tiles = uint64_t(1) << (n - k);
    for (uint64_t h = 0; h < (uint64_t(1) << (k - c)); h++) {
        uint64_t off = 0;
        for (int j = 0; j < k - c; j++)
            if ((h >> j) & 1) off |= uint64_t(1) << tpos[c + j];
        prog.hitab.push_back(off);
    }
    sw.grad_base = prog.n_grad;

    // phase scheduling
    std::vector<const SOp*> remaining = taken;
    uint32_t n_grad_sweep = 0;
    while (!remaining.empty()) {
        uint32_t R = 0;  // local-bit mask
        uint64_t blocked = 0;
        std::vector<const SOp*> ph, rest;
        for (const SOp* s : remaining) {
            if (s->dep & blocked) {
                blocked |= s->dep;
                rest.push_back(s);
                continue;
            }
            if (s->diag) {
                ph.push_back(s);
                continue;
            }
            uint32_t lm = 0;
            for (int i = 0; i < s->op->k; i++) lm |= 1u << tloc[s->op->q[i]];
            if (popc(R | lm) <= kRegBits) {
                R |= lm;
                ph.push_back(s);
            } else {
                blocked |= s->dep;
                rest.push_back(s);
            }
        }
        // pad R to 4 bits with the lowest free tile bits above the low qubits first
        for (int l = 0; l < k && popc(R) < kRegBits; l++)
            if (!((R >> l) & 1) && l >= c) R |= 1u << l;
        for (int l = 0; l < k && popc(R) < kRegBits; l++)
            if (!((R >> l) & 1)) R |= 1u << l;
        PhaseInfo pi;
        std::memset(&pi, 0, sizeof(pi));
        int slot_of[64];
        for (int l = 0; l < 64; l++) slot_of[l] = -1;
        int ns = 0;
        for (int l = 0; l < k; l++)
            if ((R >> l) & 1) {
                pi.r[ns] = uint8_t(l);
                slot_of[l] = ns++;
            }
        pi.op_begin = uint32_t(prog.ops.size());
        for (const SOp* s : ph) {
            uint32_t gslot = 0;
            if (s->grad) {
                gslot = n_grad_sweep++;
                prog.grad_param.push_back(s->param);
            }
            enc.encode(*s, tloc, slot_of, gslot);
        }
        pi.op_end = uint32_t(prog.ops.size());
        prog.phases.push_back(pi);
        remaining.swap(rest);
    }
    sw.phase_end = uint32_t(prog.phases.size());
    sw.n_grad = n_grad_sweep;
    prog.n_grad += n_grad_sweep;
    prog.steps.push_back(Step{Step::SWEEP, uint32_t(prog.sweeps.size())});
    prog.sweeps.push_back(sw);
}

void schedule(std::vector<SOp>& sops, int n, const PlanOptions& opt, Program& prog)
{
    prog.n = n;
    const bool fused = opt.fusion && n >= kRegBits;
    std::vector<const SOp*> pending;
    for (auto& s : sops)
        if (!s.identity) pending.push_back(&s);
    prog.n_lowered = pending.size();
    if (!fused) {
        for (const SOp* s : pending) {
            if (s->grad) {
                // per-op adjoint: a one-op sweep over the whole state is not
                // possible below 4 qubits; callers keep n >= 4 for adjoint.
            }
            emit_standalone(prog, *s);
        }
        return;
    }
    const int k = std::min(std::max(opt.tile_qubits, kRegBits),
                           prog.n_states == 2 ? kMaxTileQubitsAdj : kMaxTileQubits);
    const int kk = std::min(k, n);
    const int c = std::min(std::max(opt.low_qubits, 0), kk);
    const uint64_t all = n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
    const uint64_t low = c >= 64 ? ~uint64_t(0) : ((uint64_t(1) << c) - 1);
    while (!pending.empty()) {
        uint64_t T = low;
        uint64_t blocked = 0;
        long cost = 0;
        int grads = 0;
        std::vector<const SOp*> taken, rest;
        size_t i = 0;
        for (; i < pending.size(); i++) {
            const SOp* s = pending[i];
            if (blocked == all) break;
            const bool over = (opt.sweep_budget > 0 && cost >= opt.sweep_budget && !taken.empty()) ||
                              (s->grad && grads >= kMaxGradsPerSweep);
            if ((s->dep & blocked) || s->standalone || over) {
                blocked |= s->dep;
                if (over) blocked = all;
                rest.push_back(s);
                continue;
            }
            if (s->diag) {
                taken.push_back(s);
                cost += op_cost(*s->op, s->grad);
                grads += s->grad ? 1 : 0;
                continue;
            }
            const uint64_t t2 = T | s->mask;
            if (popc(t2) <= kk) {
                T = t2;
                taken.push_back(s);
                cost += op_cost(*s->op, s->grad);
                grads += s->grad ? 1 : 0;
            } else {
                blocked |= s->dep;
                rest.push_back(s);
            }
        }
        for (; i < pending.size(); i++) rest.push_back(pending[i]);
        if (taken.empty()) {
            // the first pending op cannot go in a tile
            emit_standalone(prog, *pending.front());
            pending.erase(pending.begin());
            continue;
        }
        // pad T to kk qubits with the lowest free positions
        for (int p = 0; p < n && popc(T) < kk; p++) T |= uint64_t(1) << p;
        emit_sweep(prog, taken, T, n, c);
        pending.swap(rest);
    }
}

}  // namespace

void fuse_single_qubit_runs(const std::vector<LOp>& in, std::vector<LOp>& out)
{
    out.clear();
    out.reserve(in.size());
    std::vector<int> last(64, -1);
    for (const LOp& op : in) {
        const bool mergeable = op.k == 1 && op.kind != LOp::SRN;
        if (mergeable && last[op.q[0]] >= 0) {
            LOp merged;
            lop_compose_same(out[last[op.q[0]]], op, merged);
            out[last[op.q[0]]] = merged;
            continue;
        }
        out.push_back(op);
        if (op.kind == LOp::SRN) std::fill(last.begin(), last.end(), -1);  // real-linear only
        for (int i = 0; i < op.k; i++) last[op.q[i]] = -1;
        if (mergeable) last[op.q[0]] = int(out.size()) - 1;
    }
}

void build_forward_program(const std::vector<LOp>& ops, int n, const PlanOptions& opt, Program& prog)
{
    prog = Program();
    prog.n_states = 1;
    std::vector<SOp> sops;
    sops.reserve(ops.size());
    for (const LOp& op : ops) sops.push_back(make_sop(&op, false, -1));
    schedule(sops, n, opt, prog);
}

void build_adjoint_program(const std::vector<AdjEntry>& entries, int n, const PlanOptions& opt,
                           Program& prog)
{
    prog = Program();
    prog.n_states = 2;
    std::vector<SOp> sops;
    sops.reserve(entries.size());
    for (const AdjEntry& e : entries) sops.push_back(make_sop(&e.op, e.grad, e.param));
    PlanOptions o = opt;
    o.fusion = true;  // adjoint always runs as tiled sweeps
    schedule(sops, n, o, prog);
}

std::string describe_program(const Program& prog)
{
    std::ostringstream ss;
    size_t n_sweeps = 0, n_stand = 0;
    for (const Step& st : prog.steps) (st.kind == Step::SWEEP ? n_sweeps : n_stand)++;
    ss << "n=" << prog.n << " steps=" << prog.steps.size() << " sweeps=" << n_sweeps
       << " standalone=" << n_stand << " phases=" << prog.phases.size() << " dops=" << prog.ops.size()
       << " lowered=" << prog.n_lowered << "\n";
    for (const Step& st : prog.steps) {
        if (st.kind == Step::SWEEP) {
            const SweepInfo& sw = prog.sweeps[st.index];
            ss << "  sweep k=" << sw.k << " phases=" << (sw.phase_end - sw.phase_begin) << " ops=";
            size_t no = 0;
            for (uint32_t p = sw.phase_begin; p < sw.phase_end; p++)
                no += prog.phases[p].op_end - prog.phases[p].op_begin;
            ss << no << " ext=[";
            for (uint32_t e = 0; e < sw.n_ext; e++) ss << (e ? "," : "") << int(sw.ext[e]);
            ss << "]\n";
        } else {
            const StandaloneOp& so = prog.standalone[st.index];
            ss << "  standalone kind=" << so.kind << " q=" << so.q << "\n";
        }
    }
    return ss.str();
}

}  // namespace svmi
