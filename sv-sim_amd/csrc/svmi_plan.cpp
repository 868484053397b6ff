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
#include <cmath>
#include <deque>
#include <cstring>
#include <sstream>
#include <stdexcept>

namespace svmi {

namespace {

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

uint64_t op_mask(const LOp& op)
{
    uint64_t m = 0;
    for (int i = 0; i < op.k; i++) m |= uint64_t(1) << op.q[i];
    return m;
}

bool is_register_perm(const LOp& op);

// FMA cost per amplitude (for the sweep budget).
long op_cost(const LOp& op, bool grad)
{
    if (op.kind == LOp::DIAG) return grad ? 6 : 4;
    if (op.kind == LOp::SRN) return 4;
    if (!grad && is_register_perm(op)) return 1;
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

// max_q: widest dense op a sweep may run in registers (wider: full-state pass)
SOp make_sop(const LOp* op, bool grad, int param, int max_q = kRegBits)
{
    SOp s;
    s.op = op;
    s.grad = grad;
    s.param = param;
    s.mask = op_mask(*op);
    // The reference SRN (svsim_cpu_omp.hpp:866-879) is only real-linear: it
    // does not commute with complex phases even on other qubits.
    // A controlled op must also stay ordered with every op on its controls.
    s.dep = op->kind == LOp::SRN ? ~uint64_t(0) : (s.mask | op->ctrl);
    s.diag = op->kind == LOp::DIAG;
    s.standalone = op->kind == LOp::DENSE && op->k > max_q && !grad;
    s.identity = !grad && op->dyn < 0 && is_identity(*op);
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

// The permutations the device runs as register moves (DOP_X, DOP_CX,
// DOP_SWAP): local index bit i <-> op.q[i].
bool is_register_perm(const LOp& op)
{
    if (op.kind != LOp::DENSE) return false;
    if (op.k == 1) {
        const int x[2] = {1, 0};
        return is_perm(op, x);
    }
    if (op.k == 2) {
        const int cx01[4] = {0, 3, 2, 1}, cx10[4] = {0, 1, 3, 2}, sw[4] = {0, 2, 1, 3};
        return is_perm(op, cx01) || is_perm(op, cx10) || is_perm(op, sw);
    }
    return false;
}

// diag whose entries are all 1 except the all-ones index
bool is_cphase(const SOp& s)
{
    const LOp& op = *s.op;
    if (op.kind != LOp::DIAG || op.dyn >= 0) return false;
    const size_t D = size_t(1) << op.k;
    for (size_t r = 0; r + 1 < D; r++)
        if (op.m[r].re != 1.0 || op.m[r].im != 0.0) return false;
    return true;
}

struct Encoder {
    Program& prog;
    uint32_t n_xf = 0;   // factor slots used in the current sweep
    explicit Encoder(Program& p) : prog(p) {}

    // ops [i, j) are controlled phases; emit one DOP_PHASESET per distinct
    // register condition
    void encode_phasesets(const std::vector<const SOp*>& ops, size_t i, size_t j, const int* tloc,
                          const int* slot_of)
    {
        std::vector<std::pair<uint32_t, std::vector<PhaseEntry>>> groups;
        for (size_t t = i; t < j; t++) {
            const LOp& op = *ops[t]->op;
            uint32_t rbits = 0;
            PhaseEntry e;
            std::memset(&e, 0, sizeof(e));
            cond_bits(op_mask(op) | op.ctrl, tloc, slot_of, rbits, e.gmask, e.emask);
            const c64 ph = op.m[(size_t(1) << op.k) - 1];
            e.ph_re = ph.re;
            e.ph_im = ph.im;
            uint32_t rsel = 0;
            for (int r = 0; r < kRegs; r++)
                if ((uint32_t(r) & rbits) == rbits) rsel |= 1u << r;
            size_t g = 0;
            while (g < groups.size() && groups[g].first != rsel) g++;
            if (g == groups.size()) groups.push_back({rsel, {}});
            groups[g].second.push_back(e);
        }
        for (auto& gr : groups) {
            // entries whose condition is one group bit, or one qubit outside
            // the tile, become nibble tables (DOP_PHASETAB); the rest stay a
            // list evaluated per lane
            std::vector<PhaseEntry> rest;
            encode_phasetab(gr.first, gr.second, tloc, slot_of, rest);
            if (rest.empty()) continue;
            gr.second.swap(rest);
            // entries with no condition inside the tile depend on the tile
            // only: their product is formed once per tile (factor slot)
            std::vector<PhaseEntry> lane, uni;
            for (const PhaseEntry& e : gr.second) (e.gmask == 0 ? uni : lane).push_back(e);
            if (uni.empty() || n_xf >= uint32_t(kMaxPhaseFactors)) {
                lane = gr.second;
                uni.clear();
            }
            DOp d;
            std::memset(&d, 0, sizeof(d));
            d.targets = prog.n_states == 2 ? DOP_TO_BOTH : DOP_TO_PSI;
            d.h.type = DOP_PHASESET;
            d.h.rsel = gr.first;
            d.h.s0 = uint32_t(lane.size());
            d.h.s1 = uint32_t(uni.size());
            d.h.grad = uni.empty() ? 0u : n_xf++;
            d.h.coef = uint32_t(prog.coef.size());
            for (const std::vector<PhaseEntry>* v : {&lane, &uni})
                for (const PhaseEntry& e : *v) {
                    c64 two[2];
                    std::memcpy(two, &e, sizeof(e));
                    prog.coef.push_back(two[0]);
                    prog.coef.push_back(two[1]);
                }
            prog.ops.push_back(d);
        }
    }

    // DOP_PHASETAB for the entries of one register condition (rsel) that
    // nibble tables can hold; the others are returned in `rest`.  Group bit
    // m (the m-th tile bit outside the registers, ascending) is bit m % 4 of
    // lane table m / 4; the m-th qubit outside the tile (ascending) is bit
    // m % 4 of external table m / 4; an entry without any condition outside
    // the registers multiplies every entry of the first table.
    void encode_phasetab(uint32_t rsel, const std::vector<PhaseEntry>& in, const int* tloc, const int* slot_of,
                         std::vector<PhaseEntry>& rest)
    {
        int k = 0;
        for (int p = 0; p < prog.n; p++)
            if (tloc[p] >= 0) k++;
        int gidx[64], xidx[64];
        int ng = 0, nx = 0;
        for (int l = 0; l < 64; l++) gidx[l] = (l < k && slot_of[l] < 0) ? ng++ : -1;
        for (int p = 0; p < 64; p++) xidx[p] = (p < prog.n && tloc[p] < 0) ? nx++ : -1;
        std::vector<PhaseEntry> lane, uni, cst;
        for (const PhaseEntry& e : in) {
            const int pg = __builtin_popcount(e.gmask), px = __builtin_popcountll(e.emask);
            if (pg == 1 && px == 0) lane.push_back(e);
            else if (pg == 0 && px == 1) uni.push_back(e);
            else if (pg == 0 && px == 0) cst.push_back(e);
            else rest.push_back(e);
        }
        const bool has_x = !uni.empty() || (!cst.empty() && lane.empty());
        if (lane.size() + uni.size() + cst.size() < 2 || (has_x && (nx == 0 || n_xf >= uint32_t(kMaxPhaseFactors))) ||
            (!cst.empty() && lane.empty() && nx == 0)) {
            rest = in;
            return;
        }
        int nlt = 0, nxt = 0;
        for (const PhaseEntry& e : lane) nlt = std::max(nlt, gidx[__builtin_ctz(e.gmask)] / 4 + 1);
        for (const PhaseEntry& e : uni) nxt = std::max(nxt, xidx[__builtin_ctzll(e.emask)] / 4 + 1);
        if (!cst.empty() && nlt == 0 && nxt == 0) nxt = 1;
        if (nlt > kPhaseTabMaxLane || nxt > kPhaseTabMaxExt) {
            rest = in;
            return;
        }
        std::vector<c64> T(size_t(16) * (nlt + nxt), c64{1.0, 0.0});
        auto mul_bit = [&](size_t table, int bit, double re, double im) {
            for (int v = 0; v < 16; v++)
                if ((v >> bit) & 1) T[table * 16 + v] = cmul(T[table * 16 + v], c64{re, im});
        };
        for (const PhaseEntry& e : lane) {
            const int m = gidx[__builtin_ctz(e.gmask)];
            mul_bit(size_t(m / 4), m % 4, e.ph_re, e.ph_im);
        }
        for (const PhaseEntry& e : uni) {
            const int m = xidx[__builtin_ctzll(e.emask)];
            mul_bit(size_t(nlt + m / 4), m % 4, e.ph_re, e.ph_im);
        }
        for (const PhaseEntry& e : cst)
            for (int v = 0; v < 16; v++) T[v] = cmul(T[v], c64{e.ph_re, e.ph_im});
        DOp d;
        std::memset(&d, 0, sizeof(d));
        d.targets = prog.n_states == 2 ? DOP_TO_BOTH : DOP_TO_PSI;
        d.h.type = DOP_PHASETAB;
        d.h.rsel = rsel;
        d.h.s0 = uint32_t(nlt);
        d.h.s1 = nxt ? 1u : 0u;
        d.h.smask = uint32_t(nxt);
        d.h.grad = nxt ? n_xf++ : 0u;
        d.h.gmask = kLaneTabGlobal;   // the sweep may stage the lane tables in LDS
        d.h.coef = push_coef(T);
        prog.ops.push_back(d);
    }

    // Split the qubits of a condition ("all these bits are 1") into register
    // slots, other tile-local bits and bits outside the tile.
    static void cond_bits(uint64_t bits, const int* tloc, const int* slot_of, uint32_t& rbits, uint32_t& gmask,
                          uint64_t& emask)
    {
        for (; bits; bits &= bits - 1) {
            const int p = __builtin_ctzll(bits);
            const int l = tloc[p];
            if (l < 0) emask |= uint64_t(1) << p;
            else if (slot_of[l] >= 0) rbits |= 1u << slot_of[l];
            else gmask |= 1u << l;
        }
    }

    static uint32_t rsel_of(uint32_t rbits)
    {
        uint32_t rsel = 0;
        for (int r = 0; r < kRegs; r++)
            if ((uint32_t(r) & rbits) == rbits) rsel |= 1u << r;
        return rsel;
    }

    uint32_t push_coef(const std::vector<c64>& v)
    {
        const uint32_t off = uint32_t(prog.coef.size());
        prog.coef.insert(prog.coef.end(), v.begin(), v.end());
        return off;
    }

    // A gradient op's K = U^dagger dU: components within 1e-15 of zero are
    // rounding residue of exact zeros (K of RY is real antisymmetric, of a
    // phase gate one imaginary entry) and are stored as zero, which the
    // device reduction skips.
    uint32_t push_grad_coef(std::vector<c64> v)
    {
        for (c64& x : v) {
            if (std::fabs(x.re) < 1e-15) x.re = 0.0;
            if (std::fabs(x.im) < 1e-15) x.im = 0.0;
        }
        return push_coef(v);
    }

    // DOP_GRAD_DENSE's component mask: bit 2e for Re, 2e + 1 for Im of entry e
    // (gradient ops act on at most two register slots: n <= 16 entries)
    uint32_t nonzero_mask(uint32_t off, uint32_t n) const
    {
        if (n > 16) throw std::logic_error("gradient op wider than two register slots");
        uint32_t nz = 0;
        for (uint32_t e = 0; e < n; e++) {
            if (prog.coef[off + e].re != 0.0) nz |= 1u << (2 * e);
            if (prog.coef[off + e].im != 0.0) nz |= 1u << (2 * e + 1);
        }
        return nz;
    }

    // Encode op for a phase: tloc[p] = tile-local bit of physical qubit p (or -1),
    // slot_of[l] = register slot of local bit l (or -1).
    void encode(const SOp& s, const int* tloc, const int* slot_of, uint32_t grad_slot)
    {
        const LOp& op = *s.op;
        DOp d;
        std::memset(&d, 0, sizeof(d));
        d.targets = prog.n_states == 2 ? DOP_TO_BOTH : DOP_TO_PSI;
        d.h.grad = grad_slot;
        if (op.kind == LOp::SRN) {
            d.h.type = DOP_SRN;
            d.h.s0 = uint32_t(slot_of[tloc[op.q[0]]]);
            prog.ops.push_back(d);
            return;
        }
        if (op.kind == LOp::DIAG && s.grad && op.k == 1 && op.ctrl == 0 && tloc[op.q[0]] >= 0 &&
            slot_of[tloc[op.q[0]]] >= 0) {
            // the gradient of a one-qubit diagonal on a register slot: a 2x2
            // K = diag(m0, m1) for the sparse dense-gradient reduction
            d.h.type = DOP_GRAD_DENSE;
            d.h.smask = 1u << slot_of[tloc[op.q[0]]];
            d.h.coef = push_grad_coef(std::vector<c64>{op.m[0], c64{0.0, 0.0}, c64{0.0, 0.0}, op.m[1]});
            d.h.s0 = nonzero_mask(d.h.coef, 4);
            prog.ops.push_back(d);
            return;
        }
        if (op.kind == LOp::DIAG) {
            const size_t D = size_t(1) << op.k;
            bool cphase = op.dyn < 0;
            for (size_t r = 0; r + 1 < D; r++)
                if (op.m[r].re != 1.0 || op.m[r].im != 0.0) cphase = false;
            // only controlled phases carry controls (the QIR lowering splits
            // a controlled diagonal into two of them)
            if (op.ctrl && (!cphase || s.grad)) throw std::logic_error("controlled diagonal that is not a phase");
            d.nq = uint32_t(op.k);
            for (int i = 0; i < op.k; i++) {
                d.qloc[i] = int32_t(tloc[op.q[i]]);
                d.qphys[i] = uint32_t(op.q[i]);
            }
            // CPHASE condition: every qubit of the op and every control is 1
            uint32_t rbits = 0;
            cond_bits(op_mask(op) | op.ctrl, tloc, slot_of, rbits, d.h.gmask, d.h.emask);
            d.h.rsel = rsel_of(rbits);
            if (cphase && !s.grad) {
                d.h.type = DOP_CPHASE;
                d.h.coef = push_coef(std::vector<c64>{op.m[D - 1]});
            } else {
                d.h.type = s.grad ? DOP_GRAD_DIAG : DOP_DIAG;
                d.h.coef = s.grad ? push_grad_coef(op.m) : push_coef(op.m);
                if (op.dyn >= 0) prog.measures.at(size_t(op.dyn)).coef = d.h.coef;
            }
            prog.ops.push_back(d);
            return;
        }
        if (op.ctrl) {
            // multi-controlled single-qubit gate
            if (op.k != 1 || s.grad) throw std::logic_error("controlled dense op wider than one qubit");
            uint32_t rbits = 0;
            cond_bits(op.ctrl, tloc, slot_of, rbits, d.h.gmask, d.h.emask);
            d.h.type = DOP_CDENSE;
            d.h.s0 = uint32_t(slot_of[tloc[op.q[0]]]);
            d.h.rsel = rsel_of(rbits);
            d.h.coef = push_coef(op.m);
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
                    d.h.type = DOP_X;
                    d.h.s0 = uint32_t(slots[0]);
                    prog.ops.push_back(d);
                    return;
                }
            } else if (op.k == 2) {
                // local index bit0 = q[0], bit1 = q[1]
                const int cx01[4] = {0, 3, 2, 1};   // control q0, target q1
                const int cx10[4] = {0, 1, 3, 2};   // control q1, target q0
                const int sw[4] = {0, 2, 1, 3};
                if (is_perm(op, cx01)) {
                    d.h.type = DOP_CX; d.h.s0 = uint32_t(slots[0]); d.h.s1 = uint32_t(slots[1]);
                    prog.ops.push_back(d);
                    return;
                }
                if (is_perm(op, cx10)) {
                    d.h.type = DOP_CX; d.h.s0 = uint32_t(slots[1]); d.h.s1 = uint32_t(slots[0]);
                    prog.ops.push_back(d);
                    return;
                }
                if (is_perm(op, sw)) {
                    d.h.type = DOP_SWAP;
                    d.h.s0 = uint32_t(std::min(slots[0], slots[1]));
                    d.h.s1 = uint32_t(std::max(slots[0], slots[1]));
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
        d.h.type = s.grad ? DOP_GRAD_DENSE : DOP_DENSE;
        d.h.smask = 0;
        for (int i = 0; i < op.k; i++) d.h.smask |= 1u << slots[i];
        d.h.coef = s.grad ? push_grad_coef(M) : push_coef(M);
        if (s.grad) d.h.s0 = nonzero_mask(d.h.coef, uint32_t(D * D));
        prog.ops.push_back(d);
    }
};

// ---------------------------------------------------------------------------
// Register-run merging.  Inside a phase, consecutive ops that act only on the
// 4 register slots (dense ops, and diagonal ops whose qubits are all register
// bits) are multiplied into one dense op on the union of their slots when
// that costs no more than running them one by one; every op run costs a fixed
// decode overhead on top of its arithmetic.  Ops that are not register-local
// (diagonals on other bits, gradient reads) commute with the run when they
// share no qubit with it and are emitted ahead of it; otherwise the run is
// flushed first.  SRN (real-linear) always flushes.
// ---------------------------------------------------------------------------

bool is_register_perm(const LOp& op);

// X, CX and SWAP run as register moves; other dense ops as matrix products.
// rot: the phase may run one-qubit ops as lifting rotations between merged
// register tables (about 3 per amplitude, a diagonal about 1).
long reg_cost(const LOp& op, bool rot)
{
    if (op.kind == LOp::DIAG) return rot ? 1 : 2;
    if (is_register_perm(op)) return 1;
    if (rot && op.k == 1) return 3;
    return 4L << op.k;
}

// T = E B where E is `op` embedded in the slot set S (matrix bit j <-> j-th
// lowest slot of S) and B is D x D, D = 2^|S|.  Row r of E has nonzeros only
// in the columns that differ from r in the op's bits, so each entry of T is a
// sum over those 2^k columns, taken in increasing column order -- the order a
// full D x D product would add the same (nonzero) terms in.
void left_multiply_in_slots(const LOp& op, const int* tloc, const int* slot_of, uint32_t S,
                            const std::vector<c64>& B, std::vector<c64>& T)
{
    int pos_of_slot[kRegBits];
    int q = 0;
    for (int s = 0; s < kRegBits; s++)
        if ((S >> s) & 1) pos_of_slot[s] = q++;
    const size_t D = size_t(1) << q;
    int bitpos[5];
    size_t opmask = 0;
    for (int i = 0; i < op.k; i++) {
        bitpos[i] = pos_of_slot[slot_of[tloc[op.q[i]]]];
        opmask |= size_t(1) << bitpos[i];
    }
    const size_t Dop = size_t(1) << op.k;
    std::vector<c64> A;
    lop_to_dense(op, A);
    // op-local column index cl -> offset in the D-space, in increasing offset order
    size_t order[32], offs[32];
    for (size_t cl = 0; cl < Dop; cl++) {
        size_t o = 0;
        for (int i = 0; i < op.k; i++)
            if ((cl >> i) & 1) o |= size_t(1) << bitpos[i];
        offs[cl] = o;
        order[cl] = cl;
    }
    std::sort(order, order + Dop, [&](size_t x, size_t y) { return offs[x] < offs[y]; });
    T.assign(D * D, c64{0.0, 0.0});
    for (size_t r = 0; r < D; r++) {
        size_t rl = 0;
        for (int i = 0; i < op.k; i++)
            if ((r >> bitpos[i]) & 1) rl |= size_t(1) << i;
        const size_t rest = r & ~opmask;
        for (size_t c = 0; c < D; c++) {
            c64 acc{0.0, 0.0};
            for (size_t t = 0; t < Dop; t++) {
                const size_t cl = order[t];
                acc = cmac(acc, A[rl * Dop + cl], B[(rest | offs[cl]) * D + c]);
            }
            T[r * D + c] = acc;
        }
    }
}

struct RunMerger {
    const int* tloc;
    const int* slot_of;
    const int* phys_of_slot;
    std::deque<LOp>& owned;
    std::vector<const SOp*>& out;
    std::deque<SOp>& sowned;
    int max_q;                 // widest merged op
    long overhead;             // per-op cost on top of its FMAs
    bool rot;                  // cost one-qubit ops as lifting rotations (reg_cost)
    std::vector<const SOp*> run;
    uint64_t run_mask = 0;

    bool reg_local(const SOp* s) const
    {
        if (s->grad || s->op->kind == LOp::SRN || s->op->ctrl || s->op->dyn >= 0) return false;
        for (int i = 0; i < s->op->k; i++) {
            const int l = tloc[s->op->q[i]];
            if (l < 0 || slot_of[l] < 0) return false;
        }
        return true;
    }

    void flush()
    {
        if (run.size() <= 1) {
            for (const SOp* s : run) out.push_back(s);
            run.clear();
            run_mask = 0;
            return;
        }
        uint32_t S = 0;
        long separate = 0;
        for (const SOp* s : run) {
            for (int i = 0; i < s->op->k; i++) S |= 1u << slot_of[tloc[s->op->q[i]]];
            separate += reg_cost(*s->op, rot) + overhead;
        }
        const int q = __builtin_popcount(S);
        if (q > max_q || (4L << q) + overhead > separate) {
            for (const SOp* s : run) out.push_back(s);
        } else {
            const size_t D = size_t(1) << q;
            std::vector<c64> B(D * D, c64{0.0, 0.0}), T;
            for (size_t i = 0; i < D; i++) B[i * D + i] = c64{1.0, 0.0};
            for (const SOp* s : run) {
                left_multiply_in_slots(*s->op, tloc, slot_of, S, B, T);
                B.swap(T);
            }
            LOp m;
            m.k = q;
            int qi = 0;
            for (int sl = 0; sl < kRegBits; sl++)
                if ((S >> sl) & 1) m.q[qi++] = phys_of_slot[sl];
            bool diag = true;
            for (size_t r = 0; r < D && diag; r++)
                for (size_t c = 0; c < D; c++)
                    if (r != c && (B[r * D + c].re != 0.0 || B[r * D + c].im != 0.0)) {
                        diag = false;
                        break;
                    }
            if (diag) {
                m.kind = LOp::DIAG;
                m.m.resize(D);
                for (size_t r = 0; r < D; r++) m.m[r] = B[r * D + r];
            } else {
                m.kind = LOp::DENSE;
                m.m.swap(B);
            }
            owned.push_back(std::move(m));
            sowned.push_back(make_sop(&owned.back(), false, -1));
            out.push_back(&sowned.back());
        }
        run.clear();
        run_mask = 0;
    }

    void add(const SOp* s)
    {
        if (reg_local(s)) {
            run.push_back(s);
            run_mask |= s->mask;
            return;
        }
        if ((s->dep & run_mask) || s->op->kind == LOp::SRN) flush();
        out.push_back(s);
    }
};

// ---------------------------------------------------------------------------
// Rotation/table form of a one-state phase.  A one-qubit unitary U on a
// register slot is D_a . [[c, -s], [s, c]] . D_b with D_a, D_b diagonal (the
// ZYZ form, c, s >= 0); the real rotation runs as in-place lifting steps
// (DOP_LIFT2 when c >= s, with its scale diag(c, 1/c) moved into D_a;
// DOP_LIFT3 otherwise) and every diagonal factor -- the D parts, diagonal gates
// on register slots and on qubits outside the tile -- joins the register table
// (DOP_RTAB) of its run of commuting diagonals.  The phase then alternates
// "one table" and "every rotation now free of pending diagonals".  Ops kept as
// they are (permutations, wider dense blocks, diagonals on group bits,
// controlled and collapse ops, SRN) are scheduled between them in dependency
// order.  The caller keeps this form only when its cost is lower.
// ---------------------------------------------------------------------------

// Cost (FMA per amplitude + overhead per op) of an op kept as it is.
long kept_cost(const SOp& s, long overhead)
{
    const LOp& op = *s.op;
    if (s.grad) {
        if (op.kind == LOp::DIAG && op.k > 1) return 6 + overhead;
        // sparse reduction: two FMAs per pair and nonzero component of K
        const size_t D = size_t(1) << op.k;
        long nz = 0;
        for (size_t i = 0; i < (op.kind == LOp::DIAG ? D : D * D); i++)
            nz += (std::fabs(op.m[i].re) >= 1e-15) + (std::fabs(op.m[i].im) >= 1e-15);
        return std::max(1L, 2 * nz / long(D)) + overhead;
    }
    if (op.kind == LOp::DIAG || op.kind == LOp::SRN) return 4 + overhead;
    if (is_register_perm(op)) return 1 + overhead;
    return (4L << op.k) + (op.k == 1 ? 2 : 0) + overhead;   // 2x2: plus the copies at the join
}

struct PItem {
    enum Kind { OPAQUE = 0, DIAG = 1, LIFT = 2 } kind = OPAQUE;
    const SOp* s = nullptr;   // OPAQUE, and DIAG items taken from an op
    uint64_t mask = 0;        // physical qubits the item acts on
    uint64_t dep = 0;         // physical qubits it must stay ordered with
    bool diag = false;        // commutes with every other diagonal item
    int q = -1;               // DIAG item of the decomposition: qubit q, entries d
    c64 d[2];
    int slot = 0;             // LIFT
    bool three = false;
    double u = 0.0, l = 0.0;
};

bool items_conflict(const PItem& x, const PItem& y)
{
    if (x.diag && y.diag) return false;
    return (x.dep & y.dep) != 0;
}

// U (row-major 2x2) = diag(pa0, pa1) . [[c, -s], [s, c]] . diag(1, pb1) with
// c, s >= 0 and unit pa0, pa1, pb1; false if U is diagonal or not of that form.
bool zyz_split(const c64* U, c64& pa0, c64& pa1, c64& pb1, double& c, double& s)
{
    const double c0 = std::hypot(U[0].re, U[0].im), s0 = std::hypot(U[2].re, U[2].im);
    const double nrm = std::hypot(c0, s0);
    if (!(s0 > 0.0) || !(nrm > 0.5 && nrm < 2.0)) return false;
    c = c0 / nrm;
    s = s0 / nrm;
    pa0 = c0 > 0.0 ? c64{U[0].re / c0, U[0].im / c0} : c64{1.0, 0.0};
    pa1 = c64{U[2].re / s0, U[2].im / s0};
    // -U01 = pa0 s pb1
    const c64 m01{-U[1].re / s, -U[1].im / s};
    pb1 = c64{pa0.re * m01.re + pa0.im * m01.im, pa0.re * m01.im - pa0.im * m01.re};
    const double pn = std::hypot(pb1.re, pb1.im);
    if (!(pn > 0.5 && pn < 2.0)) return false;
    pb1 = c64{pb1.re / pn, pb1.im / pn};
    // the four entries back from the factors
    const c64 w[4] = {c64{pa0.re * c, pa0.im * c}, cmul(c64{-pa0.re * s, -pa0.im * s}, pb1),
                      c64{pa1.re * s, pa1.im * s}, cmul(c64{pa1.re * c, pa1.im * c}, pb1)};
    for (int i = 0; i < 4; i++)
        if (std::fabs(w[i].re - U[i].re) > 1e-13 || std::fabs(w[i].im - U[i].im) > 1e-13) return false;
    return true;
}

// One emitted op of a compiled phase: an op encoded as usual (s), or a
// ready-made LIFT / RTAB op with its coefficients.
struct PEmit {
    const SOp* s = nullptr;
    DOp d;
    std::vector<c64> coef;
    uint32_t flip = 0;   // kept controlled phase: register slots held as quarter turns there
};

// A register mask rsel in quarter-turn coordinates f: register r holds the
// physical register r ^ f.
uint32_t flip_rsel(uint32_t rsel, uint32_t f)
{
    uint32_t out = 0;
    for (uint32_t r = 0; r < uint32_t(kRegs); r++)
        if ((rsel >> (r ^ f)) & 1u) out |= 1u << r;
    return out;
}

// Quarter turns kept in the register coordinates (one-state phases).  A
// rotation R(phi) with cos < sin runs as three shears (DOP_LIFT3).  Written
// R(phi) = Q R(phi - pi/2) with Q = R(pi/2) it needs two (DOP_LIFT2 of
// phi - pi/2, whose cosine sin(phi) is the larger) and Q, which is not run:
// from then on the slot's registers hold Q^-1 of the physical pair
// (x0 = p1, x1 = -p0; f = the slots so held), and the ops after it are
// rewritten in those coordinates -- a table entry for register r is the
// physical entry of r ^ f; a DOP_LIFT2 (diag(1/c, c) R, its diag(c, 1/c)
// in the next table) becomes the same lift plus diag(c^2, 1/c^2); a
// DOP_LIFT3 (pure rotation) stays; a second quarter turn on a slot leaves -1
// on its pair.  Such diagonal factors wait for the next table; the phase's
// last table also takes the signs (-1)^popcount(r & f), and the scatter
// writes register r to the physical index r ^ f (PhaseInfo::sxor).  Kept
// controlled phases (DOP_CPHASE / phase products: diagonal, a register mask
// rsel) apply where register r ^ f meets their condition (PEmit::flip; the
// emitter rewrites rsel).  A phase is converted only if every op from its
// first three-shear rotation on is a table, a rotation or such a phase, and
// the last one is a table.  Returns f; the number of conversions (one FMA
// per amplitude each) goes to *saved.
uint32_t quarter_turns(std::vector<PEmit>& out, long* saved)
{
    *saved = 0;
    size_t i0 = 0;
    while (i0 < out.size() && (out[i0].s || out[i0].d.h.type != DOP_LIFT3)) i0++;
    if (i0 == out.size()) return 0;
    for (size_t j = i0; j < out.size(); j++) {
        if (out[j].s) {
            if (!is_cphase(*out[j].s) || out[j].s->grad || out[j].s->op->ctrl) return 0;
            continue;
        }
        const uint32_t t = out[j].d.h.type;
        if (t != DOP_LIFT2 && t != DOP_LIFT3 && t != DOP_RTAB) return 0;
    }
    if (out.back().s || out.back().d.h.type != DOP_RTAB) return 0;
    // dry run on copies: a lift on a slot with a factor still waiting aborts
    std::vector<PEmit> res(out.begin(), out.end());
    uint32_t f = 0, pend_mask = 0;
    c64 pend[kRegBits][2];
    for (int q = 0; q < kRegBits; q++) pend[q][0] = pend[q][1] = c64{1.0, 0.0};
    auto wait = [&](int q, double d0, double d1) {
        pend[q][0] = c64{pend[q][0].re * d0, pend[q][0].im * d0};
        pend[q][1] = c64{pend[q][1].re * d1, pend[q][1].im * d1};
        pend_mask |= 1u << q;
    };
    long conv = 0;
    for (size_t j = i0; j < res.size(); j++) {
        if (res[j].s) {   // kept controlled phase: diagonal, commutes with the waiting factors
            res[j].flip = f;
            continue;
        }
        DOpHdr& h = res[j].d.h;
        if (h.type == DOP_RTAB) {
            std::vector<c64>& T = res[j].coef;
            const std::vector<c64> T0 = T;
            for (size_t v = 0; v < T.size() / kRegs; v++)
                for (uint32_t r = 0; r < uint32_t(kRegs); r++) {
                    c64 e = T0[v * kRegs + (r ^ f)];
                    for (int q = 0; q < kRegBits; q++)
                        if ((pend_mask >> q) & 1u) e = cmul(e, pend[q][(r >> q) & 1u]);
                    T[v * kRegs + r] = e;
                }
            for (int q = 0; q < kRegBits; q++) pend[q][0] = pend[q][1] = c64{1.0, 0.0};
            pend_mask = 0;
            continue;
        }
        const int q = int(h.s0);
        if ((pend_mask >> q) & 1u) return 0;
        const double u = lift_u(h), l = lift_l(h);
        if (h.type == DOP_LIFT2) {
            if ((f >> q) & 1u) {
                const double c2 = u == 0.0 ? 1.0 : -l / u;   // u = -s/c, l = s c
                wait(q, c2, 1.0 / c2);
            }
            continue;
        }
        // DOP_LIFT3: u = (c - 1)/s, l = s; as two shears of phi - pi/2
        const double sn = l, cs = 1.0 + u * l;
        const double u2 = cs / sn, l2 = -cs * sn;
        h.type = DOP_LIFT2;
        lift_set(h, u2, l2);
        res[j].coef.assign(1, c64{u2, l2});
        wait(q, sn, 1.0 / sn);
        if ((f >> q) & 1u) wait(q, -1.0, -1.0);
        f ^= 1u << q;
        conv++;
    }
    // the scatter signs on the last table (pend_mask is 0 after it)
    std::vector<c64>& T = res.back().coef;
    for (size_t v = 0; v < T.size() / kRegs; v++)
        for (uint32_t r = 0; r < uint32_t(kRegs); r++)
            if (__builtin_popcount(r & f) & 1) T[v * kRegs + r] = c64{-T[v * kRegs + r].re, -T[v * kRegs + r].im};
    out.swap(res);
    *saved = conv;
    return f;
}

// Two-state phases run every rotation as three shears, because a gradient op
// between a DOP_LIFT2 and the table that completes it would read the pair
// scaled by diag(1/c, c).  Where the next op after a three-shear rotation
// with c >= s is that table, or only lifts on other slots lead to it (no
// gradient op, nothing else on the slot), the rotation runs as two shears
// and the table takes diag(c, 1/c) on its slot.  Returns the conversions
// (one FMA per amplitude each).
long two_state_lifts(std::vector<PEmit>& out)
{
    long conv = 0;
    for (size_t j = 0; j < out.size(); j++) {
        if (out[j].s || out[j].d.h.type != DOP_LIFT3) continue;
        const uint32_t slot = out[j].d.h.s0;
        size_t k = j + 1;
        while (k < out.size() && !out[k].s && (out[k].d.h.type == DOP_LIFT2 || out[k].d.h.type == DOP_LIFT3) &&
               out[k].d.h.s0 != slot)
            k++;
        if (k == out.size() || out[k].s || out[k].d.h.type != DOP_RTAB) continue;
        DOpHdr& h = out[j].d.h;
        const double u = lift_u(h), l = lift_l(h);
        const double sn = l, cs = 1.0 + u * l;   // u = (c - 1)/s, l = s
        if (!(cs >= sn) || !(cs > 0.0)) continue;
        const double u2 = -sn / cs, l2 = sn * cs;
        h.type = DOP_LIFT2;
        lift_set(h, u2, l2);
        out[j].coef.assign(1, c64{u2, l2});
        std::vector<c64>& T = out[k].coef;
        for (size_t e = 0; e < T.size(); e++) {
            const double f = ((e >> slot) & 1u) ? 1.0 / cs : cs;
            T[e] = c64{T[e].re * f, T[e].im * f};
        }
        conv++;
    }
    return conv;
}

// Compile the phase's op list; returns the cost of the result.  two_state:
// an adjoint phase (every state op acts on phi and lambda, gradient ops read
// both).  Its rotations all use three shears: DOP_LIFT2 leaves the state
// scaled by a real diagonal until the table that follows, and a gradient op
// scheduled in between would read the scaled amplitudes.  Every other item
// keeps the pair orthogonally or unitarily equivalent, so a gradient op
// (diagonal K) commutes with every diagonal item and (dense K) with every
// item on other qubits.
long compile_phase(const std::vector<const SOp*>& ops, const int* tloc, const int* slot_of, long overhead,
                   bool two_state, std::vector<PEmit>& out, uint32_t* flip)
{
    *flip = 0;
    out.clear();
    std::vector<PItem> items;
    items.reserve(ops.size() * 3);
    auto reg_slot = [&](int p) { return tloc[p] >= 0 ? slot_of[tloc[p]] : -1; };
    for (const SOp* s : ops) {
        const LOp& op = *s->op;
        PItem it;
        it.s = s;
        it.mask = s->mask;
        it.dep = s->dep;
        it.diag = s->diag;
        const bool plain = !s->grad && op.ctrl == 0 && op.dyn < 0;
        if (plain && op.kind == LOp::DIAG) {
            bool ok = true;   // every qubit a register slot or outside the tile
            for (int i = 0; i < op.k; i++)
                if (tloc[op.q[i]] >= 0 && reg_slot(op.q[i]) < 0) ok = false;
            if (ok) it.kind = PItem::DIAG;
            items.push_back(it);
            continue;
        }
        c64 pa0, pa1, pb1;
        double c = 0.0, sn = 0.0;
        if (plain && op.kind == LOp::DENSE && op.k == 1 && reg_slot(op.q[0]) >= 0 && !is_register_perm(op) &&
            zyz_split(op.m.data(), pa0, pa1, pb1, c, sn)) {
            const uint64_t qm = uint64_t(1) << op.q[0];
            PItem db;
            db.kind = PItem::DIAG;
            db.mask = db.dep = qm;
            db.diag = true;
            db.q = op.q[0];
            db.d[0] = c64{1.0, 0.0};
            db.d[1] = pb1;
            PItem lf;
            lf.kind = PItem::LIFT;
            lf.mask = lf.dep = qm;
            lf.slot = reg_slot(op.q[0]);
            PItem da = db;
            if (c >= sn && !two_state) {
                lf.u = -sn / c;
                lf.l = sn * c;
                da.d[0] = c64{pa0.re * c, pa0.im * c};
                da.d[1] = c64{pa1.re / c, pa1.im / c};
            } else {
                lf.three = true;
                lf.u = (c - 1.0) / sn;
                lf.l = sn;
                da.d[0] = pa0;
                da.d[1] = pa1;
            }
            items.push_back(db);
            items.push_back(lf);
            items.push_back(da);
            continue;
        }
        items.push_back(it);
    }

    long cost = 0;
    std::vector<char> done(items.size(), 0);
    size_t left = items.size();
    auto ready = [&](size_t i) {
        for (size_t j = 0; j < i; j++)
            if (!done[j] && items_conflict(items[j], items[i])) return false;
        return true;
    };
    auto emit_op = [&](const SOp* s) {
        PEmit e;
        e.s = s;
        out.push_back(std::move(e));
        cost += kept_cost(*s, overhead);
    };
    while (left) {
        const size_t left0 = left;
        // diagonals free of pending non-diagonal items: one table (+ ops kept)
        std::vector<size_t> tab;
        std::vector<const SOp*> kept;
        uint64_t emask = 0;
        for (size_t i = 0; i < items.size(); i++) {
            if (done[i] || !items[i].diag || !ready(i)) continue;
            done[i] = 1;
            left--;
            const PItem& it = items[i];
            if (it.kind != PItem::DIAG) {
                kept.push_back(it.s);
                continue;
            }
            uint64_t ext = 0;
            if (it.s)
                for (int j = 0; j < it.s->op->k; j++)
                    if (tloc[it.s->op->q[j]] < 0) ext |= uint64_t(1) << it.s->op->q[j];
            if (__builtin_popcountll(emask | ext) > kMaxTabExt) {
                kept.push_back(it.s);
                continue;
            }
            emask |= ext;
            tab.push_back(i);
        }
        if (!tab.empty()) {
            const int ne = __builtin_popcountll(emask);
            int epos[kMaxTabExt];
            {
                int j = 0;
                for (uint64_t m = emask; m; m &= m - 1) epos[j++] = __builtin_ctzll(m);
            }
            std::vector<c64> T(size_t(kRegs) << ne, c64{1.0, 0.0});
            for (size_t v = 0; v < (size_t(1) << ne); v++)
                for (int r = 0; r < kRegs; r++) {
                    c64 f{1.0, 0.0};
                    for (size_t i : tab) {
                        const PItem& it = items[i];
                        if (!it.s) {
                            f = cmul(f, it.d[(r >> reg_slot(it.q)) & 1]);
                            continue;
                        }
                        const LOp& op = *it.s->op;
                        size_t idx = 0;
                        for (int j = 0; j < op.k; j++) {
                            const int p = op.q[j];
                            int bit;
                            if (tloc[p] >= 0) {
                                bit = (r >> reg_slot(p)) & 1;
                            } else {
                                int e = 0;
                                while (epos[e] != p) e++;
                                bit = int(v >> e) & 1;
                            }
                            idx |= size_t(bit) << j;
                        }
                        f = cmul(f, op.m[idx]);
                    }
                    T[v * kRegs + size_t(r)] = f;
                }
            bool ident = true;
            for (const c64& x : T) ident = ident && x.re == 1.0 && x.im == 0.0;
            if (!ident) {
                PEmit e;
                std::memset(&e.d, 0, sizeof(e.d));
                e.d.targets = two_state ? DOP_TO_BOTH : DOP_TO_PSI;
                e.d.h.type = DOP_RTAB;
                e.d.h.emask = emask;
                e.coef.swap(T);
                out.push_back(std::move(e));
                cost += 4 + overhead;
            }
        }
        for (const SOp* s : kept) emit_op(s);
        // everything non-diagonal that no pending item precedes
        for (size_t i = 0; i < items.size(); i++) {
            if (done[i] || items[i].diag || !ready(i)) continue;
            done[i] = 1;
            left--;
            const PItem& it = items[i];
            if (it.kind == PItem::LIFT) {
                PEmit e;
                std::memset(&e.d, 0, sizeof(e.d));
                e.d.targets = two_state ? DOP_TO_BOTH : DOP_TO_PSI;
                e.d.h.type = it.three ? DOP_LIFT3 : DOP_LIFT2;
                e.d.h.s0 = uint32_t(it.slot);
                lift_set(e.d.h, it.u, it.l);
                e.coef.push_back(c64{it.u, it.l});
                out.push_back(std::move(e));
                cost += (it.three ? 3 : 2) + overhead;
            } else {
                emit_op(it.s);
            }
        }
        if (left == left0) throw std::logic_error("compile_phase: no item ready");
    }
    if (!two_state) {
        long saved = 0;
        *flip = quarter_turns(out, &saved);
        cost -= saved;
    } else {
        cost -= two_state_lifts(out);
    }
    // runs of [table] + lifts on distinct slots become one DOP_ROT (one
    // dispatch instead of one per op): coef = the four {u, l} pairs (zero for
    // absent slots), then the table; the emitter points grad at the table
    std::vector<PEmit> fused;
    fused.reserve(out.size());
    auto is_lift = [](const PEmit& e) { return !e.s && (e.d.h.type == DOP_LIFT2 || e.d.h.type == DOP_LIFT3); };
    for (size_t i = 0; i < out.size();) {
        const bool tab = !out[i].s && out[i].d.h.type == DOP_RTAB;
        size_t j = i + (tab ? 1 : 0);
        uint32_t slots = 0;
        while (j < out.size() && is_lift(out[j]) && !((slots >> out[j].d.h.s0) & 1u)) slots |= 1u << out[j++].d.h.s0;
        const size_t nl = j - i - (tab ? 1 : 0);
        if (nl == 0 || (nl == 1 && !tab)) {
            fused.push_back(std::move(out[i]));
            i++;
            continue;
        }
        PEmit e;
        std::memset(&e.d, 0, sizeof(e.d));
        e.d.targets = two_state ? DOP_TO_BOTH : DOP_TO_PSI;
        e.d.h.type = DOP_ROT;
        e.coef.assign(4, c64{0.0, 0.0});
        for (size_t t = i + (tab ? 1 : 0); t < j; t++) {
            const DOpHdr& lh = out[t].d.h;
            e.d.h.smask |= 1u << lh.s0;
            if (lh.type == DOP_LIFT3) e.d.h.s1 |= 1u << lh.s0;
            e.coef[lh.s0] = c64{lift_u(lh), lift_l(lh)};
        }
        if (tab) {
            e.d.h.s0 = 1;
            e.d.h.emask = out[i].d.h.emask;
            e.coef.insert(e.coef.end(), out[i].coef.begin(), out[i].coef.end());
        }
        cost -= long(j - i - 1) * overhead;
        fused.push_back(std::move(e));
        i = j;
    }
    out.swap(fused);
    return cost;
}

// The phase's leading register permutations become its gather map: after
// them register rho holds what the gather would have put in register P(rho),
// so the map g becomes g o P (X on slot s: gxor ^= gcol[s]; CX control c,
// target t: gcol[c] ^= gcol[t]; SWAP: exchange two columns) and the ops are
// skipped.
// The tile map at the start of a phase (col[b] = M(1 << b), b < k), read
// back from its register and group columns.
void phase_map_columns(const PhaseInfo& ph, int k, int rb, uint32_t* col)
{
    uint32_t R = 0;
    for (int s = 0; s < rb; s++) {
        R |= 1u << ph.r[s];
        col[ph.r[s]] = ph.scol[s];
    }
    int gi = 0;
    for (int l = 0; l < k; l++)
        if (!((R >> l) & 1u)) col[l] = ph.gm[gi++];
}

uint32_t map_apply(const uint32_t* col, int k, uint32_t v)
{
    uint32_t o = 0;
    for (int b = 0; b < k; b++)
        if ((v >> b) & 1u) o ^= col[b];
    return o;
}

// Columns of the inverse of a bijective GF(2)-linear map of k bits
// (Gauss-Jordan on its rows).
void map_inverse(const uint32_t* col, int k, uint32_t* inv)
{
    uint32_t row[32], aug[32];
    for (int i = 0; i < k; i++) {
        row[i] = 0;
        for (int b = 0; b < k; b++) row[i] |= ((col[b] >> i) & 1u) << b;
        aug[i] = 1u << i;
    }
    for (int c = 0; c < k; c++) {
        int piv = c;
        while (piv < k && !((row[piv] >> c) & 1u)) piv++;
        if (piv == k) throw std::logic_error("tile map: not invertible");
        std::swap(row[c], row[piv]);
        std::swap(aug[c], aug[piv]);
        for (int i = 0; i < k; i++)
            if (i != c && ((row[i] >> c) & 1u)) {
                row[i] ^= row[c];
                aug[i] ^= aug[c];
            }
    }
    for (int j = 0; j < k; j++) {
        inv[j] = 0;
        for (int i = 0; i < k; i++) inv[j] |= ((aug[i] >> j) & 1u) << i;
    }
}

// A sweep cut into parts reloads the tile in the plain layout at every part
// after the first: the part's phases address LDS through N o M_p, N the
// inverse of the map at the part's first phase, and its store through
// N o (map at the part's end).  The first part loads the sweep's input, to
// which the maps already refer (N = identity; folds made before the sweep's
// first phase stay in its map).
void rebase_part(Program& prog, SweepInfo& part, int rb, bool first, const uint32_t* end_col)
{
    const int k = int(part.k);
    uint32_t start[32], N[32];
    if (first) {
        for (int l = 0; l < k; l++) N[l] = 1u << l;
    } else {
        phase_map_columns(prog.phases[part.phase_begin], k, rb, start);
        map_inverse(start, k, N);
    }
    part.mapped = 0;
    for (uint32_t p = part.phase_begin; p < part.phase_end; p++) {
        PhaseInfo& ph = prog.phases[p];
        for (int s = 0; s < rb; s++) {
            ph.scol[s] = map_apply(N, k, ph.scol[s]);
            ph.gcol[s] = map_apply(N, k, ph.gcol[s]);
        }
        for (int i = 0; i < k - rb; i++) ph.gm[i] = uint16_t(map_apply(N, k, ph.gm[i]));
        ph.gxor = map_apply(N, k, ph.gxor);
        ph.sxor = map_apply(N, k, ph.sxor);
        uint32_t col[32];
        phase_map_columns(ph, k, rb, col);
        ph.mapped = 0;
        for (int l = 0; l < k; l++) ph.mapped |= col[l] != (1u << l) ? 1u : 0u;
        part.mapped |= ph.mapped;
    }
    uint32_t fin[32];
    for (int l = 0; l < k; l++) {
        fin[l] = map_apply(N, k, end_col[l]);
        part.mapped |= fin[l] != (1u << l) ? 1u : 0u;
    }
    for (int l = 0; l < 14; l++) part.fcol[l] = 0;
    for (int j = 0; j < 16; j++) part.fj[j] = 0;
    if (part.mapped) {
        for (int l = 0; l < k; l++) part.fcol[l] = uint16_t(fin[l]);
        for (uint32_t j = 0; j < (1u << rb); j++) part.fj[j] = uint16_t(map_apply(fin + (k - rb), rb, j));
    }
}

void fold_leading_permutations(const Program& prog, PhaseInfo& pi)
{
    for (int s = 0; s < kRegBits; s++) pi.gcol[s] = pi.scol[s];
    pi.gxor = 0;
    while (pi.op_begin < pi.op_end) {
        const DOpHdr& h = prog.ops[pi.op_begin].h;
        if (h.type == DOP_X) pi.gxor ^= pi.gcol[h.s0];
        else if (h.type == DOP_CX) pi.gcol[h.s0] ^= pi.gcol[h.s1];
        else if (h.type == DOP_SWAP) std::swap(pi.gcol[h.s0], pi.gcol[h.s1]);
        else break;
        pi.op_begin++;
    }
}

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

void emit_sweep(Program& prog, const std::vector<const SOp*>& taken, uint64_t T, int n, int c,
                const PlanOptions& opt)
{
    Encoder enc(prog);
    const int rb = prog.n_states == 2 ? kRegBitsAdj : kRegBits;   // register bits of a phase
    std::deque<LOp> owned;     // merged ops (live until encoded)
    std::deque<SOp> sowned;
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
    sw.op_begin = uint32_t(prog.ops.size());
    sw.hitab = uint32_t(prog.hitab.size());
    sw.n_ext = 0;
    sw.extmask = 0;
    for (int p = 0; p < n; p++)
        if (!((T >> p) & 1)) {
            sw.extmask |= uint64_t(1) << p;
            sw.n_ext++;
        }
    sw.n_tiles = uint64_t(1) << (n - k);
    for (uint64_t h = 0; h < (uint64_t(1) << (k - c)); h++) {
        uint64_t off = 0;
        for (int j = 0; j < k - c; j++)
            if ((h >> j) & 1) off |= uint64_t(1) << tpos[c + j];
        prog.hitab.push_back(off);
    }
    sw.grad_base = prog.n_grad;

    // phase scheduling: pick the register bits R of each phase by a small
    // search that maximises the work the phase can take, then take it
    struct Pending {
        const SOp* s;
        uint32_t lm;      // tile-local bits the op touches
        bool local;       // every qubit of the op is in the tile
        int fold;         // tile-map candidate: 1 CX (control fc, target ft), 2 SWAP, 0 neither
        int fc, ft;       // tile-local bits of a candidate
    };
    // Tile map of a sweep: the amplitude of tile-local index x sits at LDS
    // index M(x) = XOR of mcol[b] over the set bits b of x.  A CX or SWAP on
    // two tile bits that are not both register bits of the phase taking it is
    // not run: after the phase, M becomes M o P (CX: mcol[c] ^= mcol[t]; SWAP:
    // two columns exchanged), and later phases and the tile's store address
    // LDS through M.
    const bool fold_map = opt.tile_map;
    uint32_t mcol[32];
    for (int l = 0; l < k; l++) mcol[l] = 1u << l;
    std::vector<Pending> remaining;
    remaining.reserve(taken.size());
    uint64_t dep_all = 0;
    for (const SOp* s : taken) {
        Pending p{s, 0u, true, 0, -1, -1};
        for (int i = 0; i < s->op->k; i++) {
            const int l = tloc[s->op->q[i]];
            if (l >= 0) p.lm |= 1u << l;
            else p.local = false;
        }
        if (fold_map && p.local && !s->grad && !s->diag && s->op->k == 2 && s->op->kind == LOp::DENSE) {
            const int cx01[4] = {0, 3, 2, 1}, cx10[4] = {0, 1, 3, 2}, sw4[4] = {0, 2, 1, 3};
            const int l0 = tloc[s->op->q[0]], l1 = tloc[s->op->q[1]];
            if (is_perm(*s->op, cx01)) p.fold = 1, p.fc = l0, p.ft = l1;
            else if (is_perm(*s->op, cx10)) p.fold = 1, p.fc = l1, p.ft = l0;
            else if (is_perm(*s->op, sw4)) p.fold = 2, p.fc = l0, p.ft = l1;
        }
        dep_all |= s->dep;
        remaining.push_back(p);
    }
    uint32_t n_grad_sweep = 0;
    // ops a phase with fixed register bits R takes; returns a score.  Without
    // output lists the scan stops once every qubit is blocked (nothing later
    // can be taken or change the score).  Tile-map folds taken by the phase
    // (folds) act after it: later ops on their qubits wait for the next
    // phase, except further folds, which compose in order.
    auto take_with = [&](uint32_t R, std::vector<const SOp*>* ph, std::vector<Pending>* rest,
                         std::vector<Pending>* folds) {
        uint64_t blocked = 0, fblocked = 0;
        long score = 0;
        for (const Pending& p : remaining) {
            const SOp* s = p.s;
            if ((s->dep & blocked) || (!p.fold && (s->dep & fblocked))) {
                blocked |= s->dep;
                if (rest) rest->push_back(p);
                else if (blocked == dep_all) break;
                continue;
            }
            const uint32_t lm = p.lm;
            if (p.fold && ((lm & ~R) != 0 || (s->dep & fblocked))) {
                if (folds) folds->push_back(p);
                score += 8 * (4L << s->op->k);
                fblocked |= s->dep;
                continue;
            }
            // (with a tile map, a diagonal with a gradient waits until its tile
            // qubits are register bits: on group bits its gradient op costs a
            // full reduction per op where register tables share one)
            if (s->diag && !(fold_map && s->grad && (lm & ~R) != 0)) {
                if (ph) ph->push_back(s);
                if (p.local && (lm & ~R) == 0) score += 1;               // merges into a run
                else if (lm & R) score -= 1;                             // splits a run
                continue;
            }
            if ((lm & ~R) == 0) {
                if (ph) ph->push_back(s);
                score += 8 * (4L << s->op->k);                           // progress first (any op
                                                                         // taken saves a phase later)
            } else {
                blocked |= s->dep;
                if (rest) rest->push_back(p);
                else if (blocked == dep_all) break;
            }
        }
        return score;
    };
    while (!remaining.empty()) {
        dep_all = 0;
        for (const Pending& p : remaining) dep_all |= p.s->dep;
        uint32_t R = 0;  // local-bit mask
        for (const Pending& p : remaining)
            if (!p.s->diag && !p.fold) {
                R = p.lm;
                break;
            }
        // candidate bits: tile bits of the first pending ops
        uint32_t cand = 0;
        for (size_t i = 0; i < remaining.size() && i < 96; i++) cand |= remaining[i].lm;
        while (popc(R) < rb) {
            long best = -(1L << 60);
            int best_b = -1;
            for (int b = 0; b < k; b++) {
                if (((R >> b) & 1) || !((cand >> b) & 1)) continue;
                const long sc = take_with(R | (1u << b), nullptr, nullptr, nullptr);
                if (sc > best) {
                    best = sc;
                    best_b = b;
                }
            }
            if (best_b < 0) break;
            R |= 1u << best_b;
        }
        std::vector<const SOp*> ph;
        std::vector<Pending> rest, folds;
        take_with(R, &ph, &rest, &folds);
        auto apply_folds = [&]() {
            for (const Pending& f : folds) {
                if (f.fold == 1) mcol[f.fc] ^= mcol[f.ft];
                else std::swap(mcol[f.fc], mcol[f.ft]);
            }
        };
        if (ph.empty() && !folds.empty()) {   // only relabelling: no phase
            apply_folds();
            remaining.swap(rest);
            continue;
        }
        // pad R to rb bits with the lowest free tile bits above the low qubits first
        for (int l = 0; l < k && popc(R) < rb; l++)
            if (!((R >> l) & 1) && l >= c) R |= 1u << l;
        for (int l = 0; l < k && popc(R) < rb; l++)
            if (!((R >> l) & 1)) R |= 1u << l;
        PhaseInfo pi;
        std::memset(&pi, 0, sizeof(pi));
        int slot_of[64];
        for (int l = 0; l < 64; l++) slot_of[l] = -1;
        int ns = 0;
        for (int l = 0; l < k; l++)
            if ((R >> l) & 1) {
                pi.r[ns] = uint32_t(l);
                slot_of[l] = ns++;
            }
        int phys_of_slot[kRegBits];
        for (int sl = 0; sl < rb; sl++) phys_of_slot[sl] = tpos[pi.r[sl]];
        if (rb < kRegBits) {   // two states: the last register bit selects lambda
            pi.r[kRegBits - 1] = uint32_t(k);
            phys_of_slot[kRegBits - 1] = -1;
        }
        // the tile map at this phase: its register and group columns
        for (int sl = 0; sl < rb; sl++) pi.scol[sl] = mcol[pi.r[sl]];
        if (rb < kRegBits) pi.scol[kRegBits - 1] = 1u << k;
        {
            int gi = 0;
            for (int l = 0; l < k; l++)
                if (!((R >> l) & 1)) {
                    if (gi >= 10) throw std::logic_error("tile map: more than 10 group bits");
                    pi.gm[gi++] = uint16_t(mcol[l]);
                }
            for (int l = 0; l < k; l++) pi.mapped |= mcol[l] != (1u << l) ? 1u : 0u;
        }
        std::vector<const SOp*> merged;
        {
            RunMerger rm{tloc, slot_of, phys_of_slot, owned, merged, sowned,
                         std::min(prog.n_states == 2 ? kMaxDenseQubitsAdj : kRegBits, opt.merge_qubits),
                         opt.op_overhead, false, {}, 0};
            for (const SOp* s : ph) rm.add(s);
            rm.flush();
        }
        pi.op_begin = uint32_t(prog.ops.size());
        if (opt.rotations) {
            const bool two = prog.n_states == 2;
            std::vector<const SOp*> merged_rot;
            RunMerger rm{tloc, slot_of, phys_of_slot, owned, merged_rot, sowned,
                         std::min(two ? kMaxDenseQubitsAdj : kRegBits, opt.merge_qubits), opt.op_overhead, true,
                         {}, 0};
            for (const SOp* s : ph) rm.add(s);
            rm.flush();
            std::vector<PEmit> em;
            uint32_t flip = 0;
            const long c_rot = compile_phase(merged_rot, tloc, slot_of, opt.op_overhead, two, em, &flip);
            long c_kept = 0;
            for (const SOp* s : merged) c_kept += kept_cost(*s, opt.op_overhead);
            if (c_rot < c_kept) {
                for (size_t i = 0; i < em.size();) {
                    if (!em[i].s) {
                        DOp d = em[i].d;
                        d.h.coef = enc.push_coef(em[i].coef);
                        if (d.h.type == DOP_ROT && d.h.s0) d.h.grad = d.h.coef + 4;
                        prog.ops.push_back(d);
                        i++;
                        continue;
                    }
                    const size_t first_op = prog.ops.size();
                    const uint32_t fl = em[i].flip;
                    size_t j = i;
                    while (j < em.size() && em[j].s && is_cphase(*em[j].s) && !em[j].s->grad && em[j].flip == fl) j++;
                    if (j - i >= 2) {
                        std::vector<const SOp*> run;
                        for (size_t t = i; t < j; t++) run.push_back(em[t].s);
                        enc.encode_phasesets(run, 0, run.size(), tloc, slot_of);
                        i = j;
                    } else {
                        const SOp* s = em[i++].s;
                        uint32_t gslot = 0;
                        if (s->grad) {
                            gslot = n_grad_sweep++;
                            prog.grad_param.push_back(s->param);
                        }
                        enc.encode(*s, tloc, slot_of, gslot);
                    }
                    if (fl)   // controlled phases under quarter turns: their register masks
                        for (size_t o = first_op; o < prog.ops.size(); o++) {
                            DOpHdr& h = prog.ops[o].h;
                            if (h.type != DOP_CPHASE && h.type != DOP_PHASESET && h.type != DOP_PHASETAB)
                                throw std::logic_error("quarter turns: a kept op without a register mask");
                            h.rsel = flip_rsel(h.rsel, fl);
                        }
                }
                merged.clear();
                for (int q = 0; q < kRegBits; q++)
                    if ((flip >> q) & 1u) pi.sxor ^= pi.scol[q];
            }
        }
        for (size_t i = 0; i < merged.size();) {
            // a run of controlled-phase ops commutes; group it by register
            // condition into phase-product ops
            size_t j = i;
            while (j < merged.size() && is_cphase(*merged[j]) && !merged[j]->grad) j++;
            if (j - i >= 2) {
                enc.encode_phasesets(merged, i, j, tloc, slot_of);
                i = j;
                continue;
            }
            const SOp* s = merged[i++];
            uint32_t gslot = 0;
            if (s->grad) {
                gslot = n_grad_sweep++;
                prog.grad_param.push_back(s->param);
            }
            enc.encode(*s, tloc, slot_of, gslot);
        }
        pi.op_end = uint32_t(prog.ops.size());
        const uint32_t ops0 = pi.op_begin;
        fold_leading_permutations(prog, pi);
        if (fold_map && pi.op_begin == pi.op_end && pi.gxor == 0 && pi.sxor == 0) {
            // the phase only permutes its registers (linearly): instead of
            // moving the amplitudes, register bit r[s] of the tile map takes
            // the gather column the phase would have read it from
            for (int sl = 0; sl < rb; sl++) mcol[pi.r[sl]] = pi.gcol[sl];
            prog.ops.resize(ops0);
            apply_folds();
            remaining.swap(rest);
            continue;
        }
        prog.phases.push_back(pi);
        apply_folds();
        remaining.swap(rest);
    }
    sw.phase_end = uint32_t(prog.phases.size());
    // mapped: some phase or the store addresses LDS through a map (fcol = the
    // final one, which may be the identity)
    for (int l = 0; l < k; l++) sw.mapped |= mcol[l] != (1u << l) ? 1u : 0u;
    for (uint32_t p = sw.phase_begin; p < sw.phase_end; p++) sw.mapped |= prog.phases[p].mapped;
    if (sw.mapped) {
        for (int l = 0; l < k; l++) sw.fcol[l] = uint16_t(mcol[l]);
        for (uint32_t j = 0; j < (1u << rb); j++) {   // M(j << (k - rb)): the store's register part
            uint32_t m = 0;
            for (int b = 0; b < rb; b++)
                if ((j >> b) & 1u) m ^= mcol[k - rb + b];
            sw.fj[j] = uint16_t(m);
        }
    }
    // qubits that select register-table variants become the tile id's high
    // bits: workgroups running at the same time then read the same variants
    for (size_t o = sw.op_begin; o < prog.ops.size(); o++)
        if (prog.ops[o].h.type == DOP_RTAB || (prog.ops[o].h.type == DOP_ROT && prog.ops[o].h.s0))
            sw.exthi |= prog.ops[o].h.emask;
    sw.n_grad = n_grad_sweep;
    // lane tables of the sweep's DOP_PHASETAB ops: one contiguous copy that
    // every tile stages in LDS after its factor slots (if the area has room)
    sw.n_xf = enc.n_xf;
    {
        uint32_t ntab = 0;
        for (size_t o = sw.op_begin; o < prog.ops.size(); o++)
            if (prog.ops[o].h.type == DOP_PHASETAB) ntab += prog.ops[o].h.s0;
        if (ntab && sw.n_xf + 16u * ntab <= uint32_t(kMaxXfSlots)) {
            sw.lt_coef = uint32_t(prog.coef.size());
            sw.n_lt = ntab;
            uint32_t slot = sw.n_xf;
            for (size_t o = sw.op_begin; o < prog.ops.size(); o++) {
                DOpHdr& h = prog.ops[o].h;
                if (h.type != DOP_PHASETAB || h.s0 == 0) continue;
                const size_t src = h.coef;
                for (size_t i = 0; i < 16u * h.s0; i++) prog.coef.push_back(prog.coef[src + i]);
                h.gmask = slot;
                slot += 16u * h.s0;
            }
            sw.n_xf = slot;
        }
    }
    sw.op_end = uint32_t(prog.ops.size());
    prog.n_grad += n_grad_sweep;
    // The coefficients a tile reads while it runs the sweep go through the
    // small scalar cache that a few CUs share; when their total exceeds
    // opt.coef_budget, run the phases as several consecutive sweeps over the
    // same tile layout (one more HBM pass each, far fewer cache misses).
    // Only sweeps of many tiles are split: with few workgroups in flight
    // the cache holds their coefficients anyway.
    auto dense_width = [&](uint32_t o0, uint32_t o1) {
        uint32_t w = 0;
        for (uint32_t o = o0; o < o1; o++)
            if (prog.ops[o].h.type == DOP_DENSE) w = std::max(w, uint32_t(__builtin_popcount(prog.ops[o].h.smask)));
        return w;
    };
    sw.max_dense = dense_width(sw.op_begin, sw.op_end);
    if (prog.n_states == 1 && opt.coef_budget > 0 && sw.n_tiles >= 512 && sw.phase_end > sw.phase_begin) {
        // bytes a tile reads through the scalar cache per phase: every op
        // record (DOp stride) and the coefficients the op loads
        auto phase_bytes = [&](const PhaseInfo& ph) {
            long b = 0;
            for (uint32_t o = ph.op_begin; o < ph.op_end; o++) {
                const DOp& d = prog.ops[o];
                const DOpHdr& h = d.h;
                b += long(sizeof(DOp));
                switch (h.type) {
                case DOP_DENSE: b += 16L << (2 * __builtin_popcount(h.smask)); break;
                case DOP_PHASESET: b += 32L * (h.s0 + h.s1); break;
                case DOP_PHASETAB: break;   // tables read with vector loads
                case DOP_RTAB: b += 16L * kRegs; break;   // one variant per tile
                case DOP_ROT: b += 64L + (h.s0 ? 16L * kRegs : 0L); break;
                case DOP_DIAG: b += 16L << d.nq; break;
                case DOP_CDENSE: b += 64; break;
                case DOP_CPHASE: b += 16; break;
                default: break;
                }
            }
            return b;
        };
        uint32_t p0 = sw.phase_begin;
        long acc = 0;
        for (uint32_t p = sw.phase_begin; p <= sw.phase_end; p++) {
            const long pb = p < sw.phase_end ? phase_bytes(prog.phases[p]) : 0;
            if (p == sw.phase_end || (p > p0 && acc + pb > opt.coef_budget)) {
                SweepInfo part = sw;
                part.phase_begin = p0;
                part.phase_end = p;
                part.op_begin = prog.phases[p0].op_begin;
                part.op_end = prog.phases[p - 1].op_end;
                part.max_dense = dense_width(part.op_begin, part.op_end);
                if (sw.mapped) {   // the map where the part ends: at phase p's start, or the final one
                    uint32_t end_col[32];
                    if (p < sw.phase_end) phase_map_columns(prog.phases[p], k, rb, end_col);
                    else
                        for (int l = 0; l < k; l++) end_col[l] = mcol[l];
                    rebase_part(prog, part, rb, p0 == sw.phase_begin, end_col);
                }
                prog.steps.push_back(Step{Step::SWEEP, uint32_t(prog.sweeps.size())});
                prog.sweeps.push_back(part);
                p0 = p;
                acc = 0;
            }
            acc += pb;
        }
        return;
    }
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
    prog.n_lowered += pending.size();
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
            // the first pending op cannot go in a tile that keeps the low
            // qubits.  A controlled op (which has no full-state kernel) gets a
            // sweep of its own around its target; anything else runs alone.
            const SOp* s = pending.front();
            if (s->op->ctrl && popc(s->mask) <= kk) {
                uint64_t T1 = s->mask;
                for (int p = 0; p < n && popc(T1) < kk; p++) T1 |= uint64_t(1) << p;
                const int c1 = std::min(c, __builtin_ctzll(~T1));
                emit_sweep(prog, std::vector<const SOp*>{s}, T1, n, c1, opt);
            } else {
                emit_standalone(prog, *s);
            }
            pending.erase(pending.begin());
            continue;
        }
        // pad T to kk qubits with the lowest free positions
        for (int p = 0; p < n && popc(T) < kk; p++) T |= uint64_t(1) << p;
        emit_sweep(prog, taken, T, n, c, opt);
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
        const bool mergeable = op.k == 1 && op.kind != LOp::SRN && op.ctrl == 0 && op.dyn < 0;
        if (mergeable && last[op.q[0]] >= 0) {
            LOp merged;
            lop_compose_same(out[last[op.q[0]]], op, merged);
            out[last[op.q[0]]] = merged;
            continue;
        }
        out.push_back(op);
        if (op.kind == LOp::SRN) std::fill(last.begin(), last.end(), -1);  // real-linear only
        for (int i = 0; i < op.k; i++) last[op.q[i]] = -1;
        for (uint64_t c = op.ctrl; c; c &= c - 1) last[__builtin_ctzll(c)] = -1;
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

void build_segmented_program(const std::vector<std::vector<LOp>>& segments, const std::vector<MeasureStep>& measures,
                             int n, const PlanOptions& opt, Program& prog)
{
    prog = Program();
    prog.n = n;
    prog.n_states = 1;
    PlanOptions o = opt;
    o.fusion = true;   // controlled ops run inside sweeps only
    for (size_t i = 0; i < segments.size(); i++) {
        std::vector<SOp> sops;
        sops.reserve(segments[i].size());
        for (const LOp& op : segments[i]) sops.push_back(make_sop(&op, false, -1));
        schedule(sops, n, o, prog);
        if (i < measures.size()) {
            prog.steps.push_back(Step{Step::MEASURE, uint32_t(prog.measures.size())});
            prog.measures.push_back(measures[i]);
        }
    }
    for (const MeasureStep& m : prog.measures)
        if (m.coef == UINT32_MAX) throw std::logic_error("measurement without a collapse op");
}

void build_adjoint_program(const std::vector<AdjEntry>& entries, int n, const PlanOptions& opt,
                           Program& prog)
{
    prog = Program();
    prog.n_states = 2;
    std::vector<SOp> sops;
    sops.reserve(entries.size());
    for (const AdjEntry& e : entries) sops.push_back(make_sop(&e.op, e.grad, e.param, kMaxDenseQubitsAdj));
    PlanOptions o = opt;
    o.fusion = true;  // adjoint always runs as tiled sweeps
    schedule(sops, n, o, prog);
}

// FP64 multiply-adds per amplitude the program's state ops execute (a
// complex multiply counted as 4, a real shear as 1 per component pair...):
// the arithmetic the sweeps do per amplitude, summed over the program.
double program_fma_per_amp(const Program& prog)
{
    double f = 0.0;
    for (const Step& st : prog.steps) {
        if (st.kind != Step::SWEEP) continue;
        const SweepInfo& sw = prog.sweeps[st.index];
        for (uint32_t p = sw.phase_begin; p < sw.phase_end; p++)
            for (uint32_t o = prog.phases[p].op_begin; o < prog.phases[p].op_end; o++) {
                const DOpHdr& h = prog.ops[o].h;
                const double sel = __builtin_popcount(h.rsel & 0xffffu) / 16.0;   // share of registers
                switch (h.type) {
                case DOP_DENSE: f += 4.0 * (1 << __builtin_popcount(h.smask)); break;
                case DOP_LIFT2: f += 2.0; break;
                case DOP_LIFT3: f += 3.0; break;
                case DOP_ROT:
                    f += (h.s0 ? 4.0 : 0.0) + 2.0 * __builtin_popcount(h.smask) + __builtin_popcount(h.s1);
                    break;
                case DOP_RTAB:
                case DOP_DIAG: f += 4.0; break;
                case DOP_CPHASE:
                case DOP_PHASESET:
                case DOP_PHASETAB: f += 4.0 * sel; break;
                case DOP_CDENSE: f += 8.0 * sel; break;
                case DOP_SRN: f += 2.0; break;
                case DOP_GRAD_DENSE: {
                    const int D = 1 << __builtin_popcount(h.smask);
                    int nz = 0;
                    for (int i = 0; i < D * D; i++)
                        nz += (prog.coef[h.coef + i].re != 0.0) + (prog.coef[h.coef + i].im != 0.0);
                    f += 2.0 * nz / D;
                    break;
                }
                case DOP_GRAD_DIAG: f += 6.0; break;
                default: break;
                }
            }
    }
    return f;
}

std::string describe_program(const Program& prog)
{
    std::ostringstream ss;
    size_t n_sweeps = 0, n_stand = 0;
    size_t n_meas = 0;
    for (const Step& st : prog.steps)
        (st.kind == Step::SWEEP ? n_sweeps : st.kind == Step::MEASURE ? n_meas : n_stand)++;
    ss << "n=" << prog.n << " steps=" << prog.steps.size() << " sweeps=" << n_sweeps
       << " standalone=" << n_stand << " measure=" << n_meas << " phases=" << prog.phases.size() << " dops=" << prog.ops.size()
       << " lowered=" << prog.n_lowered << " fma_per_amp=" << program_fma_per_amp(prog) << "\n";
    for (const Step& st : prog.steps) {
        if (st.kind == Step::SWEEP) {
            const SweepInfo& sw = prog.sweeps[st.index];
            ss << "  sweep k=" << sw.k << " phases=" << (sw.phase_end - sw.phase_begin) << " ops=";
            size_t no = 0;
            for (uint32_t p = sw.phase_begin; p < sw.phase_end; p++)
                no += prog.phases[p].op_end - prog.phases[p].op_begin;
            ss << no << " ext=[";
            bool first = true;
            for (int p = 0; p < 64; p++)
                if ((sw.extmask >> p) & 1) {
                    ss << (first ? "" : ",") << p;
                    first = false;
                }
            ss << "]";
            if (sw.mapped) ss << " [m]";   // phases or the store address LDS through a tile map
            ss << "\n";
            for (uint32_t p = sw.phase_begin; p < sw.phase_end; p++) {
                const PhaseInfo& ph = prog.phases[p];
                ss << "    R=" << ph.r[0] << "," << ph.r[1] << "," << ph.r[2] << "," << ph.r[3];
                bool plain_gather = ph.gxor == 0;
                for (int b = 0; b < kRegBits; b++) plain_gather = plain_gather && ph.gcol[b] == ph.scol[b];
                if (!plain_gather) ss << " [g]";   // leading permutations folded into the gather
                if (ph.mapped) ss << " [m]";       // LDS addressed through a tile map
                if (ph.sxor) ss << " [q]";         // quarter turns: scatter map r ^ f
                ss << " :";
                for (uint32_t o = ph.op_begin; o < ph.op_end; o++) {
                    const DOp& d = prog.ops[o];
                    static const char* nm[] = {"-", "D", "X", "CX", "SW", "SRN", "P", "G", "gD", "gG", "PS", "CD",
                                               "L", "L3", "T", "PT", "R"};
                    ss << " " << nm[d.h.type];
                    if (d.h.type == DOP_PHASETAB) {
                        ss << d.h.s0;
                        if (d.h.s1) ss << "x" << d.h.smask;
                    }
                    if (d.h.type == DOP_PHASESET) {
                        ss << d.h.s0;
                        if (d.h.s1) ss << "u" << d.h.s1;
                    }
                    if (d.h.type == DOP_DENSE || d.h.type == DOP_GRAD_DENSE) ss << __builtin_popcount(d.h.smask);
                    if (d.h.type == DOP_RTAB && d.h.emask) ss << "e" << __builtin_popcountll(d.h.emask);
                    if (d.h.type == DOP_ROT) ss << (d.h.s0 ? "t" : "") << __builtin_popcount(d.h.smask);
                }
                ss << "\n";
            }
        } else if (st.kind == Step::MEASURE) {
            ss << "  measure q=" << prog.measures[st.index].qubit << "\n";
        } else {
            const StandaloneOp& so = prog.standalone[st.index];
            ss << "  standalone kind=" << so.kind << " q=" << so.q << "\n";
        }
    }
    return ss.str();
}

}  // namespace svmi
