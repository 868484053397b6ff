// svmi_exec.hpp -- execution of one device op on one register group.
// Compiled into the HIP kernels and into the host-side program checker
// (svmi_emulate_program), so both run literally the same op code.
#pragma once

#include "svmi_internal.hpp"

namespace svmi {

// Distinct no-op markers at both ends of every switch case keep LLVM from
// hoisting/sinking the cases into one access with a runtime register index,
// which would push the whole register group to scratch memory.
#if defined(__HIP_DEVICE_COMPILE__)
#define SVMI_FENCE(N) asm volatile("; svmi case %0" ::"i"(N))
#else
#define SVMI_FENCE(N) ((void)0)
#endif
#define SVMI_CASE_ID(LABEL, ID, STMT) \
    case LABEL:                         \
        SVMI_FENCE(ID);                 \
        STMT;                           \
        SVMI_FENCE(ID);                 \
        break;
#define SVMI_CASE(N, STMT) SVMI_CASE_ID(N, N, STMT)

constexpr int cpopc(int v) { return v ? (v & 1) + cpopc(v >> 1) : 0; }

// kMaxQ: widest dense op compiled in (the planner emits no wider one for the
// kernel instantiated with it)
template <int kMaxQ, int N>
SVMI_HD inline void dispatch_dense(int smask, c64 (&a)[N], const c64* __restrict__ M)
{
    switch (smask) {
#define SVMI_DENSE_CASE(S) SVMI_CASE(S, if constexpr (cpopc(S) <= kMaxQ) (apply_dense<S, N>(a, M)))
        SVMI_DENSE_CASE(1) SVMI_DENSE_CASE(2) SVMI_DENSE_CASE(3) SVMI_DENSE_CASE(4)
        SVMI_DENSE_CASE(5) SVMI_DENSE_CASE(6) SVMI_DENSE_CASE(7) SVMI_DENSE_CASE(8)
        SVMI_DENSE_CASE(9) SVMI_DENSE_CASE(10) SVMI_DENSE_CASE(11) SVMI_DENSE_CASE(12)
        SVMI_DENSE_CASE(13) SVMI_DENSE_CASE(14) SVMI_DENSE_CASE(15)
#undef SVMI_DENSE_CASE
    default: break;
    }
}

template <int kMaxQ, int H>
SVMI_HD inline double dispatch_grad_dense(int smask, const c64 (&pl)[2 * H], const c64* __restrict__ M, uint32_t nz)
{
    double sum = 0.0;
    switch (smask) {
#define SVMI_GRAD_CASE(S) \
    SVMI_CASE(S, if constexpr (cpopc(S) <= kMaxQ && (S & ~(H - 1)) == 0) sum = (grad_dense<S, H>(pl, M, nz)))
        SVMI_GRAD_CASE(1) SVMI_GRAD_CASE(2) SVMI_GRAD_CASE(3) SVMI_GRAD_CASE(4)
        SVMI_GRAD_CASE(5) SVMI_GRAD_CASE(6) SVMI_GRAD_CASE(7) SVMI_GRAD_CASE(8)
        SVMI_GRAD_CASE(9) SVMI_GRAD_CASE(10) SVMI_GRAD_CASE(11) SVMI_GRAD_CASE(12)
        SVMI_GRAD_CASE(13) SVMI_GRAD_CASE(14) SVMI_GRAD_CASE(15)
#undef SVMI_GRAD_CASE
    default: break;
    }
    return sum;
}

template <int N>
SVMI_HD inline void dispatch_x(int s, c64 (&a)[N])
{
    switch (s) {
        SVMI_CASE(0, (apply_x<0, N>(a)))
        SVMI_CASE(1, (apply_x<1, N>(a)))
        SVMI_CASE(2, (apply_x<2, N>(a)))
        SVMI_CASE(3, (apply_x<3, N>(a)))
    default: break;
    }
}

template <int N>
SVMI_HD inline void dispatch_cx(int c, int t, c64 (&a)[N])
{
    switch (c * 4 + t) {
        SVMI_CASE(1, (apply_cx<0, 1, N>(a)))
        SVMI_CASE(2, (apply_cx<0, 2, N>(a)))
        SVMI_CASE(3, (apply_cx<0, 3, N>(a)))
        SVMI_CASE(4, (apply_cx<1, 0, N>(a)))
        SVMI_CASE(6, (apply_cx<1, 2, N>(a)))
        SVMI_CASE(7, (apply_cx<1, 3, N>(a)))
        SVMI_CASE(8, (apply_cx<2, 0, N>(a)))
        SVMI_CASE(9, (apply_cx<2, 1, N>(a)))
        SVMI_CASE(11, (apply_cx<2, 3, N>(a)))
        SVMI_CASE(12, (apply_cx<3, 0, N>(a)))
        SVMI_CASE(13, (apply_cx<3, 1, N>(a)))
        SVMI_CASE(14, (apply_cx<3, 2, N>(a)))
    default: break;
    }
}

template <int N>
SVMI_HD inline void dispatch_swap(int s0, int s1, c64 (&a)[N])
{
    switch (s0 * 4 + s1) {
        SVMI_CASE(1, (apply_swap<0, 1, N>(a)))
        SVMI_CASE(2, (apply_swap<0, 2, N>(a)))
        SVMI_CASE(3, (apply_swap<0, 3, N>(a)))
        SVMI_CASE(6, (apply_swap<1, 2, N>(a)))
        SVMI_CASE(7, (apply_swap<1, 3, N>(a)))
        SVMI_CASE(11, (apply_swap<2, 3, N>(a)))
    default: break;
    }
}

template <int N>
SVMI_HD inline void dispatch_srn(int s, c64 (&a)[N])
{
    switch (s) {
        SVMI_CASE(0, (apply_srn<0, N>(a)))
        SVMI_CASE(1, (apply_srn<1, N>(a)))
        SVMI_CASE(2, (apply_srn<2, N>(a)))
        SVMI_CASE(3, (apply_srn<3, N>(a)))
    default: break;
    }
}

template <bool kThree, int N>
SVMI_HD inline void dispatch_lift(const DOpHdr& h, c64 (&a)[N])
{
    const double u = lift_u(h), l = lift_l(h);
    switch (h.s0) {
        SVMI_CASE_ID(0, 300 + 10 * kThree, (apply_lift<0, kThree, N>(a, u, l)))
        SVMI_CASE_ID(1, 301 + 10 * kThree, (apply_lift<1, kThree, N>(a, u, l)))
        SVMI_CASE_ID(2, 302 + 10 * kThree, (apply_lift<2, kThree, N>(a, u, l)))
        SVMI_CASE_ID(3, 303 + 10 * kThree, (apply_lift<3, kThree, N>(a, u, l)))
    default: break;
    }
}

template <int N>
SVMI_HD inline void dispatch_cdense(const DOpHdr& h, c64 (&a)[N], const c64* __restrict__ coef, uint32_t goff,
                                    uint64_t base)
{
    if ((base & h.emask) != h.emask) return;
    if ((goff & h.gmask) != h.gmask) return;
    const c64* M = coef + h.coef;
    switch (h.s0) {
        SVMI_CASE_ID(0, 200, (apply_cdense<0, N>(a, M, h.rsel)))
        SVMI_CASE_ID(1, 201, (apply_cdense<1, N>(a, M, h.rsel)))
        SVMI_CASE_ID(2, 202, (apply_cdense<2, N>(a, M, h.rsel)))
        SVMI_CASE_ID(3, 203, (apply_cdense<3, N>(a, M, h.rsel)))
    default: break;
    }
}

// Index of a diagonal op's table entry for tile-local index l.
// (fixed trip count so every index stays static after unrolling)
SVMI_HD inline uint32_t diag_index(const DOp& op, uint32_t l, uint32_t eidx)
{
    uint32_t idx = eidx;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int q = op.qloc[i];
        if (i < op.nq && q >= 0) idx |= ((l >> q) & 1u) << i;
    }
    return idx;
}

// Part of a diagonal op's table index fixed by the tile (qubits outside it).
SVMI_HD inline uint32_t diag_ext_index(const DOp& op, uint64_t base)
{
    uint32_t idx = 0;
#pragma unroll
    for (int i = 0; i < 5; i++)
        if (i < op.nq && op.qloc[i] < 0) idx |= uint32_t((base >> op.qphys[i]) & 1u) << i;
    return idx;
}

// a[r] *= ph for the registers r whose bit is set in rsel (bit r & 15).
// The masks of one register slot and the full mask -- the common ones --
// run as straight code; any other mask takes one uniform branch per register.
// (s < 0: every register)
template <int s, int N>
SVMI_HD inline void mul_slot(c64 (&a)[N], c64 ph)
{
#pragma unroll
    for (int r = 0; r < N; r++)
        if (s < 0 || ((r >> (s < 0 ? 0 : s)) & 1)) cmul_to(a[r], ph);
}
template <int N>
SVMI_HD inline void apply_rsel(c64 (&a)[N], c64 ph, uint32_t rsel)
{
    switch (rsel) {
        SVMI_CASE_ID(0xAAAAu, 400, (mul_slot<0, N>(a, ph)))
        SVMI_CASE_ID(0xCCCCu, 401, (mul_slot<1, N>(a, ph)))
        SVMI_CASE_ID(0xF0F0u, 402, (mul_slot<2, N>(a, ph)))
        SVMI_CASE_ID(0xFF00u, 403, (mul_slot<3, N>(a, ph)))
        SVMI_CASE_ID(0xFFFFu, 404, (mul_slot<-1, N>(a, ph)))
    default:
#pragma unroll
        for (int r = 0; r < N; r++)
            if ((rsel >> (r & (kRegs - 1))) & 1u) cmul_to(a[r], ph);
        break;
    }
}

// Controlled phase: amp *= ph where every qubit of the op is 1.  The
// register part of the condition is the uniform mask rsel, the group part
// one compare per thread, the part outside the tile one compare per tile.
template <int N>
SVMI_HD inline void apply_cphase(const DOpHdr& op, c64 (&a)[N], const c64* __restrict__ coef, uint32_t goff,
                                 uint64_t base)
{
    if ((base & op.emask) != op.emask) return;
    if ((goff & op.gmask) != op.gmask) return;
    apply_rsel<N>(a, coef[op.coef], op.rsel);
}

template <int N>
SVMI_HD inline void apply_diag(const DOp& op, c64 (&a)[N], const c64* __restrict__ coef, uint32_t goff,
                               const uint32_t (&off)[kRegs], uint64_t base)
{
    const uint32_t e = diag_ext_index(op, base);
    const c64* tab = coef + op.h.coef;
    c64 f[kRegs];
#pragma unroll
    for (int r = 0; r < kRegs; r++) f[r] = tab[diag_index(op, goff | off[r], e)];
#pragma unroll
    for (int r = 0; r < N; r++) cmul_to(a[r], f[r & (kRegs - 1)]);
}

// Apply one non-gradient op to one state's register group.
// goff = tile-local index bits of the group (non-register bits),
// off[rho] = tile-local index bits of register rho, base = tile base index.
// Product of controlled phases (h.s0 entries) applied to registers h.rsel.
template <int N>
SVMI_HD inline void apply_phaseset(const DOpHdr& h, c64 (&a)[N], const c64* __restrict__ coef, uint32_t goff,
                                   uint64_t base, const c64* xf)
{
    const PhaseEntry* pe = reinterpret_cast<const PhaseEntry*>(coef + h.coef);
    double pr = 1.0, pi = 0.0;
    if (h.s1) {   // the tile's product of the entries without lane condition
        const c64 f = xf[h.grad];
        pr = f.re;
        pi = f.im;
    }
    const uint32_t cnt = h.s0;
#pragma unroll 4
    for (uint32_t e = 0; e < cnt; e++) {
        const PhaseEntry E = pe[e];
        if ((base & E.emask) != E.emask) continue;
        const bool on = (goff & E.gmask) == E.gmask;
        const double tr = SVMI_FMA(pr, E.ph_re, -(pi * E.ph_im));
        const double ti = SVMI_FMA(pr, E.ph_im, pi * E.ph_re);
        pr = on ? tr : pr;
        pi = on ? ti : pi;
    }
    const c64 ph{pr, pi};
    apply_rsel<N>(a, ph, h.rsel);
}

// DOP_PHASETAB: the lane's product of nibble-table entries (times the tile
// factor), applied to the registers of rsel.
template <int N>
SVMI_HD inline void apply_phasetab(const DOpHdr& h, c64 (&a)[N], const c64* __restrict__ coef, uint32_t grp,
                                   const c64* xf)
{
    c64 ph = c64{1.0, 0.0};
    if (h.s1) ph = xf[h.grad];
    // two loops, so the LDS tables are read as LDS (a pointer that may be
    // either would be read with flat loads, which also wait for global ones)
    // (at most kPhaseTabMaxLane tables: unrolled, one uniform test each)
    if (h.gmask == kLaneTabGlobal) {
        const c64* T = coef + h.coef;
#pragma unroll
        for (uint32_t j = 0; j < uint32_t(kPhaseTabMaxLane); j++)
            if (j < h.s0) ph = cmul(ph, T[16u * j + ((grp >> (4u * j)) & 15u)]);
    } else {
        const c64* T = xf + h.gmask;
#pragma unroll
        for (uint32_t j = 0; j < uint32_t(kPhaseTabMaxLane); j++)
            if (j < h.s0) ph = cmul(ph, T[16u * j + ((grp >> (4u * j)) & 15u)]);
    }
    apply_rsel<N>(a, ph, h.rsel);
}

// The tile base's bits outside the tile, packed in ascending qubit order.
SVMI_HD inline uint64_t ext_index(const SweepInfo& sw, uint64_t base) { return extract_bits(base, sw.extmask); }

// Factor slot of a DOP_PHASETAB op for the tile with external bits x.
SVMI_HD inline c64 phasetab_factor(const DOpHdr& h, const c64* __restrict__ coef, uint64_t x)
{
    const c64* E = coef + h.coef + 16u * h.s0;
    c64 f{1.0, 0.0};
    if (h.smask == 0) return f;
    // every table's entry is loaded before the first product, so the loads
    // overlap; tables past h.smask read entry 0 of the first and are unused
    c64 e[kPhaseTabMaxExt];
#pragma unroll
    for (uint32_t j = 0; j < uint32_t(kPhaseTabMaxExt); j++)
        e[j] = E[j < h.smask ? 16u * j + uint32_t((x >> (4u * j)) & 15u) : 0u];
#pragma unroll
    for (uint32_t j = 0; j < uint32_t(kPhaseTabMaxExt); j++)
        if (j < h.smask) f = cmul(f, e[j]);
    return f;
}

// Factor slot of a DOP_PHASESET op for the tile at `base`: the product of
// its h.s1 entries that follow the per-lane ones.
SVMI_HD inline c64 phase_factor(const DOpHdr& h, const c64* __restrict__ coef, uint64_t base)
{
    const PhaseEntry* pe = reinterpret_cast<const PhaseEntry*>(coef + h.coef) + h.s0;
    double pr = 1.0, pi = 0.0;
    for (uint32_t e = 0; e < h.s1; e++) {
        const PhaseEntry E = pe[e];
        if ((base & E.emask) != E.emask) continue;
        const double tr = SVMI_FMA(pr, E.ph_re, -(pi * E.ph_im));
        const double ti = SVMI_FMA(pr, E.ph_im, pi * E.ph_re);
        pr = tr;
        pi = ti;
    }
    return c64{pr, pi};
}

// The tile factor of op h (DOP_PHASESET / DOP_PHASETAB with s1) for the
// tile at `base` with external bits x.
SVMI_HD inline c64 tile_factor(const DOpHdr& h, const c64* __restrict__ coef, uint64_t base, uint64_t x)
{
    return h.type == DOP_PHASETAB ? phasetab_factor(h, coef, x) : phase_factor(h, coef, base);
}

// DOP_ROT: optional register table, then the lifts of the slots in smask.
// The four {u, l} pairs are read together (one scalar load) and each slot
// is one uniform branch.
template <int N>
SVMI_HD inline void apply_rot(const DOpHdr& h, c64 (&a)[N], const c64* __restrict__ coef, uint32_t v)
{
    if (h.s0) apply_rtab<N>(a, coef + h.grad + 16u * v);
    const c64* L = coef + h.coef;
    const c64 f0 = L[0], f1 = L[1], f2 = L[2], f3 = L[3];
#define SVMI_ROT_SLOT(S, F)                                              \
    if ((h.smask >> S) & 1u) {                                           \
        if ((h.s1 >> S) & 1u) {                                          \
            SVMI_FENCE(320 + S);                                         \
            apply_lift<S, true, N>(a, F.re, F.im);                       \
            SVMI_FENCE(320 + S);                                         \
        } else {                                                         \
            SVMI_FENCE(330 + S);                                         \
            apply_lift<S, false, N>(a, F.re, F.im);                      \
            SVMI_FENCE(330 + S);                                         \
        }                                                                \
    }
    SVMI_ROT_SLOT(0, f0)
    SVMI_ROT_SLOT(1, f1)
    SVMI_ROT_SLOT(2, f2)
    SVMI_ROT_SLOT(3, f3)
#undef SVMI_ROT_SLOT
}

// h = the op's header (prefetched), op = the op in memory (diagonal tail),
// grp = the group index (goff with the register bits squeezed out).
// vmask / vhi: the sweep's variant qubits (SweepInfo::exthi) and the tile
// base's bits there, so a table over all of them finds its variant with one
// compare (otherwise the bits of its own mask are extracted).
template <int kMaxQ = kRegBits, int N = kRegs>
SVMI_HD inline void exec_state_op(const DOpHdr& h, const DOp* __restrict__ op, c64 (&a)[N],
                                  const c64* __restrict__ coef, uint32_t goff, uint32_t grp,
                                  const uint32_t (&off)[kRegs], uint64_t base, const c64* xf,
                                  uint64_t vmask = 0, uint32_t vhi = 0)
{
    const auto variant = [&](uint64_t emask) {
        return emask == 0 ? 0u : (emask == vmask ? vhi : rtab_variant(emask, base));
    };
    // the op kinds of rotation/table phases first (one or two compares),
    // then the rest
    if (h.type == DOP_ROT) {
        SVMI_FENCE(DOP_ROT + 100);
        apply_rot<N>(h, a, coef, h.s0 ? variant(h.emask) : 0u);
        SVMI_FENCE(DOP_ROT + 100);
        return;
    }
    if (h.type == DOP_LIFT2) {
        SVMI_FENCE(DOP_LIFT2 + 100);
        dispatch_lift<false, N>(h, a);
        SVMI_FENCE(DOP_LIFT2 + 100);
        return;
    }
    if (h.type == DOP_RTAB) {
        SVMI_FENCE(DOP_RTAB + 100);
        apply_rtab<N>(a, coef + h.coef + 16u * variant(h.emask));
        SVMI_FENCE(DOP_RTAB + 100);
        return;
    }
    if (h.type == DOP_LIFT3) {
        SVMI_FENCE(DOP_LIFT3 + 100);
        dispatch_lift<true, N>(h, a);
        SVMI_FENCE(DOP_LIFT3 + 100);
        return;
    }
    if (h.type == DOP_PHASETAB) {
        SVMI_FENCE(DOP_PHASETAB + 100);
        apply_phasetab<N>(h, a, coef, grp, xf);
        SVMI_FENCE(DOP_PHASETAB + 100);
        return;
    }
    switch (h.type) {
        SVMI_CASE_ID(DOP_DENSE, DOP_DENSE + 100, (dispatch_dense<kMaxQ, N>(h.smask, a, coef + h.coef)))
        SVMI_CASE_ID(DOP_X, DOP_X + 100, dispatch_x<N>(h.s0, a))
        SVMI_CASE_ID(DOP_CX, DOP_CX + 100, dispatch_cx<N>(h.s0, h.s1, a))
        SVMI_CASE_ID(DOP_SWAP, DOP_SWAP + 100, dispatch_swap<N>(h.s0, h.s1, a))
        SVMI_CASE_ID(DOP_SRN, DOP_SRN + 100, dispatch_srn<N>(h.s0, a))
        SVMI_CASE_ID(DOP_CPHASE, DOP_CPHASE + 100, apply_cphase<N>(h, a, coef, goff, base))
        SVMI_CASE_ID(DOP_DIAG, DOP_DIAG + 100, apply_diag<N>(*op, a, coef, goff, off, base))
        SVMI_CASE_ID(DOP_PHASESET, DOP_PHASESET + 100, apply_phaseset<N>(h, a, coef, goff, base, xf))
        SVMI_CASE_ID(DOP_CDENSE, DOP_CDENSE + 100, dispatch_cdense<N>(h, a, coef, goff, base))
    default: break;
    }
}

// Gradient op: Re sum conj(lam) K phi over this group;
// pl = phi's H amplitudes of the group then lambda's.
template <int kMaxQ, int H>
SVMI_HD inline double exec_grad_op(const DOpHdr& h, const DOp* __restrict__ op, const c64 (&pl)[2 * H],
                                   const c64* __restrict__ coef, uint32_t goff, const uint32_t (&off)[kRegs],
                                   uint64_t base)
{
    if (h.type == DOP_GRAD_DENSE) return dispatch_grad_dense<kMaxQ, H>(h.smask, pl, coef + h.coef, h.s0);
    // DOP_GRAD_DIAG
    const uint32_t e = diag_ext_index(*op, base);
    const c64* tab = coef + h.coef;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < H; r++) {
        const c64 v = cmul(tab[diag_index(*op, goff | off[r], e)], pl[r]);
        s += pl[H + r].re * v.re + pl[H + r].im * v.im;
    }
    return s;
}

// Register layout of a phase: off[rho] from the 4 register bits.
SVMI_HD inline void phase_offsets(const PhaseInfo& ph, uint32_t (&off)[kRegs])
{
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        uint32_t o = 0;
#pragma unroll
        for (int s = 0; s < kRegBits; s++)
            if ((r >> s) & 1) o |= 1u << ph.r[s];
        off[r] = o;
    }
}

// Gather offsets of a phase (PhaseInfo::gcol / gxor): the register contents
// after its folded leading permutations.
SVMI_HD inline void phase_gather_offsets(const PhaseInfo& ph, uint32_t (&off)[kRegs])
{
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        uint32_t o = ph.gxor;
#pragma unroll
        for (int s = 0; s < kRegBits; s++)
            if ((r >> s) & 1) o ^= ph.gcol[s];
        off[r] = o;
    }
}

// Scatter offsets of a mapped phase: register rho is written to LDS offset
// sxor ^ (XOR of scol[s] over the set bits s of rho) within the group.
SVMI_HD inline void phase_scatter_offsets(const PhaseInfo& ph, uint32_t (&off)[kRegs])
{
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        uint32_t o = ph.sxor;
#pragma unroll
        for (int s = 0; s < kRegBits; s++)
            if ((r >> s) & 1) o ^= ph.scol[s];
        off[r] = o;
    }
}

// LDS index of group g's base under a phase's tile map: XOR of gm[i] over the
// set bits i of g (the group's register bits being zero).
SVMI_HD inline uint32_t mapped_group_offset(const PhaseInfo& ph, uint32_t g)
{
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 10; i++)
        if ((g >> i) & 1u) o ^= ph.gm[i];
    return o;
}

// LDS index of tile-local index l under a sweep's final tile map.
SVMI_HD inline uint32_t sweep_map(const SweepInfo& sw, uint32_t l)
{
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < 14; b++)
        if ((l >> b) & 1u) o ^= sw.fcol[b];
    return o;
}

// The thread part of the store's map: M(tid) for tid below the tile's
// thread count (at most 2^10 threads).
SVMI_HD inline uint32_t sweep_map_thread(const SweepInfo& sw, uint32_t tid)
{
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < 10; b++)
        if ((tid >> b) & 1u) o ^= sw.fcol[b];
    return o;
}

// Tile-local base index of group g: insert zero bits at the 4 register bits.
SVMI_HD inline uint32_t group_offset(const PhaseInfo& ph, uint32_t g)
{
#pragma unroll
    for (int s = 0; s < kRegBits; s++) {
        const uint32_t b = ph.r[s];
        g = ((g >> b) << (b + 1)) | (g & ((1u << b) - 1u));
    }
    return g;
}

// LDS slot of tile-local index l (XOR swizzle of the low 4 bits with the
// next 4, so strided register gathers spread over the LDS banks).
SVMI_HD inline uint32_t lds_slot(uint32_t l) { return l ^ ((l >> 4) & 15u); }

// Tile base: the tile id's bits deposited on the sweep's external positions,
// low bits on extmask & ~exthi, high bits on exthi (each ascending).
SVMI_HD inline uint64_t tile_base(const SweepInfo& sw, uint64_t tile)
{
    const uint64_t lo = sw.extmask & ~sw.exthi;
    return deposit_bits(tile, lo) | deposit_bits(tile >> __builtin_popcountll(lo), sw.exthi);
}

#if defined(__HIPCC__)
// ------------------------------------------------------------ LDS DMA
//
// gfx950's LDS DMA (global_load_lds_dwordx4) moves 16 bytes per lane from
// global memory straight into LDS, lane i of a wave at M0 + 16 i, without
// holding registers.  It is issued from inline asm: the compiler cannot tell
// the two tile buffers apart and would wait for the prefetch before every
// LDS read of the other buffer; the sweep waits for it explicitly (vmcnt,
// which gfx9-family GPUs decrement in issue order for loads and stores
// alike) before the barrier that publishes the buffer.
__device__ __attribute__((always_inline)) inline uint32_t lds_address(const void* p)
{
    return uint32_t(reinterpret_cast<unsigned long long>((__attribute__((address_space(3))) const void*)p));
}

__device__ __attribute__((always_inline)) inline void lds_dma16(const c64* src, uint32_t lds_byte)
{
    asm volatile("s_mov_b32 m0, %1\n\t"
                 "s_nop 0\n\t"
                 "global_load_lds_dwordx4 %0, off"
                 :
                 : "v"(src), "s"(lds_byte)
                 : "memory", "m0");
}


// Issue the DMA of the tile at `base` into the buffer at LDS byte address
// lds_buf: the j-th DMA of wave w fills slots [64 (4 j + w), +64); slot S
// holds tile-local index lds_slot(S) (the swizzle is an involution), whose
// global index is base | ht[l >> c] | (l & cmask).  All addresses are formed
// before the first DMA, so no LDS read waits behind one.
__device__ __attribute__((always_inline)) inline void
tile_dma(const c64* __restrict__ s, const uint64_t* ht, uint32_t c, uint64_t base, uint32_t lds_buf,
         uint32_t wave, uint32_t lane)
{
    const uint32_t cmask = (1u << c) - 1u;
    const c64* src[kDmaPerWave];
#pragma unroll
    for (uint32_t j = 0; j < kDmaPerWave; j++) {
        const uint32_t l = lds_slot((j * kDmaWaves + wave) * 64u + lane);
        src[j] = s + (base | ht[l >> c] | (l & cmask));
    }
#pragma unroll
    for (uint32_t j = 0; j < kDmaPerWave; j++)
        lds_dma16(src[j], __builtin_amdgcn_readfirstlane(lds_buf + (j * kDmaWaves + wave) * 64u * 16u));
}

#endif

}  // namespace svmi
