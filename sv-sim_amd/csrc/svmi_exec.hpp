// svmi_exec.hpp -- execution of one device op on one register group.
// Compiled into the HIP kernels and into the host-side program checker
// (svmi_emulate_program), so both run literally the same op code.
#pragma once

#include "svmi_internal.hpp"

namespace svmi {

SVMI_HD inline void dispatch_dense(int smask, c64 (&a)[kRegs], const c64* __restrict__ M)
{
    switch (smask) {
#define SVMI_DENSE_CASE(S) case S: apply_dense<S>(a, M); break;
        SVMI_DENSE_CASE(1) SVMI_DENSE_CASE(2) SVMI_DENSE_CASE(3) SVMI_DENSE_CASE(4)
        SVMI_DENSE_CASE(5) SVMI_DENSE_CASE(6) SVMI_DENSE_CASE(7) SVMI_DENSE_CASE(8)
        SVMI_DENSE_CASE(9) SVMI_DENSE_CASE(10) SVMI_DENSE_CASE(11) SVMI_DENSE_CASE(12)
        SVMI_DENSE_CASE(13) SVMI_DENSE_CASE(14) SVMI_DENSE_CASE(15)
#undef SVMI_DENSE_CASE
    default: break;
    }
}

SVMI_HD inline double dispatch_grad_dense(int smask, const c64 (&p)[kRegs], const c64 (&l)[kRegs],
                                          const c64* __restrict__ M)
{
    switch (smask) {
#define SVMI_GRAD_CASE(S) case S: return grad_dense<S>(p, l, M);
        SVMI_GRAD_CASE(1) SVMI_GRAD_CASE(2) SVMI_GRAD_CASE(3) SVMI_GRAD_CASE(4)
        SVMI_GRAD_CASE(5) SVMI_GRAD_CASE(6) SVMI_GRAD_CASE(7) SVMI_GRAD_CASE(8)
        SVMI_GRAD_CASE(9) SVMI_GRAD_CASE(10) SVMI_GRAD_CASE(11) SVMI_GRAD_CASE(12)
        SVMI_GRAD_CASE(13) SVMI_GRAD_CASE(14) SVMI_GRAD_CASE(15)
#undef SVMI_GRAD_CASE
    default: return 0.0;
    }
}

SVMI_HD inline void dispatch_x(int s, c64 (&a)[kRegs])
{
    switch (s) {
    case 0: apply_x<0>(a); break;
    case 1: apply_x<1>(a); break;
    case 2: apply_x<2>(a); break;
    default: apply_x<3>(a); break;
    }
}

SVMI_HD inline void dispatch_cx(int c, int t, c64 (&a)[kRegs])
{
    switch (c * 4 + t) {
    case 1: apply_cx<0, 1>(a); break;
    case 2: apply_cx<0, 2>(a); break;
    case 3: apply_cx<0, 3>(a); break;
    case 4: apply_cx<1, 0>(a); break;
    case 6: apply_cx<1, 2>(a); break;
    case 7: apply_cx<1, 3>(a); break;
    case 8: apply_cx<2, 0>(a); break;
    case 9: apply_cx<2, 1>(a); break;
    case 11: apply_cx<2, 3>(a); break;
    case 12: apply_cx<3, 0>(a); break;
    case 13: apply_cx<3, 1>(a); break;
    case 14: apply_cx<3, 2>(a); break;
    default: break;
    }
}

SVMI_HD inline void dispatch_swap(int s0, int s1, c64 (&a)[kRegs])
{
    switch (s0 * 4 + s1) {
    case 1: apply_swap<0, 1>(a); break;
    case 2: apply_swap<0, 2>(a); break;
    case 3: apply_swap<0, 3>(a); break;
    case 6: apply_swap<1, 2>(a); break;
    case 7: apply_swap<1, 3>(a); break;
    case 11: apply_swap<2, 3>(a); break;
    default: break;
    }
}

SVMI_HD inline void dispatch_srn(int s, c64 (&a)[kRegs])
{
    switch (s) {
    case 0: apply_srn<0>(a); break;
    case 1: apply_srn<1>(a); break;
    case 2: apply_srn<2>(a); break;
    default: apply_srn<3>(a); break;
    }
}

// Index of a diagonal op's table entry for tile-local index l.
SVMI_HD inline uint32_t diag_index(const DOp& op, uint32_t l, uint32_t eidx)
{
    uint32_t idx = eidx;
    for (int i = 0; i < op.nq; i++)
        if (op.qloc[i] >= 0) idx |= ((l >> op.qloc[i]) & 1u) << i;
    return idx;
}

// Part of a diagonal op's table index fixed by the tile (qubits outside it).
SVMI_HD inline uint32_t diag_ext_index(const DOp& op, uint64_t base)
{
    uint32_t idx = 0;
    for (int i = 0; i < op.nq; i++)
        if (op.qloc[i] < 0) idx |= uint32_t((base >> op.qphys[i]) & 1u) << i;
    return idx;
}

// Apply one non-gradient op to one state's register group.
// goff = tile-local index bits of the group (non-register bits),
// off[rho] = tile-local index bits of register rho, base = tile base index.
SVMI_HD inline void exec_state_op(const DOp& op, c64 (&a)[kRegs], const c64* __restrict__ coef,
                                  uint32_t goff, const uint32_t (&off)[kRegs], uint64_t base)
{
    switch (op.type) {
    case DOP_DENSE: dispatch_dense(op.smask, a, coef + op.coef); break;
    case DOP_X: dispatch_x(op.s0, a); break;
    case DOP_CX: dispatch_cx(op.s0, op.s1, a); break;
    case DOP_SWAP: dispatch_swap(op.s0, op.s1, a); break;
    case DOP_SRN: dispatch_srn(op.s0, a); break;
    case DOP_CPHASE: {
        if ((base & op.emask) != op.emask) break;
        const c64 ph = coef[op.coef];
        const uint32_t lm = op.lmask;
#pragma unroll
        for (int r = 0; r < kRegs; r++)
            if (((goff | off[r]) & lm) == lm) a[r] = cmul(a[r], ph);
        break;
    }
    case DOP_DIAG: {
        const uint32_t e = diag_ext_index(op, base);
        const c64* tab = coef + op.coef;
#pragma unroll
        for (int r = 0; r < kRegs; r++) a[r] = cmul(a[r], tab[diag_index(op, goff | off[r], e)]);
        break;
    }
    default: break;
    }
}

// Gradient op: Re sum conj(lam) K phi over this group.
SVMI_HD inline double exec_grad_op(const DOp& op, const c64 (&p)[kRegs], const c64 (&l)[kRegs],
                                   const c64* __restrict__ coef, uint32_t goff,
                                   const uint32_t (&off)[kRegs], uint64_t base)
{
    if (op.type == DOP_GRAD_DENSE) return dispatch_grad_dense(op.smask, p, l, coef + op.coef);
    // DOP_GRAD_DIAG
    const uint32_t e = diag_ext_index(op, base);
    const c64* tab = coef + op.coef;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        const c64 v = cmul(tab[diag_index(op, goff | off[r], e)], p[r]);
        s += l[r].re * v.re + l[r].im * v.im;
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

// Tile base: pdep of the tile id over the sweep's external positions.
SVMI_HD inline uint64_t tile_base(const SweepInfo& sw, uint64_t tile)
{
    uint64_t base = 0;
    for (uint32_t e = 0; e < sw.n_ext; e++)
        base |= ((tile >> e) & 1ull) << sw.ext[e];
    return base;
}

}  // namespace svmi
