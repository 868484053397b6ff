// svmi_internal.hpp -- types shared by the host planner and the HIP kernels.
//
// Layout of the state in HBM: array-of-structures, one 16-byte c64 {re, im}
// per amplitude, index = basis state (qubit q = bit q).  A 2^n state is
// 16 * 2^n bytes (16 GiB at n = 30, 128 GiB at n = 33).
//
// A circuit is lowered (svmi_lower.cpp) to a list of "lowered ops" (dense,
// diagonal or SRN on <= 5 qubits) and scheduled (svmi_plan.cpp) into sweeps.
// A sweep is one read+write pass over the state: the state is cut into tiles
// of 2^k amplitudes that share all bits outside the sweep's tile qubit set T;
// every tile is staged in LDS and a sequence of phases is applied to it.  A
// phase picks 4 tile bits as "register bits" R: each thread holds the 16
// amplitudes of one group (fixed non-R bits) in registers, applies every op of
// the phase, and writes the group back to LDS.  Non-diagonal ops of a phase
// act on R only; diagonal ops may act on any qubit (tile bit or not).
#pragma once

#if defined(__HIPCC_RTC__)   // runtime compilation: its runtime header has the integer types
using __hip_internal::int32_t;
using __hip_internal::uint16_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
using __hip_internal::uint8_t;
#else
#include <stdint.h>
#include <stddef.h>
#endif

#if defined(__HIPCC__)
#define SVMI_HD __host__ __device__ __attribute__((always_inline))
#else
#define SVMI_HD
#endif

namespace svmi {

struct alignas(16) c64 {
    double re, im;
};

#if defined(__HIP_DEVICE_COMPILE__)
#define SVMI_FMA(a, b, c) __builtin_fma((a), (b), (c))
// x * y - z with both operations rounded (never contracted into an FMA)
#define SVMI_MUL_SUB(x, y, z) __dsub_rn(__dmul_rn((x), (y)), (z))
#else
#define SVMI_FMA(a, b, c) ((a) * (b) + (c))
#define SVMI_MUL_SUB(x, y, z) ((x) * (y) - (z))
#endif

SVMI_HD inline c64 cmul(c64 a, c64 b)
{
    return c64{SVMI_FMA(a.re, b.re, -(a.im * b.im)), SVMI_FMA(a.re, b.im, a.im * b.re)};
}
// a *= b in place: the two products first, then each part's FMA is the
// last use of that part's old value, so the result can take its input's
// register in any order (the register groups of the sweep interpreter end
// every op in the registers they started in; a result that cannot is
// copied back at the join)
SVMI_HD inline void cmul_to(c64& a, c64 b)
{
    const double p = a.im * b.im;
    a.im = SVMI_FMA(a.re, b.im, a.im * b.re);
    a.re = SVMI_MUL_SUB(a.re, b.re, p);
}
SVMI_HD inline c64 cadd(c64 a, c64 b) { return c64{a.re + b.re, a.im + b.im}; }
// acc + m * v (four FMAs on the device)
SVMI_HD inline c64 cmac(c64 acc, c64 m, c64 v)
{
    acc.re = SVMI_FMA(m.re, v.re, acc.re);
    acc.re = SVMI_FMA(-m.im, v.im, acc.re);
    acc.im = SVMI_FMA(m.re, v.im, acc.im);
    acc.im = SVMI_FMA(m.im, v.re, acc.im);
    return acc;
}
// conj(a) * b
SVMI_HD inline c64 cconjmul(c64 a, c64 b) { return c64{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }

constexpr int kRegBits = 4;                 // register bits per phase
constexpr int kRegs = 1 << kRegBits;        // amplitudes per thread per group
constexpr int kMaxTileQubits = 13;          // LDS tile: 2^13 * 16 B = 128 KiB
constexpr int kMaxTileQubitsAdj = 11;       // two states (phi, lambda): 2 * 32 KiB
// Two-state phases have 3 register bits; their fourth register "bit" is k,
// the bit above the tile, which selects lambda's half of the LDS (phi at
// [0, 2^k), lambda at [2^k, 2^(k+1))).  A thread's 16-amplitude group is then
// phi's 8 amplitudes then lambda's, every op acts on both, and a tile has
// 2^(k-3) threads -- the register footprint of a one-state group.
constexpr int kRegBitsAdj = 3;
constexpr int kRegsAdj = 1 << kRegBitsAdj;   // amplitudes per state per thread
// threads of a tile of 2^k amplitudes with ns states
SVMI_HD inline uint32_t sweep_threads(uint32_t k, int ns) { return 1u << (k - (ns == 2 ? kRegBitsAdj : kRegBits)); }
constexpr int kMaxDenseQubitsAdj = 2;       // two-state sweeps run dense ops of <= 2 qubits
                                            // (wider gates: full-state passes on both states)
constexpr int kThreads = 256;               // threads of the reduction / Pauli workgroups
// double-buffered sweeps (LDS DMA of the next tile): 12-qubit tiles, 256 threads
constexpr uint32_t kDmaTileQubits = 12;
constexpr uint32_t kDmaThreads = 1u << (kDmaTileQubits - kRegBits);      // 256
constexpr uint32_t kDmaWaves = kDmaThreads / 64;                          // 4
constexpr uint32_t kDmaPerWave = (1u << kDmaTileQubits) / 64 / kDmaWaves; // 16 DMA per wave per tile
constexpr int kMaxTileThreads = 1 << (kMaxTileQubits - kRegBits);   // 512
constexpr int kMaxDenseQubits = 5;          // standalone dense ops
constexpr int kMaxGradsPerSweep = 256;      // adjoint accumulators per sweep (LDS)

// Device op types.
enum DOpType : uint16_t {
    DOP_NOP = 0,
    DOP_DENSE = 1,   // dense 2^q x 2^q on register slots `smask`; coef = matrix
    DOP_X = 2,       // bit flip on slot s0
    DOP_CX = 3,      // control slot s0, target slot s1
    DOP_SWAP = 4,    // slots s0, s1
    DOP_SRN = 5,     // reference SRN quirk on slot s0 (real-linear only)
    DOP_CPHASE = 6,  // amp *= coef[0] where all bits of the op are set
    DOP_DIAG = 7,    // amp *= coef[idx], idx from the op's nq qubits
    DOP_GRAD_DENSE = 8,  // adjoint: acc[grad] += Re sum conj(lam) (coef-matrix phi) on slots;
                         // s0 = the matrix's nonzero components (grad_dense)
    DOP_GRAD_DIAG = 9,   // adjoint: acc[grad] += Re sum conj(lam) coef[idx] phi
    DOP_PHASESET = 10,   // product of controlled phases, applied to registers rsel
    DOP_CDENSE = 11,     // 2x2 coef on register slot s0, on the register pairs in rsel
                         // where the group bits gmask and the bits emask outside the
                         // tile are all 1 (multi-controlled single-qubit gate)
    // A real plane rotation on register slot s0 as in-place lifting steps
    // (a0/a1 = the amplitudes with slot bit 0/1, each component separately):
    DOP_LIFT2 = 12,      // a0 += u*a1; a1 += l*a0.  coef[0] = {u, l}.  With u = -s/c,
                         // l = s*c this is diag(1/c, c) . [[c, -s], [s, c]]; the planner
                         // folds the diag(c, 1/c) that completes the rotation into a table
    DOP_LIFT3 = 13,      // a0 += u*a1; a1 += l*a0; a0 += u*a1 (three shears, no scale):
                         // [[c, -s], [s, c]] with u = (c - 1)/s, l = s
    DOP_RTAB = 14,       // diagonal over the register group: a[r] *= T[r], T = coef + 16*v,
                         // v = the tile base's bits at the positions of emask (ascending),
                         // so a table may depend on up to kMaxTabExt qubits outside the tile
    DOP_PHASETAB = 15,   // product of controlled phases from nibble tables, applied to the
                         // registers rsel: ph = T_0[g & 15] * ... * T_{s0-1}[(g >> 4(s0-1)) & 15]
                         // for the group index g (bit m = the m-th tile bit outside the
                         // registers), times the tile factor xf[grad] if s1; that factor is
                         // E_0[x & 15] * ... * E_{smask-1}[...] for x = the tile base's bits
                         // outside the tile (ext_index); T_j, E_j are 16 c64 each at coef;
                         // gmask = the factor-area slot where the tile keeps the T_j, or
                         // kLaneTabGlobal to read them from coef
    DOP_ROT = 16,        // a run of rotations in one dispatch: if s0, first the register
                         // table coef + grad + 16*v (v as DOP_RTAB, from emask); then the
                         // lifting steps on every slot of smask in ascending order, three
                         // shears where s1 has the slot's bit, {u, l} = coef[h.coef + slot]
};
constexpr uint32_t kLaneTabGlobal = 0xffffffffu;
constexpr int kMaxXfSlots = 1024;      // factor area per sweep: 16 KiB of LDS
constexpr int kPhaseTabMaxLane = 3;    // nibble tables over the group index (9 group bits)
constexpr int kPhaseTabMaxExt = 16;    // nibble tables over the bits outside the tile
constexpr int kMaxTabExt = 4;          // qubits outside the tile one register table may depend on

// One controlled phase of a DOP_PHASESET (2 c64 slots in the coefficient
// array): phase `ph` applies where the group bits gmask and the bits emask
// outside the tile are all 1 (the register condition is the op's rsel).
struct PhaseEntry {
    uint32_t gmask;
    uint32_t pad;
    uint64_t emask;
    double ph_re, ph_im;
};
static_assert(sizeof(PhaseEntry) == 32, "PhaseEntry layout");

// Phase products (DOP_PHASESET): h.s0 per-lane entries, then h.s1 entries
// with no condition inside the tile (gmask == 0), whose product depends on
// the tile only; it is computed once per tile into factor slot h.grad.
constexpr int kMaxPhaseFactors = 64;   // factor slots per sweep (LDS: 16 B each)

// Which state(s) an op applies to in adjoint sweeps.
enum : uint8_t { DOP_TO_PSI = 1, DOP_TO_LAM = 2, DOP_TO_BOTH = 3 };

// One device op (uniform across the workgroup).  Every field is a full
// dword so the kernels fetch it with scalar loads; the kernels prefetch the
// 40-byte header one op ahead and read the diagonal-op tail on demand.
struct DOpHdr {
    uint32_t type;
    uint32_t smask;      // DENSE / GRAD_DENSE: register slot mask
    uint32_t s0, s1;     // X/CX/SWAP/SRN slots
    uint32_t coef;       // offset into the coefficient array (c64 units)
    uint32_t grad;       // GRAD_*: accumulator slot within the sweep
    uint32_t rsel;       // CPHASE/DIAG/CDENSE: registers rho whose register bits satisfy the op
    uint32_t gmask;      // CPHASE/DIAG/CDENSE: non-register tile-local bits that must be 1
    uint64_t emask;      // CPHASE/DIAG/CDENSE: physical bits outside the tile that must be 1
};
// DOP_LIFT2 / DOP_LIFT3 also carry their two shear factors in the header
// (u in emask's bits, l in gmask | rsel << 32), so the interpreted sweeps
// apply a lift without waiting for a coefficient load; coef still points at
// {u, l} in the coefficient array for the specialised kernels.
SVMI_HD inline uint64_t lift_bits(double x)
{
    uint64_t b;
    __builtin_memcpy(&b, &x, sizeof(b));
    return b;
}
SVMI_HD inline double lift_val(uint64_t b)
{
    double x;
    __builtin_memcpy(&x, &b, sizeof(x));
    return x;
}
SVMI_HD inline void lift_set(DOpHdr& h, double u, double l)
{
    h.emask = lift_bits(u);
    const uint64_t lb = lift_bits(l);
    h.gmask = uint32_t(lb);
    h.rsel = uint32_t(lb >> 32);
}
SVMI_HD inline double lift_u(const DOpHdr& h) { return lift_val(h.emask); }
SVMI_HD inline double lift_l(const DOpHdr& h) { return lift_val(uint64_t(h.gmask) | uint64_t(h.rsel) << 32); }

// Ops whose product of phases has a per-tile factor (LDS slot h.grad).
SVMI_HD inline bool has_tile_factor(const DOpHdr& h)
{
    return (h.type == DOP_PHASESET || h.type == DOP_PHASETAB) && h.s1;
}

struct DOp {
    DOpHdr h;
    uint32_t nq;         // DIAG: number of qubits
    uint32_t targets;    // adjoint: DOP_TO_* (forward sweeps: DOP_TO_PSI)
    int32_t qloc[5];     // DIAG: tile-local bit of qubit i, or -1 if outside the tile
    uint32_t qphys[5];   // DIAG: physical position of qubit i
};
static_assert(sizeof(DOpHdr) == 40 && sizeof(DOp) == 88, "DOp layout");

struct PhaseInfo {
    uint32_t r[4];       // register bits (tile-local positions, ascending)
    uint32_t op_begin, op_end;
    // gather map: register rho reads tile-local offset gxor ^ (XOR of gcol[s]
    // over the set bits s of rho) (within the group).  The phase's leading
    // register permutations (X, CX, SWAP on register slots) are folded into it
    // instead of being run; without them gcol[s] = 1 << r[s] and gxor = 0.
    // Everything after the gather uses the plain map (1 << r[s]).
    uint32_t gcol[4];
    uint32_t gxor;
    // scatter map: register rho is written to tile-local offset off[rho] ^ sxor
    // (the quarter turns a phase keeps in its register coordinates)
    uint32_t sxor;
    // Tile map (two-state sweeps): the amplitude of tile-local index x sits at
    // LDS index M(x), M linear over GF(2) (tile-local CX and SWAP ops folded
    // into the addressing instead of run).  scol[s] = M(1 << r[s]) and gm[i] =
    // M of the group number's bit i (the i-th non-register tile bit); gcol,
    // gxor and sxor above are then LDS offsets in the span of scol.  One-state
    // phases keep M the identity.
    uint32_t scol[4];
    uint16_t gm[10];
    uint32_t mapped;     // M differs from the identity
};
static_assert(sizeof(PhaseInfo) == 88, "PhaseInfo layout");

struct SweepInfo {
    uint32_t phase_begin, phase_end;
    uint32_t k;          // tile qubits
    uint32_t c;          // low physical qubits 0..c-1 always in the tile
    uint32_t n_ext;      // number of physical qubits outside the tile
    uint32_t hitab;      // offset into the hi-offset table (2^(k-c) entries)
    uint32_t n_grad;     // adjoint: accumulators used by this sweep
    uint32_t grad_base;  // adjoint: first global accumulator of this sweep
    uint64_t extmask;    // physical qubits outside the tile
    uint64_t exthi;      // those of them that are the tile id's most significant bits (tile id
                         // bit j <-> the j-th qubit of extmask & ~exthi, then of exthi, ascending):
                         // the qubits that select register-table variants, so consecutive tiles
                         // read the same variants
    uint64_t n_tiles;
    uint32_t n_xf;       // LDS c64 slots of the sweep's factor area: tile-uniform phase
                         // factors, then (last 16 * n_lt) the lane tables of its DOP_PHASETAB ops
    uint32_t op_begin;   // the sweep's ops are [op_begin, op_end)
    uint32_t op_end;
    uint32_t max_dense;  // register qubits of the sweep's widest DOP_DENSE (0 = none): sweeps
                         // with max_dense <= 2 run a kernel compiled without wider blocks
    uint32_t lt_coef;    // coef offset of a contiguous copy of those lane tables (16 c64 each),
    uint32_t n_lt;       // which every tile copies into the end of its factor area
    uint32_t mapped;     // some phase or the tile's store addresses LDS through a tile map M
    uint16_t fcol[14];   // M after the last phase: fcol[b] = M(1 << b)
    uint16_t fj[16];     // M(j nt) (nt = the tile's threads): the store's per-register part
};

// A standalone op (run by its own kernel on the whole state, outside sweeps).
struct StandaloneOp {
    int kind;            // 0 dense, 1 diag, 2 srn
    int q;               // qubits
    int qubits[5];       // physical positions; local bit i <-> qubits[i]
    uint32_t coef;
};

// ----------------------------------------------------------------------------
// Op application on one register group (shared by kernels and host checker).
// a[rho], rho = 4 register bits; bit s of rho <-> register slot s <-> tile bit r[s].
// N = 2 * kRegs holds the same group of two states (phi in a[0..15], lambda
// in a[16..31]): slot bits never reach bit 4, so one call applies the op to
// both states and the op is dispatched once.
// ----------------------------------------------------------------------------

template <int S>
struct SlotMask {
    static constexpr int q = (S & 1) + ((S >> 1) & 1) + ((S >> 2) & 1) + ((S >> 3) & 1);
    // i-th set bit position of S
    static constexpr int bit(int i)
    {
        int c = 0;
        for (int b = 0; b < 4; b++)
            if ((S >> b) & 1) {
                if (c == i) return b;
                c++;
            }
        return -1;
    }
    // spread the low q bits of j onto the set bits of S
    static constexpr int spread(int j)
    {
        int r = 0, c = 0;
        for (int b = 0; b < 4; b++)
            if ((S >> b) & 1) {
                if ((j >> c) & 1) r |= 1 << b;
                c++;
            }
        return r;
    }
    // spread the low 4-q bits of o onto the clear bits of S
    static constexpr int spread_out(int o)
    {
        int r = 0, c = 0;
        for (int b = 0; b < 4; b++)
            if (!((S >> b) & 1)) {
                if ((o >> c) & 1) r |= 1 << b;
                c++;
            }
        return r;
    }
};

// Dense 2^q x 2^q (row-major, matrix bit i <-> i-th set slot of S).
template <int S, int N = kRegs>
SVMI_HD inline void apply_dense(c64 (&a)[N], const c64* __restrict__ M)
{
    constexpr int q = SlotMask<S>::q;
    constexpr int D = 1 << q;
#pragma unroll
    for (int h = 0; h < N; h += kRegs)
#pragma unroll
        for (int o = 0; o < (kRegs >> q); o++) {
            c64 v[D];
#pragma unroll
            for (int j = 0; j < D; j++) v[j] = a[h | SlotMask<S>::spread_out(o) | SlotMask<S>::spread(j)];
#pragma unroll
            for (int i = 0; i < D; i++) {
                c64 acc = c64{0.0, 0.0};
#pragma unroll
                for (int j = 0; j < D; j++) acc = cmac(acc, M[i * D + j], v[j]);
                a[h | SlotMask<S>::spread_out(o) | SlotMask<S>::spread(i)] = acc;
            }
        }
}

// Re sum_groups conj(l) . (M p) over the slots of S (state untouched);
// pl = phi's H amplitudes then lambda's (S within the low log2(H) slots).
// Written as sum_ij (Re M_ij R_ij - Im M_ij I_ij) with R_ij + i I_ij =
// sum_o conj(l_i) p_j: a zero component of M (the usual case: U^dagger dU
// of RY is real antisymmetric, of a phase gate a single imaginary entry)
// skips its two FMAs per pair.  nz (from the op header, so no branch waits
// for M's loads): bit 2e is set if Re M_e != 0, bit 2e + 1 if Im M_e != 0,
// e = i * D + j.
template <int S, int H>
SVMI_HD inline double grad_dense(const c64 (&pl)[2 * H], const c64* __restrict__ M, uint32_t nz)
{
    constexpr int q = SlotMask<S>::q;
    constexpr int D = 1 << q;
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            const int e = i * D + j;
            if ((nz >> (2 * e)) & 1u) {
                const double mre = M[e].re;
                double r = 0.0;
#pragma unroll
                for (int o = 0; o < (H >> q); o++) {
                    const c64 lv = pl[H + (SlotMask<S>::spread_out(o) | SlotMask<S>::spread(i))];
                    const c64 pv = pl[SlotMask<S>::spread_out(o) | SlotMask<S>::spread(j)];
                    r = SVMI_FMA(lv.re, pv.re, r);
                    r = SVMI_FMA(lv.im, pv.im, r);
                }
                sum = SVMI_FMA(mre, r, sum);
            }
            if ((nz >> (2 * e + 1)) & 1u) {
                const double mim = M[e].im;
                double t = 0.0;
#pragma unroll
                for (int o = 0; o < (H >> q); o++) {
                    const c64 lv = pl[H + (SlotMask<S>::spread_out(o) | SlotMask<S>::spread(i))];
                    const c64 pv = pl[SlotMask<S>::spread_out(o) | SlotMask<S>::spread(j)];
                    t = SVMI_FMA(lv.re, pv.im, t);
                    t = SVMI_FMA(-lv.im, pv.re, t);
                }
                sum = SVMI_FMA(-mim, t, sum);
            }
        }
    return sum;
}

template <int s, int N = kRegs>
SVMI_HD inline void apply_x(c64 (&a)[N])
{
#pragma unroll
    for (int r = 0; r < N; r++)
        if (!((r >> s) & 1)) {
            const c64 t = a[r];
            a[r] = a[r | (1 << s)];
            a[r | (1 << s)] = t;
        }
}

template <int c, int t, int N = kRegs>
SVMI_HD inline void apply_cx(c64 (&a)[N])
{
#pragma unroll
    for (int r = 0; r < N; r++)
        if (((r >> c) & 1) && !((r >> t) & 1)) {
            const c64 tmp = a[r];
            a[r] = a[r | (1 << t)];
            a[r | (1 << t)] = tmp;
        }
}

template <int s0, int s1, int N = kRegs>
SVMI_HD inline void apply_swap(c64 (&a)[N])
{
#pragma unroll
    for (int r = 0; r < N; r++)
        if (((r >> s0) & 1) && !((r >> s1) & 1)) {
            const int r2 = r ^ (1 << s0) ^ (1 << s1);
            const c64 tmp = a[r];
            a[r] = a[r2];
            a[r2] = tmp;
        }
}

// Controlled 2x2 (row-major M) on register slot s: the pairs (r, r | 1<<s)
// whose register r is in rsel (the register part of the control condition;
// rsel never tells r from r | 1<<s apart).
template <int s, int N = kRegs>
SVMI_HD inline void apply_cdense(c64 (&a)[N], const c64* __restrict__ M, uint32_t rsel)
{
    const c64 m00 = M[0], m01 = M[1], m10 = M[2], m11 = M[3];
#pragma unroll
    for (int r = 0; r < N; r++)
        if (!((r >> s) & 1) && ((rsel >> (r & (kRegs - 1))) & 1u)) {
            const c64 a0 = a[r], a1 = a[r | (1 << s)];
            a[r] = cmac(cmac(c64{0.0, 0.0}, m00, a0), m01, a1);
            a[r | (1 << s)] = cmac(cmac(c64{0.0, 0.0}, m10, a0), m11, a1);
        }
}

// Lifting steps on slot s (DOP_LIFT2 / DOP_LIFT3).  Every step writes the
// amplitude it updates in place, so an op leaves each value in the register
// it came from (no copies where the op dispatch's cases join).
template <int s, bool kThree, int N = kRegs>
SVMI_HD inline void apply_lift(c64 (&a)[N], double u, double l)
{
#pragma unroll
    for (int r = 0; r < N; r++)
        if (!((r >> s) & 1)) {
            c64& a0 = a[r];
            c64& a1 = a[r | (1 << s)];
            a0.re = SVMI_FMA(u, a1.re, a0.re);
            a0.im = SVMI_FMA(u, a1.im, a0.im);
            a1.re = SVMI_FMA(l, a0.re, a1.re);
            a1.im = SVMI_FMA(l, a0.im, a1.im);
            if constexpr (kThree) {
                a0.re = SVMI_FMA(u, a1.re, a0.re);
                a0.im = SVMI_FMA(u, a1.im, a0.im);
            }
        }
}

// Bit deposit / extract over the set bits of `mask` (ascending), one step
// per run of consecutive set bits rather than per bit.  mask < 2^63.
SVMI_HD inline uint64_t deposit_bits(uint64_t v, uint64_t mask)
{
    uint64_t r = 0;
    while (mask) {
        const int p = __builtin_ctzll(mask);
        const int len = __builtin_ctzll(~(mask >> p));
        const uint64_t run = (uint64_t(1) << len) - 1u;
        r |= (v & run) << p;
        v >>= len;
        mask &= ~(run << p);
    }
    return r;
}
SVMI_HD inline uint64_t extract_bits(uint64_t v, uint64_t mask)
{
    uint64_t r = 0;
    int j = 0;
    while (mask) {
        const int p = __builtin_ctzll(mask);
        const int len = __builtin_ctzll(~(mask >> p));
        const uint64_t run = (uint64_t(1) << len) - 1u;
        r |= ((v >> p) & run) << j;
        j += len;
        mask &= ~(run << p);
    }
    return r;
}

// Register table (DOP_RTAB): table variant of the tile at `base`.
SVMI_HD inline uint32_t rtab_variant(uint64_t emask, uint64_t base) { return uint32_t(extract_bits(base, emask)); }

template <int N = kRegs>
SVMI_HD inline void apply_rtab(c64 (&a)[N], const c64* __restrict__ T)
{
#pragma unroll
    for (int r = 0; r < N; r++) cmul_to(a[r], T[r & (kRegs - 1)]);
}

// Reference SRN_GATE (svsim_cpu_omp.hpp:866-879), as coded there.
template <int s, int N = kRegs>
SVMI_HD inline void apply_srn(c64 (&a)[N])
{
#pragma unroll
    for (int r = 0; r < N; r++)
        if (!((r >> s) & 1)) {
            const c64 a0 = a[r], a1 = a[r | (1 << s)];
            a[r] = c64{0.5 * (a0.re + a1.re), 0.5 * (a0.im - a1.im)};
            a[r | (1 << s)] = c64{0.5 * (a0.re + a1.re), 0.5 * (-a0.im + a1.im)};
        }
}

}  // namespace svmi
