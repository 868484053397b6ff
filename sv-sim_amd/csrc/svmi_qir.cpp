// svmi_qir.cpp -- QIR gate records -> engine ops (host only).
//
// Every gate of the reference's QIR-facing engine is one single-qubit 2x2
// (svsim_qsharp/svsim_cpu_omp.hpp:623-1306) or Swap (1312-1357); the
// controlled forms apply it where every bit of the record's mask is set
// (OP_HEAD_MASK, 606-615).  The matrices below are those gates' actions,
// with the angles the reference passes to cos/sin (theta/2 for RI/RX/RY/RZ,
// theta for the E* gates), evaluated on the host with the same libm.
#include "svmi_qir.hpp"

#include <cmath>

namespace svmi {

namespace {

constexpr double kS2I = 0.70710678118654752440;   // svsim_qsharp/config.hpp:67

bool is_controlled(int op)
{
    return (op >= SVMI_QIR_CONTROLLED_X && op <= SVMI_QIR_CONTROLLED_EZ) || op == SVMI_QIR_CONTROLLED_ADJOINT_S ||
           op == SVMI_QIR_CONTROLLED_ADJOINT_T;
}

// Uncontrolled counterpart of a controlled op.
int base_op(int op)
{
    if (op >= SVMI_QIR_CONTROLLED_X && op <= SVMI_QIR_CONTROLLED_EZ) return op - SVMI_QIR_CONTROLLED_X;
    if (op == SVMI_QIR_CONTROLLED_ADJOINT_S) return SVMI_QIR_ADJOINT_S;
    if (op == SVMI_QIR_CONTROLLED_ADJOINT_T) return SVMI_QIR_ADJOINT_T;
    return op;
}

bool is_diagonal(int op)
{
    switch (op) {
    case SVMI_QIR_Z: case SVMI_QIR_S: case SVMI_QIR_T: case SVMI_QIR_RI: case SVMI_QIR_RZ:
    case SVMI_QIR_EI: case SVMI_QIR_EZ: case SVMI_QIR_ADJOINT_S: case SVMI_QIR_ADJOINT_T:
        return true;
    default:
        return false;
    }
}

// Diagonal (d0, d1) of a diagonal base op.
void diag_of(int op, double t, c64& d0, c64& d1)
{
    d0 = c64{1.0, 0.0};
    switch (op) {
    case SVMI_QIR_Z: d1 = c64{-1.0, 0.0}; break;                          // 664-676
    case SVMI_QIR_S: d1 = c64{0.0, 1.0}; break;                           // 700-712
    case SVMI_QIR_T: d1 = c64{kS2I, kS2I}; break;                         // 716-728
    case SVMI_QIR_ADJOINT_S: d1 = c64{0.0, -1.0}; break;                  // 1243-1255
    case SVMI_QIR_ADJOINT_T: d1 = c64{kS2I, -kS2I}; break;                // 1259-1271
    case SVMI_QIR_RI:                                                     // 733-753
        d0 = c64{std::cos(t / 2.0), -std::sin(t / 2.0)};
        d1 = d0;
        break;
    case SVMI_QIR_RZ:                                                     // 806-829
        d0 = c64{std::cos(t / 2), -std::sin(t / 2)};
        d1 = c64{std::cos(t / 2), std::sin(t / 2)};
        break;
    case SVMI_QIR_EI:                                                     // 833-855
        d0 = c64{std::cos(t), std::sin(t)};
        d1 = d0;
        break;
    case SVMI_QIR_EZ:                                                     // 911-933
        d0 = c64{std::cos(t), std::sin(t)};
        d1 = c64{std::cos(t), -std::sin(t)};
        break;
    default: d1 = c64{1.0, 0.0}; break;
    }
}

// Row-major 2x2 of a non-diagonal base op.
void dense_of(int op, double t, c64 (&m)[4])
{
    const c64 z{0.0, 0.0};
    switch (op) {
    case SVMI_QIR_X: m[0] = z; m[1] = c64{1, 0}; m[2] = c64{1, 0}; m[3] = z; break;      // 623-640
    case SVMI_QIR_Y: m[0] = z; m[1] = c64{0, -1}; m[2] = c64{0, 1}; m[3] = z; break;     // 644-660
    case SVMI_QIR_H:                                                                     // 680-696
        m[0] = m[1] = m[2] = c64{kS2I, 0};
        m[3] = c64{-kS2I, 0};
        break;
    case SVMI_QIR_RX: {   // [[c, i s], [i s, c]], s = -sin(t/2)                          757-776
        const double c = std::cos(t / 2.0), s = -std::sin(t / 2.0);
        m[0] = m[3] = c64{c, 0};
        m[1] = m[2] = c64{0, s};
        break;
    }
    case SVMI_QIR_EX: {   // [[cos t, i sin t], [i sin t, cos t]]                          859-881
        const double c = std::cos(t), s = std::sin(t);
        m[0] = m[3] = c64{c, 0};
        m[1] = m[2] = c64{0, s};
        break;
    }
    case SVMI_QIR_RY: {   // [[c, -s], [s, c]], s = sin(t/2)                              780-802
        const double c = std::cos(t / 2.0), s = std::sin(t / 2.0);
        m[0] = m[3] = c64{c, 0};
        m[1] = c64{-s, 0};
        m[2] = c64{s, 0};
        break;
    }
    case SVMI_QIR_EY: {   // [[cos t, sin t], [-sin t, cos t]]                             885-907
        const double c = std::cos(t), s = std::sin(t);
        m[0] = m[3] = c64{c, 0};
        m[1] = c64{s, 0};
        m[2] = c64{-s, 0};
        break;
    }
    default: m[0] = m[3] = c64{1, 0}; m[1] = m[2] = z; break;
    }
}

LOp one_qubit_dense(int q, const c64 (&m)[4], uint64_t ctrl)
{
    LOp op;
    op.kind = LOp::DENSE;
    op.k = 1;
    op.q[0] = q;
    op.m.assign(m, m + 4);
    op.ctrl = ctrl;
    return op;
}

// phase ph where qubit q and every bit of ctrl are 1
LOp phase_op(int q, c64 ph, uint64_t ctrl)
{
    LOp op;
    op.kind = LOp::DIAG;
    op.k = 1;
    op.q[0] = q;
    op.m = {c64{1.0, 0.0}, ph};
    op.ctrl = ctrl;
    return op;
}

void push_basic(std::vector<LOp>& seg, int op, int q)
{
    if (op == SVMI_QIR_H) {
        c64 m[4];
        dense_of(SVMI_QIR_H, 0.0, m);
        seg.push_back(one_qubit_dense(q, m, 0));
    } else {
        c64 d0, d1;
        diag_of(op, 0.0, d0, d1);
        LOp d;
        d.kind = LOp::DIAG;
        d.k = 1;
        d.q[0] = q;
        d.m = {d0, d1};
        seg.push_back(d);
    }
}

// One non-Swap, non-Measure record as ops on qubit q with control mask
// `mask` (controlled ops only); a mask naming q or a bit outside `all` never
// acts (OP_HEAD_MASK, 606-615) and emits nothing.  conj: the entrywise
// conjugate of every matrix (the density-matrix engine's column half).
void lower_gate_qir(std::vector<LOp>& seg, int op, int q, uint64_t mask, double theta, uint64_t all, bool conj)
{
    const size_t first = seg.size();
    uint64_t ctrl = 0;
    if (is_controlled(op)) {
        ctrl = mask;
        if (((ctrl >> q) & 1) || (ctrl & ~all)) return;   // never satisfied: identity
    }
    const int b = base_op(op);
    if (is_diagonal(b)) {
        c64 d0, d1;
        diag_of(b, theta, d0, d1);
        if (!ctrl) {
            LOp d;
            d.kind = LOp::DIAG;
            d.k = 1;
            d.q[0] = q;
            d.m = {d0, d1};
            seg.push_back(d);
        } else {
            // controlled diag(d0, d1) = phase d0 on the controls times phase d1/d0
            // on controls + target
            c64 r = d1;
            if (d0.re != 1.0 || d0.im != 0.0) {
                const int lo = __builtin_ctzll(ctrl);
                seg.push_back(phase_op(lo, d0, ctrl & (ctrl - 1)));
                const double den = d0.re * d0.re + d0.im * d0.im;
                r = c64{(d1.re * d0.re + d1.im * d0.im) / den, (d1.im * d0.re - d1.re * d0.im) / den};
            }
            if (!(d0.re == d1.re && d0.im == d1.im)) seg.push_back(phase_op(q, r, ctrl));
        }
    } else {
        c64 m[4];
        dense_of(b, theta, m);
        seg.push_back(one_qubit_dense(q, m, ctrl));
    }
    if (conj)
        for (size_t i = first; i < seg.size(); i++)
            for (c64& v : seg[i].m) v.im = -v.im;
}

}  // namespace

svmi_status validate_qir_gate(const svmi_qir_gate& g, int n, std::string& err)
{
    if (g.op < 0 || g.op >= SVMI_QIR_NUM_OPS) {
        err = "unknown QIR op " + std::to_string(g.op);
        return SVMI_ERR_OP;
    }
    if (g.qubit < 0 || g.qubit >= n) {
        err = "Gate uses qubit out of range! (qubit " + std::to_string(g.qubit) + ", " + std::to_string(n) +
              " allocated)";
        return SVMI_ERR_QUBIT;
    }
    if (g.op == SVMI_QIR_SWAP && (g.mask >= uint64_t(n) || g.mask == uint64_t(g.qubit))) {
        err = "Swap needs two distinct allocated qubits (" + std::to_string(g.qubit) + ", " +
              std::to_string(g.mask) + ")";
        return SVMI_ERR_QUBIT;
    }
    return SVMI_OK;
}

void lower_qir(const svmi_qir_gate* gates, size_t n_gates, int n, std::vector<std::vector<LOp>>& segments,
               std::vector<MeasureStep>& measures)
{
    segments.assign(1, {});
    measures.clear();
    const uint64_t all = n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
    for (size_t i = 0; i < n_gates; i++) {
        const svmi_qir_gate& g = gates[i];
        std::vector<LOp>& seg = segments.back();
        const int q = g.qubit;
        if (g.op == SVMI_QIR_MEASURE) {
            // Measure_GATE (1406-1500): basis change, P(1), collapse, inverse change
            const uint64_t pauli = g.mask;
            if (pauli == 1) push_basic(seg, SVMI_QIR_H, q);
            if (pauli == 2) {
                push_basic(seg, SVMI_QIR_ADJOINT_S, q);
                push_basic(seg, SVMI_QIR_H, q);
            }
            MeasureStep m;
            m.qubit = uint32_t(q);
            m.n = uint32_t(n);
            m.draw = g.theta;
            measures.push_back(m);
            segments.emplace_back();
            std::vector<LOp>& next = segments.back();
            LOp col = phase_op(q, c64{1.0, 0.0}, 0);   // placeholder factors
            col.dyn = int(measures.size()) - 1;
            next.push_back(col);
            if (pauli == 1) push_basic(next, SVMI_QIR_H, q);
            if (pauli == 2) {
                push_basic(next, SVMI_QIR_H, q);
                push_basic(next, SVMI_QIR_S, q);
            }
            continue;
        }
        if (g.op == SVMI_QIR_SWAP) {
            LOp op;
            op.kind = LOp::DENSE;
            op.k = 2;
            op.q[0] = q;
            op.q[1] = int(g.mask);
            op.m.assign(16, c64{0.0, 0.0});
            op.m[0 * 4 + 0] = op.m[1 * 4 + 2] = op.m[2 * 4 + 1] = op.m[3 * 4 + 3] = c64{1.0, 0.0};
            seg.push_back(op);
            continue;
        }
        lower_gate_qir(seg, g.op, q, g.mask, g.theta, all, false);
    }
}

void lower_dm_qir(const svmi_qir_gate* gates, size_t n_gates, int n, std::vector<std::vector<LOp>>& segments,
                  std::vector<MeasureStep>& measures)
{
    segments.assign(1, {});
    measures.clear();
    const int n2 = 2 * n;
    const uint64_t all = n2 >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n2) - 1);
    for (size_t i = 0; i < n_gates; i++) {
        const svmi_qir_gate& g = gates[i];
        std::vector<LOp>& seg = segments.back();
        const int q = g.qubit;
        if (g.op == SVMI_QIR_MEASURE) {
            const uint64_t pauli = g.mask;
            if (pauli == 1) {
                push_basic(seg, SVMI_QIR_H, q);
                push_basic(seg, SVMI_QIR_H, q + n);
            }
            if (pauli == 2) {
                push_basic(seg, SVMI_QIR_ADJOINT_S, q);
                push_basic(seg, SVMI_QIR_S, q + n);
                push_basic(seg, SVMI_QIR_H, q);
                push_basic(seg, SVMI_QIR_H, q + n);
            }
            MeasureStep m;
            m.qubit = uint32_t(q);
            m.n = uint32_t(n2);
            m.dm_n = uint32_t(n);
            m.draw = g.theta;
            measures.push_back(m);
            segments.emplace_back();
            std::vector<LOp>& next = segments.back();
            LOp col = phase_op(q, c64{1.0, 0.0}, 0);   // collapse of the row qubit; factors set on the device
            col.dyn = int(measures.size()) - 1;
            next.push_back(col);
            if (pauli == 1) {
                push_basic(next, SVMI_QIR_H, q);
                push_basic(next, SVMI_QIR_H, q + n);
            }
            if (pauli == 2) {
                push_basic(next, SVMI_QIR_H, q);
                push_basic(next, SVMI_QIR_H, q + n);
                push_basic(next, SVMI_QIR_S, q);
                push_basic(next, SVMI_QIR_ADJOINT_S, q + n);
            }
            continue;
        }
        if (g.op == SVMI_QIR_SWAP) {   // the row qubits only
            LOp op;
            op.kind = LOp::DENSE;
            op.k = 2;
            op.q[0] = q;
            op.q[1] = int(g.mask);
            op.m.assign(16, c64{0.0, 0.0});
            op.m[0 * 4 + 0] = op.m[1 * 4 + 2] = op.m[2 * 4 + 1] = op.m[3 * 4 + 3] = c64{1.0, 0.0};
            seg.push_back(op);
            continue;
        }
        lower_gate_qir(seg, g.op, q, g.mask, g.theta, all, false);
        const uint64_t cmask = is_controlled(g.op) ? (n >= 64 ? 0 : g.mask << n) : 0;
        lower_gate_qir(seg, g.op, q + n, cmask, g.theta, all, true);
    }
}

}  // namespace svmi
