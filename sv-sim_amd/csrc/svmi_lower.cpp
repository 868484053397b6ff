// svmi_lower.cpp -- reference gates -> composite matrices on their qubits.
//
// Each reference gate is evaluated by running its reference decomposition
// (svsim/src/svsim_cpu_omp.hpp:2016-2374) on the 2^k basis states of a
// k-qubit scratch register, with the reference primitive formulas
// (svsim_cpu_omp.hpp:630-1059).  The columns of the result are the gate's
// matrix, including every global phase and quirk of the reference (RZ == U1,
// RZZ == CX.U1.CX, non-unitary R).  The sequences of primitives and the
// constants in them have to be the reference's for the results to match; the
// code that evaluates them is this engine's own (small dense register, host
// only, templated on the scalar so the same code yields derivatives).
#include "svmi_lower.hpp"

#include <cmath>
#include <cstring>

namespace svmi {

namespace {

constexpr double kS2I = 0.70710678118654752440;  // config.hpp:54
constexpr double kPI = 3.14159265358979323846;   // config.hpp:52

// Forward-mode dual number for d/d(angle).
struct Dual {
    double v, d;
    Dual() : v(0), d(0) {}
    Dual(double x) : v(x), d(0) {}
    Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual cos(Dual a) { return Dual(std::cos(a.v), -std::sin(a.v) * a.d); }
inline Dual sin(Dual a) { return Dual(std::sin(a.v), std::cos(a.v) * a.d); }
inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }
inline double der(double) { return 0.0; }
inline double der(Dual x) { return x.d; }
using std::cos;
using std::sin;

// Small register: amplitudes as separate re/im vectors like the reference.
template <class T>
struct Reg {
    int k;
    std::vector<T> re, im;
    explicit Reg(int k_) : k(k_), re(size_t(1) << k_), im(size_t(1) << k_) {}
    size_t dim() const { return size_t(1) << k; }
};

// Pair loop over local qubit q (OP_HEAD, svsim_cpu_omp.hpp:620-628).
#define PAIRS(R, q)                                                  \
    for (size_t i_ = 0; i_ < (R).dim() / 2; i_++) {                   \
        const size_t pos0 = ((i_ >> (q)) << ((q) + 1)) + (i_ & ((size_t(1) << (q)) - 1)); \
        const size_t pos1 = pos0 + (size_t(1) << (q));
#define PAIRS_END }

template <class T>
void c1(Reg<T>& s, T e0r, T e0i, T e1r, T e1i, T e2r, T e2i, T e3r, T e3i, int q)  // 631-652
{
    PAIRS(s, q)
        const T a0r = s.re[pos0], a0i = s.im[pos0], a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos0] = (e0r * a0r) - (e0i * a0i) + (e1r * a1r) - (e1i * a1i);
        s.im[pos0] = (e0r * a0i) + (e0i * a0r) + (e1r * a1i) + (e1i * a1r);
        s.re[pos1] = (e2r * a0r) - (e2i * a0i) + (e3r * a1r) - (e3i * a1i);
        s.im[pos1] = (e2r * a0i) + (e2i * a0r) + (e3r * a1i) + (e3i * a1r);
    PAIRS_END
}

template <class T>
void cx(Reg<T>& s, int c, int t)  // 753-783
{
    for (size_t x = 0; x < s.dim(); x++)
        if (((x >> c) & 1) && !((x >> t) & 1)) {
            const size_t y = x | (size_t(1) << t);
            std::swap(s.re[x], s.re[y]);
            std::swap(s.im[x], s.im[y]);
        }
}

template <class T>
void xg(Reg<T>& s, int q)  // 790-804
{
    PAIRS(s, q)
        std::swap(s.re[pos0], s.re[pos1]);
        std::swap(s.im[pos0], s.im[pos1]);
    PAIRS_END
}

template <class T>
void yg(Reg<T>& s, int q)  // 809-824
{
    PAIRS(s, q)
        const T a0r = s.re[pos0], a0i = s.im[pos0], a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos0] = a1i; s.im[pos0] = -a1r;
        s.re[pos1] = -a0i; s.im[pos1] = a0r;
    PAIRS_END
}

template <class T>
void zg(Reg<T>& s, int q)  // 829-838
{
    PAIRS(s, q)
        (void)pos0;
        s.re[pos1] = -s.re[pos1]; s.im[pos1] = -s.im[pos1];
    PAIRS_END
}

template <class T>
void hg(Reg<T>& s, int q)  // 845-858
{
    PAIRS(s, q)
        const T a0r = s.re[pos0], a0i = s.im[pos0], a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos0] = T(kS2I) * (a0r + a1r); s.im[pos0] = T(kS2I) * (a0i + a1i);
        s.re[pos1] = T(kS2I) * (a0r - a1r); s.im[pos1] = T(kS2I) * (a0i - a1i);
    PAIRS_END
}

template <class T>
void rg(Reg<T>& s, T phase, int q)  // 887-896
{
    PAIRS(s, q)
        (void)pos0;
        const T a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos1] = -(a1i * phase); s.im[pos1] = a1r * phase;
    PAIRS_END
}

template <class T>
void sg(Reg<T>& s, int q)  // 905-913
{
    PAIRS(s, q)
        (void)pos0;
        const T a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos1] = -a1i; s.im[pos1] = a1r;
    PAIRS_END
}

template <class T>
void sdg(Reg<T>& s, int q)  // 920-928
{
    PAIRS(s, q)
        (void)pos0;
        const T a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos1] = a1i; s.im[pos1] = -a1r;
    PAIRS_END
}

template <class T>
void tg(Reg<T>& s, int q)  // 935-943
{
    PAIRS(s, q)
        (void)pos0;
        const T a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos1] = T(kS2I) * (a1r - a1i); s.im[pos1] = T(kS2I) * (a1r + a1i);
    PAIRS_END
}

template <class T>
void tdg(Reg<T>& s, int q)  // 950-958
{
    PAIRS(s, q)
        (void)pos0;
        const T a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos1] = T(kS2I) * (a1r + a1i); s.im[pos1] = T(kS2I) * (-a1r + a1i);
    PAIRS_END
}

template <class T>
void rx(Reg<T>& s, T theta, int q)  // 983-998
{
    const T rxr = cos(theta / 2.0);
    const T rxi = -sin(theta / 2.0);
    PAIRS(s, q)
        const T a0r = s.re[pos0], a0i = s.im[pos0], a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos0] = (rxr * a0r) - (rxi * a1i);
        s.im[pos0] = (rxr * a0i) + (rxi * a1r);
        s.re[pos1] = -(rxi * a0i) + (rxr * a1r);
        s.im[pos1] = (rxi * a0r) + (rxr * a1i);
    PAIRS_END
}

template <class T>
void ry(Reg<T>& s, T theta, int q)  // 1002-1020
{
    const T e0 = cos(theta / 2.0), e1 = -sin(theta / 2.0);
    const T e2 = sin(theta / 2.0), e3 = cos(theta / 2.0);
    PAIRS(s, q)
        const T a0r = s.re[pos0], a0i = s.im[pos0], a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos0] = (e0 * a0r) + (e1 * a1r);
        s.im[pos0] = (e0 * a0i) + (e1 * a1i);
        s.re[pos1] = (e2 * a0r) + (e3 * a1r);
        s.im[pos1] = (e2 * a0i) + (e3 * a1i);
    PAIRS_END
}

template <class T>
void u1(Reg<T>& s, T lambda, int q)  // 1024-1040
{
    const T e3r = cos(lambda), e3i = sin(lambda);
    PAIRS(s, q)
        (void)pos0;
        const T a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos1] = (e3r * a1r) - (e3i * a1i);
        s.im[pos1] = (e3r * a1i) + (e3i * a1r);
    PAIRS_END
}

template <class T>
void wg(Reg<T>& s, int q)  // 1047-1059
{
    PAIRS(s, q)
        const T a0r = s.re[pos0], a0i = s.im[pos0], a1r = s.re[pos1], a1i = s.im[pos1];
        s.re[pos0] = T(kS2I) * (a0r + a1i);
        s.im[pos0] = T(kS2I) * (a0i - a1r);
        s.re[pos1] = T(kS2I) * (a0i + a1r);
        s.im[pos1] = T(kS2I) * (-a0r + a1i);
    PAIRS_END
}

template <class T>
void u2(Reg<T>& s, T phi, T lambda, int q)  // 2030-2043
{
    c1(s, T(kS2I), T(0.0), T(-kS2I) * cos(lambda), T(-kS2I) * sin(lambda),
       T(kS2I) * cos(phi), T(kS2I) * sin(phi),
       T(kS2I) * cos(phi + lambda), T(kS2I) * sin(phi + lambda), q);
}

template <class T>
void u3(Reg<T>& s, T theta, T phi, T lambda, int q)  // 2047-2061
{
    c1(s, cos(theta / 2.), T(0.0),
       (-cos(lambda)) * sin(theta / 2.), (-sin(lambda)) * sin(theta / 2.),
       cos(phi) * sin(theta / 2.), sin(phi) * sin(theta / 2.),
       cos(phi + lambda) * cos(theta / 2.), sin(phi + lambda) * cos(theta / 2.), q);
}

template <class T>
void cu1(Reg<T>& s, T lambda, int a, int b)  // 2122-2130
{
    u1(s, lambda / 2, a); cx(s, a, b); u1(s, -lambda / 2, b); cx(s, a, b); u1(s, lambda / 2, b);
}

// C3X (2278-2309) and C3SQRTX (2313-2343): same sequence, angle pi/4 or pi/8.
template <class T>
void c3x_pattern(Reg<T>& s, double ang, int a, int b, int c, int d)
{
    const T p(ang), m(-ang);
    hg(s, d); cu1(s, m, a, d); hg(s, d); cx(s, a, b);
    hg(s, d); cu1(s, p, b, d); hg(s, d); cx(s, a, b);
    hg(s, d); cu1(s, m, b, d); hg(s, d); cx(s, b, c);
    hg(s, d); cu1(s, p, c, d); hg(s, d); cx(s, a, c);
    hg(s, d); cu1(s, m, c, d); hg(s, d); cx(s, b, c);
    hg(s, d); cu1(s, p, c, d); hg(s, d); cx(s, a, c);
    hg(s, d); cu1(s, m, c, d); hg(s, d);
}

template <class T>
void ccx(Reg<T>& s, int a, int b, int c)  // 2150-2167
{
    hg(s, c); cx(s, b, c); tdg(s, c); cx(s, a, c); tg(s, c); cx(s, b, c); tdg(s, c);
    cx(s, a, c); tg(s, b); tg(s, c); hg(s, c); cx(s, a, b); tg(s, a); tdg(s, b); cx(s, a, b);
}

// Run reference gate `op` on local qubits 0..k-1 (gate slot i -> local bit i).
template <class T>
void run_gate(Reg<T>& s, int op, T theta, T phi, T lambda)
{
    const int a = 0, b = 1, c = 2, d = 3, e = 4;
    const T PI(kPI);
    switch (op) {
    case SVMI_U3: u3(s, theta, phi, lambda, a); break;
    case SVMI_U2: u2(s, phi, lambda, a); break;
    case SVMI_U1: u1(s, lambda, a); break;
    case SVMI_CX: cx(s, a, b); break;
    case SVMI_ID: break;
    case SVMI_X: xg(s, a); break;
    case SVMI_Y: yg(s, a); break;
    case SVMI_Z: zg(s, a); break;
    case SVMI_H: hg(s, a); break;
    case SVMI_S: sg(s, a); break;
    case SVMI_SDG: sdg(s, a); break;
    case SVMI_T: tg(s, a); break;
    case SVMI_TDG: tdg(s, a); break;
    case SVMI_RX: rx(s, theta, a); break;
    case SVMI_RY: ry(s, theta, a); break;
    case SVMI_RZ: u1(s, phi, a); break;                                  // 2065-2069
    case SVMI_CZ: hg(s, b); cx(s, a, b); hg(s, b); break;                // 2073-2079
    case SVMI_CY: sdg(s, b); cx(s, a, b); sg(s, b); break;               // 2083-2089
    case SVMI_SWAP: cx(s, a, b); cx(s, b, a); cx(s, a, b); break;        // 2171-2177
    case SVMI_CH:                                                        // 2093-2107
        hg(s, b); sdg(s, b); cx(s, a, b); hg(s, b); tg(s, b); cx(s, a, b);
        tg(s, b); hg(s, b); sg(s, b); xg(s, b); sg(s, a);
        break;
    case SVMI_CCX: ccx(s, a, b, c); break;
    case SVMI_CSWAP: cx(s, c, b); ccx(s, a, b, c); cx(s, c, b); break;   // 2180-2186
    case SVMI_CRX:                                                       // 2190-2198
        u1(s, PI / 2, b); cx(s, a, b); u3(s, -lambda / 2, T(0.0), T(0.0), b); cx(s, a, b);
        u3(s, lambda / 2, -PI / 2, T(0.0), b);
        break;
    case SVMI_CRY:                                                       // 2202-2209
        u3(s, lambda / 2, T(0.0), T(0.0), b); cx(s, a, b);
        u3(s, -lambda / 2, T(0.0), T(0.0), b); cx(s, a, b);
        break;
    case SVMI_CRZ:                                                       // 2111-2118
        u1(s, lambda / 2, b); cx(s, a, b); u1(s, -lambda / 2, b); cx(s, a, b);
        break;
    case SVMI_CU1: cu1(s, lambda, a, b); break;
    case SVMI_CU3: {                                                     // 2134-2147
        const T temp1 = (lambda - phi) / 2;
        const T temp2 = theta / 2;
        const T temp3 = -(phi + lambda) / 2;
        u1(s, -temp3, a); u1(s, temp1, b); cx(s, a, b);
        u3(s, -temp2, T(0.0), temp3, b); cx(s, a, b); u3(s, temp2, phi, T(0.0), b);
        break;
    }
    case SVMI_RXX:                                                       // 2213-2223
        u3(s, PI / 2, theta, T(0.0), a); hg(s, b); cx(s, a, b); u1(s, -theta, b);
        cx(s, a, b); hg(s, b); u2(s, -PI, PI - theta, a);
        break;
    case SVMI_RZZ: cx(s, a, b); u1(s, theta, b); cx(s, a, b); break;     // 2227-2233
    case SVMI_RCCX:                                                      // 2237-2249
        u2(s, T(0.0), PI, c); u1(s, PI / 4, c); cx(s, b, c); u1(s, -PI / 4, c);
        cx(s, a, c); u1(s, PI / 4, c); cx(s, b, c); u1(s, -PI / 4, c); u2(s, T(0.0), PI, c);
        break;
    case SVMI_RC3X:                                                      // 2253-2274
        u2(s, T(0.0), PI, d); u1(s, PI / 4, d); cx(s, c, d); u1(s, -PI / 4, d);
        u2(s, T(0.0), PI, d); cx(s, a, d); u1(s, PI / 4, d); cx(s, b, d);
        u1(s, -PI / 4, d); cx(s, a, d); u1(s, PI / 4, d); cx(s, b, d);
        u1(s, -PI / 4, d); u2(s, T(0.0), PI, d); u1(s, PI / 4, d); cx(s, c, d);
        u1(s, -PI / 4, d); u2(s, T(0.0), PI, d);
        break;
    case SVMI_C3X: c3x_pattern(s, kPI / 4, a, b, c, d); break;
    case SVMI_C3SQRTX: c3x_pattern(s, kPI / 8, a, b, c, d); break;
    case SVMI_C4X:                                                       // 2347-2360
        hg(s, e); cu1(s, T(-kPI / 2), d, e); hg(s, e);
        c3x_pattern(s, kPI / 4, a, b, c, d);
        hg(s, d); cu1(s, T(kPI / 4), d, e); hg(s, d);
        c3x_pattern(s, kPI / 4, a, b, c, d);
        c3x_pattern(s, kPI / 8, a, b, c, e);
        break;
    case SVMI_R: rg(s, theta, a); break;
    case SVMI_W: wg(s, a); break;
    case SVMI_RYY:                                                       // 2364-2374
        rx(s, PI / 2, a); rx(s, PI / 2, b); cx(s, a, b); u1(s, theta, b);
        cx(s, a, b); rx(s, -PI / 2, a); rx(s, -PI / 2, b);
        break;
    default: break;
    }
}

double snap(double x)
{
    if (std::fabs(x) < 1e-14) return 0.0;
    if (std::fabs(x - 1.0) < 1e-14) return 1.0;
    if (std::fabs(x + 1.0) < 1e-14) return -1.0;
    return x;
}

}  // namespace

int gate_arity(int op)
{
    switch (op) {
    case SVMI_CX: case SVMI_CZ: case SVMI_CY: case SVMI_SWAP: case SVMI_CH: case SVMI_CRX:
    case SVMI_CRY: case SVMI_CRZ: case SVMI_CU1: case SVMI_CU3: case SVMI_RXX: case SVMI_RZZ:
    case SVMI_RYY:
        return 2;
    case SVMI_CCX: case SVMI_CSWAP: case SVMI_RCCX: return 3;
    case SVMI_RC3X: case SVMI_C3X: case SVMI_C3SQRTX: return 4;
    case SVMI_C4X: return 5;
    default: return (op >= 0 && op < SVMI_NUM_OPS) ? 1 : -1;
    }
}

bool gate_is_unitary(int op) { return op != SVMI_R && op != SVMI_SRN; }

svmi_status validate_gate(const svmi_gate& g, int n, std::string& err)
{
    const int k = gate_arity(g.op);
    if (k < 0) {
        err = "unknown gate opcode " + std::to_string(g.op);
        return SVMI_ERR_OP;
    }
    for (int i = 0; i < 5; i++)
        if (g.qb[i] < 0 || g.qb[i] >= n) {
            err = "qubit " + std::to_string(g.qb[i]) + " out of range for " + std::to_string(n) + " qubits";
            return SVMI_ERR_QUBIT;
        }
    for (int i = 0; i < k; i++)
        for (int j = i + 1; j < k; j++)
            if (g.qb[i] == g.qb[j]) {
                err = "repeated qubit " + std::to_string(g.qb[i]) + " in one gate";
                return SVMI_ERR_QUBIT;
            }
    return SVMI_OK;
}

void lower_gate(const svmi_gate& g, LOp& out)
{
    const int k = gate_arity(g.op);
    out.k = k;
    for (int i = 0; i < 5; i++) out.q[i] = i < k ? g.qb[i] : 0;
    if (g.op == SVMI_SRN) {
        out.kind = LOp::SRN;
        out.m.clear();
        return;
    }
    const size_t D = size_t(1) << k;
    std::vector<c64> M(D * D);
    for (size_t col = 0; col < D; col++) {
        Reg<double> s(k);
        s.re[col] = 1.0;
        run_gate(s, g.op, g.theta, g.phi, g.lambda);
        for (size_t row = 0; row < D; row++)
            M[row * D + col] = c64{snap(s.re[row]), snap(s.im[row])};
    }
    bool diag = true;
    for (size_t r = 0; r < D && diag; r++)
        for (size_t c = 0; c < D; c++)
            if (r != c && (M[r * D + c].re != 0.0 || M[r * D + c].im != 0.0)) {
                diag = false;
                break;
            }
    if (diag) {
        out.kind = LOp::DIAG;
        out.m.resize(D);
        for (size_t r = 0; r < D; r++) out.m[r] = M[r * D + r];
    } else {
        out.kind = LOp::DENSE;
        out.m.swap(M);
    }
}

bool gate_derivative(const svmi_gate& g, int slot, LOp& out)
{
    const int k = gate_arity(g.op);
    out.k = k;
    for (int i = 0; i < 5; i++) out.q[i] = i < k ? g.qb[i] : 0;
    if (g.op == SVMI_SRN) return false;
    const size_t D = size_t(1) << k;
    std::vector<c64> M(D * D);
    const Dual th(g.theta, slot == 0 ? 1.0 : 0.0);
    const Dual ph(g.phi, slot == 1 ? 1.0 : 0.0);
    const Dual la(g.lambda, slot == 2 ? 1.0 : 0.0);
    bool any = false, diag = true;
    for (size_t col = 0; col < D; col++) {
        Reg<Dual> s(k);
        s.re[col] = Dual(1.0);
        run_gate(s, g.op, th, ph, la);
        for (size_t row = 0; row < D; row++) {
            const c64 v{der(s.re[row]), der(s.im[row])};
            M[row * D + col] = v;
            if (v.re != 0.0 || v.im != 0.0) {
                any = true;
                if (row != col) diag = false;
            }
        }
    }
    if (!any) return false;
    if (diag) {
        out.kind = LOp::DIAG;
        out.m.resize(D);
        for (size_t r = 0; r < D; r++) out.m[r] = M[r * D + r];
    } else {
        out.kind = LOp::DENSE;
        out.m.swap(M);
    }
    return true;
}

void lop_to_dense(const LOp& a, std::vector<c64>& M)
{
    const size_t D = size_t(1) << a.k;
    if (a.kind == LOp::DENSE) {
        M = a.m;
        return;
    }
    M.assign(D * D, c64{0.0, 0.0});
    for (size_t r = 0; r < D; r++) M[r * D + r] = a.m[r];
}

void lop_adjoint(const LOp& a, LOp& out)
{
    out.kind = a.kind;
    out.k = a.k;
    std::memcpy(out.q, a.q, sizeof(a.q));
    out.gate = a.gate;
    const size_t D = size_t(1) << a.k;
    out.m.resize(a.m.size());
    if (a.kind == LOp::DIAG) {
        for (size_t r = 0; r < D; r++) out.m[r] = c64{a.m[r].re, -a.m[r].im};
    } else if (a.kind == LOp::DENSE) {
        for (size_t r = 0; r < D; r++)
            for (size_t c = 0; c < D; c++) {
                const c64 v = a.m[c * D + r];
                out.m[r * D + c] = c64{v.re, -v.im};
            }
    }
}

void lop_compose_same(const LOp& first, const LOp& second, LOp& out)
{
    const size_t D = size_t(1) << first.k;
    out.k = first.k;
    std::memcpy(out.q, first.q, sizeof(first.q));
    out.gate = second.gate;
    if (first.kind == LOp::DIAG && second.kind == LOp::DIAG) {
        out.kind = LOp::DIAG;
        out.m.resize(D);
        for (size_t r = 0; r < D; r++) out.m[r] = cmul(second.m[r], first.m[r]);
        return;
    }
    std::vector<c64> A, B;
    lop_to_dense(first, A);
    lop_to_dense(second, B);
    out.kind = LOp::DENSE;
    out.m.assign(D * D, c64{0.0, 0.0});
    for (size_t r = 0; r < D; r++)
        for (size_t c = 0; c < D; c++) {
            c64 acc{0.0, 0.0};
            for (size_t j = 0; j < D; j++) acc = cmac(acc, B[r * D + j], A[j * D + c]);
            out.m[r * D + c] = acc;
        }
}

}  // namespace svmi
