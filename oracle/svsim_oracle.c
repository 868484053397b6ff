/*
 * svsim_oracle.c -- CPU oracle: plain-C restatement of the reference
 * CPU-OpenMP engine, svsim/src/svsim_cpu_omp.hpp.
 *
 * TEST INFRASTRUCTURE ONLY (see svsim_oracle.h).  Each function cites the
 * reference lines it restates.  The arithmetic of every primitive keeps the
 * reference's operand order so that, compiled with -ffp-contract=off and
 * linked against the same libm, amplitudes are bit-identical to the
 * reference; tests/test_oracle.py pins this against fixtures produced by the
 * reference itself (oracle/_ref).
 *
 * Differences from the reference, none of which change results:
 *   - 64-bit indices and shifts (reference: 32-bit IdxType and int shifts,
 *     config.hpp:40, cpu_omp:625-628) so n >= 29 and qubits >= 31 work;
 *   - CX pair enumeration by bit insertion instead of div/mod (cpu_omp:753-783
 *     visits the same (ctrl=1,target=0)<->(ctrl=1,target=1) pairs);
 *   - measure() draws come from the caller instead of srand(time(0))/rand(),
 *     and the interval search is a binary search over the same scan.
 */
#include "svsim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* config.hpp:52-54 */
#define S2I 0.70710678118654752440
#define PI 3.14159265358979323846

typedef struct {
    uint64_t dim, half_dim;
    double* re;
    double* im;
} sv_t;

#define ONE ((uint64_t)1)

/* OP_HEAD (cpu_omp:620-628): pair (pos0, pos1) for index i < 2^(n-1). */
#define PAIR_LOOP_BEGIN(sv, q)                                              \
    _Pragma("omp for schedule(static)")                                     \
    for (int64_t ii = 0; ii < (int64_t)(sv)->half_dim; ii++) {              \
        const uint64_t i_ = (uint64_t)ii;                                   \
        const uint64_t pos0 = ((i_ >> (q)) << ((q) + 1)) + (i_ & ((ONE << (q)) - 1)); \
        const uint64_t pos1 = pos0 + (ONE << (q));
#define PAIR_LOOP_END }

/* ---------------- primitives (cpu_omp:630-1059) ---------------- */

/* C1_GATE, cpu_omp:631-652 */
static void c1_gate(sv_t* s, double e0r, double e0i, double e1r, double e1i,
                    double e2r, double e2i, double e3r, double e3i, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        const double a0r = R[pos0], a0i = I[pos0], a1r = R[pos1], a1i = I[pos1];
        R[pos0] = (e0r * a0r) - (e0i * a0i) + (e1r * a1r) - (e1i * a1i);
        I[pos0] = (e0r * a0i) + (e0i * a0r) + (e1r * a1i) + (e1i * a1r);
        R[pos1] = (e2r * a0r) - (e2i * a0i) + (e3r * a1r) - (e3i * a1i);
        I[pos1] = (e2r * a0i) + (e2i * a0r) + (e3r * a1i) + (e3i * a1r);
    PAIR_LOOP_END
}

/* CX_GATE, cpu_omp:753-783: swap (c=1,t=0) <-> (c=1,t=1). */
static void cx_gate(sv_t* s, int c, int t)
{
    double* R = s->re; double* I = s->im;
    const int lo = c < t ? c : t, hi = c < t ? t : c;
    const int64_t quarter = (int64_t)(s->dim >> 2);
    #pragma omp for schedule(static)
    for (int64_t ii = 0; ii < quarter; ii++) {
        uint64_t b = (uint64_t)ii;
        b = ((b >> lo) << (lo + 1)) | (b & ((ONE << lo) - 1));
        b = ((b >> hi) << (hi + 1)) | (b & ((ONE << hi) - 1));
        const uint64_t p0 = b | (ONE << c);
        const uint64_t p1 = p0 | (ONE << t);
        const double r0 = R[p0], i0 = I[p0], r1 = R[p1], i1 = I[p1];
        R[p0] = r1; I[p0] = i1;
        R[p1] = r0; I[p1] = i0;
    }
}

/* X_GATE, cpu_omp:790-804 */
static void x_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        const double a0r = R[pos0], a0i = I[pos0], a1r = R[pos1], a1i = I[pos1];
        R[pos0] = a1r; I[pos0] = a1i;
        R[pos1] = a0r; I[pos1] = a0i;
    PAIR_LOOP_END
}

/* Y_GATE, cpu_omp:809-824 */
static void y_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        const double a0r = R[pos0], a0i = I[pos0], a1r = R[pos1], a1i = I[pos1];
        R[pos0] = a1i; I[pos0] = -a1r;
        R[pos1] = -a0i; I[pos1] = a0r;
    PAIR_LOOP_END
}

/* Z_GATE, cpu_omp:829-838 (touches pos1 only) */
static void z_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        (void)pos0;
        const double a1r = R[pos1], a1i = I[pos1];
        R[pos1] = -a1r; I[pos1] = -a1i;
    PAIR_LOOP_END
}

/* H_GATE, cpu_omp:845-858 */
static void h_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        const double a0r = R[pos0], a0i = I[pos0], a1r = R[pos1], a1i = I[pos1];
        R[pos0] = S2I * (a0r + a1r); I[pos0] = S2I * (a0i + a1i);
        R[pos1] = S2I * (a0r - a1r); I[pos1] = S2I * (a0i - a1i);
    PAIR_LOOP_END
}

/* SRN_GATE, cpu_omp:866-879 (code as written, which differs from its comment) */
static void srn_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        const double a0r = R[pos0], a0i = I[pos0], a1r = R[pos1], a1i = I[pos1];
        R[pos0] = 0.5 * (a0r + a1r);  I[pos0] = 0.5 * (a0i - a1i);
        R[pos1] = 0.5 * (a0r + a1r);  I[pos1] = 0.5 * (-a0i + a1i);
    PAIR_LOOP_END
}

/* R_GATE, cpu_omp:887-896: a1 <- i*phase*a1 */
static void r_gate(sv_t* s, double phase, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        (void)pos0;
        const double a1r = R[pos1], a1i = I[pos1];
        R[pos1] = -(a1i * phase); I[pos1] = a1r * phase;
    PAIR_LOOP_END
}

/* S_GATE, cpu_omp:905-913 */
static void s_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        (void)pos0;
        const double a1r = R[pos1], a1i = I[pos1];
        R[pos1] = -a1i; I[pos1] = a1r;
    PAIR_LOOP_END
}

/* SDG_GATE, cpu_omp:920-928 */
static void sdg_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        (void)pos0;
        const double a1r = R[pos1], a1i = I[pos1];
        R[pos1] = a1i; I[pos1] = -a1r;
    PAIR_LOOP_END
}

/* T_GATE, cpu_omp:935-943 */
static void t_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        (void)pos0;
        const double a1r = R[pos1], a1i = I[pos1];
        R[pos1] = S2I * (a1r - a1i); I[pos1] = S2I * (a1r + a1i);
    PAIR_LOOP_END
}

/* TDG_GATE, cpu_omp:950-958 */
static void tdg_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        (void)pos0;
        const double a1r = R[pos1], a1i = I[pos1];
        R[pos1] = S2I * (a1r + a1i); I[pos1] = S2I * (-a1r + a1i);
    PAIR_LOOP_END
}

/* RX_GATE, cpu_omp:983-998 */
static void rx_gate(sv_t* s, double theta, int q)
{
    double* R = s->re; double* I = s->im;
    const double rxr = cos(theta / 2.0);
    const double rxi = -sin(theta / 2.0);
    PAIR_LOOP_BEGIN(s, q)
        const double a0r = R[pos0], a0i = I[pos0], a1r = R[pos1], a1i = I[pos1];
        R[pos0] = (rxr * a0r) - (rxi * a1i);
        I[pos0] = (rxr * a0i) + (rxi * a1r);
        R[pos1] = -(rxi * a0i) + (rxr * a1r);
        I[pos1] = +(rxi * a0r) + (rxr * a1i);
    PAIR_LOOP_END
}

/* RY_GATE, cpu_omp:1002-1020 */
static void ry_gate(sv_t* s, double theta, int q)
{
    double* R = s->re; double* I = s->im;
    const double e0 = cos(theta / 2.0);
    const double e1 = -sin(theta / 2.0);
    const double e2 = sin(theta / 2.0);
    const double e3 = cos(theta / 2.0);
    PAIR_LOOP_BEGIN(s, q)
        const double a0r = R[pos0], a0i = I[pos0], a1r = R[pos1], a1i = I[pos1];
        R[pos0] = (e0 * a0r) + (e1 * a1r);
        I[pos0] = (e0 * a0i) + (e1 * a1i);
        R[pos1] = (e2 * a0r) + (e3 * a1r);
        I[pos1] = (e2 * a0i) + (e3 * a1i);
    PAIR_LOOP_END
}

/* U1_GATE, cpu_omp:1024-1040 (pos0 is rewritten unchanged) */
static void u1_gate(sv_t* s, double lambda, int q)
{
    double* R = s->re; double* I = s->im;
    const double e3r = cos(lambda);
    const double e3i = sin(lambda);
    PAIR_LOOP_BEGIN(s, q)
        (void)pos0;
        const double a1r = R[pos1], a1i = I[pos1];
        R[pos1] = (e3r * a1r) - (e3i * a1i);
        I[pos1] = (e3r * a1i) + (e3i * a1r);
    PAIR_LOOP_END
}

/* W_GATE, cpu_omp:1047-1059 */
static void w_gate(sv_t* s, int q)
{
    double* R = s->re; double* I = s->im;
    PAIR_LOOP_BEGIN(s, q)
        const double a0r = R[pos0], a0i = I[pos0], a1r = R[pos1], a1i = I[pos1];
        R[pos0] = S2I * (a0r + a1i);
        I[pos0] = S2I * (a0i - a1r);
        R[pos1] = S2I * (a0i + a1r);
        I[pos1] = S2I * (-a0r + a1i);
    PAIR_LOOP_END
}

/* U2_GATE, cpu_omp:2030-2043 */
static void u2_gate(sv_t* s, double phi, double lambda, int q)
{
    const double e0r = S2I, e0i = 0;
    const double e1r = -S2I * cos(lambda), e1i = -S2I * sin(lambda);
    const double e2r = S2I * cos(phi), e2i = S2I * sin(phi);
    const double e3r = S2I * cos(phi + lambda), e3i = S2I * sin(phi + lambda);
    c1_gate(s, e0r, e0i, e1r, e1i, e2r, e2i, e3r, e3i, q);
}

/* U3_GATE, cpu_omp:2047-2061 */
static void u3_gate(sv_t* s, double theta, double phi, double lambda, int q)
{
    const double e0r = cos(theta / 2.), e0i = 0;
    const double e1r = -cos(lambda) * sin(theta / 2.);
    const double e1i = -sin(lambda) * sin(theta / 2.);
    const double e2r = cos(phi) * sin(theta / 2.);
    const double e2i = sin(phi) * sin(theta / 2.);
    const double e3r = cos(phi + lambda) * cos(theta / 2.);
    const double e3i = sin(phi + lambda) * cos(theta / 2.);
    c1_gate(s, e0r, e0i, e1r, e1i, e2r, e2i, e3r, e3i, q);
}

/* ---------------- composites (cpu_omp:2065-2374) ---------------- */

static void rz_gate(sv_t* s, double phi, int q) { u1_gate(s, phi, q); }  /* 2065-2069 */

static void cz_gate(sv_t* s, int a, int b)                               /* 2073-2079 */
{ h_gate(s, b); cx_gate(s, a, b); h_gate(s, b); }

static void cy_gate(sv_t* s, int a, int b)                               /* 2083-2089 */
{ sdg_gate(s, b); cx_gate(s, a, b); s_gate(s, b); }

static void ch_gate(sv_t* s, int a, int b)                               /* 2093-2107 */
{
    h_gate(s, b); sdg_gate(s, b); cx_gate(s, a, b); h_gate(s, b); t_gate(s, b);
    cx_gate(s, a, b); t_gate(s, b); h_gate(s, b); s_gate(s, b); x_gate(s, b);
    s_gate(s, a);
}

static void crz_gate(sv_t* s, double lambda, int a, int b)               /* 2111-2118 */
{
    u1_gate(s, lambda / 2, b); cx_gate(s, a, b);
    u1_gate(s, -lambda / 2, b); cx_gate(s, a, b);
}

static void cu1_gate(sv_t* s, double lambda, int a, int b)               /* 2122-2130 */
{
    u1_gate(s, lambda / 2, a); cx_gate(s, a, b);
    u1_gate(s, -lambda / 2, b); cx_gate(s, a, b);
    u1_gate(s, lambda / 2, b);
}

static void cu3_gate(sv_t* s, double theta, double phi, double lambda, int c, int t) /* 2134-2147 */
{
    const double temp1 = (lambda - phi) / 2;
    const double temp2 = theta / 2;
    const double temp3 = -(phi + lambda) / 2;
    u1_gate(s, -temp3, c);
    u1_gate(s, temp1, t);
    cx_gate(s, c, t);
    u3_gate(s, -temp2, 0, temp3, t);
    cx_gate(s, c, t);
    u3_gate(s, temp2, phi, 0, t);
}

static void ccx_gate(sv_t* s, int a, int b, int c)                       /* 2150-2167 */
{
    h_gate(s, c);
    cx_gate(s, b, c); tdg_gate(s, c);
    cx_gate(s, a, c); t_gate(s, c);
    cx_gate(s, b, c); tdg_gate(s, c);
    cx_gate(s, a, c); t_gate(s, b); t_gate(s, c); h_gate(s, c);
    cx_gate(s, a, b); t_gate(s, a); tdg_gate(s, b);
    cx_gate(s, a, b);
}

static void swap_gate(sv_t* s, int a, int b)                             /* 2171-2177 */
{ cx_gate(s, a, b); cx_gate(s, b, a); cx_gate(s, a, b); }

static void cswap_gate(sv_t* s, int a, int b, int c)                     /* 2180-2186 */
{ cx_gate(s, c, b); ccx_gate(s, a, b, c); cx_gate(s, c, b); }

static void crx_gate(sv_t* s, double lambda, int a, int b)               /* 2190-2198 */
{
    u1_gate(s, PI / 2, b); cx_gate(s, a, b);
    u3_gate(s, -lambda / 2, 0, 0, b); cx_gate(s, a, b);
    u3_gate(s, lambda / 2, -PI / 2, 0, b);
}

static void cry_gate(sv_t* s, double lambda, int a, int b)               /* 2202-2209 */
{
    u3_gate(s, lambda / 2, 0, 0, b); cx_gate(s, a, b);
    u3_gate(s, -lambda / 2, 0, 0, b); cx_gate(s, a, b);
}

static void rxx_gate(sv_t* s, double theta, int a, int b)                /* 2213-2223 */
{
    u3_gate(s, PI / 2, theta, 0, a); h_gate(s, b); cx_gate(s, a, b);
    u1_gate(s, -theta, b); cx_gate(s, a, b); h_gate(s, b);
    u2_gate(s, -PI, PI - theta, a);
}

static void rzz_gate(sv_t* s, double theta, int a, int b)                /* 2228-2233 */
{ cx_gate(s, a, b); u1_gate(s, theta, b); cx_gate(s, a, b); }

static void rccx_gate(sv_t* s, int a, int b, int c)                      /* 2237-2249 */
{
    u2_gate(s, 0, PI, c); u1_gate(s, PI / 4, c); cx_gate(s, b, c);
    u1_gate(s, -PI / 4, c); cx_gate(s, a, c); u1_gate(s, PI / 4, c);
    cx_gate(s, b, c); u1_gate(s, -PI / 4, c); u2_gate(s, 0, PI, c);
}

static void rc3x_gate(sv_t* s, int a, int b, int c, int d)               /* 2253-2274 */
{
    u2_gate(s, 0, PI, d); u1_gate(s, PI / 4, d); cx_gate(s, c, d);
    u1_gate(s, -PI / 4, d); u2_gate(s, 0, PI, d); cx_gate(s, a, d);
    u1_gate(s, PI / 4, d); cx_gate(s, b, d); u1_gate(s, -PI / 4, d);
    cx_gate(s, a, d); u1_gate(s, PI / 4, d); cx_gate(s, b, d);
    u1_gate(s, -PI / 4, d); u2_gate(s, 0, PI, d); u1_gate(s, PI / 4, d);
    cx_gate(s, c, d); u1_gate(s, -PI / 4, d); u2_gate(s, 0, PI, d);
}

/* C3X (cpu_omp:2278-2309) and C3SQRTX (2313-2343) share one pattern; only
 * the controlled-phase angle differs (pi/4 vs pi/8). */
static void c3x_pattern(sv_t* s, double ang, int a, int b, int c, int d)
{
    h_gate(s, d); cu1_gate(s, -ang, a, d); h_gate(s, d);
    cx_gate(s, a, b);
    h_gate(s, d); cu1_gate(s, ang, b, d); h_gate(s, d);
    cx_gate(s, a, b);
    h_gate(s, d); cu1_gate(s, -ang, b, d); h_gate(s, d);
    cx_gate(s, b, c);
    h_gate(s, d); cu1_gate(s, ang, c, d); h_gate(s, d);
    cx_gate(s, a, c);
    h_gate(s, d); cu1_gate(s, -ang, c, d); h_gate(s, d);
    cx_gate(s, b, c);
    h_gate(s, d); cu1_gate(s, ang, c, d); h_gate(s, d);
    cx_gate(s, a, c);
    h_gate(s, d); cu1_gate(s, -ang, c, d); h_gate(s, d);
}

static void c4x_gate(sv_t* s, int a, int b, int c, int d, int e)         /* 2347-2360 */
{
    h_gate(s, e); cu1_gate(s, -PI / 2, d, e); h_gate(s, e);
    c3x_pattern(s, PI / 4, a, b, c, d);
    h_gate(s, d); cu1_gate(s, PI / 4, d, e); h_gate(s, d);
    c3x_pattern(s, PI / 4, a, b, c, d);
    c3x_pattern(s, PI / 8, a, b, c, e);
}

static void ryy_gate(sv_t* s, double theta, int a, int b)                /* 2364-2374 */
{
    rx_gate(s, PI / 2, a); rx_gate(s, PI / 2, b); cx_gate(s, a, b);
    rz_gate(s, theta, b); cx_gate(s, a, b);
    rx_gate(s, -PI / 2, a); rx_gate(s, -PI / 2, b);
}

/* ---------------- dispatch (cpu_omp:2378-2607) ---------------- */

int orc_gate_arity(int op)
{
    switch (op) {
    case ORC_U3: case ORC_U2: case ORC_U1: case ORC_ID: case ORC_X: case ORC_Y:
    case ORC_Z: case ORC_H: case ORC_S: case ORC_SDG: case ORC_T: case ORC_TDG:
    case ORC_RX: case ORC_RY: case ORC_RZ: case ORC_R: case ORC_SRN: case ORC_W:
        return 1;
    case ORC_CX: case ORC_CZ: case ORC_CY: case ORC_SWAP: case ORC_CH:
    case ORC_CRX: case ORC_CRY: case ORC_CRZ: case ORC_CU1: case ORC_CU3:
    case ORC_RXX: case ORC_RZZ: case ORC_RYY:
        return 2;
    case ORC_CCX: case ORC_CSWAP: case ORC_RCCX:
        return 3;
    case ORC_RC3X: case ORC_C3X: case ORC_C3SQRTX:
        return 4;
    case ORC_C4X:
        return 5;
    default:
        return -1;
    }
}

static void exe_op(sv_t* s, const orc_gate* g)
{
    const int* q = g->qb;
    switch (g->op) {
    case ORC_U3: u3_gate(s, g->theta, g->phi, g->lambda, q[0]); break;
    case ORC_U2: u2_gate(s, g->phi, g->lambda, q[0]); break;
    case ORC_U1: u1_gate(s, g->lambda, q[0]); break;
    case ORC_CX: cx_gate(s, q[0], q[1]); break;
    case ORC_ID: break;                                   /* ID_GATE, cpu_omp:2020-2025 */
    case ORC_X: x_gate(s, q[0]); break;
    case ORC_Y: y_gate(s, q[0]); break;
    case ORC_Z: z_gate(s, q[0]); break;
    case ORC_H: h_gate(s, q[0]); break;
    case ORC_S: s_gate(s, q[0]); break;
    case ORC_SDG: sdg_gate(s, q[0]); break;
    case ORC_T: t_gate(s, q[0]); break;
    case ORC_TDG: tdg_gate(s, q[0]); break;
    case ORC_RX: rx_gate(s, g->theta, q[0]); break;
    case ORC_RY: ry_gate(s, g->theta, q[0]); break;
    case ORC_RZ: rz_gate(s, g->phi, q[0]); break;          /* RZ_OP reads phi */
    case ORC_CZ: cz_gate(s, q[0], q[1]); break;
    case ORC_CY: cy_gate(s, q[0], q[1]); break;
    case ORC_SWAP: swap_gate(s, q[0], q[1]); break;
    case ORC_CH: ch_gate(s, q[0], q[1]); break;
    case ORC_CCX: ccx_gate(s, q[0], q[1], q[2]); break;
    case ORC_CSWAP: cswap_gate(s, q[0], q[1], q[2]); break;
    case ORC_CRX: crx_gate(s, g->lambda, q[0], q[1]); break;
    case ORC_CRY: cry_gate(s, g->lambda, q[0], q[1]); break;
    case ORC_CRZ: crz_gate(s, g->lambda, q[0], q[1]); break;
    case ORC_CU1: cu1_gate(s, g->lambda, q[0], q[1]); break;
    case ORC_CU3: cu3_gate(s, g->theta, g->phi, g->lambda, q[0], q[1]); break;
    case ORC_RXX: rxx_gate(s, g->theta, q[0], q[1]); break;
    case ORC_RZZ: rzz_gate(s, g->theta, q[0], q[1]); break;
    case ORC_RCCX: rccx_gate(s, q[0], q[1], q[2]); break;
    case ORC_RC3X: rc3x_gate(s, q[0], q[1], q[2], q[3]); break;
    case ORC_C3X: c3x_pattern(s, PI / 4, q[0], q[1], q[2], q[3]); break;
    case ORC_C3SQRTX: c3x_pattern(s, PI / 8, q[0], q[1], q[2], q[3]); break;
    case ORC_C4X: c4x_gate(s, q[0], q[1], q[2], q[3], q[4]); break;
    case ORC_R: r_gate(s, g->theta, q[0]); break;
    case ORC_SRN: srn_gate(s, q[0]); break;
    case ORC_W: w_gate(s, q[0]); break;
    case ORC_RYY: ryy_gate(s, g->theta, q[0], q[1]); break;
    default: break;
    }
}

static int gate_valid(int n, const orc_gate* g)
{
    const int k = orc_gate_arity(g->op);
    if (k < 0) return 0;
    for (int i = 0; i < 5; i++)
        if (g->qb[i] < 0 || g->qb[i] >= n) return 0;      /* append, cpu_omp:245-249 */
    for (int i = 0; i < k; i++)
        for (int j = i + 1; j < k; j++)
            if (g->qb[i] == g->qb[j]) return 0;            /* non-cloning asserts */
    return 1;
}

/* Simulation::sim + simulation_kernel (cpu_omp:288-319, 590-596): one OpenMP
 * team walks the gate list; each primitive is an omp-for with a barrier. */
int orc_apply_gates(int n, double* re, double* im,
                    const orc_gate* gates, size_t ng, int nthreads)
{
    if (n < 1 || n > 40) return -1;
    for (size_t i = 0; i < ng; i++)
        if (!gate_valid(n, &gates[i])) return -1;
    sv_t s;
    s.dim = ONE << n;
    s.half_dim = s.dim >> 1;
    s.re = re;
    s.im = im;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
#endif
    {
        for (size_t i = 0; i < ng; i++)
            exe_op(&s, &gates[i]);
    }
    return 0;
}

/* ---------------- measurement (cpu_omp:339-361) ---------------- */

void orc_prob_scan(int n, const double* re, const double* im, double* scan)
{
    const uint64_t dim = ONE << n;
    scan[0] = 0;
    for (uint64_t i = 1; i < dim + 1; i++)
        scan[i] = scan[i - 1] + ((re[i - 1] * re[i - 1]) + (im[i - 1] * im[i - 1]));
}

/* j with scan[j] <= r < scan[j+1]; 0 when no interval matches.  scan is
 * non-decreasing, so the intervals are disjoint and the match (if any) is
 * m-1 where m is the first index in [1, dim] with scan[m] > r. */
static uint64_t search_scan(const double* scan, uint64_t dim, double r)
{
    if (!(scan[0] <= r)) return 0;
    uint64_t lo = 1, hi = dim + 1;            /* first m in [lo, hi) with scan[m] > r */
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if (scan[mid] > r) hi = mid; else lo = mid + 1;
    }
    if (lo == dim + 1) return 0;              /* r >= scan[dim] */
    return lo - 1;
}

double orc_measure(int n, const double* re, const double* im,
                   const double* uniforms, size_t shots, uint64_t* out)
{
    const uint64_t dim = ONE << n;
    double* scan = (double*)malloc((dim + 1) * sizeof(double));
    if (!scan) return -1.0;
    orc_prob_scan(n, re, im, scan);
    for (size_t i = 0; i < shots; i++)
        out[i] = search_scan(scan, dim, uniforms[i]);
    const double total = scan[dim];
    free(scan);
    return total;
}

double orc_measure_linear(int n, const double* re, const double* im,
                          const double* uniforms, size_t shots, uint64_t* out)
{
    const uint64_t dim = ONE << n;
    double* scan = (double*)malloc((dim + 1) * sizeof(double));
    if (!scan) return -1.0;
    orc_prob_scan(n, re, im, scan);
    for (size_t i = 0; i < shots; i++) {
        const double r = uniforms[i];
        out[i] = 0;
        for (uint64_t j = 0; j < dim; j++)
            if (scan[j] <= r && r < scan[j + 1]) out[i] = j;
    }
    const double total = scan[dim];
    free(scan);
    return total;
}

/* ---------------- Pauli expectations (no reference counterpart) ---------- */

/* P|x> = i^{|xm&zm|} (-1)^{|x&zm|} |x ^ xm>, so
 * <psi|P|psi> = sum_x conj(psi[x^xm]) i^{ny} (-1)^{|x&zm|} psi[x]. */
static void pauli_phase(uint64_t x, uint64_t xm, uint64_t zm, double* pr, double* pi_)
{
    const int ny = __builtin_popcountll(xm & zm) & 3;
    const int neg = __builtin_popcountll(x & zm) & 1;
    double r = 1, i = 0;
    switch (ny) { case 0: r = 1; i = 0; break; case 1: r = 0; i = 1; break;
                  case 2: r = -1; i = 0; break; default: r = 0; i = -1; break; }
    if (neg) { r = -r; i = -i; }
    *pr = r; *pi_ = i;
}

double orc_expval(int n, const double* re, const double* im,
                  const orc_pauli_term* terms, size_t nterms, int nthreads)
{
    const uint64_t dim = ONE << n;
    double total = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
    for (size_t t = 0; t < nterms; t++) {
        const uint64_t xm = terms[t].xmask, zm = terms[t].zmask;
        double acc = 0;
#ifdef _OPENMP
        #pragma omp parallel for reduction(+:acc) num_threads(nthreads) schedule(static)
#endif
        for (int64_t xx = 0; xx < (int64_t)dim; xx++) {
            const uint64_t x = (uint64_t)xx;
            double pr, pi_;
            pauli_phase(x, xm, zm, &pr, &pi_);
            const uint64_t y = x ^ xm;
            /* (P psi)[y] = phase * psi[x]; contribution Re(conj(psi[y]) * phase * psi[x]) */
            const double vr = pr * re[x] - pi_ * im[x];
            const double vi = pr * im[x] + pi_ * re[x];
            acc += re[y] * vr + im[y] * vi;
        }
        total += terms[t].coeff * acc;
    }
    return total;
}

void orc_apply_pauli_sum(int n, const double* re, const double* im,
                         const orc_pauli_term* terms, size_t nterms,
                         double* out_re, double* out_im)
{
    const uint64_t dim = ONE << n;
    memset(out_re, 0, dim * sizeof(double));
    memset(out_im, 0, dim * sizeof(double));
    for (size_t t = 0; t < nterms; t++) {
        const uint64_t xm = terms[t].xmask, zm = terms[t].zmask;
        const double c = terms[t].coeff;
        for (uint64_t x = 0; x < dim; x++) {
            double pr, pi_;
            pauli_phase(x, xm, zm, &pr, &pi_);
            const uint64_t y = x ^ xm;
            out_re[y] += c * (pr * re[x] - pi_ * im[x]);
            out_im[y] += c * (pr * im[x] + pi_ * re[x]);
        }
    }
}
