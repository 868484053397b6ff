/*
 * qir_oracle.c -- CPU oracle: plain-C restatement of the reference's
 * QIR-facing engine, svsim_qsharp/svsim_cpu_omp.hpp.
 *
 * TEST INFRASTRUCTURE ONLY (see qir_oracle.h).  Each primitive cites the
 * reference lines it restates and keeps their expression order, so that,
 * compiled with -ffp-contract=off and linked against the same libm,
 * amplitudes are bit-identical to the reference; tests/test_qir_oracle.py
 * pins this against fixtures written by the reference itself (oracle/_ref).
 *
 * Differences from the reference, none of which change results:
 *   - the controlled and uncontrolled forms of a gate share one loop: an
 *     uncontrolled gate runs with mask 0, for which OP_HEAD_MASK's skip test
 *     (cpu_omp:606-615) never fires, i.e. it is OP_HEAD (594-603);
 *   - the Measure reduction's scratch array is allocated per measurement
 *     instead of living in the Simulation object (m_real, cpu_omp:221).
 */
#include "qir_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* svsim_qsharp/config.hpp:67 */
#define S2I 0.70710678118654752440
#define ONE ((uint64_t)1)

typedef struct {
    uint64_t dim, half_dim;
    double* re;
    double* im;
} qsv_t;

/* OP_HEAD / OP_HEAD_MASK (cpu_omp:594-615): pair (pos0, pos1) of qubit q
 * for i < 2^(n-1), skipped unless pos0 has every bit of mask set. */
#define QPAIR_BEGIN(sv, q, mask)                                             \
    _Pragma("omp for schedule(static)")                                      \
    for (int64_t ii = 0; ii < (int64_t)(sv)->half_dim; ii++) {               \
        const uint64_t i_ = (uint64_t)ii;                                    \
        const uint64_t pos0 = ((i_ >> (q)) << ((q) + 1)) + (i_ & ((ONE << (q)) - 1)); \
        const uint64_t pos1 = pos0 + (ONE << (q));                           \
        if (((~(pos0 & (mask))) & (mask)) != 0) continue;
#define QPAIR_END }

/* X_GATE / ControlledX_GATE, cpu_omp:623-640, 937-954 */
static void q_x(qsv_t* s, int q, uint64_t mask)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        const double r0 = R[pos0], i0 = I[pos0], r1 = R[pos1], i1 = I[pos1];
        R[pos0] = r1; I[pos0] = i1;
        R[pos1] = r0; I[pos1] = i0;
    QPAIR_END
}

/* Y_GATE / ControlledY_GATE, cpu_omp:644-660, 958-975 */
static void q_y(qsv_t* s, int q, uint64_t mask)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        const double r0 = R[pos0], i0 = I[pos0], r1 = R[pos1], i1 = I[pos1];
        R[pos0] = i1;  I[pos0] = -r1;
        R[pos1] = -i0; I[pos1] = r0;
    QPAIR_END
}

/* Z_GATE / ControlledZ_GATE, cpu_omp:664-676, 979-992 (pos1 only) */
static void q_z(qsv_t* s, int q, uint64_t mask)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        (void)pos0;
        const double r1 = R[pos1], i1 = I[pos1];
        R[pos1] = -r1; I[pos1] = -i1;
    QPAIR_END
}

/* H_GATE / ControlledH_GATE / measurement H, cpu_omp:680-696, 996-1013, 1361-1373 */
static void q_h(qsv_t* s, int q, uint64_t mask)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        const double r0 = R[pos0], i0 = I[pos0], r1 = R[pos1], i1 = I[pos1];
        R[pos0] = S2I * (r0 + r1); I[pos0] = S2I * (i0 + i1);
        R[pos1] = S2I * (r0 - r1); I[pos1] = S2I * (i0 - i1);
    QPAIR_END
}

/* S_GATE / ControlledS_GATE / measurement S, cpu_omp:700-712, 1017-1030, 1389-1400 */
static void q_s(qsv_t* s, int q, uint64_t mask)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        (void)pos0;
        const double r1 = R[pos1], i1 = I[pos1];
        R[pos1] = -i1; I[pos1] = r1;
    QPAIR_END
}

/* AdjointS_GATE / ControlledAdjointS_GATE / measurement AdjointS,
 * cpu_omp:1243-1255, 1275-1288, 1377-1386 */
static void q_sdg(qsv_t* s, int q, uint64_t mask)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        (void)pos0;
        const double r1 = R[pos1], i1 = I[pos1];
        R[pos1] = i1; I[pos1] = -r1;
    QPAIR_END
}

/* T_GATE / ControlledT_GATE, cpu_omp:716-728, 1034-1047 */
static void q_t(qsv_t* s, int q, uint64_t mask)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        (void)pos0;
        const double r1 = R[pos1], i1 = I[pos1];
        R[pos1] = S2I * (r1 - i1); I[pos1] = S2I * (r1 + i1);
    QPAIR_END
}

/* AdjointT_GATE / ControlledAdjointT_GATE, cpu_omp:1259-1271, 1292-1306 */
static void q_tdg(qsv_t* s, int q, uint64_t mask)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        (void)pos0;
        const double r1 = R[pos1], i1 = I[pos1];
        R[pos1] = S2I * (r1 + i1); I[pos1] = S2I * (-r1 + i1);
    QPAIR_END
}

/* a0 *= (d0r + i d0i), a1 *= (d1r + i d1i): RI, RZ, EI, EZ and their
 * controlled forms (cpu_omp:733-753, 806-829, 833-855, 911-933, 1052-1069,
 * 1117-1138, 1142-1163, 1217-1238), each written as
 * re = (x.re*e.re) - (x.im*e.im), im = (x.re*e.im) + (x.im*e.re). */
static void q_diag(qsv_t* s, int q, uint64_t mask, double d0r, double d0i, double d1r, double d1i)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        const double r0 = R[pos0], i0 = I[pos0], r1 = R[pos1], i1 = I[pos1];
        R[pos0] = (r0 * d0r) - (i0 * d0i);
        I[pos0] = (r0 * d0i) + (i0 * d0r);
        R[pos1] = (r1 * d1r) - (i1 * d1i);
        I[pos1] = (r1 * d1i) + (i1 * d1r);
    QPAIR_END
}

/* [[c, i s], [i s, c]]: RX (c = cos(t/2), s = -sin(t/2)) and EX (c = cos t,
 * s = sin t), plain and controlled (cpu_omp:757-776, 859-881, 1073-1090,
 * 1167-1188). */
static void q_xrot(qsv_t* s, int q, uint64_t mask, double c, double sn)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        const double r0 = R[pos0], i0 = I[pos0], r1 = R[pos1], i1 = I[pos1];
        R[pos0] = (c * r0) - (sn * i1);
        I[pos0] = (c * i0) + (sn * r1);
        R[pos1] = -(sn * i0) + (c * r1);
        I[pos1] = (sn * r0) + (c * i1);
    QPAIR_END
}

/* real [[e0, e1], [e2, e3]]: RY (cos/2, -sin/2, sin/2, cos/2) and EY
 * (cos, sin, -sin, cos), plain and controlled (cpu_omp:780-802, 885-907,
 * 1094-1113, 1192-1213). */
static void q_real2(qsv_t* s, int q, uint64_t mask, double e0, double e1, double e2, double e3)
{
    double* R = s->re; double* I = s->im;
    QPAIR_BEGIN(s, q, mask)
        const double r0 = R[pos0], i0 = I[pos0], r1 = R[pos1], i1 = I[pos1];
        R[pos0] = (e0 * r0) + (e1 * r1);
        I[pos0] = (e0 * i0) + (e1 * i1);
        R[pos1] = (e2 * r0) + (e3 * r1);
        I[pos1] = (e2 * i0) + (e3 * i1);
    QPAIR_END
}

/* SWAP_GATE, cpu_omp:1312-1357: exchange (q1=1, q2=0) <-> (q1=0, q2=1). */
static void q_swap(qsv_t* s, int q1, int q2)
{
    double* R = s->re; double* I = s->im;
    const int lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
    const int64_t quarter = (int64_t)(s->dim >> 2);
    #pragma omp for schedule(static)
    for (int64_t ii = 0; ii < quarter; ii++) {
        uint64_t b = (uint64_t)ii;
        b = ((b >> lo) << (lo + 1)) | (b & ((ONE << lo) - 1));
        b = ((b >> hi) << (hi + 1)) | (b & ((ONE << hi) - 1));
        const uint64_t pa = b | (ONE << q1), pb = b | (ONE << q2);
        const double ra = R[pa], ia = I[pa];
        R[pa] = R[pb]; I[pa] = I[pb];
        R[pb] = ra; I[pb] = ia;
    }
}

/* Measure_GATE, cpu_omp:1406-1500: basis change (X: H; Y: AdjointS then
 * H), m[i] = |a_i|^2 where bit q is set (else 0), halving-tree sum into
 * m[0] = P(1), collapse with the draw `rand` (rand < P(1): bit-0 amplitudes
 * zeroed, bit-1 ones divided by sqrt(P(1)); otherwise the converse with
 * sqrt(1 - P(1))), then the inverse basis change (X: H; Y: H then S). */
static void q_measure(qsv_t* s, int q, double rand_u, uint64_t pauli, double* m, double* p_out)
{
    double* R = s->re; double* I = s->im;
    const uint64_t bit = ONE << q;
    if (pauli == 1) q_h(s, q, 0);
    if (pauli == 2) {
        q_sdg(s, q, 0);
        q_h(s, q, 0);
    }
    #pragma omp for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->dim; ii++) {
        const uint64_t i = (uint64_t)ii;
        m[i] = (i & bit) == 0 ? 0. : (R[i] * R[i]) + (I[i] * I[i]);
    }
    for (uint64_t k = s->half_dim; k > 0; k >>= 1) {
        #pragma omp for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)k; ii++) m[ii] += m[ii + k];
    }
    const double p = m[0];
    #pragma omp barrier
    if (rand_u < p) {
        const double sq = sqrt(p);
        #pragma omp for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)s->dim; ii++) {
            const uint64_t i = (uint64_t)ii;
            if ((i & bit) == 0) {
                R[i] = 0.;
                I[i] = 0.;
            } else {
                R[i] /= sq;
                I[i] /= sq;
            }
        }
    } else {
        const double sq = sqrt(1.0 - p);
        #pragma omp for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)s->dim; ii++) {
            const uint64_t i = (uint64_t)ii;
            if ((i & bit) == 0) {
                R[i] /= sq;
                I[i] /= sq;
            } else {
                R[i] = 0.;
                I[i] = 0.;
            }
        }
    }
    if (pauli == 1) q_h(s, q, 0);
    if (pauli == 2) {
        q_h(s, q, 0);
        q_s(s, q, 0);
    }
    *p_out = p;   /* every thread read the same m[0] */
}

/* One gate (the switch of function pointers, cpu_omp:2378-2513). */
static void q_exec(qsv_t* s, const orc_qir_gate* g, double* m, double* p_out)
{
    const int q = g->qubit;
    const double t = g->theta;
    const int ctl = (g->op >= ORQ_CX && g->op <= ORQ_CEZ) || g->op == ORQ_CSDG || g->op == ORQ_CTDG;
    const uint64_t mask = ctl ? g->mask : 0;
    switch (g->op) {
    case ORQ_X: case ORQ_CX: q_x(s, q, mask); break;
    case ORQ_Y: case ORQ_CY: q_y(s, q, mask); break;
    case ORQ_Z: case ORQ_CZ: q_z(s, q, mask); break;
    case ORQ_H: case ORQ_CH: q_h(s, q, mask); break;
    case ORQ_S: case ORQ_CS: q_s(s, q, mask); break;
    case ORQ_T: case ORQ_CT: q_t(s, q, mask); break;
    case ORQ_SDG: case ORQ_CSDG: q_sdg(s, q, mask); break;
    case ORQ_TDG: case ORQ_CTDG: q_tdg(s, q, mask); break;
    case ORQ_RI: case ORQ_CRI: {
        const double c = cos(t / 2.0), sn = -sin(t / 2.0);
        q_diag(s, q, mask, c, sn, c, sn);
        break;
    }
    case ORQ_RZ: case ORQ_CRZ: {
        const double c = cos(t / 2), sn = sin(t / 2);
        q_diag(s, q, mask, c, -sn, c, sn);
        break;
    }
    case ORQ_EI: case ORQ_CEI: {
        const double c = cos(t), sn = sin(t);
        q_diag(s, q, mask, c, sn, c, sn);
        break;
    }
    case ORQ_EZ: case ORQ_CEZ: {
        const double c = cos(t), sn = sin(t);
        q_diag(s, q, mask, c, sn, c, -sn);
        break;
    }
    case ORQ_RX: case ORQ_CRX: q_xrot(s, q, mask, cos(t / 2.0), -sin(t / 2.0)); break;
    case ORQ_EX: case ORQ_CEX: q_xrot(s, q, mask, cos(t), sin(t)); break;
    case ORQ_RY: case ORQ_CRY: {
        const double c = cos(t / 2.0), sn = sin(t / 2.0);
        q_real2(s, q, mask, c, -sn, sn, c);
        break;
    }
    case ORQ_EY: case ORQ_CEY: {
        const double c = cos(t), sn = sin(t);
        q_real2(s, q, mask, c, sn, -sn, c);
        break;
    }
    case ORQ_SWAP: q_swap(s, q, (int)g->mask); break;
    case ORQ_MEASURE: q_measure(s, q, t, g->mask, m, p_out); break;
    default: break;
    }
}

int orc_qir_apply(int n, double* re, double* im, const orc_qir_gate* gates, size_t ng,
                  double* last_prob, double* probs, int32_t* outcomes, int nthreads)
{
    if (n < 1 || n > 40) return -1;
    size_t n_meas = 0;
    for (size_t i = 0; i < ng; i++) {
        const orc_qir_gate* g = &gates[i];
        if (g->op < 0 || g->op >= ORQ_NUM_OPS) return -1;
        if (g->qubit < 0 || g->qubit >= n) return -1;
        if (g->op == ORQ_SWAP && (g->mask >= (uint64_t)n || g->mask == (uint64_t)g->qubit)) return -1;
        if (g->op == ORQ_MEASURE) n_meas++;
    }
    qsv_t s;
    s.dim = ONE << n;
    s.half_dim = s.dim >> 1;
    s.re = re;
    s.im = im;
    double* m = NULL;
    if (n_meas) {
        m = (double*)malloc(s.dim * sizeof(double));
        if (!m) return -1;
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
#endif
    {
        size_t my_mi = 0;
        for (size_t i = 0; i < ng; i++) {
            double p = 0.0;
            q_exec(&s, &gates[i], m, &p);
            #pragma omp barrier
            if (gates[i].op == ORQ_MEASURE) {
                #pragma omp single
                {
                    if (probs) probs[my_mi] = p;
                    if (outcomes) outcomes[my_mi] = gates[i].theta < p ? 1 : 0;
                    if (last_prob) *last_prob = p;
                }
                my_mi++;
            }
        }
    }
    free(m);
    return 0;
}


/* ---- density-matrix engine: svsim_qsharp_noise/svsim_cpu_omp.hpp ----
 * rho of n qubits as a 2n-qubit vector, index = row | (col << n), from
 * rho = |0><0| (cpu_omp:218-240; update(), 467-475, sizes it 2^(2n)).  Every
 * gate is U on (q, mask) followed by conj(U) on (q + n, mask << n) (the
 * *_OP table, 2785-2990: Y -> ConjugateY, S -> AdjointS, RX -> ConjugateRX,
 * ..., real gates repeated); conj(U) is applied here as conj(U conj(x)),
 * exact sign flips around the same gate routine.  Swap exchanges the row
 * qubits only (SWAP_GATE, 1560-1606).  The noise superoperators of
 * noise_gate_*.cuh are not applied: their calls are commented out at every
 * call site (2792-2840). */

static void q_conj(qsv_t* s)
{
    #pragma omp for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->dim; ii++) s->im[ii] = -s->im[ii];
}

/* Measure_GATE (1710-1830): basis change on both halves (X: H, H; Y:
 * AdjointS(q), S(q + n), H, H), m[i] = |Re rho[i, i]| for the rows i < 2^n
 * with bit q set (0 elsewhere; the rest of m stays zero), halving-tree sum
 * over all 2^(2n) entries into m[0] = P(1), then with rand < P(1) every
 * entry whose row bit q is 0 zeroed and the others divided by P(1) (else the
 * converse with 1 - P(1)) -- the row index only -- and the inverse change
 * (X: H, H; Y: H, H, S(q), AdjointS(q + n)). */
static void q_dm_measure(qsv_t* s, int n, int q, double rand_u, uint64_t pauli, double* m, double* p_out)
{
    double* R = s->re; double* I = s->im;
    const uint64_t bit = ONE << q, rows = ONE << n;
    if (pauli == 1) {
        q_h(s, q, 0);
        q_h(s, q + n, 0);
    }
    if (pauli == 2) {
        q_sdg(s, q, 0);
        q_s(s, q + n, 0);
        q_h(s, q, 0);
        q_h(s, q + n, 0);
    }
    #pragma omp for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->dim; ii++) {
        const uint64_t i = (uint64_t)ii;
        m[i] = (i >= rows || (i & bit) == 0) ? 0. : fabs(R[(i << n) + i]);
    }
    for (uint64_t k = s->half_dim; k > 0; k >>= 1) {
        #pragma omp for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)k; ii++) m[ii] += m[ii + k];
    }
    const double p = m[0];
    #pragma omp barrier
    const int one = rand_u < p;
    const double f = one ? p : 1.0 - p;
    #pragma omp for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s->dim; ii++) {
        const uint64_t i = (uint64_t)ii;
        if (((i & bit) != 0) == one) {
            R[i] /= f;
            I[i] /= f;
        } else {
            R[i] = 0.;
            I[i] = 0.;
        }
    }
    if (pauli == 1) {
        q_h(s, q, 0);
        q_h(s, q + n, 0);
    }
    if (pauli == 2) {
        q_h(s, q, 0);
        q_h(s, q + n, 0);
        q_s(s, q, 0);
        q_sdg(s, q + n, 0);
    }
    *p_out = p;
}

int orc_dm_qir_apply(int n, double* re, double* im, const orc_qir_gate* gates, size_t ng, double* probs,
                     int32_t* outcomes, int nthreads)
{
    if (n < 1 || n > 20) return -1;
    for (size_t i = 0; i < ng; i++) {
        const orc_qir_gate* g = &gates[i];
        if (g->op < 0 || g->op >= ORQ_NUM_OPS) return -1;
        if (g->qubit < 0 || g->qubit >= n) return -1;
        if (g->op == ORQ_SWAP && (g->mask >= (uint64_t)n || g->mask == (uint64_t)g->qubit)) return -1;
    }
    qsv_t s;
    s.dim = ONE << (2 * n);
    s.half_dim = s.dim >> 1;
    s.re = re;
    s.im = im;
    double* m = (double*)calloc(s.dim, sizeof(double));
    if (!m) return -1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
#endif
    {
        size_t my_mi = 0;
        for (size_t i = 0; i < ng; i++) {
            const orc_qir_gate* g = &gates[i];
            if (g->op == ORQ_MEASURE) {
                double p = 0.0;
                q_dm_measure(&s, n, g->qubit, g->theta, g->mask, m, &p);
                #pragma omp barrier
                #pragma omp single
                {
                    if (probs) probs[my_mi] = p;
                    if (outcomes) outcomes[my_mi] = g->theta < p ? 1 : 0;
                }
                my_mi++;
                continue;
            }
            if (g->op == ORQ_SWAP) {
                q_swap(&s, g->qubit, (int)g->mask);
                continue;
            }
            q_exec(&s, g, NULL, NULL);
            orc_qir_gate c = *g;
            c.qubit += n;
            c.mask = g->mask << n;
            q_conj(&s);
            q_exec(&s, &c, NULL, NULL);
            q_conj(&s);
        }
    }
    free(m);
    return 0;
}

