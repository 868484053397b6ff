// ref_driver.cpp -- C entry points around the REAL reference CPU-OpenMP
// engine (svsim/src/svsim_cpu_omp.hpp), compiled straight from the reference
// sources by oracle/Makefile into oracle/_ref/libsvsim_ref.so.
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement and to write
// the golden fixtures under tests/golden/.  Nothing in the product links it.
//
// The engine header is included unmodified.  Two hooks, both outside it:
//   * measure() calls srand(RAND_SEED)/rand() unqualified from inside
//     namespace SVSim (cpu_omp:347-351).  Declaring SVSim::rand/SVSim::srand
//     before the include makes those calls resolve here, so the caller can
//     feed the integer draws k_i and the reference computes r = k_i/RAND_MAX
//     itself.
//   * the state is loaded/read through the public sv_real/sv_imag members
//     (cpu_omp:576-577), the only state access the reference offers.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <unistd.h>
#include <fcntl.h>

namespace SVSim {
static const int* g_draws = nullptr;
static size_t g_ndraws = 0;
static size_t g_pos = 0;
int rand() { return g_pos < g_ndraws ? g_draws[g_pos++] : 0; }
void srand(unsigned) { g_pos = 0; }
}  // namespace SVSim

#include "util_cpu.h"
#include <chrono>

#include "svsim_cpu_omp.hpp"

using namespace SVSim;

struct ref_gate {
    int32_t op;
    int32_t qb[5];
    double theta, phi, lambda;
};

static int arity(int op)
{
    switch (op) {
    case CX: case CZ: case CY: case SWAP: case CH: case CRX: case CRY: case CRZ:
    case CU1: case CU3: case RXX: case RZZ: case RYY: return 2;
    case CCX: case CSWAP: case RCCX: return 3;
    case RC3X: case C3X: case C3SQRTX: return 4;
    case C4X: return 5;
    default: return (op >= 0 && op <= RYY) ? 1 : -1;
    }
}

static bool valid(int n, const ref_gate& g)
{
    const int k = arity(g.op);
    if (k < 0) return false;
    for (int i = 0; i < 5; i++)
        if (g.qb[i] < 0 || g.qb[i] >= n) return false;
    for (int i = 0; i < k; i++)
        for (int j = i + 1; j < k; j++)
            if (g.qb[i] == g.qb[j]) return false;
    return true;
}

// Silence the per-run banner that sim() prints (config.hpp:25).
struct StdoutMute {
    int saved;
    StdoutMute() { fflush(stdout); saved = dup(1); int nul = open("/dev/null", O_WRONLY); dup2(nul, 1); close(nul); }
    ~StdoutMute() { fflush(stdout); dup2(saved, 1); close(saved); }
};

extern "C" {

// Apply gates to |0..0> (init_re == NULL) or to the given state; write the
// final state to out_re/out_im.  n <= 28: the reference's 32-bit IdxType
// overflows sv_size at n = 29 (config.hpp:40, cpu_omp:206).
int ref_run(int n, int nthreads, const ref_gate* gates, size_t ng,
            const double* init_re, const double* init_im,
            double* out_re, double* out_im)
{
    if (n < 1 || n > 28 || nthreads < 1) return -1;
    for (size_t i = 0; i < ng; i++)
        if (!valid(n, gates[i])) return -1;
    Simulation sim((IdxType)n, (IdxType)nthreads);
    if (init_re) {
        memcpy(sim.sv_real, init_re, sizeof(double) << n);
        memcpy(sim.sv_imag, init_im, sizeof(double) << n);
    }
    for (size_t i = 0; i < ng; i++) {
        const ref_gate& g = gates[i];
        Gate gate((OP)g.op, (IdxType)g.qb[0], (IdxType)g.qb[1], (IdxType)g.qb[2],
                  (IdxType)g.qb[3], (IdxType)g.qb[4], g.theta, g.phi, g.lambda);
        sim.append(&gate);
    }
    sim.upload();
    {
        StdoutMute mute;
        sim.sim();
    }
    memcpy(out_re, sim.sv_real, sizeof(double) << n);
    memcpy(out_im, sim.sv_imag, sizeof(double) << n);
    return 0;
}

// Simulation::measure on the given state with the integer draws fed to its
// rand() calls (r = draw / RAND_MAX, cpu_omp:351).
int ref_measure(int n, const double* re, const double* im,
                const int* draws, size_t shots, uint64_t* out)
{
    if (n < 1 || n > 28) return -1;
    Simulation sim((IdxType)n, 1);
    memcpy(sim.sv_real, re, sizeof(double) << n);
    memcpy(sim.sv_imag, im, sizeof(double) << n);
    g_draws = draws;
    g_ndraws = shots;
    g_pos = 0;
    IdxType* res;
    {
        StdoutMute mute;
        res = sim.measure((unsigned)shots);
    }
    for (size_t i = 0; i < shots; i++) out[i] = res[i];
    delete[] res;
    g_draws = nullptr;
    g_ndraws = 0;
    return 0;
}

// Wall time of Simulation::sim() alone (the part the reference times,
// cpu_omp:288-319) for `reps` runs of the circuit, each from |0..0>
// (reset_sv, cpu_omp:234-240): ms[r] in milliseconds.
int ref_sim_times(int n, int nthreads, const ref_gate* gates, size_t ng, int reps, double* ms)
{
    if (n < 1 || n > 28 || nthreads < 1 || reps < 1) return -1;
    for (size_t i = 0; i < ng; i++)
        if (!valid(n, gates[i])) return -1;
    Simulation sim((IdxType)n, (IdxType)nthreads);
    for (size_t i = 0; i < ng; i++) {
        const ref_gate& g = gates[i];
        Gate gate((OP)g.op, (IdxType)g.qb[0], (IdxType)g.qb[1], (IdxType)g.qb[2],
                  (IdxType)g.qb[3], (IdxType)g.qb[4], g.theta, g.phi, g.lambda);
        sim.append(&gate);
    }
    sim.upload();
    for (int r = 0; r < reps; r++) {
        sim.reset_sv();
        StdoutMute mute;
        const auto t0 = std::chrono::steady_clock::now();
        sim.sim();
        const auto t1 = std::chrono::steady_clock::now();
        ms[r] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    return 0;
}

int ref_rand_max(void) { return RAND_MAX; }

}  // extern "C"
