// ref_qir_driver.cpp -- C entry point around the REAL reference QIR-facing
// engine (svsim_qsharp/svsim_cpu_omp.hpp), compiled straight from the
// reference sources by oracle/Makefile into oracle/_ref/libsvsim_qir_ref.so;
// compiled a second time against the density-matrix engine
// (svsim_qsharp_noise/svsim_cpu_omp.hpp, same factory names) into
// oracle/_ref/libsvsim_dm_ref.so, where n qubits need capacity >= 2n (its
// update(), 467-475, runs on 2^(2n) entries).
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement
// (oracle/qir_oracle.c) and to write the golden fixtures under tests/golden/.
// Nothing in the product links it.
//
// The engine header is included unmodified, after the reference's own
// svsim_qsharp/util_cpu.h (timer and host allocation macros it uses).  The
// state is loaded and read through the public sv_real/sv_imag members
// (cpu_omp:215-232); gates go in through the public factories
// (cpu_omp:258-427); the run is Simulation::sim (467-505), called once per
// Measure as the QIR wrapper does (svsim_qsharp_wrapper.cpp:541-542).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <unistd.h>

#include "util_cpu.h"
#include "svsim_cpu_omp.hpp"

using namespace SVSim;

struct ref_qir_gate {
    int32_t op;
    int32_t qubit;
    uint64_t mask;
    double theta;
};

// Silences the per-run banner the density-matrix engine prints from sim()
// (svsim_qsharp_noise/config.hpp PRINT_MEA_PER_CIRCUIT, cpu_omp:508-514).
struct StdoutMute {
    int saved;
    StdoutMute() { fflush(stdout); saved = dup(1); int nul = open("/dev/null", O_WRONLY); dup2(nul, 1); close(nul); }
    ~StdoutMute() { fflush(stdout); dup2(saved, 1); close(saved); }
};

static void add_gate(Simulation& s, const ref_qir_gate& g)
{
    const IdxType q = IdxType(g.qubit), m = IdxType(g.mask);
    const ValType t = g.theta;
    switch (g.op) {
    case OP::X: s.X(q); break;
    case OP::Y: s.Y(q); break;
    case OP::Z: s.Z(q); break;
    case OP::H: s.H(q); break;
    case OP::S: s.S(q); break;
    case OP::T: s.T(q); break;
    case OP::RI: s.RI(t, q); break;
    case OP::RX: s.RX(t, q); break;
    case OP::RY: s.RY(t, q); break;
    case OP::RZ: s.RZ(t, q); break;
    case OP::EI: s.EI(t, q); break;
    case OP::EX: s.EX(t, q); break;
    case OP::EY: s.EY(t, q); break;
    case OP::EZ: s.EZ(t, q); break;
    case OP::ControlledX: s.ControlledX(q, m); break;
    case OP::ControlledY: s.ControlledY(q, m); break;
    case OP::ControlledZ: s.ControlledZ(q, m); break;
    case OP::ControlledH: s.ControlledH(q, m); break;
    case OP::ControlledS: s.ControlledS(q, m); break;
    case OP::ControlledT: s.ControlledT(q, m); break;
    case OP::ControlledRI: s.ControlledRI(t, q, m); break;
    case OP::ControlledRX: s.ControlledRX(t, q, m); break;
    case OP::ControlledRY: s.ControlledRY(t, q, m); break;
    case OP::ControlledRZ: s.ControlledRZ(t, q, m); break;
    case OP::ControlledEI: s.ControlledEI(t, q, m); break;
    case OP::ControlledEX: s.ControlledEX(t, q, m); break;
    case OP::ControlledEY: s.ControlledEY(t, q, m); break;
    case OP::ControlledEZ: s.ControlledEZ(t, q, m); break;
    case OP::AdjointS: s.AdjointS(q); break;
    case OP::AdjointT: s.AdjointT(q); break;
    case OP::ControlledAdjointS: s.ControlledAdjointS(q, m); break;
    case OP::ControlledAdjointT: s.ControlledAdjointT(q, m); break;
    case OP::Swap: s.Swap(q, m); break;
    case OP::Measure: s.Measure(q, t, m); break;
    default: break;
    }
}

extern "C" {

// capacity: qubits the Simulation allocates (state arrays of 2^capacity);
// n: qubits allocated through AllocateQubit (the run covers 2^n amplitudes).
// re/im: 2^capacity amplitudes in and out.  probs[m] / outcomes[m]: the m-th
// Measure's P(1) (Simulation::sim's return) and theta < P(1).
// Returns 0, or -1 on bad arguments or a gate the reference rejects.
int ref_qir_run(int capacity, int n, int nthreads, const ref_qir_gate* gates, size_t ng, double* re, double* im,
                double* probs, int32_t* outcomes)
{
    if (capacity < 1 || capacity >= N_QUBIT_SLOT || n < 1 || n > capacity || nthreads < 1) return -1;
    for (size_t i = 0; i < ng; i++) {
        if (gates[i].op < 0 || gates[i].op > OP::Measure || gates[i].qubit < 0 || gates[i].qubit >= n) return -1;
        if (gates[i].op == OP::Swap && (gates[i].mask >= uint64_t(n) || gates[i].mask == uint64_t(gates[i].qubit)))
            return -1;
    }
    Simulation s(static_cast<IdxType>(capacity), static_cast<IdxType>(nthreads));
    const size_t dim = size_t(1) << capacity;
    std::memcpy(s.sv_real, re, dim * sizeof(double));
    std::memcpy(s.sv_imag, im, dim * sizeof(double));
    for (int i = 0; i < n; i++) s.AllocateQubit();
    size_t m = 0;
    StdoutMute mute;
    try {
        for (size_t i = 0; i < ng; i++) {
            add_gate(s, gates[i]);
            if (gates[i].op == OP::Measure) {
                const double p = s.sim();
                if (probs) probs[m] = p;
                if (outcomes) outcomes[m] = gates[i].theta < p ? 1 : 0;
                m++;
            }
        }
        s.sim();
    } catch (...) {
        return -1;
    }
    std::memcpy(re, s.sv_real, dim * sizeof(double));
    std::memcpy(im, s.sv_imag, dim * sizeof(double));
    return 0;
}

}  // extern "C"
