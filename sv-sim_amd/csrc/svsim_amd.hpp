// svsim_amd.hpp -- the reference C++ API (namespace SVSim: Gate, Simulation)
// on top of the MI355X engine's C ABI (include/svmi.h).
//
// Mirrors svsim/src/svsim_cpu_omp.hpp:103-588 and svsim_amdgpu_sin.hpp:
// same enum OP, Gate record, Simulation constructor/append/upload/sim/
// measure/reset/clear_circuit and the 38 static gate factories with the
// reference's argument order and slot assignment.  Differences:
//   * errors throw std::runtime_error instead of assert/exit;
//   * IdxType is 64-bit (the reference's 32-bit IdxType stops at 28 qubits);
//   * the state lives on the GPU; get_sv/set_sv replace the public
//     sv_real/sv_imag members, and sim() makes no host copy.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <stdexcept>
#include <string>
#include <vector>

#include "svmi.h"

// config.hpp:44: the seed applications pass to srand.
#ifndef RAND_SEED
#define RAND_SEED ((unsigned)time(0))
#endif

namespace SVSim {

using IdxType = unsigned long long;
using ValType = double;

// svsim_cpu_omp.hpp:46-52
enum OP {
    U3, U2, U1, CX, ID, X, Y, Z, H, S,
    SDG, T, TDG, RX, RY, RZ, CZ, CY, SWAP, CH,
    CCX, CSWAP, CRX, CRY, CRZ, CU1, CU3, RXX, RZZ, RCCX,
    RC3X, C3X, C3SQRTX, C4X, R, SRN, W, RYY
};

inline const char* const* op_names()
{
    static const char* const names[] = {
        "U3", "U2", "U1", "CX", "ID", "X", "Y", "Z", "H", "S",
        "SDG", "T", "TDG", "RX", "RY", "RZ", "CZ", "CY", "SWAP", "CH",
        "CCX", "CSWAP", "CRX", "CRY", "CRZ", "CU1", "CU3", "RXX", "RZZ", "RCCX",
        "RC3X", "C3X", "C3SQRTX", "C4X", "R", "SRN", "W", "RYY"};
    return names;
}

inline void svmi_throw(svmi_status st, const char* what)
{
    if (st != SVMI_OK) throw std::runtime_error(std::string(what) + ": " + svmi_last_error());
}

// svsim_cpu_omp.hpp:103-194 (without the function pointer)
class Gate {
public:
    Gate(OP _op_name, IdxType _qb0, IdxType _qb1, IdxType _qb2, IdxType _qb3, IdxType _qb4,
         ValType _theta, ValType _phi, ValType _lambda)
        : op_name(_op_name), qb0(_qb0), qb1(_qb1), qb2(_qb2), qb3(_qb3), qb4(_qb4),
          theta(_theta), phi(_phi), lambda(_lambda) {}

    svmi_gate record() const
    {
        svmi_gate g;
        g.op = int32_t(op_name);
        g.qb[0] = int32_t(qb0); g.qb[1] = int32_t(qb1); g.qb[2] = int32_t(qb2);
        g.qb[3] = int32_t(qb3); g.qb[4] = int32_t(qb4);
        g.theta = theta; g.phi = phi; g.lambda = lambda;
        return g;
    }
    std::string dump() const
    {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s(%llu,%llu,%llu,%llu,%llu,%g,%g,%g);", op_names()[op_name], qb0, qb1, qb2,
                 qb3, qb4, theta, phi, lambda);
        return buf;
    }

    OP op_name;
    IdxType qb0, qb1, qb2, qb3, qb4;
    ValType theta, phi, lambda;
};

class Simulation {
public:
    // svsim_amdgpu_sin.hpp:207-239: the second argument is the GPU ordinal.
    Simulation(IdxType _n_qubits, IdxType _i_gpu = 0)
        : n_qubits(_n_qubits), i_gpu(_i_gpu), dim(IdxType(1) << _n_qubits), n_gates(0)
    {
        if (_n_qubits < 1 || _n_qubits > 40) throw std::runtime_error("n_qubits must be in [1, 40]");
        svmi_throw(svmi_create(int(_n_qubits), int(_i_gpu), &h), "Simulation");
    }
    ~Simulation() { svmi_destroy(h); }
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    // svsim_cpu_omp.hpp:229-240
    void reset()
    {
        clear_circuit();
        reset_sv();
    }
    void reset_sv() { svmi_throw(svmi_reset_state(h), "reset_sv"); }

    // svsim_cpu_omp.hpp:242-255: bounds check of all five slots, then copy.
    void append(const Gate* g)
    {
        if (!g) throw std::runtime_error("append: null gate");
        if (g->qb0 >= n_qubits || g->qb1 >= n_qubits || g->qb2 >= n_qubits || g->qb3 >= n_qubits ||
            g->qb4 >= n_qubits)
            throw std::runtime_error("append: qubit index out of range");
        circuit.push_back(*g);
        n_gates++;
    }
    // OpenQASM 2.0 source -> appended gates (no reference counterpart: the
    // reference translates .qasm files to Python scripts of append calls,
    // svsim/QASMBench/svsim_qasm.py).  Returns the number of gates appended.
    size_t append_qasm(const std::string& text)
    {
        svmi_qasm_info info;
        svmi_throw(svmi_qasm_parse(text.data(), text.size(), nullptr, 0, &info), "append_qasm");
        if (IdxType(info.n_qubits) > n_qubits)
            throw std::runtime_error("append_qasm: the circuit needs " + std::to_string(info.n_qubits) + " qubits");
        std::vector<svmi_gate> recs(info.n_gates);
        svmi_throw(svmi_qasm_parse(text.data(), text.size(), recs.data(), recs.size(), &info), "append_qasm");
        for (const svmi_gate& r : recs) {
            const Gate g(OP(r.op), IdxType(r.qb[0]), IdxType(r.qb[1]), IdxType(r.qb[2]), IdxType(r.qb[3]),
                         IdxType(r.qb[4]), r.theta, r.phi, r.lambda);
            append(&g);
        }
        return recs.size();
    }
    // svsim_cpu_omp.hpp:256-275
    Simulation* upload()
    {
        std::vector<svmi_gate> recs;
        recs.reserve(circuit.size());
        for (const Gate& g : circuit) recs.push_back(g.record());
        svmi_throw(svmi_upload(h, recs.data(), recs.size()), "upload");
        return this;
    }
    std::string dump() const
    {
        std::string s;
        for (const Gate& g : circuit) s += g.dump() + "\n";
        return s;
    }
    // svsim_cpu_omp.hpp:288-319 / svsim_amdgpu_sin.hpp:320-353
    void sim()
    {
        svmi_throw(svmi_run(h), "sim");
        svmi_stats st;
        svmi_get_stats(h, &st);
        last_ms = st.device_ms;
        const char* quiet = std::getenv("SVSIM_QUIET");
        if (!quiet || quiet[0] == '0') {
            const double mem = double(dim) * 16.0 / 1024.0 / 1024.0;
            printf("\n============== SV-Sim: AMD-MI355X ===============\n");
            printf("nqubits:%llu, ngates:%llu, ncores:%d, sim:%.3lf ms, mem:%.3lf MB, mem_per_gpu:%.3lf MB\n",
                   n_qubits, n_gates, 1, last_ms, mem, mem);
            printf("=================================================\n");
        }
    }
    // svsim_cpu_omp.hpp:321-338: the state is kept.
    void clear_circuit()
    {
        circuit.clear();
        n_gates = 0;
        svmi_throw(svmi_clear_circuit(h), "clear_circuit");
    }
    // svsim_cpu_omp.hpp:339-361: draws r = rand()/RAND_MAX after
    // srand(RAND_SEED) (config.hpp:44: time(0) unless the build defines it).
    // Returns new IdxType[repetition], owned by the caller (delete[]), as
    // the reference's measure returns its array.
    IdxType* measure(unsigned repetition = 10)
    {
        std::vector<double> u(repetition);
        srand(unsigned(RAND_SEED));
        for (unsigned i = 0; i < repetition; i++) u[i] = double(rand()) / double(RAND_MAX);
        const std::vector<IdxType> r = measure_with(u);
        IdxType* res_state = new IdxType[repetition];
        std::copy(r.begin(), r.end(), res_state);
        return res_state;
    }
    // measure with caller-supplied draws (same interval rule)
    std::vector<IdxType> measure_with(const std::vector<double>& uniforms)
    {
        std::vector<uint64_t> out(uniforms.size());
        double total = 0.0;
        svmi_throw(svmi_measure(h, uniforms.data(), uniforms.size(), out.data(), &total), "measure");
        if (std::fabs(total - 1.0) > 1e-3)  // ERROR_BAR, config.hpp:50
            printf("Sum of probability is far from 1.0 with %lf\n", total);
        return std::vector<IdxType>(out.begin(), out.end());
    }
    void get_sv(double* re, double* im) const { svmi_throw(svmi_get_state(h, re, im), "get_sv"); }
    void set_sv(const double* re, const double* im) { svmi_throw(svmi_set_state(h, re, im), "set_sv"); }
    double expval(const std::vector<svmi_pauli_term>& terms) const
    {
        double e = 0.0;
        svmi_throw(svmi_expval(h, terms.data(), terms.size(), &e), "expval");
        return e;
    }
    double adjoint(const std::vector<svmi_pauli_term>& terms, const std::vector<int32_t>& param_slots,
                   size_t n_params, std::vector<double>& grad)
    {
        grad.assign(n_params, 0.0);
        double e = 0.0;
        svmi_throw(svmi_adjoint(h, terms.data(), terms.size(), param_slots.data(), n_params, &e, grad.data()),
                   "adjoint");
        return e;
    }
    svmi_handle handle() const { return h; }

    // =============== gate factories (svsim_cpu_omp.hpp:377-564) ===============
    static Gate* U3(ValType theta, ValType phi, ValType lambda, IdxType m) { return new Gate(OP::U3, m, 0, 0, 0, 0, theta, phi, lambda); }
    static Gate* U2(ValType phi, ValType lambda, IdxType m) { return new Gate(OP::U2, m, 0, 0, 0, 0, 0., phi, lambda); }
    static Gate* U1(ValType lambda, IdxType m) { return new Gate(OP::U1, m, 0, 0, 0, 0, 0., 0., lambda); }
    static Gate* CX(IdxType m, IdxType n) { return new Gate(OP::CX, m, n, 0, 0, 0, 0., 0., 0.); }
    static Gate* ID(IdxType m) { return new Gate(OP::ID, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* X(IdxType m) { return new Gate(OP::X, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* Y(IdxType m) { return new Gate(OP::Y, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* Z(IdxType m) { return new Gate(OP::Z, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* H(IdxType m) { return new Gate(OP::H, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* S(IdxType m) { return new Gate(OP::S, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* SDG(IdxType m) { return new Gate(OP::SDG, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* T(IdxType m) { return new Gate(OP::T, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* TDG(IdxType m) { return new Gate(OP::TDG, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* RX(ValType theta, IdxType m) { return new Gate(OP::RX, m, 0, 0, 0, 0, theta, 0., 0.); }
    static Gate* RY(ValType theta, IdxType m) { return new Gate(OP::RY, m, 0, 0, 0, 0, theta, 0., 0.); }
    static Gate* RZ(ValType phi, IdxType m) { return new Gate(OP::RZ, m, 0, 0, 0, 0, 0., phi, 0.); }
    static Gate* CZ(IdxType m, IdxType n) { return new Gate(OP::CZ, m, n, 0, 0, 0, 0., 0., 0.); }
    static Gate* CY(IdxType m, IdxType n) { return new Gate(OP::CY, m, n, 0, 0, 0, 0., 0., 0.); }
    static Gate* SWAP(IdxType m, IdxType n) { return new Gate(OP::SWAP, m, n, 0, 0, 0, 0., 0., 0.); }
    static Gate* CH(IdxType m, IdxType n) { return new Gate(OP::CH, m, n, 0, 0, 0, 0., 0., 0.); }
    static Gate* CCX(IdxType l, IdxType m, IdxType n) { return new Gate(OP::CCX, l, m, n, 0, 0, 0., 0., 0.); }
    static Gate* CSWAP(IdxType l, IdxType m, IdxType n) { return new Gate(OP::CSWAP, l, m, n, 0, 0, 0., 0., 0.); }
    static Gate* CRX(ValType lambda, IdxType m, IdxType n) { return new Gate(OP::CRX, m, n, 0, 0, 0, 0., 0., lambda); }
    static Gate* CRY(ValType lambda, IdxType m, IdxType n) { return new Gate(OP::CRY, m, n, 0, 0, 0, 0., 0., lambda); }
    static Gate* CRZ(ValType lambda, IdxType m, IdxType n) { return new Gate(OP::CRZ, m, n, 0, 0, 0, 0., 0., lambda); }
    static Gate* CU1(ValType lambda, IdxType m, IdxType n) { return new Gate(OP::CU1, m, n, 0, 0, 0, 0., 0., lambda); }
    static Gate* CU3(ValType theta, ValType phi, ValType lambda, IdxType m, IdxType n) { return new Gate(OP::CU3, m, n, 0, 0, 0, theta, phi, lambda); }
    static Gate* RXX(ValType theta, IdxType m, IdxType n) { return new Gate(OP::RXX, m, n, 0, 0, 0, theta, 0., 0.); }
    static Gate* RZZ(ValType theta, IdxType m, IdxType n) { return new Gate(OP::RZZ, m, n, 0, 0, 0, theta, 0., 0.); }
    static Gate* RCCX(IdxType l, IdxType m, IdxType n) { return new Gate(OP::RCCX, l, m, n, 0, 0, 0., 0., 0.); }
    static Gate* RC3X(IdxType l, IdxType m, IdxType n, IdxType o) { return new Gate(OP::RC3X, l, m, n, o, 0, 0., 0., 0.); }
    static Gate* C3X(IdxType l, IdxType m, IdxType n, IdxType o) { return new Gate(OP::C3X, l, m, n, o, 0, 0., 0., 0.); }
    static Gate* C3SQRTX(IdxType l, IdxType m, IdxType n, IdxType o) { return new Gate(OP::C3SQRTX, l, m, n, o, 0, 0., 0., 0.); }
    static Gate* C4X(IdxType l, IdxType m, IdxType n, IdxType o, IdxType p) { return new Gate(OP::C4X, l, m, n, o, p, 0., 0., 0.); }
    static Gate* R(ValType theta, IdxType m) { return new Gate(OP::R, m, 0, 0, 0, 0, theta, 0., 0.); }
    static Gate* SRN(IdxType m) { return new Gate(OP::SRN, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* W(IdxType m) { return new Gate(OP::W, m, 0, 0, 0, 0, 0., 0., 0.); }
    static Gate* RYY(ValType theta, IdxType m, IdxType n) { return new Gate(OP::RYY, m, n, 0, 0, 0, theta, 0., 0.); }

    const IdxType n_qubits;
    const IdxType i_gpu;
    const IdxType dim;
    IdxType n_gates;
    double last_ms = 0.0;
    std::vector<Gate> circuit;

private:
    svmi_handle h = nullptr;
};

// svsim/src/util_cpu.h:70-85: a basis-state index as n_qubits binary digits
// (qubit n_qubits-1 first), and one line per measured state.
inline void print_binary(IdxType v, int width)
{
    for (int i = width - 1; i >= 0; i--) putchar('0' + int((v >> i) & 1));
}
inline void print_measurement(const IdxType* res_state, IdxType n_qubits, int repetition)
{
    if (!res_state) throw std::runtime_error("print_measurement: null results");
    printf("\n===============  Measurement (tests=%d) ================\n", repetition);
    for (int i = 0; i < repetition; i++) {
        printf("Test-%d: ", i);
        print_binary(res_state[i], int(n_qubits));
        printf("\n");
    }
}
inline void print_measurement(const std::vector<IdxType>& res_state, IdxType n_qubits, int repetition)
{
    if (res_state.size() < size_t(repetition)) throw std::runtime_error("print_measurement: too few results");
    print_measurement(res_state.data(), n_qubits, repetition);
}

}  // namespace SVSim
