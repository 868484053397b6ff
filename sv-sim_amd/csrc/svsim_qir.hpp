// svsim_qir.hpp -- QIR runtime drivers on the MI355X engine: the qubit
// bookkeeping, gate translation and measurement of the reference's two
// SVSimSimulator classes -- the state-vector one
// (svsim_qsharp/svsim_qsharp_wrapper.cpp:56-555) over the C ABI's
// svmi_qir_run, and the density-matrix one
// (svsim_qsharp_noise/svsim_qsharp_wrapper.cpp:56-425) over svmi_dm_qir_run
// (include/svmi.h).
//
// The reference classes implement the Microsoft QIR runtime interfaces
// IRuntimeDriver and IQuantumGateSet, whose headers are not part of the
// reference tree; these classes keep their method names, argument order and
// meaning with plain types (Qubit / Result are opaque handles, PauliId and
// ResultValue use the QIR runtime's numbering), so a QIR runtime binding is a
// thin subclass forwarding each virtual to the method of the same name.
//
// Gate set, shared by both (class Driver; the two wrappers' gate methods are
// the same code over different qubit lookups):
//   * gates are queued as QIR engine records with qubit indices and control
//     masks (sv wrapper:265-519); Exp/ControlledExp on several targets rotate
//     X/Y to Z (H; H S H), accumulate parity with a CX chain into
//     targets[0], apply RZ(-2 theta) (controlled by the controls) and undo
//     (300-372, 423-497);
//   * Measure(1 basis, 1 target) draws r = rand()/RAND_MAX, queues a Measure
//     record, runs the queue and reports One iff r < P(1) (520-555).
// State-vector driver (SVSimSimulator):
//   * a qubit lives in a "physical" slot (a bit of the amplitude index);
//     AllocateQubit takes the first free logical id and the next physical
//     slot (wrapper:109-129); ReleaseQubit first runs the queued gates, swaps
//     the released qubit into the last physical slot, runs that, and shrinks
//     the register (131-241); releasing the last qubit resets the state.
// Density-matrix driver (noise::SVSimSimulator):
//   * a qubit's handle is its index, n - 1 for the n-th live qubit; the
//     capacity check is 2 n <= slots (noise wrapper:93-100);
//   * ReleaseQubit runs nothing and only lowers the live count; the engine's
//     qubit count (the n of rho's index row | col << n) only grows, so a
//     release followed by an allocation runs later gates on a larger rho
//     over the same amplitudes, as the reference does (its engine's
//     ReleaseQubit is empty, noise svsim_cpu_omp.hpp:257-261); releasing the
//     last live qubit drops the queue and resets to |0..0><0..0| with no
//     qubits (noise wrapper:101-106, svsim_cpu_omp.hpp:436-457);
//   * a gate on a qubit at or above the engine's count is rejected
//     ("Gate uses qubit out of range!", svsim_cpu_omp.hpp:163-171).
// Differences: errors are exceptions (std::logic_error) where the reference
// asserts; AllocateQubit leaves the register unchanged when no slot is free;
// the draw source can be replaced (set_draw) for reproducible runs; the
// density-matrix driver also rejects control qubits at or above the engine's
// count, and runs rho of slots / 2 qubits (the reference asserts 2 n < slots
// at its run, svsim_cpu_omp.hpp:468).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <ctime>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "svmi.h"

namespace SVSim {
namespace qir {

// QIR runtime enums (numbering of the Microsoft QIR runtime).
enum PauliId : int32_t { PauliId_I = 0, PauliId_X = 1, PauliId_Z = 2, PauliId_Y = 3 };
enum ResultValue : int32_t { Result_Zero = 0, Result_One = 1, Result_Pending = 2 };

struct QubitTag;
struct ResultTag;
using Qubit = QubitTag*;
using Result = ResultTag*;

// P(1) and outcome per Measure record of one run, and the P(1) of the latest
// measurement so far (what the reference's Simulation::sim returns).
using RunResult = std::tuple<std::vector<double>, std::vector<int32_t>, double>;

// The state-vector engine a driver feeds.
class Backend {
public:
    virtual ~Backend() = default;
    // Apply the records to the first 2^n amplitudes.
    virtual RunResult run(int n, const std::vector<svmi_qir_gate>& gates) = 0;
    // |0..0>, latest P(1) = 0 (Simulation::reset).
    virtual void reset() = 0;
    // Number of physical slots (qubits of the state).
    virtual int capacity() const = 0;
};

// The GPU engine through the C ABI.
class DeviceBackend final : public Backend {
public:
    DeviceBackend(int capacity, int device)
    {
        check(svmi_create(capacity, device, &h_), "svmi_create");
    }
    ~DeviceBackend() override { svmi_destroy(h_); }
    DeviceBackend(const DeviceBackend&) = delete;
    DeviceBackend& operator=(const DeviceBackend&) = delete;

    RunResult run(int n, const std::vector<svmi_qir_gate>& gates) override
    {
        size_t nm = 0;
        for (const svmi_qir_gate& g : gates) nm += g.op == SVMI_QIR_MEASURE;
        std::vector<double> probs(nm);
        std::vector<int32_t> outs(nm);
        double last = 0.0;
        check(svmi_qir_run(h_, n, gates.data(), gates.size(), probs.data(), outs.data(), &last), "svmi_qir_run");
        return RunResult{probs, outs, last};
    }
    void reset() override { check(svmi_reset_state(h_), "svmi_reset_state"); }
    int capacity() const override { return svmi_num_qubits(h_); }
    svmi_handle handle() const { return h_; }

private:
    static void check(svmi_status st, const char* what)
    {
        if (st != SVMI_OK) throw std::runtime_error(std::string(what) + ": " + svmi_last_error());
    }
    svmi_handle h_ = nullptr;
};

// The density-matrix engine through the C ABI: run(n, ...) applies the
// records to rho of n qubits, the first 2^(2n) amplitudes of a handle of
// `capacity` qubits; the third result is the P(1) of the run's last Measure
// record (the previous value if it has none).
class DmDeviceBackend final : public Backend {
public:
    DmDeviceBackend(int capacity, int device)
    {
        check(svmi_create(capacity, device, &h_), "svmi_create");
    }
    ~DmDeviceBackend() override { svmi_destroy(h_); }
    DmDeviceBackend(const DmDeviceBackend&) = delete;
    DmDeviceBackend& operator=(const DmDeviceBackend&) = delete;

    RunResult run(int n, const std::vector<svmi_qir_gate>& gates) override
    {
        size_t nm = 0;
        for (const svmi_qir_gate& g : gates) nm += g.op == SVMI_QIR_MEASURE;
        std::vector<double> probs(nm);
        std::vector<int32_t> outs(nm);
        check(svmi_dm_qir_run(h_, n, gates.data(), gates.size(), probs.data(), outs.data()), "svmi_dm_qir_run");
        if (nm) last_ = probs.back();
        return RunResult{probs, outs, last_};
    }
    void reset() override
    {
        check(svmi_reset_state(h_), "svmi_reset_state");
        last_ = 0.0;
    }
    int capacity() const override { return svmi_num_qubits(h_); }
    svmi_handle handle() const { return h_; }

private:
    static void check(svmi_status st, const char* what)
    {
        if (st != SVMI_OK) throw std::runtime_error(std::string(what) + ": " + svmi_last_error());
    }
    svmi_handle h_ = nullptr;
    double last_ = 0.0;
};

// The gate set and measurement both drivers share; a driver supplies the
// qubit lookup (slot) and the run of the queued records (sim).
class Driver {
public:
    explicit Driver(std::unique_ptr<Backend> backend) : sim_(std::move(backend))
    {
        std::srand(unsigned(std::time(nullptr)));   // wrapper:94
        draw_ = [] { return double(std::rand()) / double(RAND_MAX); };
    }
    virtual ~Driver() = default;
    Driver(const Driver&) = delete;
    Driver& operator=(const Driver&) = delete;

    // Replace the uniform draw of Measure (default rand()/RAND_MAX).
    void set_draw(std::function<double()> f) { draw_ = std::move(f); }

    // ---------------- IRuntimeDriver (results) ----------------
    std::string QubitToString(Qubit) { throw std::logic_error("QubitToString not_implemented"); }
    void ReleaseResult(Result) {}
    bool AreEqualResults(Result r1, Result r2) { return r1 == r2; }
    ResultValue GetResultValue(Result r) { return r == UseOne() ? Result_One : Result_Zero; }
    Result UseZero() { return reinterpret_cast<Result>(&result_tags_[0]); }
    Result UseOne() { return reinterpret_cast<Result>(&result_tags_[1]); }

    // ---------------- IQuantumGateSet ----------------
    void X(Qubit t) { push(SVMI_QIR_X, slot(t)); }
    void Y(Qubit t) { push(SVMI_QIR_Y, slot(t)); }
    void Z(Qubit t) { push(SVMI_QIR_Z, slot(t)); }
    void H(Qubit t) { push(SVMI_QIR_H, slot(t)); }
    void S(Qubit t) { push(SVMI_QIR_S, slot(t)); }
    void T(Qubit t) { push(SVMI_QIR_T, slot(t)); }
    void AdjointS(Qubit t) { push(SVMI_QIR_ADJOINT_S, slot(t)); }
    void AdjointT(Qubit t) { push(SVMI_QIR_ADJOINT_T, slot(t)); }

    void R(PauliId axis, Qubit t, double theta) { push(rotation(axis, SVMI_QIR_RI), slot(t), 0, theta); }

    void Exp(long numTargets, PauliId paulis[], Qubit targets[], double theta)
    {
        exp_impl(0, false, numTargets, paulis, targets, theta);
    }

    void ControlledX(long nc, Qubit controls[], Qubit t) { push(SVMI_QIR_CONTROLLED_X, slot(t), mask(nc, controls)); }
    void ControlledY(long nc, Qubit controls[], Qubit t) { push(SVMI_QIR_CONTROLLED_Y, slot(t), mask(nc, controls)); }
    void ControlledZ(long nc, Qubit controls[], Qubit t) { push(SVMI_QIR_CONTROLLED_Z, slot(t), mask(nc, controls)); }
    void ControlledH(long nc, Qubit controls[], Qubit t) { push(SVMI_QIR_CONTROLLED_H, slot(t), mask(nc, controls)); }
    void ControlledS(long nc, Qubit controls[], Qubit t) { push(SVMI_QIR_CONTROLLED_S, slot(t), mask(nc, controls)); }
    void ControlledT(long nc, Qubit controls[], Qubit t) { push(SVMI_QIR_CONTROLLED_T, slot(t), mask(nc, controls)); }
    void ControlledAdjointS(long nc, Qubit controls[], Qubit t)
    {
        push(SVMI_QIR_CONTROLLED_ADJOINT_S, slot(t), mask(nc, controls));
    }
    void ControlledAdjointT(long nc, Qubit controls[], Qubit t)
    {
        push(SVMI_QIR_CONTROLLED_ADJOINT_T, slot(t), mask(nc, controls));
    }

    void ControlledR(long nc, Qubit controls[], PauliId axis, Qubit t, double theta)
    {
        push(rotation(axis, SVMI_QIR_CONTROLLED_RI), slot(t), mask(nc, controls), theta);
    }

    void ControlledExp(long nc, Qubit controls[], long numTargets, PauliId paulis[], Qubit targets[], double theta)
    {
        exp_impl(mask(nc, controls), true, numTargets, paulis, targets, theta);
    }

    Result Measure(long numBases, PauliId bases[], long numTargets, Qubit targets[])
    {
        if (numBases != 1 || numTargets != 1)
            throw std::logic_error("Measure supports one basis and one target");
        const uint64_t target = slot(targets[0]);
        const double r = draw_();
        uint64_t pauli = 0;   // I and Z: computational basis
        if (bases[0] == PauliId_X) pauli = 1;
        if (bases[0] == PauliId_Y) pauli = 2;
        push(SVMI_QIR_MEASURE, target, pauli, r);
        const RunResult res = sim();
        const std::vector<int32_t>& outs = std::get<1>(res);
        return outs.back() ? UseOne() : UseZero();
    }

    // ---------------- inspection (no reference counterpart) ----------------
    uint64_t slot_of(Qubit q) const { return slot(q); }
    double last_prob() const { return last_prob_; }
    const std::vector<svmi_qir_gate>& pending() const { return pending_; }
    Backend& backend() { return *sim_; }
    // Run the queued records now (the reference runs them at Measure, and
    // the state-vector driver also at ReleaseQubit).
    void flush()
    {
        if (!pending_.empty()) sim();
    }

protected:
    static constexpr uint64_t kFree = ~uint64_t(0);

    // The amplitude-index bit of a live qubit (throws for others).
    virtual uint64_t slot(Qubit q) const = 0;
    // Run and clear the queue.
    virtual RunResult sim() = 0;

    static Qubit from_id(uint64_t id) { return reinterpret_cast<Qubit>(uintptr_t(id + 1) << 4); }
    static uint64_t to_id(Qubit q)
    {
        if (!q) throw std::logic_error("null qubit");
        return uint64_t(reinterpret_cast<uintptr_t>(q) >> 4) - 1;
    }
    uint64_t mask(long nc, Qubit controls[]) const
    {
        uint64_t m = 0;
        for (long i = 0; i < nc; i++) m |= uint64_t(1) << slot(controls[i]);
        return m;
    }
    static int rotation(PauliId axis, int base)   // base + {0 I, 1 X, 2 Y, 3 Z}
    {
        switch (axis) {
        case PauliId_I: return base;
        case PauliId_X: return base + 1;
        case PauliId_Y: return base + 2;
        case PauliId_Z: return base + 3;
        }
        throw std::logic_error("unknown PauliId");
    }
    void push(int op, uint64_t qubit, uint64_t m = 0, double theta = 0.0)
    {
        pending_.push_back(svmi_qir_gate{int32_t(op), int32_t(qubit), m, theta});
    }

    // wrapper:300-372 (Exp) and 423-497 (ControlledExp)
    void exp_impl(uint64_t cmask, bool controlled, long nt, PauliId paulis[], Qubit targets[], double theta)
    {
        if (nt < 1) throw std::logic_error("Exp needs at least one target");
        if (nt == 1) {
            const int base = controlled ? SVMI_QIR_CONTROLLED_EI : SVMI_QIR_EI;
            push(rotation(paulis[0], base), slot(targets[0]), cmask, theta);
            return;
        }
        for (long i = 0; i < nt; i++) {
            const uint64_t t = slot(targets[i]);
            if (paulis[i] == PauliId_X) push(SVMI_QIR_H, t);
            if (paulis[i] == PauliId_Y) {
                push(SVMI_QIR_H, t);
                push(SVMI_QIR_S, t);
                push(SVMI_QIR_H, t);
            }
        }
        const uint64_t t0 = slot(targets[0]);
        for (long i = 1; i < nt; i++)
            if (paulis[i] != PauliId_I) push(SVMI_QIR_CONTROLLED_X, t0, uint64_t(1) << slot(targets[i]));
        if (controlled) push(SVMI_QIR_CONTROLLED_RZ, t0, cmask, theta * (-2.0));
        else push(SVMI_QIR_RZ, t0, 0, theta * (-2.0));
        for (long i = nt - 1; i > 0; i--)
            if (paulis[i] != PauliId_I) push(SVMI_QIR_CONTROLLED_X, t0, uint64_t(1) << slot(targets[i]));
        for (long i = nt - 1; i >= 0; i--) {
            const uint64_t t = slot(targets[i]);
            if (paulis[i] == PauliId_X) push(SVMI_QIR_H, t);
            if (paulis[i] == PauliId_Y) {
                push(SVMI_QIR_H, t);
                push(SVMI_QIR_ADJOINT_S, t);
                push(SVMI_QIR_H, t);
            }
        }
    }

    // Run the queue on n qubits.
    RunResult run_pending(int n)
    {
        RunResult r = sim_->run(n, pending_);
        pending_.clear();
        if (!std::get<0>(r).empty()) last_prob_ = std::get<2>(r);
        return r;
    }

    std::unique_ptr<Backend> sim_;
    std::vector<svmi_qir_gate> pending_;
    std::function<double()> draw_;
    double last_prob_ = 0.0;
    char result_tags_[2] = {0, 1};
};

// svsim_qsharp/svsim_qsharp_wrapper.cpp:56-555
class SVSimSimulator final : public Driver {
public:
    // capacity: physical slots (the reference's N_QUBIT_SLOT, config.hpp:38)
    explicit SVSimSimulator(int capacity = 28, int device = 0)
        : SVSimSimulator(std::unique_ptr<Backend>(new DeviceBackend(capacity, device)))
    {
    }
    explicit SVSimSimulator(std::unique_ptr<Backend> backend) : Driver(std::move(backend))
    {
        const size_t slots = size_t(sim_->capacity());
        logical_.assign(slots, kFree);
        physical_.assign(slots, kFree);
    }

    // ---------------- IRuntimeDriver ----------------
    Qubit AllocateQubit()
    {
        uint64_t id = kFree;
        for (size_t i = 0; i < logical_.size(); i++)
            if (logical_[i] == kFree) {
                id = i;
                break;
            }
        if (id == kFree || n_ >= int(physical_.size())) throw std::logic_error("Request qubits more than slots!");
        ++n_;
        logical_[id] = uint64_t(n_ - 1);   // the last physical slot is the next free one
        physical_[n_ - 1] = id;
        return from_id(id);
    }

    void ReleaseQubit(Qubit q)
    {
        if (!pending_.empty()) sim();
        const uint64_t id = to_id(q);
        const uint64_t phys = logical_.at(id);
        if (phys == kFree) throw std::logic_error("ReleaseQubit: qubit not allocated");
        const uint64_t last = uint64_t(n_ - 1);
        if (phys != last) {
            const uint64_t moved = physical_[last];
            push(SVMI_QIR_SWAP, phys, last, 0.0);
            sim();
            logical_[id] = kFree;
            physical_[last] = kFree;
            logical_[moved] = phys;
            physical_[phys] = moved;
        } else {
            logical_[id] = kFree;
            physical_[phys] = kFree;
        }
        --n_;
        if (n_ == 0) {
            sim_->reset();
            last_prob_ = 0.0;
        }
    }

    int num_qubits() const { return n_; }

private:
    uint64_t slot(Qubit q) const override
    {
        const uint64_t s = logical_.at(to_id(q));
        if (s == kFree) throw std::logic_error("qubit not allocated");
        return s;
    }
    RunResult sim() override { return run_pending(n_); }

    std::vector<uint64_t> logical_, physical_;   // logical id -> slot, slot -> logical id
    int n_ = 0;
};

namespace noise {

// svsim_qsharp_noise/svsim_qsharp_wrapper.cpp:56-425 on the density-matrix
// engine: rho of n qubits in a state of 2n qubits.
class SVSimSimulator final : public Driver {
public:
    // capacity: qubits of the state (the reference's N_QUBIT_SLOT,
    // svsim_qsharp_noise/config.hpp:38); rho has at most capacity / 2.
    explicit SVSimSimulator(int capacity = 28, int device = 0)
        : SVSimSimulator(std::unique_ptr<Backend>(new DmDeviceBackend(capacity, device)))
    {
    }
    explicit SVSimSimulator(std::unique_ptr<Backend> backend) : Driver(std::move(backend)) {}

    // ---------------- IRuntimeDriver ----------------
    // noise wrapper:93-100
    Qubit AllocateQubit()
    {
        if (2 * (n_ + 1) > sim_->capacity()) throw std::logic_error("Request qubits more than slots!");
        ++n_;
        ++engine_n_;
        return from_id(uint64_t(n_ - 1));
    }

    // noise wrapper:101-106: the queue stays; the last release resets.
    void ReleaseQubit(Qubit)
    {
        if (n_ == 0) throw std::logic_error("ReleaseQubit: no qubit allocated");
        --n_;
        if (n_ == 0) {
            pending_.clear();
            engine_n_ = 0;
            sim_->reset();
            last_prob_ = 0.0;
        }
    }

    // Live qubits, and the qubits of rho the engine runs on.
    int num_qubits() const { return n_; }
    int engine_qubits() const { return engine_n_; }

private:
    uint64_t slot(Qubit q) const override
    {
        const uint64_t id = to_id(q);
        if (id >= uint64_t(engine_n_)) throw std::logic_error("Gate uses qubit out of range!");
        return id;
    }
    RunResult sim() override
    {
        if (2 * engine_n_ > sim_->capacity())
            throw std::logic_error("density matrix of " + std::to_string(engine_n_) + " qubits exceeds the " +
                                   std::to_string(sim_->capacity()) + "-qubit state");
        return run_pending(engine_n_);
    }

    int n_ = 0;          // live qubits (the wrapper's n_qubits)
    int engine_n_ = 0;   // the engine's qubit count (its circuit's n_qubits)
};

}  // namespace noise

}  // namespace qir
}  // namespace SVSim
