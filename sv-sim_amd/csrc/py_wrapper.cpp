// py_wrapper.cpp -- pybind11 module `libsvsim` with the reference's Python
// surface (svsim/src/py_wrapper.cpp:53-109): Gate, Simulation(n_qubits,
// i_gpu) -- the second argument is the HIP ordinal, as in the AMD single-GPU
// build of that module (svsim_amdgpu_sin.hpp:207-220); an ordinal that is not
// present raises --, append/upload/clear_circuit/run/reset/measure and the 38 static
// gate factories.  Additions (no reference counterpart): the OP enum,
// get_sv/set_sv, measure_with(draws), expval, adjoint, last_sim_ms.
//
// QirSimulator / QirNoiseSimulator: the QIR runtime drivers (svsim_qir.hpp,
// the reference's SVSimSimulator in svsim_qsharp/svsim_qsharp_wrapper.cpp
// and in svsim_qsharp_noise/svsim_qsharp_wrapper.cpp) with the same method
// names; qubits and results are integer handles, qubit and Pauli arrays are
// Python sequences.  A QirBackend subclass written in Python can stand in for
// the GPU engine.
#include <memory>

#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "../../include/svmi.h"
#include "svsim_amd.hpp"
#include "svsim_qir.hpp"

namespace py = pybind11;
using namespace SVSim;

namespace {

std::vector<svmi_pauli_term> to_terms(const py::iterable& terms)
{
    std::vector<svmi_pauli_term> out;
    for (auto t : terms) {
        auto tup = t.cast<py::tuple>();
        if (tup.size() != 3) throw std::runtime_error("Pauli term must be (coeff, xmask, zmask)");
        out.push_back(svmi_pauli_term{tup[0].cast<double>(), tup[1].cast<uint64_t>(), tup[2].cast<uint64_t>()});
    }
    return out;
}

namespace qir = SVSim::qir;

// Python subclasses of QirBackend
struct PyQirBackend : qir::Backend {
    qir::RunResult run(int n, const std::vector<svmi_qir_gate>& gates) override
    {
        PYBIND11_OVERRIDE_PURE(qir::RunResult, qir::Backend, run, n, gates);
    }
    void reset() override { PYBIND11_OVERRIDE_PURE(void, qir::Backend, reset, ); }
    int capacity() const override { PYBIND11_OVERRIDE_PURE(int, qir::Backend, capacity, ); }
};

// Owned by the simulator; keeps the Python backend object alive.
struct HeldBackend : qir::Backend {
    py::object keep;
    qir::Backend* b;
    explicit HeldBackend(py::object o) : keep(std::move(o)), b(keep.cast<qir::Backend*>()) {}
    qir::RunResult run(int n, const std::vector<svmi_qir_gate>& gates) override { return b->run(n, gates); }
    void reset() override { b->reset(); }
    int capacity() const override { return b->capacity(); }
};

qir::Qubit to_q(uint64_t v) { return reinterpret_cast<qir::Qubit>(uintptr_t(v)); }
uint64_t from_q(qir::Qubit q) { return uint64_t(reinterpret_cast<uintptr_t>(q)); }
qir::Result to_r(uint64_t v) { return reinterpret_cast<qir::Result>(uintptr_t(v)); }
uint64_t from_r(qir::Result r) { return uint64_t(reinterpret_cast<uintptr_t>(r)); }

std::vector<qir::Qubit> to_qs(const std::vector<uint64_t>& v)
{
    std::vector<qir::Qubit> out;
    for (uint64_t x : v) out.push_back(to_q(x));
    return out;
}

// QirSimulator / QirNoiseSimulator: one driver class's methods.
template <class QS>
py::class_<QS> bind_driver(py::module_& m, const char* name)
{
    auto ctl = [](void (QS::*f)(long, qir::Qubit[], qir::Qubit)) {
        return [f](QS& s, const std::vector<uint64_t>& c, uint64_t t) {
            std::vector<qir::Qubit> cs = to_qs(c);
            (s.*f)(long(cs.size()), cs.data(), to_q(t));
        };
    };
    auto one = [](void (QS::*f)(qir::Qubit)) { return [f](QS& s, uint64_t t) { (s.*f)(to_q(t)); }; };
    py::class_<QS> cls(m, name);
    cls.def(py::init<int, int>(), py::arg("capacity") = 28, py::arg("device") = 0)
        .def(py::init([](py::object backend) { return new QS(std::unique_ptr<qir::Backend>(new HeldBackend(backend))); }),
             py::arg("backend"))
        .def("set_draw", &QS::set_draw)
        .def("AllocateQubit", [](QS& s) { return from_q(s.AllocateQubit()); })
        .def("ReleaseQubit", [](QS& s, uint64_t q) { s.ReleaseQubit(to_q(q)); })
        .def("QubitToString", [](QS& s, uint64_t q) { return s.QubitToString(to_q(q)); })
        .def("ReleaseResult", [](QS& s, uint64_t r) { s.ReleaseResult(to_r(r)); })
        .def("AreEqualResults", [](QS& s, uint64_t a, uint64_t b) { return s.AreEqualResults(to_r(a), to_r(b)); })
        .def("GetResultValue", [](QS& s, uint64_t r) { return s.GetResultValue(to_r(r)); })
        .def("UseZero", [](QS& s) { return from_r(s.UseZero()); })
        .def("UseOne", [](QS& s) { return from_r(s.UseOne()); })
        .def("X", one(&QS::X))
        .def("Y", one(&QS::Y))
        .def("Z", one(&QS::Z))
        .def("H", one(&QS::H))
        .def("S", one(&QS::S))
        .def("T", one(&QS::T))
        .def("AdjointS", one(&QS::AdjointS))
        .def("AdjointT", one(&QS::AdjointT))
        .def("R", [](QS& s, qir::PauliId axis, uint64_t t, double theta) { s.R(axis, to_q(t), theta); })
        .def("Exp", [](QS& s, std::vector<qir::PauliId> paulis, const std::vector<uint64_t>& targets, double theta) {
            if (paulis.size() != targets.size()) throw std::runtime_error("Exp: paulis and targets differ in length");
            std::vector<qir::Qubit> ts = to_qs(targets);
            s.Exp(long(ts.size()), paulis.data(), ts.data(), theta);
        })
        .def("ControlledX", ctl(&QS::ControlledX))
        .def("ControlledY", ctl(&QS::ControlledY))
        .def("ControlledZ", ctl(&QS::ControlledZ))
        .def("ControlledH", ctl(&QS::ControlledH))
        .def("ControlledS", ctl(&QS::ControlledS))
        .def("ControlledT", ctl(&QS::ControlledT))
        .def("ControlledAdjointS", ctl(&QS::ControlledAdjointS))
        .def("ControlledAdjointT", ctl(&QS::ControlledAdjointT))
        .def("ControlledR", [](QS& s, const std::vector<uint64_t>& c, qir::PauliId axis, uint64_t t, double theta) {
            std::vector<qir::Qubit> cs = to_qs(c);
            s.ControlledR(long(cs.size()), cs.data(), axis, to_q(t), theta);
        })
        .def("ControlledExp", [](QS& s, const std::vector<uint64_t>& c, std::vector<qir::PauliId> paulis,
                                 const std::vector<uint64_t>& targets, double theta) {
            if (paulis.size() != targets.size())
                throw std::runtime_error("ControlledExp: paulis and targets differ in length");
            std::vector<qir::Qubit> cs = to_qs(c), ts = to_qs(targets);
            s.ControlledExp(long(cs.size()), cs.data(), long(ts.size()), paulis.data(), ts.data(), theta);
        })
        .def("Measure", [](QS& s, std::vector<qir::PauliId> bases, const std::vector<uint64_t>& targets) {
            std::vector<qir::Qubit> ts = to_qs(targets);
            return from_r(s.Measure(long(bases.size()), bases.data(), long(ts.size()), ts.data()));
        })
        .def("num_qubits", &QS::num_qubits)
        .def("slot_of", [](const QS& s, uint64_t q) { return s.slot_of(to_q(q)); })
        .def("last_prob", &QS::last_prob)
        .def("pending", &QS::pending)
        .def("flush", &QS::flush)
        .def("get_sv", [](QS& s) {
            svmi_handle h = nullptr;
            if (auto* d = dynamic_cast<qir::DeviceBackend*>(&s.backend())) h = d->handle();
            if (auto* d = dynamic_cast<qir::DmDeviceBackend*>(&s.backend())) h = d->handle();
            if (!h) throw std::runtime_error("get_sv: the simulator does not run on the GPU engine");
            const size_t dim = size_t(1) << s.backend().capacity();
            py::array_t<double> re(dim), im(dim);
            svmi_throw(svmi_get_state(h, re.mutable_data(), im.mutable_data()), "svmi_get_state");
            return py::make_tuple(re, im);
        });
    return cls;
}

}  // namespace

PYBIND11_MODULE(libsvsim, m)
{
    m.doc() = "SV-Sim Python API on the MI355X state-vector engine";
    py::enum_<OP> op(m, "OP");
    for (int i = 0; i <= int(RYY); i++) op.value(op_names()[i], OP(i));

    py::class_<Gate>(m, "Gate")
        .def(py::init<OP, IdxType, IdxType, IdxType, IdxType, IdxType, ValType, ValType, ValType>());

    py::class_<Simulation>(m, "Simulation")
        .def(py::init<IdxType, IdxType>())  // n_qubits, i_gpu
        .def("append", &Simulation::append)
        .def("upload", [](py::object self) {
            self.cast<Simulation&>().upload();
            return self;
        })
        .def("append_qasm", &Simulation::append_qasm)
        .def("clear_circuit", &Simulation::clear_circuit)
        .def("run", &Simulation::sim, py::call_guard<py::gil_scoped_release>())
        .def("reset", &Simulation::reset)
        .def("measure", [](Simulation& s, unsigned repetition) {
            std::unique_ptr<IdxType[]> r(s.measure(repetition));
            py::list out;
            for (unsigned i = 0; i < repetition; i++) out.append(r[i]);
            return out;
        })
        .def("measure_with", [](Simulation& s, py::array_t<double, py::array::c_style | py::array::forcecast> u) {
            std::vector<double> v(u.data(), u.data() + u.size());
            std::vector<IdxType> r = s.measure_with(v);
            py::list out;
            for (IdxType x : r) out.append(x);
            return out;
        })
        .def("get_sv", [](const Simulation& s) {
            py::array_t<double> re(size_t(s.dim)), im(size_t(s.dim));
            s.get_sv(re.mutable_data(), im.mutable_data());
            return py::make_tuple(re, im);
        })
        .def("set_sv", [](Simulation& s, py::array_t<double, py::array::c_style | py::array::forcecast> re,
                          py::array_t<double, py::array::c_style | py::array::forcecast> im) {
            if (IdxType(re.size()) != s.dim || IdxType(im.size()) != s.dim)
                throw std::runtime_error("set_sv: arrays must have 2^n entries");
            s.set_sv(re.data(), im.data());
        })
        .def("expval", [](const Simulation& s, py::iterable terms) { return s.expval(to_terms(terms)); })
        .def("adjoint", [](Simulation& s, py::iterable terms, py::array_t<int32_t, py::array::c_style | py::array::forcecast> slots, size_t n_params) {
            std::vector<int32_t> ps(slots.data(), slots.data() + slots.size());
            std::vector<double> grad;
            const double e = s.adjoint(to_terms(terms), ps, n_params, grad);
            return py::make_tuple(e, py::array_t<double>(grad.size(), grad.data()));
        })
        .def("dump", &Simulation::dump)
        .def_readonly("n_qubits", &Simulation::n_qubits)
        .def_readonly("n_gates", &Simulation::n_gates)
        .def_readonly("last_sim_ms", &Simulation::last_ms)
        .def_static("U3", &Simulation::U3)
        .def_static("U2", &Simulation::U2)
        .def_static("U1", &Simulation::U1)
        .def_static("CX", &Simulation::CX)
        .def_static("ID", &Simulation::ID)
        .def_static("X", &Simulation::X)
        .def_static("Y", &Simulation::Y)
        .def_static("Z", &Simulation::Z)
        .def_static("H", &Simulation::H)
        .def_static("S", &Simulation::S)
        .def_static("SDG", &Simulation::SDG)
        .def_static("T", &Simulation::T)
        .def_static("TDG", &Simulation::TDG)
        .def_static("RX", &Simulation::RX)
        .def_static("RY", &Simulation::RY)
        .def_static("RZ", &Simulation::RZ)
        .def_static("CZ", &Simulation::CZ)
        .def_static("CY", &Simulation::CY)
        .def_static("SWAP", &Simulation::SWAP)
        .def_static("CH", &Simulation::CH)
        .def_static("CCX", &Simulation::CCX)
        .def_static("CSWAP", &Simulation::CSWAP)
        .def_static("CRX", &Simulation::CRX)
        .def_static("CRY", &Simulation::CRY)
        .def_static("CRZ", &Simulation::CRZ)
        .def_static("CU1", &Simulation::CU1)
        .def_static("CU3", &Simulation::CU3)
        .def_static("RXX", &Simulation::RXX)
        .def_static("RZZ", &Simulation::RZZ)
        .def_static("RCCX", &Simulation::RCCX)
        .def_static("RC3X", &Simulation::RC3X)
        .def_static("C3X", &Simulation::C3X)
        .def_static("C3SQRTX", &Simulation::C3SQRTX)
        .def_static("C4X", &Simulation::C4X)
        .def_static("R", &Simulation::R)
        .def_static("SRN", &Simulation::SRN)
        .def_static("W", &Simulation::W)
        .def_static("RYY", &Simulation::RYY);

    // ---------------- QIR runtime driver ----------------
    py::enum_<qir::PauliId>(m, "PauliId")
        .value("PauliId_I", qir::PauliId_I)
        .value("PauliId_X", qir::PauliId_X)
        .value("PauliId_Z", qir::PauliId_Z)
        .value("PauliId_Y", qir::PauliId_Y);
    py::enum_<qir::ResultValue>(m, "ResultValue")
        .value("Result_Zero", qir::Result_Zero)
        .value("Result_One", qir::Result_One)
        .value("Result_Pending", qir::Result_Pending);

    py::class_<svmi_qir_gate>(m, "QirGate")
        .def(py::init([](int32_t op, int32_t qubit, uint64_t mask, double theta) {
                 return svmi_qir_gate{op, qubit, mask, theta};
             }),
             py::arg("op"), py::arg("qubit"), py::arg("mask") = 0, py::arg("theta") = 0.0)
        .def_readwrite("op", &svmi_qir_gate::op)
        .def_readwrite("qubit", &svmi_qir_gate::qubit)
        .def_readwrite("mask", &svmi_qir_gate::mask)
        .def_readwrite("theta", &svmi_qir_gate::theta)
        .def("__repr__", [](const svmi_qir_gate& g) {
            return "QirGate(" + std::to_string(g.op) + ", " + std::to_string(g.qubit) + ", " +
                   std::to_string(g.mask) + ", " + std::to_string(g.theta) + ")";
        });

    py::class_<qir::Backend, PyQirBackend>(m, "QirBackend")
        .def(py::init<>())
        .def("run", &qir::Backend::run)
        .def("reset", &qir::Backend::reset)
        .def("capacity", &qir::Backend::capacity);

    bind_driver<qir::SVSimSimulator>(m, "QirSimulator");
    bind_driver<qir::noise::SVSimSimulator>(m, "QirNoiseSimulator")
        .def("engine_qubits", &qir::noise::SVSimSimulator::engine_qubits);
}
