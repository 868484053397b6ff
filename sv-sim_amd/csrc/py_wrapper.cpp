// py_wrapper.cpp -- pybind11 module `libsvsim` with the reference's Python
// surface (svsim/src/py_wrapper.cpp:53-109): Gate, Simulation(n_qubits,
// n_dev), append/upload/clear_circuit/run/reset/measure and the 38 static
// gate factories.  Additions (no reference counterpart): the OP enum,
// get_sv/set_sv, measure_with(draws), expval, adjoint, last_sim_ms.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "../../include/svmi.h"
#include "svsim_amd.hpp"

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
        .def("clear_circuit", &Simulation::clear_circuit)
        .def("run", &Simulation::sim, py::call_guard<py::gil_scoped_release>())
        .def("reset", &Simulation::reset)
        .def("measure", [](Simulation& s, unsigned repetition) {
            std::vector<IdxType> r = s.measure(repetition);
            py::list out;
            for (IdxType v : r) out.append(v);
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
}
