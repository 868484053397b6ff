// svmi_qasm.hpp -- OpenQASM 2.0 front end: source text -> reference gate
// records (svmi_gate) for svmi_upload.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/svmi.h"

namespace svmi {

struct QasmProgram {
    int n_qubits = 0;
    std::vector<svmi_gate> gates;
    uint64_t basic_gates = 0;   // gate count by the translator's table (svsim_qasm.py:30-88)
    uint64_t cx_gates = 0;      // CX count by the translator's table
    uint64_t skipped = 0;       // measure / reset / barrier / if statements (not simulated)
};

// Parse OpenQASM 2.0 text.  Returns SVMI_OK, or an error with `err` set to
// "line L: message".
svmi_status parse_qasm(const char* text, size_t len, QasmProgram& out, std::string& err);

}  // namespace svmi
