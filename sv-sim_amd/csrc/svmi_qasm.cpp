// svmi_qasm.cpp -- OpenQASM 2.0 front end (host only).
//
// Replaces the reference's translator svsim/QASMBench/svsim_qasm.py (and its
// twins under svsim/tool/), which turns a .qasm file into a Python script of
// `sim.append(sim.OP(angles..., qubits...))` calls.  Here the source is
// parsed directly into the same gate records (svmi_gate, reference factory
// slot layout) for svmi_upload:
//   * quantum registers are laid out in declaration order on one qubit index
//     space (svsim_qasm.py:298-305);
//   * the built-in gates are qelib1's, mapped to the reference factories with
//     the same names (svsim_qasm.py:30-74; u3/u2/u1 angles to theta/phi/lambda
//     as U3/U2/U1, rz to phi, crx/cry/crz/cu1 to lambda);
//   * a gate applied to whole registers repeats over their bits
//     (svsim_qasm.py:139-171);
//   * measure, reset, barrier, creg, include and `if` statements produce no
//     gates (svsim_qasm.py:98, 313-315; an `if`-guarded gate is dropped);
//   * basic_gates / cx_gates count every gate by the translator's tables
//     (svsim_qasm.py:30-88), user gates by their bodies.
// Unlike the line-based translator this is a tokenising parser: statements
// may span lines or share one, user gates may take angle parameters and call
// parameterised gates in their bodies (the translator drops those angles),
// expressions accept pi, + - * / ^ (and **), unary minus and sin, cos, tan,
// exp, ln, sqrt; unknown gates, wrong arities, repeated or out-of-range
// qubits are errors instead of warnings.  The OpenQASM primitives U and CX
// are accepted as u3 and cx.
#include "svmi_qasm.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <map>
#include <stdexcept>

namespace svmi {

namespace {

struct Tok {
    enum Kind { ID, NUM, STR, SYM, END } kind;
    std::string text;
    int line;
};

struct QasmError : std::runtime_error {
    int line;
    QasmError(int l, const std::string& m) : std::runtime_error(m), line(l) {}
};

std::vector<Tok> tokenize(const char* s, size_t n)
{
    std::vector<Tok> out;
    int line = 1;
    size_t i = 0;
    while (i < n) {
        const char c = s[i];
        if (c == '\n') {
            line++;
            i++;
            continue;
        }
        if (std::isspace(static_cast<unsigned char>(c))) {
            i++;
            continue;
        }
        if (c == '/' && i + 1 < n && s[i + 1] == '/') {
            while (i < n && s[i] != '\n') i++;
            continue;
        }
        if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
            size_t j = i + 1;
            while (j < n && (std::isalnum(static_cast<unsigned char>(s[j])) || s[j] == '_')) j++;
            out.push_back({Tok::ID, std::string(s + i, j - i), line});
            i = j;
            continue;
        }
        if (std::isdigit(static_cast<unsigned char>(c)) ||
            (c == '.' && i + 1 < n && std::isdigit(static_cast<unsigned char>(s[i + 1])))) {
            size_t j = i;
            while (j < n && std::isdigit(static_cast<unsigned char>(s[j]))) j++;
            if (j < n && s[j] == '.') {
                j++;
                while (j < n && std::isdigit(static_cast<unsigned char>(s[j]))) j++;
            }
            if (j < n && (s[j] == 'e' || s[j] == 'E')) {
                size_t k = j + 1;
                if (k < n && (s[k] == '+' || s[k] == '-')) k++;
                if (k < n && std::isdigit(static_cast<unsigned char>(s[k]))) {
                    while (k < n && std::isdigit(static_cast<unsigned char>(s[k]))) k++;
                    j = k;
                }
            }
            out.push_back({Tok::NUM, std::string(s + i, j - i), line});
            i = j;
            continue;
        }
        if (c == '"') {
            size_t j = i + 1;
            while (j < n && s[j] != '"' && s[j] != '\n') j++;
            if (j >= n || s[j] != '"') throw QasmError(line, "unterminated string");
            out.push_back({Tok::STR, std::string(s + i + 1, j - i - 1), line});
            i = j + 1;
            continue;
        }
        if (i + 1 < n) {
            const std::string two(s + i, 2);
            if (two == "->" || two == "==" || two == "**") {
                out.push_back({Tok::SYM, two, line});
                i += 2;
                continue;
            }
        }
        if (std::string(";,()[]{}+-*/^").find(c) != std::string::npos) {
            out.push_back({Tok::SYM, std::string(1, c), line});
            i++;
            continue;
        }
        throw QasmError(line, std::string("unexpected character '") + c + "'");
    }
    out.push_back({Tok::END, "", line});
    return out;
}

// Built-in gates: qelib1 names -> reference factory (svsim_qasm.py:30-88).
struct Builtin {
    int op;          // svmi_op
    int n_params;
    int slot[3];     // gate slot (0 theta, 1 phi, 2 lambda) of each parameter
    int n_qubits;
    int basic;       // translator's count of standard gates inside
    int cx;          // translator's count of CX inside
};

const std::map<std::string, Builtin>& builtins()
{
    static const std::map<std::string, Builtin> t = {
        {"u3", {SVMI_U3, 3, {0, 1, 2}, 1, 1, 0}},   {"u2", {SVMI_U2, 2, {1, 2, 0}, 1, 1, 0}},
        {"u1", {SVMI_U1, 1, {2, 0, 0}, 1, 1, 0}},   {"cx", {SVMI_CX, 0, {0, 0, 0}, 2, 1, 1}},
        {"id", {SVMI_ID, 0, {0, 0, 0}, 1, 1, 0}},   {"x", {SVMI_X, 0, {0, 0, 0}, 1, 1, 0}},
        {"y", {SVMI_Y, 0, {0, 0, 0}, 1, 1, 0}},     {"z", {SVMI_Z, 0, {0, 0, 0}, 1, 1, 0}},
        {"h", {SVMI_H, 0, {0, 0, 0}, 1, 1, 0}},     {"s", {SVMI_S, 0, {0, 0, 0}, 1, 1, 0}},
        {"sdg", {SVMI_SDG, 0, {0, 0, 0}, 1, 1, 0}}, {"t", {SVMI_T, 0, {0, 0, 0}, 1, 1, 0}},
        {"tdg", {SVMI_TDG, 0, {0, 0, 0}, 1, 1, 0}}, {"rx", {SVMI_RX, 1, {0, 0, 0}, 1, 1, 0}},
        {"ry", {SVMI_RY, 1, {0, 0, 0}, 1, 1, 0}},   {"rz", {SVMI_RZ, 1, {1, 0, 0}, 1, 1, 0}},
        {"cz", {SVMI_CZ, 0, {0, 0, 0}, 2, 3, 1}},   {"cy", {SVMI_CY, 0, {0, 0, 0}, 2, 3, 1}},
        {"swap", {SVMI_SWAP, 0, {0, 0, 0}, 2, 3, 3}}, {"ch", {SVMI_CH, 0, {0, 0, 0}, 2, 11, 2}},
        {"ccx", {SVMI_CCX, 0, {0, 0, 0}, 3, 15, 6}}, {"cswap", {SVMI_CSWAP, 0, {0, 0, 0}, 3, 17, 8}},
        {"crx", {SVMI_CRX, 1, {2, 0, 0}, 2, 5, 2}}, {"cry", {SVMI_CRY, 1, {2, 0, 0}, 2, 4, 2}},
        {"crz", {SVMI_CRZ, 1, {2, 0, 0}, 2, 4, 2}}, {"cu1", {SVMI_CU1, 1, {2, 0, 0}, 2, 5, 2}},
        {"cu3", {SVMI_CU3, 3, {0, 1, 2}, 2, 5, 2}}, {"rxx", {SVMI_RXX, 1, {0, 0, 0}, 2, 7, 2}},
        {"rzz", {SVMI_RZZ, 1, {0, 0, 0}, 2, 3, 2}}, {"rccx", {SVMI_RCCX, 0, {0, 0, 0}, 3, 9, 3}},
        {"rc3x", {SVMI_RC3X, 0, {0, 0, 0}, 4, 18, 6}}, {"c3x", {SVMI_C3X, 0, {0, 0, 0}, 4, 27, 6}},
        {"c3sqrtx", {SVMI_C3SQRTX, 0, {0, 0, 0}, 4, 27, 6}}, {"c4x", {SVMI_C4X, 0, {0, 0, 0}, 5, 87, 18}},
        // OpenQASM primitives
        {"U", {SVMI_U3, 3, {0, 1, 2}, 1, 1, 0}},    {"CX", {SVMI_CX, 0, {0, 0, 0}, 2, 1, 1}},
    };
    return t;
}

// A gate call inside a user gate body: parameter expressions as token ranges.
struct BodyCall {
    std::string name;
    std::vector<std::pair<size_t, size_t>> params;   // [begin, end) token ranges
    std::vector<std::string> qargs;
    int line;
};

struct GateDef {
    std::vector<std::string> params, qargs;
    std::vector<BodyCall> body;
    bool opaque = false;
    uint64_t basic = 0, cx = 0;   // counts of one call (filled lazily)
};

class Parser {
public:
    Parser(std::vector<Tok> toks, QasmProgram& out) : t_(std::move(toks)), out_(out) {}

    void run()
    {
        while (peek().kind != Tok::END) statement();
        out_.n_qubits = n_qubits_;
    }

private:
    using Env = std::map<std::string, double>;

    const Tok& peek(size_t k = 0) const { return t_[std::min(p_ + k, t_.size() - 1)]; }
    const Tok& next() { return t_[p_ < t_.size() - 1 ? p_++ : p_]; }
    bool is_sym(const char* s, size_t k = 0) const { return peek(k).kind == Tok::SYM && peek(k).text == s; }
    void expect(const char* s)
    {
        if (!is_sym(s)) throw QasmError(peek().line, std::string("expected '") + s + "' before '" + peek().text + "'");
        next();
    }
    std::string ident()
    {
        if (peek().kind != Tok::ID) throw QasmError(peek().line, "expected an identifier, got '" + peek().text + "'");
        return next().text;
    }
    long integer()
    {
        if (peek().kind != Tok::NUM || peek().text.find_first_of(".eE") != std::string::npos)
            throw QasmError(peek().line, "expected an integer, got '" + peek().text + "'");
        return std::strtol(next().text.c_str(), nullptr, 10);
    }

    // ---------------- expressions ----------------
    double eval_range(size_t b, size_t e, const Env& env)
    {
        const size_t save = p_;
        p_ = b;
        end_ = e;
        const double v = expr(env);
        if (p_ != e) throw QasmError(t_[p_].line, "unexpected '" + t_[p_].text + "' in expression");
        p_ = save;
        end_ = SIZE_MAX;
        return v;
    }
    bool at_end() const { return p_ >= end_; }
    double expr(const Env& env)
    {
        double v = term(env);
        while (!at_end() && (is_sym("+") || is_sym("-"))) {
            const bool plus = next().text == "+";
            const double r = term(env);
            v = plus ? v + r : v - r;
        }
        return v;
    }
    double term(const Env& env)
    {
        double v = unary(env);
        while (!at_end() && (is_sym("*") || is_sym("/"))) {
            const bool mul = next().text == "*";
            const double r = unary(env);
            v = mul ? v * r : v / r;
        }
        return v;
    }
    double unary(const Env& env)
    {
        if (!at_end() && is_sym("-")) {
            next();
            return -unary(env);
        }
        if (!at_end() && is_sym("+")) {
            next();
            return unary(env);
        }
        const double b = primary(env);
        if (!at_end() && (is_sym("^") || is_sym("**"))) {
            next();
            return std::pow(b, unary(env));
        }
        return b;
    }
    double primary(const Env& env)
    {
        if (at_end()) throw QasmError(peek().line, "incomplete expression");
        const Tok& tk = next();
        if (tk.kind == Tok::NUM) return std::strtod(tk.text.c_str(), nullptr);
        if (tk.kind == Tok::SYM && tk.text == "(") {
            const double v = expr(env);
            expect(")");
            return v;
        }
        if (tk.kind == Tok::ID) {
            if (tk.text == "pi") return 3.14159265358979323846;
            static const std::map<std::string, double (*)(double)> fn = {
                {"sin", [](double x) { return std::sin(x); }},  {"cos", [](double x) { return std::cos(x); }},
                {"tan", [](double x) { return std::tan(x); }},  {"exp", [](double x) { return std::exp(x); }},
                {"ln", [](double x) { return std::log(x); }},   {"sqrt", [](double x) { return std::sqrt(x); }}};
            auto f = fn.find(tk.text);
            if (f != fn.end()) {
                expect("(");
                const double v = expr(env);
                expect(")");
                return f->second(v);
            }
            auto e = env.find(tk.text);
            if (e != env.end()) return e->second;
            throw QasmError(tk.line, "unknown name '" + tk.text + "' in expression");
        }
        throw QasmError(tk.line, "unexpected '" + tk.text + "' in expression");
    }
    // token ranges of a parenthesised, comma-separated expression list
    std::vector<std::pair<size_t, size_t>> expr_ranges()
    {
        std::vector<std::pair<size_t, size_t>> r;
        expect("(");
        if (is_sym(")")) {
            next();
            return r;
        }
        size_t b = p_;
        int depth = 0;
        for (;;) {
            const Tok& tk = peek();
            if (tk.kind == Tok::END) throw QasmError(tk.line, "unterminated parameter list");
            if (tk.kind == Tok::SYM && tk.text == "(") depth++;
            if (tk.kind == Tok::SYM && tk.text == ")") {
                if (depth == 0) {
                    if (p_ == b) throw QasmError(tk.line, "empty parameter");
                    r.push_back({b, p_});
                    next();
                    return r;
                }
                depth--;
            }
            if (tk.kind == Tok::SYM && tk.text == "," && depth == 0) {
                if (p_ == b) throw QasmError(tk.line, "empty parameter");
                r.push_back({b, p_});
                next();
                b = p_;
                continue;
            }
            next();
        }
    }

    // ---------------- statements ----------------
    struct Arg {
        std::string reg;
        long index;   // -1: the whole register
        int line;
    };
    Arg argument()
    {
        Arg a;
        a.line = peek().line;
        a.reg = ident();
        a.index = -1;
        if (is_sym("[")) {
            next();
            a.index = integer();
            expect("]");
        }
        return a;
    }
    std::vector<Arg> arguments()
    {
        std::vector<Arg> v{argument()};
        while (is_sym(",")) {
            next();
            v.push_back(argument());
        }
        return v;
    }
    std::vector<std::string> idlist()
    {
        std::vector<std::string> v{ident()};
        while (is_sym(",")) {
            next();
            v.push_back(ident());
        }
        return v;
    }

    void statement()
    {
        const Tok& tk = peek();
        if (tk.kind != Tok::ID) throw QasmError(tk.line, "unexpected '" + tk.text + "'");
        const std::string w = tk.text;
        if (w == "OPENQASM") {
            next();
            if (peek().kind != Tok::NUM) throw QasmError(peek().line, "expected a version number");
            next();
            expect(";");
        } else if (w == "include") {
            next();
            if (peek().kind != Tok::STR) throw QasmError(peek().line, "expected a file name");
            next();
            expect(";");
        } else if (w == "qreg" || w == "creg") {
            next();
            const int line = peek().line;
            const std::string name = ident();
            expect("[");
            const long size = integer();
            expect("]");
            expect(";");
            if (size < 1) throw QasmError(line, "register '" + name + "' needs at least one bit");
            if (qregs_.count(name) || cregs_.count(name)) throw QasmError(line, "register '" + name + "' redeclared");
            if (w == "qreg") {
                if (n_qubits_ + size > 63) throw QasmError(line, "more than 63 qubits");
                qregs_[name] = {n_qubits_, int(size)};
                n_qubits_ += int(size);
            } else {
                cregs_[name] = int(size);
            }
        } else if (w == "gate" || w == "opaque") {
            gate_definition(w == "opaque");
        } else if (w == "measure") {
            next();
            argument();
            expect("->");
            argument();
            expect(";");
            out_.skipped++;
        } else if (w == "reset") {
            next();
            argument();
            expect(";");
            out_.skipped++;
        } else if (w == "barrier") {
            next();
            arguments();
            expect(";");
            out_.skipped++;
        } else if (w == "if") {
            next();
            expect("(");
            ident();
            expect("==");
            integer();
            expect(")");
            // the guarded operation is parsed and dropped
            const size_t ng = out_.gates.size();
            const uint64_t b = out_.basic_gates, c = out_.cx_gates, sk = out_.skipped;
            statement();
            out_.gates.resize(ng);
            out_.basic_gates = b;
            out_.cx_gates = c;
            out_.skipped = sk + 1;
        } else {
            gate_statement();
        }
    }

    void gate_definition(bool opaque)
    {
        next();
        const int line = peek().line;
        const std::string name = ident();
        if (builtins().count(name) || defs_.count(name)) throw QasmError(line, "gate '" + name + "' redefined");
        GateDef d;
        d.opaque = opaque;
        if (is_sym("(")) {
            next();
            if (!is_sym(")")) d.params = idlist();
            expect(")");
        }
        d.qargs = idlist();
        if (opaque) {
            expect(";");
            defs_[name] = d;
            return;
        }
        expect("{");
        while (!is_sym("}")) {
            if (peek().kind == Tok::END) throw QasmError(line, "unterminated body of gate '" + name + "'");
            BodyCall c;
            c.line = peek().line;
            c.name = ident();
            if (c.name == "barrier") {
                idlist();
                expect(";");
                continue;
            }
            if (is_sym("(")) c.params = expr_ranges();
            c.qargs = idlist();
            expect(";");
            for (const std::string& q : c.qargs)
                if (std::find(d.qargs.begin(), d.qargs.end(), q) == d.qargs.end())
                    throw QasmError(c.line, "'" + q + "' is not an argument of gate '" + name + "'");
            if (!builtins().count(c.name) && !defs_.count(c.name))
                throw QasmError(c.line, "gate '" + c.name + "' used before its definition");
            d.body.push_back(c);
        }
        next();
        defs_[name] = d;
    }

    void gate_statement()
    {
        const int line = peek().line;
        const std::string name = ident();
        std::vector<double> params;
        if (is_sym("(")) {
            for (auto& r : expr_ranges()) params.push_back(eval_range(r.first, r.second, Env{}));
        }
        const std::vector<Arg> args = arguments();
        expect(";");
        // whole-register arguments repeat the gate over their bits
        long reps = -1;
        for (const Arg& a : args) {
            auto it = qregs_.find(a.reg);
            if (it == qregs_.end()) throw QasmError(a.line, "unknown quantum register '" + a.reg + "'");
            if (a.index >= it->second.second || a.index < -1)
                throw QasmError(a.line, a.reg + "[" + std::to_string(a.index) + "] is out of range");
            if (a.index == -1 && it->second.second > 1) {
                if (reps != -1 && reps != it->second.second)
                    throw QasmError(a.line, "registers of different sizes in one gate statement");
                reps = it->second.second;
            }
        }
        if (reps == -1) reps = 1;
        for (long b = 0; b < reps; b++) {
            std::vector<int> qs;
            for (const Arg& a : args) {
                const auto& r = qregs_.at(a.reg);
                qs.push_back(r.first + int(a.index >= 0 ? a.index : (r.second > 1 ? b : 0)));
            }
            apply(name, params, qs, line, 0);
        }
    }

    void apply(const std::string& name, const std::vector<double>& params, const std::vector<int>& qs, int line,
               int depth)
    {
        if (depth > 64) throw QasmError(line, "gate definitions nest deeper than 64");
        for (size_t i = 0; i < qs.size(); i++)
            for (size_t j = i + 1; j < qs.size(); j++)
                if (qs[i] == qs[j]) throw QasmError(line, "gate '" + name + "' repeats a qubit");
        auto b = builtins().find(name);
        if (b != builtins().end()) {
            const Builtin& g = b->second;
            if (int(params.size()) != g.n_params)
                throw QasmError(line, "gate '" + name + "' takes " + std::to_string(g.n_params) + " parameters");
            if (int(qs.size()) != g.n_qubits)
                throw QasmError(line, "gate '" + name + "' acts on " + std::to_string(g.n_qubits) + " qubits");
            svmi_gate r;
            r.op = g.op;
            for (int i = 0; i < 5; i++) r.qb[i] = i < g.n_qubits ? qs[i] : 0;
            double slot[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < g.n_params; i++) slot[g.slot[i]] = params[i];
            r.theta = slot[0];
            r.phi = slot[1];
            r.lambda = slot[2];
            out_.gates.push_back(r);
            out_.basic_gates += uint64_t(g.basic);
            out_.cx_gates += uint64_t(g.cx);
            return;
        }
        auto d = defs_.find(name);
        if (d == defs_.end()) throw QasmError(line, "unknown gate '" + name + "'");
        const GateDef& def = d->second;
        if (def.opaque) throw QasmError(line, "opaque gate '" + name + "' has no definition to simulate");
        if (params.size() != def.params.size())
            throw QasmError(line, "gate '" + name + "' takes " + std::to_string(def.params.size()) + " parameters");
        if (qs.size() != def.qargs.size())
            throw QasmError(line, "gate '" + name + "' acts on " + std::to_string(def.qargs.size()) + " qubits");
        Env env;
        for (size_t i = 0; i < params.size(); i++) env[def.params[i]] = params[i];
        std::map<std::string, int> qmap;
        for (size_t i = 0; i < qs.size(); i++) qmap[def.qargs[i]] = qs[i];
        for (const BodyCall& c : def.body) {
            std::vector<double> ps;
            for (auto& r : c.params) ps.push_back(eval_range(r.first, r.second, env));
            std::vector<int> q2;
            for (const std::string& q : c.qargs) q2.push_back(qmap.at(q));
            apply(c.name, ps, q2, c.line, depth + 1);
        }
    }

    std::vector<Tok> t_;
    size_t p_ = 0;
    size_t end_ = SIZE_MAX;
    QasmProgram& out_;
    std::map<std::string, std::pair<int, int>> qregs_;   // name -> (first qubit, size)
    std::map<std::string, int> cregs_;
    std::map<std::string, GateDef> defs_;
    int n_qubits_ = 0;
};

}  // namespace

svmi_status parse_qasm(const char* text, size_t len, QasmProgram& out, std::string& err)
{
    out = QasmProgram();
    try {
        Parser p(tokenize(text, len), out);
        p.run();
    } catch (const QasmError& e) {
        err = "line " + std::to_string(e.line) + ": " + e.what();
        return SVMI_ERR_ARG;
    }
    return SVMI_OK;
}

}  // namespace svmi
