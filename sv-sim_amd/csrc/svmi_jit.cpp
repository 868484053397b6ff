// svmi_jit.cpp -- source generation, hipRTC compilation and launch of
// specialised sweep kernels (see svmi_jit.hpp).
#include "svmi_jit.hpp"

#include <hip/hiprtc.h>

#include <chrono>
#include <cstdio>
#include <sstream>
#include <stdexcept>

namespace svmi {

// The engine headers the generated source includes (embedded by build.py).
extern const char* const kJitHeaderNames[];
extern const char* const kJitHeaderText[];
extern const int kJitHeaderCount;

namespace {

constexpr uint32_t kJitMaxTileQubits = 12;   // static LDS of one tile: 64 KiB

void emit_hdr(std::ostringstream& o, const DOpHdr& h)
{
    o << "{" << h.type << "u, " << h.smask << "u, " << h.s0 << "u, " << h.s1 << "u, " << h.coef << "u, " << h.grad
      << "u, " << h.rsel << "u, " << h.gmask << "u, " << h.emask << "ull}";
}

// Registers r (0..15) whose register bits satisfy `rsel`, as a list.
std::vector<int> selected(uint32_t rsel)
{
    std::vector<int> v;
    for (int r = 0; r < kRegs; r++)
        if ((rsel >> r) & 1u) v.push_back(r);
    return v;
}

// Tile-local offset of register r in phase ph (the register bits' part).
uint32_t reg_offset(const PhaseInfo& ph, int r)
{
    uint32_t o = 0;
    for (int b = 0; b < kRegBits; b++)
        if ((r >> b) & 1) o |= 1u << ph.r[b];
    return o;
}

// One op as direct code: the templates of svmi_exec.hpp / svmi_internal.hpp
// with every slot, mask and offset a literal, so no dispatch is compiled.
void emit_state_op(std::ostringstream& o, const DOp& d, const PhaseInfo& ph)
{
    const DOpHdr& h = d.h;
    const std::string cond = "(base & " + std::to_string(h.emask) + "ull) == " + std::to_string(h.emask) +
                             "ull && (goff & " + std::to_string(h.gmask) + "u) == " + std::to_string(h.gmask) + "u";
    switch (h.type) {
    case DOP_DENSE: o << "        apply_dense<" << h.smask << ">(a, coef + " << h.coef << ");\n"; return;
    case DOP_X: o << "        apply_x<" << h.s0 << ">(a);\n"; return;
    case DOP_CX: o << "        apply_cx<" << h.s0 << ", " << h.s1 << ">(a);\n"; return;
    case DOP_SWAP: o << "        apply_swap<" << h.s0 << ", " << h.s1 << ">(a);\n"; return;
    case DOP_SRN: o << "        apply_srn<" << h.s0 << ">(a);\n"; return;
    case DOP_LIFT2:
    case DOP_LIFT3:
        o << "        apply_lift<" << h.s0 << ", " << (h.type == DOP_LIFT3 ? "true" : "false") << ">(a, coef["
          << h.coef << "].re, coef[" << h.coef << "].im);\n";
        return;
    case DOP_RTAB:
        o << "        apply_rtab(a, coef + " << h.coef << "u + 16u * rtab_variant(" << h.emask << "ull, base));\n";
        return;
    case DOP_ROT:
        if (h.s0)
            o << "        apply_rtab(a, coef + " << h.grad << "u + 16u * rtab_variant(" << h.emask << "ull, base));\n";
        for (int sl = 0; sl < kRegBits; sl++)
            if ((h.smask >> sl) & 1u)
                o << "        apply_lift<" << sl << ", " << (((h.s1 >> sl) & 1u) ? "true" : "false") << ">(a, coef["
                  << (h.coef + uint32_t(sl)) << "].re, coef[" << (h.coef + uint32_t(sl)) << "].im);\n";
        return;
    case DOP_CDENSE:
        o << "        if (" << cond << ") apply_cdense<" << h.s0 << ">(a, coef + " << h.coef << ", " << h.rsel
          << "u);\n";
        return;
    case DOP_CPHASE: {
        o << "        if (" << cond << ") {\n            const c64 ph = coef[" << h.coef << "];\n";
        for (int r : selected(h.rsel)) o << "            a[" << r << "] = cmul(a[" << r << "], ph);\n";
        o << "        }\n";
        return;
    }
    case DOP_DIAG: {
        // table index: bit i from qubit i -- outside the tile from the tile
        // base, on a register bit a constant per register, else from goff
        o << "        {\n            const uint32_t ix = 0u";
        for (uint32_t i = 0; i < d.nq; i++) {
            if (d.qloc[i] < 0) o << " | (uint32_t((base >> " << d.qphys[i] << ") & 1u) << " << i << ")";
            else {
                bool reg = false;
                for (int b = 0; b < kRegBits; b++) reg = reg || int(ph.r[b]) == d.qloc[i];
                if (!reg) o << " | (((goff >> " << d.qloc[i] << ") & 1u) << " << i << ")";
            }
        }
        o << ";\n";
        for (int r = 0; r < kRegs; r++) {
            const uint32_t ro = reg_offset(ph, r);
            uint32_t cidx = 0;
            for (uint32_t i = 0; i < d.nq; i++)
                if (d.qloc[i] >= 0 && ((ro >> d.qloc[i]) & 1u)) cidx |= 1u << i;
            o << "            a[" << r << "] = cmul(a[" << r << "], coef[" << h.coef << "u + (ix | " << cidx
              << "u)]);\n";
        }
        o << "        }\n";
        return;
    }
    case DOP_PHASESET: {
        o << "        {\n            const PhaseEntry* pe = reinterpret_cast<const PhaseEntry*>(coef + " << h.coef
          << ");\n            double pr = " << (h.s1 ? "xf[" + std::to_string(h.grad) + "].re" : "1.0")
          << ", pi = " << (h.s1 ? "xf[" + std::to_string(h.grad) + "].im" : "0.0") << ";\n"
          << "#pragma unroll 4\n            for (uint32_t e = 0; e < " << h.s0 << "u; e++) {\n"
          << "                const PhaseEntry E = pe[e];\n"
          << "                if ((base & E.emask) != E.emask) continue;\n"
          << "                const bool on = (goff & E.gmask) == E.gmask;\n"
          << "                const double tr = SVMI_FMA(pr, E.ph_re, -(pi * E.ph_im));\n"
          << "                const double ti = SVMI_FMA(pr, E.ph_im, pi * E.ph_re);\n"
          << "                pr = on ? tr : pr;\n                pi = on ? ti : pi;\n            }\n"
          << "            const c64 ph{pr, pi};\n";
        for (int r : selected(h.rsel)) o << "            a[" << r << "] = cmul(a[" << r << "], ph);\n";
        o << "        }\n";
        return;
    }
    case DOP_PHASETAB: {
        o << "        {\n            c64 ph = " << (h.s1 ? "xf[" + std::to_string(h.grad) + "]" : "c64{1.0, 0.0}")
          << ";\n";
        for (uint32_t j = 0; j < h.s0; j++)
            o << "            ph = cmul(ph, coef[" << (h.coef + 16u * j) << "u + ((tid >> " << 4 * j << ") & 15u)]);\n";
        for (int r : selected(h.rsel)) o << "            a[" << r << "] = cmul(a[" << r << "], ph);\n";
        o << "        }\n";
        return;
    }
    default: o << "        #error unsupported op type " << h.type << "\n"; return;
    }
}

// `var` = the tile base of tile id `tile` (the id's bits deposited on the
// qubits outside the tile).
void emit_base(std::ostringstream& o, const SweepInfo& sw, const char* var, const char* tile, const char* ind)
{
    o << ind << "uint64_t " << var << " = 0;\n";
    int j = 0;
    for (const uint64_t m : {sw.extmask & ~sw.exthi, sw.exthi})
        for (int p = 0; p < 64; p++)
            if ((m >> p) & 1) o << ind << var << " |= ((" << tile << " >> " << j++ << ") & 1ull) << " << p << ";\n";
}

// The tile's phase factors into xf (uses `base`, `tid`).
void emit_factors(std::ostringstream& o, const Program& prog, const SweepInfo& sw, uint32_t nt, const char* ind)
{
    if (sw.n_xf) {
        o << ind << "uint64_t xbits = 0;\n";
        int jx = 0;
        for (int p = 0; p < 64; p++)
            if ((sw.extmask >> p) & 1) o << ind << "xbits |= ((base >> " << p << ") & 1ull) << " << jx++ << ";\n";
    }
    for (uint32_t op = sw.op_begin; op < sw.op_end; op++) {
        const DOpHdr& h = prog.ops[op].h;
        if (has_tile_factor(h)) {
            o << ind << "if (tid == " << (h.grad % nt) << "u) {\n" << ind << "    constexpr DOpHdr H = ";
            emit_hdr(o, h);
            o << ";\n" << ind << "    xf[" << h.grad << "] = tile_factor(H, coef, base, xbits);\n" << ind << "}\n";
        }
    }
}

// The sweep's phases on the tile at t0 (each ends with a barrier).
void emit_phases(std::ostringstream& o, const Program& prog, const SweepInfo& sw)
{
    for (uint32_t p = sw.phase_begin; p < sw.phase_end; p++) {
        const PhaseInfo& ph = prog.phases[p];
        o << "    {\n        constexpr PhaseInfo P = {{" << ph.r[0] << "u, " << ph.r[1] << "u, " << ph.r[2] << "u, "
          << ph.r[3] << "u}, 0u, 0u, {" << ph.gcol[0] << "u, " << ph.gcol[1] << "u, " << ph.gcol[2] << "u, "
          << ph.gcol[3] << "u}, " << ph.gxor << "u, " << ph.sxor << "u};\n"
          << "        uint32_t off[kRegs], goffs[kRegs];\n        phase_offsets(P, off);\n"
          << "        phase_gather_offsets(P, goffs);\n"
          << "        const uint32_t goff = group_offset(P, tid);\n        c64 a[kRegs];\n"
          << "#pragma unroll\n        for (int r = 0; r < kRegs; r++) a[r] = t0[lds_slot(goff | off[r])];\n";
        // the phase's leading register permutations, which the interpreter
        // folds into its gather map, run here as the ops they were (ops are
        // stored contiguously phase after phase, so they sit between the
        // previous phase's end and this phase's first op): measured faster
        // than a gather through permuted literal offsets in the QFT's
        // permutation-only sweeps
        const uint32_t op0 = p ? prog.phases[p - 1].op_end : 0u;
        bool unfold = op0 <= ph.op_begin;
        for (uint32_t op = op0; unfold && op < ph.op_begin; op++) {
            const uint32_t t = prog.ops[op].h.type;
            unfold = t == DOP_X || t == DOP_CX || t == DOP_SWAP;
        }
        uint32_t gcol[kRegBits], gxor = 0;
        for (int b = 0; b < kRegBits; b++) gcol[b] = 1u << ph.r[b];
        for (uint32_t op = op0; unfold && op < ph.op_begin; op++) {   // must rebuild this phase's map
            const DOpHdr& h = prog.ops[op].h;
            if (h.type == DOP_X) gxor ^= gcol[h.s0];
            else if (h.type == DOP_CX) gcol[h.s0] ^= gcol[h.s1];
            else std::swap(gcol[h.s0], gcol[h.s1]);
        }
        for (int b = 0; b < kRegBits; b++) unfold = unfold && gcol[b] == ph.gcol[b];
        unfold = unfold && gxor == ph.gxor;
        if (!unfold) throw std::logic_error("jit: a phase's gather map is not its leading permutations");
        for (uint32_t op = op0; op < ph.op_begin; op++) emit_state_op(o, prog.ops[op], ph);
        for (uint32_t op = ph.op_begin; op < ph.op_end; op++) emit_state_op(o, prog.ops[op], ph);
        o << "#pragma unroll\n        for (int r = 0; r < kRegs; r++) t0[lds_slot((goff | off[r]) ^ " << ph.sxor
          << "u)] = a[r];\n"
          << "    }\n    __syncthreads();\n";
    }
}

void emit_sweep(std::ostringstream& o, const Program& prog, uint32_t i)
{
    const SweepInfo& sw = prog.sweeps[i];
    const uint32_t nt = 1u << (sw.k - kRegBits);
    // second bound: at least two waves per SIMD, i.e. at most 256 VGPRs
    o << "extern \"C\" __global__ void __launch_bounds__(" << nt << ", 2)\n"
      << "svmi_jit_" << i << "(c64* __restrict__ s, const c64* __restrict__ coef, const uint64_t* __restrict__ ht, "
         "uint64_t tile0)\n{\n"
      << "    __shared__ __align__(16) c64 t0[" << (1u << sw.k) << "];\n"
      << "    __shared__ __align__(16) c64 xf[" << (sw.n_xf ? sw.n_xf : 1u) << "];\n"
      << "    const uint32_t tid = threadIdx.x;\n"
      << "    const uint64_t tile = tile0 + blockIdx.x;\n";
    emit_base(o, sw, "base", "tile", "    ");
    const uint32_t cmask = (1u << sw.c) - 1u;
    o << "#pragma unroll\n    for (int j = 0; j < " << kRegs << "; j++) {\n"
      << "        const uint32_t l = tid + " << nt << "u * j;\n"
      << "        t0[lds_slot(l)] = s[base | ht[" << sw.hitab << "u + (l >> " << sw.c << ")] | (l & " << cmask
      << "u)];\n    }\n";
    emit_factors(o, prog, sw, nt, "    ");
    o << "    __syncthreads();\n";
    emit_phases(o, prog, sw);
    o << "#pragma unroll\n    for (int j = 0; j < " << kRegs << "; j++) {\n"
      << "        const uint32_t l = tid + " << nt << "u * j;\n"
      << "        s[base | ht[" << sw.hitab << "u + (l >> " << sw.c << ")] | (l & " << cmask
      << "u)] = t0[lds_slot(l)];\n    }\n}\n\n";
}

// Sweeps of 12-qubit tiles as persistent double-buffered kernels (one
// workgroup of 256 threads per CU): while the phases of tile t run in one LDS
// buffer, tile t + gridDim.x streams into the other by LDS DMA (tile_dma);
// the wait for a tile's DMA leaves only the previous tile's 16 stores per
// thread outstanding (vmcnt(16); the first tile waits for everything).
bool dma_eligible(const SweepInfo& sw) { return sw.k == kDmaTileQubits && sw.c <= kDmaTileQubits; }

void emit_sweep_dma(std::ostringstream& o, const Program& prog, uint32_t i)
{
    const SweepInfo& sw = prog.sweeps[i];
    const uint32_t nt = kDmaThreads, kt = 1u << kDmaTileQubits, rows = kt >> sw.c;
    const uint32_t cmask = (1u << sw.c) - 1u;
    o << "extern \"C\" __global__ void __launch_bounds__(" << nt << ", 1)\n"
      << "svmi_jit_" << i << "(c64* __restrict__ s, const c64* __restrict__ coef, const uint64_t* __restrict__ hitab, "
         "uint64_t n_tiles)\n{\n"
      << "    __shared__ __align__(16) c64 tb[" << 2 * kt << "];\n"
      << "    __shared__ __align__(16) uint64_t ht[" << rows << "];\n"
      << "    __shared__ __align__(16) c64 xf[" << (sw.n_xf ? sw.n_xf : 1u) << "];\n"
      << "    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;\n"
      << "    for (uint32_t i = tid; i < " << rows << "u; i += " << nt << "u) ht[i] = hitab[" << sw.hitab << "u + i];\n"
      << "    __syncthreads();\n"
      << "    const uint32_t lds0 = lds_address(tb);\n"
      << "    uint64_t t = blockIdx.x;\n"
      << "    {\n";
    emit_base(o, sw, "b0", "t", "        ");
    o << "        tile_dma(s, ht, " << sw.c << "u, b0, lds0, wave, lane);\n    }\n"
      << "    uint32_t cur = 0;\n    bool first = true;\n    for (;;) {\n"
      << "        if (first) asm volatile(\"s_waitcnt vmcnt(0)\" ::: \"memory\");\n"
      << "        else asm volatile(\"s_waitcnt vmcnt(16)\" ::: \"memory\");\n"
      << "        __syncthreads();\n";
    emit_base(o, sw, "base", "t", "        ");
    o << "        const uint64_t tn = t + gridDim.x;\n        const bool more = tn < n_tiles;\n";
    emit_factors(o, prog, sw, nt, "        ");
    o << "        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): no load of the factors still in flight\n"
      << "        if (more) {\n";
    emit_base(o, sw, "bn", "tn", "            ");
    o << "            tile_dma(s, ht, " << sw.c << "u, bn, lds0 + (cur ^ 1u) * " << kt * 16 << "u, wave, lane);\n"
      << "        }\n        __syncthreads();\n"
      << "        c64* t0 = tb + cur * " << kt << "u;\n";
    emit_phases(o, prog, sw);
    o << "#pragma unroll\n        for (int j = 0; j < " << kRegs << "; j++) {\n"
      << "            const uint32_t l = tid + " << nt << "u * j;\n"
      << "            s[base | ht[l >> " << sw.c << "] | (l & " << cmask << "u)] = t0[lds_slot(l)];\n        }\n"
      << "        if (!more) break;\n        t = tn;\n        cur ^= 1u;\n        first = false;\n    }\n}\n\n";
}

struct Rtc {
    hiprtcProgram prog = nullptr;
    ~Rtc()
    {
        if (prog) hiprtcDestroyProgram(&prog);
    }
};

bool compile(const std::string& src, const char* arch, std::vector<char>& code, double* ms, std::string& err)
{
    Rtc r;
    if (hiprtcCreateProgram(&r.prog, src.c_str(), "svmi_jit.hip", kJitHeaderCount,
                            const_cast<const char**>(kJitHeaderText),
                            const_cast<const char**>(kJitHeaderNames)) != HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return false;
    }
    const std::string a = std::string("--offload-arch=") + arch;
    const char* opts[] = {a.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops", "-mllvm", "-structurizecfg-skip-uniform-regions=true"};
    const auto t0 = std::chrono::steady_clock::now();
    const hiprtcResult rc = hiprtcCompileProgram(r.prog, int(sizeof(opts) / sizeof(opts[0])), opts);
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(r.prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(r.prog, &log[0]);
        err = std::string("hiprtc: ") + hiprtcGetErrorString(rc) + "\n" + log.substr(0, 2000);
        return false;
    }
    size_t n = 0;
    hiprtcGetCodeSize(r.prog, &n);
    code.resize(n);
    hiprtcGetCode(r.prog, code.data());
    return n > 0;
}

}  // namespace

void JitSweeps::release()
{
    if (module) hipModuleUnload(module);
    module = nullptr;
    fn.clear();
    dma.clear();
    dma_grid = 0;
}

std::string jit_source(const Program& prog, std::vector<char>& covered, uint64_t dma_grid)
{
    covered.assign(prog.sweeps.size(), 0);
    std::ostringstream o;
    o << "#include \"svmi_exec.hpp\"\nusing namespace svmi;\n\n";
    if (prog.n_states != 1) return o.str();
    for (uint32_t i = 0; i < prog.sweeps.size(); i++) {
        const SweepInfo& sw = prog.sweeps[i];
        if (sw.k > kJitMaxTileQubits || sw.mapped) continue;   // tile maps: interpreted
        if (dma_grid && dma_eligible(sw) && sw.n_tiles >= 2 * dma_grid) {
            emit_sweep_dma(o, prog, i);
            covered[i] = 2;
        } else {
            emit_sweep(o, prog, i);
            covered[i] = 1;
        }
    }
    return o.str();
}

size_t jit_compile_only(const Program& prog, const char* arch, double* ms, std::string& err, uint64_t dma_grid)
{
    std::vector<char> covered, code;
    const std::string src = jit_source(prog, covered, dma_grid);
    if (!compile(src, arch, code, ms, err)) return 0;
    return code.size();
}

bool jit_build(const Program& prog, const char* arch, JitSweeps& out, std::string& err, uint64_t dma_grid)
{
    out.release();
    std::vector<char> covered, code;
    const std::string src = jit_source(prog, covered, dma_grid);
    out.dma.assign(prog.sweeps.size(), 0);
    for (size_t i = 0; i < covered.size(); i++) out.dma[i] = covered[i] == 2;
    out.dma_grid = dma_grid;
    bool any = false;
    for (char c : covered) any = any || c;
    if (!any) return false;
    if (!compile(src, arch, code, &out.compile_ms, err)) return false;
    if (hipModuleLoadData(&out.module, code.data()) != hipSuccess) {
        err = "hipModuleLoadData failed";
        out.module = nullptr;
        return false;
    }
    out.fn.assign(prog.sweeps.size(), nullptr);
    for (uint32_t i = 0; i < prog.sweeps.size(); i++) {
        if (!covered[i]) continue;
        const std::string name = "svmi_jit_" + std::to_string(i);
        if (hipModuleGetFunction(&out.fn[i], out.module, name.c_str()) != hipSuccess) {
            err = "hipModuleGetFunction(" + name + ") failed";
            out.release();
            return false;
        }
    }
    return true;
}

hipError_t jit_launch(hipStream_t st, const JitSweeps& j, uint32_t i, const SweepInfo& sw, c64* s, const c64* coef,
                      const uint64_t* hitab)
{
    if (i < j.dma.size() && j.dma[i]) {   // persistent: one workgroup per CU
        const uint32_t grid = uint32_t(sw.n_tiles < j.dma_grid ? sw.n_tiles : j.dma_grid);
        uint64_t n_tiles = sw.n_tiles;
        void* args[] = {&s, &coef, &hitab, &n_tiles};
        return hipModuleLaunchKernel(j.fn[i], grid, 1, 1, kDmaThreads, 1, 1, 0, st, args, nullptr);
    }
    const uint32_t nt = 1u << (sw.k - kRegBits);
    const uint64_t max_grid = 1ull << 30;
    for (uint64_t t0 = 0; t0 < sw.n_tiles; t0 += max_grid) {
        const uint32_t cnt = uint32_t(sw.n_tiles - t0 < max_grid ? sw.n_tiles - t0 : max_grid);
        uint64_t tile0 = t0;
        void* args[] = {&s, &coef, &hitab, &tile0};
        const hipError_t e = hipModuleLaunchKernel(j.fn[i], cnt, 1, 1, nt, 1, 1, 0, st, args, nullptr);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace svmi
