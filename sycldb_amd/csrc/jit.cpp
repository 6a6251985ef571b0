// Fused bundle kernels compiled at run time for one program shape.
//
// The reference designed its fusion around a JIT: the SYCLDBContext slot
// operations of kernels/jit_kernels_impl.hpp:23-90 specialised per bundle
// (models/execution.hpp:546-615), switched off in its runs.  Here a fused
// program (fused_types.hpp) is turned into HIP source in which every
// instruction is a call of a device template from fused_ops.hpp with its
// code-selecting parameters (comparison, logical and arithmetic op, operand
// kind, table placement, value-table kind) fixed, and every flag and value
// register is a plain variable.  The remaining parameters -- literals,
// pointers, table bounds -- are read from the program itself, which is the
// kernel's argument exactly as for the interpreter kernel, so one compiled
// kernel serves every program of the same shape.
//
// Source text -> hipRTC (gfx950 code object, cached per architecture and
// source) -> hipModuleLoadData per device (cached) -> hipModuleLaunchKernel.
// The hipRTC library is loaded at first use, preferring the one beside the
// process's HIP runtime so that the compiler and the runtime come from the
// same ROCm.
#include <dlfcn.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <unordered_map>
#include <mutex>
#include <sstream>
#include <thread>
#include <vector>

#include "capi_util.hpp"
#include "fused.hpp"

namespace sdb {

namespace {

// the in-memory headers the generated source includes
#include "jit_headers.inc"

constexpr int JIT_BLOCK = 256;

struct Rtc {
    bool ok = false;
    std::string why, path;
    hiprtcResult (*create)(hiprtcProgram *, const char *, const char *, int, const char *const *,
                           const char *const *) = nullptr;
    hiprtcResult (*compile)(hiprtcProgram, int, const char *const *) = nullptr;
    hiprtcResult (*log_size)(hiprtcProgram, size_t *) = nullptr;
    hiprtcResult (*log)(hiprtcProgram, char *) = nullptr;
    hiprtcResult (*code_size)(hiprtcProgram, size_t *) = nullptr;
    hiprtcResult (*code)(hiprtcProgram, char *) = nullptr;
    hiprtcResult (*destroy)(hiprtcProgram *) = nullptr;
};

Rtc load_rtc()
{
    Rtc r;
    std::vector<std::string> cand;
    Dl_info info{};
    if (dladdr(reinterpret_cast<void *>(&hipModuleLoadData), &info) && info.dli_fname) {
        std::string dir(info.dli_fname);
        const size_t slash = dir.rfind('/');
        if (slash != std::string::npos) {
            dir.resize(slash);
            cand.push_back(dir + "/libhiprtc.so");
            cand.push_back(dir + "/libhiprtc.so.7");
        }
    }
    cand.push_back("libhiprtc.so.7");
    cand.push_back("/opt/rocm/lib/libhiprtc.so.7");
    void *h = nullptr;
    for (const std::string &c : cand) {
        h = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (h) {
            r.path = c;
            break;
        }
    }
    if (!h) {
        r.why = "hipRTC library not found";
        return r;
    }
    auto sym = [&](const char *name) {
        void *p = dlsym(h, name);
        if (!p && r.why.empty()) r.why = std::string("hipRTC symbol missing: ") + name;
        return p;
    };
    r.create = reinterpret_cast<decltype(r.create)>(sym("hiprtcCreateProgram"));
    r.compile = reinterpret_cast<decltype(r.compile)>(sym("hiprtcCompileProgram"));
    r.log_size = reinterpret_cast<decltype(r.log_size)>(sym("hiprtcGetProgramLogSize"));
    r.log = reinterpret_cast<decltype(r.log)>(sym("hiprtcGetProgramLog"));
    r.code_size = reinterpret_cast<decltype(r.code_size)>(sym("hiprtcGetCodeSize"));
    r.code = reinterpret_cast<decltype(r.code)>(sym("hiprtcGetCode"));
    r.destroy = reinterpret_cast<decltype(r.destroy)>(sym("hiprtcDestroyProgram"));
    r.ok = r.why.empty();
    return r;
}

Rtc &rtc()
{
    static Rtc r = load_rtc();
    return r;
}

struct Compiled {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    std::map<uint32_t, int> blocks_per_cu; // by dynamic LDS bytes
};

std::mutex g_mu;
std::map<std::string, std::vector<char>> g_code;          // arch + source -> code object
std::map<std::pair<int, std::string>, Compiled> g_loaded; // (device, arch + source)
int g_compiled = 0;

// ------------------------------------------------------------ generation

struct Gen {
    const FProgram &p;
    std::ostringstream o;
    bool prefetch_in_body; // the body issues the next tile's column loads
    // batched form: the body takes `act` (rows whose shared probe masks keep
    // this program); flags start from it and the terminal reads only its
    // rows.  For the probes marked in `decided` the masks have established
    // the match on every `act` row: a semi probe passes its input flags
    // through, a full probe only fetches its value.
    bool batch_act = false;
    std::vector<bool> decided;
    explicit Gen(const FProgram &prog, bool prefetch = true) : p(prog), prefetch_in_body(prefetch) {}

    bool is_decided(int i) const { return i < (int)decided.size() && decided[i]; }
    bool dropped(int i) const
    {
        const FInstr &I = p.ins[i];
        return is_decided(i) && (I.kind == FK_PROBE_SEMI || (I.kind == FK_PROBE_BITS && I.vkind == FV_NONE));
    }

    static std::string freg(int j) { return "f" + std::to_string(j); }
    static std::string vreg(int r) { return "v" + std::to_string(r); }

    bool cheap(int reg) const { return reg >= 0; }

    // statement filling int32_t dst[R] with the operand's rows of chunk u
    std::string operand(int reg, const std::string &mem, const std::string &dst) const
    {
        if (reg >= F_PF_BASE) return "vec_rows<R>(pf.c[" + std::to_string(reg - F_PF_BASE) + "][u], " + dst + ");";
        if (reg >= 0) return "copy_rows<R>(" + vreg(reg) + "[u], " + dst + ");";
        auto it = preload.find(mem);
        if (it != preload.end()) return "copy_rows<R>(" + it->second + "[u], " + dst + ");";
        return "load_rows<R>(" + mem + ", row0[u], " + dst + ");";
    }

    // Early reads of lazily loaded columns.  A column an instruction reads
    // from memory for the rows of its input flags F is read, instead, right
    // after the last probe through a table in global memory that narrowed F,
    // when only conjunctive instructions on F lie between (so the rows then
    // in F include every row that will need it): its read then overlaps that
    // probe's value reads and the reads of the other columns so placed,
    // instead of waiting in turn.  preload: operand expression -> register
    // array; after_finish[k]: the reads issued after instruction k.
    std::map<std::string, std::string> preload;
    std::map<int, std::string> after_finish;
    std::string preload_decls;

    bool global_probe(int k) const
    {
        const FInstr &I = p.ins[k];
        if (dropped(k)) return false;
        return I.kind == FK_PROBE_SEMI || I.kind == FK_PROBE_FULL ||
               (I.kind == FK_PROBE_BITS && (!I.in_lds || (I.vkind != FV_NONE && I.vkind != FV_RANK16)));
    }

    int preload_anchor(int j, int F) const
    {
        for (int k = j - 1; k >= 0; k--) {
            const FInstr &I = p.ins[k];
            const bool on_f = writes_flags(I.kind) && I.fout == F;
            if (!on_f) continue;
            if (!(conjunctive(I) && I.fin == F)) return -1;
            if (global_probe(k)) return k;
        }
        return -1;
    }

    void plan_preloads()
    {
        preload.clear();
        after_finish.clear();
        preload_decls.clear();
        static const bool on = [] {
            const char *e = getenv("SDB_JIT_EARLY_COLUMNS");
            return !(e && *e == '0');
        }();
        if (!on) return;
        int n = 0;
        for (int j = 0; j < p.ninstr; j++) {
            const FInstr &I = p.ins[j];
            const std::string P = "P.ins[" + std::to_string(j) + "]";
            std::vector<std::string> mems;
            switch (I.kind) {
            case FK_SEL_LIT:
            case FK_SEL_RANGE:
            case FK_SEL_IN2:
                // only an AND selection reads for the rows of its input alone
                if (I.logic == SDB_AND && I.va < 0) mems.push_back(P + ".a");
                break;
            case FK_SEL_COL:
                if (I.logic == SDB_AND) {
                    if (I.va < 0) mems.push_back(P + ".a");
                    if (I.vb < 0) mems.push_back(P + ".b");
                }
                break;
            case FK_BINOP_CC:
                if (I.va < 0) mems.push_back(P + ".a");
                if (I.vb < 0) mems.push_back(P + ".b");
                break;
            case FK_BINOP_LC:
                if (I.vb < 0) mems.push_back(P + ".b");
                break;
            case FK_BINOP_CL:
            case FK_SUM:
            case FK_PROBE_SEMI:
            case FK_PROBE_FULL:
            case FK_PROBE_BITS:
                if (I.va < 0 && !dropped(j)) mems.push_back(P + ".a");
                break;
            case FK_GROUPBY:
                for (int g = 0; g < p.gb.k; g++)
                    if (p.gb.key_reg[g] < 0) mems.push_back("P.gb.key_mem[" + std::to_string(g) + "]");
                if (p.gb.agg_reg < 0) mems.push_back("P.gb.agg_mem");
                break;
            default: break;
            }
            if (mems.empty()) continue;
            const int k = preload_anchor(j, I.fin);
            if (k < 0) continue;
            const std::string F = freg(I.fin) + "[u]";
            std::string &code = after_finish[k];
            for (const std::string &m : mems) {
                if (preload.count(m)) continue;
                const std::string name = "early" + std::to_string(n++);
                preload[m] = name;
                preload_decls += "    int32_t " + name + "[UU][R];\n";
                code += "#pragma unroll\n    for (int u = 0; u < UU; u++) if (" + F + ") load_rows<R>(" + m +
                        ", row0[u], " + name + "[u]);\n";
            }
        }
    }

    std::vector<bool> hoisted;
    std::string hoist_code;

    // Load phase of probe i (operand rows into x<i>, table reads into t<i>)
    // for the rows of flag mask `gate`.
    std::string probe_issue(int i, const std::string &gate) const
    {
        const FInstr &I = p.ins[i];
        const std::string P = "P.ins[" + std::to_string(i) + "]";
        const std::string x = "x" + std::to_string(i) + "[u]", t = "t" + std::to_string(i) + "[u]";
        const std::string U = "#pragma unroll\n        for (int u = 0; u < UU; u++) ";
        std::ostringstream a;
        if (cheap(I.va)) a << U << "{ " << operand(I.va, P + ".a", x) << " }\n";
        else a << U << "if (" << gate << ") { " << operand(I.va, P + ".a", x) << " }\n";
        a << U << "if (" << gate << ") ";
        if (is_decided(i)) {
            // the value alone (FV_U16 / FV_PAIR tables, or the reference pairs)
            const int vk = I.kind == FK_PROBE_FULL ? (int)FV_NONE : (int)I.vkind;
            const std::string vals = I.kind == FK_PROBE_FULL ? P + ".ht" : P + ".vals";
            a << "probe_value_issue<R, " << vk << ">(" << x << ", " << gate << ", " << vals << ", " << P << ".ht_min, "
              << P << ".ht_max, " << P << ".vmin, " << t << ");\n";
        } else if (I.kind == FK_PROBE_SEMI)
            a << "probe_semi_issue<R>(" << x << ", " << gate << ", (const uint8_t *)" << P << ".ht, " << P << ".ht_min, "
              << P << ".ht_max, " << t << ");\n";
        else if (I.kind == FK_PROBE_FULL)
            a << "probe_full_issue<R>(" << x << ", " << gate << ", (const int2 *)" << P << ".ht, " << P << ".ht_min, "
              << P << ".ht_max, " << t << ");\n";
        else if (unset_probe(i))
            a << "probe_unset_issue<R>(" << x << ", " << gate << ", (const uint16_t *)" << P << ".vals, " << P
              << ".ht_min, " << P << ".ht_max, " << t << ");\n";
        else {
            const std::string bits =
                I.in_lds ? "(const uint32_t *)(lds + " + P + ".lds_off)" : "(const uint32_t *)" + P + ".ht";
            a << "probe_bits_issue<R, " << (late_value(i) ? 0 : (int)I.vkind) << ", " << (I.in_lds ? "true" : "false")
              << ">(" << x << ", " << gate << ", "
              << bits << ", " << P << ".ht_min, " << P << ".ht_max, " << P << ".vals, " << P << ".vmin, " << P
              << ".ht_len, " << t << ");\n";
        }
        return a.str();
    }

    static bool is_probe(uint8_t k) { return k == FK_PROBE_SEMI || k == FK_PROBE_FULL || k == FK_PROBE_BITS; }

    // A bitmap probe with a value table in global memory fetches the value
    // of the matched rows only, after the presence test: one more dependent
    // read, but no value reads for the rows that do not match (on the q2.x
    // star join at 600M rows, 2.22 -> 1.90 ms per launch).
    // SDB_JIT_LATE_VALUES=0 fetches word and value together.
    bool late_value(int i) const
    {
        static const bool on = [] {
            const char *e = getenv("SDB_JIT_LATE_VALUES");
            return !(e && *e == '0');
        }();
        const FInstr &I = p.ins[i];
        return on && !is_decided(i) && !unset_probe(i) && I.kind == FK_PROBE_BITS &&
               (I.vkind == FV_U16 || I.vkind == FV_PAIR);
    }

    // A plain program's bitmap probe whose bitmap is in global memory and
    // whose values are FV_U16 reads the value table alone (VALS_UNSET marks
    // the slots without a match): one round trip through the caches instead
    // of the presence word's and then the value's.  SDB_JIT_UNSET_PROBE=0
    // reads the bitmap first.
    bool unset_probe(int i) const
    {
        static const bool on = [] {
            const char *e = getenv("SDB_JIT_UNSET_PROBE");
            return !(e && *e == '0');
        }();
        const FInstr &I = p.ins[i];
        return on && !batch_act && !is_decided(i) && I.kind == FK_PROBE_BITS && I.vkind == FV_U16 && !I.in_lds;
    }

    static bool writes_flags(uint8_t k)
    {
        return k == FK_SEL_LIT || k == FK_SEL_COL || k == FK_SEL_RANGE || k == FK_SEL_IN2 || k == FK_LOGICAL ||
               is_probe(k);
    }

    static bool conjunctive(const FInstr &I)
    {
        return is_probe(I.kind) || (writes_flags(I.kind) && I.logic == SDB_AND);
    }

    // Probes after the first instruction that waits for global memory (fw)
    // whose global reads can be issued together with fw's: keyed by a
    // prefetched column, and reached from fw only through conjunctive
    // instructions on fw's flag register F (so the rows they will probe are
    // a subset of F's rows at fw, which gates their early reads).
    void plan_hoisting(int fw)
    {
        hoisted.assign(p.ninstr, false);
        if (fw < 0) return;
        const FInstr &W = p.ins[fw];
        const int F = W.fin;
        if (!(conjunctive(W) && W.fout == F)) return;
        const int end = tail_start >= 0 ? tail_start : p.ninstr;
        for (int i = fw + 1; i < end; i++) {
            const FInstr &I = p.ins[i];
            if (is_probe(I.kind) && I.fin == F && I.va >= F_PF_BASE && waits(i)) {
                hoisted[i] = true;
                hoist_code += probe_issue(i, freg(F) + "[u]");
            }
            if (writes_flags(I.kind) && !(conjunctive(I) && I.fin == F && I.fout == F)) break;
            if (I.kind == FK_LOGICAL && I.floc == F) break;
        }
    }

    // Whether instruction i waits for global memory: a column operand loaded
    // now, or a table outside LDS.
    bool waits(int i) const
    {
        const FInstr &I = p.ins[i];
        if (dropped(i)) return false;
        switch (I.kind) {
        case FK_LOGICAL:
        case FK_FILL: return false;
        case FK_PROBE_SEMI:
        case FK_PROBE_FULL: return true;
        case FK_PROBE_BITS:
            if (!I.in_lds || I.vkind != FV_NONE) return true;
            break;
        case FK_GROUPBY: {
            for (int j = 0; j < p.gb.k; j++)
                if (p.gb.key_reg[j] < 0 && !preload.count("P.gb.key_mem[" + std::to_string(j) + "]")) return true;
            return p.gb.agg_reg < 0 && !preload.count("P.gb.agg_mem");
        }
        default: break;
        }
        const std::string P = "P.ins[" + std::to_string(i) + "]";
        const bool has_b = I.kind == FK_SEL_COL || I.kind == FK_BINOP_CC || I.kind == FK_BINOP_LC ||
                           I.kind == FK_BUILD_KEYVALS;
        const bool has_a = I.kind != FK_BINOP_LC;
        return (has_a && I.va < 0 && !preload.count(P + ".a")) || (has_b && I.vb < 0 && !preload.count(P + ".b"));
    }

    // Emits instruction i as { declarations; load phase; [split]; compute }.
    void instr(int i, const std::string &split)
    {
        const FInstr &I = p.ins[i];
        const std::string P = "P.ins[" + std::to_string(i) + "]";
        const std::string fin = freg(I.fin) + "[u]", fout = freg(I.fout) + "[u]";
        const std::string L = std::to_string((int)I.logic);
        const std::string U = "#pragma unroll\n        for (int u = 0; u < UU; u++) ";
        // the rows a terminal instruction acts on
        const std::string fterm = batch_act ? "(" + fin + " & act[u])" : fin;
        std::ostringstream d, a, c; // declarations, load phase, compute phase
        if (dropped(i)) {
            o << "    { // " << i << " (decided by the shared masks)\n";
            if (I.fin != I.fout) o << U << fout << " = " << fin << ";\n";
            o << (split.empty() ? "" : hoist_code + split) << "    }\n";
            return;
        }
        switch (I.kind) {
        case FK_SEL_LIT:
        case FK_SEL_COL:
        case FK_SEL_RANGE:
        case FK_SEL_IN2: {
            const bool two = I.kind == FK_SEL_COL;
            std::string pred;
            const std::string C = std::to_string((int)I.cmp);
            if (I.kind == FK_SEL_LIT) pred = "pred_lit<R, " + C + ">(x[u], " + P + ".lit)";
            else if (I.kind == FK_SEL_COL) pred = "pred_col<R, " + C + ">(x[u], y[u])";
            else if (I.kind == FK_SEL_RANGE) pred = "pred_range<R>(x[u], " + P + ".lit, " + P + ".lit2)";
            else pred = "pred_in2<R>(x[u], " + P + ".lit, " + P + ".lit2)";
            d << "        int32_t x[UU][R];" << (two ? " int32_t y[UU][R];" : "") << "\n";
            const bool lazy = !cheap(I.va) || (two && !cheap(I.vb));
            if (lazy) {
                d << "        uint32_t need[UU];\n";
                a << U << "{\n            need[u] = sel_need<" << L << ", R>(" << fin << ");\n";
                a << "            if (need[u]) { " << operand(I.va, P + ".a", "x[u]");
                if (two) a << " " << operand(I.vb, P + ".b", "y[u]");
                a << " }\n        }\n";
                c << U << fout << " = combine<" << L << ">(" << fin << ", need[u] ? " << pred << " : 0u);\n";
            } else {
                a << U << "{ " << operand(I.va, P + ".a", "x[u]");
                if (two) a << " " << operand(I.vb, P + ".b", "y[u]");
                a << " }\n";
                c << U << fout << " = combine<" << L << ">(" << fin << ", " << pred << ");\n";
            }
            break;
        }
        case FK_LOGICAL:
            c << U << fout << " = combine<" << L << ">(" << fin << ", " << freg(I.floc) << "[u]);\n";
            break;
        case FK_FILL:
            c << U << "{\n            fill_rows<R>(" << P << ".lit, " << vreg(I.vout) << "[u]);\n";
            if (I.store) c << "            store_rows<R>(" << P << ".store, row0[u], " << vreg(I.vout) << "[u], M);\n";
            c << "        }\n";
            break;
        case FK_BINOP_CC:
        case FK_BINOP_LC:
        case FK_BINOP_CL: {
            d << "        int32_t x[UU][R], y[UU][R];\n";
            a << U << "{\n            if (!" << fin << ") continue;\n";
            a << "            "
              << (I.kind == FK_BINOP_LC ? "fill_rows<R>(" + P + ".lit, x[u]);" : operand(I.va, P + ".a", "x[u]"))
              << "\n";
            a << "            "
              << (I.kind == FK_BINOP_CL ? "fill_rows<R>(" + P + ".lit, y[u]);" : operand(I.vb, P + ".b", "y[u]"))
              << "\n        }\n";
            c << U << "{\n            if (!" << fin << ") continue;\n";
            c << "            int32_t r[R];\n            binop<R, " << (int)I.op << ">(x[u], y[u], r);\n";
            c << "            set_rows<R>(" << vreg(I.vout) << "[u], r, " << fin << ");\n";
            if (I.store) c << "            store_rows<R>(" << P << ".store, row0[u], r, " << fin << ");\n";
            c << "        }\n";
            break;
        }
        case FK_PROBE_SEMI:
        case FK_PROBE_FULL:
        case FK_PROBE_BITS: {
            if (!hoisted[i]) a << probe_issue(i, fin);
            const std::string x = "x" + std::to_string(i) + "[u]", t = "t" + std::to_string(i) + "[u]";
            const bool vout = I.kind == FK_PROBE_FULL || (I.kind == FK_PROBE_BITS && I.vkind != FV_NONE);
            const std::string vdst = vout ? vreg(I.vout) + "[u]" : "none";
            if (is_decided(i)) {
                c << U << "{\n            const uint32_t o = " << fin << " & act[u];\n";
                c << "            if (o) probe_value_finish<R>(o, " << t << ", " << vdst << ");\n";
                c << "            " << fout << " = o;\n        }\n";
                break;
            }
            c << U << "{\n            uint32_t o = 0;\n";
            if (!vout || late_value(i)) c << "            int32_t none[R];\n";
            c << "            if (" << fin << ") {\n";
            if (I.kind == FK_PROBE_SEMI) c << "                o = probe_semi_finish<R>(" << fin << ", " << t << ");\n";
            else if (I.kind == FK_PROBE_FULL)
                c << "                o = probe_full_finish<R>(" << fin << ", " << t << ", " << vdst << ");\n";
            else if (unset_probe(i))
                c << "                o = probe_unset_finish<R>(" << fin << ", " << t << ", " << P << ".vmin, " << vdst
                  << ");\n";
            else {
                const std::string bits =
                    I.in_lds ? "(const uint32_t *)(lds + " + P + ".lds_off)" : "(const uint32_t *)" + P + ".ht";
                c << "                o = probe_bits_finish<R, " << (late_value(i) ? 0 : (int)I.vkind) << ", "
                  << (I.in_lds ? "true" : "false") << ">(" << x
                  << ", " << fin << ", " << P << ".ht_min, " << P << ".ht_max, " << t << ", " << bits << ", " << P
                  << ".vmin, " << P << ".ht_len, " << (late_value(i) ? std::string("none") : vdst) << ");\n";
                if (late_value(i))
                    c << "                if (o) probe_value_after<R, " << (int)I.vkind << ">(" << x << ", o, " << P
                      << ".vals, " << P << ".ht_min, " << P << ".vmin, " << vdst << ");\n";
            }
            if (vout && I.store)
                c << "                if (o) store_rows<R>(" << P << ".store, row0[u], " << vdst << ", o);\n";
            c << "            }\n            " << fout << " = o;\n        }\n";
            break;
        }
        case FK_BUILD_KEYS:
            d << "        int32_t x[UU][R];\n";
            a << U << "{ " << operand(I.va, P + ".a", "x[u]") << " }\n";
            c << U << "build_keys<R>(x[u], " << fin << ", (uint8_t *)" << P << ".ht, " << P << ".ht_min, " << P
              << ".ht_len);\n";
            break;
        case FK_BUILD_KEYVALS:
            d << "        int32_t x[UU][R], y[UU][R];\n";
            a << U << "{\n            " << operand(I.va, P + ".a", "x[u]") << "\n";
            a << "            if (" << fin << ") { " << operand(I.vb, P + ".b", "y[u]")
              << " } else fill_rows<R>(0, y[u]);\n        }\n";
            c << U << "build_keyvals<R>(x[u], y[u], " << fin << ", (int32_t *)" << P << ".ht, " << P << ".ht_min, "
              << P << ".ht_len);\n";
            break;
        case FK_SUM:
            d << "        int32_t x[UU][R];\n";
            a << U << "if (" << fterm << ") { " << operand(I.va, P + ".a", "x[u]") << " }\n";
            c << U << "if (" << fterm << ") sum_rows<R>(x[u], " << fterm << ", acc);\n";
            break;
        case FK_GROUPBY: {
            // operands of every chunk first; then one loop over the flagged
            // rows of all chunks (bit u * R + k), so the row update is
            // compiled once
            const FGroup &g = p.gb;
            d << "        typename Lanes<UU * R>::type ax = {}";
            for (int j = 0; j < g.k; j++) d << ", kx" << j << " = {}";
            d << ";\n        uint32_t rows = 0;\n";
            a << U << "{\n            rows |= " << fterm << " << (u * R);\n            if (!" << fterm << ") continue;\n";
            a << "            int32_t t[R];\n";
            for (int j = 0; j < g.k; j++)
                a << "            " << operand(g.key_reg[j], "P.gb.key_mem[" + std::to_string(j) + "]", "t")
                  << "\n#pragma unroll\n            for (int k = 0; k < R; k++) kx" << j << "[u * R + k] = t[k];\n";
            a << "            " << operand(g.agg_reg, "P.gb.agg_mem", "t")
              << "\n#pragma unroll\n            for (int k = 0; k < R; k++) ax[u * R + k] = t[k];\n        }\n";
            c << "#pragma unroll 1\n        while (rows) {\n            const int b = __builtin_ctz(rows);\n";
            c << "            rows &= rows - 1;\n            const int32_t kv[F_MAX_KEYS] = {";
            for (int j = 0; j < F_MAX_KEYS; j++) c << (j ? ", " : "") << (j < g.k ? "kx" + std::to_string(j) + "[b]" : "0");
            c << "};\n";
            c << "            fdev::gb_row_mode<" << g.lds << ">(P.gb, lds, kv, ax[b]);\n        }\n";
            break;
        }
        default: fail(SDB_ERR_RUNTIME, "jit: unknown instruction kind");
        }
        o << "    { // " << i << "\n" << d.str() << a.str() << (split.empty() ? "" : hoist_code + split) << c.str()
          << "    }\n";
        auto early = after_finish.find(i);
        if (early != after_finish.end()) o << early->second;
    }

    // Row-at-a-time tail.  The instructions after the last probe through a
    // table in global memory see only the few rows that passed every probe
    // before them (the star joins: under 1 %), yet in chunk form their code
    // runs for every row slot some lane of the wave still holds -- nearly all
    // of them.  They run instead in one loop over each lane's surviving rows
    // (as the group-by update already does), on the R = 1 forms of the same
    // templates: the per-row result is the chunk form's.  Taken only when the
    // tail is a conjunctive chain on one flag register, stores nothing, and
    // reads memory only through early reads.
    int tail_start = -1;

    static bool tail_kind(const FInstr &I)
    {
        switch (I.kind) {
        case FK_SEL_LIT:
        case FK_SEL_COL:
        case FK_SEL_RANGE:
        case FK_SEL_IN2: return I.logic == SDB_AND;
        case FK_PROBE_SEMI:
        case FK_PROBE_FULL:
        case FK_PROBE_BITS:
        case FK_BINOP_CC:
        case FK_BINOP_LC:
        case FK_BINOP_CL:
        case FK_SUM:
        case FK_GROUPBY: return true;
        default: return false;
        }
    }

    void plan_tail()
    {
        tail_start = -1;
        static const bool on = [] {
            const char *e = getenv("SDB_JIT_ROW_TAIL");
            return !(e && *e == '0');
        }();
        if (!on) return;
        int T;
        if (batch_act) {
            // a batched body sees only its `act` rows, a few per cent of the
            // tile: the whole body runs row at a time when it starts from the
            // all-true flags (which become `act`)
            T = 0;
            if (p.ninstr < 1 || p.flag_init[p.ins[0].fin] != FF_TRUE) return;
        } else {
            int last_global = -1;
            for (int i = 0; i < p.ninstr; i++)
                if (global_probe(i)) last_global = i;
            T = last_global + 1;
            // a terminal alone already runs row at a time (group-by) or is cheap
            if (last_global < 0 || T >= p.ninstr - 1) return;
        }
        for (int j = 0; j < F_MAX_FLAGS; j++)
            if (p.flag_store[j]) return;
        // one chain: every instruction reads the flags the previous one wrote
        // (a conjunctive instruction may write them to another register)
        int cur = p.ins[T].fin;
        for (int i = T; i < p.ninstr; i++) {
            const FInstr &I = p.ins[i];
            if (!tail_kind(I) || I.fin != cur || I.store) return;
            if (writes_flags(I.kind)) cur = I.fout;
            const std::string P = "P.ins[" + std::to_string(i) + "]";
            const bool has_b = I.kind == FK_SEL_COL || I.kind == FK_BINOP_CC || I.kind == FK_BINOP_LC;
            const bool has_a = I.kind != FK_BINOP_LC && I.kind != FK_GROUPBY;
            if (has_a && I.va < 0 && !preload.count(P + ".a")) return;
            if (has_b && I.vb < 0 && !preload.count(P + ".b")) return;
        }
        const FInstr &Z = p.ins[p.ninstr - 1];
        if (Z.kind == FK_GROUPBY) {
            for (int g = 0; g < p.gb.k; g++)
                if (p.gb.key_reg[g] < 0 && !preload.count("P.gb.key_mem[" + std::to_string(g) + "]")) return;
            if (p.gb.agg_reg < 0 && !preload.count("P.gb.agg_mem")) return;
        }
        // a terminal, if any, only last
        for (int i = T; i + 1 < p.ninstr; i++)
            if (p.ins[i].kind == FK_SUM || p.ins[i].kind == FK_GROUPBY) return;
        tail_start = T;
    }

    // The tail's row vectors (one element per row slot of the tile) and the
    // per-row expression of an operand.
    std::map<std::string, std::string> tail_vecs; // source statement key -> vector name
    std::vector<std::pair<std::string, std::string>> tail_fill; // vector name, statement filling t[R]

    std::string tail_operand(int reg, const std::string &mem)
    {
        if (reg >= 0 && reg < F_PF_BASE) return "s" + vreg(reg); // the row's value register
        std::string key, fill;
        if (reg >= F_PF_BASE) {
            key = "pf" + std::to_string(reg - F_PF_BASE);
            fill = "vec_rows<R>(pf.c[" + std::to_string(reg - F_PF_BASE) + "][u], t);";
        } else {
            const std::string &arr = preload.at(mem);
            key = arr;
            fill = "copy_rows<R>(" + arr + "[u], t);";
        }
        if (tail_q) {
            // queued rows: a prefetched column's value travels with the row,
            // a lazily read column is read by row index when the queue drains
            const std::string name = "tq_" + key;
            auto &list = reg >= F_PF_BASE ? q_carry : q_gather;
            const std::string src = reg >= F_PF_BASE ? "pf.c[" + std::to_string(reg - F_PF_BASE) + "][u]" : mem;
            bool seen = false;
            for (auto &e : list) seen |= e.first == name;
            if (!seen) list.push_back({name, src});
            return name;
        }
        auto it = tail_vecs.find(key);
        if (it != tail_vecs.end()) return it->second + "[b]";
        const std::string name = "tv_" + key;
        tail_vecs[key] = name;
        tail_fill.push_back({name, fill});
        return name + "[b]";
    }

    // The per-row code of instructions T.. (one row in scope; `continue`
    // drops it).
    std::string tail_loop(int T)
    {
        std::ostringstream loop;
        for (int i = T; i < p.ninstr; i++) {
            const FInstr &I = p.ins[i];
            const std::string P = "P.ins[" + std::to_string(i) + "]";
            const std::string L = std::to_string((int)I.logic), C = std::to_string((int)I.cmp);
            loop << "            { // " << i << "\n";
            switch (I.kind) {
            case FK_SEL_LIT:
                loop << "                const int32_t x[1] = {" << tail_operand(I.va, P + ".a") << "};\n"
                     << "                if (!pred_lit<1, " << C << ">(x, " << P << ".lit)) continue;\n";
                break;
            case FK_SEL_RANGE:
                loop << "                const int32_t x[1] = {" << tail_operand(I.va, P + ".a") << "};\n"
                     << "                if (!pred_range<1>(x, " << P << ".lit, " << P << ".lit2)) continue;\n";
                break;
            case FK_SEL_IN2:
                loop << "                const int32_t x[1] = {" << tail_operand(I.va, P + ".a") << "};\n"
                     << "                if (!pred_in2<1>(x, " << P << ".lit, " << P << ".lit2)) continue;\n";
                break;
            case FK_SEL_COL:
                loop << "                const int32_t x[1] = {" << tail_operand(I.va, P + ".a") << "}, y[1] = {"
                     << tail_operand(I.vb, P + ".b") << "};\n"
                     << "                if (!pred_col<1, " << C << ">(x, y)) continue;\n";
                break;
            case FK_BINOP_CC:
            case FK_BINOP_LC:
            case FK_BINOP_CL: {
                const std::string a = I.kind == FK_BINOP_LC ? P + ".lit" : tail_operand(I.va, P + ".a");
                const std::string b = I.kind == FK_BINOP_CL ? P + ".lit" : tail_operand(I.vb, P + ".b");
                loop << "                const int32_t x[1] = {" << a << "}, y[1] = {" << b << "};\n"
                     << "                int32_t r[1];\n                binop<1, " << (int)I.op << ">(x, y, r);\n"
                     << "                s" << vreg(I.vout) << " = r[0];\n";
                break;
            }
            case FK_PROBE_SEMI:
            case FK_PROBE_FULL:
            case FK_PROBE_BITS: {
                const bool vout = I.kind == FK_PROBE_FULL || (I.kind == FK_PROBE_BITS && I.vkind != FV_NONE);
                if (dropped(i)) {
                    loop << "                // decided by the shared masks\n";
                    break;
                }
                if (is_decided(i)) {
                    // the masks established the match: the value alone
                    const int vk = I.kind == FK_PROBE_FULL ? (int)FV_NONE : (int)I.vkind;
                    const std::string vals = I.kind == FK_PROBE_FULL ? P + ".ht" : P + ".vals";
                    loop << "                const int32_t x[1] = {" << tail_operand(I.va, P + ".a") << "};\n"
                         << "                ProbeTmp<1> t;\n                probe_value_issue<1, " << vk << ">(x, 1u, "
                         << vals << ", " << P << ".ht_min, " << P << ".ht_max, " << P << ".vmin, t);\n"
                         << "                s" << vreg(I.vout) << " = t.val[0];\n";
                    break;
                }
                loop << "                const int32_t x[1] = {" << tail_operand(I.va, P + ".a") << "};\n"
                     << "                ProbeTmp<1> t;\n                int32_t v[1] = {"
                     << (vout ? "s" + vreg(I.vout) : std::string("0")) << "};\n";
                if (I.kind == FK_PROBE_SEMI)
                    loop << "                probe_semi_issue<1>(x, 1u, (const uint8_t *)" << P << ".ht, " << P << ".ht_min, "
                         << P << ".ht_max, t);\n                if (!probe_semi_finish<1>(1u, t)) continue;\n";
                else if (I.kind == FK_PROBE_FULL)
                    loop << "                probe_full_issue<1>(x, 1u, (const int2 *)" << P << ".ht, " << P << ".ht_min, "
                         << P << ".ht_max, t);\n                if (!probe_full_finish<1>(1u, t, v)) continue;\n";
                else if (unset_probe(i))
                    loop << "                probe_unset_issue<1>(x, 1u, (const uint16_t *)" << P << ".vals, " << P
                         << ".ht_min, " << P << ".ht_max, t);\n                if (!probe_unset_finish<1>(1u, t, " << P
                         << ".vmin, v)) continue;\n";
                else {
                    const std::string bits =
                        I.in_lds ? "(const uint32_t *)(lds + " + P + ".lds_off)" : "(const uint32_t *)" + P + ".ht";
                    const int vk = late_value(i) ? 0 : (int)I.vkind;
                    const std::string inl = I.in_lds ? "true" : "false";
                    loop << "                probe_bits_issue<1, " << vk << ", " << inl << ">(x, 1u, " << bits << ", "
                         << P << ".ht_min, " << P << ".ht_max, " << P << ".vals, " << P << ".vmin, " << P
                         << ".ht_len, t);\n"
                         << "                if (!probe_bits_finish<1, " << vk << ", " << inl << ">(x, 1u, " << P
                         << ".ht_min, " << P << ".ht_max, t, " << bits << ", " << P << ".vmin, " << P
                         << ".ht_len, v)) continue;\n";
                    if (late_value(i))
                        loop << "                probe_value_after<1, " << (int)I.vkind << ">(x, 1u, " << P << ".vals, "
                             << P << ".ht_min, " << P << ".vmin, v);\n";
                }
                if (vout) loop << "                s" << vreg(I.vout) << " = v[0];\n";
                break;
            }
            case FK_SUM:
                loop << "                acc += " << tail_operand(I.va, P + ".a") << ";\n";
                break;
            case FK_GROUPBY: {
                const FGroup &g = p.gb;
                loop << "                const int32_t kv[F_MAX_KEYS] = {";
                for (int j = 0; j < F_MAX_KEYS; j++)
                    loop << (j ? ", " : "")
                         << (j < g.k ? tail_operand(g.key_reg[j], "P.gb.key_mem[" + std::to_string(j) + "]") : "0");
                loop << "};\n                fdev::gb_row_mode<" << g.lds << ">(P.gb, lds, kv, "
                     << tail_operand(g.agg_reg, "P.gb.agg_mem") << ");\n";
                break;
            }
            default: fail(SDB_ERR_RUNTIME, "jit: instruction kind outside the row tail");
            }
            loop << "            }\n";
        }
        return loop.str();
    }

    // queued form: values carried with the row (name, chunk-register source)
    // and lazily read columns read at drain time (name, column pointer)
    bool tail_q = false;
    std::vector<std::pair<std::string, std::string>> q_carry, q_gather;

    void tail_code()
    {
        const int T = tail_start, F = p.ins[T].fin;
        tail_vecs.clear();
        tail_fill.clear();
        const std::string body = tail_loop(T);
        // value registers: the row's value, from the chunk registers
        std::vector<bool> reads(p.nvals, false);
        for (int r = 0; r < p.nvals; r++) reads[r] = body.find("s" + vreg(r)) != std::string::npos;
        if (queue) o << "    if constexpr (Q) {\n" << tail_queued(T, reads) << "    } else\n";
        o << "    { // instructions " << T << ".. row at a time over the rows that reach them\n";
        o << "        typename Lanes<UU * R>::type";
        bool first = true;
        for (auto &f : tail_fill) {
            o << (first ? " " : ", ") << f.first << " = {}";
            first = false;
        }
        for (int r = 0; r < p.nvals; r++)
            if (reads[r]) {
                o << (first ? " " : ", ") << "tv_" << vreg(r) << " = {}";
                first = false;
            }
        if (first) o << " unused_ = {}";
        o << ";\n        uint32_t rows = 0;\n";
        o << "#pragma unroll\n        for (int u = 0; u < UU; u++) {\n";
        o << "            rows |= " << freg(F) << "[u] << (u * R);\n            if (!" << freg(F) << "[u]) continue;\n";
        o << "            int32_t t[R];\n";
        for (auto &f : tail_fill)
            o << "            " << f.second << "\n#pragma unroll\n            for (int k = 0; k < R; k++) " << f.first
              << "[u * R + k] = t[k];\n";
        for (int r = 0; r < p.nvals; r++)
            if (reads[r])
                o << "#pragma unroll\n            for (int k = 0; k < R; k++) tv_" << vreg(r) << "[u * R + k] = "
                  << vreg(r) << "[u][k];\n";
        o << "        }\n#pragma unroll 1\n        while (rows) {\n";
        o << "            const int b = __builtin_ctz(rows);\n            rows &= rows - 1;\n";
        for (int r = 0; r < p.nvals; r++)
            if (reads[r]) o << "            int32_t s" << vreg(r) << " = tv_" << vreg(r) << "[b];\n";
        o << body << "        }\n    }\n";
    }

    // Queued tail.  The rows that reach instruction T are appended to a
    // queue of 64 entries per wave in LDS (row index, then the values the
    // tail reads from registers); when the next rows would overflow it, and
    // once after the last tile, the wave drains it: lane l takes entry l,
    // reads the lazily read columns at its row and runs the tail for it.
    // The tile then never waits for those reads; a drain waits once for up
    // to 64 rows.  Field f of entry e sits at int32 index 64 * f + e.
    bool queue_ok = false; // the caller allows a queue
    bool queue = false;    // this kernel has one
    int queue_fields = 0;
    std::string queue_fn;  // queue_base / queue_flush, emitted before the body

    std::string tail_queued(int T, const std::vector<bool> &reads)
    {
        q_carry.clear();
        q_gather.clear();
        tail_q = true;
        const std::string body = tail_loop(T);
        tail_q = false;
        const int F = p.ins[T].fin;
        // fields: 1 row, then the value registers read, then carried columns
        std::vector<std::pair<std::string, std::string>> fields; // vector name, chunk source
        std::ostringstream load;                                  // drain-time declarations
        int f = 1;
        for (int r = 0; r < p.nvals; r++)
            if (reads[r]) {
                fields.push_back({"tq_vec_" + vreg(r), vreg(r) + "[u][k]"});
                load << "        int32_t s" << vreg(r) << " = qb[" << 64 * f++ << " + lane];\n";
            }
        for (auto &c : q_carry) {
            fields.push_back({"tq_vec_" + c.first.substr(3), "t[k]"});
            load << "        const int32_t " << c.first << " = qb[" << 64 * f++ << " + lane];\n";
        }
        for (auto &g : q_gather) load << "        const int32_t " << g.first << " = " << g.second << "[row];\n";
        queue_fields = f;

        std::ostringstream fn;
        fn << "__device__ __forceinline__ void wave_sync()\n{\n"
           << "    __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\");\n"
           << "    __builtin_amdgcn_wave_barrier();\n"
           << "    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");\n}\n";
        fn << "__device__ __forceinline__ int32_t *queue_base(const FProgram &P, uint8_t *lds)\n{\n"
           << "    return reinterpret_cast<int32_t *>(lds + ((P.vreg_off + 15u) & ~15u)) + (threadIdx.x >> 6) * (64 * QF);\n}\n"
           << "__device__ __forceinline__ void queue_flush(const FProgram &P, uint8_t *lds, int64_t &acc, uint32_t &qn)\n{\n"
           << "    int32_t *const qb = queue_base(P, lds);\n    wave_sync();\n"
           << "    const uint32_t lane = threadIdx.x & 63u;\n    if (lane < qn) {\n"
           << "        const int64_t row = (int64_t)(uint32_t)qb[lane];\n"
           << load.str() << "        do {\n" << body << "        } while (0);\n    }\n"
           << "    wave_sync();\n    qn = 0;\n}\n";
        queue_fn = fn.str();

        std::ostringstream a;
        a << "    { // instructions " << T << ".. queued for the rows that reach them\n";
        a << "        int32_t *const qb = queue_base(P, lds);\n";
        if (!fields.empty()) {
            a << "        typename Lanes<UU * R>::type";
            for (size_t i = 0; i < fields.size(); i++) a << (i ? ", " : " ") << fields[i].first << " = {}";
            a << ";\n";
        }
        a << "        uint32_t rows = 0;\n#pragma unroll\n        for (int u = 0; u < UU; u++) {\n";
        a << "            rows |= " << freg(F) << "[u] << (u * R);\n";
        if (!fields.empty()) {
            a << "            if (!" << freg(F) << "[u]) continue;\n            int32_t t[R];\n";
            for (size_t i = 0; i < fields.size(); i++) {
                const bool carried = i >= fields.size() - q_carry.size();
                if (carried)
                    a << "            vec_rows<R>(" << q_carry[i - (fields.size() - q_carry.size())].second << ", t);\n";
                a << "#pragma unroll\n            for (int k = 0; k < R; k++) " << fields[i].first << "[u * R + k] = "
                  << (carried ? std::string("t[k]") : fields[i].second) << ";\n";
            }
        }
        a << "        }\n#pragma unroll 1\n        while (true) {\n"
          << "            const bool has = rows != 0;\n"
          << "            const uint64_t m = __ballot(has);\n"
          << "            if (m == 0) break;\n"
          << "            const uint32_t cnt = (uint32_t)__popcll(m);\n"
          << "            if (qn + cnt > 64u) queue_flush(P, lds, acc, qn);\n"
          << "            if (has) {\n"
          << "                const int b = __builtin_ctz(rows);\n                rows &= rows - 1;\n"
          << "                const uint32_t pos = qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), "
             "__builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));\n"
          << "                int64_t rb = row0[0];\n#pragma unroll\n"
          << "                for (int u = 1; u < UU; u++) rb = b / R == u ? row0[u] : rb;\n"
          << "                qb[pos] = (int32_t)(rb + b % R);\n";
        for (size_t i = 0; i < fields.size(); i++)
            a << "                qb[" << 64 * (i + 1) << " + pos] = " << fields[i].first << "[b];\n";
        a << "            }\n            qn += cnt;\n        }\n    }\n";
        return a.str();
    }

    std::string body()
    {
        // registers the program touches
        std::vector<bool> fused(F_MAX_FLAGS, false);
        for (int j = 0; j < F_MAX_FLAGS; j++) fused[j] = p.flag_init[j] != FF_UNDEF || p.flag_store[j];
        for (int i = 0; i < p.ninstr; i++) {
            const FInstr &I = p.ins[i];
            fused[I.fin] = true;
            if (I.kind == FK_LOGICAL) fused[I.floc] = true;
            if (I.kind == FK_SEL_LIT || I.kind == FK_SEL_COL || I.kind == FK_SEL_RANGE || I.kind == FK_SEL_IN2 ||
                I.kind == FK_LOGICAL || I.kind == FK_PROBE_SEMI || I.kind == FK_PROBE_FULL || I.kind == FK_PROBE_BITS)
                fused[I.fout] = true;
        }
        o << "    constexpr uint32_t M = (1u << R) - 1;\n";
        for (int j = 0; j < F_MAX_FLAGS; j++) {
            if (!fused[j]) continue;
            o << "    uint32_t " << freg(j) << "[UU];\n";
            o << "#pragma unroll\n    for (int u = 0; u < UU; u++) " << freg(j) << "[u] = ";
            if (p.flag_init[j] == FF_TRUE) o << (batch_act ? "act[u];\n" : "M;\n");
            else if (p.flag_init[j] == FF_LOAD && p.flag_load[j])
                o << "fdev::flags_in<R>(P.flag_load[" << j << "] + row0[u])" << (batch_act ? " & act[u]" : "") << ";\n";
            else o << "0u;\n";
        }
        for (int r = 0; r < p.nvals; r++) o << "    int32_t " << vreg(r) << "[UU][R] = {};\n";
        plan_preloads();
        o << preload_decls;
        plan_tail();
        queue = queue_ok && !batch_act && prefetch_in_body && p.vec && tail_start >= 0;
        const int dense_end = tail_start >= 0 ? tail_start : p.ninstr;
        for (int i = 0; i < dense_end; i++)
            if (is_probe(p.ins[i].kind) && !dropped(i))
                o << "    int32_t x" << i << "[UU][R];\n    ProbeTmp<R> t" << i << "[UU];\n";
        // the next tile's column loads go out after the loads of the first
        // instruction that waits for global memory and before that wait, so
        // the in-order completion counter does not make it wait for them too
        const std::string prefetch =
            prefetch_in_body ? "        if (has_next) pf_load<R, UU>(P, nrow0, next);\n" : std::string();
        int first_wait = -1;
        for (int i = 0; i < dense_end && first_wait < 0; i++)
            if (waits(i)) first_wait = i;
        // early issue of chained probes trades extra table reads for one
        // fewer wait; measured slower on the star joins of a single
        // program, so opt-in there (SDB_JIT_HOIST=1); a batched body, which
        // runs between the other programs' bodies, always issues early
        static const bool hoist = [] {
            const char *e = getenv("SDB_JIT_HOIST");
            return e && *e == '1';
        }();
        plan_hoisting((hoist && prefetch_in_body) || batch_act ? first_wait : -1);
        if (first_wait < 0) o << prefetch;
        const std::string at_wait =
            prefetch_in_body ? prefetch : (hoist_code.empty() ? std::string() : "        // issued early above\n");
        if (at_wait.empty()) first_wait = -1;
        for (int i = 0; i < dense_end; i++) instr(i, i == first_wait ? at_wait : std::string());
        if (tail_start >= 0) tail_code();
        for (int j = 0; j < F_MAX_FLAGS; j++)
            if (p.flag_store[j])
                o << "#pragma unroll\n    for (int u = 0; u < UU; u++) fdev::flags_out<R>(P.flag_store[" << j
                  << "] + row0[u], " << freg(j) << "[u]);\n";
        return o.str();
    }
};

const char *KERNEL_TEMPLATE_INCLUDES = R"(#include "fused_ops.hpp"
using namespace sdb;
using namespace sdb::fops;
namespace {
)";

const char *KERNEL_TEMPLATE_HEAD = R"(template <int R, int UU> struct Pf {
    typename Vec<R>::type c[NPF > 0 ? NPF : 1][UU];
};
template <int R, int UU>
__device__ __forceinline__ void pf_load(const FProgram &P, const int64_t (&row0)[UU], Pf<R, UU> &pf)
{
#pragma unroll
    for (int j = 0; j < NPF; j++)
#pragma unroll
        for (int u = 0; u < UU; u++) pf.c[j][u] = load_vec<R>(P.pf_mem[j], row0[u]);
}
template <int R, int UU, bool Q>
__device__ __forceinline__ void body(const FProgram &P, const int64_t (&row0)[UU], const Pf<R, UU> &pf, int64_t &acc,
                                     uint8_t *lds, bool has_next, const int64_t (&nrow0)[UU], Pf<R, UU> &next,
                                     uint32_t &qn)
{
)";

const char *KERNEL_QUEUE_STUB = R"(__device__ __forceinline__ void queue_flush(const FProgram &, uint8_t *, int64_t &, uint32_t &) {}
)";

const char *KERNEL_TEMPLATE_TAIL = R"(}
} // namespace

extern "C" __global__ void __launch_bounds__(BLOCK) sdb_fused_program(const FProgram P_arg, int64_t n)
{
    // the program is read in place from the kernel-argument segment
    const FProgram &P = *(const FProgram __attribute__((address_space(4))) *)__builtin_amdgcn_kernarg_segment_ptr();
    if (P.magic != F_MAGIC) return;
    extern __shared__ uint4 dyn_lds[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(dyn_lds);
    fdev::block_prologue<BLOCK>(P, lds);
    int64_t acc = 0;
    uint32_t qn = 0; // rows in this wave's queue (queued tail)
    if constexpr (VEC) {
    // tiles of U * BLOCK 4-row chunks; chunk u of a thread in tile t starts
    // at row 4 * (t * U * BLOCK + u * BLOCK + thread); the prefetched columns
    // of the next tile load while this one runs
    const int64_t nch = n >> 2, tile = (int64_t)U * BLOCK, ntiles = nch / tile;
    auto rows_of = [&](int64_t t, int64_t (&r)[U]) {
#pragma unroll
        for (int u = 0; u < U; u++) r[u] = (t * tile + u * BLOCK + threadIdx.x) << 2;
    };
    int64_t t = blockIdx.x;
    Pf<4, U> next;
    if (t < ntiles) {
        int64_t r[U];
        rows_of(t, r);
        pf_load<4, U>(P, r, next);
    }
    while (t < ntiles) {
        const Pf<4, U> cur = next;
        int64_t r[U], rn[U];
        rows_of(t, r);
        const int64_t tn = t + gridDim.x;
        rows_of(tn, rn);
        body<4, U, QUEUE>(P, r, cur, acc, lds, tn < ntiles, rn, next, qn);
        t = tn;
    }
    // every lane of the wave is here: drain its queue
    if constexpr (QUEUE) {
        if (qn) queue_flush(P, lds, acc, qn);
    }
    for (int64_t c = ntiles * tile + (int64_t)blockIdx.x * BLOCK + threadIdx.x; c < nch; c += (int64_t)gridDim.x * BLOCK) {
        const int64_t r[1] = {c << 2};
        Pf<4, 1> pf, unused;
        pf_load<4, 1>(P, r, pf);
        body<4, 1, false>(P, r, pf, acc, lds, false, r, unused, qn);
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x < (n & 3)) {
        const int64_t r[1] = {(nch << 2) + threadIdx.x};
        Pf<1, 1> pf, unused;
        pf_load<1, 1>(P, r, pf);
        body<1, 1, false>(P, r, pf, acc, lds, false, r, unused, qn);
    }
    } else {
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    Pf<1, 1> next;
    if (i < n) {
        const int64_t r[1] = {i};
        pf_load<1, 1>(P, r, next);
    }
    while (i < n) {
        const Pf<1, 1> cur = next;
        const int64_t r[1] = {i};
        const int64_t in = i + stride;
        const int64_t rn[1] = {in};
        body<1, 1, false>(P, r, cur, acc, lds, in < n, rn, next, qn);
        i = in;
    }
    }
    fdev::block_epilogue<BLOCK>(P, lds, acc);
}
)";

std::string device_arch()
{
    static std::mutex mu;
    static std::map<int, std::string> cache;
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    hipDeviceProp_t prop;
    SDB_HIP(hipGetDeviceProperties(&prop, dev));
    std::string a(prop.gcnArchName);
    const size_t colon = a.find(':');
    if (colon != std::string::npos) a.resize(colon);
    if (a.empty()) a = "gfx950";
    cache[dev] = a;
    return a;
}

// SDB_JIT_DUMP_DIR=<dir> writes every source compiled to <dir>/sdb_fused_<n>.hip.
void dump_source(const std::string &src)
{
    static const char *dir = getenv("SDB_JIT_DUMP_DIR");
    static std::atomic<int> seq{0};
    if (!dir || !*dir) return;
    const std::string path = std::string(dir) + "/sdb_fused_" + std::to_string(seq++) + ".hip";
    if (FILE *f = fopen(path.c_str(), "w")) {
        fwrite(src.data(), 1, src.size(), f);
        fclose(f);
    }
}

std::vector<char> compile_source(const std::string &src, const std::string &arch)
{
    Rtc &r = rtc();
    dump_source(src);
    if (!r.ok) fail(SDB_ERR_UNSUPPORTED, "runtime compilation unavailable: " + r.why);
    std::vector<const char *> names, texts;
    for (const JitHeader &h : JIT_HEADERS) {
        names.push_back(h.name);
        texts.push_back(h.text);
    }
    hiprtcProgram prog = nullptr;
    if (r.create(&prog, src.c_str(), "sdb_fused_program.hip", (int)names.size(), texts.data(), names.data()) !=
        HIPRTC_SUCCESS)
        fail(SDB_ERR_RUNTIME, "hiprtcCreateProgram failed");
    const std::string arch_opt = "--offload-arch=" + arch;
    // the same code-generation flags as the library's own kernels
    const char *opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    const hiprtcResult cr = r.compile(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (cr != HIPRTC_SUCCESS) {
        size_t n = 0;
        r.log_size(prog, &n);
        std::string log(n, '\0');
        if (n) r.log(prog, &log[0]);
        r.destroy(&prog);
        fail(SDB_ERR_RUNTIME, "fused kernel compilation failed:\n" + log);
    }
    size_t n = 0;
    r.code_size(prog, &n);
    std::vector<char> code(n);
    r.code(prog, code.data());
    r.destroy(&prog);
    return code;
}

// Programs that together use every instruction form, operand kind, value
// table kind and both chunk widths (pointers are left null: they are only
// compiled, never launched).
std::vector<std::pair<FProgram, int>> coverage_programs()
{
    std::vector<std::pair<FProgram, int>> out;
    auto base = [](int vec, int npf) {
        FProgram p{};
        p.magic = F_MAGIC;
        p.vec = vec;
        p.npf = npf;
        return p;
    };
    auto ins = [](FProgram &p, FInstr I) {
        if (I.vout == 0 && I.kind != FK_FILL && I.kind != FK_BINOP_CC && I.kind != FK_BINOP_LC &&
            I.kind != FK_BINOP_CL && I.kind != FK_PROBE_FULL && I.kind != FK_PROBE_BITS)
            I.vout = -1;
        p.ins[p.ninstr++] = I;
    };
    auto mk = [](uint8_t kind, uint8_t fin, uint8_t fout, int8_t va, int8_t vb = -1, int8_t vout = -1) {
        FInstr I{};
        I.kind = kind;
        I.fin = fin;
        I.fout = fout;
        I.va = va;
        I.vb = vb;
        I.vout = vout;
        return I;
    };
    static uint8_t flag_dummy[4];
    {   // selections, logical, sum; flags loaded and stored
        FProgram p = base(1, 2);
        p.flag_init[0] = FF_TRUE;
        p.flag_init[1] = FF_LOAD;
        p.flag_load[1] = flag_dummy;
        p.flag_store[2] = flag_dummy;
        FInstr a = mk(FK_SEL_LIT, 0, 2, F_PF_BASE);
        a.cmp = SDB_LT, a.logic = SDB_AND;
        ins(p, a);
        FInstr b = mk(FK_SEL_RANGE, 2, 2, -1);
        b.logic = SDB_AND;
        ins(p, b);
        FInstr c = mk(FK_SEL_IN2, 2, 3, F_PF_BASE + 1);
        c.logic = SDB_OR;
        ins(p, c);
        FInstr d = mk(FK_SEL_COL, 3, 3, F_PF_BASE, -1);
        d.cmp = SDB_NE, d.logic = SDB_NONE;
        ins(p, d);
        FInstr e = mk(FK_LOGICAL, 1, 2, -1);
        e.floc = 3, e.logic = SDB_AND;
        ins(p, e);
        ins(p, mk(FK_SUM, 2, 0, -1));
        p.sum_acc = reinterpret_cast<uint64_t *>(flag_dummy);
        out.emplace_back(p, 2);
    }
    {   // projection: fill and the three arithmetic forms, stored
        FProgram p = base(1, 1);
        p.nvals = 4;
        p.flag_init[0] = FF_TRUE;
        FInstr f = mk(FK_FILL, 0, 0, -1, -1, 0);
        f.store = reinterpret_cast<int32_t *>(flag_dummy);
        ins(p, f);
        FInstr m = mk(FK_BINOP_CC, 0, 0, 0, -1, 1);
        m.op = SDB_MUL;
        m.store = reinterpret_cast<int32_t *>(flag_dummy);
        ins(p, m);
        FInstr l = mk(FK_BINOP_LC, 0, 0, -1, F_PF_BASE, 2);
        l.op = SDB_SUB;
        ins(p, l);
        FInstr r = mk(FK_BINOP_CL, 0, 0, 1, -1, 3);
        r.op = SDB_DIV;
        ins(p, r);
        ins(p, mk(FK_SUM, 0, 0, 3));
        p.sum_acc = reinterpret_cast<uint64_t *>(flag_dummy);
        out.emplace_back(p, 4);
    }
    {   // probes of every kind and a three-key group-by, row at a time
        FProgram p = base(0, 2);
        p.nvals = 3;
        p.flag_init[0] = FF_TRUE;
        ins(p, mk(FK_PROBE_SEMI, 0, 1, -1));
        FInstr full = mk(FK_PROBE_FULL, 1, 1, F_PF_BASE, -1, 0);
        full.store = reinterpret_cast<int32_t *>(flag_dummy);
        ins(p, full);
        FInstr b0 = mk(FK_PROBE_BITS, 1, 1, -1);
        b0.in_lds = 1;
        b0.vkind = FV_NONE;
        ins(p, b0);
        FInstr b1 = mk(FK_PROBE_BITS, 1, 1, F_PF_BASE + 1, -1, 1);
        b1.vkind = FV_U16;
        ins(p, b1);
        FInstr b2 = mk(FK_PROBE_BITS, 1, 1, 0, -1, 2);
        b2.vkind = FV_PAIR;
        b2.store = reinterpret_cast<int32_t *>(flag_dummy);
        ins(p, b2);
        ins(p, mk(FK_GROUPBY, 1, 0, -1));
        p.gb.k = 3;
        p.gb.key_reg[0] = 1, p.gb.key_reg[1] = 2, p.gb.key_reg[2] = -1;
        p.gb.agg_reg = F_PF_BASE + 1;
        out.emplace_back(p, 1);
    }
    {   // builds
        FProgram p = base(1, 1);
        p.flag_init[0] = FF_LOAD;
        p.flag_load[0] = flag_dummy;
        ins(p, mk(FK_BUILD_KEYS, 0, 0, -1));
        ins(p, mk(FK_BUILD_KEYVALS, 0, 0, F_PF_BASE, -1));
        out.emplace_back(p, 1);
    }
    {   // star join: LDS-bitmap semi probe, global and LDS probes with 16-bit
        // value tables, two-key group-by hashed in LDS
        FProgram p = base(1, 4);
        p.nvals = 2;
        p.flag_init[0] = FF_TRUE;
        FInstr a = mk(FK_PROBE_BITS, 0, 1, F_PF_BASE);
        a.in_lds = 1;
        ins(p, a);
        FInstr b = mk(FK_PROBE_BITS, 1, 1, F_PF_BASE + 1, -1, 0);
        b.vkind = FV_U16;
        ins(p, b);
        FInstr c = mk(FK_PROBE_BITS, 1, 1, F_PF_BASE + 2, -1, 1);
        c.in_lds = 1;
        c.vkind = FV_RANK16;
        ins(p, c);
        ins(p, mk(FK_GROUPBY, 1, 0, -1));
        p.gb.k = 2;
        p.gb.lds = FG_HASH_LDS;
        p.gb.key_reg[0] = 1, p.gb.key_reg[1] = 0;
        p.gb.agg_reg = F_PF_BASE + 3;
        out.emplace_back(p, 2);
    }
    {   // star join whose later columns are read from memory: the date key
        // and the aggregate are read early, after the global probe
        FProgram p = base(1, 2);
        p.nvals = 2;
        p.flag_init[0] = FF_TRUE;
        FInstr a = mk(FK_PROBE_BITS, 0, 0, F_PF_BASE);
        a.in_lds = 1;
        ins(p, a);
        FInstr b = mk(FK_PROBE_BITS, 0, 0, F_PF_BASE + 1, -1, 0);
        b.vkind = FV_U16;
        ins(p, b);
        FInstr c = mk(FK_PROBE_BITS, 0, 0, -1, -1, 1);
        c.in_lds = 1;
        c.vkind = FV_RANK16;
        ins(p, c);
        ins(p, mk(FK_GROUPBY, 0, 0, -1));
        p.gb.k = 2;
        p.gb.lds = FG_HASH_LDS;
        p.gb.key_reg[0] = 1, p.gb.key_reg[1] = 0;
        p.gb.agg_reg = -1;
        out.emplace_back(p, 2);
    }
    {   // one-key group-by over a prefetched key, memory aggregate
        FProgram p = base(1, 1);
        p.flag_init[0] = FF_TRUE;
        ins(p, mk(FK_GROUPBY, 0, 0, -1));
        p.gb.k = 1;
        p.gb.key_reg[0] = F_PF_BASE;
        p.gb.agg_reg = -1;
        out.emplace_back(p, 4);
    }
    return out;
}

Compiled &get_kernel(const std::string &src, const char *name = "sdb_fused_program")
{
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    const std::string key = device_arch() + "\n" + src;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_loaded.find({dev, key});
    if (it != g_loaded.end()) return it->second;
    auto ct = g_code.find(key);
    if (ct == g_code.end()) {
        ct = g_code.emplace(key, compile_source(src, device_arch())).first;
        g_compiled++;
    }
    Compiled c;
    SDB_HIP(hipModuleLoadData(&c.module, ct->second.data()));
    SDB_HIP(hipModuleGetFunction(&c.fn, c.module, name));
    return g_loaded.emplace(std::make_pair(dev, key), c).first->second;
}

// Compiles every source of `srcs` for `arch`, several at a time on their own
// threads (hipRTC programs are independent); the code objects in order.
std::vector<std::vector<char>> compile_all(const std::vector<std::string> &srcs, const std::string &arch)
{
    std::vector<std::vector<char>> code(srcs.size());
    std::vector<std::string> err(srcs.size());
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i; (i = next.fetch_add(1)) < srcs.size();) {
            try {
                code[i] = compile_source(srcs[i], arch);
            } catch (const std::exception &e) {
                err[i] = e.what();
            }
        }
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nthreads = std::min<size_t>(srcs.size(), std::min(hw, 16u));
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nthreads; t++) pool.emplace_back(work);
    work();
    for (std::thread &t : pool) t.join();
    for (const std::string &e : err)
        if (!e.empty()) fail(SDB_ERR_RUNTIME, e);
    return code;
}

// Compiles for `arch` the sources of `srcs` that have no code object yet, in
// parallel, and records their code objects, so that get_kernel then only
// loads them.  A new program shape's autotuner candidates are compiled this
// way at once, instead of one candidate per launch.  Returns how many were
// compiled.
int precompile(const std::vector<std::string> &srcs, const std::string &arch)
{
    std::vector<std::string> todo;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (const std::string &s : srcs)
            if (!g_code.count(arch + "\n" + s) && std::find(todo.begin(), todo.end(), s) == todo.end())
                todo.push_back(s);
    }
    if (todo.empty()) return 0;
    std::vector<std::vector<char>> code = compile_all(todo, arch);
    std::lock_guard<std::mutex> lk(g_mu);
    int n = 0;
    for (size_t i = 0; i < todo.size(); i++)
        if (g_code.emplace(arch + "\n" + todo[i], std::move(code[i])).second) {
            g_compiled++;
            n++;
        }
    return n;
}

} // namespace

bool jit_available() { return rtc().ok; }

std::string jit_unavailable_reason() { return rtc().ok ? std::string() : rtc().why; }

int jit_check(const std::string &arch)
{
    std::vector<std::string> srcs;
    for (auto &pu : coverage_programs()) {
        // with the queued row tail where the program has one, and without
        const std::string q = jit_source(pu.first, pu.second, 256, true), plain = jit_source(pu.first, pu.second, 256, false);
        srcs.push_back(q);
        if (plain != q) srcs.push_back(plain);
    }
    return (int)compile_all(srcs, arch).size() + jit_batch_check(arch);
}

int jit_kernels_compiled()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_compiled;
}

namespace {

// The fields of a program that decide its generated code: instruction kinds
// and code-selecting parameters, register numbers, which pointers are set,
// the group-by layout.  Everything else (pointers, literals, key ranges, LDS
// offsets) the kernel reads from the program at run time, so programs with
// equal keys share one kernel.  Code generation and kernel lookup by source
// text cost tens of microseconds per launch; the key is a few hundred bytes.
std::string shape_key(const FProgram &p)
{
    std::string k;
    k.reserve(32 + 14 * p.ninstr);
    auto put = [&](int v) { k.push_back((char)v); };
    put(p.ninstr);
    put(p.vec);
    put(p.nvals);
    put(p.npf);
    for (int j = 0; j < F_MAX_FLAGS; j++) {
        put(p.flag_init[j]);
        put(p.flag_load[j] != nullptr);
        put(p.flag_store[j] != nullptr);
    }
    for (int i = 0; i < p.ninstr; i++) {
        const FInstr &I = p.ins[i];
        put(I.kind);
        put(I.cmp);
        put(I.logic);
        put(I.op);
        put(I.fin);
        put(I.fout);
        put(I.floc);
        put(I.in_lds);
        put(I.va);
        put(I.vb);
        put(I.vout);
        put(I.vkind);
        put(I.store != nullptr);
    }
    put(p.gb.k);
    put(p.gb.lds);
    for (int j = 0; j < F_MAX_KEYS; j++) put(p.gb.key_reg[j]);
    put(p.gb.agg_reg);
    return k;
}

std::mutex g_src_mu;
std::unordered_map<std::string, std::string> g_src;           // shape key + launch shape -> source
std::unordered_map<std::string, Compiled *> g_kernel_by_shape; // device + shape key + launch shape

std::unordered_map<std::string, uint32_t> g_queue_bytes; // shape key + launch shape -> queue bytes per wave

std::string launch_shape(int unroll, int block, bool queue)
{
    return std::string(1, (char)unroll) + (char)(block >> 8) + (queue ? "q" : "");
}

// Queued tails (SDB_JIT_QUEUE=0 runs every tail in the tile).
bool queue_enabled()
{
    static const bool on = [] {
        const char *e = getenv("SDB_JIT_QUEUE");
        return !(e && *e == '0');
    }();
    return on;
}

} // namespace

std::string jit_source(const FProgram &p, int unroll, int block, bool queue)
{
    if (unroll != 1 && unroll != 2 && unroll != 4) fail(SDB_ERR_INVALID_ARG, "jit: unroll must be 1, 2 or 4");
    if (block != 256 && block != 512 && block != 1024) fail(SDB_ERR_INVALID_ARG, "jit: block must be 256, 512 or 1024");
    queue = queue && queue_enabled();
    const std::string key = shape_key(p) + launch_shape(unroll, block, queue);
    {
        std::lock_guard<std::mutex> lk(g_src_mu);
        auto it = g_src.find(key);
        if (it != g_src.end()) return it->second;
    }
    Gen g(p);
    g.queue_ok = queue;
    const std::string body = g.body();
    std::ostringstream s;
    s << KERNEL_TEMPLATE_INCLUDES << "constexpr int NPF = " << p.npf << ";\nconstexpr int U = " << unroll
      << ";\nconstexpr bool VEC = " << (p.vec ? "true" : "false") << ";\nconstexpr int BLOCK = " << block << ";\n";
    s << "constexpr bool QUEUE = " << (g.queue ? "true" : "false") << ";\nconstexpr int QF = " << g.queue_fields
      << ";\n";
    s << (g.queue ? g.queue_fn : std::string(KERNEL_QUEUE_STUB));
    s << KERNEL_TEMPLATE_HEAD << body << KERNEL_TEMPLATE_TAIL;
    std::lock_guard<std::mutex> lk(g_src_mu);
    g_queue_bytes[key] = g.queue ? 256u * (uint32_t)g.queue_fields : 0u;
    return g_src.emplace(key, s.str()).first->second;
}

namespace {

// The compiled kernel of a program's shape for a launch shape, and the LDS
// bytes per wave of its row queue (0: none).
Compiled &kernel_for(const FProgram &p, int unroll, int block, bool queue, uint32_t *queue_bytes)
{
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    queue = queue && queue_enabled();
    const std::string skey = shape_key(p) + launch_shape(unroll, block, queue);
    const std::string key = std::to_string(dev) + ":" + skey;
    {
        std::lock_guard<std::mutex> lk(g_src_mu);
        auto it = g_kernel_by_shape.find(key);
        if (it != g_kernel_by_shape.end()) {
            *queue_bytes = g_queue_bytes.at(skey);
            return *it->second;
        }
    }
    Compiled &k = get_kernel(jit_source(p, unroll, block, queue));
    std::lock_guard<std::mutex> lk(g_src_mu);
    g_kernel_by_shape[key] = &k;
    *queue_bytes = g_queue_bytes.at(skey);
    return k;
}

// Largest dynamic LDS of a launch with a row queue; past it (or past 2^31
// rows, whose indices the queue does not hold) the kernel without one runs.
constexpr uint64_t QUEUE_LDS_MAX = 96 * 1024;

// The kernel a launch of n rows uses (with a row queue when `queue` allows
// one and it fits), its dynamic LDS bytes, and whether it has the queue.
Compiled &pick_kernel(const FProgram &p, int64_t n, int unroll, int block, bool queue, uint32_t &lds,
                      bool *queued = nullptr)
{
    uint32_t qbytes = 0;
    if (queued) *queued = false;
    if (queue && n <= INT32_MAX) {
        Compiled &k = kernel_for(p, unroll, block, true, &qbytes);
        // the value-register file (the last LDS area) belongs to the
        // interpreter; the queues of the block's waves go there instead
        const uint64_t total = ((uint64_t)p.vreg_off + 15u) / 16u * 16u + (uint64_t)(block / 64) * qbytes;
        if (!qbytes || total <= QUEUE_LDS_MAX) {
            lds = qbytes ? (uint32_t)total : p.vreg_off;
            if (queued) *queued = qbytes != 0;
            return k;
        }
    }
    lds = p.vreg_off;
    return kernel_for(p, unroll, block, false, &qbytes);
}

} // namespace

namespace {

// Launches the compiled kernel of `src` (blocks of `block` threads) with
// `per_cu_mult` times the blocks per CU the occupancy calculator allows.
void launch_compiled(const FProgram &p, int64_t n, hipStream_t s, int unroll, int per_cu_mult, int block = JIT_BLOCK,
                     bool queue = false)
{
    uint32_t lds = 0;
    Compiled &k = pick_kernel(p, n, unroll, block, queue, lds);
    int per_cu;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = k.blocks_per_cu.find(lds);
        if (it == k.blocks_per_cu.end()) {
            int b = 0;
            SDB_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&b, k.fn, block, lds));
            it = k.blocks_per_cu.emplace(lds, std::max(1, b)).first;
        }
        per_cu = it->second * per_cu_mult;
    }
    const int64_t units = p.vec ? (n + 3) / 4 : n;
    const unsigned grid = stream_grid(units, (int64_t)block * (p.vec ? unroll : 1), per_cu);
    static const bool debug = [] {
        const char *e = getenv("SDB_DEBUG_FUSED");
        return e && *e && *e != '0';
    }();
    if (debug && n >= ((int64_t)1 << 22))
        fprintf(stderr, "[fused launch] rows %lld, unroll %d, block %d, queue %d, LDS %u bytes, %d blocks/CU, grid %u\n",
                (long long)n, unroll, block, (int)queue, lds, per_cu, grid);
    struct Args {
        FProgram p;
        int64_t n;
    } args{p, n};
    size_t size = sizeof(args);
    void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &args, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size, HIP_LAUNCH_PARAM_END};
    SDB_HIP(hipModuleLaunchKernel(k.fn, grid, 1, 1, block, 1, 1, lds, s, nullptr, config));
}

// ------------------------------------------------------------- autotuner
//
// Per program shape (and device) the launch configuration -- 4-row chunks
// per thread per iteration x threads per block, then with or without the
// queued row tail -- is chosen by timing the candidates with HIP events over
// launches of at least TUNE_MIN_ROWS rows.  Every configuration computes the
// same result; only the time differs.  Larger blocks share one copy of the
// program's LDS among more waves, so a program whose staged tables limit the
// blocks per CU runs more waves per CU.
//
// No launch of the default mode waits for a compilation.  The first time a
// shape is seen its kernels are handed to a background compiler thread
// (hipRTC compiles one program at a time per process, about a quarter of a
// second each), and until the default configuration's kernel is ready the
// program runs on the compiled-in interpreter kernel (fused.hip), which
// computes the same rows.  A launch under TUNE_MIN_ROWS needs only the
// default configuration; a larger one queues every candidate, and after
// their trials the fastest one's queued-tail variant.
//
// A program whose only outputs are its terminal accumulators (SUM, GROUP BY:
// every fact-table bundle that ends in an aggregate) runs each phase's
// trials all at once, in one launch, on a copy of the program whose
// accumulators point into a scratch sink kept per device: its own result is
// then computed once, by the kernel chosen so far.  Any other program tries
// one candidate per launch on its real outputs.
constexpr int64_t TUNE_MIN_ROWS = 4 << 20;
// timings per candidate in a one-launch tuning (the fastest counts)
constexpr int TUNE_ROUNDS = 2;

struct TuneCfg {
    int unroll, per_cu_mult, block;
};
// the first is the default configuration (F_DEFAULT_UNROLL, JIT_BLOCK)
const TuneCfg TUNE_CANDIDATES[] = {{2, 1, 256}, {1, 1, 256}, {4, 1, 256}, {2, 1, 512}, {1, 1, 512}, {2, 1, 1024},
                                   {4, 1, 512}, {4, 1, 1024}, {1, 1, 1024}};
constexpr int N_CAND = (int)(sizeof(TUNE_CANDIDATES) / sizeof(TUNE_CANDIDATES[0]));

// After the candidates' trials, a program whose tail can run queued (the
// rows that reach it drained 64 at a time per wave) times the fastest
// candidate once more with the queue, and keeps the faster of the two: the
// queue saves each tile a wait for the lazily read columns, but costs LDS and
// a drain loop, and which wins depends on the program (measured: the q3.1
// join at 600M rows 1.98 -> 1.89 ms, the q2.x join 1.39 -> 1.45 ms).
struct Tune {
    double best_ns_per_row[N_CAND];
    double queue_ns_per_row = 1e300;
    int trials = 0;
    int chosen = -1;
    int base = -1;       // fastest candidate, while its queued trial is pending
    bool chosen_queue = false;
    // background compilation of the shape's kernels
    bool queued_default = false, queued_all = false, queued_tail = false;
    bool default_ready = false, tail_ready = false;
    int cands_ready = 0;  // of the N_CAND - 1 other candidates
    std::string error;    // a failed compilation: the interpreter runs the shape
    bool tuning = false;  // a thread is running one-launch trials
    int tuned_launches = 0; // launches that ran trials
    double tune_ms = 0;   // host time of the one-launch trials
    Tune()
    {
        for (double &b : best_ns_per_row) b = 1e300;
    }
};

std::mutex g_tune_mu;
std::condition_variable g_tune_cv; // a background compilation finished
std::map<std::pair<int, std::string>, Tune> g_tune;

// One kernel for the background compiler: the program (for its source), the
// launch shape, and the tuning entry it readies.
enum JobKind { JOB_DEFAULT, JOB_CANDIDATE, JOB_TAIL };
struct CompileJob {
    FProgram p;
    int unroll, block;
    bool queue;
    std::string arch;
    std::pair<int, std::string> key;
    JobKind kind;
};

// A thread that compiles queued kernels one at a time (the default
// configurations first).  It only generates source and runs hipRTC, never
// the HIP runtime; the launching thread loads the code objects.
//
// It must be stopped before the process tears down its static objects: the
// compiler libraries behind hipRTC set up their own statics during the first
// compilation, after this object, so at exit they are destroyed before this
// object's destructor could wait for a compilation still running on the
// thread (a CPU reproduction crashed or hung at exit 3 times in 3).
// jit_shutdown (sdb_codegen_shutdown; the Python package calls it when the
// interpreter exits) stops it first.
class BackgroundCompiler {
public:
    static BackgroundCompiler &get()
    {
        static BackgroundCompiler c;
        return c;
    }
    static std::atomic<bool> &created()
    {
        static std::atomic<bool> made{false};
        return made;
    }
    void submit(std::vector<CompileJob> jobs)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (!stop_) {
                for (CompileJob &j : jobs) {
                    if (j.kind == JOB_DEFAULT) q_.push_front(std::move(j));
                    else q_.push_back(std::move(j));
                }
                jobs.clear();
            }
        }
        cv_.notify_all();
        drop(jobs);
    }
    // Drops the queued jobs and waits for the compilation in progress; later
    // submissions are dropped too.  A dropped job's program shape records an
    // error: it runs on the interpreter kernel (SDB_CODEGEN_AUTO), or its
    // execution fails where compiled kernels are required.
    void shutdown()
    {
        std::vector<CompileJob> dropped;
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            for (CompileJob &j : q_) dropped.push_back(std::move(j));
            q_.clear();
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
        drop(dropped);
    }
    ~BackgroundCompiler() { shutdown(); }

private:
    BackgroundCompiler()
    {
        rtc(); // the hipRTC library is loaded (and its exit handlers set) before this object's
        th_ = std::thread([this] { run(); });
        created() = true;
    }
    static void drop(const std::vector<CompileJob> &jobs)
    {
        if (jobs.empty()) return;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            for (const CompileJob &j : jobs) {
                Tune &t = g_tune[j.key];
                if (t.error.empty()) t.error = "the background compiler has stopped";
            }
        }
        g_tune_cv.notify_all();
    }
    void run()
    {
        for (;;) {
            CompileJob j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
                if (stop_) return;
                j = std::move(q_.front());
                q_.pop_front();
            }
            std::string err;
            try {
                const std::string src = jit_source(j.p, j.unroll, j.block, j.queue);
                const std::string key = j.arch + "\n" + src;
                bool have;
                {
                    std::lock_guard<std::mutex> lk(g_mu);
                    have = g_code.count(key) != 0;
                }
                if (!have) {
                    std::vector<char> code = compile_source(src, j.arch);
                    std::lock_guard<std::mutex> lk(g_mu);
                    if (g_code.emplace(key, std::move(code)).second) g_compiled++;
                }
            } catch (const std::exception &e) {
                err = e.what();
            }
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                Tune &t = g_tune[j.key];
                if (!err.empty() && t.error.empty()) t.error = err;
                if (j.kind == JOB_DEFAULT) t.default_ready = true;
                else if (j.kind == JOB_CANDIDATE) t.cands_ready++;
                else t.tail_ready = true;
            }
            g_tune_cv.notify_all();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<CompileJob> q_;
    bool stop_ = false;
    std::thread th_;
};

// Scratch sink of the one-launch trials: per device, grown on demand and
// kept; g_sink_mu also serialises the trials of all threads.
std::mutex g_sink_mu;
std::map<int, std::pair<void *, size_t>> g_sink;
constexpr size_t SINK_MAX = size_t(256) << 20;

// Bytes of sink a program's accumulators need, or 0 when the program has
// another output (a stored flag or value column, a built table) or none.
size_t sink_bytes(const FProgram &p)
{
    for (int j = 0; j < F_MAX_FLAGS; j++)
        if (p.flag_store[j]) return 0;
    bool terminal = false;
    for (int i = 0; i < p.ninstr; i++) {
        const FInstr &I = p.ins[i];
        if (I.store || I.kind == FK_BUILD_KEYS || I.kind == FK_BUILD_KEYVALS || I.kind == FK_FILL) return 0;
        terminal |= I.kind == FK_SUM || I.kind == FK_GROUPBY;
    }
    if (!terminal) return 0;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    size_t bytes = 256; // the SUM accumulator
    if (p.gb.k > 0) {
        if (p.gb.prod == 0 || p.gb.prod > (uint64_t)SINK_MAX) return 0;
        bytes += al(p.gb.prod * 8) + al(p.gb.prod * 4) + (size_t)p.gb.k * al(p.gb.prod * 4);
    }
    return bytes <= SINK_MAX ? bytes : 0;
}

// The program with its accumulators moved into `sink` (sink_bytes(p) bytes).
FProgram sink_program(const FProgram &p, uint8_t *sink)
{
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    FProgram q = p;
    size_t off = 0;
    if (q.sum_acc) q.sum_acc = reinterpret_cast<uint64_t *>(sink);
    off += 256;
    if (q.gb.k > 0) {
        q.gb.sums = reinterpret_cast<uint64_t *>(sink + off);
        off += al(q.gb.prod * 8);
        q.gb.seen = reinterpret_cast<uint32_t *>(sink + off);
        off += al(q.gb.prod * 4);
        for (int j = 0; j < q.gb.k; j++) {
            if (q.gb.out_keys[j]) q.gb.out_keys[j] = reinterpret_cast<int32_t *>(sink + off);
            off += al(q.gb.prod * 4);
        }
    }
    return q;
}

// Times the candidates `cands` (with the queued tail when `queue`) on the
// sink copy of p, TUNE_ROUNDS launches each, queued back to back between
// consecutive events and read after one synchronisation; the best ns/row of
// each candidate.  Caller holds g_sink_mu.
std::vector<double> time_in_sink(const FProgram &p, int64_t n, hipStream_t s, const std::vector<int> &cands, bool queue)
{
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    const size_t need = sink_bytes(p);
    auto &sk = g_sink[dev];
    if (sk.second < need) {
        if (sk.first) SDB_HIP(hipFree(sk.first));
        sk.first = nullptr;
        sk.second = 0;
        SDB_HIP(hipMalloc(&sk.first, need));
        sk.second = need;
    }
    const FProgram q = sink_program(p, static_cast<uint8_t *>(sk.first));
    for (int c : cands) { // load the modules before the first event
        uint32_t lds = 0;
        pick_kernel(q, n, TUNE_CANDIDATES[c].unroll, TUNE_CANDIDATES[c].block, queue, lds);
    }
    const int nl = (int)cands.size() * TUNE_ROUNDS;
    std::vector<hipEvent_t> ev(nl + 1, nullptr);
    hipError_t e = hipSuccess;
    for (auto &x : ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    if (e == hipSuccess) e = hipEventRecord(ev[0], s);
    for (int r = 0; r < TUNE_ROUNDS && e == hipSuccess; r++)
        for (size_t i = 0; i < cands.size() && e == hipSuccess; i++) {
            const TuneCfg &c = TUNE_CANDIDATES[cands[i]];
            launch_compiled(q, n, s, c.unroll, c.per_cu_mult, c.block, queue);
            e = hipEventRecord(ev[r * cands.size() + i + 1], s);
        }
    if (e == hipSuccess) e = hipEventSynchronize(ev[nl]);
    std::vector<double> best(cands.size(), 1e300);
    for (int l = 0; l < nl && e == hipSuccess; l++) {
        float ms = 0;
        e = hipEventElapsedTime(&ms, ev[l], ev[l + 1]);
        best[l % cands.size()] = std::min(best[l % cands.size()], ms * 1e6 / (double)n);
    }
    for (auto &x : ev)
        if (x) (void)hipEventDestroy(x);
    hip_check(e, "fused kernel tuning");
    return best;
}

std::vector<CompileJob> jobs_for(const FProgram &p, const std::pair<int, std::string> &key, const std::string &arch,
                                 JobKind kind, int cand = 0)
{
    std::vector<CompileJob> out;
    auto add = [&](int c, bool queue, JobKind k) {
        out.push_back(CompileJob{p, TUNE_CANDIDATES[c].unroll, TUNE_CANDIDATES[c].block, queue, arch, key, k});
    };
    if (kind == JOB_DEFAULT) add(0, false, kind);
    else if (kind == JOB_CANDIDATE)
        for (int c = 1; c < N_CAND; c++) add(c, false, kind);
    else add(cand, true, kind);
    return out;
}

} // namespace

FusedLaunch prepare_fused_jit(const FProgram &p, int64_t n, hipStream_t s, int unroll, bool wait)
{
    FusedLaunch L;
    if (n <= 0) return L;
    if (p.magic != F_MAGIC) fail(SDB_ERR_RUNTIME, "fused program without magic");
    if (unroll != 0) {
        // an explicit configuration: compiled now if it is not yet
        L.mode = FusedLaunch::EXPLICIT;
        L.unroll = unroll & 7;
        L.block = (unroll & SDB_UNROLL_BLOCK1024) ? 1024 : (unroll & SDB_UNROLL_BLOCK512) ? 512 : JIT_BLOCK;
        L.queue = (unroll & SDB_UNROLL_QUEUED) != 0;
        uint32_t lds = 0;
        pick_kernel(p, n, L.unroll, L.block, L.queue, lds);
        return L;
    }
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    const std::pair<int, std::string> key{dev, shape_key(p)};
    const bool big = n >= TUNE_MIN_ROWS;
    const bool sinkable = big && sink_bytes(p) > 0;
    std::unique_lock<std::mutex> lk(g_tune_mu);
    for (;;) {
        Tune &t = g_tune[key];
        std::vector<CompileJob> jobs;
        if (!t.queued_default) {
            t.queued_default = true;
            jobs = jobs_for(p, key, device_arch(), JOB_DEFAULT);
        }
        if (big && !t.queued_all) {
            t.queued_all = true;
            for (CompileJob &j : jobs_for(p, key, device_arch(), JOB_CANDIDATE)) jobs.push_back(std::move(j));
        }
        if (big && t.base >= 0 && !t.queued_tail && t.chosen < 0) {
            t.queued_tail = true;
            if (queue_enabled())
                for (CompileJob &j : jobs_for(p, key, device_arch(), JOB_TAIL, t.base)) jobs.push_back(std::move(j));
            else
                t.tail_ready = true;
        }
        if (!jobs.empty()) {
            lk.unlock();
            BackgroundCompiler::get().submit(std::move(jobs));
            lk.lock();
            continue; // the state may have moved meanwhile
        }
        if (!t.error.empty()) {
            if (wait) fail(SDB_ERR_RUNTIME, t.error);
            L.mode = FusedLaunch::INTERPRET;
            return L;
        }
        if (big && t.chosen >= 0) {
            L.mode = FusedLaunch::TUNED;
            L.cand = t.chosen;
            L.queue = t.chosen_queue;
            return L;
        }
        const bool cands_ready = t.default_ready && t.cands_ready >= N_CAND - 1;
        if (big && cands_ready && t.base < 0 && !t.tuning) {
            if (sinkable) {
                // every candidate, now, in the sink
                t.tuning = true;
                lk.unlock();
                std::vector<int> all;
                for (int c = 0; c < N_CAND; c++) all.push_back(c);
                std::vector<double> ns;
                const auto w0 = std::chrono::steady_clock::now();
                try {
                    std::lock_guard<std::mutex> sl(g_sink_mu);
                    ns = time_in_sink(p, n, s, all, false);
                } catch (...) {
                    lk.lock();
                    g_tune[key].tuning = false;
                    throw;
                }
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
                lk.lock();
                Tune &u = g_tune[key];
                u.tuning = false;
                u.tune_ms += ms;
                u.tuned_launches++;
                int best = 0;
                for (int c = 0; c < N_CAND; c++) {
                    u.best_ns_per_row[c] = std::min(u.best_ns_per_row[c], ns[c]);
                    if (u.best_ns_per_row[c] < u.best_ns_per_row[best]) best = c;
                }
                u.trials = N_CAND;
                u.base = best;
                continue; // queue its tail variant
            }
            // one candidate per launch, on the real outputs
            L.mode = FusedLaunch::TRIAL;
            L.cand = t.trials % N_CAND;
            L.key = key;
            return L;
        }
        if (big && t.base >= 0 && t.tail_ready && !t.tuning) {
            uint32_t lds = 0;
            bool queued = false;
            {
                const TuneCfg c = TUNE_CANDIDATES[t.base];
                lk.unlock();
                pick_kernel(p, n, c.unroll, c.block, true, lds, &queued); // compiled: loads it
                lk.lock();
            }
            Tune &u = g_tune[key];
            if (u.chosen >= 0 || u.tuning) continue;
            if (!queued) {
                // no queue for this program (no row tail, or its LDS does not fit)
                u.chosen = u.base;
                u.base = -1;
                continue;
            }
            if (sinkable) {
                u.tuning = true;
                const int base = u.base;
                lk.unlock();
                std::vector<double> ns;
                const auto w0 = std::chrono::steady_clock::now();
                try {
                    std::lock_guard<std::mutex> sl(g_sink_mu);
                    ns = time_in_sink(p, n, s, {base}, true);
                } catch (...) {
                    lk.lock();
                    g_tune[key].tuning = false;
                    throw;
                }
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
                lk.lock();
                Tune &v = g_tune[key];
                v.tuning = false;
                v.tune_ms += ms;
                v.tuned_launches++;
                v.queue_ns_per_row = ns[0];
                v.chosen_queue = ns[0] < v.best_ns_per_row[base];
                v.chosen = base;
                v.base = -1;
                continue;
            }
            L.mode = FusedLaunch::TRIAL;
            L.cand = u.base;
            L.queue = true;
            L.key = key;
            return L;
        }
        if (wait && !(t.default_ready && (!big || t.tuning))) {
            // compiled kernels only: wait for the background compiler
            g_tune_cv.wait(lk);
            continue;
        }
        if (t.default_ready) {
            L.mode = FusedLaunch::TUNED;
            L.cand = 0;
            return L;
        }
        L.mode = FusedLaunch::INTERPRET;
        return L;
    }
}

void run_fused_jit(const FusedLaunch &L, const FProgram &p, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    switch (L.mode) {
    case FusedLaunch::INTERPRET: launch_fused(p, n, s); return;
    case FusedLaunch::EXPLICIT: launch_compiled(p, n, s, L.unroll, 1, L.block, L.queue); return;
    case FusedLaunch::TUNED: {
        const TuneCfg &c = TUNE_CANDIDATES[L.cand];
        launch_compiled(p, n, s, c.unroll, c.per_cu_mult, c.block, L.queue);
        return;
    }
    case FusedLaunch::TRIAL: break;
    }
    // a trial on the program's real outputs, timed with HIP events
    const TuneCfg &c = TUNE_CANDIDATES[L.cand];
    hipEvent_t a, b;
    SDB_HIP(hipEventCreate(&a));
    SDB_HIP(hipEventCreate(&b));
    float ms = 0;
    hipError_t e = hipEventRecord(a, s);
    if (e == hipSuccess) {
        launch_compiled(p, n, s, c.unroll, c.per_cu_mult, c.block, L.queue);
        e = hipEventRecord(b, s);
    }
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    hip_check(e, "fused kernel tuning");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    Tune &t = g_tune[L.key];
    if (t.chosen >= 0) return;
    const double ns = ms * 1e6 / (double)n;
    t.tuned_launches++;
    if (L.queue) {
        if (t.base != L.cand) return;
        t.queue_ns_per_row = ns;
        t.chosen_queue = ns < t.best_ns_per_row[L.cand];
        t.chosen = L.cand;
        t.base = -1;
        return;
    }
    if (t.base >= 0) return;
    t.best_ns_per_row[L.cand] = std::min(t.best_ns_per_row[L.cand], ns);
    if (++t.trials >= N_CAND) {
        int best = 0;
        for (int i = 1; i < N_CAND; i++)
            if (t.best_ns_per_row[i] < t.best_ns_per_row[best]) best = i;
        t.base = best; // next: its queued trial, once compiled
    }
}

void launch_fused_jit(const FProgram &p, int64_t n, hipStream_t s, int unroll)
{
    run_fused_jit(prepare_fused_jit(p, n, s, unroll, false), p, n, s);
}

void jit_shutdown()
{
    if (BackgroundCompiler::created()) BackgroundCompiler::get().shutdown();
}

bool jit_wait_idle(double timeout_s)
{
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    std::unique_lock<std::mutex> lk(g_tune_mu);
    for (;;) {
        bool busy = false;
        for (auto &kv : g_tune) {
            const Tune &t = kv.second;
            if (!t.error.empty()) continue;
            busy |= t.queued_default && !t.default_ready;
            busy |= t.queued_all && t.cands_ready < N_CAND - 1;
            busy |= t.queued_tail && !t.tail_ready;
        }
        if (!busy) return true;
        if (g_tune_cv.wait_until(lk, until) == std::cv_status::timeout) return false;
    }
}

// ------------------------------------------------------------ batched kernel
//
// Many pipelines' fact-table programs in one launch: one scan of the union of
// their prefetched columns.  Probes on fact key columns that a program's rows
// must all pass to reach its terminal are evaluated for every program at once:
// per masked column one table lookup per row gives a word of program bits
// (FMaskCol), and the AND of a row's words says which programs the row can
// still contribute to.  Per tile, the programs of each shape (identical
// generated body) then run in a loop that reads each program from a device
// array, skipping programs no row of the wave keeps; a program sees only its
// kept rows (`act`), and the probes the masks decided are left out of its
// body.  Each program has its own LDS region (lds_base) and its own
// scalar-sum slot.

namespace {

constexpr int BATCH_MAX_PF = 16;
// threads per block of the batched kernel (the batch's programs keep little
// LDS since they stage no tables: one block's regions fit several times per
// CU); measured on 32 pipelines at 600M rows, 512-thread blocks with one tile
// of columns in registers ran 17.4 ms against 19.4 ms for 1024 threads with
// the next tile's columns loading meanwhile (the registers that double
// buffering takes cost more waves than the overlap gains)
constexpr int BATCH_BLOCK = 512;
constexpr int64_t BATCH_MIN_TILES = 8;
constexpr int BATCH_MAX_SHAPES = 32;
// a probe joins the shared masks only over a key range of at most this many
// keys (4 bytes of table per key; SDB_BATCH_MASK_MAX_KEYS overrides)
int64_t batch_mask_probe_len()
{
    static const int64_t v = [] {
        const char *e = getenv("SDB_BATCH_MASK_MAX_KEYS");
        return e && *e ? (int64_t)atoll(e) : (int64_t(1) << 22);
    }();
    return v;
}

const char *BATCH_HEAD = R"(struct BatchArgs {
    const FProgram *progs;
    int64_t n;
    int32_t nprog;
    uint32_t sums_off;
    uint32_t full;
    uint32_t lazy; // 1: mask tables read in morder, each only for rows some program still keeps
    int32_t first[NSHAPE + 1];
    const int32_t *pf[NPF > 0 ? NPF : 1];
    const uint8_t *mtab[NMASK > 0 ? NMASK : 1];
    const uint8_t *mdg[NMASK > 0 ? NMASK : 1]; // rank dictionary in global memory (header first), or null
    int32_t mlo[NMASK > 0 ? NMASK : 1];
    uint32_t mlen[NMASK > 0 ? NMASK : 1];
    uint32_t mdflt[NMASK > 0 ? NMASK : 1];
    uint32_t mshift[NMASK > 0 ? NMASK : 1];
    uint32_t mwbytes[NMASK > 0 ? NMASK : 1];
    uint32_t mlds[NMASK > 0 ? NMASK : 1]; // 1 + LDS byte offset of a staged table (0: read from memory)
    uint32_t mdict[NMASK > 0 ? NMASK : 1]; // 1 + LDS byte offset of the rank dictionary's copy (0: none)
    int32_t morder[NMASK > 0 ? NMASK : 1]; // lazy reads: masked columns, fewest rows kept first
    int32_t bitpos[32]; // program i's bit in the mask words
};
template <int R, int UU> struct Pf {
    typename Vec<R>::type c[NPF > 0 ? NPF : 1][UU];
};
template <int R, int UU>
__device__ __forceinline__ void pf_load(const BatchArgs &A, const int64_t (&row0)[UU], Pf<R, UU> &pf)
{
#pragma unroll
    for (int j = 0; j < NPF; j++)
        if (A.pf[j]) // columns this batch does not read stay unloaded
#pragma unroll
            for (int u = 0; u < UU; u++) pf.c[j][u] = load_vec<R>(A.pf[j], row0[u]);
}
// Program bits of every row: the AND of the row's words in the mask tables
// of the masked columns (a column this batch does not mask has mlen 0 and
// is skipped; the branch is uniform).  An entry is the window of the word
// that covers the column's probing programs, widened back with ones outside
// it; keys outside the table take the column's default word.
//
// Eager form: every table is read for every row, all loads issued before
// any is combined, the indexes clamped into the table so each load is an
// unconditional aligned dword holding the entry.  Lazy form (A.lazy): the
// tables are read in A.morder, the column expected to keep the fewest rows
// first, and a row reads the next table only while some program still keeps
// it, so a selective first column saves most later reads; the result is the
// same AND.
template <int R>
__device__ __forceinline__ uint32_t mask_word(const BatchArgs &A, const uint8_t *lds, int m, bool staged, int32_t x)
{
    const uint32_t idx = (uint32_t)x - (uint32_t)A.mlo[m];
    const uint32_t byte = (idx < A.mlen[m] ? idx : A.mlen[m] - 1u) * A.mwbytes[m];
    const uint32_t raw = staged ? *reinterpret_cast<const uint32_t *>(lds + (A.mlds[m] - 1u) + (byte & ~3u))
                                : *reinterpret_cast<const uint32_t *>(A.mtab[m] + (byte & ~3u));
    return raw >> ((byte & 3u) * 8u);
}
// The stored window of key x from column m's rank dictionary in LDS (0 for
// a key without a set presence bit; out-of-table keys are left to
// mask_widen, which gives them the default word).
__device__ __forceinline__ uint32_t dict_window(const BatchArgs &A, const uint8_t *lds, int m, int32_t x)
{
    const uint32_t idx = (uint32_t)x - (uint32_t)A.mlo[m], len = A.mlen[m];
    if (idx >= len) return 0u;
    const uint8_t *base = lds + (A.mdict[m] - 1u);
    const uint32_t bits = reinterpret_cast<const uint32_t *>(base)[idx >> 5], bit = idx & 31u;
    const uint32_t r = reinterpret_cast<const uint16_t *>(base + dict_prefix_off(len))[idx >> 5] +
                       (uint32_t)__popc(bits & ((1u << bit) - 1u));
    return ((bits >> bit) & 1u) ? reinterpret_cast<const uint32_t *>(base + dict_vals_off(len))[r] : 0u;
}
__device__ __forceinline__ uint32_t mask_widen(const BatchArgs &A, int m, int32_t x, uint32_t w)
{
    const uint32_t idx = (uint32_t)x - (uint32_t)A.mlo[m];
    const uint32_t wm = A.mwbytes[m] >= 4u ? ~0u : (1u << (8u * A.mwbytes[m])) - 1u;
    const uint32_t word = ((w & wm) << A.mshift[m]) | ~(wm << A.mshift[m]);
    return idx < A.mlen[m] ? word : A.mdflt[m];
}
template <int R, int UU>
__device__ __forceinline__ void mask_rows(const BatchArgs &A, const Pf<R, UU> &pf, const uint8_t *lds, uint32_t dmask,
                                          uint32_t (&mw)[UU][R])
{
#pragma unroll
    for (int u = 0; u < UU; u++)
#pragma unroll
        for (int k = 0; k < R; k++) mw[u][k] = A.full;
    if (!A.lazy) {
        uint32_t w[NMASK > 0 ? NMASK : 1][UU][R];
#pragma unroll
        for (int m = 0; m < NMASK; m++) {
            if (!A.mlen[m]) continue;
            // a table staged in LDS, or its rank dictionary there, is read
            // there (the branches are uniform)
            const bool staged = A.mlds[m] != 0, dict = (dmask >> m) & 1u;
#pragma unroll
            for (int u = 0; u < UU; u++) {
                int32_t x[R];
                vec_rows<R>(pf.c[MSLOT[m]][u], x);
#pragma unroll
                for (int k = 0; k < R; k++)
                    w[m][u][k] = dict ? dict_window(A, lds, m, x[k]) : mask_word<R>(A, lds, m, staged, x[k]);
            }
        }
#pragma unroll
        for (int m = 0; m < NMASK; m++) {
            if (!A.mlen[m]) continue;
#pragma unroll
            for (int u = 0; u < UU; u++) {
                int32_t x[R];
                vec_rows<R>(pf.c[MSLOT[m]][u], x);
#pragma unroll
                for (int k = 0; k < R; k++) mw[u][k] &= mask_widen(A, m, x[k], w[m][u][k]);
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < NMASK; j++) {
        const int m = A.morder[j];
        if (m < 0 || !A.mlen[m]) continue;
        const bool staged = A.mlds[m] != 0, dict = (dmask >> m) & 1u;
#pragma unroll
        for (int u = 0; u < UU; u++) {
            // the key column of mask m (m is uniform: one register is selected)
            int32_t x[R];
            vec_rows<R>(pf.c[MSLOT[0]][u], x);
#pragma unroll
            for (int c = 1; c < NMASK; c++)
                if (m == c) vec_rows<R>(pf.c[MSLOT[c]][u], x);
#pragma unroll
            for (int k = 0; k < R; k++)
                if (mw[u][k])
                    mw[u][k] &= mask_widen(A, m, x[k], dict ? dict_window(A, lds, m, x[k]) : mask_word<R>(A, lds, m, staged, x[k]));
        }
    }
}
// The programs of the launch, read-only while it runs: reading them through
// the constant address space lets their fields go through scalar loads.
typedef const FProgram __attribute__((address_space(4))) *ProgPtr;
template <int R, int UU>
__device__ __forceinline__ void rows_of_program(const uint32_t (&mw)[UU][R], int i, uint32_t (&act)[UU])
{
#pragma unroll
    for (int u = 0; u < UU; u++) {
        act[u] = 0;
#pragma unroll
        for (int k = 0; k < R; k++) act[u] |= __builtin_amdgcn_ubfe(mw[u][k], (uint32_t)i, 1u) << k;
    }
}
)";

const char *BATCH_TAIL = R"(
extern "C" __global__ void __launch_bounds__(BLOCK) sdb_fused_batch(const BatchArgs A_arg)
{
    // read in place from the kernel-argument segment (scalar loads; the
    // arguments preload into registers).  Passed instead as a pointer to a
    // device block, the same kernel ran 1.7x slower.
    const BatchArgs &A = *(const BatchArgs __attribute__((address_space(4))) *)__builtin_amdgcn_kernarg_segment_ptr();
    extern __shared__ uint4 dyn_lds[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(dyn_lds);
    for (int i = 0; i < A.nprog; i++) {
        const FProgram &P = ((ProgPtr)A.progs)[i];
        if (P.magic != F_MAGIC) return;
        fdev::block_prologue<BLOCK>(P, lds + P.lds_base);
    }
    unsigned long long *sums = reinterpret_cast<unsigned long long *>(lds + A.sums_off);
    for (int i = threadIdx.x; i < A.nprog; i += BLOCK) sums[i] = 0;
    // mask tables staged in LDS (their sizes are padded to 16 bytes), and
    // the rank dictionaries that fitted (bitmap, counts and set windows)
    uint32_t dmask = 0;
    for (int m = 0; m < NMASK; m++)
        if (A.mlds[m]) {
            const uint4 *src = reinterpret_cast<const uint4 *>(A.mtab[m]);
            uint4 *dst = reinterpret_cast<uint4 *>(lds + (A.mlds[m] - 1u));
            const uint32_t n16 = (A.mlen[m] * A.mwbytes[m] + 15u) / 16u;
            for (uint32_t v = threadIdx.x; v < n16; v += BLOCK) dst[v] = src[v];
        } else if (A.mdict[m]) {
            const uint32_t *hdr = reinterpret_cast<const uint32_t *>(A.mdg[m]);
            if (!hdr[0]) continue; // more set keys than F_DICT_CAP: the table is read instead
            dmask |= 1u << m;
            const uint4 *src = reinterpret_cast<const uint4 *>(A.mdg[m] + 16);
            uint4 *dst = reinterpret_cast<uint4 *>(lds + (A.mdict[m] - 1u));
            const uint32_t n16 = (dict_vals_off(A.mlen[m]) + 4u * hdr[1] + 15u) / 16u;
            for (uint32_t v = threadIdx.x; v < n16; v += BLOCK) dst[v] = src[v];
        }
    dmask = __builtin_amdgcn_readfirstlane(dmask);
    __syncthreads();
    const int64_t n = A.n;
    if constexpr (VEC) {
        const int64_t nch = n >> 2, tile = (int64_t)U * BLOCK, ntiles = nch / tile;
        auto rows_of = [&](int64_t t, int64_t (&r)[U]) {
#pragma unroll
            for (int u = 0; u < U; u++) r[u] = (t * tile + u * BLOCK + threadIdx.x) << 2;
        };
        int64_t t = blockIdx.x;
        if constexpr (DOUBLE_BUFFER) {
            Pf<4, U> next = {};
            if (t < ntiles) {
                int64_t r[U];
                rows_of(t, r);
                pf_load<4, U>(A, r, next);
            }
            while (t < ntiles) {
                const Pf<4, U> cur = next;
                int64_t r[U];
                rows_of(t, r);
                const int64_t tn = t + gridDim.x;
                if (tn < ntiles) {
                    int64_t rn[U];
                    rows_of(tn, rn);
                    pf_load<4, U>(A, rn, next);
                }
                run_all<4, U>(A, r, cur, lds, sums, dmask);
                t = tn;
            }
        } else {
            // one tile's columns in registers: fewer registers, more waves
            for (; t < ntiles; t += gridDim.x) {
                int64_t r[U];
                rows_of(t, r);
                Pf<4, U> cur = {};
                pf_load<4, U>(A, r, cur);
                run_all<4, U>(A, r, cur, lds, sums, dmask);
            }
        }
        for (int64_t c = ntiles * tile + (int64_t)blockIdx.x * BLOCK + threadIdx.x; c < nch;
             c += (int64_t)gridDim.x * BLOCK) {
            const int64_t r[1] = {c << 2};
            Pf<4, 1> pf = {};
            pf_load<4, 1>(A, r, pf);
            run_all<4, 1>(A, r, pf, lds, sums, dmask);
        }
        if (blockIdx.x == gridDim.x - 1 && threadIdx.x < (n & 3)) {
            const int64_t r[1] = {(nch << 2) + threadIdx.x};
            Pf<1, 1> pf = {};
            pf_load<1, 1>(A, r, pf);
            run_all<1, 1>(A, r, pf, lds, sums, dmask);
        }
    } else {
        for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
            const int64_t r[1] = {i};
            Pf<1, 1> pf = {};
            pf_load<1, 1>(A, r, pf);
            run_all<1, 1>(A, r, pf, lds, sums, dmask);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < A.nprog; i += BLOCK)
        if (sums[i] && ((ProgPtr)A.progs)[i].sum_acc)
            atomicAdd(reinterpret_cast<unsigned long long *>(((ProgPtr)A.progs)[i].sum_acc), sums[i]);
    for (int i = 0; i < A.nprog; i++) {
        const FProgram &P = ((ProgPtr)A.progs)[i];
        fdev::block_epilogue<BLOCK>(P, lds + P.lds_base, 0);
    }
}
)";

// Shape library of the batched kernels (per device, fact table, chunk width
// and vector mode): the bodies of every program shape met so far, a stable
// prefetch slot for every fact column met so far, and the slots masked so
// far.  A batched kernel contains all the library's bodies (absent shapes run
// empty loops), so batches of different composition reuse one compiled
// kernel once their shapes are known.
struct BatchLibrary {
    std::vector<std::string> bodies;
    std::vector<const int32_t *> slots;
    std::vector<int> mask_slots;
    std::unordered_map<std::string, std::string> body_of; // shape key + decided probes -> body
    // the kernel of the library's current state per launch shape (rebuilt
    // when the library grows)
    struct Src {
        size_t state[3] = {~size_t(0), 0, 0}; // bodies, slots, masked slots the source covers
        std::string text;
        uint64_t version = 0;             // bumped whenever text changes
        std::map<int, Compiled *> kernel; // per device, for text
    };
    std::map<int, Src> srcs; // by launch shape (BatchShape::id)
};

// Batched-kernel autotuner state per library: the best time per row and
// program measured for each candidate launch shape.
thread_local bool t_batch_trials = true;
thread_local int t_batch_waves_cap = 0; // waves per CU a batched launch may take (0: what fits)
// kinds of batch (mask-table signatures) a library tunes; batches of further
// kinds take the default shape
constexpr int BATCH_TUNE_MAX_KINDS = 16;

struct BatchTune {
    double best[16];
    int trials = 0, chosen = -1;
    BatchTune()
    {
        for (double &b : best) b = 1e300;
    }
};
std::map<std::string, BatchTune> g_batch_tune; // guarded by g_tune_mu

// Launch shape of the batched kernel: threads per block, and whether the
// next tile's columns load while the current one runs.
// Launch shape and mask-table layout of the batched kernel: threads per
// block, whether the next tile's columns load while the current one runs,
// and whether every mask table stores whole 32-bit words (wide) or only the
// 8/16/32-bit window covering the programs that probe its column (narrow:
// smaller tables, an extra shift and mask per lookup).  The layout is a
// kernel argument, so both layouts run the same compiled kernel.
struct BatchShape {
    int block;
    bool double_buffer;
    bool wide = false;
    bool stage = false; // the smallest mask tables copied into LDS per block
    bool dict = false;  // sparse mask tables as rank dictionaries in LDS
    int id() const { return block * 2 + (double_buffer ? 1 : 0); } // the compiled source
    int key() const { return ((id() * 2 + (wide ? 1 : 0)) * 2 + (stage ? 1 : 0)) * 2 + (dict ? 1 : 0); }
};
// autotuner candidates; the first is the default below the tuning size
const BatchShape BATCH_CANDIDATES[] = {{512, false, false, false}, {256, false, false, false},
                                       {1024, false, false, false}, {1024, true, false, false},
                                       {256, false, true, false},  {512, false, true, false},
                                       {256, false, false, true},  {1024, false, false, true},
                                       {512, false, false, false, true}, {256, false, false, false, true}};
// a mask table becomes a rank dictionary (with the dict layout) when it spans
// at least this many keys (and at most F_DICT_MAX_LEN): a sparse key range;
// SDB_BATCH_DICT_MIN_LEN sets another bound
constexpr uint32_t BATCH_DICT_MIN_LEN = 32768;
uint32_t batch_dict_min_len()
{
    const char *e = getenv("SDB_BATCH_DICT_MIN_LEN");
    return e && *e ? (uint32_t)std::max(1L, atol(e)) : BATCH_DICT_MIN_LEN;
}
// LDS a batched block may take with staged mask tables (gfx950: 160 KB per CU)
constexpr uint64_t BATCH_STAGE_LDS_MAX = 144 * 1024;
constexpr int N_BATCH_CAND = (int)(sizeof(BATCH_CANDIDATES) / sizeof(BATCH_CANDIDATES[0]));

std::mutex g_lib_mu;
std::map<std::string, BatchLibrary> g_libs;

// Rewrites a program's prefetch registers to library slots (adding its
// columns); a column beyond BATCH_MAX_PF slots is read from memory instead.
void remap_prefetch(FProgram &q, std::vector<const int32_t *> &slots)
{
    int map[F_MAX_PF];
    for (int j = 0; j < q.npf; j++) {
        auto it = std::find(slots.begin(), slots.end(), q.pf_mem[j]);
        if (it != slots.end()) map[j] = (int)(it - slots.begin());
        else if ((int)slots.size() < BATCH_MAX_PF) {
            map[j] = (int)slots.size();
            slots.push_back(q.pf_mem[j]);
        } else map[j] = -1;
    }
    auto fix = [&](int8_t &reg, const int32_t *&mem) {
        if (reg < F_PF_BASE) return;
        const int j = reg - F_PF_BASE;
        if (map[j] >= 0) {
            reg = (int8_t)(F_PF_BASE + map[j]);
        } else {
            mem = q.pf_mem[j];
            reg = -1;
        }
    };
    for (int i = 0; i < q.ninstr; i++) {
        FInstr &I = q.ins[i];
        fix(I.va, I.a);
        fix(I.vb, I.b);
    }
    for (int j = 0; j < q.gb.k; j++) fix(q.gb.key_reg[j], q.gb.key_mem[j]);
    fix(q.gb.agg_reg, q.gb.agg_mem);
    q.npf = 0; // the batch kernel loads the slots itself
    // columns the program reads from memory (lazily, per flagged row) come
    // from the shared scan too, while slots last: the scan streams every
    // slot once for all programs
    auto lift = [&](int8_t &reg, const int32_t *&mem) {
        if (reg >= 0 || !mem) return;
        auto it = std::find(slots.begin(), slots.end(), mem);
        if (it == slots.end()) {
            if ((int)slots.size() >= BATCH_MAX_PF) return;
            slots.push_back(mem);
            it = slots.end() - 1;
        }
        reg = (int8_t)(F_PF_BASE + (it - slots.begin()));
        mem = nullptr;
    };
    for (int i = 0; i < q.ninstr; i++) {
        FInstr &I = q.ins[i];
        switch (I.kind) {
        case FK_SEL_COL:
        case FK_BINOP_CC:
            lift(I.va, I.a);
            lift(I.vb, I.b);
            break;
        case FK_BINOP_LC: lift(I.vb, I.b); break;
        case FK_SEL_LIT:
        case FK_SEL_RANGE:
        case FK_SEL_IN2:
        case FK_BINOP_CL:
        case FK_PROBE_SEMI:
        case FK_PROBE_FULL:
        case FK_PROBE_BITS:
        case FK_SUM: lift(I.va, I.a); break;
        default: break;
        }
    }
    for (int j = 0; j < q.gb.k; j++) lift(q.gb.key_reg[j], q.gb.key_mem[j]);
    lift(q.gb.agg_reg, q.gb.agg_mem);
}

bool is_probe_kind(uint8_t k) { return k == FK_PROBE_SEMI || k == FK_PROBE_FULL || k == FK_PROBE_BITS; }

// The probes every row reaching the program's terminal has passed, keyed by
// a prefetched fact column: the flag of each register is followed as the
// AND of a set of the program's predicates (AND and probes add one, NONE
// restarts from one, a logical op joins two sets).  Only programs whose sole
// effect is their terminal (SUM or GROUPBY; no stored buffer or flags, no
// build) whose terminal flag is such a conjunction take part; for the others
// the list is empty and the program sees every row.
std::vector<int> shared_probes(const FProgram &q)
{
    std::vector<int> out;
    if (q.ninstr < 1) return out;
    const FInstr &T = q.ins[q.ninstr - 1];
    if (T.kind != FK_SUM && T.kind != FK_GROUPBY) return out;
    for (int j = 0; j < F_MAX_FLAGS; j++)
        if (q.flag_store[j]) return out;
    static_assert(F_MAX_INSTR < 31, "instruction sets are 31-bit masks");
    constexpr uint32_t BAD = 1u << 31;
    uint32_t conj[F_MAX_FLAGS];
    for (int j = 0; j < F_MAX_FLAGS; j++) conj[j] = q.flag_init[j] == FF_UNDEF ? BAD : 0u;
    for (int i = 0; i + 1 < q.ninstr; i++) {
        const FInstr &I = q.ins[i];
        if (I.store) return out;
        switch (I.kind) {
        case FK_SEL_LIT:
        case FK_SEL_COL:
        case FK_SEL_RANGE:
        case FK_SEL_IN2:
            conj[I.fout] = I.logic == SDB_AND ? (conj[I.fin] | (1u << i)) : I.logic == SDB_NONE ? (1u << i) : BAD;
            break;
        case FK_LOGICAL:
            conj[I.fout] = I.logic == SDB_AND ? (conj[I.fin] | conj[I.floc]) : I.logic == SDB_NONE ? conj[I.floc] : BAD;
            break;
        case FK_PROBE_SEMI:
        case FK_PROBE_FULL:
        case FK_PROBE_BITS: conj[I.fout] = conj[I.fin] | (1u << i); break;
        case FK_FILL:
        case FK_BINOP_CC:
        case FK_BINOP_LC:
        case FK_BINOP_CL: break;
        default: return out; // builds, a second terminal
        }
    }
    const uint32_t c = conj[T.fin];
    if (c & BAD) return out;
    for (int i = 0; i + 1 < q.ninstr; i++) {
        const FInstr &I = q.ins[i];
        // (a rank dictionary's values need its presence word: such probes
        // stay whole)
        if (((c >> i) & 1u) && is_probe_kind(I.kind) && I.va >= F_PF_BASE &&
            !(I.kind == FK_PROBE_BITS && I.vkind == FV_RANK16) &&
            (int64_t)I.ht_max - I.ht_min + 1 <= batch_mask_probe_len())
            out.push_back(i);
    }
    return out;
}

uint32_t mask_kind(const FInstr &I)
{
    return I.kind == FK_PROBE_BITS ? FM_BITS : I.kind == FK_PROBE_SEMI ? FM_SEMI : FM_PAIR;
}

struct MaskPlan {
    int slot;
    int64_t lo = 0, hi = -1;
    std::vector<FMaskEntry> ents;
    uint32_t shift = 0, wbytes = 4; // the window of the word the table stores
    uint32_t lds = 0;               // 1 + byte offset of its LDS copy (0: read from memory)
    uint32_t dict = 0;              // 1 + byte offset of its rank dictionary's LDS copy (0: none)
    uint64_t bytes() const { return ((uint64_t)std::max<int64_t>(hi - lo + 1, 1) * wbytes + 15) / 16 * 16; }
};

struct BatchPlan {
    std::vector<FProgram> progs; // in shape order
    std::vector<int> first;      // shape k: progs[first[k], first[k+1])
    std::vector<const int32_t *> pf;  // slot -> column this batch reads (null: unused)
    std::vector<int> mask_slots;      // the library's masked slots (kernel order)
    std::vector<MaskPlan> masks;      // this batch's tables, one per masked slot it uses
    std::vector<int> bitpos;          // program (batch order) -> its bit in the mask words
    bool lazy = false;                // mask tables read lazily, in morder
    std::vector<int> morder;          // kernel mask indexes, in reading order (-1: none)
    std::string src;
    BatchLibrary::Src *src_entry = nullptr; // the library source it is (never erased)
    uint64_t version = 0;                   // its version
    uint32_t lds = 0, sums_off = 0, full = 0;
    bool vec = true;
};

// A launch shape fixed by SDB_BATCH_BLOCK / SDB_BATCH_DOUBLE_BUFFER /
// SDB_BATCH_WIDE_MASKS / SDB_BATCH_STAGE_MASKS / SDB_BATCH_DICT_MASKS (no
// tuning), if any is set.
bool batch_shape_fixed(BatchShape &out)
{
    const char *b = getenv("SDB_BATCH_BLOCK");
    const char *d = getenv("SDB_BATCH_DOUBLE_BUFFER");
    const char *w = getenv("SDB_BATCH_WIDE_MASKS");
    const char *g = getenv("SDB_BATCH_STAGE_MASKS");
    const char *r = getenv("SDB_BATCH_DICT_MASKS");
    if (!(b && *b) && !(d && *d) && !(w && *w) && !(g && *g) && !(r && *r)) return false;
    const int blk = b && *b ? atoi(b) : BATCH_BLOCK;
    out.block = blk == 256 || blk == 512 || blk == 1024 ? blk : BATCH_BLOCK;
    out.double_buffer = d && *d == '1';
    out.wide = w && *w == '1';
    out.stage = g && *g == '1';
    out.dict = r && *r == '1' && !out.stage;
    return true;
}

std::string batch_source(const BatchLibrary &lib, int unroll, bool vec, const BatchShape &shape)
{
    std::ostringstream s;
    s << KERNEL_TEMPLATE_INCLUDES << "constexpr int BLOCK = " << shape.block
      << ";\nconstexpr bool DOUBLE_BUFFER = " << (shape.double_buffer ? "true" : "false")
      << ";\nconstexpr int NPF = " << std::max<size_t>(lib.slots.size(), 1)
      << ";\nconstexpr int U = " << unroll << ";\nconstexpr bool VEC = " << (vec ? "true" : "false")
      << ";\nconstexpr int NSHAPE = " << lib.bodies.size() << ";\nconstexpr int NMASK = " << lib.mask_slots.size()
      << ";\nconstexpr int MSLOT[" << std::max<size_t>(lib.mask_slots.size(), 1) << "] = {";
    for (size_t m = 0; m < lib.mask_slots.size(); m++) s << (m ? ", " : "") << lib.mask_slots[m];
    if (lib.mask_slots.empty()) s << "0";
    s << "};\n" << BATCH_HEAD;
    for (size_t k = 0; k < lib.bodies.size(); k++)
        s << "template <int R, int UU>\n__device__ __forceinline__ void body" << k
          << "(const FProgram &P, const int64_t (&row0)[UU], const Pf<R, UU> &pf, int64_t &acc, uint8_t *lds, "
             "const uint32_t (&act)[UU])\n{\n"
          << lib.bodies[k] << "}\n";
    s << "template <int R, int UU>\n__device__ __forceinline__ void run_all(const BatchArgs &A, const int64_t "
         "(&row0)[UU], const Pf<R, UU> &pf, uint8_t *lds, unsigned long long *sums, uint32_t dmask)\n{\n"
         "    uint32_t mw[UU][R];\n    mask_rows<R, UU>(A, pf, lds, dmask, mw);\n    uint32_t any = 0;\n"
         "#pragma unroll\n    for (int u = 0; u < UU; u++)\n#pragma unroll\n"
         "        for (int k = 0; k < R; k++) any |= mw[u][k];\n"
         "    // programs some row of the wave keeps (read from the first lane so\n"
         "    // the compiler sees one value for the wave)\n"
         "    const uint32_t wave = __builtin_amdgcn_readfirstlane(__ockl_wfred_or_u32(any));\n";
    for (size_t k = 0; k < lib.bodies.size(); k++)
        s << "#pragma unroll 1\n    for (int i = A.first[" << k << "]; i < A.first[" << k + 1 << "]; i++) {\n"
          << "        const int bit = A.bitpos[i];\n"
          << "        if (!((wave >> bit) & 1u)) continue;\n"
          << "        uint32_t act[UU];\n        rows_of_program<R, UU>(mw, bit, act);\n"
          << "        const FProgram &P = ((ProgPtr)A.progs)[i];\n        int64_t acc = 0;\n"
          << "        body" << k << "<R, UU>(P, row0, pf, acc, lds + P.lds_base, act);\n"
          << "        if (acc) atomicAdd(&sums[i], (unsigned long long)acc);\n    }\n";
    s << "}\n} // namespace\n" << BATCH_TAIL;
    return s.str();
}

// The batch's mask tables as the autotuner tells batches apart: per masked
// column (in slot order) its slot, entry bytes and table size (a power of
// two), and the lazy form.
std::string mask_signature(const BatchPlan &b)
{
    std::vector<const MaskPlan *> ms;
    for (const MaskPlan &m : b.masks) ms.push_back(&m);
    std::sort(ms.begin(), ms.end(), [](const MaskPlan *x, const MaskPlan *y) { return x->slot < y->slot; });
    std::ostringstream o;
    for (const MaskPlan *m : ms) {
        int lg = 0;
        while ((1ull << lg) < m->bytes()) lg++;
        o << m->slot << ":" << m->wbytes << "B:2^" << lg << " ";
    }
    o << (b.lazy ? "lazy" : "eager");
    return o.str();
}

// SDB_BATCH_LAZY_MASKS: 0 reads every mask table for every row, 1 always
// reads them lazily, unset decides by the estimate below.
int batch_lazy_mode()
{
    const char *e = getenv("SDB_BATCH_LAZY_MASKS");
    return e && *e ? (*e == '1' ? 1 : 0) : -1;
}

// Reading order of the mask tables and whether to read them lazily.  From
// the host's estimate of the share of rows each probe keeps (the share of
// its dimension's rows the build side selected), with programs and columns
// taken as independent: a row still read for column j is one that some
// program keeps after the columns before j.  Greedily, the next column is
// the one after which the fewest rows stay kept; the reads per row that
// order saves decide the lazy form (its reads wait on the previous ones, so
// it has to save a quarter of them).
void order_masks(BatchPlan &b, const std::vector<FProgram> &q, const std::vector<int> &order,
                 const std::vector<std::vector<int>> &shared)
{
    const int nm = (int)b.masks.size(), np = (int)order.size();
    b.morder.assign(std::max<size_t>(b.mask_slots.size(), 1), -1);
    b.lazy = false;
    if (nm == 0) return;
    // keep[c][pos]: share of rows program pos keeps through column c's probes
    std::vector<std::vector<double>> keep(nm, std::vector<double>(np, 1.0));
    for (int pos = 0; pos < np; pos++)
        for (int j : shared[order[pos]]) {
            const FInstr &I = q[order[pos]].ins[j];
            const int slot = I.va - F_PF_BASE;
            for (int c = 0; c < nm; c++)
                if (b.masks[c].slot == slot) keep[c][pos] *= std::min(1.0, std::max(0.0, (double)I.sel));
        }
    std::vector<double> alive(np, 1.0);
    std::vector<bool> used(nm, false);
    double reads = 0, still = 1.0;
    for (int step = 0; step < nm; step++) {
        int best = -1;
        double best_after = 2.0;
        for (int c = 0; c < nm; c++) {
            if (used[c]) continue;
            double none = 1.0;
            for (int pos = 0; pos < np; pos++) none *= 1.0 - alive[pos] * keep[c][pos];
            if (1.0 - none < best_after) {
                best_after = 1.0 - none;
                best = c;
            }
        }
        used[best] = true;
        reads += still;
        still = best_after;
        for (int pos = 0; pos < np; pos++) alive[pos] *= keep[best][pos];
        b.morder[step] = (int)(std::find(b.mask_slots.begin(), b.mask_slots.end(), b.masks[best].slot) -
                               b.mask_slots.begin());
    }
    const int mode = batch_lazy_mode();
    b.lazy = mode == 1 || (mode < 0 && reads <= 0.75 * nm);
}

BatchPlan plan_batch(const std::vector<FProgram> &in, int unroll, const std::string &lib_key, bool masks,
                     const BatchShape &shape)
{
    BatchPlan b;
    if (in.size() > 32) fail(SDB_ERR_UNSUPPORTED, "batched launch: more than 32 programs");
    std::vector<FProgram> q = in;
    for (const FProgram &p : q) b.vec = b.vec && p.vec;
    std::lock_guard<std::mutex> lk(g_lib_mu);
    BatchLibrary &lib = g_libs[lib_key + (b.vec ? "/vec" : "/row") + "/u" + std::to_string(unroll)];
    std::vector<int> shape_of(q.size());
    std::vector<std::vector<int>> shared(q.size());
    for (size_t i = 0; i < q.size(); i++) {
        remap_prefetch(q[i], lib.slots);
        if (masks) shared[i] = shared_probes(q[i]);
        Gen g(q[i], false);
        g.batch_act = true;
        g.decided.assign(q[i].ninstr, false);
        for (int j : shared[i]) {
            const FInstr &I = q[i].ins[j];
            g.decided[j] = true;
            const int slot = I.va - F_PF_BASE;
            if (std::find(lib.mask_slots.begin(), lib.mask_slots.end(), slot) == lib.mask_slots.end())
                lib.mask_slots.push_back(slot);
        }
        // tables staged only for the probes the masks decided stay unloaded
        for (int bl = 0; bl < q[i].nblobs; bl++) {
            bool used = false;
            for (int j = 0; j < q[i].ninstr; j++) {
                const FInstr &I = q[i].ins[j];
                used |= !g.decided[j] && I.in_lds && I.lds_off == q[i].blob_off[bl];
            }
            if (!used) q[i].blob_bytes[bl] = 0;
        }
        // a body depends on the program's shape and on which probes the
        // masks decided
        std::string body_key = shape_key(q[i]);
        for (int j = 0; j < q[i].ninstr; j++) body_key.push_back(g.decided[j] ? 1 : 0);
        auto cached = lib.body_of.find(body_key);
        if (cached == lib.body_of.end()) cached = lib.body_of.emplace(body_key, g.body()).first;
        const std::string &body = cached->second;
        auto it = std::find(lib.bodies.begin(), lib.bodies.end(), body);
        if (it == lib.bodies.end()) {
            if ((int)lib.bodies.size() >= BATCH_MAX_SHAPES) fail(SDB_ERR_UNSUPPORTED, "batched launch: too many program shapes");
            lib.bodies.push_back(body);
            it = lib.bodies.end() - 1;
        }
        shape_of[i] = (int)(it - lib.bodies.begin());
    }
    // programs in shape order; a program's bit in the masks is its position
    std::vector<int> order;
    for (int k = 0; k < (int)lib.bodies.size(); k++) {
        b.first.push_back((int)b.progs.size());
        for (size_t i = 0; i < q.size(); i++)
            if (shape_of[i] == k) {
                b.progs.push_back(q[i]);
                order.push_back((int)i);
            }
    }
    b.first.push_back((int)b.progs.size());
    b.full = b.progs.size() >= 32 ? ~0u : (1u << b.progs.size()) - 1u;
    b.mask_slots = lib.mask_slots;
    // Bit positions: programs ordered by which masked columns they probe
    // (the set as a Gray code, columns probed by more programs first), so
    // that each column's probing programs take nearby bits and its table
    // stores a narrow window of the word.
    {
        std::vector<int> mslot_count(lib.mask_slots.size(), 0);
        std::vector<uint32_t> sig(order.size(), 0);
        for (size_t pos = 0; pos < order.size(); pos++)
            for (int j : shared[order[pos]]) {
                const int slot = q[order[pos]].ins[j].va - F_PF_BASE;
                const int m = (int)(std::find(lib.mask_slots.begin(), lib.mask_slots.end(), slot) - lib.mask_slots.begin());
                if (!((sig[pos] >> m) & 1u)) mslot_count[m]++;
                sig[pos] |= 1u << m;
            }
        std::vector<int> rank_of(lib.mask_slots.size());
        {
            std::vector<int> by(lib.mask_slots.size());
            for (size_t m = 0; m < by.size(); m++) by[m] = (int)m;
            std::stable_sort(by.begin(), by.end(), [&](int a, int c) { return mslot_count[a] > mslot_count[c]; });
            for (size_t r = 0; r < by.size(); r++) rank_of[by[r]] = (int)r; // most probed column: top bit
        }
        const int nm = (int)lib.mask_slots.size();
        std::vector<std::pair<uint32_t, int>> key(order.size());
        for (size_t pos = 0; pos < order.size(); pos++) {
            uint32_t g = 0; // signature with the most probed column as the top bit
            for (int m = 0; m < nm; m++)
                if ((sig[pos] >> m) & 1u) g |= 1u << (nm - 1 - rank_of[m]);
            uint32_t bin = g; // Gray decode
            for (uint32_t sh = 1; sh < 32; sh <<= 1) bin ^= bin >> sh;
            key[pos] = {bin, (int)pos};
        }
        std::stable_sort(key.begin(), key.end());
        b.bitpos.assign(order.size(), 0);
        for (size_t r = 0; r < key.size(); r++) b.bitpos[key[r].second] = (int)r;
    }
    for (size_t pos = 0; pos < order.size(); pos++)
        for (int j : shared[order[pos]]) {
            const FInstr &I = q[order[pos]].ins[j];
            const int slot = I.va - F_PF_BASE;
            auto it = std::find_if(b.masks.begin(), b.masks.end(), [&](const MaskPlan &m) { return m.slot == slot; });
            if (it == b.masks.end()) {
                b.masks.push_back(MaskPlan{slot});
                it = b.masks.end() - 1;
            }
            if (I.ht_min <= I.ht_max) {
                if (it->lo > it->hi) {
                    it->lo = I.ht_min;
                    it->hi = I.ht_max;
                } else {
                    it->lo = std::min<int64_t>(it->lo, I.ht_min);
                    it->hi = std::max<int64_t>(it->hi, I.ht_max);
                }
            }
            it->ents.push_back(FMaskEntry{I.ht, I.ht_min, I.ht_max, (uint32_t)b.bitpos[pos], mask_kind(I)});
        }
    for (MaskPlan &m : b.masks) {
        if (m.hi - m.lo + 1 > F_MASK_MAX_LEN)
            fail(SDB_ERR_UNSUPPORTED, "batched launch: the probed key ranges of one column span more than 2^24 keys");
        // the narrowest window (8, 16 or 32 bits) holding every probing bit
        uint32_t lo_bit = 31, hi_bit = 0;
        for (const FMaskEntry &e : m.ents) {
            lo_bit = std::min(lo_bit, e.bit);
            hi_bit = std::max(hi_bit, e.bit);
        }
        const uint32_t span = hi_bit - lo_bit + 1;
        m.wbytes = shape.wide ? 4 : span <= 8 ? 1 : span <= 16 ? 2 : 4;
        m.shift = m.wbytes == 4 ? 0 : std::min(lo_bit, 32 - 8 * m.wbytes);
    }
    if ((int)b.masks.size() > F_MASK_MAX_COLS) fail(SDB_ERR_UNSUPPORTED, "batched launch: too many masked columns");
    order_masks(b, q, order, shared);
    // columns this batch reads (other slots stay unloaded)
    b.pf.assign(std::max<size_t>(lib.slots.size(), 1), nullptr);
    auto use = [&](int8_t reg) {
        if (reg >= F_PF_BASE) b.pf[reg - F_PF_BASE] = lib.slots[reg - F_PF_BASE];
    };
    for (const FProgram &p : b.progs) {
        for (int i = 0; i < p.ninstr; i++) {
            use(p.ins[i].va);
            use(p.ins[i].vb);
        }
        for (int j = 0; j < p.gb.k; j++) use(p.gb.key_reg[j]);
        use(p.gb.agg_reg);
    }
    // a masked slot of the library this batch does not mask is skipped by
    // the kernel (its mlen is 0)
    uint32_t off = 0;
    for (FProgram &p : b.progs) {
        p.lds_base = off;
        off += (p.vreg_off + 15u) & ~15u;
    }
    b.sums_off = off;
    b.lds = off + 8u * (uint32_t)b.progs.size();
    if (shape.stage) {
        // the smallest mask tables into LDS, while the block's LDS stays
        // within BATCH_STAGE_LDS_MAX
        std::vector<MaskPlan *> by;
        for (MaskPlan &m : b.masks) by.push_back(&m);
        std::stable_sort(by.begin(), by.end(), [](const MaskPlan *a, const MaskPlan *c) { return a->bytes() < c->bytes(); });
        uint32_t at = (b.lds + 15u) & ~15u;
        for (MaskPlan *m : by) {
            if ((uint64_t)at + m->bytes() > BATCH_STAGE_LDS_MAX) break;
            m->lds = at + 1u;
            at += (uint32_t)m->bytes();
        }
        if (at > b.lds) b.lds = at;
    }
    if (shape.dict) {
        // tables over a long, sparse key range as rank dictionaries in LDS,
        // while the block's LDS stays within BATCH_STAGE_LDS_MAX
        uint32_t at = (b.lds + 15u) & ~15u;
        for (MaskPlan &m : b.masks) {
            const int64_t len = m.hi - m.lo + 1;
            if (len < (int64_t)batch_dict_min_len() || len > (int64_t)F_DICT_MAX_LEN) continue;
            const uint32_t need = dict_lds_bytes((uint32_t)len);
            if ((uint64_t)at + need > BATCH_STAGE_LDS_MAX) continue;
            m.dict = at + 1u;
            at += need;
        }
        if (at > b.lds) b.lds = at;
    }
    BatchLibrary::Src &src = lib.srcs[shape.id()];
    if (src.state[0] != lib.bodies.size() || src.state[1] != lib.slots.size() ||
        src.state[2] != lib.mask_slots.size()) {
        src.text = batch_source(lib, unroll, b.vec, shape);
        src.state[0] = lib.bodies.size();
        src.state[1] = lib.slots.size();
        src.state[2] = lib.mask_slots.size();
        src.version++;
        src.kernel.clear();
    }
    b.src_entry = &src;
    b.version = src.version;
    b.src = src.text;
    return b;
}

} // namespace

int jit_batch_check(const std::string &arch)
{
    std::vector<FProgram> progs;
    for (auto &pu : coverage_programs()) progs.push_back(pu.first);
    // two copies of each shape, as pipelines with different constants
    std::vector<FProgram> two = progs;
    two.insert(two.end(), progs.begin(), progs.end());
    compile_source(plan_batch(two, 2, "check:" + arch, true, BATCH_CANDIDATES[0]).src, arch);
    return 1;
}

// ------------------------------------------------- batched launch and graphs
//
// Everything one batched launch needs in device memory -- the programs, the
// mask descriptors and entries, and one all-ones word -- goes to the device
// in one copy, into one block that also holds the mask tables; the kernel's
// argument block (pointers into that block, the fact columns, the programs'
// bit positions) is its kernel argument.  So a launch is three stream
// operations: copy, mask build, batched kernel.
//
// With graphs on, the block and its host staging are kept per shape library
// and launch shape, at fixed addresses.  A launch met a second time (same
// geometry, same kernel arguments) is captured into a hipGraph, which later
// launches of it replay: only the staging's contents, copied by the graph,
// change (the programs' literals, for instance).  The staging is rewritten
// only after the previous launch's copy has run (an event).  A launch met
// once runs as plain stream operations, so one-off batches pay no
// instantiation.  The graphs hold no memset node.

namespace {

struct BatchGraph {
    std::mutex mu;
    void *dev = nullptr;
    size_t dev_bytes = 0;
    uint8_t *pinned = nullptr;
    size_t pinned_bytes = 0;
    hipEvent_t done = nullptr;  // after the last replay's copy
    // captured graphs by launch geometry (the most recent first, at most
    // MAX_EXECS: a segmented table alternates a full and a last segment)
    std::vector<std::pair<std::string, hipGraphExec_t>> execs;
    std::vector<std::string> seen; // launches met once, not captured (the last 8)
    int64_t replays = 0, captures = 0;
    static constexpr size_t MAX_EXECS = 4;
    void drop_execs()
    {
        for (auto &e : execs) (void)hipGraphExecDestroy(e.second);
        execs.clear();
    }
};
std::mutex g_graph_mu;
std::map<std::string, std::unique_ptr<BatchGraph>> g_graphs;
// -1: SDB_BATCH_GRAPH decides (default off: on the 32-pipeline batch over
// 600M rows the replayed graph measured 21.2 ms against 20.0 ms for the
// same three operations issued on the stream)
std::atomic<int> g_batch_graphs{-1};

bool batch_graphs_on()
{
    int v = g_batch_graphs.load();
    if (v < 0) {
        const char *e = getenv("SDB_BATCH_GRAPH");
        v = (e && *e == '1') ? 1 : 0;
        g_batch_graphs.store(v);
    }
    return v == 1;
}

size_t al256(size_t b) { return (b + 255) / 256 * 256; }

} // namespace

void set_batch_graphs(int on) { g_batch_graphs.store(on < 0 ? -1 : (on ? 1 : 0)); }

void batch_graph_stats(int64_t *captures, int64_t *replays)
{
    std::lock_guard<std::mutex> lk(g_graph_mu);
    int64_t c = 0, r = 0;
    for (auto &kv : g_graphs) {
        std::lock_guard<std::mutex> gl(kv.second->mu);
        c += kv.second->captures;
        r += kv.second->replays;
    }
    if (captures) *captures = c;
    if (replays) *replays = r;
}

void launch_fused_batch(const std::vector<FProgram> &in, int64_t n, hipStream_t s, int unroll, const void *scope,
                        BatchAlloc alloc, void *alloc_ctx, bool masks)
{
    if (n <= 0 || in.empty()) return;
    // the shared scan holds up to BATCH_MAX_PF columns in registers, twice
    // (current and next tile): callers pass 1 (one 4-row chunk per thread)
    unroll &= ~SDB_UNROLL_QUEUED; // batched bodies have no queued tails
    if (unroll != 1 && unroll != 2 && unroll != 4) unroll = 1;
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    char scope_key[32];
    snprintf(scope_key, sizeof(scope_key), "%p", scope);
    const std::string lib_key = device_arch() + "/" + std::to_string(dev) + "/" + scope_key + (masks ? "/m" : "");
    // launch shape: fixed by the environment, else the autotuner's choice
    // for this library (trials on launches of at least TUNE_MIN_ROWS rows)
    // for this library and the batch's mask tables (their columns, entry
    // widths and sizes: what the layouts trade)
    BatchShape shape = BATCH_CANDIDATES[0];
    const bool fixed = batch_shape_fixed(shape);
    BatchPlan b = plan_batch(in, unroll, lib_key, masks, shape);
    int trial = -1;
    std::string tune_key;
    if (!fixed && n >= TUNE_MIN_ROWS) {
        tune_key = lib_key + " masks " + mask_signature(b);
        BatchShape want = shape;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_batch_tune.find(tune_key);
            if (it != g_batch_tune.end() && it->second.chosen >= 0) want = BATCH_CANDIDATES[it->second.chosen];
            else if (t_batch_trials) {
                int kinds = 0;
                const std::string prefix = lib_key + " masks ";
                for (auto &kv : g_batch_tune) kinds += kv.first.compare(0, prefix.size(), prefix) == 0;
                if (it != g_batch_tune.end() || kinds < BATCH_TUNE_MAX_KINDS) {
                    BatchTune &t = g_batch_tune[tune_key];
                    trial = t.trials % N_BATCH_CAND;
                    want = BATCH_CANDIDATES[trial];
                }
            }
        }
        if (want.key() != shape.key()) b = plan_batch(in, unroll, lib_key, masks, want);
        shape = want;
    }
    static_assert(N_BATCH_CAND <= 16, "BatchTune::best holds 16 candidates");
    if (b.lds > std::max<uint64_t>(F_BATCH_LDS + 8u * F_BATCH_MAX, BATCH_STAGE_LDS_MAX))
        fail(SDB_ERR_UNSUPPORTED, "batched launch: LDS regions exceed the budget");
    Compiled *kp = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_lib_mu);
        auto it = b.src_entry->kernel.find(dev);
        if (it != b.src_entry->kernel.end() && b.src_entry->version == b.version) kp = it->second;
    }
    if (!kp) {
        kp = &get_kernel(b.src, "sdb_fused_batch");
        std::lock_guard<std::mutex> lk(g_lib_mu);
        if (b.src_entry->version == b.version) b.src_entry->kernel[dev] = kp;
    }
    Compiled &k = *kp;

    // ---- the block: [args | programs | mask descriptors | entries | ones word, tables]
    const int nmask = (int)b.mask_slots.size();
    const int nshape = (int)b.first.size() - 1;
    size_t nents = 0, tab_bytes = 16; // the first 16 bytes: the all-ones word
    int ndicts = 0;
    for (const MaskPlan &m : b.masks) {
        nents += m.ents.size();
        tab_bytes += m.bytes(); // padded to 16 bytes, so a staged table copies in uint4s
    }
    for (const MaskPlan &m : b.masks) // rank dictionaries after the tables
        if (m.dict) {
            tab_bytes += 16 + dict_lds_bytes((uint32_t)(m.hi - m.lo + 1));
            ndicts++;
        }
    // argument block (the kernel's BatchArgs): progs, n, nprog, sums_off,
    // full, lazy, first[NSHAPE+1], pf[NPF], mtab[NMASK], mdg[NMASK],
    // mlo[NMASK], mlen[NMASK], mdflt[NMASK], mshift[NMASK], mwbytes[NMASK],
    // mlds[NMASK], mdict[NMASK], morder[NMASK], bitpos[32]
    size_t args_bytes = 0;
    {
        auto grow = [&](size_t bytes, size_t align) { args_bytes = (args_bytes + align - 1) / align * align + bytes; };
        grow(8, 8), grow(8, 8), grow(4, 4), grow(4, 4), grow(4, 4), grow(4, 4);
        for (int i = 0; i <= nshape; i++) grow(4, 4);
        for (size_t i = 0; i < b.pf.size(); i++) grow(8, 8);
        for (int f = 0; f < 2; f++)
            for (int i = 0; i < std::max(nmask, 1); i++) grow(8, 8);
        for (int f = 0; f < 8; f++)
            for (int i = 0; i < std::max(nmask, 1); i++) grow(4, 4);
        for (int i = 0; i < 32; i++) grow(4, 4);
        args_bytes = (args_bytes + 7) / 8 * 8;
    }
    const size_t off_progs = 0;
    const size_t off_cols = off_progs + al256(b.progs.size() * sizeof(FProgram));
    const size_t off_ents = off_cols + al256(std::max<size_t>(b.masks.size(), 1) * sizeof(FMaskCol));
    const size_t off_dicts = off_ents + al256(std::max<size_t>(nents, 1) * sizeof(FMaskEntry));
    const size_t off_tabs = off_dicts + al256((size_t)std::max(ndicts, 1) * sizeof(FMaskDict));
    const size_t copy_bytes = off_tabs + 16, block_bytes = off_tabs + tab_bytes;

    // tuning trials and graphs off: the block from the caller's allocator
    BatchGraph *G = nullptr;
    if (trial < 0 && batch_graphs_on()) {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        auto &slot = g_graphs[lib_key + "#" + std::to_string(shape.key())];
        if (!slot) slot.reset(new BatchGraph);
        G = slot.get();
    }
    std::unique_lock<std::mutex> glk;
    uint8_t *dblock = nullptr;
    std::vector<uint8_t> host_vec;
    uint8_t *host = nullptr;
    if (G) {
        glk = std::unique_lock<std::mutex>(G->mu);
        if (G->done) SDB_HIP(hipEventSynchronize(G->done)); // the staging's last copy has run
        if (G->dev_bytes < block_bytes || G->pinned_bytes < copy_bytes) {
            G->drop_execs();
            G->seen.clear();
            if (G->dev) (void)hipFree(G->dev);
            if (G->pinned) (void)hipHostFree(G->pinned);
            G->dev = nullptr;
            G->pinned = nullptr;
            G->dev_bytes = G->pinned_bytes = 0;
            const size_t db = std::max(block_bytes * 5 / 4, (size_t)1 << 20);
            const size_t hb = std::max(copy_bytes * 5 / 4, (size_t)1 << 16);
            SDB_HIP(hipMalloc(&G->dev, db));
            G->dev_bytes = db;
            SDB_HIP(hipHostMalloc(reinterpret_cast<void **>(&G->pinned), hb, hipHostMallocDefault));
            G->pinned_bytes = hb;
        }
        if (!G->done) SDB_HIP(hipEventCreateWithFlags(&G->done, hipEventDisableTiming));
        dblock = static_cast<uint8_t *>(G->dev);
        host = G->pinned;
    } else {
        dblock = static_cast<uint8_t *>(alloc(alloc_ctx, block_bytes));
        host_vec.assign(copy_bytes, 0);
        host = host_vec.data();
    }
    std::memset(host, 0, copy_bytes);

    // programs, masks, then the argument block, written into the staging
    std::memcpy(host + off_progs, b.progs.data(), b.progs.size() * sizeof(FProgram));
    FMaskCol *d_cols = reinterpret_cast<FMaskCol *>(dblock + off_cols);
    FMaskEntry *d_ents = reinterpret_cast<FMaskEntry *>(dblock + off_ents);
    FMaskDict *d_dicts = reinterpret_cast<FMaskDict *>(dblock + off_dicts);
    uint8_t *tabs = dblock + off_tabs;
    std::vector<const uint8_t *> mtab(std::max(nmask, 1), tabs); // unused slots (mlen 0) are skipped
    std::vector<int32_t> mlo(std::max(nmask, 1), 0);
    std::vector<uint32_t> mlen(std::max(nmask, 1), 0), mdflt(std::max(nmask, 1), ~0u); // mlen 0: unused
    std::vector<uint32_t> mshift(std::max(nmask, 1), 0), mwbytes(std::max(nmask, 1), 4);
    std::vector<uint32_t> mlds(std::max(nmask, 1), 0), mdict(std::max(nmask, 1), 0);
    std::vector<const uint8_t *> mdg(std::max(nmask, 1), nullptr);
    uint32_t max_len = 0;
    {
        size_t at = 16, ei = 0;
        for (size_t ci = 0; ci < b.masks.size(); ci++) {
            const MaskPlan &m = b.masks[ci];
            FMaskCol c{};
            const int64_t len = std::max<int64_t>(m.hi - m.lo + 1, 1);
            c.tab = tabs + at;
            c.lo = m.lo <= m.hi ? (int32_t)m.lo : 0;
            c.len = (uint32_t)len;
            c.full = b.full;
            c.first = (uint32_t)ei;
            c.count = (uint32_t)m.ents.size();
            c.shift = m.shift;
            c.wbytes = m.wbytes;
            uint32_t probing = 0;
            for (const FMaskEntry &e : m.ents) {
                std::memcpy(host + off_ents + ei * sizeof(FMaskEntry), &e, sizeof(FMaskEntry));
                ei++;
                probing |= 1u << e.bit;
            }
            std::memcpy(host + off_cols + ci * sizeof(FMaskCol), &c, sizeof(FMaskCol));
            max_len = std::max(max_len, c.len);
            const int mi = (int)(std::find(b.mask_slots.begin(), b.mask_slots.end(), m.slot) - b.mask_slots.begin());
            mtab[mi] = c.tab;
            mlo[mi] = c.lo;
            mlen[mi] = c.len;
            mdflt[mi] = b.full & ~probing;
            mshift[mi] = c.shift;
            mwbytes[mi] = c.wbytes;
            mlds[mi] = m.lds;
            at += m.bytes();
        }
        // the rank dictionaries' global buffers, after the tables
        int di = 0;
        for (size_t ci = 0; ci < b.masks.size(); ci++) {
            const MaskPlan &m = b.masks[ci];
            if (!m.dict) continue;
            const int mi = (int)(std::find(b.mask_slots.begin(), b.mask_slots.end(), m.slot) - b.mask_slots.begin());
            FMaskDict d{};
            d.tab = mtab[mi];
            d.dict = tabs + at;
            d.len = (uint32_t)(m.hi - m.lo + 1);
            d.wbytes = m.wbytes;
            std::memcpy(host + off_dicts + di * sizeof(FMaskDict), &d, sizeof(FMaskDict));
            di++;
            mdict[mi] = m.dict;
            mdg[mi] = d.dict;
            at += 16 + dict_lds_bytes(d.len);
        }
        const uint32_t ones[4] = {~0u, ~0u, ~0u, ~0u};
        std::memcpy(host + off_tabs, ones, 16);
    }
    std::vector<uint8_t> args(args_bytes, 0);
    {
        size_t at = 0;
        auto put = [&](const void *v, size_t bytes, size_t align) {
            at = (at + align - 1) / align * align;
            std::memcpy(args.data() + at, v, bytes);
            at += bytes;
        };
        const FProgram *pp = reinterpret_cast<const FProgram *>(dblock + off_progs);
        const int32_t nprog = (int32_t)b.progs.size();
        put(&pp, 8, 8);
        put(&n, 8, 8);
        put(&nprog, 4, 4);
        put(&b.sums_off, 4, 4);
        put(&b.full, 4, 4);
        const uint32_t lazy = b.lazy ? 1u : 0u;
        put(&lazy, 4, 4);
        for (int i = 0; i <= nshape; i++) {
            const int32_t f = b.first[i];
            put(&f, 4, 4);
        }
        for (const int32_t *c : b.pf) put(&c, 8, 8);
        for (const uint8_t *t : mtab) put(&t, 8, 8);
        for (const uint8_t *t : mdg) put(&t, 8, 8);
        for (int32_t v : mlo) put(&v, 4, 4);
        for (uint32_t v : mlen) put(&v, 4, 4);
        for (uint32_t v : mdflt) put(&v, 4, 4);
        for (uint32_t v : mshift) put(&v, 4, 4);
        for (uint32_t v : mwbytes) put(&v, 4, 4);
        for (uint32_t v : mlds) put(&v, 4, 4);
        for (uint32_t v : mdict) put(&v, 4, 4);
        for (int i = 0; i < std::max(nmask, 1); i++) {
            const int32_t m = i < (int)b.morder.size() ? b.morder[i] : -1;
            put(&m, 4, 4);
        }
        for (int i = 0; i < 32; i++) {
            const int32_t bp = i < (int)b.bitpos.size() ? b.bitpos[i] : i;
            put(&bp, 4, 4);
        }
        if (at > args_bytes) fail(SDB_ERR_RUNTIME, "batched launch: argument block overflow");
    }

    int per_cu;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = k.blocks_per_cu.find(b.lds);
        if (it == k.blocks_per_cu.end()) {
            int bl = 0;
            SDB_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&bl, k.fn, shape.block, b.lds));
            it = k.blocks_per_cu.emplace(b.lds, std::max(1, bl)).first;
        }
        per_cu = it->second;
    }
    // a share of each CU's wave slots, where the caller asks for one (other
    // engines' kernels then run beside this launch)
    if (t_batch_waves_cap > 0) per_cu = std::max(1, std::min(per_cu, t_batch_waves_cap / (shape.block / 64)));
    const int64_t units = b.vec ? (n + 3) / 4 : n;
    // at least BATCH_MIN_TILES tiles per block: each block also stages and
    // flushes every program's LDS state, which small tables must amortise
    const int64_t tiles = (units + (int64_t)shape.block * (b.vec ? unroll : 1) - 1) / ((int64_t)shape.block * (b.vec ? unroll : 1));
    const unsigned grid = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>(stream_grid(units, (int64_t)shape.block * (b.vec ? unroll : 1), per_cu),
                             (tiles + BATCH_MIN_TILES - 1) / BATCH_MIN_TILES));
    size_t args_size = args.size();
    void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, args.data(), HIP_LAUNCH_PARAM_BUFFER_SIZE, &args_size,
                      HIP_LAUNCH_PARAM_END};
    auto enqueue = [&]() {
        SDB_HIP(hipMemcpyAsync(dblock, host, copy_bytes, hipMemcpyHostToDevice, s));
        if (nmask > 0 && !b.masks.empty()) launch_batch_masks(d_cols, (int)b.masks.size(), max_len, d_ents, s);
        launch_batch_dict(d_dicts, ndicts, s);
        SDB_HIP(hipModuleLaunchKernel(k.fn, grid, 1, 1, shape.block, 1, 1, b.lds, s, nullptr, config));
    };
    static const bool debug = [] {
        const char *e = getenv("SDB_DEBUG_BATCH");
        return e && *e && *e != '0';
    }();
    if (trial < 0 && debug) {
        fprintf(stderr, "[batch] launch: %zu programs, block %d, grid %u (%d blocks/CU), LDS %u bytes, graph %d, %s masks",
                b.progs.size(), shape.block, grid, per_cu, b.lds, G != nullptr, b.lazy ? "lazy" : "eager");
        for (int m : b.morder)
            if (m >= 0) fprintf(stderr, " %d", b.mask_slots[m]);
        fprintf(stderr, "; mask windows");
        for (const MaskPlan &m : b.masks)
            fprintf(stderr, " [slot %d: %lld keys x %u B at bit %u%s]", m.slot, (long long)(m.hi - m.lo + 1), m.wbytes,
                    m.shift, m.lds ? ", in LDS" : m.dict ? ", rank dictionary in LDS" : "");
        fprintf(stderr, "\n");
    }
    if (G) {
        // the captured kernel node holds the argument buffer by value, so a
        // graph serves launches with the same geometry and the same
        // arguments; the programs and mask descriptors travel in the copy
        char head[256];
        snprintf(head, sizeof(head), "%p/%u/%d/%u/%zu/%u/%zu/%p/%p/", (void *)k.fn, grid, shape.block, b.lds, copy_bytes,
                 max_len, b.masks.size(), (void *)dblock, (void *)host);
        std::string topo(head);
        topo.append(reinterpret_cast<const char *>(args.data()), args.size());
        auto hit = std::find_if(G->execs.begin(), G->execs.end(), [&](const std::pair<std::string, hipGraphExec_t> &e) {
            return e.first == topo;
        });
        if (hit == G->execs.end() && std::find(G->seen.begin(), G->seen.end(), topo) == G->seen.end()) {
            // a launch seen for the first time runs as stream operations
            // (from the same buffers); it is captured when it comes again
            if (G->seen.size() >= 8) G->seen.erase(G->seen.begin());
            G->seen.push_back(topo);
            enqueue();
            SDB_HIP(hipEventRecord(G->done, s));
            return;
        }
        if (hit == G->execs.end()) {
            SDB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                enqueue();
            } catch (...) {
                hipGraph_t dead = nullptr;
                (void)hipStreamEndCapture(s, &dead);
                if (dead) (void)hipGraphDestroy(dead);
                throw;
            }
            hipGraph_t graph = nullptr;
            SDB_HIP(hipStreamEndCapture(s, &graph));
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            hip_check(e, "hipGraphInstantiate (batched launch)");
            if (G->execs.size() >= BatchGraph::MAX_EXECS) {
                (void)hipGraphExecDestroy(G->execs.back().second);
                G->execs.pop_back();
            }
            G->execs.insert(G->execs.begin(), {topo, exec});
            hit = G->execs.begin();
            G->captures++;
        }
        const hipGraphExec_t exec = hit->second;
        SDB_HIP(hipGraphLaunch(exec, s));
        SDB_HIP(hipEventRecord(G->done, s));
        G->replays++;
        return;
    }
    if (trial < 0) {
        enqueue();
        return;
    }
    // a tuning trial: the kernel alone, timed with HIP events; compared per
    // row and program, since batches of one library differ in size
    SDB_HIP(hipMemcpyAsync(dblock, host, copy_bytes, hipMemcpyHostToDevice, s));
    if (nmask > 0 && !b.masks.empty()) launch_batch_masks(d_cols, (int)b.masks.size(), max_len, d_ents, s);
    launch_batch_dict(d_dicts, ndicts, s);
    hipEvent_t ea, eb;
    SDB_HIP(hipEventCreate(&ea));
    SDB_HIP(hipEventCreate(&eb));
    float ms = 0;
    hipError_t e = hipEventRecord(ea, s);
    if (e == hipSuccess) e = hipModuleLaunchKernel(k.fn, grid, 1, 1, shape.block, 1, 1, b.lds, s, nullptr, config);
    if (e == hipSuccess) e = hipEventRecord(eb, s);
    if (e == hipSuccess) e = hipEventSynchronize(eb);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ea, eb);
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    hip_check(e, "batched kernel tuning");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    BatchTune &t = g_batch_tune[tune_key];
    if (t.chosen >= 0) return;
    const double cost = ms * 1e6 / ((double)n * (double)b.progs.size());
    t.best[trial] = std::min(t.best[trial], cost);
    if (++t.trials >= N_BATCH_CAND) {
        int best = 0;
        for (int c = 1; c < N_BATCH_CAND; c++)
            if (t.best[c] < t.best[best]) best = c;
        t.chosen = best;
    }
}

void set_batch_trials(bool on) { t_batch_trials = on; }
void set_batch_waves_per_cu(int cap) { t_batch_waves_cap = cap; }

int jit_pending()
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    int n = 0;
    for (auto &kv : g_tune) {
        const Tune &t = kv.second;
        if (!t.error.empty()) continue;
        const bool compiling = (t.queued_default && !t.default_ready) || (t.queued_all && t.cands_ready < N_CAND - 1) ||
                               (t.queued_tail && !t.tail_ready);
        n += compiling || (t.queued_all && t.chosen < 0);
    }
    for (auto &kv : g_batch_tune) n += kv.second.chosen < 0;
    return n;
}

std::string jit_tuning_report()
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::ostringstream o;
    for (auto &kv : g_batch_tune) {
        const BatchTune &t = kv.second;
        o << "batched kernel " << kv.first << ": ";
        if (t.chosen >= 0)
            o << "block " << BATCH_CANDIDATES[t.chosen].block
              << (BATCH_CANDIDATES[t.chosen].double_buffer ? ", double-buffered" : ", one tile in registers")
              << (BATCH_CANDIDATES[t.chosen].wide ? ", wide mask words" : ", narrow mask windows")
              << (BATCH_CANDIDATES[t.chosen].stage ? ", small mask tables in LDS" : "")
              << (BATCH_CANDIDATES[t.chosen].dict ? ", sparse mask tables as rank dictionaries in LDS" : "");
        else
            o << "tuning (" << t.trials << " trials)";
        o << "; ns per row and program";
        for (int c = 0; c < N_BATCH_CAND; c++) {
            o << " [b" << BATCH_CANDIDATES[c].block << (BATCH_CANDIDATES[c].double_buffer ? "d" : "")
              << (BATCH_CANDIDATES[c].wide ? "w" : "") << (BATCH_CANDIDATES[c].stage ? "s" : "")
              << (BATCH_CANDIDATES[c].dict ? "r" : "") << "] ";
            if (t.best[c] < 1e299) o << t.best[c];
            else o << "-";
        }
        o << "\n";
    }
    int shape = 0;
    for (auto &kv : g_tune) {
        const Tune &t = kv.second;
        if (!t.queued_all) continue; // never launched over the tuning size
        o << "shape " << shape++ << " device " << kv.first.first << ": ";
        if (!t.error.empty())
            o << "compilation failed, interpreted";
        else if (t.chosen >= 0)
            o << "unroll " << TUNE_CANDIDATES[t.chosen].unroll << ", block " << TUNE_CANDIDATES[t.chosen].block
              << ", blocks/CU x" << TUNE_CANDIDATES[t.chosen].per_cu_mult
              << (t.chosen_queue ? ", queued tail" : "");
        else if (!t.default_ready || t.cands_ready < N_CAND - 1 || (t.queued_tail && !t.tail_ready))
            o << "compiling (" << (t.default_ready ? 1 : 0) + t.cands_ready + (t.tail_ready ? 1 : 0) << " kernels ready)";
        else
            o << "tuning (" << t.trials << " trials)";
        o << "; " << t.tuned_launches << " tuning launches, " << t.tune_ms << " ms in one-launch trials";
        o << "; ns/row";
        for (int i = 0; i < N_CAND; i++) {
            o << " [u" << TUNE_CANDIDATES[i].unroll << " b" << TUNE_CANDIDATES[i].block << " x"
              << TUNE_CANDIDATES[i].per_cu_mult << "] ";
            if (t.best_ns_per_row[i] < 1e299) o << t.best_ns_per_row[i];
            else o << "-";
        }
        if (t.queue_ns_per_row < 1e299) o << " [queued tail] " << t.queue_ns_per_row;
        o << "\n";
    }
    return o.str();
}

} // namespace sdb
