// rtile_jit.cpp -- register-tile passes compiled at run time.
//
// The ahead-of-time kernel (regtile.hip) interprets a pass's decoded ops:
// per op it loads the descriptor and branches to the right body, and
// across that many-way branch the compiler keeps a second copy of the
// register array (two moves per amplitude per op, about as much as the
// arithmetic of a rotation).  For a pass structure that repeats -- the
// same circuit applied again, or a variational circuit with new angles --
// this file generates the pass as straight-line code: every op a call of
// the same rtile_core.h helpers, with its register bits, masks, predicates
// and remap offsets as literals.  hipRTC compiles it for the device's
// architecture with the library's own flags (no FMA contraction, no packed
// FP32), and the module is cached by its source text.  The angles are not
// in the source: each op reads its (c, s) from a table, so new angles reuse
// the compiled pass.  The arithmetic per amplitude is the interpreter's, so
// results are the same bit for bit.
#include <hip/hiprtc.h>

#include <condition_variable>
#include <cstdio>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>

#include "rtile.h"

namespace qslhip {
namespace {

#include "jit_headers.inc"  // kJitHeaders: {name, text} of the device headers

}  // namespace
extern int g_jit_compiled, g_jit_failed;
namespace {

struct Compiled {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    bool failed = false;
};

std::mutex g_mu;
std::unordered_map<std::string, int> g_seen;               // pass source -> applications
std::map<std::pair<int, std::string>, Compiled> g_cache;   // (device, source) -> module
constexpr size_t kMaxCached = 1024;
// Structures seen once are remembered so that a repeat can be recognised;
// a workload of ever-new structures would grow this without end, so it is
// cleared when it reaches this size (a structure then needs two more
// applications before it is compiled).
constexpr size_t kMaxSeen = 4096;

std::string hex(uint64_t x)
{
    char b[32];
    std::snprintf(b, sizeof b, "0x%llxull", (unsigned long long)x);
    return b;
}

// The pass as straight-line device code.  Mirrors k_rtile (regtile.hip):
// same prologue, group loop with prefetch, same helper calls per op.
std::string gen_source(const RegOp *ops, uint32_t nops, int prec, int NT, int R, bool exact,
                       const RegTileArgs &ta, bool prefetch)
{
    const uint32_t *koff0 = ta.koff0;
    const int tb = NT == 256 ? 8 : 9;
    const uint32_t rho0 = ((uint32_t)(R - 1)) << tb;  // registers on the top positions
    // zeros inserted at the positions of rho, lowest first (thread_tile_bits)
    auto insert_code = [](const char *var, uint32_t rho) {
        std::string c;
        for (; rho; rho &= rho - 1) {
            const uint32_t lowm = (rho & (0u - rho)) - 1;
            c += std::string("    ") + var + " = (" + var + " & " + std::to_string(lowm) +
                 "u) | ((" + var + " & " + std::to_string(~lowm) + "u) << 1);\n";
        }
        return c;
    };
    const int r = R == 4 ? 2 : R == 8 ? 3 : 4;
    std::ostringstream o;
    o << "#include \"rtile_core.h\"\nusing namespace qslhip;\n"
      << "typedef " << (prec == 8 ? "double" : "float") << " T;\n"
      << "constexpr int R = " << R << ", NT = " << NT << ", A = " << (prec == 8 ? 16 : 8)
      << ", r = " << r << ";\n"
      << "constexpr bool EX = " << (exact ? "true" : "false") << ";\n"
      << "extern \"C\" __global__ __launch_bounds__(NT) ";
    if (tuning().jit_waves > 0)  // caps registers at 512 / waves per lane
        o << "__attribute__((amdgpu_waves_per_eu(" << tuning().jit_waves << ", "
          << tuning().jit_waves << "))) ";
    o << "void qslhip_pass(\n"
         "    cplx<T> *__restrict__ st, unsigned n, RegTileArgs ta,\n"
         "    const double *__restrict__ cf, unsigned long long ngroups)\n{\n"
         "    extern __shared__ __align__(16) unsigned char smem[];\n"
         "    const uint32_t t = threadIdx.x;\n"
         "    const uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);\n"
         "    const uint64_t outside = full & ~ta.tile_mask;\n"
         "    const uint64_t tthread = deposit((uint64_t)t, ta.tlow);\n"
         "    uint64_t kpart[R];\n"
         "#pragma unroll\n"
         "    for (int k = 0; k < R; k++) {\n"
         "        uint64_t kp = 0;\n"
         "#pragma unroll\n"
         "        for (int i = 0; i < r; i++)\n"
         "            if ((k >> i) & 1) kp |= ta.kq[i];\n"
         "        kpart[k] = kp;\n"
         "    }\n"
         ;
    const bool same = ta.rho_in == ta.rho_out;
    o << "    uint32_t jt0 = t;\n";
    if (ta.rho_in != rho0) o << insert_code("jt0", ta.rho_in);
    if (!same) {
        o << "    const uint64_t tthread_out = deposit((uint64_t)t, ta.tlow_out);\n"
             "    uint64_t kpart_out[R];\n"
             "#pragma unroll\n"
             "    for (int k = 0; k < R; k++) {\n"
             "        uint64_t kp = 0;\n"
             "#pragma unroll\n"
             "        for (int i = 0; i < r; i++)\n"
             "            if ((k >> i) & 1) kp |= ta.kq_out[i];\n"
             "        kpart_out[k] = kp;\n"
             "    }\n"
             "    uint32_t jt1 = t;\n";
        if (ta.rho_out != rho0) o << insert_code("jt1", ta.rho_out);
    }
    o << "    const uint32_t kto[R] = {";
    for (int k = 0; k < R; k++)
        o << (k ? ", " : "") << reg_tile_bits((uint32_t)k, ta.rho_out) << "u";
    o << "};\n"
         "    uint64_t g = blockIdx.x;\n"
         "    if (g >= ngroups) return;\n";
    if (prefetch)
        o << "    uint64_t gbase_n = deposit(group_of(g, ngroups, ta.gorder), ta.gvar) | ta.gfix;\n"
             "    cplx<T> v[R], vn[R];\n"
             "    reg_load<T, R>(vn, st, gbase_n | tthread, kpart, ta);\n"
             "    wait_vmcnt<0>();\n"
             "    for (;;) {\n"
             "        const uint64_t gbase = gbase_n;\n"
             "#pragma unroll\n"
             "        for (int k = 0; k < R; k++) v[k] = vn[k];\n"
             "        const uint64_t gn = g + gridDim.x;\n"
             "        const bool more = gn < ngroups;\n"
             "        if (more) {\n"
             "            gbase_n = deposit(group_of(gn, ngroups, ta.gorder), ta.gvar) | ta.gfix;\n"
             "            reg_load<T, R>(vn, st, gbase_n | tthread, kpart, ta);\n"
             "            wait_vmcnt<2 * R>();\n"
             "        } else {\n"
             "            wait_vmcnt<R>();\n"
             "        }\n";
    else  // one group per workgroup: no second register set
        o << "    {\n"
             "        const uint64_t gbase = deposit(group_of(g, ngroups, ta.gorder), ta.gvar) | ta.gfix;\n"
             "        cplx<T> v[R];\n"
             "        reg_load<T, R>(v, st, gbase | tthread, kpart, ta);\n";
    o << "        uint32_t jt = jt0;\n"
         "        uint32_t base = lds_swz<A>(jt) * A;\n"
         "        (void)gbase; (void)jt; (void)base;\n"
         "        // a zero group of a basis state stays zero (reordered mode)\n"
         "        if (EX || ta.basis == ~0ull || (ta.basis & outside) == gbase) {\n";
    std::vector<uint32_t> kcur(koff0, koff0 + R);
    int j = 0;  // coefficient pair index
    for (uint32_t i = 0; i < nops; i++) {
        const RegOpHead &h = ops[i].h;
        if (h.code == RC_REMAP) {
            o << "        {\n";
            for (int k = 0; k < R; k++)
                o << "            *reinterpret_cast<cplx<T> *>(smem + (base ^ " << kcur[k]
                  << "u)) = v[" << k << "];\n";
            o << "            lds_barrier();\n            jt = t;\n";
            for (uint32_t rho = h.rho; rho; rho &= rho - 1) {
                const uint32_t lowm = (rho & (0u - rho)) - 1;
                o << "            jt = (jt & " << lowm << "u) | ((jt & " << ~lowm
                  << "u) << 1);\n";
            }
            o << "            base = lds_swz<A>(jt) * A;\n";
            for (int k = 0; k < R; k++) {
                kcur[k] = ops[i].koff[k];
                o << "            v[" << k << "] = *reinterpret_cast<const cplx<T> *>(smem + (base ^ "
                  << kcur[k] << "u));\n";
            }
            o << "            lds_barrier();\n        }\n";
            continue;
        }
        const bool sel = h.tmask != 0;
        o << "        {\n";
        if (h.gmask)
            o << "          if ((gbase & " << hex(h.gmask) << ") == " << hex(h.gmask) << ") {\n";
        o << "            const T c = (T)cf[" << 2 * j << "], s = (T)cf[" << 2 * j + 1 << "];\n";
        if (sel)
            o << "            const bool tp = (jt & " << h.tmask << "u) == " << h.tmask << "u;\n";
        else
            o << "            const bool tp = false;\n";
        const char *S = sel ? "true" : "false";
        switch (h.code) {
        case RC_PAIR:
            o << "            reg_pair<T, R, " << h.ia << ", " << h.body << ", " << S
              << ", false, EX>(v, " << h.kmask << "u, tp, c, s);\n";
            break;
        case RC_SWAP:
            o << "            reg_swaplike<T, R, " << h.ia << ", " << h.ib << ", " << h.body << ", "
              << ((h.flags & RF_FSWAP) ? "true" : "false") << ", EX>(v, c, s);\n";
            break;
        case RC_ONE:
            o << "            reg_one<T, R, " << h.body << ", " << S << ", false, EX>(v, "
              << h.kmask << "u, tp, c, s);\n";
            break;
        default: {  // RC_RZ
            const bool lane = h.tneg != 0;
            if (ops[i].gneg)
                o << "            const bool gneg = (gbase & " << hex(ops[i].gneg) << ") != 0;\n";
            else
                o << "            const bool gneg = false;\n";
            if (lane)
                o << "            const T sl = flip_sign(s, (jt & " << h.tneg << "u) != 0);\n";
            else
                o << "            const T sl = s;\n";
            o << "            reg_rz<T, R, " << S << ", false, " << (lane ? "true" : "false")
              << ", EX>(v, " << h.kmask << "u, " << h.kneg << "u, gneg, tp, c, sl);\n";
            break;
        }
        }
        o << "            (void)tp;\n";
        if (h.gmask) o << "          }\n";
        o << "        }\n";
        j++;
    }
    o << "        }\n";
    if (same)
        o << "        if (ta.runs) reg_run_sums<T, R, NT>(v, ta.runs, group_of(g, ngroups, ta.gorder), jt0, kto);\n"
             "        reg_store<T, R>(v, st, gbase | tthread, kpart, ta);\n";
    else
        o << "        if (ta.runs) reg_run_sums<T, R, NT>(v, ta.runs, group_of(g, ngroups, ta.gorder), jt1, kto);\n"
             "        reg_store<T, R>(v, st, gbase | tthread_out, kpart_out, ta);\n";
    if (prefetch)
        o << "        if (!more) break;\n"
             "        g = gn;\n";
    o << "    }\n}\n";
    return o.str();
}

// hipRTC compile of a pass for `arch`; the code object in `code`, the log in
// `log`.
bool rtc_compile(const std::string &src, const std::string &arch_name, std::vector<char> &code,
                 std::string &log)
{
    const int nh = (int)(sizeof(kJitHeaders) / sizeof(kJitHeaders[0]));
    std::vector<const char *> names, texts;
    for (int i = 0; i < nh; i++) {
        names.push_back(kJitHeaders[i][0]);
        texts.push_back(kJitHeaders[i][1]);
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "qslhip_pass.hip", nh, texts.data(),
                            names.data()) != HIPRTC_SUCCESS)
        return false;
    const std::string arch = std::string("--offload-arch=") + arch_name;
    const char *opts[] = {arch.c_str(), "-O3", "-std=c++20", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    size_t ls = 0;
    if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
        log.assign(ls, '\0');
        hiprtcGetProgramLog(prog, &log[0]);
    }
    bool ok = rc == HIPRTC_SUCCESS;
    if (ok) {
        size_t cs = 0;
        ok = hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0;
        if (ok) {
            code.resize(cs);
            ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
        }
    }
    hiprtcDestroyProgram(&prog);
    return ok;
}

bool compile(int dev, const std::string &src, Compiled &out)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
    std::vector<char> code;
    std::string log;
    if (!rtc_compile(src, prop.gcnArchName, code, log)) return false;
    if (hipModuleLoadData(&out.mod, code.data()) != hipSuccess) return false;
    if (hipModuleGetFunction(&out.fn, out.mod, "qslhip_pass") != hipSuccess) return false;
    return true;
}

// Compiles pass structures on a host thread of its own (jit mode 1): the
// application that finds a structure repeating queues it and runs the pass
// interpreted; a later application finds the module in the cache.  The
// thread is joined when the library unloads (a compile in progress
// finishes first).
class Compiler {
public:
    ~Compiler()
    {
        {
            std::lock_guard<std::mutex> lk(g_mu);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
    }

    // Blocks until every queued structure is compiled (or failed).
    void wait_idle()
    {
        std::unique_lock<std::mutex> lk(g_mu);
        idle_.wait(lk, [this] { return pending_.empty(); });
    }

    // Called with g_mu held.
    void enqueue(int dev, const std::string &src)
    {
        auto key = std::make_pair(dev, src);
        if (pending_.count(key)) return;
        pending_.insert(key);
        jobs_.push_back(key);
        if (!th_.joinable()) th_ = std::thread([this] { run(); });
        cv_.notify_one();
    }

private:
    void run()
    {
        std::unique_lock<std::mutex> lk(g_mu);
        for (;;) {
            cv_.wait(lk, [this] { return stop_ || !jobs_.empty(); });
            if (stop_) return;
            const auto key = jobs_.front();
            jobs_.pop_front();
            lk.unlock();
            Compiled c;
            bool ok = hipSetDevice(key.first) == hipSuccess && compile(key.first, key.second, c);
            lk.lock();
            if (!ok) {
                c.failed = true;
                g_jit_failed++;
            } else {
                g_jit_compiled++;
            }
            if (g_cache.size() < kMaxCached) g_cache.emplace(key, c);
            pending_.erase(key);
            if (pending_.empty()) idle_.notify_all();
        }
    }

    std::thread th_;
    std::condition_variable cv_, idle_;
    std::deque<std::pair<int, std::string>> jobs_;
    std::set<std::pair<int, std::string>> pending_;
    bool stop_ = false;
};

Compiler &compiler()
{
    static Compiler c;
    return c;
}

}  // namespace

std::string jit_pass_source(unsigned n, int prec, const TilePass &p, int threads, bool exact,
                            const RegOp *ops, uint32_t nops, const RegLayout &lay)
{
    const int tb = threads == 512 ? 9 : 8;
    const RegTileArgs ta = make_tile_args(n, p, tb, prec == 8 ? 16 : 8, lay);
    return gen_source(ops, nops, prec, threads, 1 << (p.logt - tb), exact, ta,
                      tuning().jit_prefetch != 0);
}

bool jit_check_source(const std::string &src, const std::string &arch, std::string &log)
{
    std::vector<char> code;
    return rtc_compile(src, arch, code, log);
}

int g_jit_compiled = 0, g_jit_failed = 0;  // counters (tuning keys jit_compiled / jit_failed)

void jit_wait_idle() { compiler().wait_idle(); }

void jit_coefficients(const RegOp *ops, uint32_t nops, std::vector<double> &coef)
{
    for (uint32_t i = 0; i < nops; i++) {
        if (ops[i].h.code == RC_REMAP) continue;
        coef.push_back(ops[i].h.c);
        coef.push_back(ops[i].h.s);
    }
}

namespace {

// The compiled function of a pass source, or nullptr (then the pass runs
// interpreted): jit mode 1 queues a source for the compiler thread the
// second time it is asked for, mode 2 compiles it now.  Caller holds g_mu.
hipFunction_t pass_function(int dev, const std::string &src, int mode)
{
    const auto key = std::make_pair(dev, src);
    auto it = g_cache.find(key);
    if (it == g_cache.end()) {
        if (g_cache.size() >= kMaxCached) return nullptr;
        if (mode == 1) {
            if (g_seen.size() >= kMaxSeen) g_seen.clear();
            if (++g_seen[src] >= 2) compiler().enqueue(dev, src);
            return nullptr;
        }
        Compiled c;
        if (!compile(dev, src, c)) {
            c.failed = true;
            g_jit_failed++;
        } else {
            g_jit_compiled++;
        }
        it = g_cache.emplace(key, c).first;
    }
    return it->second.failed ? nullptr : it->second.fn;
}

// Per-pass choice among four launch shapes of an FP64 pass that reads whole
// tiles (tuning pass_tune): one group per workgroup or the prefetching
// group loop, each in group order 0 or 1.  Which is fastest depends on the
// tile's qubits (a gate-free copy with the same addressing varies by up to
// 20% between them), so the first launches of a pass try each shape once,
// timed with events on the stream, and later launches take the fastest.
// The shapes move the same groups through the same arithmetic: results are
// bit-identical whichever runs.
struct PassTuner {
    float best[4] = {0, 0, 0, 0};
    bool tried[4] = {false, false, false, false};
    int pending = -1;  // shape whose events are in flight
    int choice = -1;
    hipEvent_t ev[2] = {nullptr, nullptr};
};
std::map<std::string, PassTuner> g_tuners;  // pass source + tile addressing -> tuner
constexpr size_t kMaxTuners = 256;
int g_tune_pending = 0;  // tuners still trying shapes

}  // namespace

int pass_tune_pending() { return g_tune_pending; }

bool jit_launch_pass(qslhip_state *h, void *amps, const TilePass &p, int threads, int gpw,
                     bool exact, const RegOp *ops, uint32_t nops, const double *dev_coef,
                     const PassExtras &x)
{
    const int mode = tuning().jit;
    if (mode == 0) return false;
    const int tb = threads == 512 ? 9 : 8;
    const int A = h->prec == 8 ? 16 : 8;
    const RegTileArgs ta = make_tile_args(h->n, p, tb, A, x.lay);
    const int R = 1 << (p.logt - tb);
    uint64_t groups = 0;
    RegTileArgs targ = pass_args(h->n, p, tb, A, x, &groups);
    const bool single_ok = groups * (uint64_t)threads <= 0xFFFFFFFFull;
    // FP64 passes that read whole tiles run one group per workgroup: without
    // the second register set more workgroups are resident (measured faster
    // at 30 qubits).  Passes that read only part of each tile (from a basis
    // state) and FP32 passes (few registers either way; measured faster at
    // 34 qubits) prefetch the next group while computing one, and so do
    // passes whose groups outnumber what one grid dimension addresses.
    const int pf = tuning().jit_prefetch;
    const bool prefetch = pf == 1 || (pf == 2 && (x.zmask != 0 || h->prec != 8)) || !single_ok;
    const bool tune = tuning().pass_tune != 0 && h->prec == 8 && x.zmask == 0 && single_ok &&
                      groups >= 64 && (groups % 8) == 0;
    hipFunction_t fn = nullptr;
    bool loop = prefetch;
    int shape = -1;
    PassTuner *tu = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!tune) {
            fn = pass_function(h->device, gen_source(ops, nops, h->prec, threads, R, exact, ta,
                                                     prefetch),
                               mode);
        } else {
            const std::string src1 = gen_source(ops, nops, h->prec, threads, R, exact, ta, false);
            const std::string src2 = gen_source(ops, nops, h->prec, threads, R, exact, ta, true);
            hipFunction_t fs = pass_function(h->device, src1, mode);
            hipFunction_t fl = pass_function(h->device, src2, mode);
            if (!fs && !fl) return false;
            std::string key = src1;
            key += "|" + std::to_string(h->device) + "|" + std::to_string(p.tile_mask) + "|" +
                   std::to_string(targ.gvar) + "|" + std::to_string(targ.gfix);
            auto it = g_tuners.find(key);
            if (it == g_tuners.end() && g_tuners.size() < kMaxTuners) {
                it = g_tuners.emplace(key, PassTuner()).first;
                g_tune_pending++;
            }
            if (it != g_tuners.end()) {
                tu = &it->second;
                if (tu->pending >= 0 && hipEventQuery(tu->ev[1]) == hipSuccess) {
                    float ms = 0;
                    if (hipEventElapsedTime(&ms, tu->ev[0], tu->ev[1]) == hipSuccess) {
                        tu->best[tu->pending] = ms;
                        tu->tried[tu->pending] = true;
                    }
                    tu->pending = -1;
                    if (tu->choice < 0 && tu->tried[0] && tu->tried[1] && tu->tried[2] &&
                        tu->tried[3]) {
                        int b = 0;
                        for (int v = 1; v < 4; v++)
                            if (tu->best[v] < tu->best[b]) b = v;
                        tu->choice = b;
                        g_tune_pending--;
                    }
                }
                if (tu->choice >= 0) {
                    shape = tu->choice;
                } else if (tu->pending < 0) {
                    for (int v = 0; v < 4 && shape < 0; v++)
                        if (!tu->tried[v] && ((v & 2) ? fl : fs)) shape = v;
                }
                if (shape < 0 || !((shape & 2) ? fl : fs)) shape = -1;
            }
            if (shape >= 0) {
                loop = (shape & 2) != 0;
                targ.gorder = (uint32_t)(shape & 1);
                fn = loop ? fl : fs;
                if (tu->choice >= 0 || tu->pending >= 0) tu = nullptr;  // nothing to time
            } else {
                tu = nullptr;
                fn = prefetch ? fl : fs;
                if (!fn) {  // the static shape is not compiled yet: the other one
                    fn = prefetch ? fs : fl;
                    loop = !prefetch;
                }
            }
        }
        if (!fn) return false;
        if (tu) {
            if (!tu->ev[0]) {
                QH_CHECK(hipEventCreate(&tu->ev[0]));
                QH_CHECK(hipEventCreate(&tu->ev[1]));
            }
            tu->pending = shape;
        }
    }
    // a pass compiled without the group loop takes one group per workgroup
    const unsigned grid = loop ? (unsigned)reg_grid(groups, gpw > 0 ? gpw : 1) : (unsigned)groups;
    if (grid % 8) targ.gorder = 0;
    const size_t shmem = (size_t)A << p.logt;
    void *st = amps;
    unsigned n = h->n;
    const double *cf = dev_coef;
    unsigned long long ng = groups;
    void *args[] = {&st, &n, &targ, &cf, &ng};
    if (tu) QH_CHECK(hipEventRecord(tu->ev[0], h->stream));
    QH_CHECK(hipModuleLaunchKernel(fn, grid, 1, 1, (unsigned)threads, 1, 1, (unsigned)shmem,
                                   h->stream, args, nullptr));
    if (tu) QH_CHECK(hipEventRecord(tu->ev[1], h->stream));
    return true;
}

}  // namespace qslhip
