// MSM window autotuner.
//
// The reference fixes the window at compile time (MSM_G1_C_BITS, default 14,
// include/msm_g1.cuh:11; C=14 in include/msm_g2.cuh:70-110). Here the window
// is chosen per MSM size class by measurement on the device: for each
// candidate c the digit sort and the MSMs that share it run on the context
// stream over the circuit's resident bases and uniform scalars, timed with HIP
// events, and the fastest c is recorded for (kind, ceil log2 n). Two kinds:
//   kind 0: the witness sort over n_vars + 4 scalars and its four MSMs
//           (A', B1', C' in G1, B2' in G2), the cost one proof pays for c_w;
//   kind 1: the sort of h over n scalars and the H MSM (c_h).
// Every proof on the context then uses the table unless a window is forced.
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <sstream>

#include "prover.hpp"
#include "synth.hpp"

namespace zk {

int size_class(size_t n) {
    int lg = 0;
    while (((size_t)1 << lg) < n) lg++;
    return lg;
}

int Context::window_for(int kind, size_t n) const {
    if (kind == 0 && force_w) return force_w;
    if (kind == 1 && force_h) return force_h;
    auto it = tuned.find({kind, size_class(n)});
    return it != tuned.end() ? it->second : msm_default_c(n);
}

std::string Context::arch() const {
    hipDeviceProp_t p;
    ZK_HIP(hipGetDeviceProperties(&p, device));
    std::string a = p.gcnArchName;
    size_t colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
}

namespace {

struct TuneBufs {
    MsmSort sort;
    MsmScratch g1, g2;
    MsmPartials part;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TuneBufs() {
        ZK_HIP(hipEventCreate(&e0));
        ZK_HIP(hipEventCreate(&e1));
    }
    ~TuneBufs() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// device ms of one sort + the MSMs of `kind` at window c (mean over reps after one warm-up)
float time_window(Context& ctx, const Circuit& circ, int kind, int c, const void* d_sc, size_t n, int reps,
                  TuneBufs& tb) {
    hipStream_t st = ctx.stream;
    float best = 0;
    for (int it = 0; it <= reps; it++) {
        ZK_HIP(hipEventRecord(tb.e0, st));
        tb.sort.run(d_sc, n, c, st);
        if (kind == 0) {
            msm_launch(tb.sort, circ.A.p, false, tb.g1, tb.part, st, circ.A29.p);
            msm_launch(tb.sort, circ.B1.p, false, tb.g1, tb.part, st, circ.B1_29.p);
            msm_launch(tb.sort, circ.C.p, false, tb.g1, tb.part, st, circ.C29.p);
            msm_launch(tb.sort, circ.B2.p, true, tb.g2, tb.part, st, circ.B2_29.p);
        } else {
            msm_launch(tb.sort, circ.H.p, false, tb.g1, tb.part, st, circ.H29.p);
        }
        ZK_HIP(hipEventRecord(tb.e1, st));
        ZK_HIP(hipEventSynchronize(tb.e1));
        float ms = 0;
        ZK_HIP(hipEventElapsedTime(&ms, tb.e0, tb.e1));
        if (it > 0) best += ms / reps;
    }
    return best;
}

}  // namespace

void Context::autotune(const Circuit& circ, bool force, int reps, int chosen[2]) {
    ZK_HIP(hipSetDevice(device));
    const size_t sizes[2] = {(size_t)circ.n_vars + 4, (size_t)circ.n};
    TuneBufs tb;
    for (int kind = 0; kind < 2; kind++) {
        const size_t n = sizes[kind];
        const int cls = size_class(n);
        auto key = std::make_pair(kind, cls);
        if (!force && tuned.count(key)) {
            chosen[kind] = tuned[key];
            continue;
        }
        std::vector<uint64_t> sc(4 * n);
        synth_uniform(0x7475'6e65ull + (uint64_t)cls, n, sc.data());
        DevBuf d_sc;
        d_sc.ensure(n * 32);
        ZK_HIP(hipMemcpy(d_sc.p, sc.data(), n * 32, hipMemcpyHostToDevice));
        int best_c = msm_default_c(n);
        float best_ms = -1;
        const int hi = std::min(TUNE_C_MAX, std::max(TUNE_C_MIN, cls + 2));
        for (int c = TUNE_C_MIN; c <= hi; c++) {
            float ms = time_window(*this, circ, kind, c, d_sc.p, n, reps, tb);
            if (best_ms < 0 || ms < best_ms) {
                best_ms = ms;
                best_c = c;
            }
        }
        tuned[key] = best_c;
        chosen[kind] = best_c;
    }
}

// text table, one entry per line: "<arch> <kind> <log2 n> <c>"; '#' starts a comment
int Context::tune_load(const std::string& path) {
    std::ifstream f(path);
    ZK_REQUIRE(f.good(), ZK_E_IO, "cannot open tuning table " + path);
    const std::string me = arch();
    std::string line;
    int n = 0;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        std::string a;
        int kind, cls, c;
        ZK_REQUIRE(bool(is >> a >> kind >> cls >> c), ZK_E_FORMAT, "bad tuning table line: " + line);
        ZK_REQUIRE(kind >= 0 && kind <= 1 && cls >= 0 && cls <= 40 && c >= 4 && c <= 22, ZK_E_FORMAT,
                   "bad tuning table entry: " + line);
        if (a != me) continue;
        tuned[{kind, cls}] = c;
        n++;
    }
    return n;
}

void Context::tune_save(const std::string& path) const {
    std::ofstream f(path, std::ios::trunc);
    ZK_REQUIRE(f.good(), ZK_E_IO, "cannot write tuning table " + path);
    f << "# zyklop_amd MSM windows: arch kind(0 witness, 1 h) ceil-log2(n) c\n";
    const std::string me = arch();
    for (auto& kv : tuned) f << me << ' ' << kv.first.first << ' ' << kv.first.second << ' ' << kv.second << '\n';
    ZK_REQUIRE(f.good(), ZK_E_IO, "write failed: " + path);
}

}  // namespace zk
