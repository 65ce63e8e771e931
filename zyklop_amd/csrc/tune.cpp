// MSM window autotuner.
//
// The reference fixes the window at compile time (MSM_G1_C_BITS, default 14,
// include/msm_g1.cuh:11; C=14 in include/msm_g2.cuh:70-110). Here the window c
// and the bucket split of the accumulation (threads per bucket, msm.hpp) are
// chosen per MSM size class by measurement on the device: for each candidate
// c the digit sort runs once and each MSM that shares it runs at every split,
// on the context stream over the circuit's resident bases and uniform
// scalars, timed with HIP events. A candidate's cost is the sort plus, per
// MSM, its fastest split; the cheapest c is recorded for (kind, ceil log2 n)
// with its G1 and G2 splits. Two kinds:
//   kind 0: the witness sort over n_vars + 4 scalars and its four MSMs
//           (A', B1', C' in G1, B2' in G2), the cost one proof pays for c_w;
//   kind 1: the sort of h over n scalars and the H MSM (c_h).
// A circuit with window tables fixes the window (the tables'): for it only
// the splits are tuned, recorded under kind 2 (witness) or 3 (H).
// Every proof on the context then uses the table unless a window is forced.
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <vector>

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
    return it != tuned.end() ? it->second.c : msm_default_c(n);
}

int Context::split_for(int kind, size_t n, int c, bool g2) const {
    const bool forced = (kind & 1) == 0 ? force_w != 0 : force_h != 0;
    auto it = tuned.find({kind, size_class(n)});
    if (!forced && it != tuned.end() && it->second.c == c) {
        const int sp = g2 ? it->second.split_g2 : it->second.split_g1;
        if (sp > 0) return sp;
    }
    return msm_default_split(n, c, g2, kind >= 2);
}

MsmPlan Context::msm_plan(const Circuit& circ) const {
    MsmPlan p;
    for (int k = 0; k < 2; k++) {
        const size_t n = k ? (size_t)circ.n : (size_t)circ.n_vars + 4;
        const int forced = k ? force_h : force_w, tc = circ.tab_c[k];
        bool merged = false;
        int c;
        if (tc && (!forced || msm_geom(n, forced, false).W == msm_geom(n, tc, false).W)) {
            c = tc;
            merged = true;
        } else {
            c = window_for(k, n);
        }
        const int kind = k + (merged ? 2 : 0);
        if (k == 0) {
            p.cw = c;
            p.mw = merged;
            p.sp_w1 = split_for(kind, n, c, false);
            p.sp_w2 = split_for(kind, n, c, true);
        } else {
            p.ch = c;
            p.mh = merged;
            p.sp_h = split_for(kind, n, c, false);
        }
    }
    return p;
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

float elapsed(TuneBufs& tb) {
    ZK_HIP(hipEventSynchronize(tb.e1));
    float ms = 0;
    ZK_HIP(hipEventElapsedTime(&ms, tb.e0, tb.e1));
    return ms;
}

// Device ms at window c, mean over reps after one warm-up: the sort, and each
// MSM of `kind` at each split 1..MSM_SPLIT_MAX (msm[m][split-1]; splits that
// would start more than 2^21 threads are skipped and left at -1).
void time_window(Context& ctx, const Circuit& circ, int kind, int c, bool merged, const void* d_sc, size_t n,
                 int reps, TuneBufs& tb, float& sort_ms, float msm[4][MSM_SPLIT_MAX]) {
    hipStream_t st = ctx.stream;
    // the witness MSMs as a proof runs them: with the circuit's wire classes,
    // A and B1, B2 read their parts of the buckets
    struct M { const void* p29; bool g2; int range; };
    const M w4[4] = {{circ.A29.p, false, MSM_RANGE_A}, {circ.B1_29.p, false, MSM_RANGE_B},
                     {circ.C29.p, false, MSM_RANGE_ALL}, {circ.B2_29.p, true, MSM_RANGE_B}};
    const M h1[1] = {{circ.H29.p, false, MSM_RANGE_ALL}};
    const uint8_t* cls = kind == 0 && circ.wire_cls.bytes ? circ.wire_cls.as<uint8_t>() : nullptr;
    const M* ms = kind == 0 ? w4 : h1;
    const int nm = kind == 0 ? 4 : 1;
    const MsmGeom g = msm_geom(n, c, false, merged);
    sort_ms = 0;
    for (int m = 0; m < 4; m++)
        for (int s = 0; s < MSM_SPLIT_MAX; s++) msm[m][s] = -1;
    for (int it = 0; it <= reps; it++) {
        ZK_HIP(hipEventRecord(tb.e0, st));
        tb.sort.run(d_sc, n, c, st, merged, cls);
        ZK_HIP(hipEventRecord(tb.e1, st));
        const float t_sort = elapsed(tb);
        if (it > 0) sort_ms += t_sort / reps;
        for (int m = 0; m < nm; m++)
            for (int s = 1; s <= MSM_SPLIT_MAX; s++) {
                if ((size_t)g.Wr * g.K * s > ((size_t)1 << 21) && s > 1) continue;
                ZK_HIP(hipEventRecord(tb.e0, st));
                MsmScratch& scr = ms[m].g2 ? tb.g2 : tb.g1;
                msm_launch_pre(tb.sort, ms[m].p29, ms[m].g2, scr, st, s);
                msm_launch_main(tb.sort, ms[m].g2, scr, tb.part, st, ms[m].p29, ms[m].range);
                ZK_HIP(hipEventRecord(tb.e1, st));
                const float t = elapsed(tb);
                if (it > 0) msm[m][s - 1] = (msm[m][s - 1] < 0 ? 0 : msm[m][s - 1]) + t / reps;
            }
    }
}

}  // namespace

constexpr float SPLIT_MARGIN = 0.97f;

void Context::autotune(const Circuit& circ, bool force, int reps, int chosen[2]) {
    ZK_HIP(hipSetDevice(device));
    const size_t sizes[2] = {(size_t)circ.n_vars + 4, (size_t)circ.n};
    TuneBufs tb;
    for (int kind = 0; kind < 2; kind++) {
        const size_t n = sizes[kind];
        const int cls = size_class(n);
        const int tc = circ.tab_c[kind];
        auto key = std::make_pair(kind + (tc ? 2 : 0), cls);
        if (!force && tuned.count(key)) {
            chosen[kind] = tuned[key].c;
            continue;
        }
        std::vector<uint64_t> sc(4 * n);
        synth_uniform(0x7475'6e65ull + (uint64_t)cls, n, sc.data());
        DevBuf d_sc;
        d_sc.ensure(n * 32);
        ZK_HIP(hipMemcpy(d_sc.p, sc.data(), n * 32, hipMemcpyHostToDevice));
        MsmChoice best{tc ? tc : msm_default_c(n), 1, 1};
        float best_ms = -1;
        // candidate windows around the size (wide for large MSMs, narrow for
        // small), one per window count: widths with the same count give the
        // same geometry (msm_geom); over window tables, the tables' window
        const int hi = std::min(TUNE_C_MAX, std::max(TUNE_C_MIN, cls - 2));
        const int lo = std::max(TUNE_C_MIN, std::min(cls - 8, hi - 5));
        std::vector<int> cands;
        if (tc) {
            cands.push_back(tc);
        } else {
            for (int c = lo; c <= hi; c++) {
                const MsmGeom g = msm_geom(n, c, false);
                if (cands.empty() || msm_geom(n, cands.back(), false).W != g.W) cands.push_back(g.c);
            }
        }
        for (int c : cands) {
            float sort_ms, msm[4][MSM_SPLIT_MAX];
            time_window(*this, circ, kind, c, tc != 0, d_sc.p, n, reps, tb, sort_ms, msm);
            float total = sort_ms;
            int sp[4] = {1, 1, 1, 1};
            // a larger split adds level-1 work (one addition per extra part
            // and bucket) that a GPU shared with other proofs pays in full:
            // it must beat the smaller split alone by SPLIT_MARGIN
            for (int m = 0; m < (kind == 0 ? 4 : 1); m++) {
                for (int s = 2; s <= MSM_SPLIT_MAX; s++)
                    if (msm[m][s - 1] >= 0 && msm[m][s - 1] < SPLIT_MARGIN * msm[m][sp[m] - 1]) sp[m] = s;
                total += msm[m][sp[m] - 1];
            }
            if (best_ms < 0 || total < best_ms) {
                best_ms = total;
                // G1 split: the one most of the G1 MSMs of this kind chose
                int g1 = sp[0];
                if (kind == 0 && sp[1] == sp[2]) g1 = sp[1];
                best = {c, g1, kind == 0 ? sp[3] : 1};
            }
        }
        tuned[key] = best;
        chosen[kind] = best.c;
    }
}

// text table, one entry per line: "<arch> <kind> <log2 n> <c> [<split G1> <split G2>]"
// (kind 0 witness, 1 h, 2 and 3 the same over window tables); '#' starts a
// comment; an entry without splits uses msm_default_split
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
        int kind, cls, c, s1 = 0, s2 = 0;
        ZK_REQUIRE(bool(is >> a >> kind >> cls >> c), ZK_E_FORMAT, "bad tuning table line: " + line);
        if (is >> s1) ZK_REQUIRE(bool(is >> s2), ZK_E_FORMAT, "bad tuning table line: " + line);
        ZK_REQUIRE(kind >= 0 && kind <= 3 && cls >= 0 && cls <= 40 && c >= 4 && c <= 22 && s1 >= 0 &&
                       s1 <= MSM_SPLIT_MAX && s2 >= 0 && s2 <= MSM_SPLIT_MAX,
                   ZK_E_FORMAT, "bad tuning table entry: " + line);
        if (a != me) continue;
        tuned[{kind, cls}] = {c, s1, s2};
        n++;
    }
    return n;
}

void Context::tune_save(const std::string& path) const {
    std::ofstream f(path, std::ios::trunc);
    ZK_REQUIRE(f.good(), ZK_E_IO, "cannot write tuning table " + path);
    f << "# zyklop_amd MSM windows: arch kind(0 witness, 1 h, 2/3 over window tables) ceil-log2(n) c split-G1 "
         "split-G2\n";
    const std::string me = arch();
    for (auto& kv : tuned)
        f << me << ' ' << kv.first.first << ' ' << kv.first.second << ' ' << kv.second.c << ' ' << kv.second.split_g1
          << ' ' << kv.second.split_g2 << '\n';
    ZK_REQUIRE(f.good(), ZK_E_IO, "write failed: " + path);
}

}  // namespace zk
