// Pippenger multi-scalar multiplication on gfx950, G1 and G2.
//
// Reference: msm_g1 (include/msm_g1.cuh:714-913), msm_g2_device
// (include/msm_g2.cuh:669-854). Same mathematics (signed c-bit digits,
// buckets, running-sum reduction), different machine design:
//   * GLV split (glv.hpp): each 254-bit scalar becomes two signed 127-bit
//     halves over P_i and lambda*P_i = (beta*x, y), so an MSM of n points is
//     one of 2n points with ceil(128/c) windows: the same bucket entries as
//     ceil(255/c) windows over n points, half the buckets to reduce and half
//     the windows to combine;
//   * digit sort: a two-pass LDS radix sort (tile histograms of bucket
//     groups + device scan + LDS-ranked scatter, then per-group LDS counting
//     sort); no global atomics, no CUB, no host synchronisation;
//   * one sort can serve several MSMs that share a scalar vector;
//   * buckets in XYZZ coordinates, one thread per bucket, accumulated in a
//     9 x 29-bit-limb Montgomery form (field29.hpp) over a converted copy of
//     the bases; oversized buckets are split into fixed chunks reduced by
//     whole workgroups (device-side work lists, no host round trip);
//   * bucket reduction: thread-level running sums over segments, then an LDS
//     suffix-scan combine per workgroup; the host finishes a few points per
//     window and runs Horner across windows;
//   * window tables for bases that stay resident (a circuit's CRS): table w
//     holds 2^(offset of window w) * P_i for every base, so the digits of all
//     windows go to ONE set of buckets (sum_w d_w 2^off_w P = sum_w d_w T_w).
//     One set of 2^(c-1) buckets for W windows makes wider windows cheap:
//     fewer windows, hence fewer entries, for the same buckets to reduce,
//     and no Horner.
#pragma once
#include <stdint.h>

#include <vector>

#include "common.hpp"
#include "host_field.hpp"

namespace zk {

constexpr int GLV_BITS = 128;  // signed half scalars: magnitude < 2^127

struct MsmGeom {
    int c;          // window bits: windows w < nlo have c - 1 bits, the others c
    int W;          // windows = ceil(GLV_BITS / requested c); then c = ceil(GLV_BITS / W)
    int nlo;        // W*c - GLV_BITS
    bool merged;    // all windows share one set of buckets (bases from window tables)
    int Wr;         // bucket sets: W, or 1 when merged
    uint32_t K;     // buckets per set = 2^(c-1)
    uint32_t l1;    // buckets per level-1 reduction thread
    uint32_t nseg;  // level-1 segments per bucket set = K / l1
    uint32_t nblk;  // level-2 blocks per bucket set
};

int msm_default_c(size_t n);
// The geometry depends on the window count only: every c with the same
// W = ceil(GLV_BITS / c) gives the same widths.
MsmGeom msm_geom(size_t n, int c, bool g2, bool merged = false);
// window of the tables of n resident bases: the c that minimises the
// accumulation (2n entries per window, one mixed addition each) plus the
// reduction of one set of 2^(c-1) buckets (two general additions each)
int msm_table_c(size_t n);

// Per-scalar-vector digit sort (shared by the MSMs over the same scalars).
// Besides the entries in bucket order it produces the accumulation order: all
// Wr*K bucket ids sorted by descending load (quantised by 2^qs), so that the
// lanes of one wave walk buckets of nearly equal length. Buckets above lmax
// entries take the chunked path and come first.
// An entry is slot | half << 30 | sign << 31: the 29-bit base slot (the
// scalar's index mod n, plus w*n in window w of a merged sort, whose bases
// come from window tables), which of its two GLV halves, and the digit's sign.
//
// Scalar classes (optional, one byte per scalar, 0..3): the entries of every
// bucket are then ordered by the class of their scalar, and cls_off[3b + k]
// is where class k + 1 starts in bucket b. With the classes of the witness
// (MSM_CLS_*), each MSM over the sort reads one contiguous range of every
// bucket (MsmRange) that leaves out the scalars whose bases are infinite.
enum : uint8_t {
    MSM_CLS_A_INF = 0,   // A' infinite, B1'/B2' not
    MSM_CLS_NONE = 1,    // no infinite base
    MSM_CLS_B_INF = 2,   // B1' and B2' infinite, A' not
    MSM_CLS_AB_INF = 3,  // A', B1' and B2' infinite
};
enum MsmRange {
    MSM_RANGE_ALL = 0,   // every entry
    MSM_RANGE_A = 1,     // classes 1..2 (A' finite)
    MSM_RANGE_B = 2,     // classes 0..1 (B1', B2' finite)
};
struct MsmSort {
    MsmGeom g{};
    size_t n = 0;                 // scalars; the sort itself runs over 2n half scalars
    uint32_t lmax = 0, nkey = 0;
    int qs = 0;
    DevBuf half, matrix, mscan, pairs, offsets, entries, order, lkeys;
    // with classes: cls_off, the accumulation orders of MSM_RANGE_A / _B (by
    // the length of those parts), and the histograms of the three orders
    DevBuf cls_off, order_a, order_b, okeys;
    bool classes = false;         // the last run() had scalar classes
    // chunk lists of the oversized buckets (above lmax entries), built by run()
    DevBuf chunks, large, counters;
    size_t n_chunk_max = 0;
    // sizes every buffer a sort of n scalars at window c needs (no launch),
    // with_classes: also cls_off, order_a, order_b and okeys; returns the
    // chunk-list capacity for MsmScratch::ensure
    size_t reserve(size_t n, int c, bool merged = false, bool with_classes = false);
    // the device bytes reserve(n, c, merged, with_classes) would add now (into
    // gc); same return value
    size_t count_reserve(size_t n, int c, GrowCount& gc, bool merged = false, bool with_classes = false) const;
    // sorts canonical scalars (n x 32 B, device, n < 2^30; W*n < 2^30 merged)
    // into bucket order, entries as above; cls: n device bytes, the scalar
    // classes (optional)
    void run(const void* d_scalars, size_t n, int c, hipStream_t st, bool merged = false,
             const uint8_t* cls = nullptr);
};

// Result of the device part of one MSM: level-2 partials per window, on the host.
struct MsmPartials {
    MsmGeom g{};
    bool g2 = false;
    HostBuf host;       // [Wr][nblk] R then [Wr][nblk] T, XYZZ each
    // caller-owned host memory used instead of `host` when set (ext_bytes long)
    void* ext = nullptr;
    size_t ext_bytes = 0;
    const uint64_t* data() const { return static_cast<const uint64_t*>(ext ? ext : host.p); }
};

// Device scratch for the bucket phase of one MSM.
struct MsmScratch {
    DevBuf buckets, r1, t1, r2, t2, partials;
    int split = 0;     // threads per bucket of the queued MSM (set by msm_launch_pre)
    size_t cap_buckets = 0;
    // optional events around the bucket-accumulation kernel (profiling)
    hipEvent_t ev_acc0 = nullptr, ev_acc1 = nullptr, ev_end = nullptr;
    void enable_timing();
    ~MsmScratch();
    void ensure(const MsmGeom& g, size_t max_chunks, bool g2, int split);
    // the device bytes ensure() would add now (into gc)
    void count_ensure(const MsmGeom& g, size_t max_chunks, bool g2, int split, GrowCount& gc) const;
};

// Launch bucket accumulation + GPU reduction against a finished sort; queue
// the D2H of the partials into out.host on st. The caller synchronises st
// before finish. d_points29: the base slots of msm_points_to29 (for a merged
// sort, the W window tables of msm_points_table, table w at slot w*n); may be
// null only for n = 0.
// split: threads per bucket in the accumulation (1..MSM_SPLIT_MAX), each over
// a consecutive part of the bucket's entries; the level-1 reduction adds the
// parts. More threads than buckets fill the GPU when the buckets are few.
constexpr int MSM_SPLIT_MAX = 4;
// split that gives about two rounds of resident waves on a 256-CU GPU
int msm_default_split(size_t n, int c, bool g2, bool merged = false);
void msm_launch(const MsmSort& sort, bool g2, MsmScratch& scr, MsmPartials& out, hipStream_t st,
                const void* d_points29, int split = 1);
// msm_launch in two parts: the chunked path of the oversized buckets, then
// the accumulation, reductions and the copy of the partials. Queueing every
// MSM's first part before any accumulation keeps those short launches from
// waiting behind the accumulations of other streams.
void msm_launch_pre(const MsmSort& sort, const void* d_points29, bool g2, MsmScratch& scr, hipStream_t st, int split);
// range: the part of every bucket this MSM reads (MsmRange; a sort without
// classes reads all of it)
void msm_launch_main(const MsmSort& sort, bool g2, MsmScratch& scr, MsmPartials& out, hipStream_t st,
                     const void* d_points29, int range = MSM_RANGE_ALL);

// n affine points (standard form, G1 64 B / G2 128 B each) -> n slots of the
// 29-bit-limb form (G1 128 B, G2 256 B; layout in field29.hpp), each holding
// P_i and the x of lambda*P_i: the 2n bases of the GLV split.
constexpr size_t PT29_BYTES = 128, PT29Q_BYTES = 256;
void msm_points_to29(const void* d_points, size_t n, void* d_out, bool g2, hipStream_t st);
// window tables of n affine points for the geometry of window c: W tables of
// n slots (the form of msm_points_to29), table w holding 2^(offset of window
// w) * P_i. d_out holds W*n slots; table 0 is msm_points_to29's output.
void msm_points_table(const void* d_points, size_t n, int c, void* d_out, bool g2, hipStream_t st);

// size out's host buffer for an MSM of geometry g (no launch)
void msm_reserve_partials(MsmPartials& out, const MsmGeom& g, bool g2);

// G1 MSM whose whole device workspace (29-bit base copies, digit sort,
// buckets, reductions) lives in a fixed per-device array allocated at load
// time, and whose host partials go to a fixed host array: no allocation of
// any kind. msm_noalloc_setup(device) resolves the array's device address
// once per device (with that device current); msm_g1_noalloc then runs on
// stream st and waits for it. Sizes whose workspace exceeds the array are
// rejected (ZK_E_ARG). Calls on one device share the array: overlapping calls
// race.
constexpr size_t MSM_NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
void msm_noalloc_setup(int device);
zkh::Xyzz<zkh::Fp> msm_g1_noalloc(int device, const void* d_points, const void* d_scalars, size_t n, int window,
                                  hipStream_t st);

// Host finish: combine the partials into the MSM value.
zkh::Xyzz<zkh::Fp> msm_finish_g1(const MsmPartials& p);
zkh::Xyzz<zkh::Fp2> msm_finish_g2(const MsmPartials& p);

}  // namespace zk
