// Pippenger multi-scalar multiplication on gfx950, G1 and G2.
//
// Reference: msm_g1 (include/msm_g1.cuh:714-913), msm_g2_device
// (include/msm_g2.cuh:669-854). Same mathematics (signed c-bit digits,
// buckets, running-sum reduction), different machine design:
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
//     window and runs Horner across windows.
#pragma once
#include <stdint.h>

#include <vector>

#include "common.hpp"
#include "host_field.hpp"

namespace zk {

struct MsmGeom {
    int c;          // window bits
    int W;          // windows = ceil(255 / c)
    uint32_t K;     // buckets per window = 2^(c-1)
    uint32_t l1;    // buckets per level-1 reduction thread
    uint32_t nseg;  // level-1 segments per window = K / l1
    uint32_t nblk;  // level-2 blocks per window
};

int msm_default_c(size_t n);
MsmGeom msm_geom(size_t n, int c, bool g2);

// Per-scalar-vector digit sort (shared by the MSMs over the same scalars).
// Besides the entries in bucket order it produces the accumulation order: all
// W*K bucket ids sorted by descending load (quantised by 2^qs), so that the
// lanes of one wave walk buckets of nearly equal length. Buckets above lmax
// entries take the chunked path and come first.
struct MsmSort {
    MsmGeom g{};
    size_t n = 0;
    uint32_t lmax = 0, nkey = 0;
    int qs = 0;
    DevBuf matrix, mscan, pairs, offsets, entries, order, lkeys;
    // sorts canonical scalars (n x 32 B, device) into bucket order
    void run(const void* d_scalars, size_t n, int c, hipStream_t st);
};

// Result of the device part of one MSM: level-2 partials per window, on the host.
struct MsmPartials {
    MsmGeom g{};
    bool g2 = false;
    HostBuf host;       // [W][nblk] R then [W][nblk] T, XYZZ each
};

// Device scratch for the bucket phase of one MSM.
struct MsmScratch {
    DevBuf buckets, r1, t1, r2, t2, chunks, large, partials, counters;
    size_t cap_buckets = 0;
    // optional events around the bucket-accumulation kernel (profiling)
    hipEvent_t ev_acc0 = nullptr, ev_acc1 = nullptr, ev_end = nullptr;
    void enable_timing();
    ~MsmScratch();
    void ensure(const MsmGeom& g, size_t n_entries, bool g2);
};

// Launch bucket accumulation + GPU reduction for points (affine, Montgomery,
// AoS: G1 64 B / G2 128 B per point) against a finished sort; queue the D2H of
// the partials into out.host on st. The caller synchronises st before finish.
// d_points29: the same points in the 29-bit-limb form of msm_points_to29, in
// which the buckets are accumulated (d_points serves the oversized-bucket
// path); may be null only for n = 0.
void msm_launch(const MsmSort& sort, const void* d_points, bool g2, MsmScratch& scr, MsmPartials& out,
                hipStream_t st, const void* d_points29);

// affine points (standard form, G1 64 B / G2 128 B each) -> 29-bit-limb form
// (G1: 80 B of limbs in a 128 B slot; G2: 144 B in a 256 B slot)
constexpr size_t PT29_BYTES = 128, PT29Q_BYTES = 256;
void msm_points_to29(const void* d_points, size_t n, void* d_out, bool g2, hipStream_t st);

// Host finish: combine the partials into the MSM value.
zkh::Xyzz<zkh::Fp> msm_finish_g1(const MsmPartials& p);
zkh::Xyzz<zkh::Fp2> msm_finish_g2(const MsmPartials& p);

}  // namespace zk
