// Pippenger MSM kernels for G1/G2 on gfx950 (see msm.hpp for the design).
#include <algorithm>
#include <future>
#include <mutex>
#include <type_traits>

#include "msm_kernels.hpp"
#include "qap.hpp"
#include "scan.hpp"

namespace zk {

// window for an untuned size: the 2n GLV halves of n scalars favour about
// log2(n) - 3 bits (2^22 points: c = 19, 7 windows, measured 7% faster than 8)
int msm_default_c(size_t n) {
    int lg = ilog2(n ? n : 1);
    int c = lg - 3;
    if (c < 4) c = 4;
    if (c > 19) c = 19;
    return c;
}

int msm_default_split(size_t n, int c, bool g2, bool merged) {
    const MsmGeom g = msm_geom(n, c, g2, merged);
    // two rounds of waves at 6 (G1) and 3 (G2) waves per SIMD, i.e. three
    // and two rounds at the 4 and 2 the accumulation kernels hold (128 VGPRs;
    // 228 VGPRs and LDS); the autotuner picks the splits of a circuit's sizes
    const size_t target = (size_t)1024 * 64 * 2 * (g2 ? 3 : 6);
    const size_t nb = (size_t)g.Wr * g.K;
    return (int)std::max<size_t>(1, std::min<size_t>(3, target / nb));
}

MsmGeom msm_geom(size_t n, int c, bool g2, bool merged) {
    if (c <= 0) c = msm_default_c(n);
    ZK_REQUIRE(c >= 4 && c <= 22, ZK_E_ARG, "MSM window must be in [4, 22]");
    MsmGeom g;
    // W windows of the requested width at most, then the 128 bits shared
    // evenly: nlo windows of c - 1 bits, the others of c bits
    g.W = (GLV_BITS + c - 1) / c;
    c = (GLV_BITS + g.W - 1) / g.W;
    g.c = c;
    g.nlo = g.W * c - GLV_BITS;
    g.merged = merged;
    g.Wr = merged ? 1 : g.W;
    g.K = 1u << (c - 1);
    // level-1 segments: 4 G1 buckets up to 2^18 buckets in all (their
    // chains, split-fold longer, set the reduction's latency), else 8, which
    // keeps the level-2 partials the host combines to a few dozen per set
    const uint32_t l1 = (!g2 && (size_t)g.Wr * g.K <= (1u << 18)) ? 4u : L1_MAX;
    g.l1 = std::min<uint32_t>(l1, g.K);
    g.nseg = g.K / g.l1;
    uint32_t per_blk = (uint32_t)(g2 ? RedBlk<Fq2>::V : RedBlk<Fp>::V) * L2;
    g.nblk = (g.nseg + per_blk - 1) / per_blk;
    return g;
}

int msm_table_c(size_t n) {
    // relative costs: a mixed addition per entry 10 products, a bucket two
    // general additions of 14
    int best = 0;
    double best_cost = 0;
    for (int c = 11; c <= 22; c++) {
        const MsmGeom g = msm_geom(n, c, false, true);
        const double cost = 2.0 * (double)n * g.W * 10 + (double)g.K * 28;
        if (!best || cost < best_cost) {
            best = g.c;
            best_cost = cost;
        }
    }
    return best;
}

// ---------------------------------------------------------------- GLV split
// Every scalar k_i becomes two half scalars (glv.hpp): half[i] = k1 for P_i and
// half[n + i] = k2 for lambda*P_i, each a 127-bit magnitude with its sign in
// bit 127 (4 x u32 LE).
__global__ void k_glv_split(const uint32_t* __restrict__ sc, size_t n, uint4* __restrict__ half) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4* p = reinterpret_cast<const uint4*>(sc) + 2 * i;
    const uint4 a = p[0], b = p[1];
    const uint64_t k[4] = {a.x | (uint64_t)a.y << 32, a.z | (uint64_t)a.w << 32, b.x | (uint64_t)b.y << 32,
                           b.z | (uint64_t)b.w << 32};
    uint64_t k1[2], k2[2];
    glv::split(k, k1, k2);
    half[i] = make_uint4((uint32_t)k1[0], (uint32_t)(k1[0] >> 32), (uint32_t)k1[1], (uint32_t)(k1[1] >> 32));
    half[n + i] = make_uint4((uint32_t)k2[0], (uint32_t)(k2[0] >> 32), (uint32_t)k2[1], (uint32_t)(k2[1] >> 32));
}

// ---------------------------------------------------------------- digits
// Signed digits of a half scalar, least significant window first, window w
// being c - 1 bits wide for w < nlo and c bits above (128 bits in all):
// v = window bits + carry; v > 2^(width-1) becomes v - 2^width with a carry
// into the next window (include/msm_g1.cuh:271-329 uses the same recoding
// with one width). The magnitude is below 2^127, so the top window's v stays
// at most 2^(width-1): no carry out. Every magnitude fits the 2^(c-1) buckets
// of a window. A negative half scalar flips the sign of every digit.
struct DigitIter {
    uint32_t s[4];
    uint32_t carry, c, left;
    bool flip;
    ZK_DEV void init(const uint32_t* hs, size_t i, int cc, int nlo) {
        const uint4 a = reinterpret_cast<const uint4*>(hs)[i];
        s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w & 0x7fffffffu;
        flip = (a.w >> 31) != 0;
        carry = 0;
        c = (uint32_t)cc;
        left = (uint32_t)nlo;
    }
    // returns magnitude (0 = skip) and sign
    ZK_DEV uint32_t next(bool& neg) {
        const uint32_t wd = left ? c - 1 : c;
        left -= left ? 1u : 0u;
        const uint32_t mask = (1u << wd) - 1, half = 1u << (wd - 1);
        uint32_t v = (s[0] & mask) + carry;
#pragma unroll
        for (int k = 0; k < 3; k++) s[k] = __builtin_amdgcn_alignbit(s[k + 1], s[k], wd);
        s[3] >>= wd;
        if (v > half) {
            carry = 1;
            neg = !flip;
            return (mask + 1) - v;
        }
        carry = 0;
        neg = flip;
        return v;
    }
};

// ---------------------------------------------------------------- digit sort
// Two-pass LDS radix sort of the flat bucket ids b = w*Kw + |digit| - 1 of all
// nonzero digits (no global atomics; Kw = K, or 0 when all windows share one
// set of buckets). Buckets are grouped by b >> s.
// Pass A, per tile of scalars: an LDS histogram of groups goes to a count
// matrix [group][tile]; a device scan of the matrix gives every tile its run
// inside every group; the tile then writes (entry, b) pairs at run + LDS rank.
// Pass B, per group: LDS counts of its 2^s buckets give the bucket offsets,
// then the entries go to offset + LDS rank.
__global__ void __launch_bounds__(1024) k_sort_tile_hist(const uint32_t* __restrict__ sc, size_t n, size_t tile, int c,
                                                         int nlo, int W, uint32_t Kw, int s, uint32_t G, uint32_t nt,
                                                         uint32_t* __restrict__ matrix) {
    extern __shared__ uint32_t lh[];
    for (uint32_t k = threadIdx.x; k < G; k += blockDim.x) lh[k] = 0;
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * tile, i1 = i0 + tile < n ? i0 + tile : n;
    for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        DigitIter it;
        it.init(sc, i, c, nlo);
        for (int w = 0; w < W; w++) {
            bool ng;
            uint32_t mag = it.next(ng);
            if (mag) atomicAdd(&lh[((uint32_t)w * Kw + mag - 1) >> s], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < G; k += blockDim.x) matrix[(size_t)k * nt + blockIdx.x] = lh[k];
}

// n here counts half scalars (2 per scalar); entry of half i in window w:
// slot (i mod n/2) + w*sw (sw = n/2 for window tables, else 0), half
// i >= n/2, sign
__global__ void __launch_bounds__(1024) k_sort_tile_scatter(const uint32_t* __restrict__ sc, size_t n, size_t tile, int c,
                                                            int nlo, int W, uint32_t Kw, uint32_t sw, int s, uint32_t G,
                                                            uint32_t nt, const uint32_t* __restrict__ mscan,
                                                            const uint8_t* __restrict__ cls, uint2* __restrict__ pairs) {
    extern __shared__ uint32_t lds[];
    uint32_t* base = lds;       // this tile's run start in each group
    uint32_t* rank = lds + G;   // next rank in each group
    for (uint32_t k = threadIdx.x; k < G; k += blockDim.x) {
        base[k] = mscan[(size_t)k * nt + blockIdx.x];
        rank[k] = 0;
    }
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * tile, i1 = i0 + tile < n ? i0 + tile : n;
    const uint32_t nh = (uint32_t)(n >> 1);
    for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        DigitIter it;
        it.init(sc, i, c, nlo);
        const bool phi = (uint32_t)i >= nh;
        const uint32_t ref0 = ((uint32_t)i - (phi ? nh : 0u)) | (phi ? 0x40000000u : 0u);
        const uint32_t ctag = cls ? (uint32_t)cls[(uint32_t)i - (phi ? nh : 0u)] << 30 : 0u;
        for (int w = 0; w < W; w++) {
            bool ng;
            uint32_t mag = it.next(ng);
            if (mag) {
                const uint32_t b = (uint32_t)w * Kw + mag - 1, g = b >> s;
                const uint32_t pos = base[g] + atomicAdd(&rank[g], 1u);
                pairs[pos] = make_uint2((ref0 + (uint32_t)w * sw) | (ng ? 0x80000000u : 0u), b | ctag);
            }
        }
    }
}

// Entries of a group go to their final positions through an LDS stage of
// `stage_cap` words when the group fits (then the block writes its whole
// range with coalesced stores), else straight to global memory: scattered
// 4-byte stores into a range that other blocks do not touch.
constexpr uint32_t SORT_GROUP_LDS = 65536 - 256;   // dynamic LDS bytes of k_sort_group by default
// gfx950 gives one workgroup up to 160 KiB of LDS once the kernel asks for
// more than the default 64 KiB
constexpr uint32_t SORT_GROUP_LDS_MAX = 160 * 1024 - 1024;
// 16 waves per group: the LDS stage often leaves room for one workgroup per
// CU, whose waves then carry the pass's LDS atomics and loads alone
constexpr int SORT_GROUP_THREADS = 1024;
//
// With scalar classes (cls_off non-null; the class is in bits 30..31 of a
// pair's bucket word) the entries go out class by class: one pass over the
// group's pairs per class, each bucket's cursor then standing where the next
// class starts, which cls_off records.
__global__ void __launch_bounds__(1024) k_sort_group(const uint2* __restrict__ pairs, const uint32_t* __restrict__ mscan,
                                                    uint32_t nt, int s, uint32_t nb, uint32_t stage_cap,
                                                    uint32_t* __restrict__ offsets, uint32_t* __restrict__ entries,
                                                    uint32_t* __restrict__ cls_off) {
    extern __shared__ uint32_t cnt[];   // 2^s counters, then the stage
    __shared__ uint32_t wsum[16];
    const uint32_t g = blockIdx.x, F = 1u << s, mask = F - 1;
    const uint32_t start = mscan[(size_t)g * nt], end = mscan[(size_t)(g + 1) * nt];
    for (uint32_t f = threadIdx.x; f < F; f += blockDim.x) cnt[f] = 0;
    __syncthreads();
    for (uint32_t e = start + threadIdx.x; e < end; e += blockDim.x) atomicAdd(&cnt[pairs[e].y & mask], 1u);
    __syncthreads();
    // exclusive scan of the counters: per-thread chunks, then the block scan
    const uint32_t q = (F + blockDim.x - 1) / blockDim.x, f0 = threadIdx.x * q;
    const uint32_t f1 = f0 + q < F ? f0 + q : F;
    uint32_t sum = 0;
    for (uint32_t f = f0; f < f1; f++) sum += cnt[f];
    uint32_t total;
    uint32_t run = start + block_exclusive_scan(sum, wsum, &total);
    for (uint32_t f = f0; f < f1; f++) {
        const uint32_t v = cnt[f];
        cnt[f] = run;
        run += v;
    }
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < F; f += blockDim.x) {
        const uint32_t b = g * F + f;
        if (b < nb) offsets[b] = cnt[f];
    }
    if (g == gridDim.x - 1 && threadIdx.x == 0) offsets[nb] = end;
    __syncthreads();
    const bool staged = end - start <= stage_cap;
    uint32_t* stage = cnt + F;
    const int ncls = cls_off ? 4 : 1;
    for (int k = 0; k < ncls; k++) {
        if (k) {
            // every entry of classes < k is placed: cnt[f] is where class k starts
            __syncthreads();
            for (uint32_t f = threadIdx.x; f < F; f += blockDim.x) {
                const uint32_t b = g * F + f;
                if (b < nb) cls_off[3 * (size_t)b + (k - 1)] = cnt[f];
            }
            __syncthreads();
        }
        for (uint32_t e = start + threadIdx.x; e < end; e += blockDim.x) {
            const uint2 p = pairs[e];
            if (cls_off && (int)(p.y >> 30) != k) continue;
            const uint32_t pos = atomicAdd(&cnt[p.y & mask], 1u);
            if (staged) stage[pos - start] = p.x;
            else entries[pos] = p.x;
        }
    }
    if (staged) {
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < end - start; e += blockDim.x) entries[start + e] = stage[e];
    }
}

// accumulation-order key of bucket b: 0 = oversized (by the whole bucket),
// then descending length of the part the MSM reads (MsmRange over a sort with
// classes: the lanes of a wave get parts of nearly equal length), nkey-1 =
// empty
__device__ inline uint32_t load_key(const uint32_t* __restrict__ off, const uint32_t* __restrict__ cls_off, int range,
                                    uint32_t b, uint32_t lmax, int qs, uint32_t nkey) {
    uint32_t s0 = off[b], e0 = off[b + 1];
    if (e0 - s0 > lmax) return 0u;
    if (range == MSM_RANGE_A) {
        s0 = cls_off[3 * (size_t)b];
        e0 = cls_off[3 * (size_t)b + 2];
    } else if (range == MSM_RANGE_B) {
        e0 = cls_off[3 * (size_t)b + 1];
    }
    return (nkey - 1) - ((e0 - s0) >> qs);
}

// With classes (cls_off non-null) blockIdx.y = 0, 1, 2 orders the buckets by
// the length of their MSM_RANGE_A, MSM_RANGE_B and whole parts, each with its
// own nkey histogram (hist + y * nkey) and cursors; every order holds all nb
// buckets, so one exclusive scan over the three histograms gives order y its
// positions offset by y * nb.
__device__ inline int order_range(const uint32_t* cls_off) {
    return !cls_off ? MSM_RANGE_ALL : blockIdx.y == 0 ? MSM_RANGE_A : blockIdx.y == 1 ? MSM_RANGE_B : MSM_RANGE_ALL;
}

__global__ void __launch_bounds__(1024) k_load_hist(const uint32_t* __restrict__ off, const uint32_t* __restrict__ cls_off,
                                                    uint32_t nb, uint32_t lmax, int qs, uint32_t nkey,
                                                    uint32_t* __restrict__ hist) {
    extern __shared__ uint32_t lh[];
    const int range = order_range(cls_off);
    hist += (size_t)blockIdx.y * nkey;
    for (uint32_t k = threadIdx.x; k < nkey; k += blockDim.x) lh[k] = 0;
    __syncthreads();
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) atomicAdd(&lh[load_key(off, cls_off, range, b, lmax, qs, nkey)], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nkey; k += blockDim.x)
        if (lh[k]) atomicAdd(&hist[k], lh[k]);
}

// each block reserves one range per key, then ranks its buckets inside it in LDS
__global__ void __launch_bounds__(1024) k_load_scatter(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ cls_off, uint32_t nb,
                                                       uint32_t lmax, int qs, uint32_t nkey,
                                                       uint32_t* __restrict__ cursor, uint32_t* __restrict__ order_a,
                                                       uint32_t* __restrict__ order_b, uint32_t* __restrict__ order) {
    extern __shared__ uint32_t lh[];
    const int range = order_range(cls_off);
    cursor += (size_t)blockIdx.y * nkey;
    uint32_t* out = range == MSM_RANGE_A ? order_a : range == MSM_RANGE_B ? order_b : order;
    const uint32_t base = blockIdx.y * nb;   // the scan's offset of order y
    for (uint32_t k = threadIdx.x; k < nkey; k += blockDim.x) lh[k] = 0;
    __syncthreads();
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = 0, rank = 0;
    if (b < nb) {
        key = load_key(off, cls_off, range, b, lmax, qs, nkey);
        rank = atomicAdd(&lh[key], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nkey; k += blockDim.x)
        if (lh[k]) lh[k] = atomicAdd(&cursor[k], lh[k]);
    __syncthreads();
    if (b < nb) out[lh[key] + rank - base] = b;
}

// Chunk lists of the oversized buckets (load key 0, first in the accumulation
// order): one thread per oversized bucket, grid-stride; hist[0] is their count.
__global__ void __launch_bounds__(256) k_chunk_list(const uint32_t* __restrict__ order, const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ hist, uint32_t* __restrict__ counters,
                                                    ChunkRec* __restrict__ chunks, LargeRec* __restrict__ large) {
    const uint32_t n_over = hist[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_over; i += gridDim.x * blockDim.x) {
        const uint32_t b = order[i];
        const uint32_t s0 = off[b], e0 = off[b + 1];
        const uint32_t nch = (e0 - s0 + CHUNK - 1) / CHUNK;
        const uint32_t first = atomicAdd(&counters[0], nch);
        for (uint32_t j = 0; j < nch; j++) {
            const uint32_t c0 = s0 + j * CHUNK;
            chunks[first + j] = {b, c0, min(CHUNK, e0 - c0), 0u};
        }
        const uint32_t li = atomicAdd(&counters[1], 1u);
        large[li] = {b, first, nch, 0u};
    }
}

// sizes of one digit sort over n scalars at window c (shared by reserve and run)
struct SortShape {
    MsmGeom g;
    size_t n2, nb, tile, mlen, ne, max_chunks, max_large;
    size_t group_max;   // expected entries of the heaviest group (uniform scalars)
    uint32_t G, nt, lmax, nkey;
    int sh, qs;
};

static SortShape sort_shape(size_t n, int c, bool merged) {
    ZK_REQUIRE(n < 0x40000000ull, ZK_E_ARG, "MSM size must be < 2^30");
    SortShape z;
    z.g = msm_geom(n, c, false, merged);
    // entries name the slot w*n + i of window w's table in 30 bits
    ZK_REQUIRE(!merged || (size_t)z.g.W * n < 0x40000000ull, ZK_E_ARG, "MSM over window tables needs W*n < 2^30");
    // the sort runs over the 2n half scalars of the GLV split
    z.n2 = 2 * n;
    z.nb = (size_t)z.g.Wr * z.g.K;
    // group size 2^s: at most 4096 groups (LDS histograms of pass A)
    z.sh = 7;
    while (((z.nb + (1ull << z.sh) - 1) >> z.sh) > 4096) z.sh++;
    z.G = (uint32_t)((z.nb + (1ull << z.sh) - 1) >> z.sh);
    // tiles of pass A: multiples of 4096 scalars, at most 512 tiles (two
    // workgroups per CU): longer runs per group and tile for the pair writes,
    // whose short runs wrote several times their bytes
    z.tile = (size_t)4096 * std::max<size_t>(1, (z.n2 + 4096 * 512 - 1) / (4096 * 512));
    z.nt = (uint32_t)std::max<size_t>(1, (z.n2 + z.tile - 1) / z.tile);
    z.mlen = (size_t)z.G * z.nt + 1;
    z.ne = std::max<size_t>(1, (size_t)z.g.W * z.n2);
    // accumulation order. Oversized threshold: 8x the mean bucket load
    // (uniform scalars put about ne/nb entries in a bucket).
    z.lmax = std::max<uint32_t>(L_MIN_LARGE, 8u * (uint32_t)((z.ne + z.nb - 1) / z.nb));
    z.qs = 0;
    while ((z.lmax >> z.qs) + 2 > 4096) z.qs++;
    z.nkey = (z.lmax >> z.qs) + 2;
    z.max_chunks = z.ne / CHUNK + z.ne / L_MIN_LARGE + 2;
    z.max_large = z.ne / L_MIN_LARGE + 2;
    // a window of width wd spreads the 2n digits of its halves over 2^(wd-1)
    // magnitudes; the lowest buckets carry the most: those of the narrower
    // windows, or of every window when they share one bucket set
    double per_b = 0;
    for (int w = 0; w < z.g.W; w++) {
        const int wd = w < z.g.nlo ? z.g.c - 1 : z.g.c;
        const double load = (double)z.n2 / (double)(1ull << (wd - 1));
        per_b = merged ? per_b + load : std::max(per_b, load);
    }
    z.group_max = (size_t)(per_b * (double)(1ull << z.sh));
    return z;
}

// every buffer of a sort with its size for shape z: f(buffer, bytes)
template <class S, class F>
static void sort_buffers(S& s, const SortShape& z, F&& f) {
    f(s.half, std::max<size_t>(1, z.n2) * 16);
    f(s.matrix, z.mlen * 4);
    f(s.mscan, z.mlen * 4 + (size_t)4 * (cdiv(z.mlen, SCAN_TILE) + 1));
    f(s.pairs, z.ne * 8);
    f(s.entries, z.ne * 4);
    f(s.offsets, (z.nb + 1) * 4);
    f(s.order, z.nb * 4);
    f(s.lkeys, (size_t)4 * (2 * z.nkey + cdiv(z.nkey, SCAN_TILE) + 1));
    f(s.chunks, z.max_chunks * sizeof(ChunkRec));
    f(s.large, z.max_large * sizeof(LargeRec));
    f(s.counters, (size_t)64);
}

// the buffers of a sort with scalar classes over nb buckets
template <class S, class F>
static void class_buffers(S& s, const SortShape& z, F&& f) {
    const size_t nb = z.nb, nkey = z.nkey;
    f(s.cls_off, nb * 12);
    f(s.order_a, nb * 4);
    f(s.order_b, nb * 4);
    f(s.okeys, (size_t)4 * (6 * (size_t)nkey + cdiv(3 * (size_t)nkey, SCAN_TILE) + 1));
}

static void sort_reserve(MsmSort& s, const SortShape& z) {
    sort_buffers(s, z, [](DevBuf& b, size_t n) { b.ensure(n); });
}

size_t MsmSort::reserve(size_t n_, int c, bool merged, bool with_classes) {
    const SortShape z = sort_shape(n_, c, merged);
    sort_reserve(*this, z);
    if (with_classes) class_buffers(*this, z, [](DevBuf& b, size_t n) { b.ensure(n); });
    return z.max_chunks;
}

size_t MsmSort::count_reserve(size_t n_, int c, GrowCount& gc, bool merged, bool with_classes) const {
    const SortShape z = sort_shape(n_, c, merged);
    sort_buffers(*this, z, [&](const DevBuf& b, size_t n) { gc.need(b, n); });
    if (with_classes) class_buffers(*this, z, [&](const DevBuf& b, size_t n) { gc.need(b, n); });
    return z.max_chunks;
}

void MsmSort::run(const void* d_scalars, size_t n_, int c, hipStream_t st, bool merged, const uint8_t* cls) {
    const SortShape z = sort_shape(n_, c, merged);
    sort_reserve(*this, z);
    if (cls) class_buffers(*this, z, [](DevBuf& b, size_t n) { b.ensure(n); });
    n = n_;
    classes = cls != nullptr && n_ != 0;
    g = z.g;
    lmax = z.lmax;
    qs = z.qs;
    nkey = z.nkey;
    n_chunk_max = z.max_chunks;
    const size_t n2 = z.n2, nb = z.nb, tile = z.tile, mlen = z.mlen;
    const uint32_t G = z.G, nt = z.nt;
    const int sh = z.sh;
    if (n) {
        k_glv_split<<<cdiv(n, 256), 256, 0, st>>>((const uint32_t*)d_scalars, n, half.as<uint4>());
        ZK_HIP(hipGetLastError());
    }
    const uint32_t* hs = half.as<uint32_t>();
    uint32_t* mat = matrix.as<uint32_t>();
    uint32_t* msc = mscan.as<uint32_t>();
    if (n) {
        k_sort_tile_hist<<<nt, 1024, G * 4, st>>>(hs, n2, tile, g.c, g.nlo, g.W, merged ? 0u : g.K, sh, G, nt, mat);
        ZK_HIP(hipGetLastError());
        fill_u32(mat + mlen - 1, 0u, 1, st);
    } else {
        fill_u32(mat, 0u, mlen, st);
    }
    exclusive_scan(mat, msc, msc + mlen, mlen, st);
    if (n) {
        k_sort_tile_scatter<<<nt, 1024, 2 * G * 4, st>>>(hs, n2, tile, g.c, g.nlo, g.W, merged ? 0u : g.K,
                                                          merged ? (uint32_t)n : 0u, sh, G, nt, msc, cls,
                                                          pairs.as<uint2>());
        ZK_HIP(hipGetLastError());
    }
    // stage for the heaviest group uniform scalars give, plus a margin
    // (group sizes are sums of many digits, so larger ones are rare and take
    // the global path); above 64 KiB of LDS the kernel is allowed up to
    // 160 KiB, or keeps 64 KiB if the runtime refuses
    ZK_REQUIRE((4u << sh) < SORT_GROUP_LDS / 2, ZK_E_INTERNAL, "sort group counters exceed the LDS budget");
    const size_t want = z.group_max + z.group_max / 16 + 256;
    uint32_t lds_budget = SORT_GROUP_LDS;
    if (4 * ((1u << sh) + want) > SORT_GROUP_LDS) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sort_group), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)SORT_GROUP_LDS_MAX) == hipSuccess)
            lds_budget = SORT_GROUP_LDS_MAX;
        else
            (void)hipGetLastError();
    }
    const uint32_t stage_cap = (uint32_t)std::min<size_t>((lds_budget - (4u << sh)) / 4, want);
    k_sort_group<<<G, SORT_GROUP_THREADS, 4u * ((1u << sh) + stage_cap), st>>>(pairs.as<uint2>(), msc, nt, sh, (uint32_t)nb, stage_cap,
                                                                offsets.as<uint32_t>(), entries.as<uint32_t>(),
                                                                cls ? cls_off.as<uint32_t>() : nullptr);
    ZK_HIP(hipGetLastError());
    const uint32_t* off = offsets.as<uint32_t>();
    uint32_t* hist = lkeys.as<uint32_t>();
    const unsigned nblk_o = (unsigned)cdiv(nb, 1024);
    const uint32_t* co = classes ? cls_off.as<uint32_t>() : nullptr;
    // accumulation orders: the whole buckets', and with classes also those of
    // the A and B parts, in one histogram, scan and scatter
    const uint32_t nord = classes ? 3 : 1;
    uint32_t* h = classes ? okeys.as<uint32_t>() : hist;
    uint32_t* cur = h + nord * nkey;
    fill_u32(h, 0u, nord * nkey, st);
    k_load_hist<<<dim3(nblk_o, nord), 1024, nkey * 4, st>>>(off, co, (uint32_t)nb, lmax, qs, nkey, h);
    ZK_HIP(hipGetLastError());
    exclusive_scan(h, cur, cur + nord * nkey, nord * nkey, st);
    k_load_scatter<<<dim3(nblk_o, nord), 1024, nkey * 4, st>>>(off, co, (uint32_t)nb, lmax, qs, nkey, cur,
                                                              order_a.as<uint32_t>(), order_b.as<uint32_t>(),
                                                              order.as<uint32_t>());
    ZK_HIP(hipGetLastError());
    hist = h + (nord - 1) * nkey;   // the whole buckets' histogram: hist[0] oversized
    // chunk lists of the oversized buckets, shared by every MSM over this sort
    fill_u32(counters.as<uint32_t>(), 0u, 16, st);
    k_chunk_list<<<64, 256, 0, st>>>(order.as<uint32_t>(), off, hist, counters.as<uint32_t>(), chunks.as<ChunkRec>(),
                                     large.as<LargeRec>());
    ZK_HIP(hipGetLastError());
}

void msm_reserve_partials(MsmPartials& out, const MsmGeom& g, bool g2) {
    out.host.ensure(2 * (size_t)g.Wr * g.nblk * (g2 ? 256 : 128));
}

void msm_points_to29(const void* d_points, size_t n, void* d_out, bool g2, hipStream_t st) {
    if (!n) return;
    ZK_REQUIRE(n < 0x40000000ull, ZK_E_ARG, "MSM size must be < 2^30");
    if (g2) msm_points_to29_g2(d_points, n, d_out, st);
    else msm_points_to29_g1(d_points, n, d_out, st);
}

void msm_points_table(const void* d_points, size_t n, int c, void* d_out, bool g2, hipStream_t st) {
    if (!n) return;
    const MsmGeom g = msm_geom(n, c, g2, true);
    ZK_REQUIRE((size_t)g.W * n < 0x40000000ull, ZK_E_ARG, "window tables need W*n < 2^30");
    msm_points_to29(d_points, n, d_out, g2, st);
    if (g2) msm_points_table_g2(d_points, n, g, d_out, st);
    else msm_points_table_g1(d_points, n, g, d_out, st);
}

void MsmScratch::enable_timing() {
    if (ev_acc0) return;
    ZK_HIP(hipEventCreate(&ev_acc0));
    ZK_HIP(hipEventCreate(&ev_acc1));
    ZK_HIP(hipEventCreate(&ev_end));
}

MsmScratch::~MsmScratch() {
    if (ev_acc0) (void)hipEventDestroy(ev_acc0);
    if (ev_acc1) (void)hipEventDestroy(ev_acc1);
    if (ev_end) (void)hipEventDestroy(ev_end);
}

template <class S, class F>
static void scratch_buffers(S& q, const MsmGeom& g, size_t max_chunks, bool g2, int split, F&& f) {
    const size_t xb = g2 ? 256 : 128;
    const size_t nb = (size_t)g.Wr * g.K;
    f(q.buckets, nb * split * xb);
    f(q.r1, (size_t)g.Wr * g.nseg * xb);
    f(q.t1, (size_t)g.Wr * g.nseg * xb);
    f(q.r2, (size_t)g.Wr * g.nblk * xb);
    f(q.t2, (size_t)g.Wr * g.nblk * xb);
    f(q.partials, max_chunks * xb);
}

void MsmScratch::ensure(const MsmGeom& g, size_t max_chunks, bool g2, int split) {
    scratch_buffers(*this, g, max_chunks, g2, split, [](DevBuf& b, size_t n) { b.ensure(n); });
}

void MsmScratch::count_ensure(const MsmGeom& g, size_t max_chunks, bool g2, int split, GrowCount& gc) const {
    scratch_buffers(*this, g, max_chunks, g2, split, [&](const DevBuf& b, size_t n) { gc.need(b, n); });
}

void msm_launch_pre(const MsmSort& sort, const void* d_points29, bool g2, MsmScratch& scr, hipStream_t st, int split) {
    ZK_REQUIRE(split >= 1 && split <= MSM_SPLIT_MAX, ZK_E_ARG, "bucket split must be in [1, 4]");
    ZK_REQUIRE(sort.n == 0 || d_points29, ZK_E_INTERNAL, "MSM launch without the 29-bit copy of its points");
    if (g2) msm_launch_pre_g2(sort, d_points29, scr, st, split);
    else msm_launch_pre_g1(sort, d_points29, scr, st, split);
}

void msm_launch_main(const MsmSort& sort, bool g2, MsmScratch& scr, MsmPartials& out, hipStream_t st,
                     const void* d_points29, int range) {
    ZK_REQUIRE(sort.n == 0 || d_points29, ZK_E_INTERNAL, "MSM launch without the 29-bit copy of its points");
    ZK_REQUIRE(scr.split >= 1, ZK_E_INTERNAL, "MSM accumulation queued before its chunked path");
    ZK_REQUIRE(range >= MSM_RANGE_ALL && range <= MSM_RANGE_B, ZK_E_INTERNAL, "unknown MSM entry range");
    if (!sort.classes) range = MSM_RANGE_ALL;
    if (g2) msm_launch_main_g2(sort, scr, out, st, d_points29, range);
    else msm_launch_main_g1(sort, scr, out, st, d_points29, range);
}

void msm_launch(const MsmSort& sort, bool g2, MsmScratch& scr, MsmPartials& out, hipStream_t st,
                const void* d_points29, int split) {
    msm_launch_pre(sort, d_points29, g2, scr, st, split);
    msm_launch_main(sort, g2, scr, out, st, d_points29);
}

// ---------------------------------------------------------------- fixed-scratch G1 MSM
// The whole device workspace of msm_g1_noalloc: allocated with the code
// object on every device the library runs on, held until the process ends.
static __device__ __align__(256) uint8_t scratch[MSM_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
// host partials of the fixed-scratch MSM: [W][nblk] R then T, XYZZ (128 B) each
static uint64_t g_noalloc_host[16384];

void msm_noalloc_setup(int device) {
    ZK_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, ZK_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return;
    void* ptr = nullptr;
    // the host-side name of a __device__ array is not its device address
    ZK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_noalloc_ptr[device] = ptr;
}

zkh::Xyzz<zkh::Fp> msm_g1_noalloc(int device, const void* d_points, const void* d_scalars, size_t n, int window,
                                  hipStream_t st) {
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        if (device >= 0 && device < NOALLOC_MAX_DEVICES) base = g_noalloc_ptr[device];
    }
    ZK_REQUIRE(base != nullptr, ZK_E_STATE, "fixed MSM scratch not set up on this device");
    const SortShape z = sort_shape(n, window, false);
    const MsmGeom& g = z.g;
    const int split = msm_default_split(n, g.c, false);
    // the same buffers the allocating path sizes, laid out in the array
    MsmSort sort;
    MsmScratch scr;
    DevBuf p29;
    size_t need = 0;
    auto count = [&](const DevBuf&, size_t b) { need = ((need + 255) & ~(size_t)255) + b; };
    sort_buffers(sort, z, count);
    scratch_buffers(scr, g, z.max_chunks, false, split, count);
    count(p29, std::max<size_t>(n, 1) * PT29_BYTES);
    ZK_REQUIRE(need <= MSM_NOALLOC_SCRATCH_BYTES, ZK_E_ARG,
               "MSM of " + std::to_string(n) + " points at window " + std::to_string(g.c) + " needs " +
                   std::to_string(need) + " bytes of workspace; the fixed scratch holds " +
                   std::to_string(MSM_NOALLOC_SCRATCH_BYTES) + " (65536 points at the default window)");
    const size_t host_need = 2 * (size_t)g.Wr * g.nblk * 128;
    ZK_REQUIRE(host_need <= sizeof g_noalloc_host, ZK_E_ARG, "MSM window too large for the fixed host partials");
    size_t used = 0;
    auto take = [&](DevBuf& b, size_t bytes) {
        used = (used + 255) & ~(size_t)255;
        b.borrow((char*)base + used, bytes);
        used += bytes;
    };
    sort_buffers(sort, z, take);
    scratch_buffers(scr, g, z.max_chunks, false, split, take);
    take(p29, std::max<size_t>(n, 1) * PT29_BYTES);
    MsmPartials part;
    part.ext = g_noalloc_host;
    part.ext_bytes = sizeof g_noalloc_host;
    if (n) msm_points_to29(d_points, n, p29.p, false, st);
    sort.run(d_scalars, n, window, st);
    msm_launch(sort, false, scr, part, st, p29.p, split);
    ZK_HIP(hipStreamSynchronize(st));
    return msm_finish_g1(part);
}

// ---------------------------------------------------------------- host finish
template <class HT>
static HT load_h(const uint64_t* p);
template <>
zkh::Fp load_h<zkh::Fp>(const uint64_t* p) { return zkh::Fp::from_raw(p); }
template <>
zkh::Fp2 load_h<zkh::Fp2>(const uint64_t* p) { return {zkh::Fp::from_raw(p), zkh::Fp::from_raw(p + 4)}; }

// value of bucket set w from its level-2 partials: sum_blk (T_blk +
// blk*U*R_blk) with U = 2^logU, by running sums over the blocks
template <class HT>
static zkh::Xyzz<HT> window_value(const MsmPartials& p, int w) {
    const MsmGeom& g = p.g;
    const int lw = sizeof(HT) / 8;  // u64 words per coordinate
    const uint64_t* R = p.data();
    const uint64_t* Tt = R + (size_t)g.Wr * g.nblk * 4 * lw;
    const int logU = ilog2(g.l1) + ilog2(L2) + ilog2(p.g2 ? RedBlk<Fq2>::V : RedBlk<Fp>::V);
    auto ld = [&](const uint64_t* base, size_t i) {
        const uint64_t* q = base + i * 4 * lw;
        return zkh::Xyzz<HT>{load_h<HT>(q), load_h<HT>(q + lw), load_h<HT>(q + 2 * lw), load_h<HT>(q + 3 * lw)};
    };
    zkh::Xyzz<HT> run = zkh::Xyzz<HT>::infinity(), acc = run, tsum = run;
    for (int b = (int)g.nblk - 1; b >= 0; b--) {
        acc = zkh::xyzz_add(acc, run);
        run = zkh::xyzz_add(run, ld(R, (size_t)w * g.nblk + b));
        tsum = zkh::xyzz_add(tsum, ld(Tt, (size_t)w * g.nblk + b));
    }
    for (int d = 0; d < logU; d++) acc = zkh::xyzz_dbl(acc);
    return zkh::xyzz_add(tsum, acc);
}

// window values (on `tasks` threads when > 1), then Horner across windows;
// over window tables the one bucket set is the whole MSM
template <class HT>
static zkh::Xyzz<HT> finish_t(const MsmPartials& p, int tasks) {
    const MsmGeom& g = p.g;
    if (g.merged) return window_value<HT>(p, 0);
    std::vector<zkh::Xyzz<HT>> S(g.W);
    if (tasks <= 1 || g.W < 2) {
        for (int w = 0; w < g.W; w++) S[w] = window_value<HT>(p, w);
    } else {
        std::vector<std::future<void>> fs;
        for (int t = 0; t < tasks; t++)
            fs.push_back(std::async(std::launch::async, [&, t] {
                for (int w = t; w < g.W; w += tasks) S[w] = window_value<HT>(p, w);
            }));
        for (auto& f : fs) f.get();
    }
    // total = sum_w 2^(offset of w) S_w; window w is c - 1 bits wide below nlo
    zkh::Xyzz<HT> total = zkh::Xyzz<HT>::infinity();
    for (int w = g.W - 1; w >= 0; w--) {
        const int width = w < g.nlo ? g.c - 1 : g.c;
        for (int d = 0; d < width; d++) total = zkh::xyzz_dbl(total);
        total = zkh::xyzz_add(total, S[w]);
    }
    return total;
}

zkh::Xyzz<zkh::Fp> msm_finish_g1(const MsmPartials& p) { return finish_t<zkh::Fp>(p, 1); }
// G2 window values cost about 4x those of G1: spread them over 4 threads
zkh::Xyzz<zkh::Fp2> msm_finish_g2(const MsmPartials& p) { return finish_t<zkh::Fp2>(p, 4); }

}  // namespace zk
