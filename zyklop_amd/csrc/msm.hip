// Pippenger MSM kernels for G1/G2 on gfx950 (see msm.hpp for the design).
#include <algorithm>
#include <future>
#include <type_traits>

#include "curve.hpp"
#include "field29.hpp"
#include "glv.hpp"
#include "msm.hpp"
#include "qap.hpp"
#include "scan.hpp"

namespace zk {

constexpr uint32_t L_MIN_LARGE = 64;  // floor of the oversized-bucket threshold
constexpr uint32_t CHUNK = 256 * 16;  // entries per chunk of an oversized bucket
constexpr uint32_t L1_MAX = 8;        // buckets per level-1 reduction thread (fewer if K is smaller)
constexpr int L2 = 4;                 // level-1 segments per level-2 thread
constexpr int LOG_L2 = 2;
template <class T> struct RedBlk;
template <> struct RedBlk<Fp> { static constexpr int V = 256; };
template <> struct RedBlk<Fq2> { static constexpr int V = 128; };

struct ChunkRec { uint32_t bucket, start, len, pad; };
struct LargeRec { uint32_t bucket, first, count, pad; };

// window for an untuned size: the 2n GLV halves of n scalars favour about
// log2(n) - 3 bits (2^22 points: c = 19, 7 windows, measured 7% faster than 8)
int msm_default_c(size_t n) {
    int lg = ilog2(n ? n : 1);
    int c = lg - 3;
    if (c < 4) c = 4;
    if (c > 19) c = 19;
    return c;
}

int msm_default_split(size_t n, int c, bool g2) {
    const MsmGeom g = msm_geom(n, c, g2);
    // resident waves per SIMD: G1 accumulation 6 (80 VGPRs), G2 about 3 (LDS)
    const size_t target = (size_t)1024 * 64 * 2 * (g2 ? 3 : 6);
    const size_t nb = (size_t)g.W * g.K;
    return (int)std::max<size_t>(1, std::min<size_t>(3, target / nb));
}

MsmGeom msm_geom(size_t n, int c, bool g2) {
    if (c <= 0) c = msm_default_c(n);
    ZK_REQUIRE(c >= 4 && c <= 22, ZK_E_ARG, "MSM window must be in [4, 22]");
    MsmGeom g;
    // W windows of the requested width at most, then the 128 bits shared
    // evenly: nlo windows of c - 1 bits, the others of c bits
    g.W = (GLV_BITS + c - 1) / c;
    c = (GLV_BITS + g.W - 1) / g.W;
    g.c = c;
    g.nlo = g.W * c - GLV_BITS;
    g.K = 1u << (c - 1);
    g.l1 = std::min<uint32_t>(L1_MAX, g.K);
    g.nseg = g.K / g.l1;
    uint32_t per_blk = (uint32_t)(g2 ? RedBlk<Fq2>::V : RedBlk<Fp>::V) * L2;
    g.nblk = (g.nseg + per_blk - 1) / per_blk;
    return g;
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
// Two-pass LDS radix sort of the flat bucket ids b = w*K + |digit| - 1 of all
// nonzero digits (no global atomics). Buckets are grouped by b >> s.
// Pass A, per tile of scalars: an LDS histogram of groups goes to a count
// matrix [group][tile]; a device scan of the matrix gives every tile its run
// inside every group; the tile then writes (entry, b) pairs at run + LDS rank.
// Pass B, per group: LDS counts of its 2^s buckets give the bucket offsets,
// then the entries go to offset + LDS rank.
__global__ void __launch_bounds__(1024) k_sort_tile_hist(const uint32_t* __restrict__ sc, size_t n, size_t tile, int c,
                                                         int nlo, int W, uint32_t K, int s, uint32_t G, uint32_t nt,
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
            if (mag) atomicAdd(&lh[((uint32_t)w * K + mag - 1) >> s], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < G; k += blockDim.x) matrix[(size_t)k * nt + blockIdx.x] = lh[k];
}

__global__ void __launch_bounds__(1024) k_sort_tile_scatter(const uint32_t* __restrict__ sc, size_t n, size_t tile, int c,
                                                            int nlo, int W, uint32_t K, int s, uint32_t G, uint32_t nt,
                                                            const uint32_t* __restrict__ mscan, uint2* __restrict__ pairs) {
    extern __shared__ uint32_t lds[];
    uint32_t* base = lds;       // this tile's run start in each group
    uint32_t* rank = lds + G;   // next rank in each group
    for (uint32_t k = threadIdx.x; k < G; k += blockDim.x) {
        base[k] = mscan[(size_t)k * nt + blockIdx.x];
        rank[k] = 0;
    }
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * tile, i1 = i0 + tile < n ? i0 + tile : n;
    for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        DigitIter it;
        it.init(sc, i, c, nlo);
        for (int w = 0; w < W; w++) {
            bool ng;
            uint32_t mag = it.next(ng);
            if (mag) {
                const uint32_t b = (uint32_t)w * K + mag - 1, g = b >> s;
                const uint32_t pos = base[g] + atomicAdd(&rank[g], 1u);
                pairs[pos] = make_uint2((uint32_t)i | (ng ? 0x80000000u : 0u), b);
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_sort_group(const uint2* __restrict__ pairs, const uint32_t* __restrict__ mscan,
                                                    uint32_t nt, int s, uint32_t nb, uint32_t* __restrict__ offsets,
                                                    uint32_t* __restrict__ entries) {
    extern __shared__ uint32_t cnt[];   // 2^s counters
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
    for (uint32_t e = start + threadIdx.x; e < end; e += blockDim.x) {
        const uint2 p = pairs[e];
        entries[atomicAdd(&cnt[p.y & mask], 1u)] = p.x;
    }
}

// accumulation-order key of bucket b: 0 = oversized, then descending load,
// nkey-1 = empty
__device__ inline uint32_t load_key(const uint32_t* __restrict__ off, uint32_t b, uint32_t lmax, int qs, uint32_t nkey) {
    uint32_t len = off[b + 1] - off[b];
    return len > lmax ? 0u : (nkey - 1) - (len >> qs);
}

__global__ void __launch_bounds__(1024) k_load_hist(const uint32_t* __restrict__ off, uint32_t nb, uint32_t lmax, int qs,
                                                    uint32_t nkey, uint32_t* __restrict__ hist) {
    extern __shared__ uint32_t lh[];
    for (uint32_t k = threadIdx.x; k < nkey; k += blockDim.x) lh[k] = 0;
    __syncthreads();
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) atomicAdd(&lh[load_key(off, b, lmax, qs, nkey)], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nkey; k += blockDim.x)
        if (lh[k]) atomicAdd(&hist[k], lh[k]);
}

// each block reserves one range per key, then ranks its buckets inside it in LDS
__global__ void __launch_bounds__(1024) k_load_scatter(const uint32_t* __restrict__ off, uint32_t nb, uint32_t lmax, int qs,
                                                       uint32_t nkey, uint32_t* __restrict__ cursor,
                                                       uint32_t* __restrict__ order) {
    extern __shared__ uint32_t lh[];
    for (uint32_t k = threadIdx.x; k < nkey; k += blockDim.x) lh[k] = 0;
    __syncthreads();
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = 0, rank = 0;
    if (b < nb) {
        key = load_key(off, b, lmax, qs, nkey);
        rank = atomicAdd(&lh[key], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nkey; k += blockDim.x)
        if (lh[k]) lh[k] = atomicAdd(&cursor[k], lh[k]);
    __syncthreads();
    if (b < nb) order[lh[key] + rank] = b;
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
    uint32_t G, nt, lmax, nkey;
    int sh, qs;
};

static SortShape sort_shape(size_t n, int c) {
    ZK_REQUIRE(n < 0x40000000ull, ZK_E_ARG, "MSM size must be < 2^30");
    SortShape z;
    z.g = msm_geom(n, c, false);
    // the sort runs over the 2n half scalars of the GLV split
    z.n2 = 2 * n;
    z.nb = (size_t)z.g.W * z.g.K;
    // group size 2^s: at most 4096 groups (LDS histograms of pass A)
    z.sh = 7;
    while (((z.nb + (1ull << z.sh) - 1) >> z.sh) > 4096) z.sh++;
    z.G = (uint32_t)((z.nb + (1ull << z.sh) - 1) >> z.sh);
    // tiles of pass A: multiples of 4096 scalars (long runs per group for the
    // pair writes), at most 2048 tiles
    z.tile = (size_t)4096 * std::max<size_t>(1, (z.n2 + 4096 * 2048 - 1) / (4096 * 2048));
    z.nt = (uint32_t)std::max<size_t>(1, (z.n2 + z.tile - 1) / z.tile);
    z.mlen = (size_t)z.G * z.nt + 1;
    z.ne = std::max<size_t>(1, (size_t)z.g.W * z.n2);
    // accumulation order. Oversized threshold: 8x the mean bucket load
    // (uniform scalars put about 2n/K entries in a bucket).
    z.lmax = std::max<uint32_t>(L_MIN_LARGE, 8u * (uint32_t)((z.n2 + z.g.K - 1) / z.g.K));
    z.qs = 0;
    while ((z.lmax >> z.qs) + 2 > 4096) z.qs++;
    z.nkey = (z.lmax >> z.qs) + 2;
    z.max_chunks = z.ne / CHUNK + z.ne / L_MIN_LARGE + 2;
    z.max_large = z.ne / L_MIN_LARGE + 2;
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

static void sort_reserve(MsmSort& s, const SortShape& z) {
    sort_buffers(s, z, [](DevBuf& b, size_t n) { b.ensure(n); });
}

size_t MsmSort::reserve(size_t n_, int c) {
    const SortShape z = sort_shape(n_, c);
    sort_reserve(*this, z);
    return z.max_chunks;
}

size_t MsmSort::count_reserve(size_t n_, int c, GrowCount& gc) const {
    const SortShape z = sort_shape(n_, c);
    sort_buffers(*this, z, [&](const DevBuf& b, size_t n) { gc.need(b, n); });
    return z.max_chunks;
}

void MsmSort::run(const void* d_scalars, size_t n_, int c, hipStream_t st) {
    const SortShape z = sort_shape(n_, c);
    sort_reserve(*this, z);
    n = n_;
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
        k_sort_tile_hist<<<nt, 1024, G * 4, st>>>(hs, n2, tile, g.c, g.nlo, g.W, g.K, sh, G, nt, mat);
        ZK_HIP(hipGetLastError());
        fill_u32(mat + mlen - 1, 0u, 1, st);
    } else {
        fill_u32(mat, 0u, mlen, st);
    }
    exclusive_scan(mat, msc, msc + mlen, mlen, st);
    if (n) {
        k_sort_tile_scatter<<<nt, 1024, 2 * G * 4, st>>>(hs, n2, tile, g.c, g.nlo, g.W, g.K, sh, G, nt, msc,
                                                          pairs.as<uint2>());
        ZK_HIP(hipGetLastError());
    }
    k_sort_group<<<G, 256, (4u << sh), st>>>(pairs.as<uint2>(), msc, nt, sh, (uint32_t)nb, offsets.as<uint32_t>(),
                                             entries.as<uint32_t>());
    ZK_HIP(hipGetLastError());
    const uint32_t* off = offsets.as<uint32_t>();
    uint32_t* hist = lkeys.as<uint32_t>();
    uint32_t* cur = hist + nkey;
    fill_u32(hist, 0u, nkey, st);
    const unsigned nblk_o = (unsigned)cdiv(nb, 1024);
    k_load_hist<<<nblk_o, 1024, nkey * 4, st>>>(off, (uint32_t)nb, lmax, qs, nkey, hist);
    ZK_HIP(hipGetLastError());
    exclusive_scan(hist, cur, cur + nkey, nkey, st);
    k_load_scatter<<<nblk_o, 1024, nkey * 4, st>>>(off, (uint32_t)nb, lmax, qs, nkey, cur, order.as<uint32_t>());
    ZK_HIP(hipGetLastError());
    // chunk lists of the oversized buckets, shared by every MSM over this sort
    fill_u32(counters.as<uint32_t>(), 0u, 16, st);
    k_chunk_list<<<64, 256, 0, st>>>(order.as<uint32_t>(), off, hist, counters.as<uint32_t>(), chunks.as<ChunkRec>(),
                                     large.as<LargeRec>());
    ZK_HIP(hipGetLastError());
}

void msm_reserve_partials(MsmPartials& out, const MsmGeom& g, bool g2) {
    out.host.ensure(2 * (size_t)g.W * g.nblk * (g2 ? 256 : 128));
}

// ---------------------------------------------------------------- buckets
// Bucket accumulation, G1. One thread per bucket, in the sort's accumulation
// order (longest first, so a wave's lanes carry near-equal loads and the
// launch has no long tail); buckets above lmax entries go to the chunked path.
// Arithmetic in the 29-bit-limb form (field29.hpp): points come from the
// converted copy of the bases, and each bucket is converted back to the
// standard form when it is stored.
// Thread tid of an accumulation launch over nb buckets, each split into
// `split` parts of consecutive entries (bucket b's parts go to slots
// b*split + h, summed by the level-1 reduction). Returns false when the thread
// has nothing to accumulate: past the end, or a bucket above lmax entries.
// Such a bucket is summed by the chunked path before this launch
// (k_msm_chunks, k_msm_merge into slot b*split); its other parts store
// infinity here. Otherwise [st, en) are the thread's entries and slot its
// bucket slot.
template <class T>
ZK_DEV bool acc_range(uint32_t tid, uint32_t nb, uint32_t split, const uint32_t* __restrict__ order,
                      const uint32_t* __restrict__ offsets, uint32_t lmax, void* __restrict__ buckets, uint32_t& st,
                      uint32_t& en, uint32_t& slot) {
    if (tid >= nb * split) return false;
    const uint32_t ob = tid / split, h = tid - ob * split;
    const uint32_t b = order[ob];
    const uint32_t s0 = offsets[b], e0 = offsets[b + 1];
    const uint32_t cnt = e0 - s0;
    slot = b * split + h;
    if (cnt > lmax) {
        if (h != 0) PointIO<T>::store_xyzz(buckets, slot, xyzz_inf<T>());
        return false;
    }
    st = s0 + (uint32_t)(((uint64_t)cnt * h) / split);
    en = s0 + (uint32_t)(((uint64_t)cnt * (h + 1)) / split);
    return true;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_msm_accum29(const uint32_t* __restrict__ entries,
                                                     const uint32_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ order, uint32_t nb, uint32_t split,
                                                     uint32_t lmax,
                                                     const void* __restrict__ points29, uint32_t nbase,
                                                     void* __restrict__ buckets) {
    uint32_t st, en, slot;
    if (!acc_range<Fp>(blockIdx.x * blockDim.x + threadIdx.x, nb, split, order, offsets, lmax, buckets, st, en, slot))
        return;
    Xyzz29 acc;
    acc.inf = true;
    for (uint32_t e = st; e < en; e++) {
        uint32_t ref = entries[e];
        F29 x, y;
        f29_load_point(points29, ref & 0x7fffffffu, nbase, x, y);
        if (f29_is_zero(x) && f29_is_zero(y)) continue;   // infinity
        if (ref >> 31) y = f29_neg(y);
        f29_madd(acc, x, y);
    }
    Xyzz<Fp> o = xyzz_inf<Fp>();
    if (!acc.inf) o = {f29_to_std(acc.X), f29_to_std(acc.Y), f29_to_std(acc.ZZ), f29_to_std(acc.ZZZ)};
    PointIO<Fp>::store_xyzz(buckets, slot, o);
}

// G2 accumulation in the 29-bit-limb form: the Fp2 state does not fit the
// register file next to the mixed add, so three coordinates of each thread's
// bucket accumulator live in LDS (field29.hpp AccLds) and are loaded where
// used. One-wave blocks let the LDS share set the waves per CU.
constexpr int ACC29Q_THREADS = 64;
__global__ void __launch_bounds__(ACC29Q_THREADS) __attribute__((amdgpu_waves_per_eu(2))) k_msm_accum29q(const uint32_t* __restrict__ entries,
                                                                 const uint32_t* __restrict__ offsets,
                                                                 const uint32_t* __restrict__ order, uint32_t nb,
                                                                 uint32_t split, uint32_t lmax,
                                                                 const void* __restrict__ points29,
                                                                 uint32_t nbase, void* __restrict__ buckets) {
    __shared__ uint32_t lds[AccLds<ACC29Q_THREADS>::LDS_WORDS];
    uint32_t st, en, slot;
    if (!acc_range<Fq2>(blockIdx.x * blockDim.x + threadIdx.x, nb, split, order, offsets, lmax, buckets, st, en, slot))
        return;
    AccLds<ACC29Q_THREADS> A{lds + threadIdx.x, {}};
    bool inf = true;
    for (uint32_t e = st; e < en; e++) {
        uint32_t ref = entries[e];
        F29q x, y;
        f29q_load_point(points29, ref & 0x7fffffffu, nbase, x, y);
        if (f29_is_zero(x.c0) && f29_is_zero(x.c1) && f29_is_zero(y.c0) && f29_is_zero(y.c1)) continue;
        if (ref >> 31) y = {f29_neg(y.c0), f29_neg(y.c1)};
        f29q_madd(A, inf, x, y);
    }
    using L = AccLds<ACC29Q_THREADS>;
    Xyzz<Fq2> o = xyzz_inf<Fq2>();
    if (!inf) o = {f29q_to_std(A.ld(L::X)), f29q_to_std(A.ld(L::Y)), f29q_to_std(A.ld(L::ZZ)), f29q_to_std(A.ld(L::ZZZ))};
    PointIO<Fq2>::store_xyzz(buckets, slot, o);
}

// endomorphism constant (Montgomery, 8 x u32) as a field element
ZK_DEV Fp beta_mont(const uint32_t* c) {
    Fp b;
#pragma unroll
    for (int k = 0; k < 8; k++) b.v[k] = c[k];
    return b;
}

// base i of an MSM over the 2n points of the GLV split: P_i for i < n, else
// lambda*P_(i-n) = (beta*x, y) (glv.hpp)
template <class T>
ZK_DEV Aff<T> load_base(const void* pts, uint32_t i, uint32_t n);
template <>
ZK_DEV Aff<Fp> load_base<Fp>(const void* pts, uint32_t i, uint32_t n) {
    if (i < n) return PointIO<Fp>::load_aff(pts, i);
    Aff<Fp> a = PointIO<Fp>::load_aff(pts, i - n);
    if (aff_is_inf(a)) return a;
    a.x = mul<FpCfg>(a.x, beta_mont(glv::BETA_G1_MONT));
    return a;
}
template <>
ZK_DEV Aff<Fq2> load_base<Fq2>(const void* pts, uint32_t i, uint32_t n) {
    if (i < n) return PointIO<Fq2>::load_aff(pts, i);
    Aff<Fq2> a = PointIO<Fq2>::load_aff(pts, i - n);
    if (aff_is_inf(a)) return a;
    const Fp b = beta_mont(glv::BETA_G2_MONT);
    a.x.c0 = mul<FpCfg>(a.x.c0, b);
    a.x.c1 = mul<FpCfg>(a.x.c1, b);
    return a;
}

// packed words of the 29-bit form of x (std Montgomery), at w[0..7]
ZK_DEV void pack29(const Fp& x, uint32_t* w) { f29_pack(f29_from_std(x), w); }

// base slots of msm_points_to29 (layout in field29.hpp); infinity stays all zero
__global__ void k_points_to29q(const void* __restrict__ pts, uint32_t n, void* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Aff<Fq2> a = PointIO<Fq2>::load_aff(pts, i);
    uint32_t w[48] = {};
    if (!aff_is_inf(a)) {
        const Fp b = beta_mont(glv::BETA_G2_MONT);
        pack29(a.y.c0, w);
        pack29(a.y.c1, w + 8);
        pack29(a.x.c0, w + 16);
        pack29(a.x.c1, w + 24);
        pack29(mul<FpCfg>(a.x.c0, b), w + 32);
        pack29(mul<FpCfg>(a.x.c1, b), w + 40);
    }
    uint4* o = reinterpret_cast<uint4*>(out) + PT29Q_SLOT_UINT4 * i;
#pragma unroll
    for (int k = 0; k < PT29Q_SLOT_UINT4; k++)
        o[k] = k < 12 ? make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]) : make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_points_to29(const void* __restrict__ pts, uint32_t n, void* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Aff<Fp> a = PointIO<Fp>::load_aff(pts, i);
    uint32_t w[24] = {};
    if (!aff_is_inf(a)) {
        pack29(a.y, w);
        pack29(a.x, w + 8);
        pack29(mul<FpCfg>(a.x, beta_mont(glv::BETA_G1_MONT)), w + 16);
    }
    uint4* o = reinterpret_cast<uint4*>(out) + PT29_SLOT_UINT4 * i;
#pragma unroll
    for (int k = 0; k < PT29_SLOT_UINT4; k++)
        o[k] = k < 6 ? make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]) : make_uint4(0u, 0u, 0u, 0u);
}

void msm_points_to29(const void* d_points, size_t n, void* d_out, bool g2, hipStream_t st) {
    if (!n) return;
    ZK_REQUIRE(n < 0x40000000ull, ZK_E_ARG, "MSM size must be < 2^30");
    if (g2) k_points_to29q<<<cdiv(n, 256), 256, 0, st>>>(d_points, (uint32_t)n, d_out);
    else k_points_to29<<<cdiv(n, 256), 256, 0, st>>>(d_points, (uint32_t)n, d_out);
    ZK_HIP(hipGetLastError());
}

// tree-sum of one XYZZ per thread across the block (blockDim.x must be a power of 2)
template <class T>
__device__ Xyzz<T> block_sum(Xyzz<T> v, Xyzz<T>* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    for (int h = blockDim.x >> 1; h >= 1; h >>= 1) {
        __syncthreads();
        if (t < h) sh[t] = xyzz_add(sh[t], sh[t + h]);
    }
    __syncthreads();
    Xyzz<T> r = sh[0];
    __syncthreads();
    return r;
}

// one workgroup per chunk of an oversized bucket (grid-stride over the chunk list)
template <class T>
__global__ void __launch_bounds__(256) k_msm_chunks(const uint32_t* __restrict__ entries, const void* __restrict__ points,
                                                    uint32_t nbase, const uint32_t* __restrict__ counters,
                                                    const ChunkRec* __restrict__ chunks, void* __restrict__ partials) {
    __shared__ Xyzz<T> sh[256];
    const uint32_t nch = counters[0];
    for (uint32_t ci = blockIdx.x; ci < nch; ci += gridDim.x) {
        ChunkRec cr = chunks[ci];
        Xyzz<T> acc = xyzz_inf<T>();
        for (uint32_t e = threadIdx.x; e < cr.len; e += blockDim.x) {
            uint32_t ref = entries[cr.start + e];
            Aff<T> P = load_base<T>(points, ref & 0x7fffffffu, nbase);
            if (aff_is_inf(P)) continue;
            if (ref >> 31) P.y = PointIO<T>::neg_y(P.y);
            xyzz_madd(acc, P);
        }
        Xyzz<T> s = block_sum<T>(acc, sh);
        if (threadIdx.x == 0) PointIO<T>::store_xyzz(partials, ci, s);
    }
}

// one workgroup per oversized bucket: sum its chunk partials
template <class T>
__global__ void __launch_bounds__(256) k_msm_merge(const uint32_t* __restrict__ counters, const LargeRec* __restrict__ large,
                                                   const void* __restrict__ partials, uint32_t split,
                                                   void* __restrict__ buckets) {
    __shared__ Xyzz<T> sh[256];
    const uint32_t nl = counters[1];
    for (uint32_t li = blockIdx.x; li < nl; li += gridDim.x) {
        LargeRec lr = large[li];
        Xyzz<T> acc = xyzz_inf<T>();
        for (uint32_t j = threadIdx.x; j < lr.count; j += blockDim.x)
            acc = xyzz_add(acc, PointIO<T>::load_xyzz(partials, lr.first + j));
        Xyzz<T> s = block_sum<T>(acc, sh);
        if (threadIdx.x == 0) PointIO<T>::store_xyzz(buckets, (size_t)lr.bucket * split, s);
    }
}

// ---------------------------------------------------------------- reduction
// The point additions here are inlined, each loop body holding one or two
// addition sites (level 2 selects its operands per step; the step index is
// uniform across the wave): out-of-line calls would pass the 128/256-byte
// operands through scratch memory on every step of these latency-bound chains.
//
// Level 1: thread j owns buckets [j*L1, (j+1)*L1) of one window; with bucket
// b holding digit b+1: R_j = sum B_b, T_j = sum (b - j*L1 + 1) B_b, so the
// window value is sum_j (T_j + j*L1*R_j). B_b is the sum of the bucket's
// `split` slots from the accumulation.
template <class T>
__global__ void __launch_bounds__(128) k_msm_reduce1(const void* __restrict__ buckets, uint32_t K, int W, uint32_t L1,
                                                     uint32_t split, void* __restrict__ R, void* __restrict__ Tt) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nseg = K / L1;
    if (j >= (uint32_t)W * nseg) return;
    uint32_t w = j / nseg, jj = j - w * nseg;
    size_t base = (size_t)w * K + (size_t)jj * L1;
    Xyzz<T> run = xyzz_inf<T>(), acc = xyzz_inf<T>();
#pragma unroll 1
    for (int b = (int)L1 - 1; b >= 0; b--) {
#pragma unroll 1
        for (uint32_t h = 0; h < split; h++) run = xyzz_add_i(run, PointIO<T>::load_xyzz(buckets, (base + b) * split + h));
        acc = xyzz_add_i(acc, run);
    }
    PointIO<T>::store_xyzz(R, j, run);
    PointIO<T>::store_xyzz(Tt, j, acc);
}

// Level 2: input per window is M pairs (R_j, T_j) worth sum_j (T_j + j*U*R_j),
// U = 2^logU. Block blk covers j in [blk*BLK*L2, (blk+1)*BLK*L2); thread t owns
// L2 consecutive j. Per thread: run = sum R_j, acc = sum (j - j0) R_j,
// tsum = sum T_j. With S_t = sum_{t' >= t} run_t' (suffix scan over the block),
// sum_t t*run_t = sum_{t >= 1} S_t, so the block's value relative to j_blk is
// sum_t V_t with V_t = tsum_t + U*(acc_t + L2*[t >= 1]*S_t): one tree sum.
// Output per block: R'' = S_0 and T'' = sum_t V_t, so the window value is
// sum_blk (T''_blk + blk * (U*L2*BLK) * R''_blk).
template <class T, int BLK>
__global__ void __launch_bounds__(BLK) k_msm_reduce2(const void* __restrict__ Rin, const void* __restrict__ Tin, uint32_t M,
                                                     int logU, uint32_t nblk, void* __restrict__ Rout,
                                                     void* __restrict__ Tout) {
    __shared__ Xyzz<T> sh[BLK];
    const int t = threadIdx.x;
    const uint32_t w = blockIdx.x / nblk, blk = blockIdx.x - w * nblk;
    const uint32_t j0 = (blk * BLK + t) * L2;
    Xyzz<T> run = xyzz_inf<T>(), acc = xyzz_inf<T>(), tsum = xyzz_inf<T>();
    // per item, highest j first: acc += run; run += R_j; tsum += T_j
#pragma unroll 1
    for (int step = 0; step < 3 * L2; step++) {
        const int q = L2 - 1 - step / 3, k = step % 3;
        const uint32_t j = j0 + q;
        if (j >= M) continue;
        Xyzz<T> x = k == 0 ? acc : (k == 1 ? run : tsum);
        Xyzz<T> y = k == 0 ? run : PointIO<T>::load_xyzz(k == 1 ? Rin : Tin, (size_t)w * M + j);
        Xyzz<T> r = xyzz_add_i(x, y);
        if (k == 0) acc = r;
        else if (k == 1) run = r;
        else tsum = r;
    }
    // suffix scan of run across the block: S_t = sum_{t' >= t} run_t'
    sh[t] = run;
#pragma unroll 1
    for (int d = 1; d < BLK; d <<= 1) {
        __syncthreads();
        Xyzz<T> mine = sh[t];
        Xyzz<T> other = (t + d < BLK) ? sh[t + d] : xyzz_inf<T>();
        __syncthreads();
        sh[t] = xyzz_add_i(mine, other);
    }
    __syncthreads();
    const Xyzz<T> S = sh[t], total_r = sh[0];
    __syncthreads();
    // V_t = tsum + U * (acc + L2 * [t >= 1] * S_t): doublings and additions
    // through one site each
    Xyzz<T> v = t >= 1 ? S : xyzz_inf<T>();
#pragma unroll 1
    for (int step = 0; step < LOG_L2 + 1 + logU + 1; step++) {
        if (step < LOG_L2 || (step > LOG_L2 && step <= LOG_L2 + logU)) {
            v = xyzz_dbl_i(v);
        } else {
            v = xyzz_add_i(v, step == LOG_L2 ? acc : tsum);
        }
    }
    // tree sum of V over the block
    sh[t] = v;
#pragma unroll 1
    for (int h = BLK >> 1; h >= 1; h >>= 1) {
        __syncthreads();
        if (t < h) sh[t] = xyzz_add_i(sh[t], sh[t + h]);
    }
    __syncthreads();
    if (t == 0) {
        PointIO<T>::store_xyzz(Rout, blockIdx.x, total_r);
        PointIO<T>::store_xyzz(Tout, blockIdx.x, sh[0]);
    }
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
    const size_t nb = (size_t)g.W * g.K;
    f(q.buckets, nb * split * xb);
    f(q.r1, (size_t)g.W * g.nseg * xb);
    f(q.t1, (size_t)g.W * g.nseg * xb);
    f(q.r2, (size_t)g.W * g.nblk * xb);
    f(q.t2, (size_t)g.W * g.nblk * xb);
    f(q.partials, max_chunks * xb);
}

void MsmScratch::ensure(const MsmGeom& g, size_t max_chunks, bool g2, int split) {
    scratch_buffers(*this, g, max_chunks, g2, split, [](DevBuf& b, size_t n) { b.ensure(n); });
}

void MsmScratch::count_ensure(const MsmGeom& g, size_t max_chunks, bool g2, int split, GrowCount& gc) const {
    scratch_buffers(*this, g, max_chunks, g2, split, [&](const DevBuf& b, size_t n) { gc.need(b, n); });
}

// The chunked path of the oversized buckets (lists built by the sort): their
// sums land in bucket slot b*split before the accumulation. Queued ahead of
// the accumulation so that these short launches run while the GPU has room,
// instead of waiting behind the large accumulation launches of other streams.
template <class T>
static void launch_pre_t(const MsmSort& sort, const void* d_points, MsmScratch& scr, hipStream_t st, int split) {
    const bool g2 = std::is_same<T, Fq2>::value;
    const MsmGeom g = msm_geom(sort.n, sort.g.c, g2);
    scr.ensure(g, sort.n_chunk_max, g2, split);
    scr.split = split;
    k_msm_chunks<T><<<512, 256, 0, st>>>(sort.entries.as<uint32_t>(), d_points, (uint32_t)sort.n,
                                          sort.counters.as<uint32_t>(), sort.chunks.as<ChunkRec>(), scr.partials.p);
    ZK_HIP(hipGetLastError());
    k_msm_merge<T><<<256, 256, 0, st>>>(sort.counters.as<uint32_t>(), sort.large.as<LargeRec>(), scr.partials.p,
                                         (uint32_t)split, scr.buckets.p);
    ZK_HIP(hipGetLastError());
}

template <class T>
static void launch_main_t(const MsmSort& sort, const void* d_points29, MsmScratch& scr, MsmPartials& out,
                          hipStream_t st) {
    const bool g2 = std::is_same<T, Fq2>::value;
    const MsmGeom g = msm_geom(sort.n, sort.g.c, g2);
    const int split = scr.split;
    const uint32_t nb = (uint32_t)((size_t)g.W * g.K), nt = nb * (uint32_t)split;
    if (scr.ev_acc0) ZK_HIP(hipEventRecord(scr.ev_acc0, st));
    if (g2)
        k_msm_accum29q<<<cdiv(nt, ACC29Q_THREADS), ACC29Q_THREADS, 0, st>>>(
            sort.entries.as<uint32_t>(), sort.offsets.as<uint32_t>(), sort.order.as<uint32_t>(), nb, (uint32_t)split,
            sort.lmax, d_points29, (uint32_t)sort.n, scr.buckets.p);
    else
        k_msm_accum29<<<cdiv(nt, 128), 128, 0, st>>>(sort.entries.as<uint32_t>(), sort.offsets.as<uint32_t>(),
                                                     sort.order.as<uint32_t>(), nb, (uint32_t)split, sort.lmax,
                                                     d_points29, (uint32_t)sort.n, scr.buckets.p);
    ZK_HIP(hipGetLastError());
    if (scr.ev_acc1) ZK_HIP(hipEventRecord(scr.ev_acc1, st));
    uint32_t nseg_all = (uint32_t)g.W * g.nseg;
    k_msm_reduce1<T><<<cdiv(nseg_all, 128), 128, 0, st>>>(scr.buckets.p, g.K, g.W, g.l1, (uint32_t)split, scr.r1.p,
                                                           scr.t1.p);
    ZK_HIP(hipGetLastError());
    constexpr int BLK = RedBlk<T>::V;
    k_msm_reduce2<T, BLK><<<(unsigned)(g.W * g.nblk), BLK, 0, st>>>(scr.r1.p, scr.t1.p, g.nseg, ilog2(g.l1), g.nblk,
                                                                     scr.r2.p, scr.t2.p);
    ZK_HIP(hipGetLastError());
    size_t xb = g2 ? 256 : 128;
    size_t half = (size_t)g.W * g.nblk * xb;
    out.g = g;
    out.g2 = g2;
    out.host.ensure(2 * half);
    ZK_HIP(hipMemcpyAsync(out.host.p, scr.r2.p, half, hipMemcpyDeviceToHost, st));
    ZK_HIP(hipMemcpyAsync((char*)out.host.p + half, scr.t2.p, half, hipMemcpyDeviceToHost, st));
    if (scr.ev_end) ZK_HIP(hipEventRecord(scr.ev_end, st));
}

void msm_launch_pre(const MsmSort& sort, const void* d_points, bool g2, MsmScratch& scr, hipStream_t st, int split) {
    ZK_REQUIRE(split >= 1 && split <= MSM_SPLIT_MAX, ZK_E_ARG, "bucket split must be in [1, 4]");
    if (g2) launch_pre_t<Fq2>(sort, d_points, scr, st, split);
    else launch_pre_t<Fp>(sort, d_points, scr, st, split);
}

void msm_launch_main(const MsmSort& sort, bool g2, MsmScratch& scr, MsmPartials& out, hipStream_t st,
                     const void* d_points29) {
    ZK_REQUIRE(sort.n == 0 || d_points29, ZK_E_INTERNAL, "MSM launch without the 29-bit copy of its points");
    ZK_REQUIRE(scr.split >= 1, ZK_E_INTERNAL, "MSM accumulation queued before its chunked path");
    if (g2) launch_main_t<Fq2>(sort, d_points29, scr, out, st);
    else launch_main_t<Fp>(sort, d_points29, scr, out, st);
}

void msm_launch(const MsmSort& sort, const void* d_points, bool g2, MsmScratch& scr, MsmPartials& out, hipStream_t st,
                const void* d_points29, int split) {
    msm_launch_pre(sort, d_points, g2, scr, st, split);
    msm_launch_main(sort, g2, scr, out, st, d_points29);
}

// ---------------------------------------------------------------- host finish
template <class HT>
static HT load_h(const uint64_t* p);
template <>
zkh::Fp load_h<zkh::Fp>(const uint64_t* p) { return zkh::Fp::from_raw(p); }
template <>
zkh::Fp2 load_h<zkh::Fp2>(const uint64_t* p) { return {zkh::Fp::from_raw(p), zkh::Fp::from_raw(p + 4)}; }

// value of window w from its level-2 partials: sum_blk (T_blk + blk*U*R_blk)
// with U = 2^logU, by running sums over the blocks
template <class HT>
static zkh::Xyzz<HT> window_value(const MsmPartials& p, int w) {
    const MsmGeom& g = p.g;
    const int lw = sizeof(HT) / 8;  // u64 words per coordinate
    const uint64_t* R = p.host.as<uint64_t>();
    const uint64_t* Tt = R + (size_t)g.W * g.nblk * 4 * lw;
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

// window values (on `tasks` threads when > 1), then Horner across windows
template <class HT>
static zkh::Xyzz<HT> finish_t(const MsmPartials& p, int tasks) {
    const MsmGeom& g = p.g;
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
