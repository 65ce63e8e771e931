// Device code shared by the per-curve MSM translation units (msm_g1.hip,
// msm_g2.hip): constants of the bucket phase, the oversized-bucket chunk path,
// the two reduction levels, and the launch sequence of one MSM's bucket phase,
// templated on the coordinate field. msm.hip holds the digit sort and the
// host finish.
#pragma once
#include <algorithm>
#include <type_traits>

#include "curve.hpp"
#include "field29.hpp"
#include "glv.hpp"
#include "msm.hpp"

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

// oversized-bucket threshold floor, chunk size and level sizes are shared with
// the sort (msm.hip), which builds the chunk lists

// ---------------------------------------------------------------- point forms
// Buckets, level-1 results and the level-2 working set are in the 29-bit form
// (field29.hpp), in which the accumulation leaves its buckets; the level-2
// partials go to the host in the standard form.
template <class T> struct Red;
template <>
struct Red<Fp> {
    using Pt = Xyzz29;
    static constexpr int WORDS = 32;  // packed words per point (xyzz29_pack)
    ZK_DEV static Pt inf() { return xyzz29_inf(); }
    ZK_DEV static Pt add(const Pt& a, const Pt& b) { return f29_xyzz_add(a, b); }
    ZK_DEV static Pt dbl(const Pt& a) { return f29_xyzz_dbl(a); }
    ZK_DEV static void pack(const Pt& p, uint32_t* w) { xyzz29_pack(p, w); }
    ZK_DEV static Pt unpack(const uint32_t* w) { return xyzz29_unpack(w); }
    ZK_DEV static Pt from_std(const Xyzz<Fp>& s) {
        if (xyzz_is_inf(s)) return xyzz29_inf();
        Xyzz29 r;
        r.X = f29_from_std(s.X);
        r.Y = f29_from_std(s.Y);
        r.ZZ = f29_from_std(s.ZZ);
        r.ZZZ = f29_from_std(s.ZZZ);
        r.inf = false;
        return r;
    }
    ZK_DEV static Xyzz<Fp> to_std(const Pt& p) {
        if (p.inf) return xyzz_inf<Fp>();
        return {f29_to_std(p.X), f29_to_std(p.Y), f29_to_std(p.ZZ), f29_to_std(p.ZZZ)};
    }
};
template <>
struct Red<Fq2> {
    using Pt = Xyzz29q;
    static constexpr int WORDS = 64;  // packed words per point (xyzz29q_pack)
    ZK_DEV static Pt inf() { return xyzz29q_inf(); }
    ZK_DEV static Pt add(const Pt& a, const Pt& b) { return f29q_xyzz_add(a, b); }
    ZK_DEV static Pt dbl(const Pt& a) { return f29q_xyzz_dbl(a); }
    ZK_DEV static void pack(const Pt& p, uint32_t* w) { xyzz29q_pack(p, w); }
    ZK_DEV static Pt unpack(const uint32_t* w) { return xyzz29q_unpack(w); }
    ZK_DEV static Pt from_std(const Xyzz<Fq2>& s) {
        if (xyzz_is_inf(s)) return xyzz29q_inf();
        Xyzz29q r;
        r.X = f29q_from_std(s.X);
        r.Y = f29q_from_std(s.Y);
        r.ZZ = f29q_from_std(s.ZZ);
        r.ZZZ = f29q_from_std(s.ZZZ);
        r.inf = false;
        return r;
    }
    ZK_DEV static Xyzz<Fq2> to_std(const Pt& p) {
        if (p.inf) return xyzz_inf<Fq2>();
        return {f29q_to_std(p.X), f29q_to_std(p.Y), f29q_to_std(p.ZZ), f29q_to_std(p.ZZZ)};
    }
};

// points of this form in memory: WORDS words each, 16-byte aligned
template <class T>
ZK_DEV typename Red<T>::Pt red_load(const void* base, size_t i) {
    const uint4* q = reinterpret_cast<const uint4*>(base) + (size_t)(Red<T>::WORDS / 4) * i;
    uint32_t w[Red<T>::WORDS];
#pragma unroll
    for (int k = 0; k < Red<T>::WORDS / 4; k++) {
        const uint4 v = q[k];
        w[4 * k] = v.x;
        w[4 * k + 1] = v.y;
        w[4 * k + 2] = v.z;
        w[4 * k + 3] = v.w;
    }
    return Red<T>::unpack(w);
}
template <class T>
ZK_DEV void red_store(void* base, size_t i, const typename Red<T>::Pt& p) {
    uint4* o = reinterpret_cast<uint4*>(base) + (size_t)(Red<T>::WORDS / 4) * i;
    uint32_t w[Red<T>::WORDS];
    Red<T>::pack(p, w);
#pragma unroll
    for (int k = 0; k < Red<T>::WORDS / 4; k++) o[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}
// a bucket slot written from the standard form (oversized buckets, empty parts)
template <class T>
ZK_DEV void store_bucket(void* buckets, size_t slot, const Xyzz<T>& p) {
    red_store<T>(buckets, slot, Red<T>::from_std(p));
}

// ---------------------------------------------------------------- buckets
// Bucket accumulation, G1. One thread per bucket, in the sort's accumulation
// order (longest first, so a wave's lanes carry near-equal loads and the
// launch has no long tail); buckets above lmax entries go to the chunked path.
// Arithmetic in the 29-bit-limb form (field29.hpp): points come from the
// converted copy of the bases; the buckets stay in that form for the
// reduction.
// Thread tid of an accumulation launch over nb buckets, each split into
// `split` parts of consecutive entries (bucket b's parts go to slots
// b*split + h, summed by the level-1 reduction). Returns false when the thread
// has nothing to accumulate: past the end, or a bucket above lmax entries.
// Such a bucket is summed by the chunked path before this launch
// (k_msm_chunks, k_msm_merge into slot b*split); its other parts store
// infinity here. Otherwise [st, en) are the thread's entries and slot its
// bucket slot. With cls_off (a sort with scalar classes) the parts cover the
// bucket's entries of MsmRange `range` only; the chunked path reads whole
// buckets and passes over the infinite bases there.
template <class T>
ZK_DEV bool acc_range(uint32_t tid, uint32_t nb, uint32_t split, const uint32_t* __restrict__ order,
                      const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ cls_off, int range,
                      uint32_t lmax, void* __restrict__ buckets, uint32_t& st, uint32_t& en, uint32_t& slot) {
    if (tid >= nb * split) return false;
    const uint32_t ob = tid / split, h = tid - ob * split;
    const uint32_t b = order[ob];
    uint32_t s0 = offsets[b], e0 = offsets[b + 1];
    slot = b * split + h;
    if (e0 - s0 > lmax) {
        if (h != 0) store_bucket<T>(buckets, slot, xyzz_inf<T>());
        return false;
    }
    if (cls_off && range == MSM_RANGE_A) {
        s0 = cls_off[3 * (size_t)b];
        e0 = cls_off[3 * (size_t)b + 2];
    } else if (cls_off && range == MSM_RANGE_B) {
        e0 = cls_off[3 * (size_t)b + 1];
    }
    const uint32_t cnt = e0 - s0;
    st = s0 + (uint32_t)(((uint64_t)cnt * h) / split);
    en = s0 + (uint32_t)(((uint64_t)cnt * (h + 1)) / split);
    return true;
}

// endomorphism constant (Montgomery, 8 x u32) as a field element
ZK_DEV Fp beta_mont(const uint32_t* c) {
    Fp b;
#pragma unroll
    for (int k = 0; k < 8; k++) b.v[k] = c[k];
    return b;
}

// the base an entry names (slot | half << 30, field29.hpp), read from the
// 29-bit slots and returned in the standard form; an all-zero slot is the
// point at infinity (0, 0)
template <class T>
ZK_DEV Aff<T> load_base(const void* pts29, uint32_t ref);
template <>
ZK_DEV Aff<Fp> load_base<Fp>(const void* pts29, uint32_t ref) {
    F29 x, y;
    f29_load_point(pts29, ref, x, y);
    return {f29_to_std(x), f29_to_std(y)};
}
template <>
ZK_DEV Aff<Fq2> load_base<Fq2>(const void* pts29, uint32_t ref) {
    F29q x, y;
    f29q_load_point(pts29, ref, x, y);
    return {f29q_to_std(x), f29q_to_std(y)};
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

// one-wave workgroups with a small LDS tree (8 KiB for G1, 16 KiB for G2):
// they must find CU slots while other streams' accumulations hold most of
// the LDS, and with uniform scalars they have nothing to do
constexpr int CHUNK_THREADS = 64;

// one workgroup per chunk of an oversized bucket (grid-stride over the chunk list)
template <class T>
__global__ void __launch_bounds__(CHUNK_THREADS) k_msm_chunks(const uint32_t* __restrict__ entries,
                                                             const void* __restrict__ points29,
                                                             const uint32_t* __restrict__ counters,
                                                             const ChunkRec* __restrict__ chunks,
                                                             void* __restrict__ partials) {
    __shared__ Xyzz<T> sh[CHUNK_THREADS];
    const uint32_t nch = counters[0];
    for (uint32_t ci = blockIdx.x; ci < nch; ci += gridDim.x) {
        ChunkRec cr = chunks[ci];
        Xyzz<T> acc = xyzz_inf<T>();
        for (uint32_t e = threadIdx.x; e < cr.len; e += blockDim.x) {
            uint32_t ref = entries[cr.start + e];
            Aff<T> P = load_base<T>(points29, ref);
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
__global__ void __launch_bounds__(CHUNK_THREADS) k_msm_merge(const uint32_t* __restrict__ counters, const LargeRec* __restrict__ large,
                                                   const void* __restrict__ partials, uint32_t split,
                                                   void* __restrict__ buckets) {
    __shared__ Xyzz<T> sh[CHUNK_THREADS];
    const uint32_t nl = counters[1];
    for (uint32_t li = blockIdx.x; li < nl; li += gridDim.x) {
        LargeRec lr = large[li];
        Xyzz<T> acc = xyzz_inf<T>();
        for (uint32_t j = threadIdx.x; j < lr.count; j += blockDim.x)
            acc = xyzz_add(acc, PointIO<T>::load_xyzz(partials, lr.first + j));
        Xyzz<T> s = block_sum<T>(acc, sh);
        if (threadIdx.x == 0) store_bucket<T>(buckets, (size_t)lr.bucket * split, s);
    }
}

// ---------------------------------------------------------------- reduction
// Arithmetic in the 29-bit form (Red<T>). The point additions are inlined,
// each loop body holding one or two addition sites (level 2 selects its operands per step; the step index is
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
    using P = Red<T>;
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nseg = K / L1;
    if (j >= (uint32_t)W * nseg) return;
    uint32_t w = j / nseg, jj = j - w * nseg;
    size_t base = (size_t)w * K + (size_t)jj * L1;
    typename P::Pt run = P::inf(), acc = P::inf();
#pragma unroll 1
    for (int b = (int)L1 - 1; b >= 0; b--) {
#pragma unroll 1
        for (uint32_t h = 0; h < split; h++) run = P::add(run, red_load<T>(buckets, (base + b) * split + h));
        acc = P::add(acc, run);
    }
    red_store<T>(R, j, run);
    red_store<T>(Tt, j, acc);
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
    using P = Red<T>;
    using Pt = typename P::Pt;
    __shared__ Pt sh[BLK];
    const int t = threadIdx.x;
    const uint32_t w = blockIdx.x / nblk, blk = blockIdx.x - w * nblk;
    const uint32_t j0 = (blk * BLK + t) * L2;
    Pt run = P::inf(), acc = P::inf(), tsum = P::inf();
    // per item, highest j first: acc += run; run += R_j; tsum += T_j
#pragma unroll 1
    for (int step = 0; step < 3 * L2; step++) {
        const int q = L2 - 1 - step / 3, k = step % 3;
        const uint32_t j = j0 + q;
        if (j >= M) continue;
        Pt x = k == 0 ? acc : (k == 1 ? run : tsum);
        Pt y = k == 0 ? run : red_load<T>(k == 1 ? Rin : Tin, (size_t)w * M + j);
        Pt r = P::add(x, y);
        if (k == 0) acc = r;
        else if (k == 1) run = r;
        else tsum = r;
    }
    // suffix scan of run across the block: S_t = sum_{t' >= t} run_t'
    sh[t] = run;
#pragma unroll 1
    for (int d = 1; d < BLK; d <<= 1) {
        __syncthreads();
        Pt mine = sh[t];
        Pt other = (t + d < BLK) ? sh[t + d] : P::inf();
        __syncthreads();
        sh[t] = P::add(mine, other);
    }
    __syncthreads();
    const Pt S = sh[t], total_r = sh[0];
    __syncthreads();
    // V_t = tsum + U * (acc + L2 * [t >= 1] * S_t): doublings and additions
    // through one site each
    Pt v = t >= 1 ? S : P::inf();
#pragma unroll 1
    for (int step = 0; step < LOG_L2 + 1 + logU + 1; step++) {
        if (step < LOG_L2 || (step > LOG_L2 && step <= LOG_L2 + logU)) {
            v = P::dbl(v);
        } else {
            v = P::add(v, step == LOG_L2 ? acc : tsum);
        }
    }
    // tree sum of V over the block
    sh[t] = v;
#pragma unroll 1
    for (int h = BLK >> 1; h >= 1; h >>= 1) {
        __syncthreads();
        if (t < h) sh[t] = P::add(sh[t], sh[t + h]);
    }
    __syncthreads();
    if (t == 0) {
        PointIO<T>::store_xyzz(Rout, blockIdx.x, P::to_std(total_r));
        PointIO<T>::store_xyzz(Tout, blockIdx.x, P::to_std(sh[0]));
    }
}

// The chunked path of the oversized buckets (lists built by the sort): their
// sums land in bucket slot b*split before the accumulation. Queued ahead of
// the accumulation so that these short launches run while the GPU has room,
// instead of waiting behind the large accumulation launches of other streams;
// a few small workgroups each (grid-stride over the lists): with uniform
// scalars the lists are empty, and a launch that needs many CU slots would
// wait behind other streams' accumulations while its own stream's
// accumulation waits behind it.
template <class T>
static void launch_pre_t(const MsmSort& sort, const void* d_points29, MsmScratch& scr, hipStream_t st, int split) {
    const bool g2 = std::is_same<T, Fq2>::value;
    const MsmGeom g = msm_geom(sort.n, sort.g.c, g2, sort.g.merged);
    scr.ensure(g, sort.n_chunk_max, g2, split);
    scr.split = split;
    k_msm_chunks<T><<<256, CHUNK_THREADS, 0, st>>>(sort.entries.as<uint32_t>(), d_points29,
                                                   sort.counters.as<uint32_t>(), sort.chunks.as<ChunkRec>(),
                                                   scr.partials.p);
    ZK_HIP(hipGetLastError());
    k_msm_merge<T><<<64, CHUNK_THREADS, 0, st>>>(sort.counters.as<uint32_t>(), sort.large.as<LargeRec>(), scr.partials.p,
                                         (uint32_t)split, scr.buckets.p);
    ZK_HIP(hipGetLastError());
}

// accumulation order of the MSMs that read MsmRange `range` of the sort
inline const uint32_t* msm_range_order(const MsmSort& sort, int range) {
    const DevBuf& o = range == MSM_RANGE_A ? sort.order_a : range == MSM_RANGE_B ? sort.order_b : sort.order;
    return o.as<uint32_t>();
}

// the bucket-accumulation kernel of curve T (msm_g1.hip / msm_g2.hip)
template <class T>
void launch_accum(const MsmSort& sort, uint32_t nb, int split, const void* d_points29, void* buckets, hipStream_t st,
                  int range);

template <class T>
static void launch_main_t(const MsmSort& sort, const void* d_points29, MsmScratch& scr, MsmPartials& out,
                          hipStream_t st, int range) {
    const bool g2 = std::is_same<T, Fq2>::value;
    const MsmGeom g = msm_geom(sort.n, sort.g.c, g2, sort.g.merged);
    const int split = scr.split;
    const uint32_t nb = (uint32_t)((size_t)g.Wr * g.K);
    if (scr.ev_acc0) ZK_HIP(hipEventRecord(scr.ev_acc0, st));
    launch_accum<T>(sort, nb, split, d_points29, scr.buckets.p, st, range);
    ZK_HIP(hipGetLastError());
    if (scr.ev_acc1) ZK_HIP(hipEventRecord(scr.ev_acc1, st));
    uint32_t nseg_all = (uint32_t)g.Wr * g.nseg;
    k_msm_reduce1<T><<<cdiv(nseg_all, 128), 128, 0, st>>>(scr.buckets.p, g.K, g.Wr, g.l1, (uint32_t)split, scr.r1.p,
                                                           scr.t1.p);
    ZK_HIP(hipGetLastError());
    constexpr int BLK = RedBlk<T>::V;
    k_msm_reduce2<T, BLK><<<(unsigned)(g.Wr * g.nblk), BLK, 0, st>>>(scr.r1.p, scr.t1.p, g.nseg, ilog2(g.l1), g.nblk,
                                                                      scr.r2.p, scr.t2.p);
    ZK_HIP(hipGetLastError());
    size_t xb = g2 ? 256 : 128;
    size_t half = (size_t)g.Wr * g.nblk * xb;
    out.g = g;
    out.g2 = g2;
    void* dst = out.ext;
    if (dst) {
        ZK_REQUIRE(out.ext_bytes >= 2 * half, ZK_E_INTERNAL, "MSM partials exceed the caller's host buffer");
    } else {
        out.host.ensure(2 * half);
        dst = out.host.p;
    }
    ZK_HIP(hipMemcpyAsync(dst, scr.r2.p, half, hipMemcpyDeviceToHost, st));
    ZK_HIP(hipMemcpyAsync((char*)dst + half, scr.t2.p, half, hipMemcpyDeviceToHost, st));
    if (scr.ev_end) ZK_HIP(hipEventRecord(scr.ev_end, st));
}


// per-curve entry points, defined in msm_g1.hip and msm_g2.hip
void msm_launch_pre_g1(const MsmSort& sort, const void* d_points29, MsmScratch& scr, hipStream_t st, int split);
void msm_launch_main_g1(const MsmSort& sort, MsmScratch& scr, MsmPartials& out, hipStream_t st, const void* d_points29,
                        int range);
void msm_points_to29_g1(const void* d_points, size_t n, void* d_out, hipStream_t st);
void msm_points_table_g1(const void* d_points, size_t n, const MsmGeom& g, void* d_out, hipStream_t st);
void msm_launch_pre_g2(const MsmSort& sort, const void* d_points29, MsmScratch& scr, hipStream_t st, int split);
void msm_launch_main_g2(const MsmSort& sort, MsmScratch& scr, MsmPartials& out, hipStream_t st, const void* d_points29,
                        int range);
void msm_points_to29_g2(const void* d_points, size_t n, void* d_out, hipStream_t st);
void msm_points_table_g2(const void* d_points, size_t n, const MsmGeom& g, void* d_out, hipStream_t st);

}  // namespace zk
