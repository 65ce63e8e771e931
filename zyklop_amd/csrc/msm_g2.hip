// G2 bucket phase of the Pippenger MSM (design in msm.hpp): accumulation in
// the 29-bit-limb form over Fp2 with the accumulator partly in LDS (the
// buckets stay in that form for the reduction), base-slot
// conversion, and the Fp2 instances of the shared chunk / reduction kernels
// (msm_kernels.hpp).
#include "msm_kernels.hpp"

namespace zk {

// G2 accumulation in the 29-bit-limb form: the Fp2 state does not fit the
// register file next to the mixed add, so three coordinates of each thread's
// bucket accumulator live in LDS (field29.hpp AccLds) and are loaded where
// used. One-wave blocks let the LDS share set the waves per CU.
constexpr int ACC29Q_THREADS = 64;
__global__ void __launch_bounds__(ACC29Q_THREADS) __attribute__((amdgpu_waves_per_eu(2))) k_msm_accum29q(const uint32_t* __restrict__ entries,
                                                                 const uint32_t* __restrict__ offsets,
                                                                 const uint32_t* __restrict__ order,
                                                                 const uint32_t* __restrict__ cls_off, int range,
                                                                 uint32_t nb, uint32_t split, uint32_t lmax,
                                                                 const void* __restrict__ points29,
                                                                 void* __restrict__ buckets) {
    __shared__ uint32_t lds[AccLds<ACC29Q_THREADS>::LDS_WORDS];
    uint32_t st, en, slot;
    if (!acc_range<Fq2>(blockIdx.x * blockDim.x + threadIdx.x, nb, split, order, offsets, cls_off, range, lmax,
                        buckets, st, en, slot))
        return;
    AccLds<ACC29Q_THREADS> A{lds + threadIdx.x, {}};
    bool inf = true;
    for (uint32_t e = st; e < en; e++) {
        uint32_t ref = entries[e];
        F29q x, y;
        f29q_load_point(points29, ref, x, y);
        if (f29_is_zero(x.c0) && f29_is_zero(x.c1) && f29_is_zero(y.c0) && f29_is_zero(y.c1)) continue;
        if (ref >> 31) y = {f29_neg_lazy2(y.c0), f29_neg_lazy2(y.c1)};   // f29q_madd takes it unnormalized
        f29q_madd(A, inf, x, y);
    }
    using L = AccLds<ACC29Q_THREADS>;
    red_store<Fq2>(buckets, slot, Xyzz29q{A.ld(L::X), A.ld(L::Y), A.ld(L::ZZ), A.ld(L::ZZZ), inf});
}

// packed words of the 29-bit form of x (std Montgomery), at w[0..7]
ZK_DEV void pack29(const Fp& x, uint32_t* w) { f29_pack(f29_from_std(x), w); }

// base slot i of msm_points_to29 (layout in field29.hpp); infinity stays all zero
ZK_DEV void store_slot29q(void* out, size_t i, const Aff<Fq2>& a) {
    uint32_t w[64] = {};
    if (!aff_is_inf(a)) {
        const Fp b = beta_mont(glv::BETA_G2_MONT);
        pack29(a.y.c0, w);
        pack29(a.y.c1, w + 8);
        pack29(a.x.c0, w + 16);
        pack29(a.x.c1, w + 24);
        pack29(a.y.c0, w + 32);
        pack29(a.y.c1, w + 40);
        pack29(mul<FpCfg>(a.x.c0, b), w + 48);
        pack29(mul<FpCfg>(a.x.c1, b), w + 56);
    }
    uint4* o = reinterpret_cast<uint4*>(out) + PT29Q_SLOT_UINT4 * i;
#pragma unroll
    for (int k = 0; k < PT29Q_SLOT_UINT4; k++) o[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}

__global__ void k_points_to29q(const void* __restrict__ pts, uint32_t n, void* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    store_slot29q(out, i, PointIO<Fq2>::load_aff(pts, i));
}

// window tables 1..W-1 of n G2 points, as k_points_table does for G1
__global__ void k_points_table_q(const void* __restrict__ pts, uint32_t n, int W, int c, int nlo,
                                 void* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Aff<Fq2> a = PointIO<Fq2>::load_aff(pts, i);
    const bool inf = aff_is_inf(a);
    Xyzz<Fq2> q = xyzz_from_aff(a);
    for (int w = 1; w < W; w++) {
        const int width = w - 1 < nlo ? c - 1 : c;
        if (!inf)
            for (int d = 0; d < width; d++) q = xyzz_dbl(q);
        Aff<Fq2> r = a;
        if (!inf) {
            const Fq2 u = inv_f(mul(q.ZZ, q.ZZZ));   // 1/ZZ = ZZZ*u, 1/ZZZ = ZZ*u
            r = {mul(q.X, mul(q.ZZZ, u)), mul(q.Y, mul(q.ZZ, u))};
        }
        store_slot29q(out, (size_t)w * n + i, r);
    }
}

template <>
void launch_accum<Fq2>(const MsmSort& sort, uint32_t nb, int split, const void* d_points29, void* buckets, hipStream_t st,
                       int range) {
    const uint32_t nt = nb * (uint32_t)split;
    k_msm_accum29q<<<cdiv(nt, ACC29Q_THREADS), ACC29Q_THREADS, 0, st>>>(
        sort.entries.as<uint32_t>(), sort.offsets.as<uint32_t>(), msm_range_order(sort, range),
        range != MSM_RANGE_ALL ? sort.cls_off.as<uint32_t>() : nullptr, range, nb, (uint32_t)split, sort.lmax,
        d_points29, buckets);
}

void msm_launch_pre_g2(const MsmSort& sort, const void* d_points29, MsmScratch& scr, hipStream_t st, int split) {
    launch_pre_t<Fq2>(sort, d_points29, scr, st, split);
}

void msm_launch_main_g2(const MsmSort& sort, MsmScratch& scr, MsmPartials& out, hipStream_t st, const void* d_points29,
                        int range) {
    launch_main_t<Fq2>(sort, d_points29, scr, out, st, range);
}

void msm_points_to29_g2(const void* d_points, size_t n, void* d_out, hipStream_t st) {
    k_points_to29q<<<cdiv(n, 256), 256, 0, st>>>(d_points, (uint32_t)n, d_out);
    ZK_HIP(hipGetLastError());
}

void msm_points_table_g2(const void* d_points, size_t n, const MsmGeom& g, void* d_out, hipStream_t st) {
    if (g.W < 2) return;
    k_points_table_q<<<cdiv(n, 128), 128, 0, st>>>(d_points, (uint32_t)n, g.W, g.c, g.nlo, d_out);
    ZK_HIP(hipGetLastError());
}

}  // namespace zk
