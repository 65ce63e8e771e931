// G1 bucket phase of the Pippenger MSM (design in msm.hpp): accumulation in
// the 9 x 29-bit-limb form, base-slot conversion, and the Fp instances of the
// shared chunk / reduction kernels (msm_kernels.hpp).
#include "msm_kernels.hpp"

namespace zk {

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_msm_accum29(const uint32_t* __restrict__ entries,
                                                     const uint32_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ order,
                                                     const uint32_t* __restrict__ cls_off, int range, uint32_t nb,
                                                     uint32_t split, uint32_t lmax,
                                                     const void* __restrict__ points29,
                                                     void* __restrict__ buckets) {
    uint32_t st, en, slot;
    if (!acc_range<Fp>(blockIdx.x * blockDim.x + threadIdx.x, nb, split, order, offsets, cls_off, range, lmax, buckets,
                       st, en, slot))
        return;
    Xyzz29 acc;
    acc.inf = true;
    for (uint32_t e = st; e < en; e++) {
        uint32_t ref = entries[e];
        F29 x, y;
        f29_load_point(points29, ref, x, y);
        if (f29_is_zero(x) && f29_is_zero(y)) continue;   // infinity
        if (ref >> 31) y = f29_neg_lazy2(y);   // f29_madd takes it unnormalized
        f29_madd(acc, x, y);
    }
    red_store<Fp>(buckets, slot, acc);
}

// packed words of the 29-bit form of x (std Montgomery), at w[0..7]
ZK_DEV void pack29(const Fp& x, uint32_t* w) { f29_pack(f29_from_std(x), w); }

// base slot i of the 29-bit form (layout in field29.hpp) from affine a;
// infinity stays all zero
ZK_DEV void store_slot29(void* out, size_t i, const Aff<Fp>& a) {
    uint32_t w[32] = {};
    if (!aff_is_inf(a)) {
        pack29(a.y, w);
        pack29(a.x, w + 8);
        pack29(a.y, w + 16);
        pack29(mul<FpCfg>(a.x, beta_mont(glv::BETA_G1_MONT)), w + 24);
    }
    uint4* o = reinterpret_cast<uint4*>(out) + PT29_SLOT_UINT4 * i;
#pragma unroll
    for (int k = 0; k < PT29_SLOT_UINT4; k++) o[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}

__global__ void k_points_to29(const void* __restrict__ pts, uint32_t n, void* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    store_slot29(out, i, PointIO<Fp>::load_aff(pts, i));
}

// Window tables 1..W-1 of n points (table 0 is k_points_to29's): slot
// w*n + i holds 2^(offset of window w) * P_i, offsets advancing by c - 1
// bits below window nlo and c above (the digit widths). Doublings in XYZZ,
// then one inversion per table entry for the affine form. Points of the
// prime-order group other than infinity never double to infinity.
template <class T>
ZK_DEV Aff<T> xyzz_to_aff_dev(const Xyzz<T>& q) {
    // 1/ZZ = ZZZ * u and 1/ZZZ = ZZ * u with u = 1/(ZZ*ZZZ)
    const T u = inv_f(mul(q.ZZ, q.ZZZ));
    return {mul(q.X, mul(q.ZZZ, u)), mul(q.Y, mul(q.ZZ, u))};
}

__global__ void k_points_table(const void* __restrict__ pts, uint32_t n, int W, int c, int nlo, void* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Aff<Fp> a = PointIO<Fp>::load_aff(pts, i);
    const bool inf = aff_is_inf(a);
    Xyzz<Fp> q = xyzz_from_aff(a);
    for (int w = 1; w < W; w++) {
        const int width = w - 1 < nlo ? c - 1 : c;
        if (!inf)
            for (int d = 0; d < width; d++) q = xyzz_dbl(q);
        store_slot29(out, (size_t)w * n + i, inf ? a : xyzz_to_aff_dev(q));
    }
}

template <>
void launch_accum<Fp>(const MsmSort& sort, uint32_t nb, int split, const void* d_points29, void* buckets, hipStream_t st,
                       int range) {
    const uint32_t nt = nb * (uint32_t)split;
    k_msm_accum29<<<cdiv(nt, 128), 128, 0, st>>>(sort.entries.as<uint32_t>(), sort.offsets.as<uint32_t>(),
                                                 msm_range_order(sort, range),
                                                 range != MSM_RANGE_ALL ? sort.cls_off.as<uint32_t>() : nullptr, range,
                                                 nb, (uint32_t)split, sort.lmax, d_points29,
                                                 buckets);
}

void msm_launch_pre_g1(const MsmSort& sort, const void* d_points29, MsmScratch& scr, hipStream_t st, int split) {
    launch_pre_t<Fp>(sort, d_points29, scr, st, split);
}

void msm_launch_main_g1(const MsmSort& sort, MsmScratch& scr, MsmPartials& out, hipStream_t st, const void* d_points29,
                        int range) {
    launch_main_t<Fp>(sort, d_points29, scr, out, st, range);
}

void msm_points_to29_g1(const void* d_points, size_t n, void* d_out, hipStream_t st) {
    k_points_to29<<<cdiv(n, 256), 256, 0, st>>>(d_points, (uint32_t)n, d_out);
    ZK_HIP(hipGetLastError());
}

void msm_points_table_g1(const void* d_points, size_t n, const MsmGeom& g, void* d_out, hipStream_t st) {
    if (g.W < 2) return;
    k_points_table<<<cdiv(n, 128), 128, 0, st>>>(d_points, (uint32_t)n, g.W, g.c, g.nlo, d_out);
    ZK_HIP(hipGetLastError());
}

}  // namespace zk
