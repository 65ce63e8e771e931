// G1 bucket phase of the Pippenger MSM (design in msm.hpp): accumulation in
// the 9 x 29-bit-limb form, base-slot conversion, and the Fp instances of the
// shared chunk / reduction kernels (msm_kernels.hpp).
#include "msm_kernels.hpp"

namespace zk {

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
    red_store<Fp>(buckets, slot, acc);
}

// packed words of the 29-bit form of x (std Montgomery), at w[0..7]
ZK_DEV void pack29(const Fp& x, uint32_t* w) { f29_pack(f29_from_std(x), w); }

__global__ void k_points_to29(const void* __restrict__ pts, uint32_t n, void* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Aff<Fp> a = PointIO<Fp>::load_aff(pts, i);
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

template <>
void launch_accum<Fp>(const MsmSort& sort, uint32_t nb, int split, const void* d_points29, void* buckets, hipStream_t st) {
    const uint32_t nt = nb * (uint32_t)split;
    k_msm_accum29<<<cdiv(nt, 128), 128, 0, st>>>(sort.entries.as<uint32_t>(), sort.offsets.as<uint32_t>(),
                                                 sort.order.as<uint32_t>(), nb, (uint32_t)split, sort.lmax, d_points29,
                                                 (uint32_t)sort.n, buckets);
}

void msm_launch_pre_g1(const MsmSort& sort, const void* d_points, MsmScratch& scr, hipStream_t st, int split) {
    launch_pre_t<Fp>(sort, d_points, scr, st, split);
}

void msm_launch_main_g1(const MsmSort& sort, MsmScratch& scr, MsmPartials& out, hipStream_t st, const void* d_points29) {
    launch_main_t<Fp>(sort, d_points29, scr, out, st);
}

void msm_points_to29_g1(const void* d_points, size_t n, void* d_out, hipStream_t st) {
    k_points_to29<<<cdiv(n, 256), 256, 0, st>>>(d_points, (uint32_t)n, d_out);
    ZK_HIP(hipGetLastError());
}

}  // namespace zk
