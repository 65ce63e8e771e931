#include <type_traits>

#include "curve.hpp"
#include "qap.hpp"

namespace zk {

__global__ void k_spmv_abc(const uint32_t* __restrict__ rowA, const uint32_t* __restrict__ colA, const void* __restrict__ valA,
                           uint32_t n_rowsA, const uint32_t* __restrict__ rowB, const uint32_t* __restrict__ colB,
                           const void* __restrict__ valB, uint32_t n_rowsB, const void* __restrict__ w, size_t n,
                           void* __restrict__ aw, void* __restrict__ bw, void* __restrict__ cw) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fr a = Fr::zero(), b = Fr::zero();
    if (i < n_rowsA) {
        for (uint32_t k = rowA[i], e = rowA[i + 1]; k < e; k++)
            a = add<FrCfg>(a, mul<FrCfg>(load<FrCfg>(valA, k), load<FrCfg>(w, colA[k])));
    }
    if (i < n_rowsB) {
        for (uint32_t k = rowB[i], e = rowB[i + 1]; k < e; k++)
            b = add<FrCfg>(b, mul<FrCfg>(load<FrCfg>(valB, k), load<FrCfg>(w, colB[k])));
    }
    store<FrCfg>(aw, i, a);
    store<FrCfg>(bw, i, b);
    store<FrCfg>(cw, i, mul<FrCfg>(a, b));
}

void spmv_abc(const uint32_t* rowA, const uint32_t* colA, const void* valA, uint32_t n_rowsA, const uint32_t* rowB,
              const uint32_t* colB, const void* valB, uint32_t n_rowsB, const void* w, size_t n, void* aw, void* bw,
              void* cw, hipStream_t st) {
    k_spmv_abc<<<cdiv(n, 256), 256, 0, st>>>(rowA, colA, valA, n_rowsA, rowB, colB, valB, n_rowsB, w, n, aw, bw, cw);
    ZK_HIP(hipGetLastError());
}

// One dependency level of the synthetic witness: aux_i = (A_i.w)(B_i.w) for
// the rows rows[0..count) (every wire they read is already final). The CSR
// values are raw c*R^2, so the row sums land in Montgomery form; their product
// times R^-1 is the Montgomery form of aux_i, stored canonical.
__global__ void k_witness_rows(const uint32_t* __restrict__ rowA, const uint32_t* __restrict__ colA,
                               const void* __restrict__ valA, const uint32_t* __restrict__ rowB,
                               const uint32_t* __restrict__ colB, const void* __restrict__ valB,
                               const uint32_t* __restrict__ rows, uint32_t count, uint32_t first_aux,
                               void* __restrict__ w) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint32_t i = rows[t];
    Fr a = Fr::zero(), b = Fr::zero();
    for (uint32_t k = rowA[i], e = rowA[i + 1]; k < e; k++)
        a = add<FrCfg>(a, mul<FrCfg>(load<FrCfg>(valA, k), load<FrCfg>(w, colA[k])));
    for (uint32_t k = rowB[i], e = rowB[i + 1]; k < e; k++)
        b = add<FrCfg>(b, mul<FrCfg>(load<FrCfg>(valB, k), load<FrCfg>(w, colB[k])));
    store<FrCfg>(w, (size_t)first_aux + i, from_mont<FrCfg>(mul<FrCfg>(a, b)));
}

void witness_rows(const uint32_t* rowA, const uint32_t* colA, const void* valA, const uint32_t* rowB,
                  const uint32_t* colB, const void* valB, const uint32_t* rows, uint32_t count, uint32_t first_aux,
                  void* w, hipStream_t st) {
    if (!count) return;
    k_witness_rows<<<cdiv(count, 256), 256, 0, st>>>(rowA, colA, valA, rowB, colB, valB, rows, count, first_aux, w);
    ZK_HIP(hipGetLastError());
}

__global__ void k_pointwise_mul(const void* __restrict__ a, const void* __restrict__ b, void* __restrict__ c, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    store<FrCfg>(c, i, mul<FrCfg>(load<FrCfg>(a, i), load<FrCfg>(b, i)));
}

void pointwise_mul(const void* a, const void* b, void* c, size_t n, hipStream_t st) {
    k_pointwise_mul<<<cdiv(n, 256), 256, 0, st>>>(a, b, c, n);
    ZK_HIP(hipGetLastError());
}

__global__ void k_h_final(const void* __restrict__ a, const void* __restrict__ b, const void* __restrict__ c,
                          void* __restrict__ h, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fr v = sub<FrCfg>(mul<FrCfg>(load<FrCfg>(a, i), load<FrCfg>(b, i)), load<FrCfg>(c, i));
    store<FrCfg>(h, i, from_mont<FrCfg>(v));
}

void h_final(const void* a, const void* b, const void* c, void* h, size_t n, hipStream_t st) {
    k_h_final<<<cdiv(n, 256), 256, 0, st>>>(a, b, c, h, n);
    ZK_HIP(hipGetLastError());
}

struct Scalars4 { uint64_t v[16]; };
__global__ void k_write4(uint64_t* dst, Scalars4 s) {
    int t = threadIdx.x;
    if (t < 16) dst[t] = s.v[t];
}
void write_scalars4(void* dst, const uint64_t v[16], hipStream_t st) {
    Scalars4 s;
    for (int i = 0; i < 16; i++) s.v[i] = v[i];
    k_write4<<<1, 64, 0, st>>>((uint64_t*)dst, s);
    ZK_HIP(hipGetLastError());
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}
void fill_u32(void* p, uint32_t v, size_t count, hipStream_t st) {
    if (!count) return;
    k_fill_u32<<<cdiv(count, 256), 256, 0, st>>>((uint32_t*)p, v, count);
    ZK_HIP(hipGetLastError());
}

// ---- fixed-base comb: e*G = sum_j T[j][byte_j(e)] ----
template <class T>
__global__ void __launch_bounds__(128) k_fixed_base(const void* __restrict__ table, const void* __restrict__ exps, size_t n,
                                                    void* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* e = reinterpret_cast<const uint32_t*>(exps) + 8 * i;
    Xyzz<T> acc = xyzz_inf<T>();
    for (int j = 0; j < 32; j++) {
        uint32_t d = (e[j >> 2] >> (8 * (j & 3))) & 0xffu;
        if (!d) continue;
        Aff<T> P = PointIO<T>::load_aff(table, (size_t)j * 256 + d);
        xyzz_madd(acc, P);
    }
    // affine: x = X / ZZ, y = Y / ZZZ using one inversion of ZZ*ZZZ
    Aff<T> r;
    if (xyzz_is_inf(acc)) {
        r.x = f_zero<T>();
        r.y = f_zero<T>();
    } else {
        T t = inv_f(mul(acc.ZZ, acc.ZZZ));
        r.x = mul(acc.X, mul(t, acc.ZZZ));
        r.y = mul(acc.Y, mul(t, acc.ZZ));
    }
    if constexpr (std::is_same<T, Fp>::value) {
        store<FpCfg>(out, 2 * i, r.x);
        store<FpCfg>(out, 2 * i + 1, r.y);
    } else {
        store<FpCfg>(out, 4 * i, r.x.c0);
        store<FpCfg>(out, 4 * i + 1, r.x.c1);
        store<FpCfg>(out, 4 * i + 2, r.y.c0);
        store<FpCfg>(out, 4 * i + 3, r.y.c1);
    }
}

void fixed_base_mul(const void* table, const void* exps, size_t n, void* out, bool g2, hipStream_t st) {
    if (!n) return;
    if (g2) k_fixed_base<Fq2><<<cdiv(n, 128), 128, 0, st>>>(table, exps, n, out);
    else k_fixed_base<Fp><<<cdiv(n, 128), 128, 0, st>>>(table, exps, n, out);
    ZK_HIP(hipGetLastError());
}

}  // namespace zk
