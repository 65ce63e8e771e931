// Fr NTT kernels (design in ntt.hpp): register-radix passes of the four-step
// recursion. Each thread owns one column of R = 2^lr points (lr <= 4), does
// the whole radix-R DFT in registers, applies the inter-pass twiddle and
// writes back; the last pass writes the digit-reversed (natural) order.
#include <vector>

#include "field.hpp"
#include "field29.hpp"
#include "host_field.hpp"
#include "ntt.hpp"

namespace zk {

constexpr int NTT_THREADS = 256;
constexpr int NTT_MAX_BATCH = 4;

struct NttPtrs {
    void* in[NTT_MAX_BATCH];
    void* out[NTT_MAX_BATCH];
};

// w_R^i for i < R/2 (standard form, prescaled by 2^5: field29.hpp "Fr"),
// passed by value so they live in SGPRs
struct RootsR {
    Fr w[8];
};

struct NttDigits {
    int np;
    int lr[NTT_MAX_PASSES];
};

constexpr int bitrev_c(int x, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

// ---- elements: 9 x 29-bit limbs in registers, 8 canonical words in memory ----
// A product by a prescaled table value keeps the standard form (field29.hpp).
ZK_DEV F29 ld29(const void* base, size_t i) {
    const uint4* p = reinterpret_cast<const uint4*>(base) + 2 * i;
    const uint4 a = p[0], b = p[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return f29_unpack(w);
}
// x normalized and below 2r
ZK_DEV void st29(void* base, size_t i, const F29& x) {
    uint32_t w[8];
    f29_pack(fr29_canon(x), w);
    uint4* p = reinterpret_cast<uint4*>(base) + 2 * i;
    p[0] = make_uint4(w[0], w[1], w[2], w[3]);
    p[1] = make_uint4(w[4], w[5], w[6], w[7]);
}
// limbs below 2^31, value below 2^261 -> normalized, below r + 2^240
ZK_DEV F29 settle29(const F29& x) { return fr29_reduce(f29_norm(x)); }

// one radix-2 DIT stage (span 2^(S-1)) of a 2^LR-point DFT held in registers;
// every loop has a constant trip count so the whole DFT unrolls to registers.
// Bounds (worst cases followed stage by stage in an exact model): the DFT's
// inputs are below 1.2r (canonical loads, products by a prescaled table
// value, below 1.01r, or settled values, below r + 2^240); stage S then
// sees b below 2^S r with a top limb below that of the borrowed 2^S r, so
// a - b + 2^S r has no negative limb; outputs stay below 32r (radix 16),
// limbs below 2^31. Between stages the limbs are normalized; the last
// stage's outputs go straight to a product (limbs below 2^31 are fine
// there) or to settle29.
template <int LR, int S>
struct DftStage {
    ZK_DEV static void run(F29 (&x)[1 << LR], const F29 (&w)[8]) {
        constexpr int R = 1 << LR, half = 1 << (S - 1);
#pragma unroll
        for (int q = 0; q < R / 2; q++) {
            const int p = q & (half - 1);
            const int i0 = (q >> (S - 1)) * 2 * half + p, i1 = i0 + half;
            const F29 a = x[i0];
            F29 b = x[i1];
            if (p) b = fr29_mul(b, w[p * (R >> S)]);
            x[i0] = f29_add(a, b);
            x[i1] = fr29_sub<S>(a, b);
        }
        if (S < LR) {
#pragma unroll
            for (int i = 0; i < R; i++) x[i] = f29_norm(x[i]);
        }
        DftStage<LR, S + 1>::run(x, w);
    }
};
template <int LR>
struct DftStage<LR, LR + 1> {
    ZK_DEV static void run(F29 (&)[1 << LR], const F29 (&)[8]) {}
};

// radix-2^LR DFT in registers, X_k = sum_b x_b w_R^(b k); the caller loads
// input b into x[bitrev(b)] and gets natural-order output
template <int LR>
ZK_DEV void dft_reg(F29 (&x)[1 << LR], const RootsR& rt) {
    F29 w[8];
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = f29_unpack(rt.w[i].v);
    DftStage<LR, 1>::run(x, w);
}

// In-place pass j < np-1 over blocks of M = R*S elements: thread (beta, c)
// transforms x[beta*M + b*S + c], b < R, then multiplies output k by w_M^(c k)
// (the table holds w_M^0 for c = 0, so every output k >= 1 is a product).
template <int LR>
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_pass(NttPtrs ptrs, int logM, size_t n, RootsR rt,
                                                          const void* __restrict__ tw, const void* __restrict__ pre) {
    constexpr int R = 1 << LR;
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (n >> LR)) return;
    void* data = ptrs.in[blockIdx.y];
    const int logS = logM - LR;
    const size_t c = tid & (((size_t)1 << logS) - 1);
    const size_t base = (tid >> logS) << logM;
    F29 x[R];
#pragma unroll
    for (int b = 0; b < R; b++) {
        const size_t i = base + ((size_t)b << logS) + c;
        F29 v = ld29(data, i);
        if (pre) v = fr29_mul(v, ld29(pre, i));
        x[bitrev_c(b, LR)] = v;
    }
    dft_reg<LR>(x, rt);
#pragma unroll
    for (int k = 0; k < R; k++) {
        const F29 y = k ? fr29_mul(x[k], ld29(tw, ((size_t)k << logS) + c)) : settle29(x[k]);
        st29(data, base + ((size_t)k << logS) + c, y);
    }
}

// Two consecutive passes j, j+1 (radix R1 = 2^LR1, then R2 = 2^LR2) in one
// kernel, the intermediate result staged in LDS. With M = R1*S (pass j),
// S = R2*S' (pass j+1), the R1*R2 points x[beta*M + b*S + b2*S' + c2]
// (b < R1, b2 < R2) of one unit u = (beta, c2) touch no other unit in either
// pass. A workgroup takes U consecutive units:
//   step 1: thread (b2, u) runs pass j's DFT over b for column c = b2*S' + c2,
//           multiplies output k by w_M^(c k) and writes LDS[k][b2][u];
//   step 2: thread (k, u) runs pass j+1's DFT over b2 for sub-block (beta, k),
//           multiplies output k2 by w_S^(c2 k2) and stores x[beta*M + k*S +
//           k2*S' + c2].
// Consecutive lanes take consecutive units, so global accesses stay coalesced
// (runs of min(S', U) elements). The LDS holds the 9 limbs of each staged
// element (normalized, below 1.2r), limb-major.
template <int LR1, int LR2, int U>
__global__ void __launch_bounds__((1 << LR2) * U) k_ntt_pass2(NttPtrs ptrs, int logM, size_t n, RootsR rt1, RootsR rt2,
                                                              const void* __restrict__ tw1, const void* __restrict__ tw2,
                                                              const void* __restrict__ pre) {
    constexpr int R1 = 1 << LR1, R2 = 1 << LR2, NS = R1 * R2 * U;
    static_assert(R1 == R2, "one task per thread in both steps");
    __shared__ uint32_t lds[9 * NS];
    void* data = ptrs.in[blockIdx.y];
    const int logS = logM - LR1, logS2 = logS - LR2;
    const size_t S2 = (size_t)1 << logS2;
    const int ul = threadIdx.x % U, q = threadIdx.x / U;   // unit lane, digit (b2 in step 1, k in step 2)
    const size_t u = (size_t)blockIdx.x * U + ul;
    const size_t beta = u >> logS2, c2 = u & (S2 - 1);
    const size_t base = beta << logM;
    const bool live = u < (n >> (LR1 + LR2));
    if (live) {
        // step 1: pass j over b for column c = q*S' + c2
        const size_t c = ((size_t)q << logS2) + c2;
        F29 x[R1];
#pragma unroll
        for (int b = 0; b < R1; b++) {
            const size_t i = base + ((size_t)b << logS) + c;
            F29 v = ld29(data, i);
            if (pre) v = fr29_mul(v, ld29(pre, i));
            x[bitrev_c(b, LR1)] = v;
        }
        dft_reg<LR1>(x, rt1);
#pragma unroll
        for (int k = 0; k < R1; k++) {
            const F29 y = k ? fr29_mul(x[k], ld29(tw1, ((size_t)k << logS) + c)) : settle29(x[k]);
#pragma unroll
            for (int l = 0; l < 9; l++) lds[l * NS + (k * R2 + q) * U + ul] = y.l[l];
        }
    }
    __syncthreads();
    if (!live) return;
    // step 2: pass j+1 over b2 for sub-block (beta, k = q), column c2
    F29 x[R2];
#pragma unroll
    for (int b2 = 0; b2 < R2; b2++) {
        F29 v;
#pragma unroll
        for (int l = 0; l < 9; l++) v.l[l] = lds[l * NS + (q * R2 + b2) * U + ul];
        x[bitrev_c(b2, LR2)] = v;
    }
    dft_reg<LR2>(x, rt2);
    const size_t sub = base + ((size_t)q << logS);
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) {
        const F29 y = k2 ? fr29_mul(x[k2], ld29(tw2, ((size_t)k2 << logS2) + c2)) : settle29(x[k2]);
        st29(data, sub + ((size_t)k2 << logS2) + c2, y);
    }
}

// Last pass: thread = output column rv (digit-reversed row index); the row of
// R contiguous points at unrev(rv)*R gives out[rv + O*k], k < R.
template <int LR>
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_final(NttPtrs ptrs, NttDigits dg, size_t n, RootsR rt,
                                                           const void* __restrict__ pre, const void* __restrict__ post) {
    constexpr int R = 1 << LR;
    const size_t O = n >> LR;
    size_t rv = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (rv >= O) return;
    const void* in = ptrs.in[blockIdx.y];
    void* out = ptrs.out[blockIdx.y];
    size_t r = rv, row = 0;
    for (int d = 0; d < dg.np - 1; d++) {
        size_t digit = r & (((size_t)1 << dg.lr[d]) - 1);
        r >>= dg.lr[d];
        row = (row << dg.lr[d]) | digit;
    }
    F29 x[R];
#pragma unroll
    for (int b = 0; b < R; b++) {
        const size_t i = (row << LR) + b;
        F29 v = ld29(in, i);
        if (pre && dg.np == 1) v = fr29_mul(v, ld29(pre, i));
        x[bitrev_c(b, LR)] = v;
    }
    dft_reg<LR>(x, rt);
#pragma unroll
    for (int k = 0; k < R; k++) {
        const size_t o = rv + (size_t)k * O;
        const F29 y = post ? fr29_mul(x[k], ld29(post, o)) : settle29(x[k]);
        st29(out, o, y);
    }
}

// table[i] = lo[e & 4095] * hi[e >> 12] for e = exponent(i)
//   mode 0: e = i * step (mod 2n)
//   mode 1: pass twiddle, i = k*S + c -> e = c*k*step
__global__ void k_pow_table(void* __restrict__ out, size_t count, const void* __restrict__ lo, const void* __restrict__ hi,
                            int mode, uint64_t step, int logS, uint64_t mask, const void* __restrict__ scale) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint64_t e;
    if (mode == 0) e = (i * step) & mask;
    else {
        uint64_t k = i >> logS, c = i & (((uint64_t)1 << logS) - 1);
        e = (c * k * step) & mask;
    }
    Fr v = load<FrCfg>(lo, e & 4095);
    if (e >> 12) v = mul<FrCfg>(v, load<FrCfg>(hi, e >> 12));
    if (scale) v = mul<FrCfg>(v, load<FrCfg>(scale, 0));
    store<FrCfg>(out, i, v);
}

// host: powers tables of g (order 2^log_ord): lo[i] = g^i (i < 4096), hi[j] = g^(4096 j)
static void host_pow_tables(const zkh::Fr& g, int log_ord, std::vector<zkh::Fr>& lo, std::vector<zkh::Fr>& hi) {
    lo.assign(4096, zkh::Fr::one());
    for (int i = 1; i < 4096; i++) lo[i] = lo[i - 1] * g;
    zkh::Fr g4096 = lo[4095] * g;
    size_t nh = log_ord > 12 ? ((size_t)1 << (log_ord - 12)) : 1;
    hi.assign(nh, zkh::Fr::one());
    for (size_t j = 1; j < nh; j++) hi[j] = hi[j - 1] * g4096;
}

static zkh::Fr root_of_unity(int log_n) {
    // w_{2^log_n} = 5^((r-1) >> log_n)
    uint64_t e[4];
    memcpy(e, zkh::FrP::M, 32);
    e[0] -= 1;
    for (int s = 0; s < log_n; s++) {
        for (int k = 0; k < 3; k++) e[k] = (e[k] >> 1) | (e[k + 1] << 63);
        e[3] >>= 1;
    }
    return zkh::Fr::from_u64(5).pow(e);
}

// pass radices of a size-2^ln transform: radix-8 passes; a radix-16 pass
// absorbs one leftover bit, radix-4/2 the rest. Returns the pass count.
static int ntt_layout(int ln, int lr[NTT_MAX_PASSES]) {
    int np = 0, rem = ln;
    while (rem > 0) {
        int take = rem >= 3 ? 3 : rem;
        lr[np++] = take;
        rem -= take;
    }
    if (np >= 2 && lr[np - 1] == 1) { lr[np - 2] += 1; np--; }  // ...3,1 -> ...4
    for (int j = np; j < NTT_MAX_PASSES; j++) lr[j] = 0;
    return np;
}

// entries of the concatenated pass twiddle tables (passes 0..np-2), and each pass's offset
static size_t ntt_tw_entries(int ln, int np, const int lr[NTT_MAX_PASSES], size_t* off) {
    size_t tot = 0;
    int consumed = 0;
    for (int j = 0; j < np; j++) {
        if (off) off[j] = tot;
        if (j < np - 1) tot += (size_t)1 << (ln - consumed);
        consumed += lr[j];
    }
    return tot;
}

size_t NttPlan::bytes_for(int ln) {
    int lr_[NTT_MAX_PASSES];
    const int np_ = ntt_layout(ln, lr_);
    const size_t tot = ntt_tw_entries(ln, np_, lr_, nullptr);
    return 2 * std::max<size_t>(tot, 1) * 32 + 4 * ((size_t)1 << ln) * 32;
}

void NttPlan::build(int ln, hipStream_t st) {
    ZK_REQUIRE(ln >= 1 && ln <= 27, ZK_E_ARG, "NTT size must be 2^1..2^27");
    log_n = ln;
    n = (size_t)1 << ln;
    np = ntt_layout(ln, lr);

    // base power tables for w_2n and w_2n^-1 (cover every w_M, w_R and the coset factors)
    zkh::Fr w2n = root_of_unity(ln + 1), w2n_inv = w2n.inv();
    std::vector<zkh::Fr> lo_f, hi_f, lo_i, hi_i;
    host_pow_tables(w2n, ln + 1, lo_f, hi_f);
    host_pow_tables(w2n_inv, ln + 1, lo_i, hi_i);
    DevBuf d_lo_f(4096 * 32), d_hi_f(hi_f.size() * 32), d_lo_i(4096 * 32), d_hi_i(hi_i.size() * 32), d_scale(32),
        d_s32(32);
    ZK_HIP(hipMemcpyAsync(d_lo_f.p, lo_f.data(), 4096 * 32, hipMemcpyHostToDevice, st));
    ZK_HIP(hipMemcpyAsync(d_hi_f.p, hi_f.data(), hi_f.size() * 32, hipMemcpyHostToDevice, st));
    ZK_HIP(hipMemcpyAsync(d_lo_i.p, lo_i.data(), 4096 * 32, hipMemcpyHostToDevice, st));
    ZK_HIP(hipMemcpyAsync(d_hi_i.p, hi_i.data(), hi_i.size() * 32, hipMemcpyHostToDevice, st));
    // every table value the kernels multiply by is prescaled by 2^5 (the
    // 29-bit product's R'/2^256, field29.hpp "Fr")
    const zkh::Fr s32 = zkh::Fr::from_u64(32);
    const zkh::Fr ninv = zkh::Fr::from_u64(n).inv() * s32;
    ZK_HIP(hipMemcpyAsync(d_scale.p, &ninv, 32, hipMemcpyHostToDevice, st));
    ZK_HIP(hipMemcpyAsync(d_s32.p, &s32, 32, hipMemcpyHostToDevice, st));
    const uint64_t mask2n = (2 * (uint64_t)n) - 1;

    // radix roots w_R^i = w_2n^(i * 2n / R), forward and inverse, on the host
    for (int j = 0; j < np; j++) {
        int R = 1 << lr[j];
        for (int i = 0; i < 8; i++) {
            uint64_t e = i < R / 2 ? ((uint64_t)i * ((2 * (uint64_t)n) >> lr[j])) & mask2n : 0;
            zkh::Fr f = lo_f[e & 4095] * hi_f[e >> 12] * s32, v = lo_i[e & 4095] * hi_i[e >> 12] * s32;
            memcpy(roots_fwd[j] + 4 * i, f.v, 32);  // 4 words per element
            memcpy(roots_inv[j] + 4 * i, v.v, 32);
        }
    }
    // pass twiddle tables w_M^(c k), layout [k][c], for passes 0..np-2
    const size_t tot = ntt_tw_entries(ln, np, lr, tw_off);
    tw_fwd.ensure(std::max<size_t>(tot, 1) * 32);
    tw_inv.ensure(std::max<size_t>(tot, 1) * 32);
    int consumed = 0;
    for (int j = 0; j < np - 1; j++) {
        int logM = ln - consumed;
        int logS = logM - lr[j];
        uint64_t step_M = (uint64_t)1 << (ln + 1 - logM);  // w_M = w_2n^(2n/M)
        size_t cnt = (size_t)1 << logM;
        k_pow_table<<<cdiv(cnt, 256), 256, 0, st>>>((char*)tw_fwd.p + tw_off[j] * 32, cnt, d_lo_f.p, d_hi_f.p, 1, step_M,
                                                   logS, mask2n, d_s32.p);
        k_pow_table<<<cdiv(cnt, 256), 256, 0, st>>>((char*)tw_inv.p + tw_off[j] * 32, cnt, d_lo_i.p, d_hi_i.p, 1, step_M,
                                                   logS, mask2n, d_s32.p);
        consumed += lr[j];
    }
    // element-wise tables
    coset_fwd.ensure(n * 32);
    inv_coset.ensure(n * 32);
    inv_plain.ensure(n * 32);
    inv_uncoset.ensure(n * 32);
    k_pow_table<<<cdiv(n, 256), 256, 0, st>>>(coset_fwd.p, n, d_lo_f.p, d_hi_f.p, 0, 1, 0, mask2n, d_s32.p);
    k_pow_table<<<cdiv(n, 256), 256, 0, st>>>(inv_coset.p, n, d_lo_f.p, d_hi_f.p, 0, 1, 0, mask2n, d_scale.p);
    k_pow_table<<<cdiv(n, 256), 256, 0, st>>>(inv_plain.p, n, d_lo_f.p, d_hi_f.p, 0, 0, 0, mask2n, d_scale.p);
    k_pow_table<<<cdiv(n, 256), 256, 0, st>>>(inv_uncoset.p, n, d_lo_i.p, d_hi_i.p, 0, 1, 0, mask2n, d_scale.p);
    ZK_HIP(hipGetLastError());
    ZK_HIP(hipStreamSynchronize(st));  // the temporaries above are freed on return
}

template <int LR>
static void launch_pass(const NttPtrs& ptrs, int batch, int logM, size_t n, const RootsR& rt, const void* tw,
                        const void* pre, hipStream_t st) {
    unsigned grid = cdiv(n >> LR, NTT_THREADS);
    k_ntt_pass<LR><<<dim3(grid, batch), NTT_THREADS, 0, st>>>(ptrs, logM, n, rt, tw, pre);
}

constexpr int NTT2_UNITS = 16;   // 36 KiB of LDS per workgroup (9 limbs x 64 points x 16 units)

static void launch_pass2(const NttPtrs& ptrs, int batch, int logM, size_t n, const RootsR& rt1, const RootsR& rt2,
                         const void* tw1, const void* tw2, const void* pre, hipStream_t st) {
    unsigned grid = cdiv(n >> 6, NTT2_UNITS);
    k_ntt_pass2<3, 3, NTT2_UNITS><<<dim3(grid, batch), 8 * NTT2_UNITS, 0, st>>>(ptrs, logM, n, rt1, rt2, tw1, tw2, pre);
}

template <int LR>
static void launch_final(const NttPtrs& ptrs, int batch, const NttDigits& dg, size_t n, const RootsR& rt, const void* pre,
                         const void* post, hipStream_t st) {
    unsigned grid = cdiv(n >> LR, NTT_THREADS);
    k_ntt_final<LR><<<dim3(grid, batch), NTT_THREADS, 0, st>>>(ptrs, dg, n, rt, pre, post);
}

void ntt_run(const NttPlan& p, void* const* in, void* const* out, int batch, bool inverse, const void* pre, const void* post,
             hipStream_t st) {
    ZK_REQUIRE(batch >= 1 && batch <= NTT_MAX_BATCH, ZK_E_ARG, "NTT batch must be 1..4");
    NttPtrs ptrs;
    for (int b = 0; b < NTT_MAX_BATCH; b++) {
        ptrs.in[b] = b < batch ? in[b] : nullptr;
        ptrs.out[b] = b < batch ? out[b] : nullptr;
        if (b < batch && p.np > 1) ZK_REQUIRE(in[b] != out[b], ZK_E_ARG, "multi-pass NTT needs distinct in/out");
    }
    NttDigits dg;
    dg.np = p.np;
    for (int j = 0; j < NTT_MAX_PASSES; j++) dg.lr[j] = p.lr[j];
    const char* tw = (const char*)(inverse ? p.tw_inv.p : p.tw_fwd.p);
    int consumed = 0;
    for (int j = 0; j < p.np; j++) {
        RootsR rt;
        memcpy(&rt, inverse ? p.roots_inv[j] : p.roots_fwd[j], sizeof rt);
        const int lrj = p.lr[j];
        if (j < p.np - 1) {
            const int logM = p.log_n - consumed;
            const void* pj = j == 0 ? pre : nullptr;
            const void* twj = tw + p.tw_off[j] * 32;
            // two radix-8 passes through LDS when both are regular passes
            if (lrj == 3 && j + 1 < p.np - 1 && p.lr[j + 1] == 3) {
                RootsR rt2;
                memcpy(&rt2, inverse ? p.roots_inv[j + 1] : p.roots_fwd[j + 1], sizeof rt2);
                launch_pass2(ptrs, batch, logM, p.n, rt, rt2, twj, tw + p.tw_off[j + 1] * 32, pj, st);
                ZK_HIP(hipGetLastError());
                consumed += 6;
                j++;
                continue;
            }
            switch (lrj) {
                case 1: launch_pass<1>(ptrs, batch, logM, p.n, rt, twj, pj, st); break;
                case 2: launch_pass<2>(ptrs, batch, logM, p.n, rt, twj, pj, st); break;
                case 3: launch_pass<3>(ptrs, batch, logM, p.n, rt, twj, pj, st); break;
                default: launch_pass<4>(ptrs, batch, logM, p.n, rt, twj, pj, st); break;
            }
        } else {
            switch (lrj) {
                case 1: launch_final<1>(ptrs, batch, dg, p.n, rt, pre, post, st); break;
                case 2: launch_final<2>(ptrs, batch, dg, p.n, rt, pre, post, st); break;
                case 3: launch_final<3>(ptrs, batch, dg, p.n, rt, pre, post, st); break;
                default: launch_final<4>(ptrs, batch, dg, p.n, rt, pre, post, st); break;
            }
        }
        ZK_HIP(hipGetLastError());
        consumed += lrj;
    }
}

}  // namespace zk
