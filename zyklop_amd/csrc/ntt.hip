// Fr NTT kernels (design in ntt.hpp).
#include <vector>

#include "field.hpp"
#include "host_field.hpp"
#include "ntt.hpp"

namespace zk {

constexpr int NTT_TILE = 2048;  // elements resident in LDS per workgroup
constexpr int NTT_THREADS = 256;
constexpr int NTT_MAX_BATCH = 4;

struct NttPtrs {
    void* in[NTT_MAX_BATCH];
    void* out[NTT_MAX_BATCH];
};

struct NttShape {
    int np, j;
    int lr[4];
};

// LDS tile: limb pairs, structure-of-arrays, so lanes touching consecutive
// elements hit consecutive banks.
struct LdsTile {
    uint2 w[4][NTT_TILE];
    ZK_DEV Fr ld(int i) const {
        Fr r;
#pragma unroll
        for (int k = 0; k < 4; k++) { uint2 v = w[k][i]; r.v[2 * k] = v.x; r.v[2 * k + 1] = v.y; }
        return r;
    }
    ZK_DEV void st(int i, const Fr& a) {
#pragma unroll
        for (int k = 0; k < 4; k++) w[k][i] = make_uint2(a.v[2 * k], a.v[2 * k + 1]);
    }
};

ZK_DEV uint32_t bitrev(uint32_t x, int bits) { return bits ? (__brev(x) >> (32 - bits)) : 0u; }

// radix-2 DIT over T columns of R = 2^lr points held in LDS in bit-reversed order;
// tw[i] = w_R^i for i < R/2. Leaves natural order.
ZK_DEV void lds_dft(LdsTile& L, int lr, int T, const void* __restrict__ tw) {
    const int R = 1 << lr, halfR = R >> 1;
    const int nbf = T * halfR;
    for (int s = 1; s <= lr; s++) {
        const int half = 1 << (s - 1);
        const int tstride = R >> s;  // w_{2^s}^pos = w_R^(pos * R / 2^s)
        __syncthreads();
        for (int q = threadIdx.x; q < nbf; q += NTT_THREADS) {
            int ct = q >> (lr - 1);
            int u = q & (halfR - 1);
            int grp = u >> (s - 1), pos = u & (half - 1);
            int i0 = ct * R + grp * 2 * half + pos, i1 = i0 + half;
            Fr a = L.ld(i0), b = L.ld(i1);
            if (pos) b = mul<FrCfg>(b, load<FrCfg>(tw, (size_t)pos * tstride));
            L.st(i0, add<FrCfg>(a, b));
            L.st(i1, sub<FrCfg>(a, b));
        }
    }
    __syncthreads();
}

// In-place pass j < np-1: block beta of size M = R*S, columns [c0, c0+T).
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_pass(NttPtrs ptrs, NttShape sh, int log_n, const void* __restrict__ tw,
                                                          const void* __restrict__ small, const void* __restrict__ pre) {
    __shared__ LdsTile L;
    void* data = ptrs.in[blockIdx.y];
    int consumed = 0;
    for (int q = 0; q < sh.j; q++) consumed += sh.lr[q];
    const int lr = sh.lr[sh.j];
    const int logM = log_n - consumed, logS = logM - lr;
    const size_t S = (size_t)1 << logS;
    const int R = 1 << lr;
    const int T = (int)min((size_t)(NTT_TILE >> lr), S);
    const size_t tiles_per_block = S / T;
    const size_t beta = blockIdx.x / tiles_per_block, gamma = blockIdx.x - beta * tiles_per_block;
    const size_t base = beta << logM;
    const size_t c0 = gamma * T;
    const int nel = T * R;
    for (int q = threadIdx.x; q < nel; q += NTT_THREADS) {
        int b = q / T, ct = q - b * T;
        size_t g = base + ((size_t)b << logS) + c0 + ct;
        Fr x = load<FrCfg>(data, g);
        if (pre) x = mul<FrCfg>(x, load<FrCfg>(pre, g));
        L.st(ct * R + (int)bitrev((uint32_t)b, lr), x);
    }
    lds_dft(L, lr, T, small);
    for (int q = threadIdx.x; q < nel; q += NTT_THREADS) {
        int k = q / T, ct = q - k * T;
        size_t c = c0 + ct;
        Fr y = L.ld(ct * R + k);
        if (k && c) y = mul<FrCfg>(y, load<FrCfg>(tw, ((size_t)k << logS) + c));
        store<FrCfg>(data, base + ((size_t)k << logS) + c, y);
    }
}

// Last pass: rows of R contiguous points; the workgroup takes T rows with
// consecutive digit-reversed index and writes out[rev(row) + O*k].
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_final(NttPtrs ptrs, NttShape sh, int log_n, const void* __restrict__ small,
                                                           const void* __restrict__ pre, const void* __restrict__ post) {
    __shared__ LdsTile L;
    const void* in = ptrs.in[blockIdx.y];
    void* out = ptrs.out[blockIdx.y];
    const int lr = sh.lr[sh.np - 1];
    const int R = 1 << lr;
    const int logO = log_n - lr;
    const size_t O = (size_t)1 << logO;
    const int T = (int)min((size_t)(NTT_TILE >> lr), O);
    const size_t rev0 = (size_t)blockIdx.x * T;
    const int nel = T * R;
    for (int q = threadIdx.x; q < nel; q += NTT_THREADS) {
        int ri = q >> lr, b = q & (R - 1);
        // un-reverse: rev = d0 + R0*d1 + ... ; row = ((d0*R1 + d1)*R2 + ...)
        size_t rv = rev0 + ri, row = 0;
        for (int d = 0; d < sh.np - 1; d++) {
            size_t digit = rv & (((size_t)1 << sh.lr[d]) - 1);
            rv >>= sh.lr[d];
            row = (row << sh.lr[d]) | digit;
        }
        size_t g = (row << lr) + b;
        Fr x = load<FrCfg>(in, g);
        if (pre && sh.np == 1) x = mul<FrCfg>(x, load<FrCfg>(pre, g));
        L.st(ri * R + (int)bitrev((uint32_t)b, lr), x);
    }
    lds_dft(L, lr, T, small);
    for (int q = threadIdx.x; q < nel; q += NTT_THREADS) {
        int k = q / T, ri = q - k * T;
        size_t o = rev0 + ri + (size_t)k * O;
        Fr y = L.ld(ri * R + k);
        if (post) y = mul<FrCfg>(y, load<FrCfg>(post, o));
        store<FrCfg>(out, o, y);
    }
}

// table[i] = lo[e & 4095] * hi[e >> 12] for e = exponent(i)
//   mode 0: e = i * step (mod n_mod)
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

void NttPlan::build(int ln, hipStream_t st) {
    ZK_REQUIRE(ln >= 1 && ln <= 27, ZK_E_ARG, "NTT size must be 2^1..2^27");
    log_n = ln;
    n = (size_t)1 << ln;
    np = (ln + 7) / 8;
    for (int j = 0; j < 4; j++) lr[j] = 0;
    for (int j = 0; j < np; j++) lr[j] = ln / np + (j < ln % np ? 1 : 0);

    // base power tables for w_n, w_n^-1, w_2n, w_2n^-1 (the 2n-th root covers all)
    zkh::Fr w2n = root_of_unity(ln + 1), w2n_inv = w2n.inv();
    std::vector<zkh::Fr> lo_f, hi_f, lo_i, hi_i;
    host_pow_tables(w2n, ln + 1, lo_f, hi_f);
    host_pow_tables(w2n_inv, ln + 1, lo_i, hi_i);
    DevBuf d_lo_f(4096 * 32), d_hi_f(hi_f.size() * 32), d_lo_i(4096 * 32), d_hi_i(hi_i.size() * 32), d_scale(32);
    ZK_HIP(hipMemcpyAsync(d_lo_f.p, lo_f.data(), 4096 * 32, hipMemcpyHostToDevice, st));
    ZK_HIP(hipMemcpyAsync(d_hi_f.p, hi_f.data(), hi_f.size() * 32, hipMemcpyHostToDevice, st));
    ZK_HIP(hipMemcpyAsync(d_lo_i.p, lo_i.data(), 4096 * 32, hipMemcpyHostToDevice, st));
    ZK_HIP(hipMemcpyAsync(d_hi_i.p, hi_i.data(), hi_i.size() * 32, hipMemcpyHostToDevice, st));
    zkh::Fr ninv = zkh::Fr::from_u64(n).inv();
    ZK_HIP(hipMemcpyAsync(d_scale.p, &ninv, 32, hipMemcpyHostToDevice, st));
    const uint64_t mask2n = (2 * (uint64_t)n) - 1;

    // pass tables: w_M^(c k) with w_M = w_2n^(2n/M)
    size_t tot = 0, stot = 0;
    int consumed = 0;
    for (int j = 0; j < np; j++) {
        int logM = ln - consumed;
        tw_off[j] = tot;
        small_off[j] = stot;
        tot += (j < np - 1) ? ((size_t)1 << logM) : 0;
        stot += (size_t)1 << (lr[j] - 1);
        consumed += lr[j];
    }
    tw_fwd.ensure(std::max<size_t>(tot, 1) * 32);
    tw_inv.ensure(std::max<size_t>(tot, 1) * 32);
    small_fwd.ensure(stot * 32);
    small_inv.ensure(stot * 32);
    consumed = 0;
    for (int j = 0; j < np; j++) {
        int logM = ln - consumed;
        int logS = logM - lr[j];
        uint64_t step_M = (uint64_t)1 << (ln + 1 - logM);  // w_M = w_2n^(2n/M)
        if (j < np - 1) {
            size_t cnt = (size_t)1 << logM;
            k_pow_table<<<cdiv(cnt, 256), 256, 0, st>>>((char*)tw_fwd.p + tw_off[j] * 32, cnt, d_lo_f.p, d_hi_f.p, 1, step_M,
                                                       logS, mask2n, nullptr);
            k_pow_table<<<cdiv(cnt, 256), 256, 0, st>>>((char*)tw_inv.p + tw_off[j] * 32, cnt, d_lo_i.p, d_hi_i.p, 1, step_M,
                                                       logS, mask2n, nullptr);
        }
        size_t sc = (size_t)1 << (lr[j] - 1);
        uint64_t step_R = (uint64_t)1 << (ln + 1 - lr[j]);  // w_R = w_2n^(2n/R)
        k_pow_table<<<cdiv(sc, 256), 256, 0, st>>>((char*)small_fwd.p + small_off[j] * 32, sc, d_lo_f.p, d_hi_f.p, 0, step_R, 0,
                                                  mask2n, nullptr);
        k_pow_table<<<cdiv(sc, 256), 256, 0, st>>>((char*)small_inv.p + small_off[j] * 32, sc, d_lo_i.p, d_hi_i.p, 0, step_R, 0,
                                                  mask2n, nullptr);
        consumed += lr[j];
    }
    // element-wise tables
    coset_fwd.ensure(n * 32);
    inv_coset.ensure(n * 32);
    inv_plain.ensure(n * 32);
    inv_uncoset.ensure(n * 32);
    k_pow_table<<<cdiv(n, 256), 256, 0, st>>>(coset_fwd.p, n, d_lo_f.p, d_hi_f.p, 0, 1, 0, mask2n, nullptr);
    k_pow_table<<<cdiv(n, 256), 256, 0, st>>>(inv_coset.p, n, d_lo_f.p, d_hi_f.p, 0, 1, 0, mask2n, d_scale.p);
    k_pow_table<<<cdiv(n, 256), 256, 0, st>>>(inv_plain.p, n, d_lo_f.p, d_hi_f.p, 0, 0, 0, mask2n, d_scale.p);
    k_pow_table<<<cdiv(n, 256), 256, 0, st>>>(inv_uncoset.p, n, d_lo_i.p, d_hi_i.p, 0, 1, 0, mask2n, d_scale.p);
    ZK_HIP(hipGetLastError());
    ZK_HIP(hipStreamSynchronize(st));  // the temporaries above are freed on return
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
    NttShape sh;
    sh.np = p.np;
    for (int j = 0; j < 4; j++) sh.lr[j] = p.lr[j];
    const char* tw = (const char*)(inverse ? p.tw_inv.p : p.tw_fwd.p);
    const char* sm = (const char*)(inverse ? p.small_inv.p : p.small_fwd.p);
    int consumed = 0;
    for (int j = 0; j < p.np - 1; j++) {
        sh.j = j;
        int logM = p.log_n - consumed;
        size_t S = (size_t)1 << (logM - p.lr[j]);
        size_t T = std::min<size_t>((size_t)NTT_TILE >> p.lr[j], S);
        unsigned grid = (unsigned)(p.n / (((size_t)1 << p.lr[j]) * T));
        consumed += p.lr[j];
        k_ntt_pass<<<dim3(grid, batch), NTT_THREADS, 0, st>>>(ptrs, sh, p.log_n, tw + p.tw_off[j] * 32,
                                                              sm + p.small_off[j] * 32, j == 0 ? pre : nullptr);
        ZK_HIP(hipGetLastError());
    }
    sh.j = p.np - 1;
    int lr_last = p.lr[p.np - 1];
    size_t O = p.n >> lr_last;
    size_t T = std::min<size_t>((size_t)NTT_TILE >> lr_last, O);
    unsigned grid = (unsigned)(O / T);
    k_ntt_final<<<dim3(grid, batch), NTT_THREADS, 0, st>>>(ptrs, sh, p.log_n, sm + p.small_off[p.np - 1] * 32, pre, post);
    ZK_HIP(hipGetLastError());
}

}  // namespace zk
