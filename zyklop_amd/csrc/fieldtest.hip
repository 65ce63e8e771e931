// Device field arithmetic, element-wise over arrays, for tests of the field
// code itself: the standard 8 x 32-bit Montgomery form of field.hpp (the
// product-scanning asm product of mont_asm.hpp included) for Fp and Fr, and
// the 9 x 29-bit form of field29.hpp (Fp) on raw limbs, so that inputs can sit
// at the lazy-reduction bounds that form documents.
#include "field29.hpp"
#include "fieldtest.hpp"

namespace zk {

template <class C>
__global__ void k_field_std(int op, const void* __restrict__ a, const void* __restrict__ b, void* __restrict__ out,
                            size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Fe<C> x = load<C>(a, i);
    const Fe<C> y = b ? load<C>(b, i) : Fe<C>::zero();
    Fe<C> r;
    switch (op) {
        case FOP_ADD: r = add<C>(x, y); break;
        case FOP_SUB: r = sub<C>(x, y); break;
        case FOP_NEG: r = neg<C>(x); break;
        case FOP_MUL: r = mul<C>(x, y); break;
        case FOP_SQR: r = sqr<C>(x); break;
        case FOP_INV: r = inv<C>(x); break;
        case FOP_TO_MONT: r = to_mont<C>(x); break;
        case FOP_FROM_MONT: r = from_mont<C>(x); break;
        default: r = mul_cios<C>(x, y); break;   // FOP_MUL_CIOS
    }
    store<C>(out, i, r);
}

// raw 29-bit limbs (9 words per element); b may be null for unary ops
__global__ void k_field29(int op, const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                          uint32_t* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (op == F29Q_MUL || op == F29Q_MSM) {   // element pairs (c0, c1)
        if (i >= n / 2) return;
        F29q x, y;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            x.c0.l[k] = a[18 * i + k];
            x.c1.l[k] = a[18 * i + 9 + k];
            y.c0.l[k] = b[18 * i + k];
            y.c1.l[k] = b[18 * i + 9 + k];
        }
        // T*A - B*C with T = x and A = B = C = y, as in the G2 Y3
        const F29q r = op == F29Q_MUL ? f29q_mul(x, y) : f29q_mul_sub_mul(x, y, y, y);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            out[18 * i + k] = r.c0.l[k];
            out[18 * i + 9 + k] = r.c1.l[k];
        }
        return;
    }
    if (i >= n) return;
    F29 x, y;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        x.l[k] = a[9 * i + k];
        y.l[k] = b ? b[9 * i + k] : 0u;
    }
    F29 r;
    switch (op) {
        case F29_MUL: r = f29_mul(x, y); break;
        case F29_SQR: r = f29_sqr(x); break;
        case F29_MULSUM2: {            // x*y + y*x
            const F29 p[2] = {x, y}, q[2] = {y, x};
            r = f29_mulsum<2>(p, q);
            break;
        }
        case F29_MULSUM4: {            // x*x + x*y + y*x + y*y
            const F29 p[4] = {x, x, y, y}, q[4] = {x, y, x, y};
            r = f29_mulsum<4>(p, q);
            break;
        }
        case F29_NEG: r = f29_neg(x); break;
        case F29_SUB4: r = f29_norm(f29_sub4(x, y)); break;
        case F29_REDUCE: r = f29_reduce(x); break;
        case F29_NEG4: r = f29_neg_lazy4(x); break;
        case F29_Y3: {                 // the G1 Y3 form: R*(Q - X3) + Y1*(4p - PPP), R = Q = x, X3 = Y1 = PPP = y
            const F29 p[2] = {x, y}, q[2] = {f29_sub4(x, y), f29_neg_lazy4(y)};
            r = f29_mulsum<2>(p, q);
            break;
        }
        default: {                      // F29_ROUNDTRIP: to the standard form and back
            Fp s = f29_to_std(x);
            r = f29_from_std(s);
            break;
        }
    }
#pragma unroll
    for (int k = 0; k < 9; k++) out[9 * i + k] = r.l[k];
}

// VALU rate probe: every thread runs 8 independent dependent chains of
// `iters` x 16 steps, each step one v_mad_u64_u32 acc = lo32(acc) * a + acc
// (kind 0) or the pair x += y, y += x of v_add_u32 (kind 1; two ops per
// step). Each step reads the previous one's result, so no step folds away;
// the results are stored so nothing is dead code.
__global__ void __launch_bounds__(256) k_valu_probe(int kind, uint32_t iters, uint32_t seed, uint64_t* __restrict__ sink) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t acc[8];
    uint32_t a[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        acc[j] = t + j;
        a[j] = seed * (t + 17 * j + 1) | 1u;
    }
    if (kind == 0) {
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int r = 0; r < 16; r++)
#pragma unroll
                for (int j = 0; j < 8; j++) acc[j] = (uint64_t)(uint32_t)acc[j] * a[j] + acc[j];
        }
    } else if (kind == 1) {
        uint32_t x[8], y[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            x[j] = (uint32_t)acc[j];
            y[j] = a[j];
        }
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int r = 0; r < 16; r++)
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    x[j] += y[j];
                    y[j] += x[j];
                }
        }
#pragma unroll
        for (int j = 0; j < 8; j++) acc[j] = x[j] ^ ((uint64_t)y[j] << 32);
    } else if (kind == 2) {
        // v_fma_f64, 8 independent chains
        double x[8], y[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            x[j] = (double)(acc[j] & 0xffff) * 0x1p-20;
            y[j] = 1.0 - (double)(a[j] & 0xff) * 0x1p-30;
        }
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int r = 0; r < 16; r++)
#pragma unroll
                for (int j = 0; j < 8; j++) x[j] = __builtin_fma(x[j], y[j], 0x1p-40);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) acc[j] = (uint64_t)__double_as_longlong(x[j]);
    } else {
        // one dependent chain of v_mad_u64_u32 (8 steps per inner round)
        uint64_t c = acc[0];
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int r = 0; r < 16; r++)
#pragma unroll
                for (int j = 0; j < 8; j++) c = (uint64_t)(uint32_t)c * a[j] + c;
        }
        acc[0] = c;
#pragma unroll
        for (int j = 1; j < 8; j++) acc[j] = 0;
    }
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) s ^= acc[j];
    sink[t] = s;
}

// dependent chains of point additions in the 29-bit form, 16 per iteration:
// op 4 mixed XYZZ additions (the bucket accumulation's), op 5 general XYZZ
// additions (the bucket reduction's). The operands are limb patterns below p,
// not curve points: only the instruction stream counts.
__global__ void __launch_bounds__(64) k_point_probe(int op, uint32_t iters, uint32_t seed, uint64_t* __restrict__ sink) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    Xyzz29 p, q;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint32_t h = seed * (t * 9 + i + 1), m = i < 8 ? f29c::MASK : 0xffffu;
        p.X.l[i] = (h * 2654435761u) & m;
        p.Y.l[i] = (h * 40503u + 7) & m;
        p.ZZ.l[i] = (h * 69069u + 1) & m;
        p.ZZZ.l[i] = (h * 1103515245u + 12345u) & m;
        q.X.l[i] = p.Y.l[i] ^ 0x5555u;
        q.Y.l[i] = p.ZZ.l[i] ^ 0x3333u;
        q.ZZ.l[i] = p.X.l[i] ^ 0x0f0fu;
        q.ZZZ.l[i] = p.ZZZ.l[i] ^ 0x00ffu;
    }
    p.inf = q.inf = false;
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll 1
        for (int r = 0; r < 16; r++) {
            if (op == 4) f29_madd(p, q.X, q.Y);
            else p = f29_xyzz_add(p, q);
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) s ^= p.X.l[i] ^ p.Y.l[i] ^ p.ZZ.l[i] ^ p.ZZZ.l[i];
    sink[t] = s;
}

// kind = op + 16 * waves: op 0 v_mad_u64_u32 (8 independent chains), 1
// v_add_u32 (2 per step), 2 v_fma_f64 (8 chains), 3 one dependent chain of
// v_mad_u64_u32, 4 / 5 point additions (k_point_probe); waves = waves per
// SIMD on a 1024-SIMD chip (0: 8)
double valu_probe(int kind, hipStream_t st) {
    const int op = kind & 15, waves = kind >> 4 ? kind >> 4 : 8;
    const bool pts = op >= 4;
    const uint32_t threads = 64, blocks = 1024 * (uint32_t)waves, iters = pts ? 16 : 2048;
    DevBuf sink((size_t)blocks * threads * 8);
    hipEvent_t e0, e1;
    ZK_HIP(hipEventCreate(&e0));
    ZK_HIP(hipEventCreate(&e1));
    auto launch = [&](uint32_t it) {
        if (pts) k_point_probe<<<blocks, threads, 0, st>>>(op, it, 12345u, sink.as<uint64_t>());
        else k_valu_probe<<<blocks, threads, 0, st>>>(op, it, 12345u, sink.as<uint64_t>());
    };
    launch(1);  // warm-up
    ZK_HIP(hipEventRecord(e0, st));
    launch(iters);
    ZK_HIP(hipEventRecord(e1, st));
    ZK_HIP(hipEventSynchronize(e1));
    float ms = 0;
    ZK_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double ops_per_step = op == 1 ? 2.0 : (pts ? 1.0 / 8 : 1.0);
    return (double)blocks * threads * iters * 16 * 8 * ops_per_step / (ms * 1e-3);   // lane-ops per second
}

void field_op_std(int field, int op, const void* a, const void* b, void* out, size_t n, hipStream_t st) {
    if (!n) return;
    if (field == 0) k_field_std<FpCfg><<<cdiv(n, 128), 128, 0, st>>>(op, a, b, out, n);
    else k_field_std<FrCfg><<<cdiv(n, 128), 128, 0, st>>>(op, a, b, out, n);
    ZK_HIP(hipGetLastError());
}

void field29_op(int op, const uint32_t* a, const uint32_t* b, uint32_t* out, size_t n, hipStream_t st) {
    if (!n) return;
    k_field29<<<cdiv(n, 128), 128, 0, st>>>(op, a, b, out, n);
    ZK_HIP(hipGetLastError());
}

}  // namespace zk
