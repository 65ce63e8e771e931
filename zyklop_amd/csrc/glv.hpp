// GLV scalar split for BN254 (host and device).
//
// BN254 has a = 0, so (x, y) -> (beta*x, y) with beta a primitive cube root of
// unity in Fp is an endomorphism of both E(Fp) and the twist E'(Fp2). On the
// order-r subgroups it is multiplication by lambda, a cube root of unity mod r:
//   G1: lambda*P = (BETA_G1*x, y), G2: lambda*Q = (BETA_G2*x, y)
// with BETA_G1 = beta^2 and BETA_G2 = beta for the pair below (checked against
// textbook scalar multiplication of the generators).
//
// A canonical scalar k < r is written k = k1 + k2*lambda (mod r) with
// |k1|, |k2| < 2^127, so an MSM over n points with 254-bit scalars becomes one
// over 2n points (P_i and its image) with 128-bit signed scalars: half the
// windows, the same number of bucket entries, half the buckets.
//
// Split: (a1, b1), (a2, b2) is the short basis of {(x, y) : x + y*lambda = 0
// mod r} from the extended Euclidean algorithm on (r, lambda):
//   a1 = b2 = 0x89d3256894d213e3, a2 = 0x6f4d8248eeb859fd0be4e1541221250b,
//   b1 = -0x6f4d8248eeb859fc8211bbeb7d4f1128.
// c1 = round(b2*k/r) and c2 = round(-b1*k/r) are computed as
// (k*g + 2^255) >> 256 with g1 = round(2^256*b2/r), g2 = round(2^256*(-b1)/r);
// the truncation error of g moves c by less than 1/8, so
// |k1|, |k2| <= (1/2 + 1/8)(|a1| + |a2|) < 2^126.2. Then
//   k1 = k - c1*a1 - c2*a2,   k2 = -c1*b1 - c2*b2,
// both evaluated mod 2^128 (their true values are far inside +-2^127).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ZK_GLV_FN __host__ __device__ inline
#else
#define ZK_GLV_FN inline
#endif

namespace zk {
namespace glv {

constexpr uint64_t G1W[2] = {0xd91d232ec7e0b3d7ULL, 0x2ULL};
constexpr uint64_t G2W[3] = {0x7a7bd9d4391eb18eULL, 0x4ccef014a773d2cfULL, 0x2ULL};
constexpr uint64_t A1 = 0x89d3256894d213e3ULL;                                  // = b2
constexpr uint64_t A2[2] = {0x0be4e1541221250bULL, 0x6f4d8248eeb859fdULL};
constexpr uint64_t B1ABS[2] = {0x8211bbeb7d4f1128ULL, 0x6f4d8248eeb859fcULL};   // -b1
constexpr uint64_t B2 = 0x89d3256894d213e3ULL;

// Montgomery (R = 2^256) forms of the endomorphism constants, 8 x u32 LE
constexpr uint32_t BETA_G1_MONT[8] = {0xd782e155u, 0x71930c11u, 0xffbe3323u, 0xa6bb947cu,
                                      0xd4741444u, 0xaa303344u, 0x26594943u, 0x2c3b3f0du};
constexpr uint32_t BETA_G2_MONT[8] = {0x13e80b9cu, 0x3350c88eu, 0xdb5e56b9u, 0x7dce557cu,
                                      0xb615564au, 0x6001b4b8u, 0x020217e0u, 0x2682e617u};

ZK_GLV_FN uint64_t mulhi(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// out[0..m+1] = a[0..m-1] * b[0..1] (schoolbook, m <= 4)
template <int M, int NB>
ZK_GLV_FN void mul_wide(const uint64_t* a, const uint64_t* b, uint64_t* out) {
    for (int i = 0; i < M + NB; i++) out[i] = 0;
    for (int j = 0; j < NB; j++) {
        uint64_t carry = 0;
        for (int i = 0; i < M; i++) {
            const uint64_t lo = a[i] * b[j], hi = mulhi(a[i], b[j]);
            uint64_t s = out[i + j] + lo;
            uint64_t c1 = s < lo;
            s += carry;
            c1 += s < carry;
            out[i + j] = s;
            carry = hi + c1;   // hi <= 2^64 - 2, so no overflow
        }
        out[M + j] = carry;
    }
}

// 128-bit helpers on (lo, hi) pairs, all mod 2^128
ZK_GLV_FN void mul128_lo(uint64_t al, uint64_t ah, uint64_t bl, uint64_t bh, uint64_t& rl, uint64_t& rh) {
    rl = al * bl;
    rh = mulhi(al, bl) + al * bh + ah * bl;
}
ZK_GLV_FN void sub128(uint64_t& al, uint64_t& ah, uint64_t bl, uint64_t bh) {
    const uint64_t l = al - bl;
    ah = ah - bh - (al < bl ? 1 : 0);
    al = l;
}

// signed 128-bit value (two's complement, |v| < 2^127) -> magnitude with the
// sign in bit 127
ZK_GLV_FN void sign_mag(uint64_t l, uint64_t h, uint64_t out[2]) {
    if (h >> 63) {
        l = ~l + 1;
        h = ~h + (l == 0 ? 1 : 0);
        out[0] = l;
        out[1] = h | 0x8000000000000000ULL;
    } else {
        out[0] = l;
        out[1] = h;
    }
}

// k: canonical scalar (4 x u64 LE, < r). k1, k2: magnitude | sign << 127.
ZK_GLV_FN void split(const uint64_t k[4], uint64_t k1[2], uint64_t k2[2]) {
    uint64_t p1[6], p2[7];
    mul_wide<4, 2>(k, G1W, p1);
    mul_wide<4, 3>(k, G2W, p2);
    // + 2^255, keep bits >= 256
    uint64_t c1, c2l, c2h;
    {
        uint64_t t = p1[3] + 0x8000000000000000ULL;
        uint64_t cy = t < p1[3];
        c1 = p1[4] + cy;   // p1[5] == 0 and no carry out: c1 < 2^64
    }
    {
        uint64_t t = p2[3] + 0x8000000000000000ULL;
        uint64_t cy = t < p2[3];
        c2l = p2[4] + cy;
        c2h = p2[5] + (c2l < cy ? 1 : 0);   // c2 < 2^128
    }
    // k1 = k - c1*a1 - c2*a2 (mod 2^128)
    uint64_t xl = k[0], xh = k[1];
    sub128(xl, xh, c1 * A1, mulhi(c1, A1));
    uint64_t tl, th;
    mul128_lo(c2l, c2h, A2[0], A2[1], tl, th);
    sub128(xl, xh, tl, th);
    sign_mag(xl, xh, k1);
    // k2 = c1*|b1| - c2*b2 (mod 2^128)
    uint64_t yl, yh;
    mul128_lo(c1, 0, B1ABS[0], B1ABS[1], yl, yh);
    mul128_lo(c2l, c2h, B2, 0, tl, th);
    sub128(yl, yh, tl, th);
    sign_mag(yl, yh, k2);
}

}  // namespace glv
}  // namespace zk
