// BN254 base field Fp and scalar field Fr for gfx950 (CDNA4, wave64).
//
// Elements are 8 x 32-bit little-endian limbs in Montgomery form with
// R = 2^256, so a device value is bit-identical to the reference's 4 x u64
// limb layout (include/fp_bn254.cuh:48-69, include/fr_bn254.cuh:34-69) when
// the two limb arrays are viewed as the same 32 bytes.
//
// Multiplication is Montgomery on 32-bit words; every partial product is one
// v_mad_u64_u32 (32x32 + 64 -> 64). mul() uses the product-scanning form in
// mont_asm.hpp (column accumulators, carry-outs through VCC); mul_cios() is the
// same product in operand-scanning C++. Both moduli have a top word below
// 0x7FFFFFFE, so the running value stays below 2p and one conditional
// subtraction finishes the product.
//
// Reference counterparts: fpMul (include/fp_bn254.cuh:187-277), frMul
// (include/fr_bn254.cuh:173-210), fpAdd/fpSub/fpNeg (fp_bn254.cuh:87-140),
// frAdd/frSub/frNeg (fr_bn254.cuh:215-277).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mont_asm.hpp"

#ifndef ZK_DEV
#define ZK_DEV __device__ __forceinline__
#endif

namespace zk {

struct FpCfg {
    static constexpr uint32_t M0 = 0xd87cfd47u, M1 = 0x3c208c16u, M2 = 0x6871ca8du, M3 = 0x97816a91u,
                              M4 = 0x8181585du, M5 = 0xb85045b6u, M6 = 0xe131a029u, M7 = 0x30644e72u;
    static constexpr uint32_t NPRIME = 0xe4866389u;  // -p^-1 mod 2^32
    // R mod p (Montgomery one)
    static constexpr uint32_t O0 = 0xc58f0d9du, O1 = 0xd35d438du, O2 = 0xf5c70b3du, O3 = 0x0a78eb28u,
                              O4 = 0x7879462cu, O5 = 0x666ea36fu, O6 = 0x9a07df2fu, O7 = 0x0e0a77c1u;
    // R^2 mod p
    static constexpr uint32_t Q0 = 0x538afa89u, Q1 = 0xf32cfc5bu, Q2 = 0xd44501fbu, Q3 = 0xb5e71911u,
                              Q4 = 0x0a417ff6u, Q5 = 0x47ab1effu, Q6 = 0xcab8351fu, Q7 = 0x06d89f71u;
};

struct FrCfg {
    static constexpr uint32_t M0 = 0xf0000001u, M1 = 0x43e1f593u, M2 = 0x79b97091u, M3 = 0x2833e848u,
                              M4 = 0x8181585du, M5 = 0xb85045b6u, M6 = 0xe131a029u, M7 = 0x30644e72u;
    static constexpr uint32_t NPRIME = 0xefffffffu;  // -r^-1 mod 2^32
    static constexpr uint32_t O0 = 0x4ffffffbu, O1 = 0xac96341cu, O2 = 0x9f60cd29u, O3 = 0x36fc7695u,
                              O4 = 0x7879462eu, O5 = 0x666ea36fu, O6 = 0x9a07df2fu, O7 = 0x0e0a77c1u;
    static constexpr uint32_t Q0 = 0xae216da7u, Q1 = 0x1bb8e645u, Q2 = 0xe35c59e3u, Q3 = 0x53fe3ab1u,
                              Q4 = 0x53bb8085u, Q5 = 0x8c49833du, Q6 = 0x7f4e44a5u, Q7 = 0x0216d0b1u;
};

template <class C>
ZK_DEV uint32_t mod_word(int i) {
    switch (i) {
        case 0: return C::M0; case 1: return C::M1; case 2: return C::M2; case 3: return C::M3;
        case 4: return C::M4; case 5: return C::M5; case 6: return C::M6; default: return C::M7;
    }
}
template <class C>
ZK_DEV uint32_t one_word(int i) {
    switch (i) {
        case 0: return C::O0; case 1: return C::O1; case 2: return C::O2; case 3: return C::O3;
        case 4: return C::O4; case 5: return C::O5; case 6: return C::O6; default: return C::O7;
    }
}
template <class C>
ZK_DEV uint32_t r2_word(int i) {
    switch (i) {
        case 0: return C::Q0; case 1: return C::Q1; case 2: return C::Q2; case 3: return C::Q3;
        case 4: return C::Q4; case 5: return C::Q5; case 6: return C::Q6; default: return C::Q7;
    }
}

template <class C>
struct Fe {
    uint32_t v[8];

    ZK_DEV static Fe zero() {
        Fe r;
#pragma unroll
        for (int i = 0; i < 8; i++) r.v[i] = 0;
        return r;
    }
    ZK_DEV static Fe one() {
        Fe r;
#pragma unroll
        for (int i = 0; i < 8; i++) r.v[i] = one_word<C>(i);
        return r;
    }
    ZK_DEV bool is_zero() const {
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) acc |= v[i];
        return acc == 0;
    }
    ZK_DEV bool eq(const Fe& o) const {
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) acc |= v[i] ^ o.v[i];
        return acc == 0;
    }
};

using Fp = Fe<FpCfg>;
using Fr = Fe<FrCfg>;

// t (< 2m) -> t mod m, branch-free (select on the borrow of t - m).
template <class C>
ZK_DEV void reduce_once(uint32_t t[8]) {
    uint32_t d[8];
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d[i] = __builtin_subc(t[i], mod_word<C>(i), br, &br);
#pragma unroll
    for (int i = 0; i < 8; i++) t[i] = br ? t[i] : d[i];
}

template <class C>
ZK_DEV Fe<C> add(const Fe<C>& a, const Fe<C>& b) {
    Fe<C> r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = __builtin_addc(a.v[i], b.v[i], c, &c);
    // a + b < 2m < 2^255: no carry out of the top word.
    reduce_once<C>(r.v);
    return r;
}

template <class C>
ZK_DEV Fe<C> sub(const Fe<C>& a, const Fe<C>& b) {
    Fe<C> r;
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = __builtin_subc(a.v[i], b.v[i], br, &br);
    // on borrow add m back (masked add keeps it branch-free)
    uint32_t mask = 0u - br;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = __builtin_addc(r.v[i], mod_word<C>(i) & mask, c, &c);
    return r;
}

template <class C>
ZK_DEV Fe<C> neg(const Fe<C>& a) {
    Fe<C> r;
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = __builtin_subc(mod_word<C>(i), a.v[i], br, &br);
    // neg(0) must be 0, not m
    uint32_t nz = a.is_zero() ? 0u : 0xffffffffu;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] &= nz;
    return r;
}

template <class C>
ZK_DEV Fe<C> dbl(const Fe<C>& a) { return add<C>(a, a); }

// Montgomery product a*b*2^-256 mod m, operand-scanning CIOS in plain C++.
// Kept as the readable statement of the algorithm and as the comparison
// point for the assembly product below; mul() does not use it.
template <class C>
ZK_DEV Fe<C> mul_cios(const Fe<C>& a, const Fe<C>& b) {
    uint32_t t[8];
#pragma unroll
    for (int j = 0; j < 8; j++) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t bi = b.v[i];
        uint64_t s = (uint64_t)a.v[0] * bi + t[0];
        uint32_t A = (uint32_t)(s >> 32);
        uint32_t t0 = (uint32_t)s;
        uint32_t m = t0 * C::NPRIME;
        uint64_t u = (uint64_t)m * mod_word<C>(0) + t0;
        uint32_t Cc = (uint32_t)(u >> 32);
#pragma unroll
        for (int j = 1; j < 8; j++) {
            s = (uint64_t)a.v[j] * bi + t[j] + A;
            A = (uint32_t)(s >> 32);
            u = (uint64_t)m * mod_word<C>(j) + (uint32_t)s + Cc;
            Cc = (uint32_t)(u >> 32);
            t[j - 1] = (uint32_t)u;
        }
        t[7] = Cc + A;
    }
    reduce_once<C>(t);
    Fe<C> r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = t[i];
    return r;
}

// Montgomery product a*b*2^-256 mod m; inputs in [0,m), output in [0,m).
// Product scanning with v_mad_u64_u32 carry chains (mont_asm.hpp): about 2
// VALU ops per 32x32 partial product instead of ~5 for the compiled CIOS.
template <class C>
ZK_DEV Fe<C> mul(const Fe<C>& a, const Fe<C>& b) {
    Fe<C> r;
    mont_mul_ps<C>(r.v, a.v, b.v);
    reduce_once<C>(r.v);
    return r;
}

template <class C>
ZK_DEV Fe<C> sqr(const Fe<C>& a) { return mul<C>(a, a); }

// canonical -> Montgomery
template <class C>
ZK_DEV Fe<C> to_mont(const Fe<C>& a) {
    Fe<C> r2;
#pragma unroll
    for (int i = 0; i < 8; i++) r2.v[i] = r2_word<C>(i);
    return mul<C>(a, r2);
}

// Montgomery -> canonical (multiply by 1)
template <class C>
ZK_DEV Fe<C> from_mont(const Fe<C>& a) {
    Fe<C> one;
    one.v[0] = 1;
#pragma unroll
    for (int i = 1; i < 8; i++) one.v[i] = 0;
    return mul<C>(a, one);
}

// a^e for a 256-bit exponent given as 8 words (square-and-multiply, MSB first)
template <class C>
ZK_DEV Fe<C> pow_words(const Fe<C>& a, const uint32_t e[8]) {
    Fe<C> r = Fe<C>::one();
    for (int w = 7; w >= 0; w--) {
        for (int b = 31; b >= 0; b--) {
            r = sqr<C>(r);
            if ((e[w] >> b) & 1u) r = mul<C>(r, a);
        }
    }
    return r;
}

// a^-1 = a^(m-2) (Fermat). Only used off the hot loops (affine conversion).
template <class C>
ZK_DEV Fe<C> inv(const Fe<C>& a) {
    uint32_t e[8];
    uint32_t br = 0;
    e[0] = __builtin_subc(mod_word<C>(0), 2u, br, &br);
#pragma unroll
    for (int i = 1; i < 8; i++) e[i] = __builtin_subc(mod_word<C>(i), 0u, br, &br);
    return pow_words<C>(a, e);
}

// ---- global memory helpers: elements are stored as 32 contiguous bytes ----
template <class C>
ZK_DEV Fe<C> load(const void* base, size_t idx) {
    const uint4* p = reinterpret_cast<const uint4*>(base) + 2 * idx;
    uint4 lo = p[0], hi = p[1];
    Fe<C> r;
    r.v[0] = lo.x; r.v[1] = lo.y; r.v[2] = lo.z; r.v[3] = lo.w;
    r.v[4] = hi.x; r.v[5] = hi.y; r.v[6] = hi.z; r.v[7] = hi.w;
    return r;
}
template <class C>
ZK_DEV void store(void* base, size_t idx, const Fe<C>& a) {
    uint4* p = reinterpret_cast<uint4*>(base) + 2 * idx;
    p[0] = make_uint4(a.v[0], a.v[1], a.v[2], a.v[3]);
    p[1] = make_uint4(a.v[4], a.v[5], a.v[6], a.v[7]);
}

}  // namespace zk
