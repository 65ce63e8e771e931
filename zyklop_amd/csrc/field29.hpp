// BN254 Fp in 9 limbs of 29 bits, Montgomery with R' = 2^261, for the MSM
// bucket phase (accumulation and reduction, G1 and G2).
//
// Why: with 32-bit limbs a column of the Montgomery product sums up to 16
// products of 64 bits, so every v_mad_u64_u32 needs a second instruction to
// catch its carry (about 2 VALU ops per 32x32 product). With 29-bit limbs a
// column of 9 products a_i*b_j and 9 products m_i*p_j (each < 2^60 for inputs
// with limbs < 2^31 and 2^29) plus the incoming carry stays below 2^64: every
// product is one v_mad_u64_u32 and a column ends with one shift. The product
// costs about 210 VALU ops instead of about 320.
//
// Forms (every bound below is checked for BN254's p ~ 2^253.6):
//   N: limbs < 2^29, value < 2p       (products, buckets, converted points)
//   L: limbs < 3*2^29, value < 6p     (sub4 results; normalize() -> limbs < 2^29)
// mul(a, b) needs value(a)*value(b) < R'*p ~ 168 p^2 and per-column sums
// < 2^64; it accepts (L or N) x N and normalized(<6p) x normalized(<6p), and
// returns N (t = (ab + mp) / R' < ab/R' + p < 2p).
//
// Points enter this form once per circuit (from the standard 8 x 32-bit,
// R = 2^256 Montgomery form); the buckets stay in it through the reduction
// and only the level-2 partials leave it, so every other kernel keeps the
// standard form.
#pragma once
#include <stdint.h>

#include "curve.hpp"

namespace zk {

struct F29 {
    uint32_t l[9];
};

namespace f29c {
constexpr uint32_t MASK = (1u << 29) - 1;
// p, and -p^-1 mod 2^29
constexpr uint32_t P[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                           0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
constexpr uint32_t NP = 0x04866389u;
// 2p and 4p with every limb but the top one in [2^29, 2^30) (borrowed from
// the next limb), so that kp - b never borrows for normalized b < kp; 3 x 4p
constexpr uint32_t P2[9] = {0x30f9fa8eu, 0x2208c16cu, 0x38e5469du, 0x25aa45a0u, 0x2b0bb2efu,
                            0x25b68180u, 0x214dc281u, 0x3cb84c67u, 0x0060c89bu};
constexpr uint32_t P4[9] = {0x21f3f51cu, 0x241182dau, 0x31ca8d3bu, 0x2b548b42u, 0x361765dfu,
                            0x2b6d0301u, 0x229b8503u, 0x397098cfu, 0x00c19138u};
constexpr uint32_t P12[9] = {0x65dbdf54u, 0x6c34888eu, 0x955fa7b1u, 0x81fda1c6u, 0xa246319du,
                             0x82470903u, 0x67d28f09u, 0xac51ca6du, 0x0244b3a8u};
// 8p with every limb but the top one in [2^30, 2^31): 8p - a - b never
// borrows for normalized a, b
constexpr uint32_t P8[9] = {0x43e7ea38u, 0x482305b4u, 0x43951a76u, 0x56a91685u, 0x4c2ecbbeu,
                            0x56da0603u, 0x45370a06u, 0x52e1319eu, 0x01832271u};
// 3p, 4p, 5p normalized (a sub4 result is 0 mod p iff it equals one of them)
constexpr uint32_t P3N[9] = {0x0976f7d5u, 0x030d2224u, 0x1557e9edu, 0x087f6872u, 0x00918c68u,
                             0x0891c242u, 0x01f4a3c3u, 0x0b14729cu, 0x00912cebu};
constexpr uint32_t P4N[9] = {0x01f3f51cu, 0x041182dbu, 0x11ca8d3cu, 0x0b548b43u, 0x161765e0u,
                             0x0b6d0302u, 0x029b8504u, 0x197098d0u, 0x00c19139u};
constexpr uint32_t P5N[9] = {0x1a70f263u, 0x0515e391u, 0x0e3d308bu, 0x0e29ae14u, 0x0b9d3f58u,
                             0x0e4843c3u, 0x03426645u, 0x07ccbf04u, 0x00f1f588u};
// R' mod p (one); 2^256 mod p and 2^266 mod p (form conversions by one product)
constexpr uint32_t ONE[9] = {0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u,
                             0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};
constexpr uint32_t TO_STD[9] = {0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u,
                                0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};
constexpr uint32_t FROM_STD[9] = {0x13349ca1u, 0x1a5d84a8u, 0x0a3e5cacu, 0x100249e0u, 0x12b951e8u,
                                  0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};
constexpr uint32_t PTOP1 = 3171407u;  // (p >> 232) + 1
}  // namespace f29c

ZK_DEV F29 f29_const(const uint32_t (&c)[9]) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = c[i];
    return r;
}

// (lo * np) mod 2^29 with one full-rate v_mad_u64_u32 (a 32-bit
// v_mul_lo_u32 would issue at quarter rate); its carry-out goes to an SGPR
// pair that nothing reads. np = -m^-1 mod 2^29 for the modulus m.
template <uint32_t NPV = f29c::NP>
ZK_DEV uint32_t f29_qword(uint32_t lo) {
    uint64_t t, cc;
    asm("v_mad_u64_u32 %0, %1, %2, %3, 0" : "=v"(t), "=s"(cc) : "v"(lo), "s"(NPV));
    (void)cc;
    return (uint32_t)t & f29c::MASK;
}

// moduli of the 29-bit Montgomery product: limbs of m and -m^-1 mod 2^29
struct F29ModP {
    static constexpr const uint32_t* L = f29c::P;
    static constexpr uint32_t NEG_INV = f29c::NP;
};

// Montgomery product a*b/R' mod m (product scanning, one 64-bit accumulator)
template <class M>
ZK_DEV F29 f29_mul_mod(const F29& a, const F29& b) {
    using namespace f29c;
    uint32_t m[9];
    F29 t;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const int j = k - i;
            if (j >= 0 && j < 9) acc += (uint64_t)a.l[i] * b.l[j];
        }
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const int j = k - i;
            if (i < k && j >= 0 && j < 9) acc += (uint64_t)m[i] * M::L[j];
        }
        if (k < 9) {
            m[k] = f29_qword<M::NEG_INV>((uint32_t)acc);
            acc += (uint64_t)m[k] * M::L[0];
        } else {
            t.l[k - 9] = (uint32_t)acc & MASK;
        }
        acc >>= 29;
        // keeps the columns of consecutive products in program order: without it
        // the scheduler interleaves independent products (such as the three
        // Karatsuba products of Fp2) and runs out of registers
        asm volatile("" : "+v"(acc));
    }
    t.l[8] = (uint32_t)acc;
    return t;
}

ZK_DEV F29 f29_mul(const F29& a, const F29& b) { return f29_mul_mod<F29ModP>(a, b); }

// Montgomery square a*a/R' mod p for normalized a: 45 partial products
// (the off-diagonal a_i*a_j, i < j, taken once against 2*a_j, which stays
// below 2^30) instead of 81; a column then sums at most 5 products below
// 2^59, 9 reduction products below 2^58 and the carry, under 2^64.
ZK_DEV F29 f29_sqr(const F29& a) {
    using namespace f29c;
    uint32_t m[9], d[9];
#pragma unroll
    for (int i = 0; i < 9; i++) d[i] = a.l[i] << 1;
    F29 t;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const int j = k - i;
            if (j > i && j < 9) acc += (uint64_t)a.l[i] * d[j];
            if (j == i) acc += (uint64_t)a.l[i] * a.l[i];
        }
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const int j = k - i;
            if (i < k && j >= 0 && j < 9) acc += (uint64_t)m[i] * P[j];
        }
        if (k < 9) {
            m[k] = f29_qword((uint32_t)acc);
            acc += (uint64_t)m[k] * P[0];
        } else {
            t.l[k - 9] = (uint32_t)acc & MASK;
        }
        acc >>= 29;
        asm volatile("" : "+v"(acc));
    }
    t.l[8] = (uint32_t)acc;
    return t;
}

// sum of N Montgomery products, sum_k a_k*b_k/R' mod p, with one reduction
// (product scanning as in f29_mul). Operands normalized (limbs < 2^29); a
// column then sums at most (2N + 1)*9 products below 2^58 plus the carry,
// under 2^64 for N <= 4. With operands below 6p and N*36p^2 < 2^261*p/2 the
// result is below 2p. A difference a*b - c*d is the sum with c replaced by
// f29_neg(c).
template <int N>
ZK_DEV F29 f29_mulsum(const F29 (&a)[N], const F29 (&b)[N]) {
    static_assert(N >= 1 && N <= 4, "column sums must stay below 2^64");
    using namespace f29c;
    uint32_t m[9];
    F29 t;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
#pragma unroll
        for (int q = 0; q < N; q++)
#pragma unroll
            for (int i = 0; i < 9; i++) {
                const int j = k - i;
                if (j >= 0 && j < 9) acc += (uint64_t)a[q].l[i] * b[q].l[j];
            }
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const int j = k - i;
            if (i < k && j >= 0 && j < 9) acc += (uint64_t)m[i] * P[j];
        }
        if (k < 9) {
            m[k] = f29_qword((uint32_t)acc);
            acc += (uint64_t)m[k] * P[0];
        } else {
            t.l[k - 9] = (uint32_t)acc & MASK;
        }
        acc >>= 29;
        asm volatile("" : "+v"(acc));
    }
    t.l[8] = (uint32_t)acc;
    return t;
}

// carry propagation: limbs < 2^32 -> limbs < 2^29 (same value)
ZK_DEV F29 f29_norm(const F29& a) {
    F29 r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint32_t x = a.l[i] + c;
        r.l[i] = i < 8 ? (x & f29c::MASK) : x;
        c = x >> 29;
    }
    return r;
}

// a + 4p - b for normalized a, b (value < 2p): form L, value in (2p, 6p)
ZK_DEV F29 f29_sub4(const F29& a, const F29& b) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = a.l[i] + f29c::P4[i] - b.l[i];
    return r;
}

// normalized value < 16p -> value < 2p (subtract q*p, q from the top limb)
ZK_DEV F29 f29_reduce(const F29& a) {
    const uint32_t q = a.l[8] / f29c::PTOP1;
    F29 r;
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int64_t x = (int64_t)a.l[i] - (int64_t)((uint64_t)q * f29c::P[i]) + c;
        r.l[i] = (uint32_t)x & f29c::MASK;
        c = x >> 29;
    }
    return r;
}

ZK_DEV bool f29_eq(const F29& a, const uint32_t (&c)[9]) {
    uint32_t d = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) d |= a.l[i] ^ c[i];
    return d == 0;
}

// for normalized x = u + 4p - v (u, v < 2p): x = 0 mod p iff x is 3p, 4p or 5p
ZK_DEV bool f29_sub4_is_zero(const F29& x) {
    using namespace f29c;
    const uint32_t l0 = x.l[0];
    if (l0 != P3N[0] && l0 != P4N[0] && l0 != P5N[0]) return false;
    return f29_eq(x, P3N) || f29_eq(x, P4N) || f29_eq(x, P5N);
}

ZK_DEV bool f29_is_zero(const F29& a) {
    uint32_t d = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) d |= a.l[i];
    return d == 0;
}

// -y for y < p: 2p - y normalized (value in (p, 2p])
ZK_DEV F29 f29_neg(const F29& y) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = f29c::P2[i] - y.l[i];
    return f29_norm(r);
}

// -y without carry propagation, as an operand of the products below: 2p - y
// for y < p (value in (p, 2p]) or 4p - y for normalized y < 2p (value in
// (2p, 4p]); every limb lies in (0, 2^30) (the constants borrow into
// [2^29, 2^30) and the top limb of y stays below theirs), so a product
// column that takes one such factor per term sums terms below 2^59
ZK_DEV F29 f29_neg_lazy2(const F29& y) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = f29c::P2[i] - y.l[i];
    return r;
}
ZK_DEV F29 f29_neg_lazy4(const F29& y) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = f29c::P4[i] - y.l[i];
    return r;
}

// limb-wise sum (no carry propagation)
ZK_DEV F29 f29_add(const F29& a, const F29& b) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = a.l[i] + b.l[i];
    return r;
}

// 8 x 32-bit words (value < 2^256) <-> 9 x 29-bit limbs
ZK_DEV F29 f29_unpack(const uint32_t w[8]) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int bit = 29 * i, k = bit >> 5, sh = bit & 31;
        const uint32_t lo = w[k], hi = k + 1 < 8 ? w[k + 1] : 0u;
        r.l[i] = __builtin_amdgcn_alignbit(hi, lo, sh) & f29c::MASK;
    }
    return r;
}
ZK_DEV void f29_pack(const F29& a, uint32_t w[8]) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int bit = 32 * k, i = bit / 29, sh = bit % 29;
        uint64_t v = (uint64_t)a.l[i] >> sh;
        if (i + 1 < 9) v |= (uint64_t)a.l[i + 1] << (29 - sh);
        if (i + 2 < 9 && 58 - sh < 64) v |= (uint64_t)a.l[i + 2] << (58 - sh);
        w[k] = (uint32_t)v;
    }
}

// standard Montgomery (x*2^256, < p) -> this form (x*2^261, < p)
ZK_DEV F29 f29_from_std(const Fp& x) {
    F29 t = f29_mul(f29_unpack(x.v), f29_const(f29c::FROM_STD));
    Fp s;
    f29_pack(t, s.v);
    reduce_once<FpCfg>(s.v);  // t < 2p -> < p
    return f29_unpack(s.v);
}

// this form (x*2^261, value < 2p) -> standard Montgomery (x*2^256, < p)
ZK_DEV Fp f29_to_std(const F29& v) {
    F29 t = f29_mul(v, f29_const(f29c::TO_STD));
    Fp s;
    f29_pack(t, s.v);
    reduce_once<FpCfg>(s.v);
    return s;
}

// XYZZ bucket in this form; inf = no point added yet (or the sum is infinity)
struct Xyzz29 {
    F29 X, Y, ZZ, ZZZ;
    bool inf;
};

// madd-2008-s, acc += (x2, y2) with x2 < p and y2 < 2p, normalized or the
// unnormalized negation of a value below p (f29_neg_lazy2). The equal-x
// cases (doubling, or the sum is infinity) are rare: the doubling goes through
// the standard-form formulas.
ZK_DEV void f29_madd(Xyzz29& A, const F29& x2, const F29& y2) {
    using namespace f29c;
    if (A.inf) {
        A.X = x2;
        A.Y = f29_norm(y2);
        A.ZZ = f29_const(ONE);
        A.ZZZ = f29_const(ONE);
        A.inf = false;
        return;
    }
    const F29 U2 = f29_mul(x2, A.ZZ);
    const F29 S2 = f29_mul(y2, A.ZZZ);
    const F29 P_ = f29_norm(f29_sub4(U2, A.X));
    const F29 R_ = f29_norm(f29_sub4(S2, A.Y));
    if (f29_sub4_is_zero(P_)) {
        if (f29_sub4_is_zero(R_)) {
            Aff<Fp> a;
            a.x = f29_to_std(x2);
            a.y = f29_to_std(y2);
            const Xyzz<Fp> d = xyzz_mdbl(a);
            A.X = f29_from_std(d.X);
            A.Y = f29_from_std(d.Y);
            A.ZZ = f29_from_std(d.ZZ);
            A.ZZZ = f29_from_std(d.ZZZ);
        } else {
            A.inf = true;
        }
        return;
    }
    const F29 PP = f29_sqr(P_);
    const F29 PPP = f29_mul(P_, PP);
    const F29 Q = f29_mul(A.X, PP);
    const F29 RR = f29_sqr(R_);
    // X3 = RR - PPP - 2Q + 12p: limbs stay in [0, 2^32)
    F29 x3;
#pragma unroll
    for (int i = 0; i < 9; i++) x3.l[i] = (P12[i] - PPP.l[i] - 2 * Q.l[i]) + RR.l[i];
    const F29 X3 = f29_reduce(f29_norm(x3));
    // Y3 = R*(Q - X3) - Y1*PPP with one reduction: R and T = Q - X3 below 6p,
    // Y1 below 2p and -PPP <= 4p (f29_neg_lazy4), so the products sum below
    // 44p^2. T stays unnormalized (limbs below 2^31: X3 < 2p keeps every limb
    // of Q + 4p - X3 non-negative), so a column sums 9 terms below 2^60, 9
    // below 2^59 and 9 below 2^58: under 63 * 2^58.
    const F29 T = f29_sub4(Q, X3);
    const F29 ya[2] = {R_, A.Y}, yb[2] = {T, f29_neg_lazy4(PPP)};
    A.Y = f29_mulsum<2>(ya, yb);
    A.X = X3;
    A.ZZ = f29_mul(A.ZZ, PP);
    A.ZZZ = f29_mul(A.ZZZ, PPP);
}

ZK_DEV Xyzz29 xyzz29_inf() {
    Xyzz29 r;
    r.X = r.Y = r.ZZ = r.ZZZ = F29{};
    r.inf = true;
    return r;
}

// General XYZZ operations of this form, for the bucket reduction.
// dbl-2008-s-1 (a = 0) on an N point: U = 2Y, V = U^2, W = U*V, S = X*V,
// M = 3X^2, X3 = M^2 - 2S, Y3 = M*(S - X3) - W*Y, ZZ3 = V*ZZ, ZZZ3 = W*ZZZ.
// Bounds: U normalized < 4p, M normalized < 6p (both fine for f29_sqr /
// f29_mul); M^2 + 8p - 2S lies in (4p, 10p) with every limb of 8p - 2S
// positive; Y3 sums M*(S - X3) < 36p^2 and W*(4p - Y) < 8p^2 (f29_mulsum<2>
// bound, columns as in f29_madd). BN254 G1 has no point of order 2, so Y = 0 only at infinity.
ZK_DEV Xyzz29 f29_xyzz_dbl(const Xyzz29& p) {
    using namespace f29c;
    if (p.inf) return p;
    const F29 U = f29_norm(f29_add(p.Y, p.Y));
    const F29 V = f29_sqr(U);
    const F29 Wv = f29_mul(U, V);
    const F29 S = f29_mul(p.X, V);
    const F29 XX = f29_sqr(p.X);
    const F29 M = f29_norm(f29_add(f29_add(XX, XX), XX));
    const F29 MM = f29_sqr(M);
    F29 x3;
#pragma unroll
    for (int i = 0; i < 9; i++) x3.l[i] = (P8[i] - S.l[i] - S.l[i]) + MM.l[i];
    const F29 X3 = f29_reduce(f29_norm(x3));
    const F29 T = f29_sub4(S, X3);   // unnormalized, as in f29_madd
    const F29 ya[2] = {M, Wv}, yb[2] = {T, f29_neg_lazy4(p.Y)};
    Xyzz29 r;
    r.X = X3;
    r.Y = f29_mulsum<2>(ya, yb);
    r.ZZ = f29_mul(V, p.ZZ);
    r.ZZZ = f29_mul(Wv, p.ZZZ);
    r.inf = false;
    return r;
}

// add-2008-s: p + q for N points, either of which may be infinity; equal
// points go through the doubling, opposite points give infinity. Same bounds
// as f29_madd; the result is N.
ZK_DEV Xyzz29 f29_xyzz_add(const Xyzz29& p, const Xyzz29& q) {
    using namespace f29c;
    if (p.inf) return q;
    if (q.inf) return p;
    const F29 U1 = f29_mul(p.X, q.ZZ), U2 = f29_mul(q.X, p.ZZ);
    const F29 S1 = f29_mul(p.Y, q.ZZZ), S2 = f29_mul(q.Y, p.ZZZ);
    const F29 P_ = f29_norm(f29_sub4(U2, U1));
    const F29 R_ = f29_norm(f29_sub4(S2, S1));
    if (f29_sub4_is_zero(P_)) {
        if (f29_sub4_is_zero(R_)) return f29_xyzz_dbl(p);
        return xyzz29_inf();
    }
    const F29 PP = f29_sqr(P_);
    const F29 PPP = f29_mul(P_, PP);
    const F29 Q = f29_mul(U1, PP);
    const F29 RR = f29_sqr(R_);
    F29 x3;
#pragma unroll
    for (int i = 0; i < 9; i++) x3.l[i] = (P12[i] - PPP.l[i] - 2 * Q.l[i]) + RR.l[i];
    const F29 X3 = f29_reduce(f29_norm(x3));
    const F29 T = f29_sub4(Q, X3);   // unnormalized, as in f29_madd
    const F29 ya[2] = {R_, S1}, yb[2] = {T, f29_neg_lazy4(PPP)};
    Xyzz29 r;
    r.X = X3;
    r.Y = f29_mulsum<2>(ya, yb);
    r.ZZ = f29_mul(f29_mul(p.ZZ, q.ZZ), PP);
    r.ZZZ = f29_mul(f29_mul(p.ZZZ, q.ZZZ), PPP);
    r.inf = false;
    return r;
}

// An XYZZ point of this form in 32 words: X, Y, ZZ, ZZZ packed (values < 2p
// fit 256 bits). Infinity is stored with ZZ = 0, which no finite point has
// (its ZZ is a nonzero square mod p, so neither 0 nor p).
ZK_DEV void xyzz29_pack(const Xyzz29& p, uint32_t w[32]) {
    f29_pack(p.X, w);
    f29_pack(p.Y, w + 8);
    f29_pack(p.ZZ, w + 16);
    f29_pack(p.ZZZ, w + 24);
    if (p.inf) {
#pragma unroll
        for (int k = 16; k < 24; k++) w[k] = 0;
    }
}
ZK_DEV Xyzz29 xyzz29_unpack(const uint32_t w[32]) {
    Xyzz29 r;
    r.X = f29_unpack(w);
    r.Y = f29_unpack(w + 8);
    r.ZZ = f29_unpack(w + 16);
    r.ZZZ = f29_unpack(w + 24);
    uint32_t z = 0;
#pragma unroll
    for (int k = 16; k < 24; k++) z |= w[k];
    r.inf = z == 0;
    return r;
}

// One 128-byte slot per G1 base P = (x, y): the two bases of the GLV split,
// P and its endomorphism image lambda*P = (beta*x, y), each as a 64-byte half
// holding its coordinates in this form packed into 8 words (canonical, < p):
//   words 0..7 y, 8..15 x | 16..23 y, 24..31 beta*x.
// A bucket entry names its base as slot | half << 30 (bit 31 is the digit's
// sign): half 0 reads the first half of the slot, half 1 the second. A
// gather thus touches one aligned 64-byte half.
constexpr int PT29_SLOT_UINT4 = 8;
constexpr uint32_t PT29_SLOT_MASK = 0x3fffffffu;

ZK_DEV F29 f29_load_packed(const uint4* q) {
    const uint4 a = q[0], b = q[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return f29_unpack(w);
}

ZK_DEV void f29_load_point(const void* base, uint32_t ref, F29& x, F29& y) {
    const uint4* p = reinterpret_cast<const uint4*>(base) + (size_t)PT29_SLOT_UINT4 * (ref & PT29_SLOT_MASK) +
                     ((ref >> 28) & 4u);
    y = f29_load_packed(p);
    x = f29_load_packed(p + 2);
}

// ---------------------------------------------------------------- Fr (NTT)
// BN254's group order r in the same limbs, for the NTT. Its data stay in the
// standard form (x*2^256 mod r, canonical, 8 words in memory); its tables are
// prescaled by 2^5 = R'/2^256, so that a*b/R' of a standard-form a and a
// prescaled b is the standard-form product and no element changes form.
// Bounds (r ~ 2^253.6, R'/r ~ 169): for b < r and any a < 2^261 whose limbs
// are below 2^31, a*b/R' < r*(a/R') + r < 2r, and a column sums at most 9
// products below 2^60, 9 below 2^58 and the carry, under 2^64.
namespace fr29c {
constexpr uint32_t R[9] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u,
                           0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
constexpr uint32_t NR = 0x0fffffffu;   // -r^-1 mod 2^29
// 2^k r with every limb but the top one in [2^29, 2^30) (borrowed from the
// next limb), so that 2^k r - b never borrows for normalized b < 2^k r
constexpr uint32_t R2[9] = {0x20000002u, 0x3e1f593eu, 0x3cb848a0u, 0x2fa121e5u, 0x2b0ba505u,
                            0x25b68180u, 0x214dc281u, 0x3cb84c67u, 0x0060c89bu};
constexpr uint32_t R4[9] = {0x20000004u, 0x3c3eb27du, 0x39709142u, 0x3f4243ccu, 0x36174a0bu,
                            0x2b6d0301u, 0x229b8503u, 0x397098cfu, 0x00c19138u};
constexpr uint32_t R8[9] = {0x20000008u, 0x387d64fbu, 0x32e12286u, 0x3e84879au, 0x2c2e9418u,
                            0x36da0604u, 0x25370a07u, 0x32e1319fu, 0x01832272u};
constexpr uint32_t R16[9] = {0x20000010u, 0x30fac9f7u, 0x25c2450eu, 0x3d090f36u, 0x385d2832u,
                             0x2db40c09u, 0x2a6e1410u, 0x25c2633fu, 0x030644e6u};
constexpr uint32_t RTOP1 = 3171407u;   // (r >> 232) + 1
template <int K> struct ROff;          // 2^K r, borrowed
template <> struct ROff<1> { static constexpr const uint32_t* L = R2; };
template <> struct ROff<2> { static constexpr const uint32_t* L = R4; };
template <> struct ROff<3> { static constexpr const uint32_t* L = R8; };
template <> struct ROff<4> { static constexpr const uint32_t* L = R16; };
}  // namespace fr29c

struct F29ModR {
    static constexpr const uint32_t* L = fr29c::R;
    static constexpr uint32_t NEG_INV = fr29c::NR;
};

ZK_DEV F29 fr29_mul(const F29& a, const F29& b) { return f29_mul_mod<F29ModR>(a, b); }

// a - b + 2^K r limb-wise, for normalized b < 2^K r: limbs below 2^31 (the
// top one wraps below zero only when a carry from below makes it up)
template <int K>
ZK_DEV F29 fr29_sub(const F29& a, const F29& b) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = a.l[i] + fr29c::ROff<K>::L[i] - b.l[i];
    return r;
}

// normalized value < 2^261 -> value < 2r (subtract q*r, q from the top limb:
// q <= value/r and value - q*r < r + value/(r >> 232) < 2r)
ZK_DEV F29 fr29_reduce(const F29& a) {
    const uint32_t q = a.l[8] / fr29c::RTOP1;
    F29 r;
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int64_t x = (int64_t)a.l[i] - (int64_t)((uint64_t)q * fr29c::R[i]) + c;
        r.l[i] = (uint32_t)x & f29c::MASK;
        c = x >> 29;
    }
    return r;
}

// normalized value < 2r -> canonical (< r)
ZK_DEV F29 fr29_canon(const F29& a) {
    F29 d;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int32_t x = (int32_t)a.l[i] - (int32_t)fr29c::R[i] + c;
        d.l[i] = (uint32_t)x & f29c::MASK;
        c = x >> 29;
    }
    return c < 0 ? a : d;
}

// ---------------------------------------------------------------- Fq2 (G2)
// Fp2 = Fp[u]/(u^2 + 1) over this form, component-wise bounds as above.
struct F29q {
    F29 c0, c1;
};

// a: components normalized with value < 6p, or lazy negations of values below
// p (f29_neg_lazy2); b: components N. Each component is a sum of two products
// with one reduction (f29_mulsum): c0 = a0*b0 + a1*(-b1), c1 = a0*b1 + a1*b0,
// -b1 = 4p - b1 unnormalized (a column stays below 63 * 2^58); result N. Four products and two reductions
// cost fewer instructions than Karatsuba's three products, three reductions
// and the normalisations of its differences.
ZK_DEV F29q f29q_mul(const F29q& a, const F29q& b) {
    const F29 x[2] = {a.c0, a.c1};
    const F29 y0[2] = {b.c0, f29_neg_lazy4(b.c1)}, y1[2] = {b.c1, b.c0};
    return {f29_mulsum<2>(x, y0), f29_mulsum<2>(x, y1)};
}

// a: components N. (a0+a1)(a0-a1) + (2 a0 a1) u; result N.
ZK_DEV F29q f29q_sqr(const F29q& a) {
    const F29 d = f29_norm(f29_sub4(a.c0, a.c1));
    return {f29_mul(f29_add(a.c0, a.c1), d), f29_mul(f29_add(a.c0, a.c0), a.c1)};
}

// a - b component-wise for N inputs: reduced to N, or only normalized (< 6p)
ZK_DEV F29q f29q_sub(const F29q& a, const F29q& b) {
    return {f29_reduce(f29_norm(f29_sub4(a.c0, b.c0))), f29_reduce(f29_norm(f29_sub4(a.c1, b.c1)))};
}
ZK_DEV F29q f29q_sub_nr(const F29q& a, const F29q& b) {
    return {f29_norm(f29_sub4(a.c0, b.c0)), f29_norm(f29_sub4(a.c1, b.c1))};
}

ZK_DEV F29q f29q_from_std(const Fq2& x) { return {f29_from_std(x.c0), f29_from_std(x.c1)}; }
ZK_DEV Fq2 f29q_to_std(const F29q& v) { return {f29_to_std(v.c0), f29_to_std(v.c1)}; }

// ---- general XYZZ operations over Fp2 in this form (the G2 bucket
// reduction); bounds as in f29q_madd
ZK_DEV F29q f29q_reduce(const F29q& a) { return {f29_reduce(a.c0), f29_reduce(a.c1)}; }
ZK_DEV bool f29q_sub_is_zero(const F29q& x) { return f29_sub4_is_zero(x.c0) && f29_sub4_is_zero(x.c1); }

// T*A - B*C, each component one sum of four products with one reduction:
// T normalized < 6p, A, B, C N; the negations are 4p - x unnormalized (two of
// the four terms of a column carry one, so a column stays below 63 * 2^58),
// and the products sum below 52p^2
ZK_DEV F29q f29q_mul_sub_mul(const F29q& T, const F29q& A, const F29q& B, const F29q& C) {
    const F29 nb0 = f29_neg_lazy4(B.c0), nb1 = f29_neg_lazy4(B.c1), na1 = f29_neg_lazy4(A.c1);
    const F29 a0[4] = {T.c0, T.c1, nb0, B.c1}, b0[4] = {A.c0, na1, C.c0, C.c1};
    const F29 a1[4] = {T.c0, T.c1, nb0, nb1}, b1[4] = {A.c1, A.c0, C.c1, C.c0};
    return {f29_mulsum<4>(a0, b0), f29_mulsum<4>(a1, b1)};
}

struct Xyzz29q {
    F29q X, Y, ZZ, ZZZ;
    bool inf;
};

ZK_DEV Xyzz29q xyzz29q_inf() {
    Xyzz29q r;
    r.X = r.Y = r.ZZ = r.ZZZ = F29q{};
    r.inf = true;
    return r;
}

// dbl-2008-s-1 over Fp2 (G2 has no point of order 2 either)
ZK_DEV Xyzz29q f29q_xyzz_dbl(const Xyzz29q& p) {
    using namespace f29c;
    if (p.inf) return p;
    const F29q U = {f29_reduce(f29_norm(f29_add(p.Y.c0, p.Y.c0))), f29_reduce(f29_norm(f29_add(p.Y.c1, p.Y.c1)))};
    const F29q V = f29q_sqr(U);
    const F29q Wv = f29q_mul(U, V);
    const F29q S = f29q_mul(p.X, V);
    const F29q XX = f29q_sqr(p.X);
    const F29q M = {f29_reduce(f29_norm(f29_add(f29_add(XX.c0, XX.c0), XX.c0))),
                    f29_reduce(f29_norm(f29_add(f29_add(XX.c1, XX.c1), XX.c1)))};
    const F29q MM = f29q_sqr(M);
    F29 x0, x1;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        x0.l[i] = (P8[i] - S.c0.l[i] - S.c0.l[i]) + MM.c0.l[i];
        x1.l[i] = (P8[i] - S.c1.l[i] - S.c1.l[i]) + MM.c1.l[i];
    }
    Xyzz29q r;
    r.X = {f29_reduce(f29_norm(x0)), f29_reduce(f29_norm(x1))};
    r.Y = f29q_mul_sub_mul(f29q_sub_nr(S, r.X), M, Wv, p.Y);
    r.ZZ = f29q_mul(V, p.ZZ);
    r.ZZZ = f29q_mul(Wv, p.ZZZ);
    r.inf = false;
    return r;
}

// add-2008-s over Fp2
ZK_DEV Xyzz29q f29q_xyzz_add(const Xyzz29q& p, const Xyzz29q& q) {
    using namespace f29c;
    if (p.inf) return q;
    if (q.inf) return p;
    const F29q U1 = f29q_mul(p.X, q.ZZ), U2 = f29q_mul(q.X, p.ZZ);
    const F29q S1 = f29q_mul(p.Y, q.ZZZ), S2 = f29q_mul(q.Y, p.ZZZ);
    const F29q Pn = f29q_sub_nr(U2, U1), Rn = f29q_sub_nr(S2, S1);
    if (f29q_sub_is_zero(Pn)) {
        if (f29q_sub_is_zero(Rn)) return f29q_xyzz_dbl(p);
        return xyzz29q_inf();
    }
    const F29q P_ = f29q_reduce(Pn), R_ = f29q_reduce(Rn);
    const F29q PP = f29q_sqr(P_);
    const F29q PPP = f29q_mul(P_, PP);
    const F29q Q = f29q_mul(U1, PP);
    const F29q RR = f29q_sqr(R_);
    F29 x0, x1;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        x0.l[i] = (P12[i] - PPP.c0.l[i] - 2 * Q.c0.l[i]) + RR.c0.l[i];
        x1.l[i] = (P12[i] - PPP.c1.l[i] - 2 * Q.c1.l[i]) + RR.c1.l[i];
    }
    Xyzz29q r;
    r.X = {f29_reduce(f29_norm(x0)), f29_reduce(f29_norm(x1))};
    r.Y = f29q_mul_sub_mul(f29q_sub_nr(Q, r.X), R_, S1, PPP);
    r.ZZ = f29q_mul(f29q_mul(p.ZZ, q.ZZ), PP);
    r.ZZZ = f29q_mul(f29q_mul(p.ZZZ, q.ZZZ), PPP);
    r.inf = false;
    return r;
}

// 64 words: X, Y, ZZ, ZZZ, each c0 then c1 packed; infinity has ZZ = 0
ZK_DEV void xyzz29q_pack(const Xyzz29q& p, uint32_t w[64]) {
    const F29q* c[4] = {&p.X, &p.Y, &p.ZZ, &p.ZZZ};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        f29_pack(c[k]->c0, w + 16 * k);
        f29_pack(c[k]->c1, w + 16 * k + 8);
    }
    if (p.inf) {
#pragma unroll
        for (int k = 32; k < 48; k++) w[k] = 0;
    }
}
ZK_DEV Xyzz29q xyzz29q_unpack(const uint32_t w[64]) {
    Xyzz29q r;
    F29q* c[4] = {&r.X, &r.Y, &r.ZZ, &r.ZZZ};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        c[k]->c0 = f29_unpack(w + 16 * k);
        c[k]->c1 = f29_unpack(w + 16 * k + 8);
    }
    uint32_t z = 0;
#pragma unroll
    for (int k = 32; k < 48; k++) z |= w[k];
    r.inf = z == 0;
    return r;
}

// One 256-byte slot per G2 base, halves as for G1 (one 128-byte line per
// gather): words 0..15 y (c0, c1), 16..31 x | 32..47 y, 48..63 beta*x;
// entries name slot and half as for G1.
constexpr int PT29Q_SLOT_UINT4 = 16;

ZK_DEV void f29q_load_point(const void* base, uint32_t ref, F29q& x, F29q& y) {
    const uint4* p = reinterpret_cast<const uint4*>(base) + (size_t)PT29Q_SLOT_UINT4 * (ref & PT29_SLOT_MASK) +
                     ((ref >> 27) & 8u);
    y.c0 = f29_load_packed(p);
    y.c1 = f29_load_packed(p + 2);
    x.c0 = f29_load_packed(p + 4);
    x.c1 = f29_load_packed(p + 6);
}

// A G2 XYZZ accumulator: X, Y and ZZZ in LDS, 3 coordinates x 18 limbs laid
// out [coordinate][limb][lane] over a block of NT threads (consecutive lanes
// hit consecutive banks), and ZZ in registers. The LDS coordinates are loaded
// where they are used, so they do not occupy 54 registers for the whole loop;
// with ZZ in registers the LDS share per thread is 216 B, which leaves room
// for more waves per CU than 288 B would.
template <int NT>
struct AccLds {
    uint32_t* base;   // &lds[0][0][lane]
    F29q zz;
    enum { X = 0, Y = 1, ZZZ = 2, ZZ = 3 };
    static constexpr int LDS_WORDS = 3 * 18 * NT;
    ZK_DEV F29q ld(int c) const {
        if (c == ZZ) return zz;
        asm volatile("" ::: "memory");
        F29q v;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            v.c0.l[i] = base[(c * 18 + i) * NT];
            v.c1.l[i] = base[(c * 18 + 9 + i) * NT];
        }
        return v;
    }
    ZK_DEV void st(int c, const F29q& v) {
        if (c == ZZ) {
            zz = v;
            return;
        }
#pragma unroll
        for (int i = 0; i < 9; i++) {
            base[(c * 18 + i) * NT] = v.c0.l[i];
            base[(c * 18 + 9 + i) * NT] = v.c1.l[i];
        }
        asm volatile("" ::: "memory");
    }
};

// madd-2008-s over Fp2 on that accumulator, acc += (x2, y2) with components
// of x2 < p and y2 < 2p, normalized or f29_neg_lazy2 of values below p; inf
// tracks the empty/infinite sum. The
// equal-x cases go through the standard-form formulas.
template <int NT>
ZK_DEV void f29q_madd(AccLds<NT>& A, bool& inf, const F29q& x2, const F29q& y2) {
    using namespace f29c;
    using L = AccLds<NT>;
    if (inf) {
        const F29q one = {f29_const(ONE), F29{}};
        A.st(L::X, x2);
        A.st(L::Y, F29q{f29_norm(y2.c0), f29_norm(y2.c1)});
        A.st(L::ZZ, one);
        A.st(L::ZZZ, one);
        inf = false;
        return;
    }
    const F29q U2 = f29q_mul(x2, A.ld(L::ZZ));
    const F29q S2 = f29q_mul(y2, A.ld(L::ZZZ));
    const F29q Pn = f29q_sub_nr(U2, A.ld(L::X));   // normalized, in (2p, 6p): zero test
    const F29q Rn = f29q_sub_nr(S2, A.ld(L::Y));
    if (f29_sub4_is_zero(Pn.c0) && f29_sub4_is_zero(Pn.c1)) {
        if (f29_sub4_is_zero(Rn.c0) && f29_sub4_is_zero(Rn.c1)) {
            Aff<Fq2> a;
            a.x = f29q_to_std(x2);
            a.y = f29q_to_std(y2);
            const Xyzz<Fq2> d = xyzz_mdbl(a);
            A.st(L::X, f29q_from_std(d.X));
            A.st(L::Y, f29q_from_std(d.Y));
            A.st(L::ZZ, f29q_from_std(d.ZZ));
            A.st(L::ZZZ, f29q_from_std(d.ZZZ));
        } else {
            inf = true;
        }
        return;
    }
    const F29q P_ = {f29_reduce(Pn.c0), f29_reduce(Pn.c1)};
    const F29q PP = f29q_sqr(P_);
    const F29q PPP = f29q_mul(P_, PP);
    A.st(L::ZZ, f29q_mul(A.ld(L::ZZ), PP));
    A.st(L::ZZZ, f29q_mul(A.ld(L::ZZZ), PPP));
    const F29q Q = f29q_mul(A.ld(L::X), PP);
    const F29q R_ = {f29_reduce(Rn.c0), f29_reduce(Rn.c1)};
    F29q X3;
    {
        const F29q RR = f29q_sqr(R_);
        F29 x0, x1;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            x0.l[i] = (P12[i] - PPP.c0.l[i] - 2 * Q.c0.l[i]) + RR.c0.l[i];
            x1.l[i] = (P12[i] - PPP.c1.l[i] - 2 * Q.c1.l[i]) + RR.c1.l[i];
        }
        X3 = {f29_reduce(f29_norm(x0)), f29_reduce(f29_norm(x1))};
    }
    A.st(L::X, X3);
    // Y3 = (Q - X3)*R - Y1*PPP, each Fp2 component as one sum of four
    // products with one reduction: T = Q - X3 < 6p, R, Y1, PPP < 2p and the
    // negations 4p - x unnormalized (f29_neg_lazy4), so the products sum
    // below 48p^2 and a column, two of whose four terms carry a negation,
    // below 63 * 2^58
    const F29q T = f29q_sub_nr(Q, X3);
    const F29q Y1 = A.ld(L::Y);
    const F29 nr1 = f29_neg_lazy4(R_.c1), ny0 = f29_neg_lazy4(Y1.c0), ny1 = f29_neg_lazy4(Y1.c1);
    const F29 a0[4] = {T.c0, T.c1, ny0, Y1.c1}, b0[4] = {R_.c0, nr1, PPP.c0, PPP.c1};
    const F29 a1[4] = {T.c0, T.c1, ny0, ny1}, b1[4] = {R_.c1, R_.c0, PPP.c1, PPP.c0};
    A.st(L::Y, F29q{f29_mulsum<4>(a0, b0), f29_mulsum<4>(a1, b1)});
}

}  // namespace zk
