// Device curve arithmetic for BN254 G1 (over Fp) and G2 (over Fp2 = Fp[u]/(u^2+1)).
//
// Buckets are XYZZ points (x = X/ZZ, y = Y/ZZZ): a mixed add costs 8M+2S and a
// full add 12M+2S, against 7M+4S / 11M+5S for the reference's Jacobian
// formulas (include/fp_bn254.cuh:931-1022, include/fp2_bn254.cuh:296-403).
// Affine inputs use the zkey convention: (0,0) is the point at infinity
// (include/msm_g1.cuh:243-244). The P == Q and P == -Q cases are handled.
#pragma once
#include "field.hpp"

#ifndef ZK_NOINL
#define ZK_NOINL __device__ __attribute__((noinline))
#endif

namespace zk {

struct Fq2 {
    Fp c0, c1;
};

ZK_DEV Fq2 add(const Fq2& a, const Fq2& b) { return {add<FpCfg>(a.c0, b.c0), add<FpCfg>(a.c1, b.c1)}; }
ZK_DEV Fq2 sub(const Fq2& a, const Fq2& b) { return {sub<FpCfg>(a.c0, b.c0), sub<FpCfg>(a.c1, b.c1)}; }
ZK_DEV Fq2 neg(const Fq2& a) { return {neg<FpCfg>(a.c0), neg<FpCfg>(a.c1)}; }
// Karatsuba: 3 Fp products (include/fp2_bn254.cuh:130-142 uses the same count).
// Inline: out-of-line calls in the G2 bucket loop cost more in spills than the
// larger code.
ZK_DEV Fq2 mul(const Fq2& a, const Fq2& b) {
    Fp t0 = mul<FpCfg>(a.c0, b.c0);
    Fp t1 = mul<FpCfg>(a.c1, b.c1);
    Fp t2 = mul<FpCfg>(add<FpCfg>(a.c0, a.c1), add<FpCfg>(b.c0, b.c1));
    return {sub<FpCfg>(t0, t1), sub<FpCfg>(sub<FpCfg>(t2, t0), t1)};
}
// (a0 + a1 u)^2 = (a0+a1)(a0-a1) + 2 a0 a1 u : 2 Fp products
ZK_DEV Fq2 sqr(const Fq2& a) {
    Fp t = mul<FpCfg>(a.c0, a.c1);
    return {mul<FpCfg>(add<FpCfg>(a.c0, a.c1), sub<FpCfg>(a.c0, a.c1)), add<FpCfg>(t, t)};
}
ZK_DEV bool is_zero(const Fq2& a) { return a.c0.is_zero() && a.c1.is_zero(); }
// inversions (Fermat; only used off the hot loops)
ZK_NOINL Fp inv_f(const Fp& a) { return inv<FpCfg>(a); }
ZK_NOINL Fq2 inv_f(const Fq2& a) {
    Fp n = inv<FpCfg>(add<FpCfg>(mul<FpCfg>(a.c0, a.c0), mul<FpCfg>(a.c1, a.c1)));
    return {mul<FpCfg>(a.c0, n), neg<FpCfg>(mul<FpCfg>(a.c1, n))};
}
ZK_DEV bool is_zero(const Fp& a) { return a.is_zero(); }
ZK_DEV Fp sqr(const Fp& a) { return mul<FpCfg>(a, a); }

template <class T> ZK_DEV T f_one();
template <> ZK_DEV Fp f_one<Fp>() { return Fp::one(); }
template <> ZK_DEV Fq2 f_one<Fq2>() { return {Fp::one(), Fp::zero()}; }
template <class T> ZK_DEV T f_zero();
template <> ZK_DEV Fp f_zero<Fp>() { return Fp::zero(); }
template <> ZK_DEV Fq2 f_zero<Fq2>() { return {Fp::zero(), Fp::zero()}; }

template <class T>
struct Aff {
    T x, y;
};

template <class T>
struct Xyzz {
    T X, Y, ZZ, ZZZ;
};

template <class T>
ZK_DEV Xyzz<T> xyzz_inf() {
    return {f_one<T>(), f_one<T>(), f_zero<T>(), f_zero<T>()};
}
template <class T>
ZK_DEV bool xyzz_is_inf(const Xyzz<T>& p) { return is_zero(p.ZZ); }

template <class T>
ZK_DEV bool aff_is_inf(const Aff<T>& a) { return is_zero(a.x) && is_zero(a.y); }

template <class T>
ZK_DEV Xyzz<T> xyzz_from_aff(const Aff<T>& a) {
    return {a.x, a.y, f_one<T>(), f_one<T>()};
}

// mdbl-2008-s-1: 2 * (x, y) for an affine point, a = 0. Inline: it is the
// rare branch of the mixed add, and an out-of-line call there would make the
// caller keep its point in memory on every iteration of the bucket loop.
template <class T>
ZK_DEV Xyzz<T> xyzz_mdbl(const Aff<T>& a) {
    T U = add(a.y, a.y);
    T V = sqr(U);
    T W = mul(U, V);
    T S = mul(a.x, V);
    T XX = sqr(a.x);
    T M = add(add(XX, XX), XX);
    T X3 = sub(sub(sqr(M), S), S);
    T Y3 = sub(mul(M, sub(S, X3)), mul(W, a.y));
    return {X3, Y3, V, W};
}

// dbl-2008-s-1, a = 0 (inline form; the rare equal-points branch of the
// additions uses it so that their operands never need an address)
template <class T>
ZK_DEV Xyzz<T> xyzz_dbl_i(const Xyzz<T>& p) {
    T U = add(p.Y, p.Y);
    T V = sqr(U);
    T W = mul(U, V);
    T S = mul(p.X, V);
    T XX = sqr(p.X);
    T M = add(add(XX, XX), XX);
    T X3 = sub(sub(sqr(M), S), S);
    T Y3 = sub(mul(M, sub(S, X3)), mul(W, p.Y));
    // doubling infinity keeps ZZ = 0
    return {X3, Y3, mul(V, p.ZZ), mul(W, p.ZZZ)};
}
// out-of-line form for code outside the hot loops (one copy per kernel)
template <class T>
ZK_NOINL Xyzz<T> xyzz_dbl(const Xyzz<T>& p) { return xyzz_dbl_i(p); }

// madd-2008-s: acc += a (a affine, not infinity)
template <class T>
ZK_DEV void xyzz_madd(Xyzz<T>& acc, const Aff<T>& a) {
    if (xyzz_is_inf(acc)) {
        acc = xyzz_from_aff(a);
        return;
    }
    T U2 = mul(a.x, acc.ZZ);
    T S2 = mul(a.y, acc.ZZZ);
    T P = sub(U2, acc.X);
    T R = sub(S2, acc.Y);
    if (is_zero(P)) {
        if (is_zero(R)) acc = xyzz_mdbl(a);
        else acc = xyzz_inf<T>();
        return;
    }
    T PP = sqr(P);
    T PPP = mul(P, PP);
    T Q = mul(acc.X, PP);
    T X3 = sub(sub(sub(sqr(R), PPP), Q), Q);
    T Y3 = sub(mul(R, sub(Q, X3)), mul(acc.Y, PPP));
    acc.X = X3;
    acc.Y = Y3;
    acc.ZZ = mul(acc.ZZ, PP);
    acc.ZZZ = mul(acc.ZZZ, PPP);
}

// add-2008-s: p + q, both XYZZ
template <class T>
ZK_DEV Xyzz<T> xyzz_add_i(const Xyzz<T>& p, const Xyzz<T>& q) {
    if (xyzz_is_inf(p)) return q;
    if (xyzz_is_inf(q)) return p;
    T U1 = mul(p.X, q.ZZ);
    T U2 = mul(q.X, p.ZZ);
    T S1 = mul(p.Y, q.ZZZ);
    T S2 = mul(q.Y, p.ZZZ);
    T P = sub(U2, U1);
    T R = sub(S2, S1);
    if (is_zero(P)) {
        if (is_zero(R)) return xyzz_dbl_i(p);
        return xyzz_inf<T>();
    }
    T PP = sqr(P);
    T PPP = mul(P, PP);
    T Q = mul(U1, PP);
    T X3 = sub(sub(sub(sqr(R), PPP), Q), Q);
    T Y3 = sub(mul(R, sub(Q, X3)), mul(S1, PPP));
    return {X3, Y3, mul(mul(p.ZZ, q.ZZ), PP), mul(mul(p.ZZZ, q.ZZZ), PPP)};
}

// out-of-line form for code outside the hot loops (one copy per kernel)
template <class T>
ZK_NOINL Xyzz<T> xyzz_add(const Xyzz<T>& p, const Xyzz<T>& q) { return xyzz_add_i(p, q); }

// ---- global memory access ----
// G1 affine point: 64 B = x (32 B) then y (32 B). G2: 128 B = x.c0, x.c1, y.c0, y.c1.
template <class T> struct PointIO;

template <>
struct PointIO<Fp> {
    static constexpr int AFF_BYTES = 64;
    static constexpr int XYZZ_BYTES = 128;
    ZK_DEV static Aff<Fp> load_aff(const void* base, size_t i) {
        return {load<FpCfg>(base, 2 * i), load<FpCfg>(base, 2 * i + 1)};
    }
    ZK_DEV static Xyzz<Fp> load_xyzz(const void* base, size_t i) {
        return {load<FpCfg>(base, 4 * i), load<FpCfg>(base, 4 * i + 1), load<FpCfg>(base, 4 * i + 2),
                load<FpCfg>(base, 4 * i + 3)};
    }
    ZK_DEV static void store_xyzz(void* base, size_t i, const Xyzz<Fp>& p) {
        store<FpCfg>(base, 4 * i, p.X);
        store<FpCfg>(base, 4 * i + 1, p.Y);
        store<FpCfg>(base, 4 * i + 2, p.ZZ);
        store<FpCfg>(base, 4 * i + 3, p.ZZZ);
    }
    ZK_DEV static Fp neg_y(const Fp& y) { return neg<FpCfg>(y); }
};

template <>
struct PointIO<Fq2> {
    static constexpr int AFF_BYTES = 128;
    static constexpr int XYZZ_BYTES = 256;
    ZK_DEV static Fq2 ld2(const void* base, size_t i) { return {load<FpCfg>(base, 2 * i), load<FpCfg>(base, 2 * i + 1)}; }
    ZK_DEV static void st2(void* base, size_t i, const Fq2& a) {
        store<FpCfg>(base, 2 * i, a.c0);
        store<FpCfg>(base, 2 * i + 1, a.c1);
    }
    ZK_DEV static Aff<Fq2> load_aff(const void* base, size_t i) { return {ld2(base, 2 * i), ld2(base, 2 * i + 1)}; }
    ZK_DEV static Xyzz<Fq2> load_xyzz(const void* base, size_t i) {
        return {ld2(base, 4 * i), ld2(base, 4 * i + 1), ld2(base, 4 * i + 2), ld2(base, 4 * i + 3)};
    }
    ZK_DEV static void store_xyzz(void* base, size_t i, const Xyzz<Fq2>& p) {
        st2(base, 4 * i, p.X);
        st2(base, 4 * i + 1, p.Y);
        st2(base, 4 * i + 2, p.ZZ);
        st2(base, 4 * i + 3, p.ZZZ);
    }
    ZK_DEV static Fq2 neg_y(const Fq2& y) { return neg(y); }
};

}  // namespace zk
