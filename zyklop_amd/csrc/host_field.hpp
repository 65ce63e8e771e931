// Host-side BN254 arithmetic: Montgomery Fp/Fr on 4 x u64 limbs (R = 2^256),
// Fp2 = Fp[u]/(u^2+1), and XYZZ / affine curve arithmetic for G1 and G2.
//
// The device does the bulk work; the host uses this for the short serial
// tails of a proof (window combination, blinding terms, affine conversion),
// for proof serialization and for the synthetic-circuit generator.
// Values are bit-compatible with the device layout (8 x u32 == 4 x u64).
#pragma once
#include <stdint.h>
#include <string.h>

#include <string>

namespace zkh {

typedef unsigned __int128 u128;

struct FpP {
    static constexpr uint64_t M[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
    static constexpr uint64_t NP = 0x87d20782e4866389ULL;
    static constexpr uint64_t ONE[4] = {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL};
    static constexpr uint64_t R2[4] = {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL};
};
struct FrP {
    static constexpr uint64_t M[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
    static constexpr uint64_t NP = 0xc2e1f593efffffffULL;
    static constexpr uint64_t ONE[4] = {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL};
    static constexpr uint64_t R2[4] = {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL};
};

template <class P>
struct F {
    static_assert(P::M[3] < (1ULL << 62), "the no-carry product needs a modulus below 2^254");
    uint64_t v[4];

    static F zero() { F r; memset(r.v, 0, 32); return r; }
    static F one() { F r; memcpy(r.v, P::ONE, 32); return r; }
    static F from_raw(const uint64_t* p) { F r; memcpy(r.v, p, 32); return r; }
    bool is_zero() const { return (v[0] | v[1] | v[2] | v[3]) == 0; }
    bool operator==(const F& o) const { return memcmp(v, o.v, 32) == 0; }
    bool operator!=(const F& o) const { return !(*this == o); }

    static bool geq_mod(const uint64_t* a) {
        for (int i = 3; i >= 0; i--) {
            if (a[i] != P::M[i]) return a[i] > P::M[i];
        }
        return true;
    }
    friend F operator+(const F& a, const F& b) {
        F r;
        u128 c = 0;
        for (int i = 0; i < 4; i++) { c += (u128)a.v[i] + b.v[i]; r.v[i] = (uint64_t)c; c >>= 64; }
        if (c || geq_mod(r.v)) {
            u128 br = 0;
            for (int i = 0; i < 4; i++) {
                u128 d = (u128)r.v[i] - P::M[i] - br;
                r.v[i] = (uint64_t)d; br = (d >> 64) ? 1 : 0;
            }
        }
        return r;
    }
    friend F operator-(const F& a, const F& b) {
        F r;
        u128 br = 0;
        for (int i = 0; i < 4; i++) {
            u128 d = (u128)a.v[i] - b.v[i] - br;
            r.v[i] = (uint64_t)d; br = (d >> 64) ? 1 : 0;
        }
        if (br) {
            u128 c = 0;
            for (int i = 0; i < 4; i++) { c += (u128)r.v[i] + P::M[i]; r.v[i] = (uint64_t)c; c >>= 64; }
        }
        return r;
    }
    F operator-() const { return zero() - *this; }
    // Montgomery CIOS product a*b*2^-256. Both BN254 moduli have a top word
    // below 2^62, so the running value stays below 2m after every outer step
    // and fits four words: no fifth carry word is kept ("no-carry" CIOS).
    friend F operator*(const F& a, const F& b) {
        uint64_t t0 = 0, t1 = 0, t2 = 0, t3 = 0;
        for (int i = 0; i < 4; i++) {
            const uint64_t bi = b.v[i];
            u128 x = (u128)a.v[0] * bi + t0;
            uint64_t A = (uint64_t)(x >> 64);
            t0 = (uint64_t)x;
            const uint64_t m = t0 * P::NP;
            u128 y = (u128)m * P::M[0] + t0;
            uint64_t C = (uint64_t)(y >> 64);
            x = (u128)a.v[1] * bi + t1 + A; A = (uint64_t)(x >> 64); t1 = (uint64_t)x;
            y = (u128)m * P::M[1] + t1 + C; C = (uint64_t)(y >> 64); t0 = (uint64_t)y;
            x = (u128)a.v[2] * bi + t2 + A; A = (uint64_t)(x >> 64); t2 = (uint64_t)x;
            y = (u128)m * P::M[2] + t2 + C; C = (uint64_t)(y >> 64); t1 = (uint64_t)y;
            x = (u128)a.v[3] * bi + t3 + A; A = (uint64_t)(x >> 64); t3 = (uint64_t)x;
            y = (u128)m * P::M[3] + t3 + C; C = (uint64_t)(y >> 64); t2 = (uint64_t)y;
            t3 = C + A;
        }
        F r;
        r.v[0] = t0; r.v[1] = t1; r.v[2] = t2; r.v[3] = t3;
        if (geq_mod(r.v)) {
            u128 br = 0;
            for (int i = 0; i < 4; i++) {
                u128 d = (u128)r.v[i] - P::M[i] - br;
                r.v[i] = (uint64_t)d; br = (d >> 64) ? 1 : 0;
            }
        }
        return r;
    }
    F sqr() const { return *this * *this; }
    static F to_mont(const F& a) { F r2; memcpy(r2.v, P::R2, 32); return a * r2; }
    static F from_mont(const F& a) { F o = zero(); o.v[0] = 1; return a * o; }
    static F from_u64(uint64_t x) { F a = zero(); a.v[0] = x; return to_mont(a); }
    F pow(const uint64_t e[4]) const {
        F r = one();
        for (int w = 3; w >= 0; w--)
            for (int b = 63; b >= 0; b--) {
                r = r.sqr();
                if ((e[w] >> b) & 1) r = r * *this;
            }
        return r;
    }
    F inv() const {  // Fermat a^(m-2); inv(0) = 0
        uint64_t e[4];
        memcpy(e, P::M, 32);
        e[0] -= 2;  // m is odd and > 2: no borrow
        return pow(e);
    }
    // canonical (non-Montgomery) value of a Montgomery element
    void canonical(uint64_t out[4]) const { F c = from_mont(*this); memcpy(out, c.v, 32); }
    static bool is_canonical(const uint64_t* a) { return !geq_mod(a); }
};

using Fp = F<FpP>;
using Fr = F<FrP>;

struct Fp2 {
    Fp c0, c1;
    static Fp2 zero() { return {Fp::zero(), Fp::zero()}; }
    static Fp2 one() { return {Fp::one(), Fp::zero()}; }
    bool is_zero() const { return c0.is_zero() && c1.is_zero(); }
    bool operator==(const Fp2& o) const { return c0 == o.c0 && c1 == o.c1; }
    bool operator!=(const Fp2& o) const { return !(*this == o); }
    friend Fp2 operator+(const Fp2& a, const Fp2& b) { return {a.c0 + b.c0, a.c1 + b.c1}; }
    friend Fp2 operator-(const Fp2& a, const Fp2& b) { return {a.c0 - b.c0, a.c1 - b.c1}; }
    Fp2 operator-() const { return {-c0, -c1}; }
    friend Fp2 operator*(const Fp2& a, const Fp2& b) {
        Fp t0 = a.c0 * b.c0, t1 = a.c1 * b.c1;
        Fp t2 = (a.c0 + a.c1) * (b.c0 + b.c1);
        return {t0 - t1, t2 - t0 - t1};
    }
    Fp2 sqr() const { return *this * *this; }
    Fp2 inv() const {
        Fp n = (c0.sqr() + c1.sqr()).inv();
        return {c0 * n, -(c1 * n)};
    }
};

inline Fp fsqr(const Fp& a) { return a.sqr(); }
inline Fp2 fsqr(const Fp2& a) { return a.sqr(); }
inline Fp finv(const Fp& a) { return a.inv(); }
inline Fp2 finv(const Fp2& a) { return a.inv(); }
template <class T> inline T fone();
template <> inline Fp fone<Fp>() { return Fp::one(); }
template <> inline Fp2 fone<Fp2>() { return Fp2::one(); }
template <class T> inline T fzero();
template <> inline Fp fzero<Fp>() { return Fp::zero(); }
template <> inline Fp2 fzero<Fp2>() { return Fp2::zero(); }

// Affine point; infinity flagged explicitly (stored as all-zero coordinates).
template <class T>
struct Aff {
    T x, y;
    bool inf;
};

// XYZZ point: x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2; infinity <=> ZZ = 0.
template <class T>
struct Xyzz {
    T X, Y, ZZ, ZZZ;
    static Xyzz infinity() { return {fone<T>(), fone<T>(), fzero<T>(), fzero<T>()}; }
    bool is_inf() const { return ZZ.is_zero(); }
    static Xyzz from_aff(const Aff<T>& a) {
        if (a.inf) return infinity();
        return {a.x, a.y, fone<T>(), fone<T>()};
    }
};

// dbl-2008-s-1 (a = 0)
template <class T>
Xyzz<T> xyzz_dbl(const Xyzz<T>& p) {
    if (p.is_inf()) return p;
    T U = p.Y + p.Y, V = fsqr(U), W = U * V, S = p.X * V;
    T XX = fsqr(p.X), M = XX + XX + XX;
    T X3 = fsqr(M) - S - S;
    T Y3 = M * (S - X3) - W * p.Y;
    return {X3, Y3, V * p.ZZ, W * p.ZZZ};
}

// add-2008-s with the P == Q and P == -Q cases
template <class T>
Xyzz<T> xyzz_add(const Xyzz<T>& p, const Xyzz<T>& q) {
    if (p.is_inf()) return q;
    if (q.is_inf()) return p;
    T U1 = p.X * q.ZZ, U2 = q.X * p.ZZ, S1 = p.Y * q.ZZZ, S2 = q.Y * p.ZZZ;
    T P = U2 - U1, R = S2 - S1;
    if (P.is_zero()) {
        if (R.is_zero()) return xyzz_dbl(p);
        return Xyzz<T>::infinity();
    }
    T PP = fsqr(P), PPP = P * PP, Q = U1 * PP;
    T X3 = fsqr(R) - PPP - Q - Q;
    T Y3 = R * (Q - X3) - S1 * PPP;
    return {X3, Y3, p.ZZ * q.ZZ * PP, p.ZZZ * q.ZZZ * PPP};
}

template <class T>
Xyzz<T> xyzz_neg(const Xyzz<T>& p) { return {p.X, -p.Y, p.ZZ, p.ZZZ}; }

template <class T>
Aff<T> xyzz_to_aff(const Xyzz<T>& p) {
    if (p.is_inf()) return {fzero<T>(), fzero<T>(), true};
    // one inversion: 1/(ZZ*ZZZ)
    T t = finv(p.ZZ * p.ZZZ);
    T izz = t * p.ZZZ, izzz = t * p.ZZ;
    return {p.X * izz, p.Y * izzz, false};
}

// scalar multiplication by a canonical 256-bit scalar (4-bit fixed window)
template <class T>
Xyzz<T> xyzz_mul(const Xyzz<T>& p, const uint64_t k[4]) {
    Xyzz<T> tab[16];
    tab[0] = Xyzz<T>::infinity();
    for (int i = 1; i < 16; i++) tab[i] = xyzz_add(tab[i - 1], p);
    Xyzz<T> acc = Xyzz<T>::infinity();
    for (int nib = 63; nib >= 0; nib--) {
        for (int d = 0; d < 4; d++) acc = xyzz_dbl(acc);
        unsigned v = (unsigned)((k[nib >> 4] >> (4 * (nib & 15))) & 15);
        if (v) acc = xyzz_add(acc, tab[v]);
    }
    return acc;
}

// multiply by a small unsigned integer (double-and-add)
template <class T>
Xyzz<T> xyzz_mul_small(const Xyzz<T>& p, uint64_t k) {
    Xyzz<T> acc = Xyzz<T>::infinity();
    for (int b = 63; b >= 0; b--) {
        acc = xyzz_dbl(acc);
        if ((k >> b) & 1) acc = xyzz_add(acc, p);
    }
    return acc;
}

// canonical 256-bit (4 x u64 LE) -> decimal string
inline std::string to_decimal(const uint64_t a[4]) {
    uint64_t n[4];
    memcpy(n, a, 32);
    char tmp[96];
    int len = 0;
    do {
        uint64_t rem = 0;
        for (int i = 3; i >= 0; i--) {
            u128 d = ((u128)rem << 64) | n[i];
            n[i] = (uint64_t)(d / 10);
            rem = (uint64_t)(d % 10);
        }
        tmp[len++] = (char)('0' + rem);
    } while (n[0] | n[1] | n[2] | n[3]);
    std::string s(len, '0');
    for (int i = 0; i < len; i++) s[i] = tmp[len - 1 - i];
    return s;
}

}  // namespace zkh
