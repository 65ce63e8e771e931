// Optimal ate pairing on BN254 and the Groth16 verifier (host). Construction
// and conventions are described in pairing.hpp.
#include "pairing.hpp"

#include <stdexcept>

namespace zkh {

namespace {

// ---------------------------------------------------------------- big integers
// Unsigned integers of up to 1536 bits, only for deriving the exponents of the
// pairing from p and r at start-up.
constexpr int BW = 24;
struct BigU {
    uint64_t w[BW] = {};
};

BigU big_from(const uint64_t* v, int n) {
    BigU r;
    for (int i = 0; i < n; i++) r.w[i] = v[i];
    return r;
}

BigU big_mul(const BigU& a, const BigU& b) {
    BigU r;
    for (int i = 0; i < BW; i++) {
        if (!a.w[i]) continue;
        u128 c = 0;
        for (int j = 0; i + j < BW; j++) {
            c += (u128)a.w[i] * b.w[j] + r.w[i + j];
            r.w[i + j] = (uint64_t)c;
            c >>= 64;
        }
        if (c) throw std::overflow_error("pairing exponent overflow");
    }
    return r;
}

BigU big_add_u64(BigU a, uint64_t x) {
    u128 c = x;
    for (int i = 0; i < BW && c; i++) {
        c += a.w[i];
        a.w[i] = (uint64_t)c;
        c >>= 64;
    }
    return a;
}

BigU big_sub_u64(BigU a, uint64_t x) {
    uint64_t br = x;
    for (int i = 0; i < BW && br; i++) {
        uint64_t t = a.w[i];
        a.w[i] = t - br;
        br = t < br ? 1 : 0;
    }
    return a;
}

int big_cmp(const BigU& a, const BigU& b) {
    for (int i = BW - 1; i >= 0; i--)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
}

void big_sub(BigU& a, const BigU& b) {
    uint64_t br = 0;
    for (int i = 0; i < BW; i++) {
        u128 d = (u128)a.w[i] - b.w[i] - br;
        a.w[i] = (uint64_t)d;
        br = (d >> 64) ? 1 : 0;
    }
}

int big_bits(const BigU& a) {
    for (int i = BW - 1; i >= 0; i--)
        if (a.w[i]) return 64 * i + 64 - __builtin_clzll(a.w[i]);
    return 0;
}

bool big_bit(const BigU& a, int i) { return (a.w[i >> 6] >> (i & 63)) & 1; }

// a / d by shift and subtract; throws unless the division is exact
BigU big_div_exact(const BigU& a, const BigU& d) {
    BigU q, rem;
    for (int i = big_bits(a) - 1; i >= 0; i--) {
        for (int k = BW - 1; k > 0; k--) rem.w[k] = (rem.w[k] << 1) | (rem.w[k - 1] >> 63);
        rem.w[0] = (rem.w[0] << 1) | (big_bit(a, i) ? 1 : 0);
        if (big_cmp(rem, d) >= 0) {
            big_sub(rem, d);
            q.w[i >> 6] |= 1ULL << (i & 63);
        }
    }
    for (uint64_t x : rem.w)
        if (x) throw std::logic_error("pairing exponent is not an exact quotient");
    return q;
}

// ---------------------------------------------------------------- Fp2 / Fp6 / Fp12
Fp2 f2_scale(const Fp2& a, const Fp& s) { return {a.c0 * s, a.c1 * s}; }
Fp2 f2_conj(const Fp2& a) { return {a.c0, -a.c1}; }
// a * xi, xi = 9 + u: (9 a0 - a1) + (a0 + 9 a1) u
Fp2 f2_mul_xi(const Fp2& a) {
    Fp t0 = a.c0 + a.c0, t1 = a.c1 + a.c1;
    t0 = t0 + t0;
    t1 = t1 + t1;
    t0 = t0 + t0 + a.c0;
    t1 = t1 + t1 + a.c1;
    return {t0 - a.c1, a.c0 + t1};
}
Fp2 f2_pow(const Fp2& a, const BigU& e) {
    Fp2 r = Fp2::one();
    for (int i = big_bits(e) - 1; i >= 0; i--) {
        r = r.sqr();
        if (big_bit(e, i)) r = r * a;
    }
    return r;
}

Fp6 f6_zero() { return {Fp2::zero(), Fp2::zero(), Fp2::zero()}; }
Fp6 f6_add(const Fp6& a, const Fp6& b) { return {a.c0 + b.c0, a.c1 + b.c1, a.c2 + b.c2}; }
Fp6 f6_sub(const Fp6& a, const Fp6& b) { return {a.c0 - b.c0, a.c1 - b.c1, a.c2 - b.c2}; }
Fp6 f6_neg(const Fp6& a) { return {-a.c0, -a.c1, -a.c2}; }
// schoolbook product with v^3 = xi
Fp6 f6_mul(const Fp6& a, const Fp6& b) {
    Fp2 c0 = a.c0 * b.c0 + f2_mul_xi(a.c1 * b.c2 + a.c2 * b.c1);
    Fp2 c1 = a.c0 * b.c1 + a.c1 * b.c0 + f2_mul_xi(a.c2 * b.c2);
    Fp2 c2 = a.c0 * b.c2 + a.c1 * b.c1 + a.c2 * b.c0;
    return {c0, c1, c2};
}
Fp6 f6_mul_v(const Fp6& a) { return {f2_mul_xi(a.c2), a.c0, a.c1}; }
// (c0 + c1 v + c2 v^2)^-1 = (t0 + t1 v + t2 v^2) / (c0 t0 + xi (c2 t1 + c1 t2))
Fp6 f6_inv(const Fp6& c) {
    Fp2 t0 = c.c0.sqr() - f2_mul_xi(c.c1 * c.c2);
    Fp2 t1 = f2_mul_xi(c.c2.sqr()) - c.c0 * c.c1;
    Fp2 t2 = c.c1.sqr() - c.c0 * c.c2;
    Fp2 di = (c.c0 * t0 + f2_mul_xi(c.c2 * t1 + c.c1 * t2)).inv();
    return {t0 * di, t1 * di, t2 * di};
}

// conjugation over Fp6 (w -> -w) is the p^6-power Frobenius
Fp12 f12_conj(const Fp12& a) { return {a.c0, f6_neg(a.c1)}; }
Fp12 f12_inv(const Fp12& a) {
    Fp6 t = f6_sub(f6_mul(a.c0, a.c0), f6_mul_v(f6_mul(a.c1, a.c1)));
    Fp6 ti = f6_inv(t);
    return {f6_mul(a.c0, ti), f6_neg(f6_mul(a.c1, ti))};
}
Fp12 f12_pow(const Fp12& a, const BigU& e) {
    Fp12 r = fp12_one();
    for (int i = big_bits(e) - 1; i >= 0; i--) {
        r = fp12_mul(r, r);
        if (big_bit(e, i)) r = fp12_mul(r, a);
    }
    return r;
}

// ---------------------------------------------------------------- constants
constexpr uint64_t BN_X = 4965661367192848881ULL;  // curve parameter; the loop runs over 6x + 2

struct PairConsts {
    Fp2 frob1_x, frob1_y;  // xi^((p-1)/3), xi^((p-1)/2)
    Fp2 frob2_x, frob2_y;  // xi^((p^2-1)/3), xi^((p^2-1)/2)
    Fp2 b_twist;           // 3 / xi
    BigU e_final;          // (p^6 + 1) / r
};

PairConsts make_consts() {
    PairConsts k;
    const BigU p = big_from(FpP::M, 4), r = big_from(FrP::M, 4);
    const BigU p2 = big_mul(p, p);
    const uint64_t w3[1] = {3}, w2[1] = {2};
    const BigU three = big_from(w3, 1), two = big_from(w2, 1);
    const Fp2 xi = {Fp::from_u64(9), Fp::from_u64(1)};
    k.frob1_x = f2_pow(xi, big_div_exact(big_sub_u64(p, 1), three));
    k.frob1_y = f2_pow(xi, big_div_exact(big_sub_u64(p, 1), two));
    k.frob2_x = f2_pow(xi, big_div_exact(big_sub_u64(p2, 1), three));
    k.frob2_y = f2_pow(xi, big_div_exact(big_sub_u64(p2, 1), two));
    k.b_twist = Fp2{Fp::from_u64(3), Fp::zero()} * xi.inv();
    k.e_final = big_div_exact(big_add_u64(big_mul(big_mul(p2, p2), p2), 1), r);
    return k;
}

const PairConsts& consts() {
    static const PairConsts k = make_consts();
    return k;
}

// ---------------------------------------------------------------- Miller loop
// Line through T and Q (or the tangent at T) on the twist, slope lam, at P:
// psi maps the line to Y - lam w X - (y_T - lam x_T) w^3, so its value at P is
// y_P - lam x_P w + (lam x_T - y_T) v w (w^2 = v).
Fp12 line_at(const Fp2& lam, const Fp2& xT, const Fp2& yT, const Aff<Fp>& P) {
    Fp12 l = {f6_zero(), f6_zero()};
    l.c0.c0 = {P.y, Fp::zero()};
    l.c1.c0 = -f2_scale(lam, P.x);
    l.c1.c1 = lam * xT - yT;
    return l;
}

struct Twist {
    Fp2 x, y;
};

// f <- f * line(T, Q)(P); T <- T + Q. False if T = -Q (a vertical line, which
// no valid input reaches).
bool add_step(Fp12& f, Twist& T, const Twist& Q, const Aff<Fp>& P) {
    const Fp2 dx = Q.x - T.x;
    if (dx.is_zero()) return false;
    const Fp2 lam = (Q.y - T.y) * dx.inv();
    f = fp12_mul(f, line_at(lam, T.x, T.y, P));
    const Fp2 x3 = lam.sqr() - T.x - Q.x;
    T.y = lam * (T.x - x3) - T.y;
    T.x = x3;
    return true;
}

bool dbl_step(Fp12& f, Twist& T, const Aff<Fp>& P) {
    if (T.y.is_zero()) return false;
    const Fp2 xx = T.x.sqr();
    const Fp2 lam = (xx + xx + xx) * (T.y + T.y).inv();
    f = fp12_mul(fp12_mul(f, f), line_at(lam, T.x, T.y, P));
    const Fp2 x3 = lam.sqr() - T.x - T.x;
    T.y = lam * (T.x - x3) - T.y;
    T.x = x3;
    return true;
}

// f_{6x+2,Q}(P) * l_{[6x+2]Q, pi(Q)}(P) * l_{[6x+2]Q + pi(Q), -pi^2(Q)}(P)
bool miller(const Aff<Fp>& P, const Aff<Fp2>& Qa, Fp12& f) {
    f = fp12_one();
    if (P.inf || Qa.inf) return true;
    const PairConsts& k = consts();
    const u128 s = (u128)6 * BN_X + 2;
    int top = 127;
    while (!((s >> top) & 1)) top--;
    const Twist Q = {Qa.x, Qa.y};
    Twist T = Q;
    for (int i = top - 1; i >= 0; i--) {
        if (!dbl_step(f, T, P)) return false;
        if (((s >> i) & 1) && !add_step(f, T, Q, P)) return false;
    }
    const Twist Q1 = {f2_conj(Q.x) * k.frob1_x, f2_conj(Q.y) * k.frob1_y};
    const Twist Q2n = {Q.x * k.frob2_x, -(Q.y * k.frob2_y)};
    if (!add_step(f, T, Q1, P)) return false;
    return add_step(f, T, Q2n, P);
}

Fp12 final_exp(const Fp12& f) { return f12_pow(fp12_mul(f12_conj(f), f12_inv(f)), consts().e_final); }

}  // namespace

// ---------------------------------------------------------------- public
Fp12 fp12_one() {
    Fp12 r = {f6_zero(), f6_zero()};
    r.c0.c0 = Fp2::one();
    return r;
}

bool fp12_eq(const Fp12& a, const Fp12& b) {
    return a.c0.c0 == b.c0.c0 && a.c0.c1 == b.c0.c1 && a.c0.c2 == b.c0.c2 && a.c1.c0 == b.c1.c0 &&
           a.c1.c1 == b.c1.c1 && a.c1.c2 == b.c1.c2;
}

Fp12 fp12_mul(const Fp12& a, const Fp12& b) {
    const Fp6 t0 = f6_mul(a.c0, b.c0), t1 = f6_mul(a.c1, b.c1);
    return {f6_add(t0, f6_mul_v(t1)), f6_add(f6_mul(a.c0, b.c1), f6_mul(a.c1, b.c0))};
}

bool g1_on_curve(const Aff<Fp>& p) {
    if (p.inf) return true;
    return p.y.sqr() == p.x.sqr() * p.x + Fp::from_u64(3);
}

bool g2_on_curve(const Aff<Fp2>& q) {
    if (q.inf) return true;
    return q.y.sqr() == q.x.sqr() * q.x + consts().b_twist;
}

bool g2_in_subgroup(const Aff<Fp2>& q) {
    if (q.inf) return true;
    return xyzz_mul(Xyzz<Fp2>::from_aff(q), FrP::M).is_inf();
}

Fp12 pairing(const Aff<Fp>& p, const Aff<Fp2>& q) {
    Fp12 f;
    if (!miller(p, q, f)) throw std::invalid_argument("pairing input reaches a vertical line");
    return final_exp(f);
}

bool pairing_product_is_one(const Aff<Fp>* p, const Aff<Fp2>* q, int n) {
    Fp12 acc = fp12_one();
    for (int i = 0; i < n; i++) {
        Fp12 f;
        if (!miller(p[i], q[i], f)) return false;
        acc = fp12_mul(acc, f);
    }
    return fp12_eq(final_exp(acc), fp12_one());
}

Aff<Fp> g1_from_words(const uint64_t w[8]) {
    Aff<Fp> a = {Fp::from_raw(w), Fp::from_raw(w + 4), false};
    a.inf = a.x.is_zero() && a.y.is_zero();
    return a;
}

Aff<Fp2> g2_from_words(const uint64_t w[16]) {
    Aff<Fp2> a = {{Fp::from_raw(w), Fp::from_raw(w + 4)}, {Fp::from_raw(w + 8), Fp::from_raw(w + 12)}, false};
    a.inf = a.x.is_zero() && a.y.is_zero();
    return a;
}

Xyzz<Fp> g1_msm_host(const Aff<Fp>* pts, const uint64_t* k, size_t n) {
    int c = 4;
    while (c < 12 && ((size_t)1 << (c + 2)) < n) c++;
    const int nw = (256 + c - 1) / c;
    const size_t nb = ((size_t)1 << c) - 1;
    std::vector<Xyzz<Fp>> bk(nb);
    Xyzz<Fp> total = Xyzz<Fp>::infinity();
    for (int w = nw - 1; w >= 0; w--) {
        for (int d = 0; d < c; d++) total = xyzz_dbl(total);
        for (auto& b : bk) b = Xyzz<Fp>::infinity();
        for (size_t i = 0; i < n; i++) {
            if (pts[i].inf) continue;
            uint64_t d = 0;
            for (int q = 0; q < c; q++) {
                const int bit = w * c + q;
                if (bit < 256 && ((k[4 * i + (bit >> 6)] >> (bit & 63)) & 1)) d |= 1ULL << q;
            }
            if (d) bk[d - 1] = xyzz_add(bk[d - 1], Xyzz<Fp>::from_aff(pts[i]));
        }
        Xyzz<Fp> run = Xyzz<Fp>::infinity(), sum = Xyzz<Fp>::infinity();
        for (size_t b = nb; b-- > 0;) {
            run = xyzz_add(run, bk[b]);
            sum = xyzz_add(sum, run);
        }
        total = xyzz_add(total, sum);
    }
    return total;
}

bool groth16_verify(const uint64_t alpha1[8], const uint64_t beta2[16], const uint64_t gamma2[16],
                    const uint64_t delta2[16], const uint64_t* ic, uint32_t n_public, const uint64_t* publics,
                    const uint64_t pi_a[8], const uint64_t pi_b[16], const uint64_t pi_c[8]) {
    const Aff<Fp> A = g1_from_words(pi_a), C = g1_from_words(pi_c);
    const Aff<Fp2> B = g2_from_words(pi_b);
    if (!g1_on_curve(A) || !g1_on_curve(C) || !g2_on_curve(B) || !g2_in_subgroup(B)) return false;
    std::vector<Aff<Fp>> icp(n_public);
    for (uint32_t i = 0; i < n_public; i++) icp[i] = g1_from_words(ic + 8ull * (i + 1));
    const Xyzz<Fp> vk = xyzz_add(Xyzz<Fp>::from_aff(g1_from_words(ic)), g1_msm_host(icp.data(), publics, n_public));
    Aff<Fp> P[4] = {A, g1_from_words(alpha1), xyzz_to_aff(vk), C};
    P[0].y = -P[0].y;  // e(-A, B) e(alpha1, beta2) e(vk_x, gamma2) e(C, delta2) == 1
    const Aff<Fp2> Q[4] = {B, g2_from_words(beta2), g2_from_words(gamma2), g2_from_words(delta2)};
    return pairing_product_is_one(P, Q, 4);
}

}  // namespace zkh
