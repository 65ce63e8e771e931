#include "synth.hpp"

#include <string.h>

namespace zk {

using zkh::Fr;

namespace {
constexpr uint64_t TAG_R1CS = 0x5231435352314353ULL;
constexpr uint64_t TAG_PUB = 0x5055424c49435055ULL;
constexpr uint64_t TAG_TOXIC = 0x544f584943574153ULL;
constexpr uint64_t TAG_BLIND = 0x424c494e44524e47ULL;

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) {
            uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // canonical value < r
    void fr_canon(uint64_t v[4]) {
        for (;;) {
            for (int i = 0; i < 4; i++) v[i] = next();
            v[3] &= 0x3fffffffffffffffULL;
            if (Fr::is_canonical(v)) return;
        }
    }
    Fr fr() {
        uint64_t v[4];
        fr_canon(v);
        return Fr::to_mont(Fr::from_raw(v));
    }
    uint32_t below(uint32_t b) { return (uint32_t)(((unsigned __int128)next() * b) >> 64); }
};

// L_k(tau) = (tau^N - 1)/N * w^k / (tau - w^k) for the k in idx (size-N domain)
void lagrange(const Fr& tau, int log_N, const std::vector<uint64_t>& idx, std::vector<Fr>& out) {
    uint64_t N = 1ull << log_N;
    // w_N = 5^((r-1) >> log_N)
    uint64_t e[4];
    memcpy(e, zkh::FrP::M, 32);
    e[0] -= 1;
    for (int s = 0; s < log_N; s++) {
        for (int k = 0; k < 3; k++) e[k] = (e[k] >> 1) | (e[k + 1] << 63);
        e[3] >>= 1;
    }
    Fr w = Fr::from_u64(5).pow(e);
    Fr tn = tau;
    for (int i = 0; i < log_N; i++) tn = tn.sqr();
    Fr zf = (tn - Fr::one()) * Fr::from_u64(N).inv();
    size_t cnt = idx.size();
    // w^k via a 2-level table
    std::vector<Fr> lo(4096), hi;
    lo[0] = Fr::one();
    for (int i = 1; i < 4096; i++) lo[i] = lo[i - 1] * w;
    size_t nh = (N >> 12) + 1;
    hi.resize(nh);
    hi[0] = Fr::one();
    Fr w4096 = lo[4095] * w;
    for (size_t j = 1; j < nh; j++) hi[j] = hi[j - 1] * w4096;
    std::vector<Fr> wk(cnt), den(cnt), pre(cnt);
    Fr acc = Fr::one();
    for (size_t i = 0; i < cnt; i++) {
        uint64_t k = idx[i];
        wk[i] = lo[k & 4095] * hi[k >> 12];
        den[i] = tau - wk[i];
        pre[i] = acc;
        acc = acc * den[i];
    }
    Fr inv = acc.inv();
    out.resize(cnt);
    for (size_t i = cnt; i-- > 0;) {
        Fr di = inv * pre[i];
        inv = inv * den[i];
        out[i] = zf * wk[i] * di;
    }
}

void put_canon(const Fr& x, uint64_t* dst) { x.canonical(dst); }
}  // namespace

void synth_r1cs(uint64_t seed, uint32_t m, uint32_t n_pub, SynthR1cs& c) {
    c.seed = seed;
    c.m = m;
    c.l = n_pub;
    c.n_vars = 1 + n_pub + m;
    uint32_t need = m + n_pub + 1;
    c.n = 1;
    c.log_n = 0;
    while (c.n < need) { c.n <<= 1; c.log_n++; }
    c.a_wire.resize(3ull * m); c.b_wire.resize(3ull * m);
    c.a_coef.resize(3ull * m); c.b_coef.resize(3ull * m);
    Xoshiro g(seed ^ TAG_R1CS);
    for (uint32_t i = 0; i < m; i++) {
        uint32_t bound = 1 + n_pub + i;
        for (int t = 0; t < 3; t++) { c.a_wire[3ull * i + t] = g.below(bound); c.a_coef[3ull * i + t] = g.fr(); }
        for (int t = 0; t < 3; t++) { c.b_wire[3ull * i + t] = g.below(bound); c.b_coef[3ull * i + t] = g.fr(); }
    }
}

std::vector<Fr> synth_public(uint64_t seed, uint32_t n_pub) {
    Xoshiro g(seed ^ TAG_PUB);
    std::vector<Fr> p(n_pub);
    for (uint32_t k = 0; k < n_pub; k++) p[k] = g.fr();
    return p;
}

void synth_witness_from_public(const SynthR1cs& c, const std::vector<Fr>& pub, std::vector<uint64_t>& out) {
    std::vector<Fr> w(c.n_vars);
    w[0] = Fr::one();
    for (uint32_t k = 0; k < c.l; k++) w[1 + k] = pub[k];
    for (uint32_t i = 0; i < c.m; i++) {
        Fr sa = Fr::zero(), sb = Fr::zero();
        for (int q = 0; q < 3; q++) {
            sa = sa + c.a_coef[3ull * i + q] * w[c.a_wire[3ull * i + q]];
            sb = sb + c.b_coef[3ull * i + q] * w[c.b_wire[3ull * i + q]];
        }
        w[1 + c.l + i] = sa * sb;
    }
    out.resize(4ull * c.n_vars);
    for (uint32_t i = 0; i < c.n_vars; i++) put_canon(w[i], out.data() + 4ull * i);
}

void synth_blinding(uint64_t seed, uint64_t r[4], uint64_t s[4]) {
    Xoshiro g(seed ^ TAG_BLIND);
    g.fr_canon(r);
    g.fr_canon(s);
}

void synth_uniform(uint64_t seed, size_t count, uint64_t* out) {
    Xoshiro g(seed);
    for (size_t i = 0; i < count; i++) g.fr_canon(out + 4 * i);
}

void synth_setup(const SynthR1cs& c, SynthSetupExps& o) {
    Xoshiro g(c.seed ^ TAG_TOXIC);
    Fr tau = g.fr(), alpha = g.fr(), beta = g.fr(), gamma = g.fr(), delta = g.fr();
    const uint32_t nV = c.n_vars, l = c.l, n = c.n;
    std::vector<uint64_t> idx(n);
    for (uint32_t k = 0; k < n; k++) idx[k] = k;
    std::vector<Fr> lag;
    lagrange(tau, c.log_n, idx, lag);
    std::vector<Fr> u(nV, Fr::zero()), v(nV, Fr::zero()), ww(nV, Fr::zero());
    for (uint32_t i = 0; i < c.m; i++)
        for (int q = 0; q < 3; q++) {
            u[c.a_wire[3ull * i + q]] = u[c.a_wire[3ull * i + q]] + c.a_coef[3ull * i + q] * lag[i];
            v[c.b_wire[3ull * i + q]] = v[c.b_wire[3ull * i + q]] + c.b_coef[3ull * i + q] * lag[i];
        }
    for (uint32_t k = 0; k <= l; k++) u[k] = u[k] + lag[c.m + k];
    for (uint32_t i = 0; i < c.m; i++) ww[1 + l + i] = lag[i];
    Fr gi = gamma.inv(), di = delta.inv();

    const size_t nC = nV - l - 1;
    o.off_ic = 3;
    o.off_a = o.off_ic + (l + 1);
    o.off_b1 = o.off_a + nV;
    o.off_c = o.off_b1 + nV;
    o.off_h = o.off_c + nC;
    size_t n1 = o.off_h + n;
    o.g1.assign(4 * n1, 0);
    o.off_b2 = 3;
    o.g2.assign(4 * (3 + (size_t)nV), 0);
    auto put1 = [&](size_t i, const Fr& x) { put_canon(x, o.g1.data() + 4 * i); };
    auto put2 = [&](size_t i, const Fr& x) { put_canon(x, o.g2.data() + 4 * i); };
    put1(0, alpha); put1(1, beta); put1(2, delta);
    put2(0, beta); put2(1, gamma); put2(2, delta);
    for (uint32_t i = 0; i < nV; i++) {
        put1(o.off_a + i, u[i]);
        put1(o.off_b1 + i, v[i]);
        put2(o.off_b2 + i, v[i]);
        Fr e = beta * u[i] + alpha * v[i] + ww[i];
        if (i <= l) put1(o.off_ic + i, e * gi);
        else put1(o.off_c + (i - l - 1), e * di);
    }
    // H_j = L^(2n)_{2j+1}(tau) / delta
    std::vector<uint64_t> odd(n);
    for (uint32_t j = 0; j < n; j++) odd[j] = 2ull * j + 1;
    std::vector<Fr> lag2;
    lagrange(tau, c.log_n + 1, odd, lag2);
    for (uint32_t j = 0; j < n; j++) put1(o.off_h + j, lag2[j] * di);
}

void synth_coef_bytes(const SynthR1cs& c, std::vector<uint8_t>& out) {
    size_t nco = 3ull * c.m + c.l + 1 + 3ull * c.m;
    out.resize(44 * nco);
    uint8_t* p = out.data();
    auto put = [&](uint32_t mat, uint32_t row, uint32_t wire, const Fr& coef_mont) {
        Fr raw = Fr::to_mont(coef_mont);  // (c R) R = c R^2
        memcpy(p, &mat, 4); memcpy(p + 4, &row, 4); memcpy(p + 8, &wire, 4);
        memcpy(p + 12, raw.v, 32);
        p += 44;
    };
    for (uint32_t i = 0; i < c.m; i++)
        for (int q = 0; q < 3; q++) put(0, i, c.a_wire[3ull * i + q], c.a_coef[3ull * i + q]);
    for (uint32_t k = 0; k <= c.l; k++) put(0, c.m + k, k, Fr::one());
    for (uint32_t i = 0; i < c.m; i++)
        for (int q = 0; q < 3; q++) put(1, i, c.b_wire[3ull * i + q], c.b_coef[3ull * i + q]);
}

}  // namespace zk
