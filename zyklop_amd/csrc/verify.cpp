#include "verify.hpp"

#include <stdio.h>
#include <string.h>

#include "common.hpp"
#include "host_field.hpp"
#include "pairing.hpp"

namespace zk {

bool verify_proof(const ZkeyHeader& h, const uint64_t* ic, const uint64_t* publics, uint32_t n_public, const zk_proof& p) {
    ZK_REQUIRE(n_public == h.n_public, ZK_E_ARG,
               "expected " + std::to_string(h.n_public) + " public inputs, got " + std::to_string(n_public));
    for (uint32_t i = 0; i < n_public; i++)
        ZK_REQUIRE(zkh::Fr::is_canonical(publics + 4ull * i), ZK_E_ARG,
                   "public input " + std::to_string(i) + " is not canonical (>= r)");
    return zkh::groth16_verify(h.alpha1, h.beta2, h.gamma2, h.delta2, ic, n_public, publics, p.pi_a, p.pi_b, p.pi_c);
}

// the quoted strings of a JSON text, in order (no escapes in these formats)
static std::vector<std::string> quoted(const std::string& t, size_t from = 0) {
    std::vector<std::string> out;
    size_t i = t.find('"', from);
    while (i != std::string::npos) {
        size_t j = t.find('"', i + 1);
        ZK_REQUIRE(j != std::string::npos, ZK_E_FORMAT, "unterminated string in JSON");
        out.push_back(t.substr(i + 1, j - i - 1));
        i = t.find('"', j + 1);
    }
    return out;
}

// decimal string -> canonical value below modulus m (4 words)
static void parse_dec(const std::string& s, const uint64_t m[4], uint64_t out[4]) {
    ZK_REQUIRE(!s.empty() && s.size() <= 78, ZK_E_FORMAT, "bad decimal field element \"" + s + "\"");
    uint64_t v[5] = {0, 0, 0, 0, 0};
    for (char ch : s) {
        ZK_REQUIRE(ch >= '0' && ch <= '9', ZK_E_FORMAT, "bad decimal field element \"" + s + "\"");
        zkh::u128 c = (uint64_t)(ch - '0');
        for (int k = 0; k < 5; k++) {
            c += (zkh::u128)v[k] * 10;
            v[k] = (uint64_t)c;
            c >>= 64;
        }
    }
    bool below = v[4] == 0;
    if (below) {
        below = false;
        for (int k = 3; k >= 0; k--)
            if (v[k] != m[k]) { below = v[k] < m[k]; break; }
    }
    ZK_REQUIRE(below, ZK_E_FORMAT, "field element \"" + s + "\" is not below the modulus");
    memcpy(out, v, 32);
}

std::vector<uint64_t> parse_public_json(const std::string& text) {
    std::vector<uint64_t> w;
    for (const std::string& s : quoted(text)) {
        uint64_t v[4];
        parse_dec(s, zkh::FrP::M, v);
        w.insert(w.end(), v, v + 4);
    }
    return w;
}

// the n quoted strings after "key"
static std::vector<std::string> field(const std::string& t, const char* key, size_t n) {
    const std::string k = std::string("\"") + key + "\"";
    size_t at = t.find(k);
    ZK_REQUIRE(at != std::string::npos, ZK_E_FORMAT, std::string("proof JSON has no ") + key);
    std::vector<std::string> q = quoted(t, at + k.size());
    ZK_REQUIRE(q.size() >= n, ZK_E_FORMAT, std::string("proof JSON: short ") + key);
    q.resize(n);
    return q;
}

static void put_mont(const std::string& dec, uint64_t* dst) {
    uint64_t c[4];
    parse_dec(dec, zkh::FpP::M, c);
    zkh::Fp m = zkh::Fp::to_mont(zkh::Fp::from_raw(c));
    memcpy(dst, m.v, 32);
}

zk_proof parse_proof_json(const std::string& text) {
    zk_proof p;
    memset(&p, 0, sizeof p);
    const auto a = field(text, "pi_a", 3), b = field(text, "pi_b", 6), c = field(text, "pi_c", 3);
    ZK_REQUIRE(a[2] == "1" && c[2] == "1" && b[4] == "1" && b[5] == "0", ZK_E_FORMAT,
               "proof JSON: only affine points (third coordinate 1) are supported");
    put_mont(a[0], p.pi_a);
    put_mont(a[1], p.pi_a + 4);
    for (int k = 0; k < 4; k++) put_mont(b[k], p.pi_b + 4 * k);
    put_mont(c[0], p.pi_c);
    put_mont(c[1], p.pi_c + 4);
    return p;
}

std::string read_text_file(const std::string& path) {
    FILE* f = fopen(path.c_str(), "rb");
    ZK_REQUIRE(f != nullptr, ZK_E_IO, "cannot open " + path);
    std::string s;
    char buf[65536];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
    fclose(f);
    return s;
}

}  // namespace zk
