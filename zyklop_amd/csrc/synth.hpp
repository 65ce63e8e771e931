// Seeded synthetic Groth16 workload (R1CS, witness, trapdoor setup exponents).
//
// Recipe (shared with the oracle, which implements it independently):
//   PRNG: SplitMix64 seeding xoshiro256**, one stream per purpose seeded with
//   seed ^ TAG; uniform Fr = rejection sampling of 254-bit candidates (four
//   outputs, top word masked to 62 bits) below r; uniform integer below b =
//   high 64 bits of next() * b.
//   R1CS (m constraints, l public inputs): wire 0 = 1, wires 1..l public,
//   wire l+1+i = aux_i. Row i: A_i and B_i have 3 terms each on wires
//   < l+1+i, drawn (wire, coef) in that order, A first; C_i = 1 * aux_i.
//   Public-input rows A[m+k][k] = 1 for k = 0..l. Domain n = next pow2 of m+l+1.
//   Witness: public inputs uniform (own stream), aux_i = (A_i.w)(B_i.w).
//   Toxic waste tau, alpha, beta, gamma, delta (own stream); blinding r, s.
#pragma once
#include <stdint.h>

#include <vector>

#include "host_field.hpp"

namespace zk {

struct SynthR1cs {
    uint64_t seed = 0;
    uint32_t m = 0, l = 0, n_vars = 0, n = 0;
    int log_n = 0;
    std::vector<uint32_t> a_wire, b_wire;    // 3*m
    std::vector<zkh::Fr> a_coef, b_coef;     // Montgomery
};

void synth_r1cs(uint64_t seed, uint32_t m, uint32_t n_pub, SynthR1cs& out);
// satisfying witness for the given public inputs (Montgomery) -> canonical n_vars x 4
void synth_witness_from_public(const SynthR1cs& c, const std::vector<zkh::Fr>& pub, std::vector<uint64_t>& w);
// public inputs from the seed's stream (the circuit's default witness)
std::vector<zkh::Fr> synth_public(uint64_t seed, uint32_t n_pub);
void synth_blinding(uint64_t seed, uint64_t r[4], uint64_t s[4]);
// uniform canonical Fr values from an arbitrary seed (fresh witnesses)
void synth_uniform(uint64_t seed, size_t count, uint64_t* out);

struct SynthSetupExps {
    // canonical exponents; G1 list: alpha, beta, delta, IC[0..l], A[nV], B1[nV],
    // C[nV-l-1], H[n]; G2 list: beta, gamma, delta, B2[nV]
    std::vector<uint64_t> g1, g2;
    size_t off_ic, off_a, off_b1, off_c, off_h, off_b2;
};
void synth_setup(const SynthR1cs& c, SynthSetupExps& out);

// §4 entries (44 B each, raw c*R^2): A rows, public rows, B rows
void synth_coef_bytes(const SynthR1cs& c, std::vector<uint8_t>& out);

}  // namespace zk
