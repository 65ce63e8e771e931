// Groth16 verification of proofs against a verification key (zkey §2/§3) and
// the snarkjs text formats it reads: public.json (decimal public inputs) and
// proof.json (decimal affine coordinates). The reference leaves verification
// to snarkjs (`snarkjs groth16 verify`, Makefile:135); the formats are those
// the reference writes (src/groth16.cu:1467-1502, src/main.cu:28-110).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/zk_abi.h"
#include "zkey.hpp"

namespace zk {

// pairing check of proof p for the canonical public inputs (n_public x 4 words)
bool verify_proof(const ZkeyHeader& h, const uint64_t* ic, const uint64_t* publics, uint32_t n_public, const zk_proof& p);

// public.json: every quoted decimal in order, canonical Fr (< r) each
std::vector<uint64_t> parse_public_json(const std::string& text);
// proof.json: pi_a, pi_b, pi_c as written by proof_to_json (projective "1" /
// ["1","0"] third coordinates), to Montgomery affine
zk_proof parse_proof_json(const std::string& text);

std::string read_text_file(const std::string& path);

}  // namespace zk
