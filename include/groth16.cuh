/*
 * groth16.cuh — drop-in replacement of the reference's public prover API
 * (include/groth16.cuh of the CUDA prover), implemented on the MI355X C ABI
 * in zk_abi.h / libzkp.so.
 *
 * Same types and signatures as the reference:
 *   Groth16Proof, Groth16ZkeyHeader, Groth16CoefEntry        (groth16.cuh:57-98)
 *   groth16Prove                                              (:119-123)
 *   groth16ProofToSnarkjsJson, groth16ProofToCalldata         (:149-167)
 *   groth16ReadZkeyHeader                                     (:177-180)
 *   groth16SparseMvm (declared, never defined there)          (:197-205)
 * Error behaviour matches the reference's: a failure prints a message to
 * stderr and exits with status 1. Code that wants status codes calls zk_abi.h.
 *
 * groth16Prove computes the reference's proof (no blinding: r = s = 0). The
 * blinded snarkjs proof is zk_prove_files with nonzero r, s.
 */
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "zk_abi.h"

struct Groth16Proof {
    uint64_t pi_A[8];   // G1 affine [x[4], y[4]], Montgomery Fp
    uint64_t pi_B[16];  // G2 affine [x.a0, x.a1, y.a0, y.a1], Montgomery Fp
    uint64_t pi_C[8];   // G1 affine, Montgomery Fp
};

struct Groth16ZkeyHeader {
    uint32_t n_constraints;
    uint32_t n_vars;
    uint32_t n_public;
    uint32_t domain_size;
    int log_domain;
    uint64_t prime[4];
};

// One §4 entry. The reference's header documents a 1-byte matrix field; its
// parser (src/groth16.cu:759-760) and every zkey use a u32, as here.
struct Groth16CoefEntry {
    uint8_t matrix;
    uint32_t constraint;
    uint32_t signal;
    uint64_t value[4];
};

namespace groth16_detail {
inline void die(const char* what) {
    fprintf(stderr, "[groth16] %s failed: %s\n", what, zk_last_error());
    exit(1);
}
inline zk_ctx* ctx() {
    static zk_ctx* c = nullptr;
    if (!c && zk_ctx_create(0, &c) != ZK_OK) die("zk_ctx_create");
    return c;
}
}  // namespace groth16_detail

inline void groth16Prove(const char* zkey_path, const char* wtns_path, Groth16Proof& proof) {
    zk_proof p;
    if (zk_prove_files(groth16_detail::ctx(), zkey_path, wtns_path, nullptr, nullptr, &p) != ZK_OK)
        groth16_detail::die("groth16Prove");
    memcpy(proof.pi_A, p.pi_a, sizeof proof.pi_A);
    memcpy(proof.pi_B, p.pi_b, sizeof proof.pi_B);
    memcpy(proof.pi_C, p.pi_c, sizeof proof.pi_C);
}

inline void groth16ProofToSnarkjsJson(const Groth16Proof& proof, char* out, size_t out_size) {
    zk_proof p;
    memcpy(p.pi_a, proof.pi_A, sizeof p.pi_a);
    memcpy(p.pi_b, proof.pi_B, sizeof p.pi_b);
    memcpy(p.pi_c, proof.pi_C, sizeof p.pi_c);
    if (zk_proof_to_json(&p, out, out_size) != ZK_OK) groth16_detail::die("groth16ProofToSnarkjsJson");
}

inline void groth16ProofToCalldata(const Groth16Proof& proof, uint8_t* out, size_t out_size) {
    if (out_size < 256) {  // reference: prints and returns (src/groth16.cu:1519)
        fprintf(stderr, "groth16ProofToCalldata: buffer too small\n");
        return;
    }
    zk_proof p;
    memcpy(p.pi_a, proof.pi_A, sizeof p.pi_a);
    memcpy(p.pi_b, proof.pi_B, sizeof p.pi_b);
    memcpy(p.pi_c, proof.pi_C, sizeof p.pi_c);
    if (zk_proof_to_calldata(&p, out, out_size) != ZK_OK) groth16_detail::die("groth16ProofToCalldata");
}

inline void groth16ReadZkeyHeader(const char* zkey_path, Groth16ZkeyHeader& hdr) {
    zk_zkey_header h;
    if (zk_read_zkey_header(zkey_path, &h) != ZK_OK) groth16_detail::die("groth16ReadZkeyHeader");
    hdr.n_constraints = h.n_constraints;
    hdr.n_vars = h.n_vars;
    hdr.n_public = h.n_public;
    hdr.domain_size = h.domain_size;
    hdr.log_domain = h.log_domain;
    memcpy(hdr.prime, h.prime, sizeof hdr.prime);
}

// coefs: entries as Groth16CoefEntry with canonical values (the reference
// header's contract, :93-98), entries of other matrices skipped; witness
// canonical; out canonical, zero-padded to domain_size.
inline void groth16SparseMvm(const Groth16CoefEntry* coefs, size_t n_coefs, uint8_t matrix_sel, const uint64_t* witness,
                             uint32_t n_vars, uint64_t* out, uint32_t domain_size) {
    std::vector<uint8_t> raw(44 * n_coefs);
    for (size_t i = 0; i < n_coefs; i++) {
        uint32_t mat = coefs[i].matrix;
        memcpy(&raw[44 * i], &mat, 4);
        memcpy(&raw[44 * i + 4], &coefs[i].constraint, 4);
        memcpy(&raw[44 * i + 8], &coefs[i].signal, 4);
        memcpy(&raw[44 * i + 12], coefs[i].value, 32);
    }
    if (zk_sparse_mvm_canonical(raw.data(), n_coefs, matrix_sel, witness, n_vars, out, domain_size) != ZK_OK)
        groth16_detail::die("groth16SparseMvm");
}
