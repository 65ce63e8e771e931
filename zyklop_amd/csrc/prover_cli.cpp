// zkprover — command-line prover with the reference CLI's contract
// (src/main.cu:115-171):
//   zkprover <circuit.zkey> <witness.wtns> <proof.json> <public.json>
// writes proof.json (snarkjs JSON, also printed) and public.json (w[1..nPublic]).
// Optional: --blind <seed> uses snarkjs blinding with r, s derived from seed
// (the reference proves unblinded, r = s = 0, which stays the default).
//   zkprover verify <circuit.zkey> <public.json> <proof.json>
// checks a proof with the pairing (the reference's `snarkjs groth16 verify`,
// Makefile:135), reading the verification key from the zkey; prints OK or
// INVALID and exits 0 or 2.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/groth16.cuh"

static int fail(const char* what) {
    fprintf(stderr, "[zkprover] %s: %s\n", what, zk_last_error());
    return 1;
}

static bool slurp(const char* path, std::string& out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    char buf[65536];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, k);
    fclose(f);
    return true;
}

static int verify_main(int argc, char** argv) {
    if (argc != 5) {
        fprintf(stderr, "Usage:   %s verify <circuit.zkey> <public.json> <proof.json>\n", argv[0]);
        return 1;
    }
    std::string pub, prf;
    if (!slurp(argv[3], pub) || !slurp(argv[4], prf)) {
        fprintf(stderr, "[zkprover] cannot read %s or %s\n", argv[3], argv[4]);
        return 1;
    }
    int ok = 0;
    if (zk_verify_json(argv[2], pub.c_str(), prf.c_str(), &ok) != ZK_OK) return fail("verify");
    printf("%s\n", ok ? "OK" : "INVALID");
    return ok ? 0 : 2;
}

int main(int argc, char** argv) {
    if (argc >= 2 && strcmp(argv[1], "verify") == 0) return verify_main(argc, argv);
    if (argc < 5) {
        fprintf(stderr, "Usage:   %s <circuit.zkey> <witness.wtns> <proof.json> <public.json> [--blind SEED]\n", argv[0]);
        return 1;
    }
    const char *zkey = argv[1], *wtns = argv[2], *proof_path = argv[3], *public_path = argv[4];
    Groth16Proof proof;
    if (argc >= 7 && strcmp(argv[5], "--blind") == 0) {
        uint64_t r[4], s[4];
        if (zk_synth_blinding(strtoull(argv[6], nullptr, 0), r, s) != ZK_OK) return fail("blinding");
        zk_ctx* ctx = nullptr;
        if (zk_ctx_create(0, &ctx) != ZK_OK) return fail("context");
        zk_proof p;
        if (zk_prove_files(ctx, zkey, wtns, r, s, &p) != ZK_OK) return fail("prove");
        memcpy(proof.pi_A, p.pi_a, 64);
        memcpy(proof.pi_B, p.pi_b, 128);
        memcpy(proof.pi_C, p.pi_c, 64);
        zk_ctx_destroy(ctx);
    } else {
        groth16Prove(zkey, wtns, proof);
    }
    char json[4096];
    groth16ProofToSnarkjsJson(proof, json, sizeof json);
    printf("%s\n", json);
    FILE* f = fopen(proof_path, "w");
    if (!f) {
        fprintf(stderr, "[zkprover] Cannot write to %s\n", proof_path);
        return 1;
    }
    fprintf(f, "%s\n", json);
    fclose(f);
    Groth16ZkeyHeader hdr;
    groth16ReadZkeyHeader(zkey, hdr);
    std::vector<char> pub(96ull * (hdr.n_public + 2) + 16);
    if (zk_public_json(wtns, hdr.n_public, pub.data(), pub.size()) != ZK_OK) return fail("public.json");
    f = fopen(public_path, "w");
    if (!f) {
        fprintf(stderr, "[zkprover] Cannot write to %s\n", public_path);
        return 1;
    }
    fputs(pub.data(), f);
    fclose(f);
    return 0;
}
