// snarkjs container files (.zkey, .wtns): mmap reader and writer.
//
// Layout (what the reference reads: src/binfile_utils.cpp:32-80,
// src/zkey_utils.cpp:19-53, src/wtns_utils.cpp:13-26, src/groth16.cu:737-805,
// src/main.cu:28-110):
//   magic[4] | u32 version | u32 nSections | {u32 type, u64 size, payload}*
//   zkey: version <= 1; §1 u32 protocol (1 = groth16); §2 n8q, q, n8r, r,
//         nVars, nPublic, domainSize, alpha1, beta1, beta2, gamma2, delta1,
//         delta2 (G1 64 B / G2 128 B, Montgomery Fp); §3 IC; §4 u32 nCoefs then
//         44-byte entries {u32 matrix, u32 row, u32 wire, 32 B c*R^2 mod r};
//         §5 A, §6 B1 (G1, nVars), §7 B2 (G2, nVars), §8 C (G1,
//         nVars-nPublic-1), §9 H (G1, domainSize).
//   wtns: version <= 2; §1 u32 n8, prime, u32 nVars; §2 nVars x 32 B canonical.
// Sections may appear in any order; the first occurrence of a type is used.
#pragma once
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

namespace zk {

struct MappedFile {
    const uint8_t* data = nullptr;
    size_t size = 0;
    int fd = -1;
    explicit MappedFile(const std::string& path);
    ~MappedFile();
    MappedFile(const MappedFile&) = delete;
    MappedFile& operator=(const MappedFile&) = delete;
};

struct BinSections {
    uint32_t version = 0;
    std::map<uint32_t, std::pair<size_t, size_t>> sec;  // type -> (offset, size)
    void parse(const MappedFile& f, const char* magic, uint32_t max_version);
    const uint8_t* at(const MappedFile& f, uint32_t type, size_t* size) const;
};

struct ZkeyHeader {
    uint32_t n_vars = 0, n_public = 0, domain_size = 0;
    int log_domain = 0;
    uint64_t q[4], r[4];
    uint64_t alpha1[8], beta1[8], delta1[8];
    uint64_t beta2[16], gamma2[16], delta2[16];
};

// parsed view of a .zkey (pointers into the mapping)
struct ZkeyView {
    MappedFile file;
    BinSections bs;
    ZkeyHeader hdr;
    uint32_t n_coefs = 0;
    const uint8_t* coefs = nullptr;  // n_coefs x 44 B
    const uint8_t *A = nullptr, *B1 = nullptr, *B2 = nullptr, *C = nullptr, *H = nullptr, *IC = nullptr;
    explicit ZkeyView(const std::string& path);
};

void read_zkey_header(const std::string& path, ZkeyHeader& h);

// witness: canonical values, n_vars x 4 u64
std::vector<uint64_t> read_wtns(const std::string& path, uint32_t* n_vars);

// CSR of matrices A and B (rows = domain) built from §4 entries, in file order.
struct QapCsr {
    std::vector<uint32_t> rowA, colA, rowB, colB;
    std::vector<uint64_t> valA, valB;  // raw c*R^2, 4 words each
};
void build_csr(const uint8_t* coefs, uint32_t n_coefs, uint32_t domain, uint32_t n_vars, QapCsr& out);

// writer
struct ZkeyWriterInput {
    ZkeyHeader hdr;
    const uint64_t* IC;          // (n_public+1) x 8
    uint32_t n_coefs;
    const uint8_t* coefs;        // n_coefs x 44
    const uint64_t *A, *B1, *B2, *C, *H;  // point arrays in file layout
};
void write_zkey(const std::string& path, const ZkeyWriterInput& in);
void write_wtns(const std::string& path, const uint64_t* w, uint32_t n_vars);

}  // namespace zk
