#include "zkey.hpp"

#include <fcntl.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>

#include "common.hpp"
#include "host_field.hpp"

namespace zk {

MappedFile::MappedFile(const std::string& path) {
    fd = ::open(path.c_str(), O_RDONLY);
    ZK_REQUIRE(fd >= 0, ZK_E_IO, "cannot open " + path);
    struct stat st;
    if (fstat(fd, &st) != 0) {
        ::close(fd);
        throw Error(ZK_E_IO, "cannot stat " + path);
    }
    size = (size_t)st.st_size;
    if (size) {
        void* p = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (p == MAP_FAILED) {
            ::close(fd);
            throw Error(ZK_E_IO, "cannot mmap " + path);
        }
        madvise(p, size, MADV_SEQUENTIAL);
        data = (const uint8_t*)p;
    }
}

MappedFile::~MappedFile() {
    if (data) munmap((void*)data, size);
    if (fd >= 0) ::close(fd);
}

static uint32_t rd32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t rd64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }

void BinSections::parse(const MappedFile& f, const char* magic, uint32_t max_version) {
    ZK_REQUIRE(f.size >= 12, ZK_E_FORMAT, "file is too short");
    ZK_REQUIRE(memcmp(f.data, magic, 4) == 0, ZK_E_FORMAT, std::string("invalid file type, expected ") + magic);
    version = rd32(f.data + 4);
    ZK_REQUIRE(version <= max_version, ZK_E_FORMAT,
               "invalid version " + std::to_string(version) + " (max " + std::to_string(max_version) + ")");
    uint32_t ns = rd32(f.data + 8);
    ZK_REQUIRE((uint64_t)f.size >= 12 + (uint64_t)ns * 12, ZK_E_FORMAT, "file too short for its section table");
    size_t pos = 12;
    for (uint32_t i = 0; i < ns; i++) {
        ZK_REQUIRE(pos + 12 <= f.size, ZK_E_FORMAT, "truncated section header");
        uint32_t type = rd32(f.data + pos);
        uint64_t sz = rd64(f.data + pos + 4);
        pos += 12;
        ZK_REQUIRE(sz <= f.size - pos, ZK_E_FORMAT, "section " + std::to_string(type) + " runs past end of file");
        if (!sec.count(type)) sec[type] = {pos, (size_t)sz};
        pos += sz;
    }
}

const uint8_t* BinSections::at(const MappedFile& f, uint32_t type, size_t* size) const {
    auto it = sec.find(type);
    ZK_REQUIRE(it != sec.end(), ZK_E_FORMAT, "missing section " + std::to_string(type));
    if (size) *size = it->second.second;
    return f.data + it->second.first;
}

static const uint64_t P_LE[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
static const uint64_t R_LE[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};

static void parse_header(const MappedFile& f, const BinSections& bs, ZkeyHeader& h) {
    size_t s1;
    const uint8_t* p1 = bs.at(f, 1, &s1);
    ZK_REQUIRE(s1 >= 4, ZK_E_FORMAT, "section 1 too short");
    ZK_REQUIRE(rd32(p1) == 1, ZK_E_FORMAT, "zkey file is not groth16");
    size_t s2;
    const uint8_t* p = bs.at(f, 2, &s2);
    const size_t need = 4 + 32 + 4 + 32 + 12 + 3 * 64 + 3 * 128;
    ZK_REQUIRE(s2 >= need, ZK_E_FORMAT, "section 2 too short");
    uint32_t n8q = rd32(p);
    ZK_REQUIRE(n8q == 32, ZK_E_FORMAT, "only 32-byte base field supported (BN254)");
    memcpy(h.q, p + 4, 32);
    p += 36;
    uint32_t n8r = rd32(p);
    ZK_REQUIRE(n8r == 32, ZK_E_FORMAT, "only 32-byte scalar field supported (BN254)");
    memcpy(h.r, p + 4, 32);
    p += 36;
    ZK_REQUIRE(memcmp(h.q, P_LE, 32) == 0 && memcmp(h.r, R_LE, 32) == 0, ZK_E_FORMAT, "zkey is not over BN254");
    h.n_vars = rd32(p);
    h.n_public = rd32(p + 4);
    h.domain_size = rd32(p + 8);
    p += 12;
    memcpy(h.alpha1, p, 64); p += 64;
    memcpy(h.beta1, p, 64); p += 64;
    memcpy(h.beta2, p, 128); p += 128;
    memcpy(h.gamma2, p, 128); p += 128;
    memcpy(h.delta1, p, 64); p += 64;
    memcpy(h.delta2, p, 128);
    ZK_REQUIRE(h.domain_size >= 1 && (h.domain_size & (h.domain_size - 1)) == 0, ZK_E_FORMAT,
               "domain size must be a power of two");
    ZK_REQUIRE(h.n_vars >= h.n_public + 1, ZK_E_FORMAT, "nVars < nPublic + 1");
    h.log_domain = ilog2(h.domain_size);
}

void read_zkey_header(const std::string& path, ZkeyHeader& h) {
    MappedFile f(path);
    BinSections bs;
    bs.parse(f, "zkey", 1);
    parse_header(f, bs, h);
}

ZkeyView::ZkeyView(const std::string& path) : file(path) {
    bs.parse(file, "zkey", 1);
    parse_header(file, bs, hdr);
    size_t sz;
    const uint8_t* p4 = bs.at(file, 4, &sz);
    ZK_REQUIRE(sz >= 4, ZK_E_FORMAT, "section 4 too short");
    n_coefs = rd32(p4);
    ZK_REQUIRE((uint64_t)sz >= 4 + 44ull * n_coefs, ZK_E_FORMAT, "section 4 shorter than its coefficient count");
    coefs = p4 + 4;
    const uint64_t nv = hdr.n_vars, nc = hdr.n_vars - hdr.n_public - 1, nh = hdr.domain_size;
    A = bs.at(file, 5, &sz);  ZK_REQUIRE(sz >= 64 * nv, ZK_E_FORMAT, "section 5 too short");
    B1 = bs.at(file, 6, &sz); ZK_REQUIRE(sz >= 64 * nv, ZK_E_FORMAT, "section 6 too short");
    B2 = bs.at(file, 7, &sz); ZK_REQUIRE(sz >= 128 * nv, ZK_E_FORMAT, "section 7 too short");
    C = bs.at(file, 8, &sz);  ZK_REQUIRE(sz >= 64 * nc, ZK_E_FORMAT, "section 8 too short");
    H = bs.at(file, 9, &sz);  ZK_REQUIRE(sz >= 64 * nh, ZK_E_FORMAT, "section 9 too short");
    if (bs.sec.count(3)) {
        IC = bs.at(file, 3, &sz);
        ZK_REQUIRE(sz >= 64ull * (hdr.n_public + 1), ZK_E_FORMAT, "section 3 shorter than nPublic + 1 points");
    }
}

std::vector<uint64_t> read_wtns(const std::string& path, uint32_t* n_vars) {
    MappedFile f(path);
    BinSections bs;
    bs.parse(f, "wtns", 2);
    size_t s1, s2;
    const uint8_t* p1 = bs.at(f, 1, &s1);
    ZK_REQUIRE(s1 >= 40, ZK_E_FORMAT, "wtns section 1 too short");
    ZK_REQUIRE(rd32(p1) == 32, ZK_E_FORMAT, "wtns field width must be 32 bytes");
    ZK_REQUIRE(memcmp(p1 + 4, R_LE, 32) == 0, ZK_E_FORMAT, "wtns prime is not BN254 r");
    uint32_t n = rd32(p1 + 36);
    const uint8_t* p2 = bs.at(f, 2, &s2);
    ZK_REQUIRE((uint64_t)s2 >= 32ull * n, ZK_E_FORMAT, "wtns section 2 shorter than nVars values");
    std::vector<uint64_t> w((size_t)n * 4);
    memcpy(w.data(), p2, 32ull * n);
    *n_vars = n;
    return w;
}

void build_csr(const uint8_t* coefs, uint32_t n_coefs, uint32_t domain, uint32_t n_vars, QapCsr& out) {
    // counts per row (matrix 2 = C is not needed: C*w is taken as A*w o B*w)
    std::vector<uint32_t> cntA(domain + 1, 0), cntB(domain + 1, 0);
    for (uint32_t i = 0; i < n_coefs; i++) {
        const uint8_t* e = coefs + 44ull * i;
        uint32_t mat = rd32(e), row = rd32(e + 4), wire = rd32(e + 8);
        ZK_REQUIRE(mat <= 2, ZK_E_FORMAT, "coefficient matrix id out of range");
        if (mat == 2) continue;
        ZK_REQUIRE(row < domain, ZK_E_FORMAT, "coefficient row out of the domain");
        ZK_REQUIRE(wire < n_vars, ZK_E_FORMAT, "coefficient wire out of range");
        (mat == 0 ? cntA : cntB)[row + 1]++;
    }
    for (uint32_t i = 0; i < domain; i++) { cntA[i + 1] += cntA[i]; cntB[i + 1] += cntB[i]; }
    out.rowA = cntA;
    out.rowB = cntB;
    out.colA.assign(cntA[domain], 0);
    out.colB.assign(cntB[domain], 0);
    out.valA.assign((size_t)cntA[domain] * 4, 0);
    out.valB.assign((size_t)cntB[domain] * 4, 0);
    std::vector<uint32_t> posA(cntA.begin(), cntA.end() - 1), posB(cntB.begin(), cntB.end() - 1);
    for (uint32_t i = 0; i < n_coefs; i++) {
        const uint8_t* e = coefs + 44ull * i;
        uint32_t mat = rd32(e), row = rd32(e + 4), wire = rd32(e + 8);
        if (mat == 2) continue;
        uint32_t k = (mat == 0 ? posA : posB)[row]++;
        (mat == 0 ? out.colA : out.colB)[k] = wire;
        memcpy((mat == 0 ? out.valA : out.valB).data() + 4ull * k, e + 12, 32);
    }
}

namespace {
struct Out {
    FILE* f;
    explicit Out(const std::string& path) : f(fopen(path.c_str(), "wb")) {
        ZK_REQUIRE(f, ZK_E_IO, "cannot write " + path);
    }
    ~Out() { if (f) fclose(f); }
    void put(const void* p, size_t n) { ZK_REQUIRE(fwrite(p, 1, n, f) == n, ZK_E_IO, "short write"); }
    void u32(uint32_t v) { put(&v, 4); }
    void u64(uint64_t v) { put(&v, 8); }
    void section(uint32_t type, uint64_t size) { u32(type); u64(size); }
};
}  // namespace

void write_zkey(const std::string& path, const ZkeyWriterInput& in) {
    const ZkeyHeader& h = in.hdr;
    Out o(path);
    o.put("zkey", 4);
    o.u32(1);
    o.u32(9);
    o.section(1, 4);
    o.u32(1);
    o.section(2, 4 + 32 + 4 + 32 + 12 + 3 * 64 + 3 * 128);
    o.u32(32); o.put(P_LE, 32); o.u32(32); o.put(R_LE, 32);
    o.u32(h.n_vars); o.u32(h.n_public); o.u32(h.domain_size);
    o.put(h.alpha1, 64); o.put(h.beta1, 64); o.put(h.beta2, 128); o.put(h.gamma2, 128); o.put(h.delta1, 64); o.put(h.delta2, 128);
    o.section(3, 64ull * (h.n_public + 1));
    o.put(in.IC, 64ull * (h.n_public + 1));
    o.section(4, 4 + 44ull * in.n_coefs);
    o.u32(in.n_coefs);
    o.put(in.coefs, 44ull * in.n_coefs);
    o.section(5, 64ull * h.n_vars); o.put(in.A, 64ull * h.n_vars);
    o.section(6, 64ull * h.n_vars); o.put(in.B1, 64ull * h.n_vars);
    o.section(7, 128ull * h.n_vars); o.put(in.B2, 128ull * h.n_vars);
    uint64_t nc = h.n_vars - h.n_public - 1;
    o.section(8, 64ull * nc); o.put(in.C, 64ull * nc);
    o.section(9, 64ull * h.domain_size); o.put(in.H, 64ull * h.domain_size);
}

void write_wtns(const std::string& path, const uint64_t* w, uint32_t n_vars) {
    Out o(path);
    o.put("wtns", 4);
    o.u32(2);
    o.u32(2);
    o.section(1, 40);
    o.u32(32); o.put(R_LE, 32); o.u32(n_vars);
    o.section(2, 32ull * n_vars);
    o.put(w, 32ull * n_vars);
}

}  // namespace zk
