#include "prover.hpp"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <future>
#include <thread>

#include "qap.hpp"
#include "synth.hpp"

namespace zk {

using zkh::Fp;
using zkh::Fp2;
using zkh::Fr;

// ------------------------------------------------------------------ context
Context::Context(int dev) : device(dev) {
    int count = 0;
    ZK_HIP(hipGetDeviceCount(&count));
    ZK_REQUIRE(dev >= 0 && dev < count, ZK_E_ARG, "no such HIP device: " + std::to_string(dev));
    ZK_HIP(hipSetDevice(dev));
    ZK_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    ZK_HIP(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
    msm_noalloc_setup(dev);   // device address of the fixed MSM scratch, once per device
    // default window-table mode for circuits loaded on this context
    if (const char* e = getenv("ZK_WINDOW_TABLES")) {
        const int m = atoi(e);
        ZK_REQUIRE(m >= -1 && m <= 1, ZK_E_ARG, "ZK_WINDOW_TABLES must be -1, 0 or 1");
        tables_mode = m;
    }
}

Context::~Context() {
    (void)hipSetDevice(device);
    slot.reset();
    ntt_plans.clear();
    if (stream) (void)hipStreamDestroy(stream);
    if (aux_stream) (void)hipStreamDestroy(aux_stream);
}

NttPlan& Context::ntt_plan(int log_n) {
    auto it = ntt_plans.find(log_n);
    if (it != ntt_plans.end()) return *it->second;
    auto p = std::make_unique<NttPlan>();
    p->build(log_n, stream);
    NttPlan& ref = *p;
    ntt_plans[log_n] = std::move(p);
    return ref;
}

// fixed-base tables: T[j][d] = d * 2^(8j) * G, affine Montgomery
template <class T>
static void comb_table(const zkh::Aff<T>& g, std::vector<uint64_t>& out) {
    const int lw = sizeof(T) / 8;
    out.assign((size_t)32 * 256 * 2 * lw, 0);
    zkh::Xyzz<T> base = zkh::Xyzz<T>::from_aff(g);
    for (int j = 0; j < 32; j++) {
        zkh::Xyzz<T> acc = zkh::Xyzz<T>::infinity();
        for (int d = 0; d < 256; d++) {
            if (d) {
                zkh::Aff<T> a = zkh::xyzz_to_aff(acc);
                uint64_t* dst = out.data() + ((size_t)j * 256 + d) * 2 * lw;
                memcpy(dst, &a.x, 8 * lw);
                memcpy(dst + lw, &a.y, 8 * lw);
            }
            acc = zkh::xyzz_add(acc, base);
        }
        for (int k = 0; k < 8; k++) base = zkh::xyzz_dbl(base);
    }
}

static zkh::Aff<Fp> g1_generator() {
    return {Fp::from_u64(1), Fp::from_u64(2), false};
}
static zkh::Aff<Fp2> g2_generator() {
    // canonical coordinates of the BN254 G2 generator (EIP-197)
    static const uint64_t x0[4] = {0x46debd5cd992f6edULL, 0x674322d4f75edaddULL, 0x426a00665e5c4479ULL, 0x1800deef121f1e76ULL};
    static const uint64_t x1[4] = {0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL, 0x7260bfb731fb5d25ULL, 0x198e9393920d483aULL};
    static const uint64_t y0[4] = {0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL, 0x4aab71808dcb408fULL, 0x12c85ea5db8c6debULL};
    static const uint64_t y1[4] = {0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL, 0xec9e99ad690c3395ULL, 0x090689d0585ff075ULL};
    auto m = [](const uint64_t* v) { return Fp::to_mont(Fp::from_raw(v)); };
    return {{m(x0), m(x1)}, {m(y0), m(y1)}, false};
}

void Context::ensure_comb() {
    if (comb_g1.p && comb_g2.p) return;
    std::vector<uint64_t> t1, t2;
    comb_table(g1_generator(), t1);
    comb_table(g2_generator(), t2);
    comb_g1.alloc(t1.size() * 8);
    comb_g2.alloc(t2.size() * 8);
    ZK_HIP(hipMemcpy(comb_g1.p, t1.data(), t1.size() * 8, hipMemcpyHostToDevice));
    ZK_HIP(hipMemcpy(comb_g2.p, t2.data(), t2.size() * 8, hipMemcpyHostToDevice));
}

// ------------------------------------------------------------------ circuits
size_t Circuit::device_bytes() const {
    return rowA.bytes + colA.bytes + valA.bytes + rowB.bytes + colB.bytes + valB.bytes + A.bytes + B1.bytes + B2.bytes +
           C.bytes + H.bytes + A29.bytes + B1_29.bytes + B2_29.bytes + C29.bytes + H29.bytes + wit_rows.bytes +
           wire_cls.bytes;
}

static void upload(DevBuf& d, const void* src, size_t bytes) {
    d.alloc(bytes);
    if (bytes) ZK_HIP(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
}

static void upload_csr(Circuit& c, const QapCsr& csr) {
    upload(c.rowA, csr.rowA.data(), csr.rowA.size() * 4);
    upload(c.colA, csr.colA.data(), csr.colA.size() * 4);
    upload(c.valA, csr.valA.data(), csr.valA.size() * 8);
    upload(c.rowB, csr.rowB.data(), csr.rowB.size() * 4);
    upload(c.colB, csr.colB.data(), csr.colB.size() * 4);
    upload(c.valB, csr.valB.data(), csr.valB.size() * 8);
    c.n_rowsA = (uint32_t)csr.rowA.size() - 1;
    c.n_rowsB = (uint32_t)csr.rowB.size() - 1;
    c.nnzA = csr.colA.size();
    c.nnzB = csr.colB.size();
}

// extended base arrays: [bases..., e0, e1, e2, e3] with per-slot extra points
static void build_ext(DevBuf& d, const uint8_t* lead_zeros_then, size_t lead_zero_pts, const uint8_t* pts, size_t npts,
                      const uint64_t* const extras[4], size_t pt_bytes) {
    (void)lead_zeros_then;
    size_t total = lead_zero_pts + npts + 4;
    std::vector<uint8_t> h(total * pt_bytes, 0);
    if (npts) memcpy(h.data() + lead_zero_pts * pt_bytes, pts, npts * pt_bytes);
    for (int k = 0; k < 4; k++)
        if (extras[k]) memcpy(h.data() + (lead_zero_pts + npts + k) * pt_bytes, extras[k], pt_bytes);
    upload(d, h.data(), h.size());
}

static bool words_zero(const uint8_t* p, size_t bytes) {
    uint64_t v[16];
    memcpy(v, p, bytes);
    uint64_t o = 0;
    for (size_t k = 0; k < bytes / 8; k++) o |= v[k];
    return o == 0;
}

// Circuit::wire_cls from the extended A', B1', B2' arrays (bases, then the
// four extra points, a null extra being infinity)
static void build_wire_cls(Circuit& c, const uint8_t* A, const uint8_t* B1, const uint8_t* B2,
                           const uint64_t* const exA[4], const uint64_t* const exB1[4], const uint64_t* const exB2[4]) {
    const size_t nv = c.n_vars, ne = nv + 4;
    auto inf = [&](const uint8_t* base, const uint64_t* const ex[4], size_t bytes, size_t i) {
        const uint8_t* p = i < nv ? base + bytes * i : reinterpret_cast<const uint8_t*>(ex[i - nv]);
        return !p || words_zero(p, bytes);
    };
    std::vector<uint8_t> cls(ne);
    size_t n_inf = 0;
    for (size_t i = 0; i < ne; i++) {
        const bool a = inf(A, exA, 64, i), b = inf(B1, exB1, 64, i) && inf(B2, exB2, 128, i);
        cls[i] = a ? (b ? MSM_CLS_AB_INF : MSM_CLS_A_INF) : (b ? MSM_CLS_B_INF : MSM_CLS_NONE);
        n_inf += (size_t)a + (size_t)b;
    }
    if (n_inf * 8 >= ne) upload(c.wire_cls, cls.data(), ne);
    else c.wire_cls.release();
}

static void finish_circuit_bases(Circuit& c, const uint8_t* A, const uint8_t* B1, const uint8_t* B2, const uint8_t* C,
                                 const uint8_t* H) {
    const ZkeyHeader& h = c.hdr;
    const size_t nv = c.n_vars, l = c.n_public;
    const uint64_t* exA[4] = {h.alpha1, h.delta1, nullptr, nullptr};
    const uint64_t* exB1[4] = {h.beta1, nullptr, h.delta1, nullptr};
    const uint64_t* exB2[4] = {h.beta2, nullptr, h.delta2, nullptr};
    const uint64_t* exC[4] = {nullptr, nullptr, nullptr, h.delta1};
    build_ext(c.A, nullptr, 0, A, nv, exA, 64);
    build_ext(c.B1, nullptr, 0, B1, nv, exB1, 64);
    build_ext(c.B2, nullptr, 0, B2, nv, exB2, 128);
    build_wire_cls(c, A, B1, B2, exA, exB1, exB2);
    build_ext(c.C, nullptr, l + 1, C, nv - l - 1, exC, 64);
    upload(c.H, H, (size_t)c.n * 64);
    circuit_set_tables(c, c.ctx->tables_mode);
}

void circuit_set_tables(Circuit& c, int mode) {
    const size_t ne = (size_t)c.n_vars + 4, n = c.n;
    const int cw = msm_table_c(ne), ch = msm_table_c(n);
    const int Ww = msm_geom(ne, cw, false, true).W, Wh = msm_geom(n, ch, false, true).W;
    bool on = mode > 0;
    if (mode < 0) {
        const size_t extra = (size_t)(Ww - 1) * ne * (3 * PT29_BYTES + PT29Q_BYTES) + (size_t)(Wh - 1) * n * PT29_BYTES;
        size_t free_b = 0, total_b = 0;
        ZK_HIP(hipMemGetInfo(&free_b, &total_b));
        // tables the circuit holds now are freed before the new ones are made
        const size_t held = c.A29.bytes + c.B1_29.bytes + c.C29.bytes + c.B2_29.bytes + c.H29.bytes;
        const size_t base = ne * (3 * PT29_BYTES + PT29Q_BYTES) + n * PT29_BYTES;
        on = extra <= (free_b + (held > base ? held - base : 0)) / 3;
    }
    // entries name a table slot in 30 bits
    const bool fits = (size_t)Ww * ne < 0x40000000ull && (size_t)Wh * n < 0x40000000ull;
    ZK_REQUIRE(!(mode > 0 && !fits), ZK_E_ARG, "window tables need W*n < 2^30 slots");
    on = on && fits;
    // the arrays are reallocated: nothing may still read them
    ZK_HIP(hipDeviceSynchronize());
    hipStream_t st = c.ctx->stream;
    const size_t kw = on ? (size_t)Ww : 1, kh = on ? (size_t)Wh : 1;
    for (auto pr : {std::make_pair(&c.A, &c.A29), std::make_pair(&c.B1, &c.B1_29), std::make_pair(&c.C, &c.C29)}) {
        pr.second->alloc(kw * ne * PT29_BYTES);
        if (on) msm_points_table(pr.first->p, ne, cw, pr.second->p, false, st);
        else msm_points_to29(pr.first->p, ne, pr.second->p, false, st);
    }
    c.B2_29.alloc(kw * ne * PT29Q_BYTES);
    if (on) msm_points_table(c.B2.p, ne, cw, c.B2_29.p, true, st);
    else msm_points_to29(c.B2.p, ne, c.B2_29.p, true, st);
    c.H29.alloc(kh * n * PT29_BYTES);
    if (on) msm_points_table(c.H.p, n, ch, c.H29.p, false, st);
    else msm_points_to29(c.H.p, n, c.H29.p, false, st);
    ZK_HIP(hipStreamSynchronize(st));
    c.tab_c[0] = on ? cw : 0;
    c.tab_c[1] = on ? ch : 0;
}

std::unique_ptr<Circuit> circuit_from_zkey(Context& ctx, const std::string& path) {
    ZkeyView zv(path);
    auto c = std::make_unique<Circuit>();
    c->ctx = &ctx;
    c->hdr = zv.hdr;
    c->n_vars = zv.hdr.n_vars;
    c->n_public = zv.hdr.n_public;
    c->n = zv.hdr.domain_size;
    c->log_n = zv.hdr.log_domain;
    ZK_REQUIRE(c->log_n >= 1 && c->log_n <= 27, ZK_E_FORMAT, "domain size must be 2^1..2^27");
    QapCsr csr;
    build_csr(zv.coefs, zv.n_coefs, c->n, c->n_vars, csr);
    upload_csr(*c, csr);
    c->coefs.assign(zv.coefs, zv.coefs + 44ull * zv.n_coefs);
    if (zv.IC) c->ic.assign((const uint64_t*)zv.IC, (const uint64_t*)zv.IC + 8ull * (c->n_public + 1));
    else c->ic.assign(8ull * (c->n_public + 1), 0);
    finish_circuit_bases(*c, zv.A, zv.B1, zv.B2, zv.C, zv.H);
    return c;
}

std::unique_ptr<Circuit> circuit_synthetic(Context& ctx, uint64_t seed, uint32_t m, uint32_t n_pub) {
    ZK_REQUIRE(m >= 1, ZK_E_ARG, "synthetic circuit needs m >= 1");
    SynthR1cs r1;
    synth_r1cs(seed, m, n_pub, r1);
    ZK_REQUIRE(r1.log_n <= 27, ZK_E_ARG, "synthetic circuit too large");
    SynthSetupExps ex;
    synth_setup(r1, ex);
    auto c = std::make_unique<Circuit>();
    c->ctx = &ctx;
    c->n_vars = r1.n_vars;
    c->n_public = n_pub;
    c->n = r1.n;
    c->log_n = r1.log_n;
    synth_coef_bytes(r1, c->coefs);
    QapCsr csr;
    build_csr(c->coefs.data(), (uint32_t)(c->coefs.size() / 44), c->n, c->n_vars, csr);
    upload_csr(*c, csr);
    // CRS points on the device: e * G for every exponent
    ctx.ensure_comb();
    size_t n1 = ex.g1.size() / 4, n2 = ex.g2.size() / 4;
    DevBuf de1(n1 * 32), de2(n2 * 32), dp1(n1 * 64), dp2(n2 * 128);
    ZK_HIP(hipMemcpy(de1.p, ex.g1.data(), n1 * 32, hipMemcpyHostToDevice));
    ZK_HIP(hipMemcpy(de2.p, ex.g2.data(), n2 * 32, hipMemcpyHostToDevice));
    fixed_base_mul(ctx.comb_g1.p, de1.p, n1, dp1.p, false, ctx.stream);
    fixed_base_mul(ctx.comb_g2.p, de2.p, n2, dp2.p, true, ctx.stream);
    ZK_HIP(hipStreamSynchronize(ctx.stream));
    std::vector<uint8_t> h1(n1 * 64), h2(n2 * 128);
    ZK_HIP(hipMemcpy(h1.data(), dp1.p, h1.size(), hipMemcpyDeviceToHost));
    ZK_HIP(hipMemcpy(h2.data(), dp2.p, h2.size(), hipMemcpyDeviceToHost));
    ZkeyHeader& h = c->hdr;
    memcpy(h.q, zkh::FpP::M, 32);
    memcpy(h.r, zkh::FrP::M, 32);
    h.n_vars = c->n_vars;
    h.n_public = n_pub;
    h.domain_size = c->n;
    h.log_domain = c->log_n;
    memcpy(h.alpha1, h1.data() + 0 * 64, 64);
    memcpy(h.beta1, h1.data() + 1 * 64, 64);
    memcpy(h.delta1, h1.data() + 2 * 64, 64);
    memcpy(h.beta2, h2.data() + 0 * 128, 128);
    memcpy(h.gamma2, h2.data() + 1 * 128, 128);
    memcpy(h.delta2, h2.data() + 2 * 128, 128);
    c->ic.assign((const uint64_t*)(h1.data() + ex.off_ic * 64),
                 (const uint64_t*)(h1.data() + ex.off_ic * 64) + 8ull * (n_pub + 1));
    finish_circuit_bases(*c, h1.data() + ex.off_a * 64, h1.data() + ex.off_b1 * 64, h2.data() + ex.off_b2 * 128,
                         h1.data() + ex.off_c * 64, h1.data() + ex.off_h * 64);
    // dependency levels of the constraint rows (device witness generator)
    {
        const uint32_t first_aux = 1 + n_pub;
        std::vector<uint32_t> lev(m);
        uint32_t top = 0;
        for (uint32_t i = 0; i < m; i++) {
            uint32_t mx = 0;
            for (int q = 0; q < 3; q++) {
                const uint32_t wa = r1.a_wire[3ull * i + q], wb = r1.b_wire[3ull * i + q];
                if (wa >= first_aux) mx = std::max(mx, lev[wa - first_aux]);
                if (wb >= first_aux) mx = std::max(mx, lev[wb - first_aux]);
            }
            lev[i] = mx + 1;
            top = std::max(top, lev[i]);
        }
        c->wit_levels.assign(top + 2, 0);
        for (uint32_t i = 0; i < m; i++) c->wit_levels[lev[i] + 1]++;
        for (uint32_t L = 1; L <= top + 1; L++) c->wit_levels[L] += c->wit_levels[L - 1];
        std::vector<uint32_t> pos(c->wit_levels.begin(), c->wit_levels.end() - 1), rows(m);
        for (uint32_t i = 0; i < m; i++) rows[pos[lev[i]]++] = i;
        upload(c->wit_rows, rows.data(), (size_t)m * 4);
    }
    c->synth = std::make_unique<SynthR1cs>(std::move(r1));
    return c;
}

void circuit_witness_device(const Circuit& c, uint64_t pub_seed, void* d_w, hipStream_t st) {
    ZK_REQUIRE(c.synth != nullptr, ZK_E_STATE, "circuit was not built by zk_synth_circuit");
    const SynthR1cs& r1 = *c.synth;
    // wire 0 and the public inputs from the host, the aux wires level by level
    std::vector<uint64_t> head(4ull * (1 + r1.l), 0);
    head[0] = 1;
    const std::vector<Fr> pub = synth_public(pub_seed, r1.l);
    for (uint32_t k = 0; k < r1.l; k++) pub[k].canonical(head.data() + 4ull * (1 + k));
    ZK_HIP(hipMemcpyAsync(d_w, head.data(), head.size() * 8, hipMemcpyHostToDevice, st));
    const uint32_t* rows = c.wit_rows.as<uint32_t>();
    for (size_t L = 1; L + 1 < c.wit_levels.size(); L++)
        witness_rows(c.rowA.as<uint32_t>(), c.colA.as<uint32_t>(), c.valA.p, c.rowB.as<uint32_t>(),
                     c.colB.as<uint32_t>(), c.valB.p, rows + c.wit_levels[L], c.wit_levels[L + 1] - c.wit_levels[L],
                     1 + r1.l, d_w, st);
    ZK_HIP(hipStreamSynchronize(st));   // the host copy of wire 0 and the inputs is freed on return
}

void circuit_write_zkey(const Circuit& c, const std::string& path) {
    const size_t nv = c.n_vars, l = c.n_public;
    std::vector<uint64_t> A(nv * 8), B1(nv * 8), B2(nv * 16), C((nv - l - 1) * 8 + 8), H((size_t)c.n * 8);
    ZK_HIP(hipMemcpy(A.data(), c.A.p, nv * 64, hipMemcpyDeviceToHost));
    ZK_HIP(hipMemcpy(B1.data(), c.B1.p, nv * 64, hipMemcpyDeviceToHost));
    ZK_HIP(hipMemcpy(B2.data(), c.B2.p, nv * 128, hipMemcpyDeviceToHost));
    if (nv - l - 1) ZK_HIP(hipMemcpy(C.data(), (const char*)c.C.p + (l + 1) * 64, (nv - l - 1) * 64, hipMemcpyDeviceToHost));
    ZK_HIP(hipMemcpy(H.data(), c.H.p, (size_t)c.n * 64, hipMemcpyDeviceToHost));
    ZkeyWriterInput in;
    in.hdr = c.hdr;
    in.IC = c.ic.data();
    in.n_coefs = (uint32_t)(c.coefs.size() / 44);
    in.coefs = c.coefs.data();
    in.A = A.data(); in.B1 = B1.data(); in.B2 = B2.data(); in.C = C.data(); in.H = H.data();
    write_zkey(path, in);
}

// ------------------------------------------------------------------ proving
ProveSlot::ProveSlot() {
    for (hipStream_t* p : {&s0, &s1, &s2, &s3, &s4}) ZK_HIP(hipStreamCreateWithFlags(p, hipStreamNonBlocking));
    ZK_HIP(hipEventCreateWithFlags(&ev_scal, hipEventDisableTiming));
    ZK_HIP(hipEventCreateWithFlags(&ev_done1, hipEventDisableTiming));
    for (auto& e : ev_join) ZK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ZK_HIP(hipEventCreateWithFlags(&ev_h, hipEventDisableTiming));
}

void ProveSlot::enable_timing() {
    if (timed) return;
    for (auto& e : ev) ZK_HIP(hipEventCreate(&e));
    for (MsmScratch* m : {&scrA, &scrB1, &scrC, &scrB2, &scrH}) m->enable_timing();
    timed = true;
}

ProveSlot::~ProveSlot() {
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t p : {s0, s1, s2, s3, s4})
        if (p) { (void)hipStreamSynchronize(p); (void)hipStreamDestroy(p); }
    if (ev_scal) (void)hipEventDestroy(ev_scal);
    if (ev_done1) (void)hipEventDestroy(ev_done1);
    if (ev_h) (void)hipEventDestroy(ev_h);
    for (auto& e : ev_join)
        if (e) (void)hipEventDestroy(e);
}

size_t ProveSlot::device_bytes() const {
    size_t b = scal.bytes + aw.bytes + bw.bytes + cw.bytes + t0.bytes + t1.bytes + t2.bytes + h.bytes;
    for (const MsmSort* q : {&sortW, &sortH})
        b += q->half.bytes + q->matrix.bytes + q->mscan.bytes + q->pairs.bytes + q->offsets.bytes + q->entries.bytes +
             q->order.bytes + q->lkeys.bytes + q->chunks.bytes + q->large.bytes + q->counters.bytes + q->cls_off.bytes +
             q->order_a.bytes + q->order_b.bytes + q->okeys.bytes;
    for (const MsmScratch* s : {&scrA, &scrB1, &scrC, &scrB2, &scrH})
        b += s->buckets.bytes + s->r1.bytes + s->t1.bytes + s->r2.bytes + s->t2.bytes + s->partials.bytes;
    return b;
}

static bool is_zero4(const uint64_t* v) { return !(v[0] | v[1] | v[2] | v[3]); }

// the four scalars after the witness: [1, r, s, -r*s], canonical
static void blinding_tail(const uint64_t r[4], const uint64_t s[4], uint64_t ex[16]) {
    memset(ex, 0, 16 * 8);
    ex[0] = 1;
    memcpy(ex + 4, r, 32);
    memcpy(ex + 8, s, 32);
    Fr rm = Fr::to_mont(Fr::from_raw(r)), sm = Fr::to_mont(Fr::from_raw(s));
    (-(rm * sm)).canonical(ex + 12);
}

void ProveSlot::prepare(Context& ctx, const Circuit& c) {
    const size_t nS = (size_t)c.n_vars + 4, n = c.n;
    ctx.ntt_plan(c.log_n);
    scal.ensure(nS * 32);
    for (DevBuf* b : {&aw, &bw, &cw, &t0, &t1, &t2, &h}) b->ensure(n * 32);
}

void ProveSlot::reserve(Context& ctx, const Circuit& c, bool host_witness) {
    const size_t nS = (size_t)c.n_vars + 4, n = c.n;
    prepare(ctx, c);
    if (host_witness) staging.ensure((size_t)c.n_vars * 32);
    const MsmPlan pl = ctx.msm_plan(c);
    const size_t kw = sortW.reserve(nS, pl.cw, pl.mw, c.wire_cls.bytes != 0), kh = sortH.reserve(n, pl.ch, pl.mh);
    const MsmGeom gw1 = msm_geom(nS, pl.cw, false, pl.mw), gw2 = msm_geom(nS, pl.cw, true, pl.mw),
                  gh = msm_geom(n, pl.ch, false, pl.mh);
    for (MsmScratch* q : {&scrA, &scrB1, &scrC}) q->ensure(gw1, kw, false, pl.sp_w1);
    scrB2.ensure(gw2, kw, true, pl.sp_w2);
    scrH.ensure(gh, kh, false, pl.sp_h);
    for (MsmPartials* q : {&parts.a, &parts.b1, &parts.c}) msm_reserve_partials(*q, gw1, false);
    msm_reserve_partials(parts.b2, gw2, true);
    msm_reserve_partials(parts.h, gh, false);
}

void ProveSlot::enqueue(Context& ctx, const Circuit& c, const uint64_t* witness, bool on_dev, const uint64_t r_[4],
                        const uint64_t s_[4], bool alone) {
    circ = &c;
    for (int i = 0; i < 4; i++) { r[i] = r_ ? r_[i] : 0; s[i] = s_ ? s_[i] : 0; }
    ZK_REQUIRE(Fr::is_canonical(r) && Fr::is_canonical(s), ZK_E_ARG, "blinding scalars must be canonical (< r)");
    const size_t nv = c.n_vars;
    prepare(ctx, c);
    if (timed) ZK_HIP(hipEventRecord(ev[0], s0));
    // scalars: witness, then [1, r, s, -r*s]
    if (on_dev) {
        ZK_HIP(hipMemcpyAsync(scal.p, witness, nv * 32, hipMemcpyDeviceToDevice, s0));
    } else {
        // pinned staging makes the upload asynchronous; the previous proof of
        // this slot has finished (finish() ran), so the staging buffer is free
        staging.ensure(nv * 32);
        memcpy(staging.p, witness, nv * 32);
        ZK_HIP(hipMemcpyAsync(scal.p, staging.p, nv * 32, hipMemcpyHostToDevice, s0));
    }
    uint64_t ex[16];
    blinding_tail(r, s, ex);
    write_scalars4((char*)scal.p + nv * 32, ex, s0);
    body(ctx, c, parts, alone);
}

void ProveSlot::body(Context& ctx, const Circuit& c, ProofParts& out, bool alone) {
    const size_t nS = (size_t)c.n_vars + 4, n = c.n;
    NttPlan& plan = ctx.ntt_plan(c.log_n);
    const MsmPlan pl = ctx.msm_plan(c);
    window_w = pl.cw;
    window_h = pl.ch;
    const int sp_w1 = pl.sp_w1, sp_w2 = pl.sp_w2, sp_h = pl.sp_h;
    ZK_HIP(hipEventRecord(ev_scal, s0));
    // witness MSMs first: the G2 one is the longest chain of the proof, so its
    // sort and accumulation are queued before the H pipeline. One shared
    // sort, then the four MSMs read the same sorted entries on their own
    // streams (B2 s4, B1 s2, C s3). Two layouts for the sort and A:
    //  - batches (several proofs in flight): sort and A on s1, so each MSM
    //    has a stream and one MSM's latency-bound reduction overlaps
    //    another's accumulation;
    //  - a proof alone on the device: sort on s2, A after B1 on s2, and s1
    //    left empty. With more streams than hardware queues s1 shares its
    //    queue with s0 (creation order), and work on it holds the H
    //    pipeline, the proof's longest chain, behind it (plain proofs
    //    13.8 -> 13.4 ms; in batches the other slots fill that gap and the
    //    extra MSM stream does better).
    //
    // With wire classes (Circuit::wire_cls) the A MSM and the B1, B2 MSMs
    // read only the entries of their finite bases. An infinite base in the
    // entries would cost its lane a whole mixed addition, since the other
    // lanes of the wave are adding; a quarter of the A' and of the B' bases
    // of the synthetic circuits are infinite (row i draws its wires uniformly
    // from the first ones, so late wires are rarely used).
    alone_layout = alone;
    hipStream_t s_sort = alone ? s2 : s1, s_a = alone ? s2 : s1;
    const uint8_t* cls = c.wire_cls.bytes ? c.wire_cls.as<uint8_t>() : nullptr;
    ZK_HIP(hipStreamWaitEvent(s_sort, ev_scal, 0));
    if (timed) ZK_HIP(hipEventRecord(ev[4], s_sort));
    sortW.run(scal.p, nS, window_w, s_sort, pl.mw, cls);
    ZK_HIP(hipEventRecord(ev_done1, s_sort));
    if (timed) ZK_HIP(hipEventRecord(ev[5], s_sort));
    for (hipStream_t p : {s1, s2, s3, s4})
        if (p != s_sort) ZK_HIP(hipStreamWaitEvent(p, ev_done1, 0));
    msm_launch_pre(sortW, c.B2_29.p, true, scrB2, s4, sp_w2);
    msm_launch_pre(sortW, c.B1_29.p, false, scrB1, s2, sp_w1);
    msm_launch_pre(sortW, c.C29.p, false, scrC, s3, sp_w1);
    msm_launch_pre(sortW, c.A29.p, false, scrA, s_a, sp_w1);
    msm_launch_main(sortW, true, scrB2, out.b2, s4, c.B2_29.p, MSM_RANGE_B);
    if (!alone) msm_launch_main(sortW, false, scrA, out.a, s_a, c.A29.p, MSM_RANGE_A);
    msm_launch_main(sortW, false, scrB1, out.b1, s2, c.B1_29.p, MSM_RANGE_B);
    msm_launch_main(sortW, false, scrC, out.c, s3, c.C29.p);
    if (alone) msm_launch_main(sortW, false, scrA, out.a, s_a, c.A29.p, MSM_RANGE_A);
    // H pipeline on s0
    spmv_abc(c.rowA.as<uint32_t>(), c.colA.as<uint32_t>(), c.valA.p, c.n_rowsA, c.rowB.as<uint32_t>(),
             c.colB.as<uint32_t>(), c.valB.p, c.n_rowsB, scal.p, n, aw.p, bw.p, cw.p, s0);
    if (timed) ZK_HIP(hipEventRecord(ev[1], s0));
    void* abc[3] = {aw.p, bw.p, cw.p};
    void* tmp[3] = {t0.p, t1.p, t2.p};
    ntt_run(plan, abc, tmp, 3, true, nullptr, plan.inv_coset.p, s0);  // coefficients * w_2n^i
    if (timed) ZK_HIP(hipEventRecord(ev[2], s0));
    ntt_run(plan, tmp, abc, 3, false, nullptr, nullptr, s0);          // values at w_2n^(2j+1)
    if (timed) ZK_HIP(hipEventRecord(ev[3], s0));
    h_final(aw.p, bw.p, cw.p, h.p, n, s0);
    sortH.run(h.p, n, window_h, s0, pl.mh);
    msm_launch(sortH, false, scrH, out.h, s0, c.H29.p, sp_h);
    ZK_HIP(hipEventRecord(ev_h, s0));
    hipStream_t br[4] = {s1, s2, s3, s4};
    for (int k = 0; k < 4; k++) {
        ZK_HIP(hipEventRecord(ev_join[k], br[k]));
        ZK_HIP(hipStreamWaitEvent(s0, ev_join[k], 0));
    }
}

void ProveSlot::drain() noexcept {
    for (hipStream_t p : {s0, s1, s2, s3, s4})
        if (p) (void)hipStreamSynchronize(p);
}

bool ProveSlot::device_done() const {
    for (hipStream_t p : {s0, s1, s2, s3, s4}) {
        hipError_t e = hipStreamQuery(p);
        if (e == hipErrorNotReady) return false;
        if (e != hipSuccess) ZK_HIP(e);
    }
    return true;
}

// The buffers reserve(ctx, c, ...) sizes, counted instead of allocated: sl
// null counts a slot not yet created. The NTT plan is the context's (shared
// by every slot) and counts while it does not exist.
size_t ProveSlot::count_bytes(const Context& ctx, const Circuit& c, const ProveSlot* sl) {
    static const DevBuf no_buf;
    static const MsmSort no_sort;
    static const MsmScratch no_scr;
    GrowCount g;
    g.fresh = sl == nullptr;
    const size_t nS = (size_t)c.n_vars + 4, n = c.n;
    g.need(sl ? sl->scal : no_buf, nS * 32);
    if (sl) {
        for (const DevBuf* b : {&sl->aw, &sl->bw, &sl->cw, &sl->t0, &sl->t1, &sl->t2, &sl->h}) g.need(*b, n * 32);
    } else {
        g.bytes += 7 * n * 32;
    }
    const MsmPlan pl = ctx.msm_plan(c);
    const size_t kw = (sl ? sl->sortW : no_sort).count_reserve(nS, pl.cw, g, pl.mw, c.wire_cls.bytes != 0);
    const size_t kh = (sl ? sl->sortH : no_sort).count_reserve(n, pl.ch, g, pl.mh);
    const MsmGeom gw1 = msm_geom(nS, pl.cw, false, pl.mw), gw2 = msm_geom(nS, pl.cw, true, pl.mw),
                  gh = msm_geom(n, pl.ch, false, pl.mh);
    (sl ? sl->scrA : no_scr).count_ensure(gw1, kw, false, pl.sp_w1, g);
    (sl ? sl->scrB1 : no_scr).count_ensure(gw1, kw, false, pl.sp_w1, g);
    (sl ? sl->scrC : no_scr).count_ensure(gw1, kw, false, pl.sp_w1, g);
    (sl ? sl->scrB2 : no_scr).count_ensure(gw2, kw, true, pl.sp_w2, g);
    (sl ? sl->scrH : no_scr).count_ensure(gh, kh, false, pl.sp_h, g);
    if (!ctx.ntt_plans.count(c.log_n)) g.bytes += NttPlan::bytes_for(c.log_n);
    return g.bytes;
}

size_t ProveSlot::bytes_needed(const Context& ctx, const Circuit& c) { return count_bytes(ctx, c, nullptr); }

size_t ProveSlot::grow_bytes(const Context& ctx, const Circuit& c) const { return count_bytes(ctx, c, this); }

template <class T>
static void put_aff(const zkh::Xyzz<T>& p, uint64_t* dst) {
    const int lw = sizeof(T) / 8;
    zkh::Aff<T> a = zkh::xyzz_to_aff(p);
    if (a.inf) { memset(dst, 0, 16 * lw); return; }
    memcpy(dst, &a.x, 8 * lw);
    memcpy(dst + lw, &a.y, 8 * lw);
}

double assemble_proof(const ProofParts& p, const uint64_t r[4], const uint64_t s[4], zk_proof* out,
                      const hipEvent_t* ready) {
    using clk = std::chrono::steady_clock;
    using G1 = zkh::Xyzz<zkh::Fp>;
    std::atomic<int64_t> t_last{0};   // latest completion seen, ns since t0
    const auto t0 = clk::now();
    auto wait = [&](int k) {
        if (ready && ready[k]) ZK_HIP(hipEventSynchronize(ready[k]));
        const int64_t t = std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now() - t0).count();
        int64_t prev = t_last.load();
        while (t > prev && !t_last.compare_exchange_weak(prev, t)) {}
    };
    // pi_A and s*pi_A; r*B1; C' + H: one thread each. pi_B (G2, the longest) here.
    auto fa = std::async(std::launch::async, [&] {
        wait(0);
        G1 A = msm_finish_g1(p.a);
        return std::make_pair(A, is_zero4(s) ? G1::infinity() : zkh::xyzz_mul(A, s));
    });
    auto fb = std::async(std::launch::async, [&] {
        wait(1);
        return is_zero4(r) ? G1::infinity() : zkh::xyzz_mul(msm_finish_g1(p.b1), r);
    });
    auto fc = std::async(std::launch::async, [&] {
        wait(2);
        G1 Cw = msm_finish_g1(p.c);
        wait(4);
        return zkh::xyzz_add(Cw, msm_finish_g1(p.h));
    });
    wait(3);
    put_aff(msm_finish_g2(p.b2), out->pi_b);
    auto A = fa.get();
    put_aff(A.first, out->pi_a);
    G1 C = zkh::xyzz_add(zkh::xyzz_add(fc.get(), A.second), fb.get());
    put_aff(C, out->pi_c);
    const double total = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    return total - t_last.load() / 1e6;
}

void ProveSlot::finish(zk_proof* out) {
    // A on s1 (ev_join[0]), or after B1 on s2 (ev_join[1]) in the layout of a
    // proof alone; B1 s2, C s3, B2 s4, H s0
    const hipEvent_t ready[5] = {alone_layout ? ev_join[1] : ev_join[0], ev_join[1], ev_join[2], ev_join[3], ev_h};
    const double tail_ms = assemble_proof(parts, r, s, out, ready);
    ZK_HIP(hipStreamSynchronize(s0));   // s1..s4 are joined into s0
    if (timed) {
        collect_stats();
        stats[23] = tail_ms;
    }
}

void ProveSlot::collect_stats() {
    auto el = [](hipEvent_t a, hipEvent_t b) {
        float ms = 0;
        ZK_HIP(hipEventElapsedTime(&ms, a, b));
        return (double)ms;
    };
    for (double& d : stats) d = 0;
    stats[0] = el(ev[0], ev[1]);   // scalar upload + SpMV
    stats[1] = el(ev[1], ev[2]);   // 3 x (iNTT + coset), batched
    stats[2] = el(ev[2], ev[3]);   // 3 x NTT, batched
    stats[3] = el(ev[3], scrH.ev_acc0);  // h + sort of h
    stats[4] = el(ev[4], ev[5]);   // shared sort of [w | 1, r, s, -rs]
    MsmScratch* m[5] = {&scrA, &scrB1, &scrC, &scrB2, &scrH};
    for (int i = 0; i < 5; i++) {
        stats[5 + i] = el(m[i]->ev_acc0, m[i]->ev_acc1);
        stats[10 + i] = el(m[i]->ev_acc0, m[i]->ev_end);
    }
    uint32_t ew = 0, eh = 0;
    size_t nbw = (size_t)sortW.g.Wr * sortW.g.K, nbh = (size_t)sortH.g.Wr * sortH.g.K;
    ZK_HIP(hipMemcpy(&ew, sortW.offsets.as<uint32_t>() + nbw, 4, hipMemcpyDeviceToHost));
    ZK_HIP(hipMemcpy(&eh, sortH.offsets.as<uint32_t>() + nbh, 4, hipMemcpyDeviceToHost));
    stats[15] = ew;
    stats[16] = eh;
    stats[17] = sortW.g.c;
    stats[18] = sortH.g.c;
    stats[19] = sortW.g.W;
    stats[20] = sortH.g.W;
    stats[21] = el(ev[0], scrH.ev_end);   // stream 0 span
    stats[22] = 0;                         // witness-MSM span (latest of the four streams)
    for (MsmScratch* q : {&scrA, &scrB1, &scrC, &scrB2}) stats[22] = std::max(stats[22], el(ev[4], q->ev_end));
}

// ------------------------------------------------------------------ batch graph
BatchGraph::BatchGraph(Context& c, int count, const Circuit* const* cs, const void* const* d_wits, int nslots)
    : ctx(c) {
    ZK_REQUIRE(count >= 1, ZK_E_ARG, "empty batch");
    ZK_REQUIRE(nslots >= 1 && nslots <= 16, ZK_E_ARG, "slots must be 1..16");
    for (int i = 0; i < count; i++) {
        ZK_REQUIRE(cs[i] && d_wits[i], ZK_E_ARG, "NULL circuit or witness in batch");
        circs.push_back(cs[i]);
        wits.push_back(d_wits[i]);
    }
    const int S = std::min(nslots, count);
    for (int k = 0; k < S; k++) slots.push_back(std::make_unique<ProveSlot>());
    parts.reset(new ProofParts[count]);
    params.ensure((size_t)count * 128);
    ZK_HIP(hipStreamCreateWithFlags(&origin, hipStreamNonBlocking));
    ZK_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    ev_end.resize(S);
    for (auto& e : ev_end) ZK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // one eager pass sizes every buffer (slots, sorts, scratch, partials, NTT
    // plans) so that the captured pass allocates nothing
    fill_params(nullptr, nullptr);
    record(false);
    ZK_HIP(hipStreamSynchronize(origin));
    ZK_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal));
    try {
        record(true);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(origin, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    ZK_HIP(hipStreamEndCapture(origin, &graph));
    ZK_HIP(hipGraphGetNodes(graph, nullptr, &n_nodes));
    ZK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
}

BatchGraph::~BatchGraph() {
    if (origin) (void)hipStreamSynchronize(origin);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& e : ev_end)
        if (e) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    slots.clear();
    if (origin) (void)hipStreamDestroy(origin);
}

void BatchGraph::fill_params(const uint64_t* r, const uint64_t* s) {
    static const uint64_t zero[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < circs.size(); i++) {
        const uint64_t* ri = r ? r + 4 * i : zero;
        const uint64_t* si = s ? s + 4 * i : zero;
        ZK_REQUIRE(Fr::is_canonical(ri) && Fr::is_canonical(si), ZK_E_ARG, "blinding scalars must be canonical (< r)");
        blinding_tail(ri, si, params.as<uint64_t>() + 16 * i);
    }
}

// queue the whole batch: proof i on slot i % S; a slot's proofs are ordered by
// its s0 (each proof's end joins s1..s4 into s0)
void BatchGraph::record(bool capture) {
    const size_t S = slots.size();
    ZK_HIP(hipEventRecord(ev_fork, origin));
    for (auto& sl : slots) ZK_HIP(hipStreamWaitEvent(sl->s0, ev_fork, 0));
    for (size_t i = 0; i < circs.size(); i++) {
        ProveSlot& sl = *slots[i % S];
        const Circuit& c = *circs[i];
        if (!capture) {
            // sizing pass: buffers may grow (free + allocate), so the slot's
            // previous proof must have drained first
            if (i >= S) ZK_HIP(hipStreamSynchronize(sl.s0));
            sl.prepare(ctx, c);
        }
        const size_t nv = c.n_vars;
        ZK_HIP(hipMemcpyAsync(sl.scal.p, wits[i], nv * 32, hipMemcpyDeviceToDevice, sl.s0));
        ZK_HIP(hipMemcpyAsync((char*)sl.scal.p + nv * 32, params.as<uint64_t>() + 16 * i, 128, hipMemcpyHostToDevice,
                              sl.s0));
        sl.body(ctx, c, parts[i]);
    }
    for (size_t k = 0; k < S; k++) {
        ZK_HIP(hipEventRecord(ev_end[k], slots[k]->s0));
        ZK_HIP(hipStreamWaitEvent(origin, ev_end[k], 0));
    }
}

void BatchGraph::run(const uint64_t* r, const uint64_t* s, zk_proof* out) {
    fill_params(r, s);
    ZK_HIP(hipGraphLaunch(exec, origin));
    ZK_HIP(hipStreamSynchronize(origin));
    // host tails in parallel (each proof's partials are its own)
    const size_t count = circs.size();
    unsigned nt = std::max(1u, std::min<unsigned>(8u, std::thread::hardware_concurrency()));
    nt = (unsigned)std::min<size_t>(nt, count);
    std::vector<std::thread> th;
    std::vector<std::string> errs(nt);
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            try {
                for (size_t i = t; i < count; i += nt)
                    assemble_proof(parts[i], params.as<uint64_t>() + 16 * i + 4, params.as<uint64_t>() + 16 * i + 8,
                                   &out[i]);
            } catch (const std::exception& e) {
                errs[t] = e.what();
            }
        });
    for (auto& x : th) x.join();
    for (auto& e : errs) ZK_REQUIRE(e.empty(), ZK_E_INTERNAL, e);
}

// ------------------------------------------------------------------ serialization
static std::string dec_of_mont(const uint64_t* m) {
    uint64_t c[4];
    Fp::from_raw(m).canonical(c);
    return zkh::to_decimal(c);
}

std::string proof_to_json(const zk_proof& p) {
    char buf[2048];
    snprintf(buf, sizeof buf,
             "{\n"
             " \"pi_a\": [\"%s\",\"%s\",\"1\"],\n"
             " \"pi_b\": [[\"%s\",\"%s\"],[\"%s\",\"%s\"],[\"1\",\"0\"]],\n"
             " \"pi_c\": [\"%s\",\"%s\",\"1\"],\n"
             " \"protocol\": \"groth16\",\n"
             " \"curve\": \"bn128\"\n"
             "}",
             dec_of_mont(p.pi_a).c_str(), dec_of_mont(p.pi_a + 4).c_str(), dec_of_mont(p.pi_b).c_str(),
             dec_of_mont(p.pi_b + 4).c_str(), dec_of_mont(p.pi_b + 8).c_str(), dec_of_mont(p.pi_b + 12).c_str(),
             dec_of_mont(p.pi_c).c_str(), dec_of_mont(p.pi_c + 4).c_str());
    return buf;
}

static void be32(const uint64_t* mont, uint8_t* out) {
    uint64_t c[4];
    Fp::from_raw(mont).canonical(c);
    for (int i = 0; i < 4; i++) {
        uint64_t v = c[3 - i];
        for (int j = 0; j < 8; j++) out[8 * i + j] = (uint8_t)(v >> (8 * (7 - j)));
    }
}

void proof_to_calldata(const zk_proof& p, uint8_t out[256]) {
    be32(p.pi_a, out);
    be32(p.pi_a + 4, out + 32);
    be32(p.pi_b + 4, out + 64);   // x.a1
    be32(p.pi_b, out + 96);       // x.a0
    be32(p.pi_b + 12, out + 128); // y.a1
    be32(p.pi_b + 8, out + 160);  // y.a0
    be32(p.pi_c, out + 192);
    be32(p.pi_c + 4, out + 224);
}

std::string public_json(const uint64_t* w, uint32_t n_public) {
    std::string s = "[\n";
    for (uint32_t i = 0; i < n_public; i++) {
        s += " \"" + zkh::to_decimal(w + 4ull * (i + 1)) + "\"";
        s += (i + 1 < n_public) ? ",\n" : "\n";
    }
    s += "]\n";
    return s;
}

}  // namespace zk
