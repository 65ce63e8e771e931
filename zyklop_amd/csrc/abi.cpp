// extern "C" entry points (include/zk_abi.h). Exceptions stop here and become
// status codes plus a per-thread message.
#include <string.h>

#include <set>
#include <string>

#include "../../include/zk_abi.h"
#include "fieldtest.hpp"
#include "glv.hpp"
#include "msm.hpp"
#include "ntt.hpp"
#include "pairing.hpp"
#include "prover.hpp"
#include "qap.hpp"
#include "server.hpp"
#include "synth.hpp"
#include "verify.hpp"
#include "zkey.hpp"

struct zk_server;

struct zk_ctx {
    zk::Context impl;
    std::vector<std::unique_ptr<zk::ProveSlot>> batch_slots;
    // servers created on this context: destroying the context finishes and
    // detaches them first (their worker threads use the context)
    std::mutex servers_mu;
    std::set<zk_server*> servers;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    float last_ms[5] = {0, 0, 0, 0, 0};
    bool profiling = false;
    int batch_slots_n = 3;
    explicit zk_ctx(int dev) : impl(dev) {
        ZK_HIP(hipEventCreate(&ev0));
        ZK_HIP(hipEventCreate(&ev1));
        ZK_HIP(hipEventCreate(&ev2));
    }
    ~zk_ctx() {
        batch_slots.clear();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ev2) (void)hipEventDestroy(ev2);
    }
};

struct zk_circuit {
    std::unique_ptr<zk::Circuit> impl;
};

struct zk_server {
    zk_ctx* ctx = nullptr;               // null once the context was destroyed
    std::unique_ptr<zk::Server> impl;    // null once the context was destroyed
};

struct zk_batch {
    zk_ctx* ctx = nullptr;
    std::unique_ptr<zk::BatchGraph> impl;
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        g_err.clear();
        return ZK_OK;
    } catch (const zk::Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host out of memory";
        return ZK_E_NOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return ZK_E_INTERNAL;
    } catch (...) {
        g_err = "unknown error";
        return ZK_E_INTERNAL;
    }
}

void need(bool ok, const char* what) {
    if (!ok) throw zk::Error(ZK_E_ARG, what);
}

void check_scalars(const uint64_t* sc, size_t n) {
    for (size_t i = 0; i < n; i++)
        if (!zkh::Fr::is_canonical(sc + 4 * i))
            throw zk::Error(ZK_E_ARG, "scalar " + std::to_string(i) + " is not canonical (>= r)");
}

zk::ProveSlot& ctx_slot(zk_ctx* ctx) {
    if (!ctx->impl.slot) ctx->impl.slot = std::make_unique<zk::ProveSlot>();
    return *ctx->impl.slot;
}

// Several workspace slots in flight: while the host finishes proof i, the
// device already runs the next ones queued on the other slots' streams.
void run_batch(zk_ctx* ctx, int count, const zk_circuit* const* circuits, const uint64_t* const* witnesses, bool on_dev,
               const uint64_t* r, const uint64_t* s, zk_proof* out) {
    const int nslots = ctx->batch_slots_n;
    while ((int)ctx->batch_slots.size() < nslots) ctx->batch_slots.push_back(std::make_unique<zk::ProveSlot>());
    for (int i = 0; i < count; i++) need(circuits[i] && witnesses[i], "NULL circuit or witness");
    int next = 0;
    const int inflight = std::min(count, nslots);
    // every slot sized for every circuit of the batch before any proof is
    // queued: a buffer that grew mid-batch would free and allocate device
    // memory, which waits for the work of all the other slots
    std::set<const zk::Circuit*> distinct;
    for (int i = 0; i < count; i++) distinct.insert(circuits[i]->impl.get());
    for (int k = 0; k < inflight; k++)
        for (const zk::Circuit* c : distinct) ctx->batch_slots[k]->reserve(ctx->impl, *c, !on_dev);
    try {
        for (; next < inflight; next++)
            ctx->batch_slots[next % nslots]->enqueue(ctx->impl, *circuits[next]->impl, witnesses[next], on_dev,
                                                     r ? r + 4 * next : nullptr, s ? s + 4 * next : nullptr);
        for (int i = 0; i < count; i++) {
            zk::ProveSlot& sl = *ctx->batch_slots[i % nslots];
            sl.finish(&out[i]);
            if (next < count) {
                sl.enqueue(ctx->impl, *circuits[next]->impl, witnesses[next], on_dev, r ? r + 4 * next : nullptr,
                           s ? s + 4 * next : nullptr);
                next++;
            }
        }
    } catch (...) {
        // drain what the other slots still have queued before reporting, so
        // no device work of this batch outlives the call
        for (int k = 0; k < inflight; k++) ctx->batch_slots[k]->drain();
        throw;
    }
}

template <class HT>
void put_result(const zkh::Xyzz<HT>& p, uint64_t* out, int* is_inf) {
    const int lw = sizeof(HT) / 8;
    zkh::Aff<HT> a = zkh::xyzz_to_aff(p);
    if (is_inf) *is_inf = a.inf ? 1 : 0;
    if (a.inf) { memset(out, 0, 16 * lw); return; }
    memcpy(out, &a.x, 8 * lw);
    memcpy(out + lw, &a.y, 8 * lw);
}

// MSM over device points; slots29, when given, are the points' 29-bit slots
// already made (a circuit's resident arrays, W window tables when merged)
// cls / range: scalar classes and the part of every bucket to read, as a
// proof's witness MSMs do (MsmSort::run, msm_launch_main); split: threads per
// bucket, 0 for msm_default_split
void msm_device(zk_ctx* ctx, const void* d_points, const void* d_scalars, size_t n, int window, bool g2, uint64_t* out,
                int* is_inf, const void* slots29 = nullptr, bool merged = false, const uint8_t* cls = nullptr,
                int range = zk::MSM_RANGE_ALL, int split = 0) {
    zk::Context& C = ctx->impl;
    hipStream_t st = C.stream;
    zk::MsmSort sort;
    zk::MsmScratch scr;
    zk::MsmPartials part;
    scr.enable_timing();
    // buckets are accumulated in the 29-bit-limb form: convert the points first
    zk::DevBuf p29;
    if (n && !slots29) {
        p29.ensure(n * (g2 ? zk::PT29Q_BYTES : zk::PT29_BYTES));
        zk::msm_points_to29(d_points, n, p29.p, g2, st);
    }
    ZK_HIP(hipEventRecord(ctx->ev0, st));
    sort.run(d_scalars, n, window, st, merged, cls);
    ZK_HIP(hipEventRecord(ctx->ev1, st));
    const void* pts = slots29 ? slots29 : p29.p;
    zk::msm_launch_pre(sort, pts, g2, scr, st, split > 0 ? split : zk::msm_default_split(n, sort.g.c, g2, merged));
    zk::msm_launch_main(sort, g2, scr, part, st, pts, range);
    ZK_HIP(hipEventRecord(ctx->ev2, st));
    ZK_HIP(hipStreamSynchronize(st));
    ZK_HIP(hipEventElapsedTime(&ctx->last_ms[0], ctx->ev0, ctx->ev2));
    ZK_HIP(hipEventElapsedTime(&ctx->last_ms[1], ctx->ev0, ctx->ev1));
    ZK_HIP(hipEventElapsedTime(&ctx->last_ms[2], scr.ev_acc0, scr.ev_acc1));
    ZK_HIP(hipEventElapsedTime(&ctx->last_ms[4], scr.ev_acc1, scr.ev_end));
    // entries the accumulation read: the sort's, or the parts of the range
    const size_t nb = (size_t)sort.g.Wr * sort.g.K;
    uint32_t entries = 0;
    ZK_HIP(hipMemcpy(&entries, sort.offsets.as<uint32_t>() + nb, 4, hipMemcpyDeviceToHost));
    if (sort.classes && range != zk::MSM_RANGE_ALL) {
        std::vector<uint32_t> off(nb + 1), co(3 * nb);
        ZK_HIP(hipMemcpy(off.data(), sort.offsets.p, 4 * (nb + 1), hipMemcpyDeviceToHost));
        ZK_HIP(hipMemcpy(co.data(), sort.cls_off.p, 12 * nb, hipMemcpyDeviceToHost));
        entries = 0;
        for (size_t b = 0; b < nb; b++)
            entries += range == zk::MSM_RANGE_A ? co[3 * b + 2] - co[3 * b] : co[3 * b + 1] - off[b];
    }
    ctx->last_ms[3] = (float)entries;
    if (g2) put_result(zk::msm_finish_g2(part), out, is_inf);
    else put_result(zk::msm_finish_g1(part), out, is_inf);
}

void ntt_device(zk_ctx* ctx, void* const* in, void* const* out, int batch, int log_n, int kind) {
    need(kind >= 0 && kind <= 3, "NTT kind must be 0..3");
    zk::Context& C = ctx->impl;
    zk::NttPlan& p = C.ntt_plan(log_n);
    const void* pre = kind == 2 ? p.coset_fwd.p : nullptr;
    const void* post = kind == 1 ? p.inv_plain.p : kind == 3 ? p.inv_uncoset.p : nullptr;
    ZK_HIP(hipEventRecord(ctx->ev0, C.stream));
    zk::ntt_run(p, in, out, batch, kind == 1 || kind == 3, pre, post, C.stream);
    ZK_HIP(hipEventRecord(ctx->ev2, C.stream));
    ZK_HIP(hipStreamSynchronize(C.stream));
    ZK_HIP(hipEventElapsedTime(&ctx->last_ms[0], ctx->ev0, ctx->ev2));
    ctx->last_ms[1] = ctx->last_ms[2] = ctx->last_ms[3] = ctx->last_ms[4] = 0;
}
}  // namespace

extern "C" {

const char* zk_last_error(void) { return g_err.c_str(); }
const char* zk_version(void) { return "zyklop_amd 0.1 (gfx950)"; }

int zk_ctx_create(int device, zk_ctx** out) {
    return guard([&] {
        need(out != nullptr, "out is NULL");
        *out = new zk_ctx(device);
    });
}

void zk_ctx_destroy(zk_ctx* ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::mutex> lk(ctx->servers_mu);
        for (zk_server* s : ctx->servers) {
            s->impl.reset();   // finishes every request the server accepted
            s->ctx = nullptr;
        }
        ctx->servers.clear();
    }
    delete ctx;
}

int zk_circuit_load(zk_ctx* ctx, const char* zkey_path, zk_circuit** out) {
    return guard([&] {
        need(ctx && zkey_path && out, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        auto c = new zk_circuit;
        try {
            c->impl = zk::circuit_from_zkey(ctx->impl, zkey_path);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

void zk_circuit_free(zk_circuit* c) { delete c; }

int zk_circuit_info(const zk_circuit* c, uint64_t info[7]) {
    return guard([&] {
        need(c && info, "NULL argument");
        const zk::Circuit& k = *c->impl;
        info[0] = k.n_vars; info[1] = k.n_public; info[2] = k.n; info[3] = (uint64_t)k.log_n;
        info[4] = k.nnzA; info[5] = k.nnzB; info[6] = k.device_bytes();
    });
}

int zk_circuit_write_zkey(const zk_circuit* c, const char* path) {
    return guard([&] {
        need(c && path, "NULL argument");
        zk::circuit_write_zkey(*c->impl, path);
    });
}

int zk_prove(zk_ctx* ctx, const zk_circuit* c, const uint64_t* witness, uint32_t n_witness, const uint64_t r[4],
             const uint64_t s[4], zk_proof* out) {
    return guard([&] {
        need(ctx && c && witness && out, "NULL argument");
        need(n_witness >= c->impl->n_vars, "witness shorter than nVars");
        check_scalars(witness, c->impl->n_vars);
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        zk::ProveSlot& sl = ctx_slot(ctx);
        sl.enqueue(ctx->impl, *c->impl, witness, false, r, s, true);
        sl.finish(out);
    });
}

int zk_prove_device(zk_ctx* ctx, const zk_circuit* c, const void* d_witness, const uint64_t r[4], const uint64_t s[4],
                    zk_proof* out) {
    return guard([&] {
        need(ctx && c && d_witness && out, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        zk::ProveSlot& sl = ctx_slot(ctx);
        sl.enqueue(ctx->impl, *c->impl, (const uint64_t*)d_witness, true, r, s, true);
        sl.finish(out);
    });
}

int zk_set_windows(zk_ctx* ctx, int window_witness, int window_h) {
    return guard([&] {
        need(ctx != nullptr, "NULL ctx");
        need(window_witness == 0 || (window_witness >= 4 && window_witness <= 22), "window must be 0 or 4..22");
        need(window_h == 0 || (window_h >= 4 && window_h <= 22), "window must be 0 or 4..22");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ctx->impl.force_w = window_witness;
        ctx->impl.force_h = window_h;
    });
}

int zk_set_window_tables(zk_ctx* ctx, int mode) {
    return guard([&] {
        need(ctx != nullptr, "NULL ctx");
        need(mode >= -1 && mode <= 1, "mode must be -1 (auto), 0 (off) or 1 (on)");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ctx->impl.tables_mode = mode;
    });
}

int zk_circuit_window_tables(zk_ctx* ctx, zk_circuit* c, int mode, int windows[2]) {
    return guard([&] {
        need(ctx && c, "NULL argument");
        need(mode >= -2 && mode <= 1, "mode must be -2 (query), -1 (auto), 0 (off) or 1 (on)");
        need(c->impl->ctx == &ctx->impl, "circuit belongs to another context");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        if (mode != -2) zk::circuit_set_tables(*c->impl, mode);
        if (windows) { windows[0] = c->impl->tab_c[0]; windows[1] = c->impl->tab_c[1]; }
    });
}

int zk_autotune(zk_ctx* ctx, const zk_circuit* c, int force, int windows[2]) {
    return guard([&] {
        need(ctx && c, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        int w[2] = {0, 0};
        ctx->impl.autotune(*c->impl, force != 0, 2, w);
        if (windows) { windows[0] = w[0]; windows[1] = w[1]; }
    });
}

int zk_tune_get(zk_ctx* ctx, int kind, uint64_t n, int* window) {
    return guard([&] {
        need(ctx && window, "NULL argument");
        need(kind == 0 || kind == 1, "kind must be 0 (witness) or 1 (h)");
        need(n >= 1, "n must be >= 1");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        *window = ctx->impl.window_for(kind, n);
    });
}

int zk_tune_load(zk_ctx* ctx, const char* path, int* n_entries) {
    return guard([&] {
        need(ctx && path, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        int k = ctx->impl.tune_load(path);
        if (n_entries) *n_entries = k;
    });
}

int zk_tune_save(zk_ctx* ctx, const char* path) {
    return guard([&] {
        need(ctx && path, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ctx->impl.tune_save(path);
    });
}

int zk_batch_graph_create(zk_ctx* ctx, int count, const zk_circuit* const* circuits, const void* const* d_witnesses,
                          int slots, zk_batch** out) {
    return guard([&] {
        need(ctx && circuits && d_witnesses && out, "NULL argument");
        need(count >= 1, "count must be >= 1");
        *out = nullptr;
        // Disabled: the first run of graph creation on the MI355X ended in a
        // host segmentation fault inside this call whose cause has not been
        // found; the stream-ordered batch (zk_prove_batch_device) is the
        // batched path until it is.
        throw zk::Error(ZK_E_STATE, "captured batch graphs are disabled in this build (zk_prove_batch_device is the "
                                    "batched path)");
        std::vector<const zk::Circuit*> cs(count);
        for (int i = 0; i < count; i++) {
            need(circuits[i] != nullptr, "NULL circuit");
            cs[i] = circuits[i]->impl.get();
        }
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        auto b = std::make_unique<zk_batch>();
        b->ctx = ctx;
        b->impl = std::make_unique<zk::BatchGraph>(ctx->impl, count, cs.data(), d_witnesses, slots);
        *out = b.release();
    });
}

int zk_batch_graph_run(zk_batch* b, const uint64_t* r, const uint64_t* s, zk_proof* out) {
    return guard([&] {
        need(b && out, "NULL argument");
        std::lock_guard<std::mutex> lk(b->ctx->impl.mu);
        ZK_HIP(hipSetDevice(b->ctx->impl.device));
        b->impl->run(r, s, out);
    });
}

int zk_batch_graph_info(const zk_batch* b, uint64_t* n_nodes, int* n_slots) {
    return guard([&] {
        need(b != nullptr, "NULL argument");
        if (n_nodes) *n_nodes = b->impl->n_nodes;
        if (n_slots) *n_slots = (int)b->impl->slots.size();
    });
}

void zk_batch_graph_destroy(zk_batch* b) {
    if (!b) return;
    std::lock_guard<std::mutex> lk(b->ctx->impl.mu);
    (void)hipSetDevice(b->ctx->impl.device);
    delete b;
}

int zk_prove_files(zk_ctx* ctx, const char* zkey_path, const char* wtns_path, const uint64_t r[4], const uint64_t s[4],
                   zk_proof* out) {
    return guard([&] {
        need(ctx && zkey_path && wtns_path && out, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        auto c = zk::circuit_from_zkey(ctx->impl, zkey_path);
        uint32_t nw = 0;
        std::vector<uint64_t> w = zk::read_wtns(wtns_path, &nw);
        ZK_REQUIRE(nw >= c->n_vars, ZK_E_FORMAT,
                   "witness has " + std::to_string(nw) + " entries, need " + std::to_string(c->n_vars));
        check_scalars(w.data(), c->n_vars);
        zk::ProveSlot& sl = ctx_slot(ctx);
        sl.enqueue(ctx->impl, *c, w.data(), false, r, s, true);
        sl.finish(out);
    });
}

int zk_prove_batch(zk_ctx* ctx, int count, const zk_circuit* const* circuits, const uint64_t* const* witnesses,
                   const uint64_t* r, const uint64_t* s, zk_proof* out) {
    return guard([&] {
        need(ctx && circuits && witnesses && out && count >= 0, "bad argument");
        for (int i = 0; i < count; i++) {
            need(circuits[i] && witnesses[i], "NULL circuit or witness");
            check_scalars(witnesses[i], circuits[i]->impl->n_vars);
        }
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        run_batch(ctx, count, circuits, witnesses, false, r, s, out);
    });
}

int zk_prove_batch_device(zk_ctx* ctx, int count, const zk_circuit* const* circuits, const void* const* d_witnesses,
                          const uint64_t* r, const uint64_t* s, zk_proof* out) {
    return guard([&] {
        need(ctx && circuits && d_witnesses && out && count >= 0, "bad argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        run_batch(ctx, count, circuits, (const uint64_t* const*)d_witnesses, true, r, s, out);
    });
}

int zk_set_batch_slots(zk_ctx* ctx, int n) {
    return guard([&] {
        need(ctx != nullptr, "NULL ctx");
        need(n >= 1 && n <= 16, "slots must be 1..16");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ctx->batch_slots_n = n;
        if ((int)ctx->batch_slots.size() > n) ctx->batch_slots.resize(n);
    });
}

int zk_set_profiling(zk_ctx* ctx, int on) {
    return guard([&] {
        need(ctx != nullptr, "NULL ctx");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ctx->profiling = on != 0;
        if (on) ctx_slot(ctx).enable_timing();
    });
}

int zk_prove_stats(zk_ctx* ctx, double* out, int n) {
    return guard([&] {
        need(ctx && out, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        zk::ProveSlot& sl = ctx_slot(ctx);
        need(sl.timed, "profiling is off (zk_set_profiling)");
        for (int i = 0; i < n && i < zk::ProveSlot::N_STATS; i++) out[i] = sl.stats[i];
    });
}

int zk_server_create(zk_ctx* ctx, int max_slots, uint64_t reserve_bytes, zk_server** out) {
    return guard([&] {
        need(ctx && out, "NULL argument");
        auto s = new zk_server;
        try {
            s->impl = std::make_unique<zk::Server>(ctx->impl, max_slots, (size_t)reserve_bytes);
        } catch (...) {
            delete s;
            throw;
        }
        s->ctx = ctx;
        std::lock_guard<std::mutex> lk(ctx->servers_mu);
        ctx->servers.insert(s);
        *out = s;
    });
}

void zk_server_destroy(zk_server* s) {
    if (!s) return;
    if (s->ctx) {
        std::lock_guard<std::mutex> lk(s->ctx->servers_mu);
        s->ctx->servers.erase(s);
    }
    delete s;
}

static zk::Server& live(zk_server* s) {
    need(s != nullptr, "NULL server");
    if (!s->impl) throw zk::Error(ZK_E_STATE, "the server's context was destroyed");
    return *s->impl;
}

int zk_server_prepare(zk_server* s, const zk_circuit* const* circuits, int count, int* n_slots) {
    return guard([&] {
        need(circuits != nullptr, "NULL circuits");
        std::vector<const zk::Circuit*> cs;
        for (int i = 0; i < count; i++) {
            need(circuits[i] != nullptr, "NULL circuit");
            cs.push_back(circuits[i]->impl.get());
        }
        int k = live(s).prepare(cs.data(), count);
        if (n_slots) *n_slots = k;
    });
}

int zk_server_submit(zk_server* s, const zk_circuit* c, const uint64_t* witness, uint32_t n_witness, const uint64_t r[4],
                     const uint64_t sb[4], uint64_t* ticket) {
    return guard([&] {
        need(c && witness && ticket, "NULL argument");
        *ticket = live(s).submit(*c->impl, witness, n_witness, r, sb);
    });
}

int zk_server_submit_device(zk_server* s, const zk_circuit* c, const void* d_witness, const uint64_t r[4],
                            const uint64_t sb[4], uint64_t* ticket) {
    return guard([&] {
        need(c && d_witness && ticket, "NULL argument");
        *ticket = live(s).submit_device(*c->impl, d_witness, r, sb);
    });
}

static int server_result(int status, const std::string& err) {
    if (status != ZK_OK) throw zk::Error(status, err);
    return ZK_OK;
}

int zk_server_poll(zk_server* s, uint64_t ticket, int* done, zk_proof* out, double times[3]) {
    return guard([&] {
        need(done != nullptr, "NULL argument");
        int st = ZK_OK;
        std::string err;
        *done = live(s).poll(ticket, out, times, &st, &err) ? 1 : 0;
        if (*done) server_result(st, err);
    });
}

int zk_server_wait(zk_server* s, uint64_t ticket, zk_proof* out, double times[3]) {
    return guard([&] {
        int st = ZK_OK;
        std::string err;
        live(s).wait(ticket, out, times, &st, &err);
        server_result(st, err);
    });
}

int zk_server_cancel(zk_server* s, uint64_t ticket, int* cancelled) {
    return guard([&] {
        need(cancelled != nullptr, "NULL argument");
        *cancelled = live(s).cancel(ticket) ? 1 : 0;
    });
}

int zk_server_stats(zk_server* s, double out[8]) {
    return guard([&] {
        need(out != nullptr, "NULL argument");
        live(s).stats(out);
    });
}

int zk_proof_to_json(const zk_proof* p, char* out, size_t out_size) {
    return guard([&] {
        need(p && out && out_size > 0, "bad argument");
        std::string j = zk::proof_to_json(*p);
        need(j.size() + 1 <= out_size, "output buffer too small");
        memcpy(out, j.c_str(), j.size() + 1);
    });
}

int zk_proof_to_calldata(const zk_proof* p, uint8_t* out, size_t out_size) {
    return guard([&] {
        need(p && out, "NULL argument");
        need(out_size >= 256, "calldata buffer must hold 256 bytes");
        zk::proof_to_calldata(*p, out);
    });
}

int zk_public_json(const char* wtns_path, uint32_t n_public, char* out, size_t out_size) {
    return guard([&] {
        need(wtns_path && out, "NULL argument");
        uint32_t nw = 0;
        std::vector<uint64_t> w = zk::read_wtns(wtns_path, &nw);
        ZK_REQUIRE(nw >= n_public + 1, ZK_E_FORMAT, "witness has fewer than nPublic + 1 entries");
        std::string j = zk::public_json(w.data(), n_public);
        need(j.size() + 1 <= out_size, "output buffer too small");
        memcpy(out, j.c_str(), j.size() + 1);
    });
}

int zk_read_zkey_header(const char* zkey_path, zk_zkey_header* out) {
    return guard([&] {
        need(zkey_path && out, "NULL argument");
        zk::ZkeyHeader h;
        zk::read_zkey_header(zkey_path, h);
        out->n_constraints = h.domain_size;
        out->n_vars = h.n_vars;
        out->n_public = h.n_public;
        out->domain_size = h.domain_size;
        out->log_domain = h.log_domain;
        memcpy(out->prime, h.r, 32);
    });
}

int zk_read_wtns(const char* wtns_path, uint64_t* w, uint32_t cap, uint32_t* n_out) {
    return guard([&] {
        need(wtns_path && n_out, "NULL argument");
        uint32_t n = 0;
        std::vector<uint64_t> v = zk::read_wtns(wtns_path, &n);
        *n_out = n;
        if (w) memcpy(w, v.data(), 32ull * std::min(n, cap));
    });
}

int zk_write_wtns(const char* wtns_path, const uint64_t* w, uint32_t n_vars) {
    return guard([&] {
        need(wtns_path && w, "NULL argument");
        zk::write_wtns(wtns_path, w, n_vars);
    });
}

namespace {
// out = M_sel * w over 44-byte entries {u32 matrix, u32 row, u32 wire, 32 B
// value}; value_raw: the zkey's c*R^2 encoding, else canonical c
void sparse_mvm(const uint8_t* coefs44, size_t n_coefs, uint8_t matrix_sel, const uint64_t* witness, uint32_t n_vars,
                uint64_t* out, uint32_t domain_size, bool value_raw) {
    need(coefs44 && witness && out, "NULL argument");
    need(matrix_sel <= 2, "matrix_sel must be 0, 1 or 2");
    for (uint32_t i = 0; i < n_vars; i++) need(zkh::Fr::is_canonical(witness + 4ull * i), "witness value not canonical");
    const zkh::Fr r2 = zkh::Fr::from_raw(zkh::FrP::R2);
    std::vector<zkh::Fr> acc(domain_size, zkh::Fr::zero());
    for (size_t i = 0; i < n_coefs; i++) {
        const uint8_t* e = coefs44 + 44 * i;
        uint32_t mat, row, wire;
        memcpy(&mat, e, 4); memcpy(&row, e + 4, 4); memcpy(&wire, e + 8, 4);
        if (mat != matrix_sel) continue;
        ZK_REQUIRE(row < domain_size && wire < n_vars, ZK_E_FORMAT, "coefficient index out of range");
        uint64_t v[4];
        memcpy(v, e + 12, 32);
        ZK_REQUIRE(zkh::Fr::is_canonical(v), ZK_E_ARG, "coefficient value not canonical (>= r)");
        // raw c*R^2: frMul(c R^2, w) = (c w) R; canonical c: (c R) from c * R^2 * R^-1 first
        const zkh::Fr c = value_raw ? zkh::Fr::from_raw(v) : zkh::Fr::from_raw(v) * r2 * r2;
        acc[row] = acc[row] + c * zkh::Fr::from_raw(witness + 4ull * wire);
    }
    for (uint32_t i = 0; i < domain_size; i++) acc[i].canonical(out + 4ull * i);
}
}  // namespace

int zk_sparse_mvm(const uint8_t* coefs44, size_t n_coefs, uint8_t matrix_sel, const uint64_t* witness, uint32_t n_vars,
                  uint64_t* out, uint32_t domain_size) {
    return guard([&] { sparse_mvm(coefs44, n_coefs, matrix_sel, witness, n_vars, out, domain_size, true); });
}

int zk_sparse_mvm_canonical(const uint8_t* coefs44, size_t n_coefs, uint8_t matrix_sel, const uint64_t* witness,
                            uint32_t n_vars, uint64_t* out, uint32_t domain_size) {
    return guard([&] { sparse_mvm(coefs44, n_coefs, matrix_sel, witness, n_vars, out, domain_size, false); });
}

int zk_ntt(zk_ctx* ctx, uint64_t* data, int log_n, int kind) {
    return guard([&] {
        need(ctx && data, "NULL argument");
        need(log_n >= 1 && log_n <= 27, "log_n must be 1..27");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        size_t bytes = 32ull << log_n;
        zk::DevBuf a(bytes), b(bytes);
        ZK_HIP(hipMemcpy(a.p, data, bytes, hipMemcpyHostToDevice));
        void* in[1] = {a.p};
        void* outp[1] = {b.p};
        ntt_device(ctx, in, outp, 1, log_n, kind);
        ZK_HIP(hipMemcpy(data, b.p, bytes, hipMemcpyDeviceToHost));
    });
}

int zk_ntt_device(zk_ctx* ctx, void* const* d_in, void* const* d_out, int batch, int log_n, int kind) {
    return guard([&] {
        need(ctx && d_in && d_out, "NULL argument");
        need(log_n >= 1 && log_n <= 27, "log_n must be 1..27");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        ntt_device(ctx, d_in, d_out, batch, log_n, kind);
    });
}

static int msm_host(zk_ctx* ctx, const uint64_t* points, const uint64_t* scalars, size_t n, int window, bool g2,
                    uint64_t* out, int* is_inf) {
    return guard([&] {
        need(ctx && (points || !n) && (scalars || !n) && out, "NULL argument");
        check_scalars(scalars, n);
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        size_t pb = (g2 ? 128 : 64) * n;
        zk::DevBuf dp(pb), ds(32 * n);
        if (n) {
            ZK_HIP(hipMemcpy(dp.p, points, pb, hipMemcpyHostToDevice));
            ZK_HIP(hipMemcpy(ds.p, scalars, 32 * n, hipMemcpyHostToDevice));
        }
        msm_device(ctx, dp.p, ds.p, n, window, g2, out, is_inf);
    });
}

int zk_msm_g1(zk_ctx* ctx, const uint64_t* points, const uint64_t* scalars, size_t n, int window, uint64_t out[8],
              int* is_inf) {
    return msm_host(ctx, points, scalars, n, window, false, out, is_inf);
}

int zk_msm_g2(zk_ctx* ctx, const uint64_t* points, const uint64_t* scalars, size_t n, int window, uint64_t out[16],
              int* is_inf) {
    return msm_host(ctx, points, scalars, n, window, true, out, is_inf);
}

int zk_msm_g1_device(zk_ctx* ctx, const void* d_points, const void* d_scalars, size_t n, int window, uint64_t out[8],
                     int* is_inf) {
    return guard([&] {
        need(ctx && out, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        msm_device(ctx, d_points, d_scalars, n, window, false, out, is_inf);
    });
}

int zk_msm_g1_device_noalloc(zk_ctx* ctx, const void* d_points, const void* d_scalars, size_t n, int window,
                             uint64_t out[8], int* is_inf) {
    return guard([&] {
        need(ctx && out && (n == 0 || (d_points && d_scalars)), "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        put_result(zk::msm_g1_noalloc(ctx->impl.device, d_points, d_scalars, n, window, ctx->impl.stream), out, is_inf);
    });
}

int zk_msm_g2_device(zk_ctx* ctx, const void* d_points, const void* d_scalars, size_t n, int window, uint64_t out[16],
                     int* is_inf) {
    return guard([&] {
        need(ctx && out, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        msm_device(ctx, d_points, d_scalars, n, window, true, out, is_inf);
    });
}

int zk_circuit_msm_device(zk_ctx* ctx, const zk_circuit* c, int which, const void* d_scalars, int window,
                          uint64_t out[16], int* is_inf) {
    return guard([&] {
        need(ctx && c && out && d_scalars, "NULL argument");
        need(which >= 0 && which <= 4, "which must be 0..4 (A, B1, B2, C, H)");
        need(window == 0 || (window >= 4 && window <= 22), "window must be 0 or 4..22");
        need(c->impl->ctx == &ctx->impl, "circuit belongs to another context");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        const zk::Circuit& k = *c->impl;
        const zk::DevBuf* std_pts[5] = {&k.A, &k.B1, &k.B2, &k.C, &k.H};
        const zk::DevBuf* slots[5] = {&k.A29, &k.B1_29, &k.B2_29, &k.C29, &k.H29};
        const int kind = which == 4 ? 1 : 0;
        const size_t n = which == 4 ? (size_t)k.n : (size_t)k.n_vars + 4;
        const int tc = k.tab_c[kind];
        // over the window tables unless a window with another window count is asked for
        const bool merged = tc && (!window || zk::msm_geom(n, window, false).W == zk::msm_geom(n, tc, false).W);
        const int cw = merged ? tc : (window ? window : ctx->impl.window_for(kind, n));
        // A', B1', B2': the prover's wire classes, which leave out infinite bases
        const uint8_t* cls = which <= 2 && k.wire_cls.bytes ? k.wire_cls.as<uint8_t>() : nullptr;
        const int range = !cls ? zk::MSM_RANGE_ALL : which == 0 ? zk::MSM_RANGE_A : zk::MSM_RANGE_B;
        // and the bucket split a proof would use (tuned, else the default)
        const int split = ctx->impl.split_for(kind + (merged ? 2 : 0), n, cw, which == 2);
        msm_device(ctx, std_pts[which]->p, d_scalars, n, cw, which == 2, out, is_inf, slots[which]->p, merged, cls,
                   range, split);
    });
}

int zk_spmv(zk_ctx* ctx, const uint32_t* row_ptr, const uint32_t* cols, const uint64_t* vals, uint32_t n_rows,
            const uint64_t* witness, uint32_t n_vars, uint64_t* out) {
    return guard([&] {
        need(ctx && row_ptr && witness && out, "NULL argument");
        uint32_t nnz = row_ptr[n_rows];
        for (uint32_t i = 0; i < n_rows; i++) need(row_ptr[i] <= row_ptr[i + 1], "row_ptr must be non-decreasing");
        for (uint32_t k = 0; k < nnz; k++) need(cols[k] < n_vars, "column index out of range");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        zk::Context& C = ctx->impl;
        zk::DevBuf dr(4ull * (n_rows + 1)), dc(4ull * nnz + 4), dv(32ull * nnz + 32), dw(32ull * n_vars + 32);
        zk::DevBuf a(32ull * n_rows + 32), b(32ull * n_rows + 32), c(32ull * n_rows + 32);
        ZK_HIP(hipMemcpy(dr.p, row_ptr, 4ull * (n_rows + 1), hipMemcpyHostToDevice));
        if (nnz) {
            ZK_HIP(hipMemcpy(dc.p, cols, 4ull * nnz, hipMemcpyHostToDevice));
            ZK_HIP(hipMemcpy(dv.p, vals, 32ull * nnz, hipMemcpyHostToDevice));
        }
        ZK_HIP(hipMemcpy(dw.p, witness, 32ull * n_vars, hipMemcpyHostToDevice));
        zk::spmv_abc(dr.as<uint32_t>(), dc.as<uint32_t>(), dv.p, n_rows, dr.as<uint32_t>(), dc.as<uint32_t>(), dv.p, 0,
                     dw.p, n_rows, a.p, b.p, c.p, C.stream);
        ZK_HIP(hipStreamSynchronize(C.stream));
        ZK_HIP(hipMemcpy(out, a.p, 32ull * n_rows, hipMemcpyDeviceToHost));
    });
}

int zk_hpoly(zk_ctx* ctx, const uint64_t* aw, const uint64_t* bw, int log_n, uint64_t* h) {
    return guard([&] {
        need(ctx && aw && bw && h, "NULL argument");
        need(log_n >= 1 && log_n <= 27, "log_n must be 1..27");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        zk::Context& C = ctx->impl;
        zk::NttPlan& p = C.ntt_plan(log_n);
        size_t n = (size_t)1 << log_n, bytes = 32 * n;
        zk::DevBuf a(bytes), b(bytes), c(bytes), t0(bytes), t1(bytes), t2(bytes), dh(bytes);
        ZK_HIP(hipMemcpy(a.p, aw, bytes, hipMemcpyHostToDevice));
        ZK_HIP(hipMemcpy(b.p, bw, bytes, hipMemcpyHostToDevice));
        zk::pointwise_mul(a.p, b.p, c.p, n, C.stream);
        void* abc[3] = {a.p, b.p, c.p};
        void* tmp[3] = {t0.p, t1.p, t2.p};
        ZK_HIP(hipEventRecord(ctx->ev0, C.stream));
        zk::ntt_run(p, abc, tmp, 3, true, nullptr, p.inv_coset.p, C.stream);
        zk::ntt_run(p, tmp, abc, 3, false, nullptr, nullptr, C.stream);
        zk::h_final(a.p, b.p, c.p, dh.p, n, C.stream);
        ZK_HIP(hipEventRecord(ctx->ev2, C.stream));
        ZK_HIP(hipStreamSynchronize(C.stream));
        ZK_HIP(hipEventElapsedTime(&ctx->last_ms[0], ctx->ev0, ctx->ev2));
        ZK_HIP(hipMemcpy(h, dh.p, bytes, hipMemcpyDeviceToHost));
    });
}

int zk_dev_alloc(zk_ctx* ctx, size_t bytes, void** out) {
    return guard([&] {
        need(ctx && out, "NULL argument");
        ZK_HIP(hipSetDevice(ctx->impl.device));
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw zk::Error(ZK_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
        *out = p;
    });
}

int zk_dev_free(zk_ctx* ctx, void* p) {
    return guard([&] {
        need(ctx != nullptr, "NULL ctx");
        if (p) ZK_HIP(hipFree(p));
    });
}

int zk_dev_upload(zk_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guard([&] {
        need(ctx && dst && src, "NULL argument");
        ZK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    });
}

int zk_dev_copy(zk_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guard([&] {
        need(ctx && dst && src, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        ZK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->impl.stream));
        ZK_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

int zk_dev_download(zk_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guard([&] {
        need(ctx && dst && src, "NULL argument");
        ZK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    });
}

int zk_last_timing(zk_ctx* ctx, float ms[5]) {
    return guard([&] {
        need(ctx && ms, "NULL argument");
        for (int i = 0; i < 5; i++) ms[i] = ctx->last_ms[i];
    });
}

int zk_circuit_bases(const zk_circuit* c, int which, void** d_ptr, uint64_t* count) {
    return guard([&] {
        need(c && d_ptr && count, "NULL argument");
        const zk::Circuit& k = *c->impl;
        const zk::DevBuf* b[5] = {&k.A, &k.B1, &k.B2, &k.C, &k.H};
        need(which >= 0 && which <= 4, "which must be 0..4 (A, B1, B2, C, H)");
        *d_ptr = b[which]->p;
        *count = which == 4 ? k.n : (uint64_t)k.n_vars + 4;
    });
}

int zk_synth_circuit(zk_ctx* ctx, uint64_t seed, uint32_t m, uint32_t n_pub, zk_circuit** out) {
    return guard([&] {
        need(ctx && out, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        auto c = new zk_circuit;
        try {
            c->impl = zk::circuit_synthetic(ctx->impl, seed, m, n_pub);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

int zk_synth_witness(uint64_t seed, uint32_t m, uint32_t n_pub, uint64_t pub_seed, uint64_t* w_out) {
    return guard([&] {
        need(w_out != nullptr, "NULL argument");
        zk::SynthR1cs r1;
        zk::synth_r1cs(seed, m, n_pub, r1);
        std::vector<zkh::Fr> pub = zk::synth_public(pub_seed, n_pub);
        std::vector<uint64_t> w;
        zk::synth_witness_from_public(r1, pub, w);
        memcpy(w_out, w.data(), w.size() * 8);
    });
}

int zk_circuit_witness(const zk_circuit* c, uint64_t pub_seed, uint64_t* w_out) {
    return guard([&] {
        need(c && w_out, "NULL argument");
        ZK_REQUIRE(c->impl->synth != nullptr, ZK_E_STATE, "circuit was not built by zk_synth_circuit");
        const zk::SynthR1cs& r1 = *c->impl->synth;
        std::vector<zkh::Fr> pub = zk::synth_public(pub_seed, r1.l);
        std::vector<uint64_t> w;
        zk::synth_witness_from_public(r1, pub, w);
        memcpy(w_out, w.data(), w.size() * 8);
    });
}

int zk_circuit_witness_device(zk_ctx* ctx, const zk_circuit* c, uint64_t pub_seed, void* d_w) {
    return guard([&] {
        need(ctx && c && d_w, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.aux_mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        zk::circuit_witness_device(*c->impl, pub_seed, d_w, ctx->impl.aux_stream);
    });
}

int zk_synth_blinding(uint64_t seed, uint64_t r[4], uint64_t s[4]) {
    return guard([&] {
        need(r && s, "NULL argument");
        zk::synth_blinding(seed, r, s);
    });
}

int zk_synth_uniform(uint64_t seed, size_t count, uint64_t* out) {
    return guard([&] {
        need(out != nullptr, "NULL argument");
        zk::synth_uniform(seed, count, out);
    });
}

int zk_glv_split(const uint64_t* k, size_t n, uint64_t* out) {
    return guard([&] {
        need(n == 0 || (k && out), "NULL argument");
        for (size_t i = 0; i < n; i++) {
            need(zkh::Fr::is_canonical(k + 4 * i), "scalar not canonical (>= r)");
            zk::glv::split(k + 4 * i, out + 4 * i, out + 4 * i + 2);
        }
    });
}

int zk_synth_sizes(uint32_t m, uint32_t n_pub, uint32_t* n_vars, uint32_t* domain) {
    return guard([&] {
        need(n_vars && domain, "NULL argument");
        *n_vars = 1 + n_pub + m;
        uint64_t need_rows = (uint64_t)m + n_pub + 1, d = 1;
        while (d < need_rows) d <<= 1;
        *domain = (uint32_t)d;
    });
}

}  // extern "C"

// ---------------------------------------------------------------- verification
static void check_points_words(const uint64_t* w, size_t n, const char* what) {
    for (size_t i = 0; i < n; i++)
        if (!zkh::Fp::is_canonical(w + 4 * i)) throw zk::Error(ZK_E_ARG, std::string(what) + " coordinate not below p");
}

int zk_verify(const zk_circuit* c, const uint64_t* publics, uint32_t n_public, const zk_proof* proof, int* ok) {
    return guard([&] {
        need(c && proof && ok && (publics || n_public == 0), "NULL argument");
        check_points_words(&proof->pi_a[0], 8, "proof");
        const zk::Circuit& ci = *c->impl;
        *ok = zk::verify_proof(ci.hdr, ci.ic.data(), publics, n_public, *proof) ? 1 : 0;
    });
}

int zk_verify_zkey(const char* zkey_path, const uint64_t* publics, uint32_t n_public, const zk_proof* proof, int* ok) {
    return guard([&] {
        need(zkey_path && proof && ok && (publics || n_public == 0), "NULL argument");
        check_points_words(&proof->pi_a[0], 8, "proof");
        zk::ZkeyView zv(zkey_path);
        ZK_REQUIRE(zv.IC != nullptr, ZK_E_FORMAT, "zkey has no section 3 (IC)");
        std::vector<uint64_t> ic((const uint64_t*)zv.IC, (const uint64_t*)zv.IC + 8ull * (zv.hdr.n_public + 1));
        *ok = zk::verify_proof(zv.hdr, ic.data(), publics, n_public, *proof) ? 1 : 0;
    });
}

int zk_verify_json(const char* zkey_path, const char* public_json, const char* proof_json, int* ok) {
    return guard([&] {
        need(zkey_path && public_json && proof_json && ok, "NULL argument");
        const std::vector<uint64_t> pub = zk::parse_public_json(public_json);
        const zk_proof p = zk::parse_proof_json(proof_json);
        zk::ZkeyView zv(zkey_path);
        ZK_REQUIRE(zv.IC != nullptr, ZK_E_FORMAT, "zkey has no section 3 (IC)");
        std::vector<uint64_t> ic((const uint64_t*)zv.IC, (const uint64_t*)zv.IC + 8ull * (zv.hdr.n_public + 1));
        *ok = zk::verify_proof(zv.hdr, ic.data(), pub.data(), (uint32_t)(pub.size() / 4), p) ? 1 : 0;
    });
}

int zk_pairing(const uint64_t p[8], const uint64_t q[16], uint64_t out[48]) {
    return guard([&] {
        need(p && q && out, "NULL argument");
        check_points_words(p, 2, "G1");
        check_points_words(q, 4, "G2");
        const zkh::Aff<zkh::Fp> P = zkh::g1_from_words(p);
        const zkh::Aff<zkh::Fp2> Q = zkh::g2_from_words(q);
        need(zkh::g1_on_curve(P) && zkh::g2_on_curve(Q), "point not on its curve");
        const zkh::Fp12 e = zkh::pairing(P, Q);
        const zkh::Fp2* c[6] = {&e.c0.c0, &e.c0.c1, &e.c0.c2, &e.c1.c0, &e.c1.c1, &e.c1.c2};
        for (int k = 0; k < 6; k++) {
            memcpy(out + 8 * k, c[k]->c0.v, 32);
            memcpy(out + 8 * k + 4, c[k]->c1.v, 32);
        }
    });
}

// ---------------------------------------------------------------- field tests
int zk_field_op_device(zk_ctx* ctx, int field, int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    return guard([&] {
        need(ctx && a && out, "NULL argument");
        need(field == 0 || field == 1, "field must be 0 (Fp) or 1 (Fr)");
        need(op >= 0 && op < zk::FOP_COUNT, "unknown field op");
        need(b || op == zk::FOP_NEG || op == zk::FOP_SQR || op == zk::FOP_INV || op == zk::FOP_TO_MONT ||
                 op == zk::FOP_FROM_MONT, "binary op without b");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        const size_t bytes = std::max<size_t>(n, 1) * 32;
        zk::DevBuf da(bytes), db(b ? bytes : 0), dout(bytes);
        hipStream_t st = ctx->impl.stream;
        ZK_HIP(hipMemcpyAsync(da.p, a, n * 32, hipMemcpyHostToDevice, st));
        if (b) ZK_HIP(hipMemcpyAsync(db.p, b, n * 32, hipMemcpyHostToDevice, st));
        zk::field_op_std(field, op, da.p, b ? db.p : nullptr, dout.p, n, st);
        ZK_HIP(hipMemcpyAsync(out, dout.p, n * 32, hipMemcpyDeviceToHost, st));
        ZK_HIP(hipStreamSynchronize(st));
    });
}

int zk_field29_op_device(zk_ctx* ctx, int op, const uint32_t* a, const uint32_t* b, uint32_t* out, size_t n) {
    return guard([&] {
        need(ctx && a && out, "NULL argument");
        need(op >= 0 && op < zk::F29_COUNT, "unknown 29-bit op");
        need(b || op == zk::F29_SQR || op == zk::F29_NEG || op == zk::F29_REDUCE || op == zk::F29_ROUNDTRIP ||
                 op == zk::F29_NEG4,
             "binary op without b");
        need((op != zk::F29Q_MUL && op != zk::F29Q_MSM) || n % 2 == 0, "Fp2 ops take element pairs (even n)");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        const size_t bytes = std::max<size_t>(n, 1) * 36;
        zk::DevBuf da(bytes), db(b ? bytes : 0), dout(bytes);
        hipStream_t st = ctx->impl.stream;
        ZK_HIP(hipMemcpyAsync(da.p, a, n * 36, hipMemcpyHostToDevice, st));
        if (b) ZK_HIP(hipMemcpyAsync(db.p, b, n * 36, hipMemcpyHostToDevice, st));
        zk::field29_op(op, da.as<uint32_t>(), b ? db.as<uint32_t>() : nullptr, dout.as<uint32_t>(), n, st);
        ZK_HIP(hipMemcpyAsync(out, dout.p, n * 36, hipMemcpyDeviceToHost, st));
        ZK_HIP(hipStreamSynchronize(st));
    });
}

int zk_probe_valu(zk_ctx* ctx, int kind, double* lane_ops_per_s) {
    return guard([&] {
        need(ctx && lane_ops_per_s, "NULL argument");
        need(kind >= 0 && (kind & 15) <= 5 && (kind >> 4) <= 16, "kind: op 0..5 + 16 * waves per SIMD (0..16)");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        *lane_ops_per_s = zk::valu_probe(kind, ctx->impl.stream);
    });
}

int zk_slot_bytes_check(zk_ctx* ctx, const zk_circuit* c, uint64_t out[3]) {
    return guard([&] {
        need(ctx && c && out, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        const zk::Circuit& ci = *c->impl;
        ctx->impl.ntt_plan(ci.log_n);
        zk::ProveSlot sl;
        out[0] = zk::ProveSlot::bytes_needed(ctx->impl, ci);
        sl.reserve(ctx->impl, ci, false);
        out[1] = sl.device_bytes();
        out[2] = sl.grow_bytes(ctx->impl, ci);
    });
}

int zk_circuit_spmv_device(zk_ctx* ctx, const zk_circuit* c, const void* d_witness, void* d_aw, void* d_bw, void* d_cw) {
    return guard([&] {
        need(ctx && c && d_witness && d_aw && d_bw && d_cw, "NULL argument");
        std::lock_guard<std::mutex> lk(ctx->impl.mu);
        ZK_HIP(hipSetDevice(ctx->impl.device));
        const zk::Circuit& ci = *c->impl;
        hipStream_t st = ctx->impl.stream;
        ZK_HIP(hipEventRecord(ctx->ev0, st));
        zk::spmv_abc(ci.rowA.as<uint32_t>(), ci.colA.as<uint32_t>(), ci.valA.p, ci.n_rowsA, ci.rowB.as<uint32_t>(),
                     ci.colB.as<uint32_t>(), ci.valB.p, ci.n_rowsB, d_witness, ci.n, d_aw, d_bw, d_cw, st);
        ZK_HIP(hipEventRecord(ctx->ev2, st));
        ZK_HIP(hipEventSynchronize(ctx->ev2));
        ZK_HIP(hipEventElapsedTime(&ctx->last_ms[0], ctx->ev0, ctx->ev2));
        ctx->last_ms[1] = ctx->last_ms[2] = ctx->last_ms[3] = ctx->last_ms[4] = 0;
    });
}
