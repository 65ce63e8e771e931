#include "server.hpp"

#include <string.h>

#include <algorithm>

namespace zk {

using clk = std::chrono::steady_clock;

static double ms_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

Server::Server(Context& ctx, int max_slots, size_t reserve_bytes)
    : ctx_(ctx), max_slots_(max_slots), reserve_(reserve_bytes) {
    ZK_REQUIRE(max_slots >= 1 && max_slots <= 16, ZK_E_ARG, "max_slots must be 1..16");
    worker_ = std::thread([this] { loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    if (worker_.joinable()) worker_.join();
}

uint64_t Server::submit(const Circuit& c, const uint64_t* w, uint32_t nw, const uint64_t* r, const uint64_t* s) {
    ZK_REQUIRE(w != nullptr, ZK_E_ARG, "NULL witness");
    ZK_REQUIRE(nw >= c.n_vars, ZK_E_ARG, "witness shorter than nVars");
    for (uint32_t i = 0; i < c.n_vars; i++)
        ZK_REQUIRE(zkh::Fr::is_canonical(w + 4ull * i), ZK_E_ARG, "witness value " + std::to_string(i) + " is not canonical");
    auto req = std::make_shared<ServerRequest>();
    req->witness.assign(w, w + 4ull * c.n_vars);
    return admit_request(std::move(req), c, r, s);
}

uint64_t Server::submit_device(const Circuit& c, const void* d_w, const uint64_t* r, const uint64_t* s) {
    ZK_REQUIRE(d_w != nullptr, ZK_E_ARG, "NULL device witness");
    auto req = std::make_shared<ServerRequest>();
    req->d_witness = d_w;
    return admit_request(std::move(req), c, r, s);
}

uint64_t Server::admit_request(std::shared_ptr<ServerRequest> req, const Circuit& c, const uint64_t* r,
                               const uint64_t* s) {
    if (r) memcpy(req->r, r, 32);
    if (s) memcpy(req->s, s, 32);
    ZK_REQUIRE(zkh::Fr::is_canonical(req->r) && zkh::Fr::is_canonical(req->s), ZK_E_ARG, "blinding must be canonical");
    size_t free_b = 0, total_b = 0;
    ZK_HIP(hipMemGetInfo(&free_b, &total_b));
    size_t need;
    {
        std::lock_guard<std::mutex> g(ctx_.mu);
        need = ProveSlot::bytes_needed(ctx_, c);
    }
    if (need + reserve_ > total_b) {
        std::lock_guard<std::mutex> lk(mu_);
        n_rejected_++;
        throw Error(ZK_E_NOMEM, "request needs more device memory than the GPU has");
    }
    req->circ = &c;
    req->t_submit = clk::now();
    uint64_t t;
    {
        std::lock_guard<std::mutex> lk(mu_);
        ZK_REQUIRE(!stop_, ZK_E_STATE, "server is stopping");
        t = req->ticket = next_ticket_++;
        reqs_[t] = req;
        queue_.push_back(req);
    }
    cv_work_.notify_one();
    return t;
}

int Server::prepare(const Circuit* const* cs, int count) {
    ZK_REQUIRE(count >= 1, ZK_E_ARG, "no circuits to prepare for");
    std::vector<const Circuit*> order;
    for (int i = 0; i < count; i++) {
        ZK_REQUIRE(cs[i] != nullptr, ZK_E_ARG, "NULL circuit");
        order.push_back(cs[i]);
    }
    // largest first: its reserve allocates the largest buffers once, and the
    // smaller circuits then find them big enough
    std::sort(order.begin(), order.end(), [](const Circuit* a, const Circuit* b) {
        return (size_t)a->n + a->n_vars > (size_t)b->n + b->n_vars;
    });
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& r : running_) ZK_REQUIRE(!r, ZK_E_STATE, "zk_server_prepare needs an idle server");
    ZK_REQUIRE(queue_.empty(), ZK_E_STATE, "zk_server_prepare needs an idle server");
    std::lock_guard<std::mutex> g(ctx_.mu);
    ZK_HIP(hipSetDevice(ctx_.device));
    for (int k = 0; k < max_slots_; k++) {
        if (k == (int)slots_.size()) {
            size_t need = 0;
            for (const Circuit* c : order) need = std::max(need, ProveSlot::bytes_needed(ctx_, *c));
            size_t free_b = 0, total_b = 0;
            ZK_HIP(hipMemGetInfo(&free_b, &total_b));
            if (k > 0 && free_b < need + reserve_) break;
            slots_.push_back(std::make_unique<ProveSlot>());
            running_.push_back(nullptr);
        }
        for (const Circuit* c : order) {
            size_t free_b = 0, total_b = 0;
            ZK_HIP(hipMemGetInfo(&free_b, &total_b));
            ZK_REQUIRE(free_b >= slots_[k]->grow_bytes(ctx_, *c) + reserve_, ZK_E_NOMEM,
                       "not enough free device memory to prepare a slot above the reserve");
            slots_[k]->reserve(ctx_, *c, true);
        }
    }
    return (int)slots_.size();
}

static void copy_out(const ServerRequest& q, zk_proof* out, double times[3], int* status, std::string* err) {
    if (out) *out = q.proof;
    if (times) {
        times[0] = q.t_start.time_since_epoch().count() ? ms_between(q.t_submit, q.t_start) : 0;
        times[1] = q.t_start.time_since_epoch().count() ? ms_between(q.t_start, q.t_done) : 0;
        times[2] = ms_between(q.t_submit, q.t_done);
    }
    if (status) *status = q.status;
    if (err) *err = q.error;
}

bool Server::poll(uint64_t ticket, zk_proof* out, double times[3], int* status, std::string* err) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    ZK_REQUIRE(it != reqs_.end(), ZK_E_ARG, "unknown ticket " + std::to_string(ticket));
    if (!it->second->done) return false;
    copy_out(*it->second, out, times, status, err);
    reqs_.erase(it);
    return true;
}

void Server::wait(uint64_t ticket, zk_proof* out, double times[3], int* status, std::string* err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    ZK_REQUIRE(it != reqs_.end(), ZK_E_ARG, "unknown ticket " + std::to_string(ticket));
    auto req = it->second;
    cv_done_.wait(lk, [&] { return req->done; });
    copy_out(*req, out, times, status, err);
    reqs_.erase(ticket);
}

bool Server::cancel(uint64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = std::find_if(queue_.begin(), queue_.end(), [&](const std::shared_ptr<ServerRequest>& q) { return q->ticket == ticket; });
    if (it == queue_.end()) return false;
    auto req = *it;
    queue_.erase(it);
    req->status = ZK_E_STATE;
    req->error = "cancelled";
    req->t_done = clk::now();
    req->done = true;
    n_cancelled_++;
    cv_done_.notify_all();
    return true;
}

void Server::stats(double out[8]) {
    std::lock_guard<std::mutex> lk(mu_);
    double inflight = 0;
    for (auto& r : running_) inflight += r ? 1 : 0;
    out[0] = n_completed_;
    out[1] = n_failed_;
    out[2] = n_rejected_;
    out[3] = n_cancelled_;
    out[4] = (double)queue_.size();
    out[5] = inflight;
    out[6] = (double)slots_.size();
    out[7] = peak_inflight_;
}

// A proof over a domain this large (m >= 2^22, 45 ms and more alone) keeps
// the GPU busy on its own: two of them in flight only share it, each taking
// about twice as long. They are admitted one at a time, and smaller requests
// may pass a waiting one, so that the smaller ones are not held behind them.
// Measured with Poisson arrivals over 2^18..2^24 circuits at 0.4 / 0.7 / 1.0
// of capacity: the p99 of 2^18 and 2^20 requests 216-820 ms -> 165-322 ms,
// the median at full load 359 -> 158 ms, the overall p99 unchanged (the
// largest proofs set it); at 2^21 as the threshold the 2^20 requests waited
// behind the 2^24 ones instead.
constexpr uint32_t SERVE_FULL_GPU_N = 1u << 23;

// Admit queued requests into slots, oldest first, at most one request whose
// domain fills the GPU in flight (SERVE_FULL_GPU_N). A free slot whose buffers
// already fit the request (grow_bytes == 0: nothing is allocated) is used as
// is; otherwise a new slot is created, or else a free slot grows, only if free
// HBM stays above the reserve afterwards. A new slot comes first: growing
// frees device memory, which waits for the proofs the other slots have in
// flight. zk_server_prepare sizes every slot up front, so that admission
// never allocates while proofs run.
bool Server::try_admit(std::unique_lock<std::mutex>& lk) {
    bool any = false;
    for (;;) {
        bool big_running = false;
        for (auto& r : running_) big_running = big_running || (r && r->circ->n >= SERVE_FULL_GPU_N);
        auto qi = queue_.begin();
        while (qi != queue_.end() && big_running && (*qi)->circ->n >= SERVE_FULL_GPU_N) ++qi;
        if (qi == queue_.end()) break;
        auto req = *qi;
        // the sizes depend on the context's windows and NTT plans: read under its lock
        auto grow = [&](size_t i) {
            std::lock_guard<std::mutex> g(ctx_.mu);
            return slots_[i]->grow_bytes(ctx_, *req->circ);
        };
        int pick = -1;
        for (size_t i = 0; i < slots_.size() && pick < 0; i++)
            if (!running_[i] && grow(i) == 0) pick = (int)i;
        if (pick < 0) {
            size_t free_b = 0, total_b = 0;
            ZK_HIP(hipMemGetInfo(&free_b, &total_b));
            size_t need;
            {
                std::lock_guard<std::mutex> g(ctx_.mu);
                need = ProveSlot::bytes_needed(ctx_, *req->circ);
            }
            if ((int)slots_.size() < max_slots_ && free_b >= need + reserve_) {
                slots_.push_back(std::make_unique<ProveSlot>());
                running_.push_back(nullptr);
                pick = (int)slots_.size() - 1;
            }
            for (size_t i = 0; i < slots_.size() && pick < 0; i++)
                if (!running_[i] && free_b >= grow(i) + reserve_) pick = (int)i;
        }
        if (pick < 0) break;  // wait for a slot or for memory
        queue_.erase(qi);
        running_[pick] = req;
        double inflight = 0;
        for (auto& r : running_) inflight += r ? 1 : 0;
        peak_inflight_ = std::max(peak_inflight_, inflight);
        req->t_start = clk::now();
        lk.unlock();
        int st = ZK_OK;
        std::string err;
        try {
            std::lock_guard<std::mutex> g(ctx_.mu);
            if (req->d_witness)
                slots_[pick]->enqueue(ctx_, *req->circ, (const uint64_t*)req->d_witness, true, req->r, req->s);
            else
                slots_[pick]->enqueue(ctx_, *req->circ, req->witness.data(), false, req->r, req->s);
        } catch (const Error& e) {
            st = e.code;
            err = e.what();
        } catch (const std::exception& e) {
            st = ZK_E_INTERNAL;
            err = e.what();
        }
        if (st != ZK_OK) slots_[pick]->drain();  // drop whatever part of the proof was queued before the failure
        lk.lock();
        if (st != ZK_OK) {
            req->status = st;
            req->error = err;
            req->t_done = clk::now();
            req->done = true;
            running_[pick] = nullptr;
            n_failed_++;
            cv_done_.notify_all();
        }
        any = true;
    }
    return any;
}

void Server::loop() {
    (void)hipSetDevice(ctx_.device);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        bool progress = false;
        for (size_t i = 0; i < slots_.size(); i++) {
            auto req = running_[i];
            if (!req) continue;
            lk.unlock();
            bool done = false;
            int st = ZK_OK;
            std::string err;
            try {
                if (slots_[i]->device_done()) {
                    // the host tail reads only the slot's own buffers: no context lock,
                    // so other callers of the context are not held up by it
                    slots_[i]->finish(&req->proof);
                    done = true;
                }
            } catch (const Error& e) {
                done = true;
                st = e.code;
                err = e.what();
            } catch (const std::exception& e) {
                done = true;
                st = ZK_E_INTERNAL;
                err = e.what();
            }
            lk.lock();
            if (done) {
                req->status = st;
                req->error = err;
                req->t_done = clk::now();
                req->done = true;
                running_[i] = nullptr;
                (st == ZK_OK ? n_completed_ : n_failed_) += 1;
                progress = true;
                cv_done_.notify_all();
            }
        }
        try {
            if (try_admit(lk)) progress = true;
        } catch (const std::exception&) {
            // hipMemGetInfo failed: retry on the next iteration
        }
        bool busy = std::any_of(running_.begin(), running_.end(), [](const std::shared_ptr<ServerRequest>& r) { return r != nullptr; });
        if (stop_ && queue_.empty() && !busy) break;
        if (!progress) {
            if (busy) cv_work_.wait_for(lk, std::chrono::microseconds(50));
            else cv_work_.wait(lk, [&] { return stop_ || !queue_.empty(); });
        }
    }
}

}  // namespace zk
