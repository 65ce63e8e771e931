// Continuous-batching proof server on one MI355X.
//
// Requests (circuit resident in HBM, witness, r, s) enter a queue. A worker
// thread admits them into workspace slots as slots free up and as free HBM
// allows, at most one proof over a domain of 2^23 or more at a time (smaller
// requests may pass a waiting one) (a slot is created or grown only if hipMemGetInfo leaves the
// reserve intact), so requests that arrive while others run on the GPU are
// overlapped with them. Completion is detected without blocking (stream
// queries); the host tail of a finished proof runs while other slots keep the
// device busy. Each request records its queueing and service time.
//
// The reference has no serving mode: it proves one zkey/wtns pair per
// process (src/main.cu:115-171).
#pragma once
#include <stdint.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "prover.hpp"

namespace zk {

struct ServerRequest {
    uint64_t ticket = 0;
    const Circuit* circ = nullptr;
    std::vector<uint64_t> witness;      // host copy (empty for a device witness)
    const void* d_witness = nullptr;    // caller's device witness
    uint64_t r[4] = {0, 0, 0, 0}, s[4] = {0, 0, 0, 0};
    std::chrono::steady_clock::time_point t_submit, t_start, t_done;
    zk_proof proof{};
    int status = ZK_OK;
    std::string error;
    bool done = false;
};

class Server {
  public:
    Server(Context& ctx, int max_slots, size_t reserve_bytes);
    ~Server();
    uint64_t submit(const Circuit& c, const uint64_t* w, uint32_t nw, const uint64_t* r, const uint64_t* s);
    // witness already in device memory (n_vars x 32 B, canonical), read when
    // the request runs: it must stay valid and unchanged until completion
    uint64_t submit_device(const Circuit& c, const void* d_w, const uint64_t* r, const uint64_t* s);
    // done: copies the result out and forgets the request; returns false while pending
    bool poll(uint64_t ticket, zk_proof* out, double times[3], int* status, std::string* err);
    void wait(uint64_t ticket, zk_proof* out, double times[3], int* status, std::string* err);
    bool cancel(uint64_t ticket);
    // on an idle server: create up to max_slots slots (while free HBM stays
    // above the reserve) and size each for every circuit of cs, so that no
    // request on them allocates device memory later; returns the slot count
    int prepare(const Circuit* const* cs, int count);
    void stats(double out[8]);

  private:
    uint64_t admit_request(std::shared_ptr<ServerRequest> req, const Circuit& c, const uint64_t* r, const uint64_t* s);
    void loop();
    bool try_admit(std::unique_lock<std::mutex>& lk);
    Context& ctx_;
    const int max_slots_;
    const size_t reserve_;
    std::vector<std::unique_ptr<ProveSlot>> slots_;
    std::vector<std::shared_ptr<ServerRequest>> running_;  // per slot, null when free
    std::deque<std::shared_ptr<ServerRequest>> queue_;
    std::unordered_map<uint64_t, std::shared_ptr<ServerRequest>> reqs_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::thread worker_;
    bool stop_ = false;
    uint64_t next_ticket_ = 1;
    double n_completed_ = 0, n_failed_ = 0, n_rejected_ = 0, n_cancelled_ = 0, peak_inflight_ = 0;
};

}  // namespace zk
