// Groth16 prover orchestration on one MI355X.
//
// Reference: groth16Prove (src/groth16.cu:1143-1397). Differences by design:
//   * a circuit is loaded once (CSR of A/B and every CRS base resident in
//     HBM) and then proves any number of witnesses;
//   * the four MSMs over [w | 1, r, s, -rs] (A, B1, C in G1, B2 in G2) share
//     one digit sort: the base arrays carry alpha1/beta1/beta2/delta points in
//     four extra slots, so the blinding terms and the alpha/beta offsets fall
//     out of the MSMs themselves (snarkjs formula; r = s = 0 reproduces the
//     reference's unblinded proof bit for bit);
//   * the H pipeline (SpMV, iNTT + coset, NTT, h) and the witness MSMs run on
//     two streams of one workspace slot; several slots overlap proofs;
//   * only short serial tails run on the host: per-window combination of the
//     MSM partials, s*pi_A + r*B1, and the affine conversion.
#pragma once
#include <stdint.h>

#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/zk_abi.h"
#include "common.hpp"
#include "msm.hpp"
#include "ntt.hpp"
#include "synth.hpp"
#include "zkey.hpp"

namespace zk {

struct Context;

struct Circuit {
    Context* ctx = nullptr;
    uint32_t n_vars = 0, n_public = 0, n = 0;
    int log_n = 0;
    uint32_t n_rowsA = 0, n_rowsB = 0;
    size_t nnzA = 0, nnzB = 0;
    DevBuf rowA, colA, valA, rowB, colB, valB;
    // extended bases, n_vars + 4 points each (see header comment); H: n points
    DevBuf A, B1, B2, C, H;
    // the bases again in the 29-bit-limb form of the bucket accumulation,
    // each slot also holding the x of its endomorphism image (msm.hpp); with
    // window tables, W tables of the array one after the other
    DevBuf A29, B1_29, B2_29, C29, H29;
    // window of the tables (msm_points_table) of the witness arrays (A', B1',
    // C', B2') and of H; 0: no tables, one slot per base
    int tab_c[2] = {0, 0};
    // class of every witness scalar (n_vars + 4 bytes, MSM_CLS_*): which of
    // its A', B1'/B2' bases are infinite (a wire that no row of A, or of B,
    // uses). The witness sort orders each bucket by class, so that the A MSM
    // and the B1, B2 MSMs read only entries with finite bases. Empty when
    // fewer than 1/8 of the A' and B' bases are infinite.
    DevBuf wire_cls;
    ZkeyHeader hdr;               // VK in file layout
    std::vector<uint8_t> coefs;   // §4 bytes, kept to write the circuit back out
    std::vector<uint64_t> ic;     // (l+1) x 8
    std::unique_ptr<SynthR1cs> synth;  // host R1CS of synthetic circuits (witness generation)
    // synthetic circuits: constraint rows grouped by dependency level (a row's
    // level is one more than the highest level among the aux wires it reads),
    // levels[L]..levels[L+1] indexing wit_rows, for the device witness generator
    DevBuf wit_rows;
    std::vector<uint32_t> wit_levels;
    size_t device_bytes() const;
};

// Window tables of c's resident bases (msm.hpp): mode 1 builds them (window
// msm_table_c per array size), 0 frees them, -1 builds them when they take
// at most a third of the device memory still free. Waits for the device
// first: no proof of c may be in flight.
void circuit_set_tables(Circuit& c, int mode);

// Device results of one proof's five MSMs (level-2 partials, host memory).
struct ProofParts {
    MsmPartials a, b1, c, b2, h;
};

// Host tail of a proof: finish the MSMs and assemble pi_A, pi_B,
// pi_C = C' + H + s*pi_A + r*B1 (Montgomery affine). ready, when given, holds
// one event per MSM (a, b1, c, b2, h) after which its partials are on the
// host; each MSM is finished on its own thread as soon as its event
// completes, so the host work overlaps the MSMs still running. Returns the
// host milliseconds spent after the last of them completed.
double assemble_proof(const ProofParts& p, const uint64_t r[4], const uint64_t s[4], zk_proof* out,
                      const hipEvent_t* ready = nullptr);

struct ProveSlot {
    // s0: H pipeline + H MSM; s1: witness sort + A; s2: B1; s3: C; s4: B2.
    // Each proof forks s1..s4 from s0 and joins them back into s0 at its end.
    hipStream_t s0 = nullptr, s1 = nullptr, s2 = nullptr, s3 = nullptr, s4 = nullptr;
    hipEvent_t ev_scal = nullptr, ev_done1 = nullptr, ev_join[4] = {}, ev_h = nullptr;
    bool alone_layout = false;   // the layout body() used for the queued proof
    DevBuf scal, aw, bw, cw, t0, t1, t2, h, wtmp;
    HostBuf staging;                 // pinned copy of a host witness (async upload)
    MsmSort sortW, sortH;
    MsmScratch scrA, scrB1, scrC, scrB2, scrH;
    ProofParts parts;
    uint64_t r[4], s[4];
    const Circuit* circ = nullptr;
    int window_w = 0, window_h = 0;  // windows of the last proof
    // profiling (off by default): events on both streams, read after finish()
    bool timed = false;
    hipEvent_t ev[8] = {};
    double stats[24] = {};
    void enable_timing();
    void collect_stats();
    ProveSlot();
    ~ProveSlot();
    ProveSlot(const ProveSlot&) = delete;
    ProveSlot& operator=(const ProveSlot&) = delete;
    // queue all device work for one proof; witness is host (canonical) or
    // device; alone: no other proof shares the device (the stream layout for
    // latency, see body)
    void enqueue(Context& ctx, const Circuit& c, const uint64_t* witness, bool witness_on_device, const uint64_t r[4],
                 const uint64_t s[4], bool alone = false);
    // size every device buffer a proof of c needs in this slot
    void prepare(Context& ctx, const Circuit& c);
    // size every buffer a proof of c needs in this slot, sorts, MSM scratch
    // and host partials included, without queueing work; a batch reserves its
    // slots for all of its circuits first, so that no buffer grows (and no
    // allocation waits for the device) while proofs are in flight
    void reserve(Context& ctx, const Circuit& c, bool host_witness);
    // the device work of one proof once scal holds [w | 1, r, s, -rs] (queued
    // on s0 before the call): everything up to the partials in out, ending
    // with s1..s4 joined into s0. Stream-capturable after prepare().
    void body(Context& ctx, const Circuit& c, ProofParts& out, bool alone = false);
    // wait for the device work and run the host tail
    void finish(zk_proof* out);
    // true when every stream of the slot has drained (non-blocking)
    bool device_done() const;
    // device bytes a proof of c needs in a new slot (its whole workspace at
    // the windows and splits ctx would use, plus the NTT plan if missing)
    static size_t bytes_needed(const Context& ctx, const Circuit& c);
    // device bytes reserve(ctx, c, ...) would add to this slot now (0: a
    // proof of c queues without allocating)
    size_t grow_bytes(const Context& ctx, const Circuit& c) const;
    // wait for every stream of the slot, ignoring errors (cleanup after a failure)
    void drain() noexcept;
    // number of per-proof statistics in stats[] (layout in zk_abi.h: zk_prove_stats)
    static constexpr int N_STATS = 24;
    size_t device_bytes() const;

  private:
    static size_t count_bytes(const Context& ctx, const Circuit& c, const ProveSlot* sl);
};

constexpr int TUNE_C_MIN = 11, TUNE_C_MAX = 19;  // autotuner window range (c = 19: 7 windows)
int size_class(size_t n);                         // ceil(log2 n)

// windows and bucket splits of one proof's MSMs on a circuit
struct MsmPlan {
    int cw = 0, ch = 0;            // windows of the witness sort and of the h sort
    bool mw = false, mh = false;   // over the circuit's window tables (one bucket set)
    int sp_w1 = 1, sp_w2 = 1, sp_h = 1;   // splits: witness G1, witness G2, H
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    // side work that needs no other context state (synthetic witness
    // generation): its own stream and lock, so it does not wait for provers
    hipStream_t aux_stream = nullptr;
    std::mutex aux_mu;
    std::map<int, std::unique_ptr<NttPlan>> ntt_plans;
    std::unique_ptr<ProveSlot> slot;   // slot used by the one-shot prove entry point
    DevBuf comb_g1, comb_g2;           // fixed-base tables for synthetic setup
    // MSM windows: forced (nonzero) or tuned per (kind, ceil log2 n), else the
    // size default. kind 0 = witness sort (A', B1', C', B2'), 1 = H; kinds 2
    // and 3 are the same MSMs over window tables, whose window is the
    // tables' (only their splits are tuned). With a tuned window come the
    // tuned bucket splits of its G1 and G2 MSMs; else msm_default_split.
    struct MsmChoice {
        int c, split_g1, split_g2;   // split 0: msm_default_split
    };
    int force_w = 0, force_h = 0;
    std::map<std::pair<int, int>, MsmChoice> tuned;
    int tables_mode = -1;   // circuit_set_tables mode for circuits loaded from now on
    explicit Context(int dev);
    ~Context();
    NttPlan& ntt_plan(int log_n);
    void ensure_comb();
    int window_for(int kind, size_t n) const;
    int split_for(int kind, size_t n, int c, bool g2) const;
    // windows and splits of a proof of c: its window tables unless a window
    // with another window count is forced
    MsmPlan msm_plan(const Circuit& c) const;
    std::string arch() const;
    void autotune(const Circuit& c, bool force, int reps, int chosen[2]);
    int tune_load(const std::string& path);
    void tune_save(const std::string& path) const;
};

// A fixed sequence of proofs captured once as one HIP graph: every proof's
// device work on a few slots (round-robin), witnesses read from caller device
// buffers whose addresses stay fixed, blinding read from pinned memory that
// run() fills, so the graph is replayed for new witness contents and new r, s.
struct BatchGraph {
    Context& ctx;
    std::vector<const Circuit*> circs;
    std::vector<const void*> wits;
    std::vector<std::unique_ptr<ProveSlot>> slots;
    std::unique_ptr<ProofParts[]> parts;  // one per proof
    HostBuf params;                     // count x [1, r, s, -rs] canonical
    hipStream_t origin = nullptr;
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_end;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t n_nodes = 0;
    BatchGraph(Context& ctx, int count, const Circuit* const* circs, const void* const* d_wits, int nslots);
    ~BatchGraph();
    // r, s: count x 4 words each (NULL = zero blinding); out: count proofs
    void run(const uint64_t* r, const uint64_t* s, zk_proof* out);

  private:
    void fill_params(const uint64_t* r, const uint64_t* s);
    void record(bool capture);
};

std::unique_ptr<Circuit> circuit_from_zkey(Context& ctx, const std::string& path);
// satisfying witness of a synthetic circuit with public inputs from pub_seed's
// stream (the same values as synth_witness_from_public), computed on the
// device into d_w (n_vars x 32 B canonical) level by level on stream st
void circuit_witness_device(const Circuit& c, uint64_t pub_seed, void* d_w, hipStream_t st);
std::unique_ptr<Circuit> circuit_synthetic(Context& ctx, uint64_t seed, uint32_t m, uint32_t n_pub);
void circuit_write_zkey(const Circuit& c, const std::string& path);

// serialization (src/groth16.cu:1432-1557, src/main.cu:28-110)
std::string proof_to_json(const zk_proof& p);
void proof_to_calldata(const zk_proof& p, uint8_t out[256]);
std::string public_json(const uint64_t* w, uint32_t n_public);

}  // namespace zk
