/*
 * zk_abi.h — C ABI of the MI355X Groth16/BN254 prover (libzkp.so).
 *
 * Plain C types only: pointers, sizes, status codes. Every function returns
 * ZK_OK (0) or a nonzero status; zk_last_error() gives the message of the last
 * failure on the calling thread. No exception crosses this boundary.
 *
 * Data conventions (same bytes as the reference, include/groth16.cuh:57-73):
 *   field elements: 32 bytes = 4 x uint64 little-endian;
 *   G1 affine: x[4], y[4]; G2 affine: x.a0, x.a1, y.a0, y.a1 (Montgomery Fp,
 *   R = 2^256); the point at infinity is all zeros;
 *   scalars (witness, MSM scalars, r, s): canonical Fr (< r);
 *   NTT data: Montgomery Fr.
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef ZK_ABI_H
#define ZK_ABI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ZK_OK 0
#define ZK_E_ARG 1
#define ZK_E_HIP 2
#define ZK_E_IO 3
#define ZK_E_FORMAT 4
#define ZK_E_NOMEM 5
#define ZK_E_STATE 6
#define ZK_E_INTERNAL 7

typedef struct zk_ctx zk_ctx;
typedef struct zk_circuit zk_circuit;
typedef struct zk_server zk_server;
typedef struct zk_batch zk_batch;

/* Same layout as Groth16Proof (include/groth16.cuh:69-73). */
typedef struct zk_proof {
    uint64_t pi_a[8];
    uint64_t pi_b[16];
    uint64_t pi_c[8];
} zk_proof;

/* Same fields as Groth16ZkeyHeader (include/groth16.cuh:79-86). */
typedef struct zk_zkey_header {
    uint32_t n_constraints; /* domain size as stored in the zkey */
    uint32_t n_vars;
    uint32_t n_public;
    uint32_t domain_size;
    int32_t log_domain;
    uint64_t prime[4];
} zk_zkey_header;

const char* zk_last_error(void);
const char* zk_version(void);

/* Context: device, streams, NTT tables, workspace. Replaces the reference's
 * global state (g_ntt_* in include/ntt_bn254.cuh:50-53); contexts are
 * independent, calls on one context are serialised. */
int zk_ctx_create(int device, zk_ctx** out);
void zk_ctx_destroy(zk_ctx* ctx);

/* ---- prove (include/groth16.cuh:119-123 groth16Prove) ----
 * r, s: canonical blinding scalars; NULL means 0 (the reference's unblinded
 * proof). */
int zk_prove_files(zk_ctx* ctx, const char* zkey_path, const char* wtns_path, const uint64_t r[4], const uint64_t s[4],
                   zk_proof* out);

/* Resident circuit: the zkey is parsed and its CRS uploaded once
 * (the reference re-reads it per proof, src/groth16.cu:1151,1248-1281). */
int zk_circuit_load(zk_ctx* ctx, const char* zkey_path, zk_circuit** out);
void zk_circuit_free(zk_circuit* c);
/* info: [0]=n_vars [1]=n_public [2]=domain n [3]=log n [4]=nnzA [5]=nnzB [6]=device bytes */
int zk_circuit_info(const zk_circuit* c, uint64_t info[7]);
int zk_circuit_write_zkey(const zk_circuit* c, const char* path);
/* witness: n_witness >= n_vars canonical values (host memory) */
int zk_prove(zk_ctx* ctx, const zk_circuit* c, const uint64_t* witness, uint32_t n_witness, const uint64_t r[4],
             const uint64_t s[4], zk_proof* out);
/* same, witness already in device memory (n_vars x 32 bytes) */
int zk_prove_device(zk_ctx* ctx, const zk_circuit* c, const void* d_witness, const uint64_t r[4], const uint64_t s[4],
                    zk_proof* out);
/* MSM window override for the next proofs on ctx (0 = tuned table, else the
 * per-size default). The reference fixes the window at compile time
 * (include/msm_g1.cuh:11, MSM_G1_C_BITS). */
int zk_set_windows(zk_ctx* ctx, int window_witness, int window_h);

/* Window tables (no reference counterpart: the reference runs every MSM over
 * its bases as given, include/msm_g1.cuh:714-913). A circuit's resident
 * bases can also be held as W tables, table w holding 2^(offset of window w)
 * times every base, so that the MSMs of a proof put the digits of all windows
 * into one set of buckets: fewer, wider windows for the same reduction work.
 * Costs (W - 1) more copies of the 29-bit base arrays in device memory
 * (W = 6..10, by size). Proofs are the same either way.
 * zk_set_window_tables: mode for circuits loaded on ctx from now on: 1 on,
 * 0 off, -1 on when the tables take at most a third of the device memory
 * still free. A new context takes its mode from the environment variable
 * ZK_WINDOW_TABLES (-1, 0 or 1) when it is set, else -1.
 * zk_circuit_window_tables: builds (1) or frees (0) c's tables, or decides
 * as above (-1), or only reports (-2); windows (may be NULL) receives the
 * tables' (witness, h) windows, 0 when c has none. Waits for the device: no
 * proof of c may be in flight. A window forced with zk_set_windows whose
 * window count differs from the tables' proves without them. */
int zk_set_window_tables(zk_ctx* ctx, int mode);
int zk_circuit_window_tables(zk_ctx* ctx, zk_circuit* c, int mode, int windows[2]);

/* MSM window autotuner (no reference counterpart: the reference's window is a
 * compile-time constant, include/msm_g1.cuh:11). For circuit c it measures on
 * the device, over c's resident bases and uniform scalars, the digit sort plus
 * the MSMs that share it, for one window per window count in [11, 19] around
 * the size (at most ceil(log2 n) - 2) and every bucket split 1..4: kind 0 =
 * the witness sort over n_vars + 4 scalars and the A', B1', C' (G1) and B2'
 * (G2) MSMs; kind 1 = the sort of h and the H MSM (n scalars). The fastest
 * window, with its G1 and G2 splits, is recorded per (kind, ceil log2 n) in
 * the context's table, which every later proof on ctx uses unless
 * zk_set_windows forces one. force = 0 keeps size classes already in the
 * table. For a circuit with window tables the window is the tables' and only
 * the splits are tuned.
 * windows (may be NULL) receives the (witness, h) windows now in effect. */
int zk_autotune(zk_ctx* ctx, const zk_circuit* c, int force, int windows[2]);
/* window a proof on ctx would use for an MSM of kind (0 witness, 1 h) and size n
 * over a circuit without window tables (with them: zk_circuit_window_tables) */
int zk_tune_get(zk_ctx* ctx, int kind, uint64_t n, int* window);
/* merge / write the tuning table as text ("<arch> <kind> <ceil log2 n> <c>"
 * per line); entries for other GPU architectures are skipped on load */
int zk_tune_load(zk_ctx* ctx, const char* path, int* n_entries);
int zk_tune_save(zk_ctx* ctx, const char* path);

/* Captured batch (no reference counterpart: the reference proves one pair per
 * process, src/main.cu:115-171). A fixed sequence of count proofs over
 * circuits[i] and the device witnesses at d_witnesses[i] (n_vars x 32 bytes,
 * canonical) is recorded once as one HIP graph over `slots` workspace slots
 * (proof i on slot i % slots). zk_batch_graph_run replays it: the witnesses
 * are read from the same device addresses (the caller may change their
 * contents between runs), r and s (count x 4 words, NULL = zeros) are new per
 * run. Results are identical to zk_prove_device on the same inputs. */
int zk_batch_graph_create(zk_ctx* ctx, int count, const zk_circuit* const* circuits, const void* const* d_witnesses,
                          int slots, zk_batch** out);
int zk_batch_graph_run(zk_batch* b, const uint64_t* r, const uint64_t* s, zk_proof* out);
int zk_batch_graph_info(const zk_batch* b, uint64_t* n_nodes, int* n_slots);
void zk_batch_graph_destroy(zk_batch* b);

/* Batched proving: count independent proofs, possibly over different
 * circuits, overlapped on the device (several workspace slots and streams).
 * witnesses[i] is host memory; r, s are count x 4 words (NULL = zeros). */
int zk_prove_batch(zk_ctx* ctx, int count, const zk_circuit* const* circuits, const uint64_t* const* witnesses,
                   const uint64_t* r, const uint64_t* s, zk_proof* out);

/* Same, witnesses already in device memory (n_vars x 32 bytes each). */
int zk_prove_batch_device(zk_ctx* ctx, int count, const zk_circuit* const* circuits, const void* const* d_witnesses,
                          const uint64_t* r, const uint64_t* s, zk_proof* out);

/* number of workspace slots (proofs in flight) used by the batch entry points, 1..16 */
int zk_set_batch_slots(zk_ctx* ctx, int n);

/* Per-phase device timing of proofs (profiling events on the prover's own
 * streams; off by default). zk_prove_stats fills up to n values for the last
 * proof finished by zk_prove / zk_prove_device on ctx:
 *  [0] SpMV (+ witness copy) ms   [1] 3 x iNTT+coset ms   [2] 3 x NTT ms
 *  [3] h + sort(h) ms             [4] shared witness sort ms
 *  [5..9]  bucket-accumulation kernel ms for A, B1, C, B2, H
 *  [10..14] whole MSM (accumulation .. partials on host) ms, same order
 *  [15] digit entries of the witness sort  [16] of the h sort
 *  [17] window c (witness) [18] c (h) [19] windows W (witness) [20] W (h)
 *  [21] stream-0 span ms  [22] witness-MSM span ms (shared sort start to
 *  the last of the four MSMs)  [23] host tail ms (host work left once the
 *  last MSM's partials reached the host) */
int zk_set_profiling(zk_ctx* ctx, int on);
int zk_prove_stats(zk_ctx* ctx, double* out, int n);

/* ---- serving (no reference counterpart: the reference proves one file pair
 * per process, src/main.cu:115-171) ----
 * A worker thread admits submitted requests into workspace slots while other
 * proofs run (continuous batching), creating or growing a slot only if free
 * HBM stays above reserve_bytes. Requests over a domain of 2^23 or more
 * (m >= 2^22), each enough to fill the GPU, run one at a time; smaller ones
 * may pass a waiting one. Circuits must outlive their requests.
 * zk_server_destroy finishes every submitted request first. Destroying the
 * context first does the same for each of its servers, after which their
 * handles only accept zk_server_destroy (other calls return ZK_E_STATE). */
int zk_server_create(zk_ctx* ctx, int max_slots, uint64_t reserve_bytes, zk_server** out);
void zk_server_destroy(zk_server* s);
/* On an idle server (nothing queued or running; else ZK_E_STATE): create up
 * to max_slots workspace slots, while free HBM stays above the reserve, and
 * size each for every one of the count circuits, so that requests on them
 * never allocate device memory while other proofs run. *n_slots (may be
 * NULL) receives the slot count. */
int zk_server_prepare(zk_server* s, const zk_circuit* const* circuits, int count, int* n_slots);
/* witness: host, canonical, copied at submit; r, s NULL = 0 */
int zk_server_submit(zk_server* s, const zk_circuit* c, const uint64_t* witness, uint32_t n_witness, const uint64_t r[4],
                     const uint64_t sb[4], uint64_t* ticket);
/* witness already in device memory (n_vars x 32 bytes, canonical, as for
 * zk_prove_device); it is read when the request runs and must stay valid and
 * unchanged until the request completes */
int zk_server_submit_device(zk_server* s, const zk_circuit* c, const void* d_witness, const uint64_t r[4],
                            const uint64_t sb[4], uint64_t* ticket);
/* *done = 1 when finished (then the ticket is released); times (ms):
 * [0] queued, [1] on the device + host tail, [2] total latency. A failed
 * request returns its status here without affecting other requests. */
int zk_server_poll(zk_server* s, uint64_t ticket, int* done, zk_proof* out, double times[3]);
int zk_server_wait(zk_server* s, uint64_t ticket, zk_proof* out, double times[3]);
/* removes a request that has not started; *cancelled = 1 if it was removed */
int zk_server_cancel(zk_server* s, uint64_t ticket, int* cancelled);
/* [0] completed [1] failed [2] rejected [3] cancelled [4] queued [5] in flight [6] slots [7] peak in flight */
int zk_server_stats(zk_server* s, double out[8]);

/* ---- serialization (groth16ProofToSnarkjsJson / groth16ProofToCalldata,
 * include/groth16.cuh:149-167; public.json: src/main.cu:28-110) ---- */
int zk_proof_to_json(const zk_proof* p, char* out, size_t out_size);
int zk_proof_to_calldata(const zk_proof* p, uint8_t* out, size_t out_size);
int zk_public_json(const char* wtns_path, uint32_t n_public, char* out, size_t out_size);
/* groth16ReadZkeyHeader (include/groth16.cuh:177-180) */
int zk_read_zkey_header(const char* zkey_path, zk_zkey_header* out);
/* .wtns reader: *n_out = nVars; fills up to cap values (canonical) */
int zk_read_wtns(const char* wtns_path, uint64_t* w, uint32_t cap, uint32_t* n_out);
int zk_write_wtns(const char* wtns_path, const uint64_t* w, uint32_t n_vars);

/* groth16SparseMvm (include/groth16.cuh:183-205; declared but never defined
 * by the reference): out = M * w on the host for matrix_sel (0 = A, 1 = B,
 * 2 = C) over 44-byte entries {u32 matrix, u32 row, u32 wire, 32-byte value},
 * entries of other matrices skipped; witness canonical in, result canonical
 * out (domain_size x 4 words, zero-padded). zk_sparse_mvm_canonical takes
 * canonical values, the reference header's contract (:93-98); zk_sparse_mvm
 * takes the zkey §4 encoding c*R^2 mod r, so §4 bytes can be passed as they
 * are. Values and witness must be below r (ZK_E_ARG). */
int zk_sparse_mvm(const uint8_t* coefs44, size_t n_coefs, uint8_t matrix_sel, const uint64_t* witness,
                  uint32_t n_vars, uint64_t* out, uint32_t domain_size);
int zk_sparse_mvm_canonical(const uint8_t* coefs44, size_t n_coefs, uint8_t matrix_sel, const uint64_t* witness,
                            uint32_t n_vars, uint64_t* out, uint32_t domain_size);

/* ---- kernel-level entry points (host buffers), used by parity tests ----
 * NTT (include/ntt_bn254.cuh:418 ntt): data n = 2^log_n Montgomery Fr, in
 * place. kind 0 forward, 1 inverse, 2 coset forward (x_i * w_2n^i then NTT,
 * snarkjs convention), 3 coset inverse. */
int zk_ntt(zk_ctx* ctx, uint64_t* data, int log_n, int kind);
/* MSM (include/msm_g1.cuh:918 msm_g1, include/msm_g2.cuh:859 msmG2):
 * points AoS affine Montgomery (G1 8 words, G2 16 words per point), scalars
 * canonical; window 0 = default. out affine Montgomery, *is_inf set. */
int zk_msm_g1(zk_ctx* ctx, const uint64_t* points, const uint64_t* scalars, size_t n, int window, uint64_t out[8],
              int* is_inf);
int zk_msm_g2(zk_ctx* ctx, const uint64_t* points, const uint64_t* scalars, size_t n, int window, uint64_t out[16],
              int* is_inf);
/* SpMV over CSR with raw c*R^2 values (src/groth16.cu:144-167):
 * out[i] = sum frMul(val, w[col]) (Montgomery), rows = n_rows. */
int zk_spmv(zk_ctx* ctx, const uint32_t* row_ptr, const uint32_t* cols, const uint64_t* vals, uint32_t n_rows,
            const uint64_t* witness, uint32_t n_vars, uint64_t* out);
/* H pipeline (src/groth16.cu:1037-1090): aw, bw Montgomery, h canonical. */
int zk_hpoly(zk_ctx* ctx, const uint64_t* aw, const uint64_t* bw, int log_n, uint64_t* h);

/* ---- device-buffer variants (inputs resident in HBM; for benchmarks) ---- */
int zk_dev_alloc(zk_ctx* ctx, size_t bytes, void** out);
int zk_dev_free(zk_ctx* ctx, void* p);
int zk_dev_upload(zk_ctx* ctx, void* dst, const void* src, size_t bytes);
int zk_dev_download(zk_ctx* ctx, void* dst, const void* src, size_t bytes);
/* device-to-device copy on the context stream (synchronous) */
int zk_dev_copy(zk_ctx* ctx, void* dst, const void* src, size_t bytes);
/* NTT of `batch` arrays d_in[i] -> d_out[i] (distinct buffers; d_in clobbered) */
int zk_ntt_device(zk_ctx* ctx, void* const* d_in, void* const* d_out, int batch, int log_n, int kind);
int zk_msm_g1_device(zk_ctx* ctx, const void* d_points, const void* d_scalars, size_t n, int window, uint64_t out[8],
                     int* is_inf);
int zk_msm_g2_device(zk_ctx* ctx, const void* d_points, const void* d_scalars, size_t n, int window, uint64_t out[16],
                     int* is_inf);
/* zk_msm_g1_device without any allocation. No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Its whole device workspace
 * is a fixed 64 MiB array that the library holds on every device it runs on
 * (enough for 65536 points at the default window); larger workspaces are
 * refused with ZK_E_ARG. The per-block partial sums of every window go to a
 * fixed host array. Calls on one device share both arrays: calls that
 * overlap on one device race.
 * Results are identical to zk_msm_g1_device's. */
int zk_msm_g1_device_noalloc(zk_ctx* ctx, const void* d_points, const void* d_scalars, size_t n, int window,
                             uint64_t out[8], int* is_inf);
/* device time of the last kernel-level call on ctx (HIP events on its stream):
 * MSM: [0] total ms [1] digit sort ms [2] bucket-accumulation kernel ms
 *      [3] number of (point, bucket) entries accumulated
 *      [4] bucket reduction ms (end of the accumulation to the window values
 *          on the host);
 * NTT: [0] total ms, others 0 */
int zk_last_timing(zk_ctx* ctx, float ms[5]);
/* device pointer and point count of a resident base array: which = 0 A',
 * 1 B1', 2 B2' (G2), 3 C', 4 H; the primed arrays carry the 4 extra slots
 * described in the prover (alpha/beta/delta terms) after the n_vars bases */
int zk_circuit_bases(const zk_circuit* c, int which, void** d_ptr, uint64_t* count);
/* MSM of d_scalars (count of zk_circuit_bases, canonical, device) over the
 * resident base array `which` (as in zk_circuit_bases; 2 is G2, out[16]),
 * through the arrays the prover itself reads: the window tables when c has
 * them (window 0, or one with the tables' window count), else the 29-bit
 * copy at `window` (0 = tuned or default); for A', B1' and B2' of a circuit
 * with wire classes (enough infinite bases), the prover's class-ordered sort,
 * reading only the entries of finite bases; at the bucket split a proof of c
 * would use (autotuned, else the default). Same result as zk_msm_g1_device /
 * zk_msm_g2_device over zk_circuit_bases' points; timing in zk_last_timing
 * ([3]: the entries the accumulation read).
 * No counterpart in the reference (its MSMs run inside groth16Prove,
 * src/groth16.cu:1143-1397). */
int zk_circuit_msm_device(zk_ctx* ctx, const zk_circuit* c, int which, const void* d_scalars, int window,
                          uint64_t out[16], int* is_inf);

/* ---- synthetic workload (seeded; see the generator's recipe) ---- */
/* builds the R1CS + trapdoor setup for (seed, m, n_pub) and makes it resident */
int zk_synth_circuit(zk_ctx* ctx, uint64_t seed, uint32_t m, uint32_t n_pub, zk_circuit** out);
/* satisfying witness; public inputs from the seed (pub_seed == seed) or fresh (any other value) */
int zk_synth_witness(uint64_t seed, uint32_t m, uint32_t n_pub, uint64_t pub_seed, uint64_t* w_out);
/* satisfying witness of a circuit built by zk_synth_circuit, public inputs
 * drawn from pub_seed's stream (n_vars x 4 words); reuses the resident R1CS */
int zk_circuit_witness(const zk_circuit* c, uint64_t pub_seed, uint64_t* w_out);
/* the same witness computed on the device into d_w (n_vars x 32 bytes,
 * canonical), the constraint rows level by level in dependency order; returns
 * when d_w is written. For a fresh witness per request at any circuit size. */
int zk_circuit_witness_device(zk_ctx* ctx, const zk_circuit* c, uint64_t pub_seed, void* d_w);
int zk_synth_blinding(uint64_t seed, uint64_t r[4], uint64_t s[4]);
int zk_synth_uniform(uint64_t seed, size_t count, uint64_t* out);
/* n_vars / domain of the synthetic circuit (m, n_pub) */
int zk_synth_sizes(uint32_t m, uint32_t n_pub, uint32_t* n_vars, uint32_t* domain);

/* ---- Groth16 verification (host; no GPU) ----
 * Replaces `snarkjs groth16 verify` (the reference's check, Makefile:135):
 * the optimal ate pairing on BN254, e(A, B) == e(alpha1, beta2) e(vk_x, gamma2)
 * e(C, delta2) with vk_x = IC_0 + sum_i pub_i IC_i, after checking that A and
 * C lie on G1 and B on the twist in the order-r subgroup. *ok = 1 when the
 * proof verifies, 0 when it does not; the status is nonzero only for bad
 * arguments or files. publics: n_public canonical Fr values (4 words each),
 * the statement's public inputs w_1..w_l; the proof in the zk_proof layout
 * (Montgomery affine). The verification key comes from a resident circuit,
 * from zkey §2/§3, or, with public.json and proof.json texts as written by
 * zk_public_json / zk_proof_to_json, from a zkey file. */
int zk_verify(const zk_circuit* c, const uint64_t* publics, uint32_t n_public, const zk_proof* proof, int* ok);
int zk_verify_zkey(const char* zkey_path, const uint64_t* publics, uint32_t n_public, const zk_proof* proof, int* ok);
int zk_verify_json(const char* zkey_path, const char* public_json, const char* proof_json, int* ok);
/* e(P, Q) for P in G1 (8 words) and Q in G2 (16 words), Montgomery affine:
 * the Fp12 value as 6 Fp2 coefficients c0.c0, c0.c1, c0.c2, c1.c0, c1.c1,
 * c1.c2 of Fp12 = Fp6[w]/(w^2 - v), Fp6 = Fp2[v]/(v^3 - (9 + u)), each Fp2
 * as (a0, a1) Montgomery (48 words). For tests of the pairing itself. */
int zk_pairing(const uint64_t p[8], const uint64_t q[16], uint64_t out[48]);

/* ---- device field arithmetic, element-wise (tests of the field code) ----
 * No counterpart entry point in the reference, whose unit tests call fpMul,
 * frMul etc. from test kernels (test/unit/). zk_field_op_device runs the
 * standard form (8 x 32-bit, Montgomery R = 2^256) of field (0 = Fp, 1 = Fr)
 * on n elements of 4 words: op 0 add, 1 sub, 2 neg, 3 Montgomery product (the
 * product-scanning asm), 4 square, 5 inverse (Montgomery in and out), 6 to
 * Montgomery, 7 from Montgomery, 8 Montgomery product by the C++ CIOS form.
 * zk_field29_op_device runs the bucket loops' 9 x 29-bit form of Fp
 * (Montgomery R' = 2^261) on raw limbs, 9 words per element, limbs as given
 * (no range check: inputs outside the form's documented bounds give
 * unspecified values): op 0 product, 1 square, 2 a*b + b*a with one
 * reduction, 3 a*a + a*b + b*a + b*b with one reduction, 4 negation of a
 * value < p (2p - a, unnormalized), 5 a + 4p - b normalized, 6 reduction of a
 * value < 16p below 2p, 7 conversion to the standard form and back, 8
 * negation 4p - a of a normalized a < 2p (unnormalized), 9 the G1 mixed
 * addition's Y3 sum a*(a + 4p - b) + b*(4p - b) with one reduction (the
 * difference and the negation unnormalized); on element pairs (c0, c1) of
 * Fp2, n even: 10 the Fp2 product a*b, 11 a*b - b*b as the G2 Y3's one
 * reduction per component. */
int zk_field_op_device(zk_ctx* ctx, int field, int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n);
int zk_field29_op_device(zk_ctx* ctx, int op, const uint32_t* a, const uint32_t* b, uint32_t* out, size_t n);

/* ---- measurement probes (no reference counterpart) ----
 * zk_probe_valu: lane operations per second over the whole GPU of
 * kind & 15: 0 v_mad_u64_u32 (8 independent chains per thread), 1 v_add_u32,
 * 2 v_fma_f64 (8 chains), 3 v_mad_u64_u32 in one dependent chain, 4 / 5
 * dependent chains of the bucket accumulation's mixed XYZZ additions / the
 * bucket reduction's general ones (both in the 29-bit form), in point
 * additions per second; with
 * kind >> 4 waves per SIMD (0: 8). Kind 0 is the issue-rate ceiling behind
 * the MSM roofline; the others show what a lone wave can issue.
 * zk_circuit_spmv_device: the circuit's fused A*w, B*w, A*w o B*w kernel
 * (src/groth16.cu:144-167) alone on the device from d_witness (n_vars x 32
 * bytes canonical) into three domain-size arrays; zk_last_timing()[0] gives
 * its device time. */
int zk_probe_valu(zk_ctx* ctx, int kind, double* lane_ops_per_s);
int zk_circuit_spmv_device(zk_ctx* ctx, const zk_circuit* c, const void* d_witness, void* d_aw, void* d_bw, void* d_cw);
/* Check of the admission arithmetic (no reference counterpart): for a fresh
 * proof slot on ctx, out[0] = the device bytes counted for a proof of c at the
 * context's windows and splits (what serving admission uses), out[1] = the
 * bytes the slot holds once reserved for c, out[2] = the bytes still counted
 * after that (0 when the count covers every buffer). The context's NTT plan
 * for c's domain is created first and is in neither figure. */
int zk_slot_bytes_check(zk_ctx* ctx, const zk_circuit* c, uint64_t out[3]);

/* ---- host-side check of the MSM's scalar split (no GPU) ---- */
/* No counterpart in the reference, whose MSMs run over full 254-bit scalars
 * (include/msm_g1.cuh:271-329). The GLV split the MSM kernels apply to every
 * scalar, run on the host from the same source: canonical k (4 words each) ->
 * out[4i..4i+1] = k1, out[4i+2..4i+3] = k2, each a magnitude < 2^127 with the
 * sign in bit 127, k = k1 + k2*lambda (mod r). For tests and tools. */
int zk_glv_split(const uint64_t* k, size_t n, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif /* ZK_ABI_H */
