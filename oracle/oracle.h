/*
 * CPU oracle for the Groth16/BN254 prove path — TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product (zyklop_amd) never
 * links or calls it.
 *
 * Arithmetic is exact and deliberately independent of the device code: field
 * products are full 512-bit products reduced by GMP long division (mpn_mul_n +
 * mpn_tdiv_qr), inverses use mpz_invert, and Montgomery form is only applied at
 * the boundaries (value * 2^256 mod m).
 *
 * Conventions (all restated from the reference, file:line cited per function
 * in oracle.c):
 *   - field elements: 4 x uint64 little-endian limbs;
 *   - G1 affine point: x[4], y[4]; G2 affine: x.a0, x.a1, y.a0, y.a1;
 *     infinity is (0,0) plus an explicit flag where one is returned;
 *   - "mont" I/O means coordinates are x*2^256 mod p (snarkjs .zkey layout).
 *
 * Parity status of each function is listed in oracle.c's header.
 */
#ifndef ZK_ORACLE_H
#define ZK_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- field ops, vectorised over count elements ---- */
enum { OR_FP = 0, OR_FR = 1 };
enum {
    OR_ADD = 0, OR_SUB = 1, OR_NEG = 2, OR_MONT_MUL = 3, OR_MONT_SQR = 4,
    OR_INV = 5,        /* canonical in/out, inv(0) = 0 */
    OR_TO_MONT = 6, OR_FROM_MONT = 7,
    OR_MUL = 8,        /* canonical product */
    OR_MONT_POW = 9    /* r = a^e (a, r Montgomery), e = b[0] (u64) */
};
int or_field_op(int field, int op, const uint64_t* a, const uint64_t* b, uint64_t* r, size_t count);

/* ---- curve ops (canonical affine coordinates) ---- */
int or_g1_add(const uint64_t p[8], int p_inf, const uint64_t q[8], int q_inf, uint64_t r[8], int* r_inf);
int or_g1_mul(const uint64_t p[8], int p_inf, const uint64_t k[4], uint64_t r[8], int* r_inf);
int or_g1_on_curve(const uint64_t p[8]);
int or_g2_mul(const uint64_t p[16], int p_inf, const uint64_t k[4], uint64_t r[16], int* r_inf);
int or_g2_on_curve(const uint64_t p[16]);
void or_g1_generator(uint64_t g[8]);
void or_g2_generator(uint64_t g[16]);

/* ---- MSM: points Montgomery affine SoA (px[n][4], py[n][4]); scalars
 * canonical; result Montgomery affine (G1: x,y; G2: x.a0,x.a1,y.a0,y.a1). ---- */
int or_msm_g1(const uint64_t* px, const uint64_t* py, const uint64_t* scalars, size_t n,
              uint64_t out[8], int* out_inf);
int or_msm_g2(const uint64_t* px, const uint64_t* py, const uint64_t* scalars, size_t n,
              uint64_t out[16], int* out_inf);
/* brute force: sum of double-and-add products (independent of Pippenger) */
int or_msm_g1_naive(const uint64_t* px, const uint64_t* py, const uint64_t* scalars, size_t n,
                    uint64_t out[8], int* out_inf);

/* ---- NTT over Fr, natural order in/out, Montgomery I/O when mont != 0 ----
 * kind 0: forward   X_k = sum_j x_j w^{jk},          w = w_n from generator 5
 * kind 1: inverse   x_j = n^-1 sum_k X_k w^{-jk}
 * kind 2: coset     X_k = sum_j (x_j w_2n^j) w^{jk}  (snarkjs odd coset)
 * kind 3: coset inverse (undo kind 2)
 * kind 4: naive O(n^2) forward DFT (independent check, small n only) */
int or_ntt(uint64_t* data, int log_n, int kind, int mont);
int or_root_of_unity(int log_n, uint64_t w[4]); /* canonical w_{2^log_n} */

/* ---- QAP: SpMV with raw = c*R^2 coefficients (Montgomery result) ----
 * rows: CSR row_ptr[n_rows+1], cols[nnz], vals[nnz][4] (raw c*R^2);
 * w: canonical witness; out[n_rows][4] Montgomery. */
int or_spmv(const uint32_t* row_ptr, const uint32_t* cols, const uint64_t* vals, size_t n_rows,
            const uint64_t* w, size_t n_vars, uint64_t* out);
/* H pipeline: aw, bw Montgomery [n]; h out canonical [n] */
int or_hpoly(const uint64_t* aw, const uint64_t* bw, int log_n, uint64_t* h);

/* ---- synthetic circuit (seeded R1CS + witness + trapdoor setup) ---- */
typedef struct or_circuit or_circuit;
or_circuit* or_circuit_create(uint64_t seed, uint32_t m, uint32_t n_pub, int with_crs);
void or_circuit_free(or_circuit* c);
/* sizes: [0]=m, [1]=n_pub, [2]=n_vars, [3]=domain n, [4]=log n, [5]=nnzA, [6]=nnzB */
void or_circuit_sizes(const or_circuit* c, uint64_t out[7]);
/* witness canonical [n_vars][4] */
void or_circuit_witness(const or_circuit* c, uint64_t* w);
/* blinding (canonical) derived from the seed */
void or_circuit_blinding(const or_circuit* c, uint64_t r[4], uint64_t s[4]);
int or_circuit_write_zkey(const or_circuit* c, const char* path);
int or_circuit_write_wtns(const or_circuit* c, const char* path);
/* proof Montgomery affine: piA[8], piB[16], piC[8] concatenated (32 u64) */
int or_prove(const or_circuit* c, const uint64_t r[4], const uint64_t s[4], uint64_t proof[32]);
/* closed-form proof from the trapdoor (works without the CRS) */
int or_prove_trapdoor(const or_circuit* c, const uint64_t r[4], const uint64_t s[4], uint64_t proof[32]);
/* witness-replaced variant: use the given canonical witness instead */
int or_prove_trapdoor_w(const or_circuit* c, const uint64_t* w, const uint64_t r[4], const uint64_t s[4],
                        uint64_t proof[32]);
/* Groth16 check with the trapdoor: proof points == closed form and the
 * verification equation holds in the exponent. 1 = ok. */
int or_verify_trapdoor(const or_circuit* c, const uint64_t* w, const uint64_t r[4], const uint64_t s[4],
                       const uint64_t proof[32]);
/* exponents (canonical Fr) of the toxic waste: tau, alpha, beta, gamma, delta */
void or_circuit_toxic(const or_circuit* c, uint64_t out[20]);

/* ---- helpers for checks at full size ---- */
/* out[i] = (k0 + i*d)*G, i < n: Montgomery affine AoS x[4], y[4] (known logs) */
int or_g1_progression(const uint64_t k0[4], const uint64_t d[4], size_t n, uint64_t* out);
/* out[q] = sum_j coeffs[j] * zs[q]^j mod r (z canonical; coefficient form carries through) */
int or_fr_poly_eval(const uint64_t* coeffs, size_t n, const uint64_t* zs, size_t nz, uint64_t* out);
/* sum_i a_i * b_i mod r, canonical */
int or_fr_dot(const uint64_t* a, const uint64_t* b, size_t n, uint64_t out[4]);

/* threads of the parallel loops (MSM windows, NTT butterflies, SpMV rows, CRS
 * points); n <= 0 keeps the setting. Returns the count now in effect. */
int or_set_threads(int n);

/* decimal string of a canonical 256-bit value */
void or_to_dec(const uint64_t a[4], char* buf, size_t buf_size);

#ifdef __cplusplus
}
#endif
#endif
