/*
 * CPU oracle for the Groth16/BN254 prove path — TEST INFRASTRUCTURE ONLY.
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker; never linked or called by the product.
 *
 * It restates the reference's algorithm with independent arithmetic (GMP
 * long-division reduction instead of Montgomery CIOS, Jacobian formulas
 * instead of the device's XYZZ buckets, unsigned Pippenger windows instead of
 * signed digits).
 *
 * Parity pins (what checks each function against something other than itself):
 *   field ops      test/unit/test_vectors.json (fp_add/sub/neg/mul/sqr/inv),
 *                  test/unit/test_vectors_fr.json (frAdd..frPow, 2697 vectors),
 *                  plus Python big-int closed forms in tests/test_oracle.py
 *   G1 add / mul   test_vectors.json curve_g1 (34 vectors incl. k*G)
 *   MSM G1         test_vectors_msm.json (16) + brute force + discrete-log
 *                  closed form (points k_i*G, result (sum s_i k_i)*G)
 *   MSM G2         test_vectors_msm_g2.json (12) + discrete-log closed form
 *   NTT fwd/inv    test_vectors_ntt.json (7) + O(n^2) DFT + round trips
 *   coset NTT, H   no reference fixture: brute-force evaluation of the
 *                  interpolant at w_2n^(2j+1) (tests/test_oracle.py)
 *   SpMV           closed form with Python big ints (c*R^2 convention)
 *   prove          algorithmic prover == trapdoor closed form, and the
 *                  Groth16 equation holds in the exponent (or_verify_trapdoor)
 *   zkey/wtns I/O  tests/test_zkey_layout.py parses a written zkey/wtns at
 *                  the reference reader's offsets and checks every field
 *                  against values fixed independently (the R1CS regenerated
 *                  by a Python restatement of the generator, CRS points from
 *                  exponents computed with Python big ints); no snarkjs-made
 *                  file exists in the reference to compare bytes with
 *   g1_progression, fr_poly_eval, fr_dot (helpers of the full-size GPU
 *                  checks): textbook double-and-add, the DFT definition and
 *                  Python big ints (tests/test_oracle.py)
 * Several loops run under OpenMP (MSM windows, NTT butterflies, SpMV rows,
 * CRS generation); or_set_threads chooses the thread count.
 */
#include "oracle.h"

#include <gmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

_Static_assert(sizeof(mp_limb_t) == 8, "64-bit GMP limbs required");

typedef struct { uint64_t v[4]; } fe;
typedef struct { fe a0, a1; } fe2;

/* BN254 moduli (include/fp_bn254.cuh:48-51, include/fr_bn254.cuh:34-37) */
static const fe P_MOD = {{0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}};
static const fe R_MOD = {{0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}};

static fe RINV_P, RINV_R;   /* 2^-256 mod p, mod r */
static fe R2_R;             /* 2^512 mod r */
static int g_init = 0;

/* ---------------- raw field arithmetic (canonical values) ---------------- */
static int fe_is_zero(const fe* a) { return (a->v[0] | a->v[1] | a->v[2] | a->v[3]) == 0; }
static int fe_eq(const fe* a, const fe* b) { return memcmp(a, b, sizeof(fe)) == 0; }
static fe fe_zero(void) { fe z; memset(&z, 0, sizeof z); return z; }
static fe fe_u64(uint64_t x) { fe z = fe_zero(); z.v[0] = x; return z; }

static void f_add(fe* r, const fe* a, const fe* b, const fe* m) {
    mp_limb_t c = mpn_add_n((mp_limb_t*)r->v, (const mp_limb_t*)a->v, (const mp_limb_t*)b->v, 4);
    if (c || mpn_cmp((const mp_limb_t*)r->v, (const mp_limb_t*)m->v, 4) >= 0)
        mpn_sub_n((mp_limb_t*)r->v, (const mp_limb_t*)r->v, (const mp_limb_t*)m->v, 4);
}
static void f_sub(fe* r, const fe* a, const fe* b, const fe* m) {
    mp_limb_t br = mpn_sub_n((mp_limb_t*)r->v, (const mp_limb_t*)a->v, (const mp_limb_t*)b->v, 4);
    if (br) mpn_add_n((mp_limb_t*)r->v, (const mp_limb_t*)r->v, (const mp_limb_t*)m->v, 4);
}
static void f_neg(fe* r, const fe* a, const fe* m) {
    if (fe_is_zero(a)) { *r = fe_zero(); return; }
    mpn_sub_n((mp_limb_t*)r->v, (const mp_limb_t*)m->v, (const mp_limb_t*)a->v, 4);
}
static void f_mul(fe* r, const fe* a, const fe* b, const fe* m) {
    mp_limb_t prod[8], q[5], rem[4];
    mpn_mul_n(prod, (const mp_limb_t*)a->v, (const mp_limb_t*)b->v, 4);
    mpn_tdiv_qr(q, rem, 0, prod, 8, (const mp_limb_t*)m->v, 4);
    memcpy(r->v, rem, 32);
}
/* reduce an arbitrary 256-bit value mod m */
static void f_reduce(fe* r, const fe* a, const fe* m) {
    mp_limb_t q[2], rem[4];
    mpn_tdiv_qr(q, rem, 0, (const mp_limb_t*)a->v, 4, (const mp_limb_t*)m->v, 4);
    memcpy(r->v, rem, 32);
}
static void f_inv(fe* r, const fe* a, const fe* m) {
    if (fe_is_zero(a)) { *r = fe_zero(); return; }
    mpz_t x, mm, y;
    mpz_init(x); mpz_init(mm); mpz_init(y);
    mpz_import(x, 4, -1, 8, 0, 0, a->v);
    mpz_import(mm, 4, -1, 8, 0, 0, m->v);
    mpz_invert(y, x, mm);
    *r = fe_zero();
    mpz_export(r->v, NULL, -1, 8, 0, 0, y);
    mpz_clear(x); mpz_clear(mm); mpz_clear(y);
}
static void f_pow(fe* r, const fe* a, const fe* e, const fe* m) {
    fe acc = fe_u64(1), base = *a;
    for (int w = 3; w >= 0; w--)
        for (int b = 63; b >= 0; b--) {
            f_mul(&acc, &acc, &acc, m);
            if ((e->v[w] >> b) & 1) f_mul(&acc, &acc, &base, m);
        }
    *r = acc;
}
/* a * 2^256 mod m */
static void f_to_mont(fe* r, const fe* a, const fe* m) {
    mp_limb_t num[8] = {0, 0, 0, 0, a->v[0], a->v[1], a->v[2], a->v[3]}, q[5], rem[4];
    mpn_tdiv_qr(q, rem, 0, num, 8, (const mp_limb_t*)m->v, 4);
    memcpy(r->v, rem, 32);
}
static void f_from_mont(fe* r, const fe* a, const fe* m) {
    f_mul(r, a, (m == &P_MOD) ? &RINV_P : &RINV_R, m);
}

static void or_init(void) {
    if (g_init) return;
    fe two256p, two256r;
    fe one = fe_u64(1);
    f_to_mont(&two256p, &one, &P_MOD);
    f_to_mont(&two256r, &one, &R_MOD);
    f_inv(&RINV_P, &two256p, &P_MOD);
    f_inv(&RINV_R, &two256r, &R_MOD);
    f_mul(&R2_R, &two256r, &two256r, &R_MOD);
    g_init = 1;
}

/* Fp / Fr shorthands */
#define PADD(r, a, b) f_add(r, a, b, &P_MOD)
#define PSUB(r, a, b) f_sub(r, a, b, &P_MOD)
#define PMUL(r, a, b) f_mul(r, a, b, &P_MOD)
#define RADD(r, a, b) f_add(r, a, b, &R_MOD)
#define RSUB(r, a, b) f_sub(r, a, b, &R_MOD)
#define RMUL(r, a, b) f_mul(r, a, b, &R_MOD)

/* ---------------- Fp2 = Fp[u]/(u^2+1)  (include/fp2_bn254.cuh:96-180) ---------------- */
static fe2 fe2_zero(void) { fe2 z; memset(&z, 0, sizeof z); return z; }
static int fe2_is_zero(const fe2* a) { return fe_is_zero(&a->a0) && fe_is_zero(&a->a1); }
static int fe2_eq(const fe2* a, const fe2* b) { return fe_eq(&a->a0, &b->a0) && fe_eq(&a->a1, &b->a1); }
static void fe2_add(fe2* r, const fe2* a, const fe2* b) { PADD(&r->a0, &a->a0, &b->a0); PADD(&r->a1, &a->a1, &b->a1); }
static void fe2_sub(fe2* r, const fe2* a, const fe2* b) { PSUB(&r->a0, &a->a0, &b->a0); PSUB(&r->a1, &a->a1, &b->a1); }
static void fe2_mul(fe2* r, const fe2* a, const fe2* b) {
    fe t0, t1, t2, t3;
    PMUL(&t0, &a->a0, &b->a0);
    PMUL(&t1, &a->a1, &b->a1);
    PMUL(&t2, &a->a0, &b->a1);
    PMUL(&t3, &a->a1, &b->a0);
    PSUB(&r->a0, &t0, &t1);
    PADD(&r->a1, &t2, &t3);
}
static void fe2_inv(fe2* r, const fe2* a) {
    fe t0, t1, n, ni;
    PMUL(&t0, &a->a0, &a->a0);
    PMUL(&t1, &a->a1, &a->a1);
    PADD(&n, &t0, &t1);
    f_inv(&ni, &n, &P_MOD);
    fe neg1;
    f_neg(&neg1, &a->a1, &P_MOD);
    PMUL(&r->a0, &a->a0, &ni);
    PMUL(&r->a1, &neg1, &ni);
}

/* ---------------- generic Jacobian EC over Fp (G1) and Fp2 (G2), a = 0 ---------------- */
/* Textbook Jacobian formulas (x = X/Z^2, y = Y/Z^3), infinity <=> Z = 0. */
#define DEFINE_EC(PFX, F, FADD, FSUB, FMUL, FINV, FISZERO, FEQ, FZERO, FONE)                  \
    typedef struct { F x, y; int inf; } PFX##_aff;                                            \
    typedef struct { F X, Y, Z; } PFX##_jac;                                                  \
    static PFX##_jac PFX##_jac_inf(void) { PFX##_jac r; r.X = FONE(); r.Y = FONE(); r.Z = FZERO(); return r; } \
    static int PFX##_jac_is_inf(const PFX##_jac* p) { return FISZERO(&p->Z); }               \
    static PFX##_jac PFX##_from_aff(const PFX##_aff* a) {                                     \
        if (a->inf) return PFX##_jac_inf();                                                   \
        PFX##_jac r; r.X = a->x; r.Y = a->y; r.Z = FONE(); return r;                          \
    }                                                                                         \
    static PFX##_jac PFX##_dbl(const PFX##_jac* p) {                                          \
        if (PFX##_jac_is_inf(p) || FISZERO(&p->Y)) return PFX##_jac_inf();                   \
        F yy, s, m, t, x3, y3, z3, yyyy;                                                      \
        FMUL(&yy, &p->Y, &p->Y);                                                              \
        FMUL(&s, &p->X, &yy); FADD(&s, &s, &s); FADD(&s, &s, &s);       /* S = 4XY^2 */       \
        FMUL(&m, &p->X, &p->X); FADD(&t, &m, &m); FADD(&m, &t, &m);     /* M = 3X^2 */        \
        FMUL(&x3, &m, &m); FSUB(&x3, &x3, &s); FSUB(&x3, &x3, &s);      /* X3 = M^2-2S */     \
        FMUL(&yyyy, &yy, &yy); FADD(&yyyy, &yyyy, &yyyy); FADD(&yyyy, &yyyy, &yyyy);          \
        FADD(&yyyy, &yyyy, &yyyy);                                      /* 8Y^4 */            \
        FSUB(&t, &s, &x3); FMUL(&y3, &m, &t); FSUB(&y3, &y3, &yyyy);                          \
        FMUL(&z3, &p->Y, &p->Z); FADD(&z3, &z3, &z3);                                         \
        PFX##_jac r; r.X = x3; r.Y = y3; r.Z = z3; return r;                                  \
    }                                                                                         \
    static PFX##_jac PFX##_add(const PFX##_jac* p, const PFX##_jac* q) {                      \
        if (PFX##_jac_is_inf(p)) return *q;                                                   \
        if (PFX##_jac_is_inf(q)) return *p;                                                   \
        F z1z1, z2z2, u1, u2, s1, s2, t;                                                      \
        FMUL(&z1z1, &p->Z, &p->Z); FMUL(&z2z2, &q->Z, &q->Z);                                 \
        FMUL(&u1, &p->X, &z2z2); FMUL(&u2, &q->X, &z1z1);                                     \
        FMUL(&t, &q->Z, &z2z2); FMUL(&s1, &p->Y, &t);                                         \
        FMUL(&t, &p->Z, &z1z1); FMUL(&s2, &q->Y, &t);                                         \
        if (FEQ(&u1, &u2)) {                                                                  \
            if (FEQ(&s1, &s2)) return PFX##_dbl(p);                                           \
            return PFX##_jac_inf();                                                           \
        }                                                                                     \
        F h, rr, hh, hhh, v, x3, y3, z3;                                                      \
        FSUB(&h, &u2, &u1); FSUB(&rr, &s2, &s1);                                              \
        FMUL(&hh, &h, &h); FMUL(&hhh, &hh, &h); FMUL(&v, &u1, &hh);                           \
        FMUL(&x3, &rr, &rr); FSUB(&x3, &x3, &hhh); FSUB(&x3, &x3, &v); FSUB(&x3, &x3, &v);    \
        FSUB(&t, &v, &x3); FMUL(&y3, &rr, &t); FMUL(&t, &s1, &hhh); FSUB(&y3, &y3, &t);       \
        FMUL(&z3, &p->Z, &q->Z); FMUL(&z3, &z3, &h);                                          \
        PFX##_jac r; r.X = x3; r.Y = y3; r.Z = z3; return r;                                  \
    }                                                                                         \
    static PFX##_aff PFX##_to_aff(const PFX##_jac* p) {                                       \
        PFX##_aff r;                                                                          \
        if (PFX##_jac_is_inf(p)) { r.x = FZERO(); r.y = FZERO(); r.inf = 1; return r; }       \
        F zi, zi2, zi3;                                                                       \
        FINV(&zi, &p->Z); FMUL(&zi2, &zi, &zi); FMUL(&zi3, &zi2, &zi);                        \
        FMUL(&r.x, &p->X, &zi2); FMUL(&r.y, &p->Y, &zi3); r.inf = 0; return r;                \
    }                                                                                         \
    static PFX##_jac PFX##_mul_jac(const PFX##_jac* p, const fe* k) {                         \
        PFX##_jac acc = PFX##_jac_inf();                                                      \
        for (int w = 3; w >= 0; w--)                                                          \
            for (int b = 63; b >= 0; b--) {                                                   \
                acc = PFX##_dbl(&acc);                                                        \
                if ((k->v[w] >> b) & 1) acc = PFX##_add(&acc, p);                             \
            }                                                                                 \
        return acc;                                                                           \
    }                                                                                         \
    /* Pippenger with unsigned c-bit windows over canonical scalars < 2^256;  \
     * the window sums are independent and run in parallel */                               \
    static PFX##_jac PFX##_msm(const PFX##_aff* pts, const fe* sc, size_t n) {                \
        int c = 1;                                                                            \
        while (c < 16 && ((size_t)1 << (c + 2)) < n) c++;                                     \
        size_t nb = ((size_t)1 << c) - 1;                                                     \
        int nw = (256 + c - 1) / c;                                                           \
        PFX##_jac* S = (PFX##_jac*)malloc(sizeof(PFX##_jac) * nw);                            \
        _Pragma("omp parallel for schedule(dynamic, 1)")                                      \
        for (int w = 0; w < nw; w++) {                                                        \
            PFX##_jac* bk = (PFX##_jac*)malloc(sizeof(PFX##_jac) * nb);                       \
            for (size_t b = 0; b < nb; b++) bk[b] = PFX##_jac_inf();                          \
            for (size_t i = 0; i < n; i++) {                                                  \
                if (pts[i].inf) continue;                                                     \
                unsigned bit = (unsigned)(w * c);                                             \
                uint64_t d = 0;                                                               \
                for (int q = 0; q < c; q++) {                                                 \
                    unsigned bb = bit + q;                                                    \
                    if (bb < 256 && ((sc[i].v[bb >> 6] >> (bb & 63)) & 1)) d |= (uint64_t)1 << q; \
                }                                                                             \
                if (!d) continue;                                                             \
                PFX##_jac pj = PFX##_from_aff(&pts[i]);                                       \
                bk[d - 1] = PFX##_add(&bk[d - 1], &pj);                                       \
            }                                                                                 \
            PFX##_jac run = PFX##_jac_inf(), sum = PFX##_jac_inf();                           \
            for (size_t b = nb; b-- > 0;) {                                                   \
                run = PFX##_add(&run, &bk[b]);                                                \
                sum = PFX##_add(&sum, &run);                                                  \
            }                                                                                 \
            S[w] = sum;                                                                       \
            free(bk);                                                                         \
        }                                                                                     \
        PFX##_jac total = PFX##_jac_inf();                                                    \
        for (int w = nw - 1; w >= 0; w--) {                                                   \
            for (int d = 0; d < c; d++) total = PFX##_dbl(&total);                            \
            total = PFX##_add(&total, &S[w]);                                                 \
        }                                                                                     \
        free(S);                                                                              \
        return total;                                                                         \
    }

static void fp_add_(fe* r, const fe* a, const fe* b) { PADD(r, a, b); }
static void fp_sub_(fe* r, const fe* a, const fe* b) { PSUB(r, a, b); }
static void fp_mul_(fe* r, const fe* a, const fe* b) { PMUL(r, a, b); }
static void fp_inv_(fe* r, const fe* a) { f_inv(r, a, &P_MOD); }
static fe fp_one_(void) { return fe_u64(1); }
static fe2 fe2_one(void) { fe2 r = fe2_zero(); r.a0 = fe_u64(1); return r; }

DEFINE_EC(g1, fe, fp_add_, fp_sub_, fp_mul_, fp_inv_, fe_is_zero, fe_eq, fe_zero, fp_one_)
DEFINE_EC(g2, fe2, fe2_add, fe2_sub, fe2_mul, fe2_inv, fe2_is_zero, fe2_eq, fe2_zero, fe2_one)

/* generators (canonical). G1 = (1, 2); G2 from EIP-197 / snarkjs bn128. */
static g1_aff G1_GEN(void) { g1_aff g; g.x = fe_u64(1); g.y = fe_u64(2); g.inf = 0; return g; }
static g2_aff G2_GEN(void) {
    g2_aff g;
    g.x.a0 = (fe){{0x46debd5cd992f6edULL, 0x674322d4f75edaddULL, 0x426a00665e5c4479ULL, 0x1800deef121f1e76ULL}};
    g.x.a1 = (fe){{0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL, 0x7260bfb731fb5d25ULL, 0x198e9393920d483aULL}};
    g.y.a0 = (fe){{0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL, 0x4aab71808dcb408fULL, 0x12c85ea5db8c6debULL}};
    g.y.a1 = (fe){{0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL, 0xec9e99ad690c3395ULL, 0x090689d0585ff075ULL}};
    g.inf = 0;
    return g;
}

/* curve equations: y^2 = x^3 + 3 ; twist y^2 = x^3 + 3/(9+u) */
static int g1_on_curve_(const g1_aff* p) {
    if (p->inf) return 1;
    fe l, r, t, three = fe_u64(3);
    PMUL(&l, &p->y, &p->y);
    PMUL(&t, &p->x, &p->x); PMUL(&r, &t, &p->x); PADD(&r, &r, &three);
    return fe_eq(&l, &r);
}
static int g2_on_curve_(const g2_aff* p) {
    if (p->inf) return 1;
    fe2 l, r, t, b, nine_u;
    nine_u.a0 = fe_u64(9); nine_u.a1 = fe_u64(1);
    fe2 three = fe2_zero(); three.a0 = fe_u64(3);
    fe2 inv9u; fe2_inv(&inv9u, &nine_u);
    fe2_mul(&b, &three, &inv9u);
    fe2_mul(&l, &p->y, &p->y);
    fe2_mul(&t, &p->x, &p->x); fe2_mul(&r, &t, &p->x); fe2_add(&r, &r, &b);
    return fe2_eq(&l, &r);
}

/* ---------------- packing helpers ---------------- */
static g1_aff g1_unpack(const uint64_t* p, int inf) {
    g1_aff a; memcpy(a.x.v, p, 32); memcpy(a.y.v, p + 4, 32); a.inf = inf; return a;
}
static void g1_pack(const g1_aff* a, uint64_t* p) {
    if (a->inf) { memset(p, 0, 64); return; }
    memcpy(p, a->x.v, 32); memcpy(p + 4, a->y.v, 32);
}
static g2_aff g2_unpack(const uint64_t* p, int inf) {
    g2_aff a;
    memcpy(a.x.a0.v, p, 32); memcpy(a.x.a1.v, p + 4, 32);
    memcpy(a.y.a0.v, p + 8, 32); memcpy(a.y.a1.v, p + 12, 32);
    a.inf = inf; return a;
}
static void g2_pack(const g2_aff* a, uint64_t* p) {
    if (a->inf) { memset(p, 0, 128); return; }
    memcpy(p, a->x.a0.v, 32); memcpy(p + 4, a->x.a1.v, 32);
    memcpy(p + 8, a->y.a0.v, 32); memcpy(p + 12, a->y.a1.v, 32);
}
/* Montgomery (zkey layout) <-> canonical; (0,0) means infinity (include/msm_g1.cuh:243-244) */
static g1_aff g1_from_mont(const uint64_t* x, const uint64_t* y) {
    g1_aff a;
    fe X, Y; memcpy(X.v, x, 32); memcpy(Y.v, y, 32);
    if (fe_is_zero(&X) && fe_is_zero(&Y)) { a.x = fe_zero(); a.y = fe_zero(); a.inf = 1; return a; }
    f_from_mont(&a.x, &X, &P_MOD); f_from_mont(&a.y, &Y, &P_MOD); a.inf = 0; return a;
}
static void g1_to_mont(const g1_aff* a, uint64_t* x, uint64_t* y) {
    if (a->inf) { memset(x, 0, 32); memset(y, 0, 32); return; }
    fe X, Y; f_to_mont(&X, &a->x, &P_MOD); f_to_mont(&Y, &a->y, &P_MOD);
    memcpy(x, X.v, 32); memcpy(y, Y.v, 32);
}
static g2_aff g2_from_mont(const uint64_t* x, const uint64_t* y) {
    g2_aff a;
    fe t[4];
    memcpy(t[0].v, x, 32); memcpy(t[1].v, x + 4, 32); memcpy(t[2].v, y, 32); memcpy(t[3].v, y + 4, 32);
    if (fe_is_zero(&t[0]) && fe_is_zero(&t[1]) && fe_is_zero(&t[2]) && fe_is_zero(&t[3])) {
        a.x = fe2_zero(); a.y = fe2_zero(); a.inf = 1; return a;
    }
    f_from_mont(&a.x.a0, &t[0], &P_MOD); f_from_mont(&a.x.a1, &t[1], &P_MOD);
    f_from_mont(&a.y.a0, &t[2], &P_MOD); f_from_mont(&a.y.a1, &t[3], &P_MOD);
    a.inf = 0; return a;
}
static void g2_to_mont(const g2_aff* a, uint64_t* x, uint64_t* y) {
    if (a->inf) { memset(x, 0, 64); memset(y, 0, 64); return; }
    fe t;
    f_to_mont(&t, &a->x.a0, &P_MOD); memcpy(x, t.v, 32);
    f_to_mont(&t, &a->x.a1, &P_MOD); memcpy(x + 4, t.v, 32);
    f_to_mont(&t, &a->y.a0, &P_MOD); memcpy(y, t.v, 32);
    f_to_mont(&t, &a->y.a1, &P_MOD); memcpy(y + 4, t.v, 32);
}

/* ---------------- exported field / curve ops ---------------- */
int or_field_op(int field, int op, const uint64_t* a, const uint64_t* b, uint64_t* r, size_t count) {
    or_init();
    const fe* m = field == OR_FP ? &P_MOD : &R_MOD;
    const fe* rinv = field == OR_FP ? &RINV_P : &RINV_R;
    for (size_t i = 0; i < count; i++) {
        fe x, y, z;
        memcpy(x.v, a + 4 * i, 32);
        if (b) memcpy(y.v, b + 4 * i, 32); else y = fe_zero();
        switch (op) {
            case OR_ADD: f_add(&z, &x, &y, m); break;
            case OR_SUB: f_sub(&z, &x, &y, m); break;
            case OR_NEG: f_neg(&z, &x, m); break;
            case OR_MONT_MUL: f_mul(&z, &x, &y, m); f_mul(&z, &z, rinv, m); break;
            case OR_MONT_SQR: f_mul(&z, &x, &x, m); f_mul(&z, &z, rinv, m); break;
            case OR_INV: f_inv(&z, &x, m); break;
            case OR_TO_MONT: f_to_mont(&z, &x, m); break;
            case OR_FROM_MONT: f_from_mont(&z, &x, m); break;
            case OR_MUL: f_mul(&z, &x, &y, m); break;
            case OR_MONT_POW: {
                fe xc, e = fe_u64(b[4 * i]);
                f_from_mont(&xc, &x, m);
                f_pow(&z, &xc, &e, m);
                f_to_mont(&z, &z, m);
                break;
            }
            default: return -1;
        }
        memcpy(r + 4 * i, z.v, 32);
    }
    return 0;
}

int or_g1_add(const uint64_t p[8], int p_inf, const uint64_t q[8], int q_inf, uint64_t r[8], int* r_inf) {
    or_init();
    g1_aff a = g1_unpack(p, p_inf), b = g1_unpack(q, q_inf);
    g1_jac ja = g1_from_aff(&a), jb = g1_from_aff(&b);
    g1_jac s = g1_add(&ja, &jb);
    g1_aff o = g1_to_aff(&s);
    g1_pack(&o, r); *r_inf = o.inf;
    return 0;
}
int or_g1_mul(const uint64_t p[8], int p_inf, const uint64_t k[4], uint64_t r[8], int* r_inf) {
    or_init();
    g1_aff a = g1_unpack(p, p_inf);
    g1_jac ja = g1_from_aff(&a);
    fe kk; memcpy(kk.v, k, 32);
    g1_jac s = g1_mul_jac(&ja, &kk);
    g1_aff o = g1_to_aff(&s);
    g1_pack(&o, r); *r_inf = o.inf;
    return 0;
}
int or_g1_on_curve(const uint64_t p[8]) { or_init(); g1_aff a = g1_unpack(p, 0); return g1_on_curve_(&a); }
int or_g2_mul(const uint64_t p[16], int p_inf, const uint64_t k[4], uint64_t r[16], int* r_inf) {
    or_init();
    g2_aff a = g2_unpack(p, p_inf);
    g2_jac ja = g2_from_aff(&a);
    fe kk; memcpy(kk.v, k, 32);
    g2_jac s = g2_mul_jac(&ja, &kk);
    g2_aff o = g2_to_aff(&s);
    g2_pack(&o, r); *r_inf = o.inf;
    return 0;
}
int or_g2_on_curve(const uint64_t p[16]) { or_init(); g2_aff a = g2_unpack(p, 0); return g2_on_curve_(&a); }
void or_g1_generator(uint64_t g[8]) { g1_aff a = G1_GEN(); g1_pack(&a, g); }
void or_g2_generator(uint64_t g[16]) { g2_aff a = G2_GEN(); g2_pack(&a, g); }

/* ---------------- MSM (include/msm_g1.cuh:714-913, include/msm_g2.cuh:669-854) ---------------- */
int or_msm_g1(const uint64_t* px, const uint64_t* py, const uint64_t* scalars, size_t n,
              uint64_t out[8], int* out_inf) {
    or_init();
    g1_aff* pts = (g1_aff*)malloc(sizeof(g1_aff) * (n ? n : 1));
    fe* sc = (fe*)malloc(sizeof(fe) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) {
        pts[i] = g1_from_mont(px + 4 * i, py + 4 * i);
        memcpy(sc[i].v, scalars + 4 * i, 32);
    }
    g1_jac s = g1_msm(pts, sc, n);
    g1_aff o = g1_to_aff(&s);
    g1_to_mont(&o, out, out + 4);
    *out_inf = o.inf;
    free(pts); free(sc);
    return 0;
}
int or_msm_g1_naive(const uint64_t* px, const uint64_t* py, const uint64_t* scalars, size_t n,
                    uint64_t out[8], int* out_inf) {
    or_init();
    g1_jac acc = g1_jac_inf();
    for (size_t i = 0; i < n; i++) {
        g1_aff a = g1_from_mont(px + 4 * i, py + 4 * i);
        g1_jac ja = g1_from_aff(&a);
        fe k; memcpy(k.v, scalars + 4 * i, 32);
        g1_jac t = g1_mul_jac(&ja, &k);
        acc = g1_add(&acc, &t);
    }
    g1_aff o = g1_to_aff(&acc);
    g1_to_mont(&o, out, out + 4);
    *out_inf = o.inf;
    return 0;
}
int or_msm_g2(const uint64_t* px, const uint64_t* py, const uint64_t* scalars, size_t n,
              uint64_t out[16], int* out_inf) {
    or_init();
    g2_aff* pts = (g2_aff*)malloc(sizeof(g2_aff) * (n ? n : 1));
    fe* sc = (fe*)malloc(sizeof(fe) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) {
        pts[i] = g2_from_mont(px + 8 * i, py + 8 * i);
        memcpy(sc[i].v, scalars + 4 * i, 32);
    }
    g2_jac s = g2_msm(pts, sc, n);
    g2_aff o = g2_to_aff(&s);
    g2_to_mont(&o, out, out + 8);
    *out_inf = o.inf;
    free(pts); free(sc);
    return 0;
}

/* ---------------- NTT (include/ntt_bn254.cuh:190-464) ----------------
 * w_n = 5^((r-1)/n): equal to the reference's w_28^(2^(28-log n)) with
 * w_28 = 5^((r-1)/2^28) (ntt_bn254.cuh:195-198). */
static void root_of_unity(fe* w, int log_n) {
    if (log_n == 0) { *w = fe_u64(1); return; }
    fe e = R_MOD, five = fe_u64(5);
    e.v[0] -= 1;                                  /* r - 1 (r is odd, no borrow) */
    mpn_rshift((mp_limb_t*)e.v, (const mp_limb_t*)e.v, 4, (unsigned)log_n);
    f_pow(w, &five, &e, &R_MOD);
}
int or_root_of_unity(int log_n, uint64_t w[4]) {
    or_init();
    if (log_n < 0 || log_n > 28) return -1;
    fe x; root_of_unity(&x, log_n); memcpy(w, x.v, 32); return 0;
}

static void ntt_core(fe* a, int log_n, const fe* w) {
    size_t n = (size_t)1 << log_n;
    for (size_t i = 1, j = 0; i < n; i++) {      /* bit reversal */
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { fe t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    enum { TW_CHUNK = 4096 };
    for (int s = 1; s <= log_n; s++) {
        size_t len = (size_t)1 << s, half = len >> 1;
        fe wl = *w;                                 /* w_len = w^(n/len) */
        for (int k = s; k < log_n; k++) RMUL(&wl, &wl, &wl);
        fe* tw = (fe*)malloc(sizeof(fe) * half);
        /* tw[k] = wl^k, in chunks that start from a power */
#pragma omp parallel for schedule(static)
        for (size_t c0 = 0; c0 < half; c0 += TW_CHUNK) {
            fe e = fe_u64(c0);
            f_pow(&tw[c0], &wl, &e, &R_MOD);
            size_t c1 = c0 + TW_CHUNK < half ? c0 + TW_CHUNK : half;
            for (size_t k = c0 + 1; k < c1; k++) RMUL(&tw[k], &tw[k - 1], &wl);
        }
#pragma omp parallel for schedule(static)
        for (size_t t = 0; t < n / 2; t++) {
            size_t k = t & (half - 1), i = (t >> (s - 1)) << s;
            fe u = a[i + k], v;
            RMUL(&v, &a[i + k + half], &tw[k]);
            RADD(&a[i + k], &u, &v);
            RSUB(&a[i + k + half], &u, &v);
        }
        free(tw);
    }
}

static int ntt_canon(fe* a, int log_n, int kind) {
    size_t n = (size_t)1 << log_n;
    fe w; root_of_unity(&w, log_n);
    fe w2n; root_of_unity(&w2n, log_n + 1);
    if (kind == 0) { ntt_core(a, log_n, &w); return 0; }
    if (kind == 1 || kind == 3) {
        fe wi, ninv, nn = fe_u64(n);
        f_inv(&wi, &w, &R_MOD);
        f_inv(&ninv, &nn, &R_MOD);
        ntt_core(a, log_n, &wi);
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n; i++) RMUL(&a[i], &a[i], &ninv);
        if (kind == 3) {
            fe g, gi = fe_u64(1);
            f_inv(&g, &w2n, &R_MOD);
            for (size_t i = 0; i < n; i++) { RMUL(&a[i], &a[i], &gi); RMUL(&gi, &gi, &g); }
        }
        return 0;
    }
    if (kind == 2) {
        fe g = fe_u64(1);
        for (size_t i = 0; i < n; i++) { RMUL(&a[i], &a[i], &g); RMUL(&g, &g, &w2n); }
        ntt_core(a, log_n, &w);
        return 0;
    }
    if (kind == 4) {
        fe* out = (fe*)malloc(sizeof(fe) * n);
        for (size_t k = 0; k < n; k++) {
            fe acc = fe_zero(), wk, e = fe_u64(k), wp = fe_u64(1);
            f_pow(&wk, &w, &e, &R_MOD);
            for (size_t j = 0; j < n; j++) {
                fe t; RMUL(&t, &a[j], &wp); RADD(&acc, &acc, &t); RMUL(&wp, &wp, &wk);
            }
            out[k] = acc;
        }
        memcpy(a, out, sizeof(fe) * n);
        free(out);
        return 0;
    }
    return -1;
}

int or_ntt(uint64_t* data, int log_n, int kind, int mont) {
    or_init();
    if (log_n < 0 || log_n > 27) return -1;
    size_t n = (size_t)1 << log_n;
    fe* a = (fe*)malloc(sizeof(fe) * n);
    for (size_t i = 0; i < n; i++) {
        memcpy(a[i].v, data + 4 * i, 32);
        if (mont) f_from_mont(&a[i], &a[i], &R_MOD);
    }
    int rc = ntt_canon(a, log_n, kind);
    for (size_t i = 0; i < n; i++) {
        if (mont) f_to_mont(&a[i], &a[i], &R_MOD);
        memcpy(data + 4 * i, a[i].v, 32);
    }
    free(a);
    return rc;
}

/* ---------------- SpMV (src/groth16.cu:139-167, 737-861) ----------------
 * raw = c*R^2 mod r; frMul(raw, w) = raw*w*R^-1 = (c*w)*R, i.e. the Montgomery
 * form of c*w, accumulated with plain modular additions. */
int or_spmv(const uint32_t* row_ptr, const uint32_t* cols, const uint64_t* vals, size_t n_rows,
            const uint64_t* w, size_t n_vars, uint64_t* out) {
    or_init();
    for (size_t i = 0; i < n_rows; i++) {
        fe acc = fe_zero();
        for (uint32_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            if (cols[k] >= n_vars) return -1;
            fe c, x, t;
            memcpy(c.v, vals + 4 * (size_t)k, 32);
            memcpy(x.v, w + 4 * (size_t)cols[k], 32);
            RMUL(&t, &c, &x); RMUL(&t, &t, &RINV_R);
            RADD(&acc, &acc, &t);
        }
        memcpy(out + 4 * i, acc.v, 32);
    }
    return 0;
}

/* ---------------- H pipeline (src/groth16.cu:1037-1090) ----------------
 * Cw = Aw o Bw; iNTT_n each; multiply coefficient i by w_2n^i; NTT_n;
 * h_j = A_j*B_j - C_j (no division by Z); canonical output. */
static void hpoly_canon(fe* a, fe* b, int log_n, fe* h) {
    size_t n = (size_t)1 << log_n;
    fe* c = (fe*)malloc(sizeof(fe) * n);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) RMUL(&c[i], &a[i], &b[i]);
    ntt_canon(a, log_n, 1); ntt_canon(b, log_n, 1); ntt_canon(c, log_n, 1);
    ntt_canon(a, log_n, 2); ntt_canon(b, log_n, 2); ntt_canon(c, log_n, 2);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        fe t; RMUL(&t, &a[i], &b[i]); RSUB(&h[i], &t, &c[i]);
    }
    free(c);
}
int or_hpoly(const uint64_t* aw, const uint64_t* bw, int log_n, uint64_t* h) {
    or_init();
    size_t n = (size_t)1 << log_n;
    fe* a = (fe*)malloc(sizeof(fe) * n);
    fe* b = (fe*)malloc(sizeof(fe) * n);
    fe* hh = (fe*)malloc(sizeof(fe) * n);
    for (size_t i = 0; i < n; i++) {
        memcpy(a[i].v, aw + 4 * i, 32); f_from_mont(&a[i], &a[i], &R_MOD);
        memcpy(b[i].v, bw + 4 * i, 32); f_from_mont(&b[i], &b[i], &R_MOD);
    }
    hpoly_canon(a, b, log_n, hh);
    memcpy(h, hh, sizeof(fe) * n);
    free(a); free(b); free(hh);
    return 0;
}

void or_to_dec(const uint64_t a[4], char* buf, size_t buf_size) {
    mpz_t x; mpz_init(x);
    mpz_import(x, 4, -1, 8, 0, 0, a);
    char* s = mpz_get_str(NULL, 10, x);
    snprintf(buf, buf_size, "%s", s);
    free(s);
    mpz_clear(x);
}

/* ====================================================================
 * Synthetic circuit: seeded R1CS + witness + trapdoor setup.
 *
 * PRNG: SplitMix64 -> xoshiro256**; one independent stream per purpose,
 * seeded with seed ^ STREAM_TAG. Uniform Fr = rejection sampling of 254-bit
 * candidates (four outputs, top word masked to 62 bits) below r. Uniform
 * integer below b = high 64 bits of (next() * b).
 * ==================================================================== */
typedef struct { uint64_t s[4]; } xo_t;
static uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static void xo_seed(xo_t* g, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) g->s[i] = splitmix64(&x);
}
static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t xo_next(xo_t* g) {
    uint64_t* s = g->s;
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl64(s[3], 45);
    return result;
}
static fe xo_fr(xo_t* g) {
    for (;;) {
        fe x;
        for (int i = 0; i < 4; i++) x.v[i] = xo_next(g);
        x.v[3] &= 0x3fffffffffffffffULL;
        if (mpn_cmp((const mp_limb_t*)x.v, (const mp_limb_t*)R_MOD.v, 4) < 0) return x;
    }
}
static uint32_t xo_below(xo_t* g, uint32_t b) {
    return (uint32_t)(((unsigned __int128)xo_next(g) * b) >> 64);
}

#define TAG_R1CS 0x5231435352314353ULL
#define TAG_PUB 0x5055424c49435055ULL
#define TAG_TOXIC 0x544f584943574153ULL
#define TAG_BLIND 0x424c494e44524e47ULL

struct or_circuit {
    uint32_t m, l, n_vars, n, log_n;
    uint32_t* a_wire; fe* a_coef;     /* 3*m entries, row-major */
    uint32_t* b_wire; fe* b_coef;
    fe* w;                             /* witness */
    fe tau, alpha, beta, gamma, delta, r, s;
    fe *u, *v, *ww;                    /* QAP polys at tau per wire */
    fe* lag;                           /* L_k(tau), k < n */
    int with_crs;
    g1_aff *A, *B1, *C, *H, *IC;
    g2_aff* B2;
    g1_aff alpha1, beta1, delta1;
    g2_aff beta2, gamma2, delta2;
};

/* fixed-base comb: 32 windows of 8 bits, table[j][d] = d * 2^(8j) * G */
typedef struct { g1_aff t[32][256]; } g1_comb;
typedef struct { g2_aff t[32][256]; } g2_comb;
static g1_comb* g1_comb_make(const g1_aff* g) {
    g1_comb* c = (g1_comb*)malloc(sizeof(g1_comb));
    g1_jac base = g1_from_aff(g);
    for (int j = 0; j < 32; j++) {
        g1_jac acc = g1_jac_inf();
        for (int d = 0; d < 256; d++) { c->t[j][d] = g1_to_aff(&acc); acc = g1_add(&acc, &base); }
        for (int k = 0; k < 8; k++) base = g1_dbl(&base);
    }
    return c;
}
static g2_comb* g2_comb_make(const g2_aff* g) {
    g2_comb* c = (g2_comb*)malloc(sizeof(g2_comb));
    g2_jac base = g2_from_aff(g);
    for (int j = 0; j < 32; j++) {
        g2_jac acc = g2_jac_inf();
        for (int d = 0; d < 256; d++) { c->t[j][d] = g2_to_aff(&acc); acc = g2_add(&acc, &base); }
        for (int k = 0; k < 8; k++) base = g2_dbl(&base);
    }
    return c;
}
static g1_aff g1_comb_mul(const g1_comb* c, const fe* k) {
    g1_jac acc = g1_jac_inf();
    for (int j = 0; j < 32; j++) {
        unsigned d = (unsigned)((k->v[j >> 3] >> (8 * (j & 7))) & 0xff);
        if (d) { g1_jac t = g1_from_aff(&c->t[j][d]); acc = g1_add(&acc, &t); }
    }
    return g1_to_aff(&acc);
}
static g2_aff g2_comb_mul(const g2_comb* c, const fe* k) {
    g2_jac acc = g2_jac_inf();
    for (int j = 0; j < 32; j++) {
        unsigned d = (unsigned)((k->v[j >> 3] >> (8 * (j & 7))) & 0xff);
        if (d) { g2_jac t = g2_from_aff(&c->t[j][d]); acc = g2_add(&acc, &t); }
    }
    return g2_to_aff(&acc);
}

/* L_k(tau) over the size-n domain, k < n: (tau^n - 1)/n * w^k / (tau - w^k) */
static void lagrange_all(fe* out, const fe* tau, int log_n) {
    size_t n = (size_t)1 << log_n;
    fe w; root_of_unity(&w, log_n);
    fe tn = *tau;
    for (int i = 0; i < log_n; i++) RMUL(&tn, &tn, &tn);
    fe one = fe_u64(1), nn = fe_u64(n), ninv, zf;
    f_inv(&ninv, &nn, &R_MOD);
    RSUB(&zf, &tn, &one); RMUL(&zf, &zf, &ninv);
    /* batch inversion of (tau - w^k) */
    fe* den = (fe*)malloc(sizeof(fe) * n);
    fe* wk = (fe*)malloc(sizeof(fe) * n);
    fe cur = one;
    for (size_t k = 0; k < n; k++) { wk[k] = cur; RSUB(&den[k], tau, &cur); RMUL(&cur, &cur, &w); }
    fe* pre = (fe*)malloc(sizeof(fe) * n);
    fe acc = one;
    for (size_t k = 0; k < n; k++) { pre[k] = acc; RMUL(&acc, &acc, &den[k]); }
    fe inv; f_inv(&inv, &acc, &R_MOD);
    for (size_t k = n; k-- > 0;) {
        fe dk; RMUL(&dk, &inv, &pre[k]); RMUL(&inv, &inv, &den[k]);
        fe t; RMUL(&t, &zf, &wk[k]); RMUL(&out[k], &t, &dk);
    }
    free(den); free(wk); free(pre);
}

or_circuit* or_circuit_create(uint64_t seed, uint32_t m, uint32_t n_pub, int with_crs) {
    or_init();
    or_circuit* c = (or_circuit*)calloc(1, sizeof(or_circuit));
    c->m = m; c->l = n_pub; c->n_vars = 1 + n_pub + m;
    uint32_t need = m + n_pub + 1, n = 1, ln = 0;
    while (n < need) { n <<= 1; ln++; }
    c->n = n; c->log_n = ln;
    c->a_wire = (uint32_t*)malloc(4ull * 3 * m); c->a_coef = (fe*)malloc(sizeof(fe) * 3ull * m);
    c->b_wire = (uint32_t*)malloc(4ull * 3 * m); c->b_coef = (fe*)malloc(sizeof(fe) * 3ull * m);
    c->w = (fe*)malloc(sizeof(fe) * c->n_vars);

    /* R1CS rows: A_i, B_i each 3 terms on wires < 1 + l + i with uniform coefficients */
    xo_t gr; xo_seed(&gr, seed ^ TAG_R1CS);
    for (uint32_t i = 0; i < m; i++) {
        uint32_t bound = 1 + n_pub + i;
        for (int t = 0; t < 3; t++) { c->a_wire[3 * i + t] = xo_below(&gr, bound); c->a_coef[3 * i + t] = xo_fr(&gr); }
        for (int t = 0; t < 3; t++) { c->b_wire[3 * i + t] = xo_below(&gr, bound); c->b_coef[3 * i + t] = xo_fr(&gr); }
    }
    /* witness: w0 = 1, public inputs uniform, aux_i = (A_i.w)(B_i.w) */
    xo_t gp; xo_seed(&gp, seed ^ TAG_PUB);
    c->w[0] = fe_u64(1);
    for (uint32_t k = 1; k <= n_pub; k++) c->w[k] = xo_fr(&gp);
    for (uint32_t i = 0; i < m; i++) {
        fe sa = fe_zero(), sb = fe_zero(), t;
        for (int q = 0; q < 3; q++) {
            RMUL(&t, &c->a_coef[3 * i + q], &c->w[c->a_wire[3 * i + q]]); RADD(&sa, &sa, &t);
            RMUL(&t, &c->b_coef[3 * i + q], &c->w[c->b_wire[3 * i + q]]); RADD(&sb, &sb, &t);
        }
        RMUL(&c->w[1 + n_pub + i], &sa, &sb);
    }
    /* toxic waste and blinding */
    xo_t gt; xo_seed(&gt, seed ^ TAG_TOXIC);
    c->tau = xo_fr(&gt); c->alpha = xo_fr(&gt); c->beta = xo_fr(&gt); c->gamma = xo_fr(&gt); c->delta = xo_fr(&gt);
    xo_t gb; xo_seed(&gb, seed ^ TAG_BLIND);
    c->r = xo_fr(&gb); c->s = xo_fr(&gb);

    /* QAP polynomials at tau */
    c->lag = (fe*)malloc(sizeof(fe) * n);
    lagrange_all(c->lag, &c->tau, (int)ln);
    c->u = (fe*)calloc(c->n_vars, sizeof(fe));
    c->v = (fe*)calloc(c->n_vars, sizeof(fe));
    c->ww = (fe*)calloc(c->n_vars, sizeof(fe));
    for (uint32_t i = 0; i < m; i++)
        for (int q = 0; q < 3; q++) {
            fe t;
            RMUL(&t, &c->a_coef[3 * i + q], &c->lag[i]); RADD(&c->u[c->a_wire[3 * i + q]], &c->u[c->a_wire[3 * i + q]], &t);
            RMUL(&t, &c->b_coef[3 * i + q], &c->lag[i]); RADD(&c->v[c->b_wire[3 * i + q]], &c->v[c->b_wire[3 * i + q]], &t);
        }
    for (uint32_t k = 0; k <= n_pub; k++) RADD(&c->u[k], &c->u[k], &c->lag[m + k]);   /* public rows A[m+k][k] = 1 */
    for (uint32_t i = 0; i < m; i++) c->ww[1 + n_pub + i] = c->lag[i];                 /* C_i = aux_i */

    c->with_crs = with_crs;
    if (with_crs) {
        g1_aff G = G1_GEN(); g2_aff G2 = G2_GEN();
        g1_comb* t1 = g1_comb_make(&G);
        g2_comb* t2 = g2_comb_make(&G2);
        c->alpha1 = g1_comb_mul(t1, &c->alpha);
        c->beta1 = g1_comb_mul(t1, &c->beta);
        c->delta1 = g1_comb_mul(t1, &c->delta);
        c->beta2 = g2_comb_mul(t2, &c->beta);
        c->gamma2 = g2_comb_mul(t2, &c->gamma);
        c->delta2 = g2_comb_mul(t2, &c->delta);
        fe gi, di; f_inv(&gi, &c->gamma, &R_MOD); f_inv(&di, &c->delta, &R_MOD);
        c->A = (g1_aff*)malloc(sizeof(g1_aff) * c->n_vars);
        c->B1 = (g1_aff*)malloc(sizeof(g1_aff) * c->n_vars);
        c->B2 = (g2_aff*)malloc(sizeof(g2_aff) * c->n_vars);
        c->IC = (g1_aff*)malloc(sizeof(g1_aff) * (n_pub + 1));
        c->C = (g1_aff*)malloc(sizeof(g1_aff) * (c->n_vars - n_pub - 1 + 1));
        c->H = (g1_aff*)malloc(sizeof(g1_aff) * n);
#pragma omp parallel for schedule(dynamic, 256)
        for (uint32_t i = 0; i < c->n_vars; i++) {
            c->A[i] = g1_comb_mul(t1, &c->u[i]);
            c->B1[i] = g1_comb_mul(t1, &c->v[i]);
            c->B2[i] = g2_comb_mul(t2, &c->v[i]);
            fe e, t;
            RMUL(&e, &c->beta, &c->u[i]); RMUL(&t, &c->alpha, &c->v[i]); RADD(&e, &e, &t); RADD(&e, &e, &c->ww[i]);
            if (i <= n_pub) { RMUL(&t, &e, &gi); c->IC[i] = g1_comb_mul(t1, &t); }
            else { RMUL(&t, &e, &di); c->C[i - n_pub - 1] = g1_comb_mul(t1, &t); }
        }
        /* H_j = L^(2n)_{2j+1}(tau) / delta */
        fe* lag2 = (fe*)malloc(sizeof(fe) * 2 * n);
        lagrange_all(lag2, &c->tau, (int)ln + 1);
#pragma omp parallel for schedule(dynamic, 256)
        for (uint32_t j = 0; j < n; j++) { fe t; RMUL(&t, &lag2[2 * j + 1], &di); c->H[j] = g1_comb_mul(t1, &t); }
        free(lag2);
        free(t1); free(t2);
    }
    return c;
}

void or_circuit_free(or_circuit* c) {
    if (!c) return;
    free(c->a_wire); free(c->a_coef); free(c->b_wire); free(c->b_coef); free(c->w);
    free(c->u); free(c->v); free(c->ww); free(c->lag);
    free(c->A); free(c->B1); free(c->B2); free(c->C); free(c->H); free(c->IC);
    free(c);
}
void or_circuit_sizes(const or_circuit* c, uint64_t out[7]) {
    out[0] = c->m; out[1] = c->l; out[2] = c->n_vars; out[3] = c->n; out[4] = c->log_n;
    out[5] = 3ull * c->m + c->l + 1; out[6] = 3ull * c->m;
}
void or_circuit_witness(const or_circuit* c, uint64_t* w) { memcpy(w, c->w, sizeof(fe) * c->n_vars); }
void or_circuit_blinding(const or_circuit* c, uint64_t r[4], uint64_t s[4]) { memcpy(r, c->r.v, 32); memcpy(s, c->s.v, 32); }
void or_circuit_toxic(const or_circuit* c, uint64_t out[20]) {
    memcpy(out, c->tau.v, 32); memcpy(out + 4, c->alpha.v, 32); memcpy(out + 8, c->beta.v, 32);
    memcpy(out + 12, c->gamma.v, 32); memcpy(out + 16, c->delta.v, 32);
}

/* ---------------- zkey / wtns writers (layout read by src/zkey_utils.cpp:19-53,
 * src/binfile_utils.cpp:32-80, src/groth16.cu:737-805, src/main.cu:28-110) ---------------- */
static void w32(FILE* f, uint32_t x) { fwrite(&x, 4, 1, f); }
static void w64(FILE* f, uint64_t x) { fwrite(&x, 8, 1, f); }
static void wg1(FILE* f, const g1_aff* p) { uint64_t b[8]; g1_to_mont(p, b, b + 4); fwrite(b, 8, 8, f); }
static void wg2(FILE* f, const g2_aff* p) { uint64_t b[16]; g2_to_mont(p, b, b + 8); fwrite(b, 8, 16, f); }
static void wcoef(FILE* f, uint32_t mat, uint32_t row, uint32_t wire, const fe* c) {
    fe raw; RMUL(&raw, c, &R2_R);   /* c * R^2 mod r */
    w32(f, mat); w32(f, row); w32(f, wire); fwrite(raw.v, 8, 4, f);
}

int or_circuit_write_zkey(const or_circuit* c, const char* path) {
    if (!c->with_crs) return -1;
    FILE* f = fopen(path, "wb");
    if (!f) return -2;
    fwrite("zkey", 1, 4, f); w32(f, 1); w32(f, 9);
    /* §1 */ w32(f, 1); w64(f, 4); w32(f, 1);
    /* §2 */
    w32(f, 2); w64(f, 4 + 32 + 4 + 32 + 12 + 64 * 3 + 128 * 3);
    w32(f, 32); fwrite(P_MOD.v, 8, 4, f); w32(f, 32); fwrite(R_MOD.v, 8, 4, f);
    w32(f, c->n_vars); w32(f, c->l); w32(f, c->n);
    wg1(f, &c->alpha1); wg1(f, &c->beta1); wg2(f, &c->beta2); wg2(f, &c->gamma2); wg1(f, &c->delta1); wg2(f, &c->delta2);
    /* §3 IC */
    w32(f, 3); w64(f, 64ull * (c->l + 1));
    for (uint32_t i = 0; i <= c->l; i++) wg1(f, &c->IC[i]);
    /* §4 coefficients: A rows, public rows, B rows */
    uint64_t ncoef = 3ull * c->m + c->l + 1 + 3ull * c->m;
    w32(f, 4); w64(f, 4 + 44 * ncoef); w32(f, (uint32_t)ncoef);
    for (uint32_t i = 0; i < c->m; i++)
        for (int q = 0; q < 3; q++) wcoef(f, 0, i, c->a_wire[3 * i + q], &c->a_coef[3 * i + q]);
    fe one = fe_u64(1);
    for (uint32_t k = 0; k <= c->l; k++) wcoef(f, 0, c->m + k, k, &one);
    for (uint32_t i = 0; i < c->m; i++)
        for (int q = 0; q < 3; q++) wcoef(f, 1, i, c->b_wire[3 * i + q], &c->b_coef[3 * i + q]);
    /* §5 A, §6 B1, §7 B2, §8 C, §9 H */
    w32(f, 5); w64(f, 64ull * c->n_vars); for (uint32_t i = 0; i < c->n_vars; i++) wg1(f, &c->A[i]);
    w32(f, 6); w64(f, 64ull * c->n_vars); for (uint32_t i = 0; i < c->n_vars; i++) wg1(f, &c->B1[i]);
    w32(f, 7); w64(f, 128ull * c->n_vars); for (uint32_t i = 0; i < c->n_vars; i++) wg2(f, &c->B2[i]);
    uint32_t nc = c->n_vars - c->l - 1;
    w32(f, 8); w64(f, 64ull * nc); for (uint32_t i = 0; i < nc; i++) wg1(f, &c->C[i]);
    w32(f, 9); w64(f, 64ull * c->n); for (uint32_t j = 0; j < c->n; j++) wg1(f, &c->H[j]);
    fclose(f);
    return 0;
}

int or_circuit_write_wtns(const or_circuit* c, const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) return -2;
    fwrite("wtns", 1, 4, f); w32(f, 2); w32(f, 2);
    w32(f, 1); w64(f, 40); w32(f, 32); fwrite(R_MOD.v, 8, 4, f); w32(f, c->n_vars);
    w32(f, 2); w64(f, 32ull * c->n_vars);
    fwrite(c->w, 32, c->n_vars, f);
    fclose(f);
    return 0;
}

/* ---------------- prove (src/groth16.cu:1143-1397 + snarkjs blinding) ---------------- */
static void proof_pack(const g1_aff* A, const g2_aff* B, const g1_aff* C, uint64_t proof[32]) {
    g1_to_mont(A, proof, proof + 4);
    g2_to_mont(B, proof + 8, proof + 16);
    g1_to_mont(C, proof + 24, proof + 28);
}

int or_prove(const or_circuit* c, const uint64_t rr[4], const uint64_t ss[4], uint64_t proof[32]) {
    or_init();
    if (!c->with_crs) return -1;
    fe r, s; memcpy(r.v, rr, 32); memcpy(s.v, ss, 32);
    size_t n = c->n;
    /* SpMV through the zkey convention: raw = c*R^2, frMul(raw, w) */
    fe* aw = (fe*)calloc(n, sizeof(fe));
    fe* bw = (fe*)calloc(n, sizeof(fe));
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < c->m; i++)
        for (int q = 0; q < 3; q++) {
            fe raw, t;
            RMUL(&raw, &c->a_coef[3 * i + q], &R2_R);
            RMUL(&t, &raw, &c->w[c->a_wire[3 * i + q]]); RMUL(&t, &t, &RINV_R); RADD(&aw[i], &aw[i], &t);
            RMUL(&raw, &c->b_coef[3 * i + q], &R2_R);
            RMUL(&t, &raw, &c->w[c->b_wire[3 * i + q]]); RMUL(&t, &t, &RINV_R); RADD(&bw[i], &bw[i], &t);
        }
    for (uint32_t k = 0; k <= c->l; k++) {
        fe raw, t, one = fe_u64(1);
        RMUL(&raw, &one, &R2_R); RMUL(&t, &raw, &c->w[k]); RMUL(&t, &t, &RINV_R);
        RADD(&aw[c->m + k], &aw[c->m + k], &t);
    }
    /* Montgomery -> canonical, H pipeline */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) { f_from_mont(&aw[i], &aw[i], &R_MOD); f_from_mont(&bw[i], &bw[i], &R_MOD); }
    fe* h = (fe*)malloc(sizeof(fe) * n);
    hpoly_canon(aw, bw, (int)c->log_n, h);
    free(aw); free(bw);
    /* MSMs */
    g1_jac a_part = g1_msm(c->A, c->w, c->n_vars);
    g1_jac b1_part = g1_msm(c->B1, c->w, c->n_vars);
    g2_jac b2_part = g2_msm(c->B2, c->w, c->n_vars);
    uint32_t nc = c->n_vars - c->l - 1;
    g1_jac c_part = g1_msm(c->C, c->w + c->l + 1, nc);
    g1_jac h_part = g1_msm(c->H, h, n);
    free(h);
    /* pi_A = alpha1 + A + r*delta1 */
    g1_jac t, jd1 = g1_from_aff(&c->delta1), ja1 = g1_from_aff(&c->alpha1), jb1 = g1_from_aff(&c->beta1);
    g1_jac piA = g1_add(&ja1, &a_part);
    t = g1_mul_jac(&jd1, &r); piA = g1_add(&piA, &t);
    /* pi_B = beta2 + B2 + s*delta2 ; B1 = beta1 + B1 + s*delta1 */
    g2_jac jb2 = g2_from_aff(&c->beta2), jd2 = g2_from_aff(&c->delta2);
    g2_jac piB = g2_add(&jb2, &b2_part);
    g2_jac t2 = g2_mul_jac(&jd2, &s); piB = g2_add(&piB, &t2);
    g1_jac B1 = g1_add(&jb1, &b1_part);
    t = g1_mul_jac(&jd1, &s); B1 = g1_add(&B1, &t);
    /* pi_C = C + H + s*pi_A + r*B1 - r*s*delta1 */
    g1_jac piC = g1_add(&c_part, &h_part);
    t = g1_mul_jac(&piA, &s); piC = g1_add(&piC, &t);
    t = g1_mul_jac(&B1, &r); piC = g1_add(&piC, &t);
    fe rs, nrs; RMUL(&rs, &r, &s); f_neg(&nrs, &rs, &R_MOD);
    t = g1_mul_jac(&jd1, &nrs); piC = g1_add(&piC, &t);
    g1_aff oA = g1_to_aff(&piA), oC = g1_to_aff(&piC);
    g2_aff oB = g2_to_aff(&piB);
    proof_pack(&oA, &oB, &oC, proof);
    return 0;
}

/* exponents a, b, c of the proof (see or_prove for the group formulas):
 *   a = alpha + U + r*delta, b = beta + V + s*delta,
 *   c = (sum_{i>l} w_i (beta u_i + alpha v_i + ww_i) + U*V - Cp)/delta + s*a + r*b - r*s*delta
 * where U = sum w_i u_i(tau), V = sum w_i v_i(tau) and Cp = sum_k (Aw_k Bw_k) L_k(tau)
 * is the value at tau of the interpolant of Aw o Bw (the reference computes
 * Cw as that pointwise product, src/groth16.cu:1059-1060). */
static void trapdoor_exps(const or_circuit* c, const fe* w, const fe* r, const fe* s, fe* a, fe* b, fe* cc) {
    fe U = fe_zero(), V = fe_zero(), K = fe_zero(), t, e;
    for (uint32_t i = 0; i < c->n_vars; i++) {
        RMUL(&t, &w[i], &c->u[i]); RADD(&U, &U, &t);
        RMUL(&t, &w[i], &c->v[i]); RADD(&V, &V, &t);
        if (i > c->l) {
            RMUL(&e, &c->beta, &c->u[i]); RMUL(&t, &c->alpha, &c->v[i]); RADD(&e, &e, &t); RADD(&e, &e, &c->ww[i]);
            RMUL(&t, &w[i], &e); RADD(&K, &K, &t);
        }
    }
    /* Aw, Bw per row and Cp */
    fe Cp = fe_zero();
    for (uint32_t i = 0; i < c->m; i++) {
        fe sa = fe_zero(), sb = fe_zero();
        for (int q = 0; q < 3; q++) {
            RMUL(&t, &c->a_coef[3 * i + q], &w[c->a_wire[3 * i + q]]); RADD(&sa, &sa, &t);
            RMUL(&t, &c->b_coef[3 * i + q], &w[c->b_wire[3 * i + q]]); RADD(&sb, &sb, &t);
        }
        RMUL(&t, &sa, &sb); RMUL(&t, &t, &c->lag[i]); RADD(&Cp, &Cp, &t);
    }
    /* public rows have B = 0: no contribution to Cp */
    RMUL(&t, r, &c->delta); RADD(a, &c->alpha, &U); RADD(a, a, &t);
    RMUL(&t, s, &c->delta); RADD(b, &c->beta, &V); RADD(b, b, &t);
    fe di; f_inv(&di, &c->delta, &R_MOD);
    fe num; RMUL(&num, &U, &V); RSUB(&num, &num, &Cp); RADD(&num, &num, &K);
    RMUL(cc, &num, &di);
    RMUL(&t, s, a); RADD(cc, cc, &t);
    RMUL(&t, r, b); RADD(cc, cc, &t);
    RMUL(&t, r, s); RMUL(&t, &t, &c->delta); RSUB(cc, cc, &t);
}

int or_prove_trapdoor_w(const or_circuit* c, const uint64_t* wv, const uint64_t rr[4], const uint64_t ss[4],
                        uint64_t proof[32]) {
    or_init();
    fe r, s, a, b, cc;
    memcpy(r.v, rr, 32); memcpy(s.v, ss, 32);
    trapdoor_exps(c, (const fe*)wv, &r, &s, &a, &b, &cc);
    g1_aff G = G1_GEN(); g2_aff G2 = G2_GEN();
    g1_jac jg = g1_from_aff(&G); g2_jac jg2 = g2_from_aff(&G2);
    g1_jac pa = g1_mul_jac(&jg, &a), pc = g1_mul_jac(&jg, &cc);
    g2_jac pb = g2_mul_jac(&jg2, &b);
    g1_aff oA = g1_to_aff(&pa), oC = g1_to_aff(&pc);
    g2_aff oB = g2_to_aff(&pb);
    proof_pack(&oA, &oB, &oC, proof);
    return 0;
}
int or_prove_trapdoor(const or_circuit* c, const uint64_t r[4], const uint64_t s[4], uint64_t proof[32]) {
    return or_prove_trapdoor_w(c, (const uint64_t*)c->w, r, s, proof);
}

/* Groth16 verification with the trapdoor: the proof points must equal a*G1,
 * b*G2, c*G1 for the exponents recomputed from (w, r, s), and those exponents
 * must satisfy the verification equation in the exponent
 *   a*b = alpha*beta + sum_{i<=l} w_i (beta u_i + alpha v_i + ww_i)/gamma * gamma + c*delta,
 * which is e(A,B) = e(alpha1,beta2) e(vk_x,gamma2) e(C,delta2) with all points
 * written as multiples of the generators. Returns 1 when both hold. */
int or_verify_trapdoor(const or_circuit* c, const uint64_t* wv, const uint64_t rr[4], const uint64_t ss[4],
                       const uint64_t proof[32]) {
    or_init();
    fe r, s, a, b, cc, lhs, rhs, t, e;
    memcpy(r.v, rr, 32); memcpy(s.v, ss, 32);
    const fe* w = (const fe*)wv;
    trapdoor_exps(c, w, &r, &s, &a, &b, &cc);
    RMUL(&lhs, &a, &b);
    RMUL(&rhs, &c->alpha, &c->beta);
    for (uint32_t i = 0; i <= c->l; i++) {
        RMUL(&e, &c->beta, &c->u[i]); RMUL(&t, &c->alpha, &c->v[i]); RADD(&e, &e, &t); RADD(&e, &e, &c->ww[i]);
        RMUL(&t, &w[i], &e); RADD(&rhs, &rhs, &t);
    }
    RMUL(&t, &cc, &c->delta); RADD(&rhs, &rhs, &t);
    if (!fe_eq(&lhs, &rhs)) return 0;
    uint64_t ref[32];
    or_prove_trapdoor_w(c, wv, rr, ss, ref);
    return memcmp(ref, proof, sizeof ref) == 0;
}

/* ====================================================================
 * Helpers for checks at full size (2^20 .. 2^22 elements).
 * ==================================================================== */

/* Points with known discrete logs: out[i] = (k0 + i*d)*G for i < n, Montgomery
 * affine AoS (x[4], y[4]; infinity all zero). Point i = a*B + b is A_a + D_b with
 * A_a = (k0 + a*B*d)*G and D_b = (b*d)*G; every block of B sums shares one
 * inversion (Montgomery's simultaneous-inversion trick), blocks run in parallel.
 * Pinned by: the exponents are known, so any sample point equals a textbook
 * double-and-add product (tests/test_oracle.py). */
#define PROG_B 4096
int or_g1_progression(const uint64_t k0v[4], const uint64_t dv[4], size_t n, uint64_t* out) {
    or_init();
    if (!n) return 0;
    fe k0, d;
    memcpy(k0.v, k0v, 32); memcpy(d.v, dv, 32);
    f_reduce(&k0, &k0, &R_MOD); f_reduce(&d, &d, &R_MOD);
    g1_aff G = G1_GEN();
    g1_jac jg = g1_from_aff(&G);
    size_t nblk = (n + PROG_B - 1) / PROG_B;
    g1_aff* D = (g1_aff*)malloc(sizeof(g1_aff) * PROG_B);
    g1_aff* A = (g1_aff*)malloc(sizeof(g1_aff) * nblk);
    g1_jac jd = g1_mul_jac(&jg, &d), acc = g1_jac_inf();
    for (size_t b = 0; b < PROG_B; b++) { D[b] = g1_to_aff(&acc); acc = g1_add(&acc, &jd); }
    fe bd, bb = fe_u64(PROG_B);
    RMUL(&bd, &bb, &d);
    g1_jac step = g1_mul_jac(&jg, &bd);
    acc = g1_mul_jac(&jg, &k0);
    for (size_t a = 0; a < nblk; a++) { A[a] = g1_to_aff(&acc); acc = g1_add(&acc, &step); }
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t a = 0; a < nblk; a++) {
        fe* den = (fe*)malloc(sizeof(fe) * PROG_B);
        fe* pre = (fe*)malloc(sizeof(fe) * PROG_B);
        unsigned char gen[PROG_B];   /* 1 = generic affine sum, 0 = handled by the Jacobian path */
        size_t i0 = a * PROG_B, cnt = n - i0 < PROG_B ? n - i0 : PROG_B;
        fe prod = fe_u64(1);
        for (size_t b = 0; b < cnt; b++) {
            gen[b] = !A[a].inf && !D[b].inf && !fe_eq(&A[a].x, &D[b].x);
            if (gen[b]) PSUB(&den[b], &D[b].x, &A[a].x); else den[b] = fe_u64(1);
            pre[b] = prod;
            PMUL(&prod, &prod, &den[b]);
        }
        fe inv;
        f_inv(&inv, &prod, &P_MOD);
        for (size_t b = cnt; b-- > 0;) {
            fe di;
            PMUL(&di, &inv, &pre[b]);
            PMUL(&inv, &inv, &den[b]);
            g1_aff o;
            if (gen[b]) {
                fe lam, t, x3, y3;
                PSUB(&t, &D[b].y, &A[a].y); PMUL(&lam, &t, &di);
                PMUL(&x3, &lam, &lam); PSUB(&x3, &x3, &A[a].x); PSUB(&x3, &x3, &D[b].x);
                PSUB(&t, &A[a].x, &x3); PMUL(&y3, &lam, &t); PSUB(&y3, &y3, &A[a].y);
                o.x = x3; o.y = y3; o.inf = 0;
            } else {
                g1_jac p = g1_from_aff(&A[a]), q = g1_from_aff(&D[b]);
                g1_jac s = g1_add(&p, &q);
                o = g1_to_aff(&s);
            }
            uint64_t* dst = out + 8 * (i0 + b);
            g1_to_mont(&o, dst, dst + 4);
        }
        free(den); free(pre);
    }
    free(D); free(A);
    return 0;
}

/* Horner evaluation of sum_j c_j z^j mod r at nz points z (canonical) in one
 * pass per point, points in parallel. The coefficients' representation carries
 * through (Montgomery coefficients give the Montgomery value), so NTT output
 * rows are checked by definition: X_k = sum_j x_j w^(jk). */
int or_fr_poly_eval(const uint64_t* coeffs, size_t n, const uint64_t* zs, size_t nz, uint64_t* out) {
    or_init();
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t q = 0; q < nz; q++) {
        fe z, acc = fe_zero(), c;
        memcpy(z.v, zs + 4 * q, 32);
        for (size_t j = n; j-- > 0;) {
            memcpy(c.v, coeffs + 4 * j, 32);
            RMUL(&acc, &acc, &z);
            RADD(&acc, &acc, &c);
        }
        memcpy(out + 4 * q, acc.v, 32);
    }
    return 0;
}

/* sum_i a_i b_i mod r (canonical in and out), in parallel chunks */
int or_fr_dot(const uint64_t* a, const uint64_t* b, size_t n, uint64_t out[4]) {
    or_init();
    enum { CH = 1 << 14 };
    size_t nch = (n + CH - 1) / CH;
    fe* part = (fe*)calloc(nch ? nch : 1, sizeof(fe));
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < nch; k++) {
        fe acc = fe_zero(), x, y, t;
        size_t e = (k + 1) * CH < n ? (k + 1) * CH : n;
        for (size_t i = k * CH; i < e; i++) {
            memcpy(x.v, a + 4 * i, 32); memcpy(y.v, b + 4 * i, 32);
            RMUL(&t, &x, &y); RADD(&acc, &acc, &t);
        }
        part[k] = acc;
    }
    fe tot = fe_zero();
    for (size_t k = 0; k < nch; k++) RADD(&tot, &tot, &part[k]);
    memcpy(out, tot.v, 32);
    free(part);
    return 0;
}

/* worker threads of the parallel loops (OpenMP); 0 leaves the current setting */
#include <omp.h>
int or_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}
