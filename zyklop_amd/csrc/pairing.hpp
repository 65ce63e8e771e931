// Host-side optimal ate pairing on BN254 and the Groth16 verifier built on it.
//
// The reference verifies with snarkjs (`snarkjs groth16 verify`, Makefile:135);
// it has no pairing of its own. This is a textbook construction (the optimal
// ate pairing for BN curves of Beuchat, Gonzalez-Diaz, Mitsunari, Okamoto,
// Rodriguez-Henriquez and Teruya, "High-Speed Software Implementation of the
// Optimal Ate Pairing over Barreto-Naehrig Curves", Pairing 2010), written for
// clarity rather than speed: verification is one multi-pairing per proof.
//   * tower: Fp2 = Fp[u]/(u^2 + 1), Fp6 = Fp2[v]/(v^3 - xi), xi = 9 + u,
//     Fp12 = Fp6[w]/(w^2 - v), so w^6 = xi;
//   * G2 lives on the D-type twist y^2 = x^3 + 3/xi; psi(x, y) = (x w^2, y w^3)
//     maps it into E(Fp12);
//   * Miller loop over 6x + 2 (x = 4965661367192848881) with affine points on
//     the twist (one Fp2 inversion per step), then the two Frobenius lines;
//   * final exponentiation f^((p^12 - 1)/r) as f^(p^6 - 1) (conjugate over
//     inverse) raised to (p^6 + 1)/r, with that exponent computed from p and r
//     at start-up (no stored constant).
#pragma once
#include <stdint.h>

#include <vector>

#include "host_field.hpp"

namespace zkh {

struct Fp6 {
    Fp2 c0, c1, c2;
};
struct Fp12 {
    Fp6 c0, c1;
};

Fp12 fp12_one();
bool fp12_eq(const Fp12& a, const Fp12& b);
Fp12 fp12_mul(const Fp12& a, const Fp12& b);

bool g1_on_curve(const Aff<Fp>& p);
bool g2_on_curve(const Aff<Fp2>& q);
// q * r == infinity (G2 has a cofactor; G1 has none)
bool g2_in_subgroup(const Aff<Fp2>& q);

// e(P, Q); 1 when either point is infinity. Points affine Montgomery.
Fp12 pairing(const Aff<Fp>& p, const Aff<Fp2>& q);
// prod_i e(P_i, Q_i) == 1, with one final exponentiation
bool pairing_product_is_one(const Aff<Fp>* p, const Aff<Fp2>* q, int n);

// Groth16 verification: e(A, B) == e(alpha1, beta2) e(vk_x, gamma2) e(C, delta2)
// with vk_x = IC_0 + sum_i pub_i IC_i. Points affine Montgomery (zkey layout:
// G1 x[4] y[4], G2 x.a0 x.a1 y.a0 y.a1, all zero = infinity); publics
// canonical Fr (n_public x 4 words); ic has n_public + 1 points. Returns false
// for a proof point off its curve or B outside the order-r subgroup.
bool groth16_verify(const uint64_t alpha1[8], const uint64_t beta2[16], const uint64_t gamma2[16],
                    const uint64_t delta2[16], const uint64_t* ic, uint32_t n_public, const uint64_t* publics,
                    const uint64_t pi_a[8], const uint64_t pi_b[16], const uint64_t pi_c[8]);

// sum_i k_i P_i over G1 on the host (windowed bucket method), k canonical
Xyzz<Fp> g1_msm_host(const Aff<Fp>* pts, const uint64_t* k, size_t n);

Aff<Fp> g1_from_words(const uint64_t w[8]);
Aff<Fp2> g2_from_words(const uint64_t w[16]);

}  // namespace zkh
