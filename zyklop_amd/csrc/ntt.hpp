// Fr NTT on gfx950: natural order in and out, w_n = 5^((r-1)/n) (the
// reference's convention, include/ntt_bn254.cuh:190-261,300-464).
//
// Design: four-step recursion n = R_0 * R_1 * ... * R_{p-1} with R_j = 8
// (one radix-16 or radix-4 pass absorbs the leftover bits). In every pass
// each thread owns one column of R_j points: it loads them (consecutive
// lanes on consecutive columns, so every load is coalesced), runs the
// radix-R_j DFT in registers, multiplies by the inter-pass twiddle
// w_{M_j}^{c k} from a precomputed [k][c] table (also coalesced) and stores.
// Two consecutive radix-8 passes run in one kernel, the intermediate points
// staged in LDS. Passes 0..p-2 run in place; the last pass writes the
// digit-reversed (natural) order out of place with lanes on consecutive
// output indices. An optional element-wise table multiplies the input (first
// pass) or the output (last pass): that is how the coset shift and n^-1 are
// fused instead of separate element-wise passes (src/groth16.cu:914-921,
// 1066-1076).
// Arithmetic: elements are loaded into the 9 x 29-bit limbs of field29.hpp
// and multiplied there (no carry instruction per partial product); memory
// holds the standard form (x*2^256 mod r, canonical), and every table value
// is prescaled by 2^5 so that the 29-bit product keeps that form.
#pragma once
#include <stdint.h>

#include "common.hpp"

namespace zk {

constexpr int NTT_MAX_PASSES = 10;

struct NttPlan {
    int log_n = -1;
    int np = 0;
    int lr[NTT_MAX_PASSES] = {};
    size_t n = 0;
    // per-pass twiddle tables (forward / inverse), concatenated; offsets per pass
    DevBuf tw_fwd, tw_inv;
    size_t tw_off[NTT_MAX_PASSES] = {};
    // radix roots w_R^i (i < R/2, Montgomery, prescaled by 2^5) per pass, forward and inverse
    uint64_t roots_fwd[NTT_MAX_PASSES][32] = {};
    uint64_t roots_inv[NTT_MAX_PASSES][32] = {};
    // element-wise tables (Montgomery Fr prescaled by 2^5, n entries each; the
    // pass twiddle tables above too)
    DevBuf coset_fwd;   // w_2n^i           (coset forward prologue)
    DevBuf inv_coset;   // n^-1 * w_2n^i    (iNTT then coset shift: H pipeline)
    DevBuf inv_plain;   // n^-1             (plain inverse)
    DevBuf inv_uncoset; // n^-1 * w_2n^-i   (inverse of the coset NTT)
    void build(int log_n, hipStream_t st);
    // device bytes a plan of size 2^log_n holds
    static size_t bytes_for(int log_n);
};

// in -> out (in is clobbered unless in == out and np == 1). Each of the
// `batch` arrays has n Montgomery Fr elements. pre/post: optional n-entry
// element-wise multipliers (device), indexed by natural element index.
void ntt_run(const NttPlan& p, void* const* in, void* const* out, int batch, bool inverse, const void* pre,
             const void* post, hipStream_t st);

}  // namespace zk
