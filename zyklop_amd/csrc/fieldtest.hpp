// Element-wise device field operations for tests of the field code (see
// fieldtest.hip).
#pragma once
#include <stdint.h>

#include "common.hpp"

namespace zk {

// standard form (8 x 32-bit words, Montgomery R = 2^256); field 0 = Fp, 1 = Fr
enum { FOP_ADD = 0, FOP_SUB, FOP_NEG, FOP_MUL, FOP_SQR, FOP_INV, FOP_TO_MONT, FOP_FROM_MONT, FOP_MUL_CIOS, FOP_COUNT };
// 9 x 29-bit limbs (Fp, Montgomery R' = 2^261); the F29Q ops take element
// pairs (c0, c1) of Fp2
enum {
    F29_MUL = 0, F29_SQR, F29_MULSUM2, F29_MULSUM4, F29_NEG, F29_SUB4, F29_REDUCE, F29_ROUNDTRIP,
    F29_NEG4, F29_Y3, F29Q_MUL, F29Q_MSM, F29_COUNT
};

void field_op_std(int field, int op, const void* a, const void* b, void* out, size_t n, hipStream_t st);
void field29_op(int op, const uint32_t* a, const uint32_t* b, uint32_t* out, size_t n, hipStream_t st);
// lane-operations per second of v_mad_u64_u32 (kind 0) or v_add_u32 (kind 1)
// over the whole GPU (8 independent chains per thread, 8 waves per SIMD)
double valu_probe(int kind, hipStream_t st);

}  // namespace zk
