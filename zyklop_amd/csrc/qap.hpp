// QAP kernels: sparse A*w / B*w (with C*w = A*w o B*w fused), the H-pipeline
// epilogue, the prover's scalar-vector extras and the fixed-base kernel used
// to build synthetic CRS points.
#pragma once
#include <stdint.h>

#include "common.hpp"

namespace zk {

// Aw[i] = sum_k frMul(valA[k], w[colA[k]]) for k in row i (raw = c*R^2, so the
// product lands in Montgomery form), likewise Bw; Cw = Aw o Bw. Rows at or
// beyond n_rowsA / n_rowsB are zero. Outputs n Montgomery Fr.
// Reference: k_sparse_mvm_direct + k_fr_pointwise_mul (src/groth16.cu:144-186).
void spmv_abc(const uint32_t* rowA, const uint32_t* colA, const void* valA, uint32_t n_rowsA, const uint32_t* rowB,
              const uint32_t* colB, const void* valB, uint32_t n_rowsB, const void* w, size_t n, void* aw, void* bw,
              void* cw, hipStream_t st);

// c[i] = a[i] * b[i] (Montgomery)
void pointwise_mul(const void* a, const void* b, void* c, size_t n, hipStream_t st);

// h[i] = canonical(A[i]*B[i] - C[i]) (src/groth16.cu:935-948 + k_fr_decode_batch :218)
void h_final(const void* a, const void* b, const void* c, void* h, size_t n, hipStream_t st);

// writes 4 canonical scalars at dst (32 B each)
void write_scalars4(void* dst, const uint64_t v[16], hipStream_t st);

// out[i] = e_i * G for canonical exponents e (n x 32 B); table: 32 windows x
// 256 affine Montgomery multiples of the generator (T[j][d] = d*2^(8j)*G).
// Output affine Montgomery (G1 64 B / G2 128 B), infinity as zeros.
void fixed_base_mul(const void* table, const void* exps, size_t n, void* out, bool g2, hipStream_t st);

// Synthetic witness, one dependency level: w[first_aux + i] = canonical
// (A_i.w)(B_i.w) for each row i in rows[0..count), CSR values raw c*R^2, w
// canonical (the wires the rows read must already hold their final values).
void witness_rows(const uint32_t* rowA, const uint32_t* colA, const void* valA, const uint32_t* rowB,
                  const uint32_t* colB, const void* valB, const uint32_t* rows, uint32_t count, uint32_t first_aux,
                  void* w, hipStream_t st);

// fill a u32 buffer (capture-safe alternative to hipMemsetAsync)
void fill_u32(void* p, uint32_t v, size_t count, hipStream_t st);

}  // namespace zk
