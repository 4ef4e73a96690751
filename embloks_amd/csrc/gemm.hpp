// gemm.hpp - dense products of the DST-I Poisson solve (gemm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace emb {

// C = A . B, all row-major FP32: A is M x K (row pitch lda), B is K x N (ldb),
// C is M x N (ldc). When `scale` is set, C is multiplied elementwise by it
// (an M x N array with pitch ldc).
struct GemmArgs {
    const float *A = nullptr, *B = nullptr;
    float *C = nullptr;
    const float *scale = nullptr;
    int M = 0, N = 0, K = 0, lda = 0, ldb = 0, ldc = 0;
};

// One FP32 product with 64 x 64 (tile 64) or 128 x 64 (tile 128) output tiles.
hipError_t launch_gemm(const GemmArgs &g, int tile, hipStream_t stream);

// BF16x3: split an FP32 rows x cols matrix (pitch ld) into bf16 halves whose
// rows have pitch ldd (a multiple of 32, zero-padded), transposed or not ...
hipError_t launch_split(const float *src, int rows, int cols, int ld, bool transpose, __bf16 *hi, __bf16 *lo, int ldd,
                        hipStream_t stream);
// ... and C = A . B^T from K-contiguous halves (A: M x K, B: N x K, K a
// multiple of 32), FP32 accumulation, optional elementwise scale.
hipError_t launch_gemm_nt_bf16x3(const __bf16 *ah, const __bf16 *al, int lda, const __bf16 *bh, const __bf16 *bl,
                                 int ldb, float *C, const float *scale, int ldc, int M, int N, int K, int tile,
                                 hipStream_t stream);

// The four products of the tick's solve for every scene of a batch.
hipError_t launch_dst_gemms(const Batch &b, hipStream_t stream);

}  // namespace emb
