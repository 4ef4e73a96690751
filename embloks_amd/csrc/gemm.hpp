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

enum { GEMM_FP32 = 0, GEMM_BF16X3 = 1 };

// One product with 64 x 64 or 128 x 128 output tiles.
hipError_t launch_gemm(const GemmArgs &g, int precision, int tile, hipStream_t stream);

// The four products of the tick's solve for every scene of a batch.
hipError_t launch_dst_gemms(const Batch &b, hipStream_t stream);

}  // namespace emb
