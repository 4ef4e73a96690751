// gemm.hpp - dense products of the DST-I Poisson solve (gemm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace emb {

// C = A . B, all row-major FP32: A is M x K (row pitch lda), B is K x N (ldb),
// C is M x N (ldc). When `scale` is set, C is multiplied elementwise by it
// (an M x N array with pitch ldc).
//
// Folded (one half of a DST-I product through the matrix's symmetry, see
// sine_halves): with fold = 1, row r of B is read as B[r] + s B[fold_n - 1 - r]
// (B[r] alone where the two are the same row), and result row i goes to row
// 2 i + odd of C; with fold = 2, column c of A is read as A[.][c] + s
// A[.][fold_n - 1 - c] and result column j goes to column 2 j + odd. s = +1 for
// the even half (odd = 0), -1 for the odd half; `scale` is read at C's position.
//
// Np (when set) is how many columns of B the loads may touch, a multiple of 4
// >= N: B's columns past N up to Np are padding (their results are not
// stored). K may include zero padding of the sine half (its rows or columns
// past the half's size are zero), so that K is a whole number of slices.
struct GemmArgs {
    const float *A = nullptr, *B = nullptr;
    float *C = nullptr;
    const float *scale = nullptr;
    int M = 0, N = 0, K = 0, lda = 0, ldb = 0, ldc = 0;
    int fold = 0, fold_n = 0, odd = 0;
    int Np = 0;
};

// K of a folded half product over a half of side mh: padded to whole slices
// when the padding stays inside the folded operand (it multiplies the sine
// half's zero padding).
__host__ __device__ inline int fold_k(int mh, int fold_n) {
    const int kp = sine_half_pitch(mh);
    return kp <= fold_n ? kp : mh;
}

// One FP32 product with 64 x 64 (tile 64) or 128 x 64 (tile 128) output tiles.
hipError_t launch_gemm(const GemmArgs &g, int tile, hipStream_t stream);

// A whole DST-I product through its two folded halves (same fold, odd = 0 and 1)
// in one launch.
hipError_t launch_gemm_fold(const GemmArgs &even, const GemmArgs &odd, int tile, hipStream_t stream);

// BF16x3: split an FP32 rows x cols matrix (pitch ld) into bf16 halves whose
// rows have pitch ldd (a multiple of 32, zero-padded), transposed or not ...
hipError_t launch_split(const float *src, int rows, int cols, int ld, bool transpose, __bf16 *hi, __bf16 *lo, int ldd,
                        hipStream_t stream);
// ... and C = A . B^T from K-contiguous halves (A: M x K, B: N x K, K a
// multiple of 32), FP32 accumulation, optional elementwise scale.
hipError_t launch_gemm_nt_bf16x3(const __bf16 *ah, const __bf16 *al, int lda, const __bf16 *bh, const __bf16 *bl,
                                 int ldb, float *C, const float *scale, int ldc, int M, int N, int K, int tile,
                                 hipStream_t stream);

// The four products of the tick's solve for every scene of a batch, and the
// blocks per scene they need (KC_GEMM).
// Which channels a solve runs: both (scene set-up), the tick's (the residual
// always, the charge's potential where the scene has battery stacks), or the
// potential alone (read on demand).
enum { DST_BOTH = 0, DST_TICK = 1, DST_POT = 2 };
hipError_t launch_dst_gemms(const Batch &b, hipStream_t stream, int chans);
// (KC_GEMM: 64 x 64 tiles; KC_GEMM_SMALL: the small tiles a batch with few
// 64 x 64 tiles runs instead)
int dst_gemm_blocks(int nx, int ny);
int dst_gemm_small_blocks(int nx, int ny);

}  // namespace emb
