// gemm.hip - the DST-I Poisson solve as dense matrix products on MFMA.
//
// Maxwell::poisson (crates/embloks-core/src/maxwell.rs:139-168) transforms the
// right-hand side with a type-I sine transform along both axes, scales by the
// inverse Laplacian eigenvalues, and transforms back. A DST-I along an axis is a
// product with the symmetric sine matrix S[k][n] = sin(pi (k+1)(n+1) / (N+1))
// (poisson.wgsl:1-6), so the solve of both right-hand sides (the charge's
// potential and the divergence correction, maxwell.rs:200-204) is four
// products per channel:
//   stage 0: w2 = Sx . w1              stage 1: w1 = (w2 . Sy) * inv_lap
//   stage 2: w2 = Sx . w1              stage 3: phi = w2 . Sy
// Each block computes a 64 x 64 output tile with FP32 MFMA (v_mfma_f32_16x16x4f32):
// four 64-wide waves, each a 32 x 32 sub-tile of 2 x 2 MFMA tiles, operands
// staged through LDS in K slices of 16.

#include "device.hpp"

namespace emb {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GM = 64, GN = 64, GK = 16;
constexpr int LDA = GM + 4, LDB = GN + 4;

template <int STAGE>
__global__ __launch_bounds__(256) void k_dst_gemm(Batch b) {
    constexpr bool LEFT = (STAGE == 0 || STAGE == 2);
    __shared__ float As[2][GK * LDA];
    __shared__ float Bs[2][GK * LDB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int l16 = lane & 15, kq = lane >> 4;
    for (int blk = blockIdx.x; blk < b.total[KC_GEMM]; blk += gridDim.x) {
        const int s = scene_of(b, KC_GEMM, blk);
        const DevScene &S = b.scenes[s];
        const int nx = S.nx, ny = S.ny;
        const int local = block_in_scene(b, KC_GEMM, s, blk);
        const int tm_n = (nx + GM - 1) / GM, tn_n = (ny + GN - 1) / GN;
        const int ch = local / (tm_n * tn_n);
        const int rem = local - ch * tm_n * tn_n;
        const int m0 = (rem / tn_n) * GM, n0 = (rem % tn_n) * GN;
        const size_t nodes = (size_t)nx * ny;
        const float *in = (STAGE == 0 || STAGE == 2) ? S.w1 : S.w2;
        in += ch * nodes;
        float *out = STAGE == 3 ? S.phi : (STAGE == 1 ? S.w1 : S.w2);
        out += ch * nodes;
        const int K = LEFT ? nx : ny;
        // A (M x K): LEFT Sx[row][k], RIGHT in[row][k]. B (K x N): LEFT in[k][col], RIGHT Sy[k][col].
        const float *A = LEFT ? S.sine_x : in;
        const int lda = LEFT ? nx : ny;
        const float *B = LEFT ? in : S.sine_y;
        const int ldb = ny;
        floatx4 acc[2][2];
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++) acc[x][y] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        // Global -> register staging of one K slice.
        const int arow = tid >> 2, akq = (tid & 3) * 4;
        const int bk = tid >> 4, bcol = (tid & 15) * 4;
        float ra[4], rb[4];
        auto load = [&](int k0) {
            const int row = m0 + arow;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int k = k0 + akq + e;
                ra[e] = (row < nx && k < K) ? A[(size_t)row * lda + k] : 0.0f;
            }
            const int k = k0 + bk;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int col = n0 + bcol + e;
                rb[e] = (k < K && col < ny) ? B[(size_t)k * ldb + col] : 0.0f;
            }
        };
        auto store = [&](int buf) {
#pragma unroll
            for (int e = 0; e < 4; e++) As[buf][(akq + e) * LDA + arow] = ra[e];
#pragma unroll
            for (int e = 0; e < 4; e++) Bs[buf][bk * LDB + bcol + e] = rb[e];
        };
        load(0);
        store(0);
        __syncthreads();
        int buf = 0;
        for (int k0 = 0; k0 < K; k0 += GK) {
            const bool more = k0 + GK < K;
            if (more) load(k0 + GK);
#pragma unroll
            for (int kk = 0; kk < GK; kk += 4) {
                const float a0 = As[buf][(kk + kq) * LDA + wm + l16];
                const float a1 = As[buf][(kk + kq) * LDA + wm + 16 + l16];
                const float b0 = Bs[buf][(kk + kq) * LDB + wn + l16];
                const float b1 = Bs[buf][(kk + kq) * LDB + wn + 16 + l16];
                acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
            }
            if (more) store(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
        // D[row][col]: lane holds rows 4 * (lane / 16) + i of column lane % 16.
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int row = m0 + wm + x * 16 + kq * 4 + i;
                    const int col = n0 + wn + y * 16 + l16;
                    if (row < nx && col < ny) {
                        float v = acc[x][y][i];
                        if (STAGE == 1) v *= S.inv_lap[(size_t)row * ny + col];
                        out[(size_t)row * ny + col] = v;
                    }
                }
        __syncthreads();
    }
}

template __global__ void k_dst_gemm<0>(Batch);
template __global__ void k_dst_gemm<1>(Batch);
template __global__ void k_dst_gemm<2>(Batch);
template __global__ void k_dst_gemm<3>(Batch);

}  // namespace emb

namespace emb {

// The four products of one Poisson solve for every scene of the batch.
hipError_t launch_dst_gemms(const Batch &b, hipStream_t stream) {
    const int total = b.total[KC_GEMM];
    if (total == 0) return hipSuccess;
    const int grid = total < 8192 ? total : 8192;
    k_dst_gemm<0><<<grid, 256, 0, stream>>>(b);
    k_dst_gemm<1><<<grid, 256, 0, stream>>>(b);
    k_dst_gemm<2><<<grid, 256, 0, stream>>>(b);
    k_dst_gemm<3><<<grid, 256, 0, stream>>>(b);
    return hipGetLastError();
}

}  // namespace emb
