// gemm.hip - the DST-I Poisson solve as dense matrix products on MFMA.
//
// Maxwell::poisson (crates/embloks-core/src/maxwell.rs:139-168) transforms the
// right-hand side with a type-I sine transform along both axes, scales by the
// inverse Laplacian eigenvalues, and transforms back. A DST-I along an axis is a
// product with the symmetric sine matrix S[k][n] = sin(pi (k+1)(n+1) / (N+1))
// (poisson.wgsl:1-6), so one solve is four products:
//   stage 0: w2 = Sx . w1              stage 1: w1 = (w2 . Sy) * inv_lap
//   stage 2: w2 = Sx . w1              stage 3: phi = w2 . Sy
// The tick solves two right-hand sides (the charge's potential and the
// divergence correction, maxwell.rs:200-204) side by side.
//
// Two tile kernels compute C = A . B (row-major, optional elementwise scale):
//   * FP32: v_mfma_f32_16x16x4f32, FP32 in and out; what the tick runs.
//   * BF16x3: each operand split as x = hi + lo (hi = bf16(x), lo = bf16(x - hi)),
//     A.B ~ lo.hi + hi.lo + hi.hi with v_mfma_f32_16x16x32_bf16 and FP32
//     accumulation (the lo.lo term is dropped). Offered for the isolated solve;
//     its error is measured against FP32 and FP64. The operands are split (and
//     transposed where the contraction runs over rows) by an O(n^2) pass before
//     each product, so the O(n^3) kernel only loads bf16 and multiplies.
// Both use 256 threads as a 2 x 2 grid of waves, operands staged through LDS
// in K slices with the next slice prefetched into registers.

#include "device.hpp"
#include "gemm.hpp"

namespace emb {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Four consecutive floats of a row, zero past `limit` (vector load when whole and aligned).
__device__ __forceinline__ float4 load4(const float *p, int avail, bool vec) {
    if (vec && avail >= 4) return *reinterpret_cast<const float4 *>(p);
    float4 r;
    r.x = avail > 0 ? p[0] : 0.0f;
    r.y = avail > 1 ? p[1] : 0.0f;
    r.z = avail > 2 ? p[2] : 0.0f;
    r.w = avail > 3 ? p[3] : 0.0f;
    return r;
}

// ---------------------------------------------------------------------------------------
// FP32 tile: BM x BN output, K slices of 16, waves of (BM/2) x (BN/2).

template <int BM, int BN>
__device__ __forceinline__ void tile_f32(const GemmArgs &g, int m0, int n0, float *lds) {
    constexpr int GK = 16, LA = BM + 4, LB = BN + 4;
    constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
    constexpr int AV = BM * GK / 4 / 256, BV = BN * GK / 4 / 256;
    static_assert(AV >= 1 && BV >= 1, "tile too small for 256 threads");
    float *As = lds, *Bs = lds + 2 * GK * LA;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
    const int l16 = lane & 15, kq = lane >> 4;
    const bool avec = (g.lda & 3) == 0 && (reinterpret_cast<uintptr_t>(g.A) & 15) == 0;
    const bool bvec = (g.ldb & 3) == 0 && (reinterpret_cast<uintptr_t>(g.B) & 15) == 0;
    floatx4 acc[TM][TN];
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
        for (int y = 0; y < TN; y++) acc[x][y] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    float4 ra[AV], rb[BV];
    auto load = [&](int k0) {
#pragma unroll
        for (int v = 0; v < AV; v++) {
            const int idx = tid + 256 * v, row = idx >> 2, kk = (idx & 3) * 4;
            const int gr = m0 + row, gk = k0 + kk;
            ra[v] = gr < g.M ? load4(g.A + (size_t)gr * g.lda + gk, g.K - gk, avec) : float4{0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int v = 0; v < BV; v++) {
            const int idx = tid + 256 * v, kr = idx / (BN / 4), col = (idx % (BN / 4)) * 4;
            const int gk = k0 + kr, gc = n0 + col;
            rb[v] = gk < g.K ? load4(g.B + (size_t)gk * g.ldb + gc, g.N - gc, bvec) : float4{0.0f, 0.0f, 0.0f, 0.0f};
        }
    };
    auto store = [&](int buf) {
        float *a = As + buf * GK * LA, *b = Bs + buf * GK * LB;
#pragma unroll
        for (int v = 0; v < AV; v++) {
            const int idx = tid + 256 * v, row = idx >> 2, kk = (idx & 3) * 4;
            a[(kk + 0) * LA + row] = ra[v].x;
            a[(kk + 1) * LA + row] = ra[v].y;
            a[(kk + 2) * LA + row] = ra[v].z;
            a[(kk + 3) * LA + row] = ra[v].w;
        }
#pragma unroll
        for (int v = 0; v < BV; v++) {
            const int idx = tid + 256 * v, kr = idx / (BN / 4), col = (idx % (BN / 4)) * 4;
            *reinterpret_cast<float4 *>(b + kr * LB + col) = rb[v];
        }
    };
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < g.K; k0 += GK) {
        const bool more = k0 + GK < g.K;
        if (more) load(k0 + GK);
        const float *a = As + buf * GK * LA, *b = Bs + buf * GK * LB;
#pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            float af[TM], bf[TN];
#pragma unroll
            for (int x = 0; x < TM; x++) af[x] = a[(kk + kq) * LA + wm + x * 16 + l16];
#pragma unroll
            for (int y = 0; y < TN; y++) bf[y] = b[(kk + kq) * LB + wn + y * 16 + l16];
#pragma unroll
            for (int x = 0; x < TM; x++)
#pragma unroll
                for (int y = 0; y < TN; y++)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[x], bf[y], acc[x][y], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // D[row][col]: lane holds rows 4 * (lane / 16) + i of column lane % 16.
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
        for (int y = 0; y < TN; y++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = m0 + wm + x * 16 + kq * 4 + i, col = n0 + wn + y * 16 + l16;
                if (row < g.M && col < g.N) {
                    float v = acc[x][y][i];
                    if (g.scale) v *= g.scale[(size_t)row * g.ldc + col];
                    g.C[(size_t)row * g.ldc + col] = v;
                }
            }
}

// ---------------------------------------------------------------------------------------
// The tick's solve over a batch of scenes: blocks walk (scene, channel, tile).

constexpr int TICK_BM = 64, TICK_BN = 64;

template <int STAGE>
__global__ __launch_bounds__(256) void k_dst_gemm(Batch b) {
    constexpr bool LEFT = (STAGE == 0 || STAGE == 2);
    __shared__ __align__(16) float lds[2 * 16 * (TICK_BM + 4 + TICK_BN + 4)];
    for (int blk = blockIdx.x; blk < b.total[KC_GEMM]; blk += gridDim.x) {
        const int s = scene_of(b, KC_GEMM, blk);
        const DevScene &S = b.scenes[s];
        const int nx = S.nx, ny = S.ny;
        const int local = block_in_scene(b, KC_GEMM, s, blk);
        const int tm_n = (nx + TICK_BM - 1) / TICK_BM, tn_n = (ny + TICK_BN - 1) / TICK_BN;
        const int ch = local / (tm_n * tn_n);
        const int rem = local - ch * tm_n * tn_n;
        const size_t nodes = (size_t)nx * ny;
        const float *in = (LEFT ? S.w1 : S.w2) + ch * nodes;
        float *out = (STAGE == 3 ? S.phi : (STAGE == 1 ? S.w1 : S.w2)) + ch * nodes;
        GemmArgs g;
        g.A = LEFT ? S.sine_x : in;
        g.lda = LEFT ? nx : ny;
        g.B = LEFT ? in : S.sine_y;
        g.ldb = ny;
        g.C = out;
        g.ldc = ny;
        g.scale = STAGE == 1 ? S.inv_lap : nullptr;
        g.M = nx;
        g.N = ny;
        g.K = LEFT ? nx : ny;
        tile_f32<TICK_BM, TICK_BN>(g, (rem / tn_n) * TICK_BM, (rem % tn_n) * TICK_BN, lds);
        __syncthreads();
    }
}

template __global__ void k_dst_gemm<0>(Batch);
template __global__ void k_dst_gemm<1>(Batch);
template __global__ void k_dst_gemm<2>(Batch);
template __global__ void k_dst_gemm<3>(Batch);

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

// ---------------------------------------------------------------------------------------
// One product on plain pointers (the isolated solve).

template <int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm_f32(GemmArgs g) {
    __shared__ __align__(16) float lds[2 * 16 * (BM + 4 + BN + 4)];
    const int tn = (g.N + BN - 1) / BN, tiles = ((g.M + BM - 1) / BM) * tn;
    for (int blk = blockIdx.x; blk < tiles; blk += gridDim.x) {
        tile_f32<BM, BN>(g, (blk / tn) * BM, (blk % tn) * BN, lds);
        __syncthreads();
    }
}

hipError_t launch_gemm(const GemmArgs &g, int tile, hipStream_t stream) {
    const int t = tile == 128 ? 128 : 64;
    const long long tiles = (long long)((g.M + t - 1) / t) * ((g.N + t - 1) / t);
    const int grid = (int)(tiles < 8192 ? tiles : 8192);
    if (grid == 0) return hipSuccess;
    if (t == 128) k_gemm_f32<128, 128><<<grid, 256, 0, stream>>>(g);
    else k_gemm_f32<64, 64><<<grid, 256, 0, stream>>>(g);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// BF16x3: x = hi + lo with hi = bf16(x) and lo = bf16(x - hi); a product keeps
// hi.hi + hi.lo + lo.hi (the lo.lo term, ~2^-16 of it, is dropped).

// Split an FP32 rows x cols matrix (pitch ld) into bf16 halves with rows of
// pitch ldd (a multiple of 32, zero past the data): straight, dst[r][c], or
// TRANSPOSE, dst[c][r]. 32 x 32 tiles through LDS keep both sides coalesced.
template <bool TRANSPOSE>
__global__ __launch_bounds__(256) void k_split(const float *src, int rows, int cols, int ld, __bf16 *hi, __bf16 *lo,
                                               int ldd) {
    __shared__ float tile[32][33];
    const int out_rows = TRANSPOSE ? cols : rows, out_cols = ldd;
    const int tr = (out_rows + 31) / 32, tc = (out_cols + 31) / 32;
    for (int blk = blockIdx.x; blk < tr * tc; blk += gridDim.x) {
        const int r0 = (blk / tc) * 32, c0 = (blk % tc) * 32;   // output tile
        for (int e = threadIdx.x; e < 1024; e += 256) {
            const int i = e >> 5, j = e & 31;
            // source element for output (r0 + i, c0 + j); read along the source's rows
            const int sr = TRANSPOSE ? c0 + i : r0 + i, sc = TRANSPOSE ? r0 + j : c0 + j;
            tile[i][j] = (sr < rows && sc < cols) ? src[(size_t)sr * ld + sc] : 0.0f;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 1024; e += 256) {
            const int i = e >> 5, j = e & 31;
            const int r = r0 + i, c = c0 + j;
            if (r < out_rows && c < out_cols) {
                const float x = TRANSPOSE ? tile[j][i] : tile[i][j];
                const __bf16 h = (__bf16)x;
                hi[(size_t)r * ldd + c] = h;
                lo[(size_t)r * ldd + c] = (__bf16)(x - (float)h);
            }
        }
        __syncthreads();
    }
}

// C = A . B^T with A (M x K) and B (N x K) given as K-contiguous bf16 halves of
// pitch lda / ldb (multiples of 32, zero past K), FP32 accumulation, optional
// elementwise scale. BM x BN tiles, K slices of 32 (one v_mfma_f32_16x16x32_bf16
// per 16 x 16 x 32 block and half pairing), the next slice prefetched into
// registers while the current one is multiplied.
struct NtArgs {
    const __bf16 *ah = nullptr, *al = nullptr, *bh = nullptr, *bl = nullptr;
    int lda = 0, ldb = 0;
    float *C = nullptr;
    const float *scale = nullptr;
    int ldc = 0, M = 0, N = 0, K = 0;
};

template <int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm_nt_bf16x3(NtArgs g) {
    constexpr int GK = 32, P = GK + 8;   // LDS row pitch in bf16 (80 bytes: 16-byte reads stay aligned)
    constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
    constexpr int CA = BM * (GK / 8) / 256, CB = BN * (GK / 8) / 256;   // 16-byte chunks per thread per half
    static_assert(CA >= 1 && CB >= 1, "tile too small for 256 threads");
    __shared__ __align__(16) __bf16 lds[2 * (BM + BN) * P];
    __bf16 *Ah = lds, *Al = Ah + BM * P, *Bh = Al + BM * P, *Bl = Bh + BN * P;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
    const int l16 = lane & 15, kq = lane >> 4;
    const int tn = (g.N + BN - 1) / BN, tiles = ((g.M + BM - 1) / BM) * tn;
    for (int blk = blockIdx.x; blk < tiles; blk += gridDim.x) {
        const int m0 = (blk / tn) * BM, n0 = (blk % tn) * BN;
        floatx4 acc[TM][TN];
#pragma unroll
        for (int x = 0; x < TM; x++)
#pragma unroll
            for (int y = 0; y < TN; y++) acc[x][y] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        uint4 rah[CA], ral[CA], rbh[CB], rbl[CB];
        auto load = [&](int k0) {
#pragma unroll
            for (int v = 0; v < CA; v++) {
                const int c = tid + 256 * v, row = c / (GK / 8), part = c % (GK / 8);
                const bool ok = m0 + row < g.M;
                const size_t off = (size_t)(m0 + row) * g.lda + k0 + part * 8;
                rah[v] = ok ? *reinterpret_cast<const uint4 *>(g.ah + off) : uint4{0u, 0u, 0u, 0u};
                ral[v] = ok ? *reinterpret_cast<const uint4 *>(g.al + off) : uint4{0u, 0u, 0u, 0u};
            }
#pragma unroll
            for (int v = 0; v < CB; v++) {
                const int c = tid + 256 * v, row = c / (GK / 8), part = c % (GK / 8);
                const bool ok = n0 + row < g.N;
                const size_t off = (size_t)(n0 + row) * g.ldb + k0 + part * 8;
                rbh[v] = ok ? *reinterpret_cast<const uint4 *>(g.bh + off) : uint4{0u, 0u, 0u, 0u};
                rbl[v] = ok ? *reinterpret_cast<const uint4 *>(g.bl + off) : uint4{0u, 0u, 0u, 0u};
            }
        };
        auto store = [&]() {
#pragma unroll
            for (int v = 0; v < CA; v++) {
                const int c = tid + 256 * v, row = c / (GK / 8), part = c % (GK / 8);
                *reinterpret_cast<uint4 *>(Ah + row * P + part * 8) = rah[v];
                *reinterpret_cast<uint4 *>(Al + row * P + part * 8) = ral[v];
            }
#pragma unroll
            for (int v = 0; v < CB; v++) {
                const int c = tid + 256 * v, row = c / (GK / 8), part = c % (GK / 8);
                *reinterpret_cast<uint4 *>(Bh + row * P + part * 8) = rbh[v];
                *reinterpret_cast<uint4 *>(Bl + row * P + part * 8) = rbl[v];
            }
        };
        load(0);
        for (int k0 = 0; k0 < g.K; k0 += GK) {
            __syncthreads();
            store();
            __syncthreads();
            if (k0 + GK < g.K) load(k0 + GK);
            bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
            for (int x = 0; x < TM; x++) {
                const int r = (wm + x * 16 + l16) * P + kq * 8;
                ah[x] = *reinterpret_cast<const bf16x8 *>(Ah + r);
                al[x] = *reinterpret_cast<const bf16x8 *>(Al + r);
            }
#pragma unroll
            for (int y = 0; y < TN; y++) {
                const int r = (wn + y * 16 + l16) * P + kq * 8;
                bh[y] = *reinterpret_cast<const bf16x8 *>(Bh + r);
                bl[y] = *reinterpret_cast<const bf16x8 *>(Bl + r);
            }
#pragma unroll
            for (int x = 0; x < TM; x++)
#pragma unroll
                for (int y = 0; y < TN; y++) {
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
                }
        }
#pragma unroll
        for (int x = 0; x < TM; x++)
#pragma unroll
            for (int y = 0; y < TN; y++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int row = m0 + wm + x * 16 + kq * 4 + i, col = n0 + wn + y * 16 + l16;
                    if (row < g.M && col < g.N) {
                        float v = acc[x][y][i];
                        if (g.scale) v *= g.scale[(size_t)row * g.ldc + col];
                        g.C[(size_t)row * g.ldc + col] = v;
                    }
                }
        __syncthreads();
    }
}

hipError_t launch_split(const float *src, int rows, int cols, int ld, bool transpose, __bf16 *hi, __bf16 *lo, int ldd,
                        hipStream_t stream) {
    const int out_rows = transpose ? cols : rows;
    const long long tiles = (long long)((out_rows + 31) / 32) * ((ldd + 31) / 32);
    const int grid = (int)(tiles < 8192 ? tiles : 8192);
    if (transpose) k_split<true><<<grid, 256, 0, stream>>>(src, rows, cols, ld, hi, lo, ldd);
    else k_split<false><<<grid, 256, 0, stream>>>(src, rows, cols, ld, hi, lo, ldd);
    return hipGetLastError();
}

hipError_t launch_gemm_nt_bf16x3(const __bf16 *ah, const __bf16 *al, int lda, const __bf16 *bh, const __bf16 *bl,
                                 int ldb, float *C, const float *scale, int ldc, int M, int N, int K, int tile,
                                 hipStream_t stream) {
    NtArgs g;
    g.ah = ah; g.al = al; g.bh = bh; g.bl = bl; g.lda = lda; g.ldb = ldb;
    g.C = C; g.scale = scale; g.ldc = ldc; g.M = M; g.N = N; g.K = K;
    const int t = tile == 128 ? 128 : 64;
    const long long tiles = (long long)((M + t - 1) / t) * ((N + t - 1) / t);
    const int grid = (int)(tiles < 8192 ? tiles : 8192);
    if (grid == 0) return hipSuccess;
    if (t == 128) k_gemm_nt_bf16x3<128, 128><<<grid, 256, 0, stream>>>(g);
    else k_gemm_nt_bf16x3<64, 64><<<grid, 256, 0, stream>>>(g);
    return hipGetLastError();
}

}  // namespace emb
