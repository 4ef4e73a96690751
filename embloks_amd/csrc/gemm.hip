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
//   * FP32: v_mfma_f32_16x16x4f32, FP32 in and out; what the tick runs. Four
//     K slices are in flight per block, which is what the tick's small K needs.
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

#include <type_traits>
#include <utility>

namespace emb {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>), so that an
// index into a register array stays a compile-time constant.
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

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
// FP32 tile: BM x BN output on a GM x GN grid of waves (each (BM/GM) x (BN/GN)),
// K slices of GK. `lds` holds 2 * GK * (BM + 4 + BN + 4) floats.
//
// At the tick's sizes (K of a few hundred) a product is bound by the latency of
// its operand loads, not by the MFMA rate, so PF slices are in flight at once:
// register slot u holds slice s + 1 while slice s is multiplied out of LDS, and
// is refilled with slice s + 1 + PF once stored. The loop is unrolled by PF so
// every slot is a fixed set of registers.
//
// FAST (K a multiple of GK, N of 4, lda/ldb of 4, 16-byte aligned operands)
// loads whole float4s with no per-element guards: rows of A past M and columns
// of B past N are clamped to the last valid ones, since they only reach output
// rows and columns that are never stored.

template <int BM, int BN, int GM = 2, int GN = 2, int GK = 16, int PF = 1, bool FAST = false>
__device__ __forceinline__ void tile_f32(const GemmArgs &g, int m0, int n0, float *lds) {
    constexpr int NT = 64 * GM * GN;
    constexpr int LA = BM + 4, LB = BN + 4, KV = GK / 4;
    constexpr int WM = BM / GM, WN = BN / GN, TM = WM / 16, TN = WN / 16;
    constexpr int AV = BM * GK / 4 / NT, BV = BN * GK / 4 / NT;
    static_assert(AV >= 1 && BV >= 1 && AV * NT * 4 == BM * GK && BV * NT * 4 == BN * GK, "tile and threads mismatch");
    float *As = lds, *Bs = lds + 2 * GK * LA;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / GN) * WM, wn = (wave % GN) * WN;
    const int l16 = lane & 15, kq = lane >> 4;
    const bool avec = (g.lda & 3) == 0 && (reinterpret_cast<uintptr_t>(g.A) & 15) == 0;
    const bool bvec = (g.ldb & 3) == 0 && (reinterpret_cast<uintptr_t>(g.B) & 15) == 0;
    floatx4 acc[TM][TN];
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
        for (int y = 0; y < TN; y++) acc[x][y] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // Per-thread operand addresses at slice 0 (FAST: clamped, so every load is valid).
    const float *pa[AV];
    const float *pb[BV];
    int ra_row[AV], ra_k[AV], rb_k[BV], rb_col[BV];
#pragma unroll
    for (int v = 0; v < AV; v++) {
        const int idx = tid + NT * v;
        ra_row[v] = idx / KV;
        ra_k[v] = (idx % KV) * 4;
        const int gr = FAST ? min(m0 + ra_row[v], g.M - 1) : m0 + ra_row[v];
        pa[v] = g.A + (size_t)gr * g.lda + ra_k[v];
    }
#pragma unroll
    for (int v = 0; v < BV; v++) {
        const int idx = tid + NT * v;
        rb_k[v] = idx / (BN / 4);
        rb_col[v] = (idx % (BN / 4)) * 4;
        const int gc = FAST ? min(n0 + rb_col[v], g.N - 4) : n0 + rb_col[v];
        pb[v] = g.B + (size_t)rb_k[v] * g.ldb + gc;
    }
    // Native vectors, not float4 structs: a struct copy between address spaces
    // goes through private memory.
    floatx4 ra[PF][AV], rb[PF][BV];
    auto vec4 = [](float4 f) { return floatx4{f.x, f.y, f.z, f.w}; };
    auto load = [&](auto slot, int k0) {
#pragma unroll
        for (int v = 0; v < AV; v++) {
            if (FAST) {
                ra[slot][v] = *reinterpret_cast<const floatx4 *>(pa[v] + k0);
            } else {
                const int gr = m0 + ra_row[v], gk = k0 + ra_k[v];
                ra[slot][v] = gr < g.M ? vec4(load4(pa[v] + k0, g.K - gk, avec)) : floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            }
        }
#pragma unroll
        for (int v = 0; v < BV; v++) {
            if (FAST) {
                rb[slot][v] = *reinterpret_cast<const floatx4 *>(pb[v] + (size_t)k0 * g.ldb);
            } else {
                const int gk = k0 + rb_k[v], gc = n0 + rb_col[v];
                rb[slot][v] = gk < g.K ? vec4(load4(pb[v] + (size_t)k0 * g.ldb, g.N - gc, bvec))
                                       : floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            }
        }
    };
    auto store = [&](auto slot, int buf) {
        float *a = As + buf * GK * LA, *b = Bs + buf * GK * LB;
#pragma unroll
        for (int v = 0; v < AV; v++) {
            const int row = ra_row[v], kk = ra_k[v];
            a[(kk + 0) * LA + row] = ra[slot][v][0];
            a[(kk + 1) * LA + row] = ra[slot][v][1];
            a[(kk + 2) * LA + row] = ra[slot][v][2];
            a[(kk + 3) * LA + row] = ra[slot][v][3];
        }
#pragma unroll
        for (int v = 0; v < BV; v++)
            *reinterpret_cast<floatx4 *>(b + rb_k[v] * LB + rb_col[v]) = rb[slot][v];
    };
    auto multiply = [&](int buf) {
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
    };
    // Slices past the last are clamped to it, so every load and store runs
    // unconditionally in the unrolled body (a repeated slice is never multiplied).
    const int steps = (g.K + GK - 1) / GK, last = (steps - 1) * GK;
    load(std::integral_constant<int, 0>{}, 0);
    store(std::integral_constant<int, 0>{}, 0);
    static_for<PF>([&](auto u) { load(u, min((1 + u) * GK, last)); });
    __syncthreads();
    int s0 = 0;
    for (; s0 + PF <= steps; s0 += PF) {
        static_for<PF>([&](auto u) {
            const int st = s0 + u;
            multiply(st & 1);
            store(u, (st + 1) & 1);
            load(u, min((st + 1 + PF) * GK, last));
            __syncthreads();
        });
    }
    // Tail: only the multiply is conditional (a branch around the register
    // slots would make their index dynamic); stores past the end are unread.
    static_for<PF>([&](auto u) {
        const int st = s0 + u;
        if (st < steps) multiply(st & 1);
        store(u, (st + 1) & 1);
        __syncthreads();
    });
    // D[row][col]: lane holds rows 4 * (lane / 16) + i of column lane % 16.
    // Scale factors are read at clamped indices, all before the first is used.
    if (g.scale) {
#pragma unroll
        for (int x = 0; x < TM; x++)
#pragma unroll
            for (int y = 0; y < TN; y++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int row = min(m0 + wm + x * 16 + kq * 4 + i, g.M - 1);
                    const int col = min(n0 + wn + y * 16 + l16, g.N - 1);
                    acc[x][y][i] *= g.scale[(size_t)row * g.ldc + col];
                }
    }
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
        for (int y = 0; y < TN; y++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = m0 + wm + x * 16 + kq * 4 + i, col = n0 + wn + y * 16 + l16;
                if (row < g.M && col < g.N) g.C[(size_t)row * g.ldc + col] = acc[x][y][i];
            }
}

// Whether a product can take tile_f32's FAST path for slices of GK.
__host__ __device__ inline bool gemm_fast(const GemmArgs &g, int gk) {
    return g.K % gk == 0 && g.N % 4 == 0 && g.N >= 4 && g.lda % 4 == 0 && g.ldb % 4 == 0 &&
           (reinterpret_cast<uintptr_t>(g.A) & 15) == 0 && (reinterpret_cast<uintptr_t>(g.B) & 15) == 0;
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
        const int m0 = (rem / tn_n) * TICK_BM, n0 = (rem % tn_n) * TICK_BN;
        if (gemm_fast(g, 16)) tile_f32<TICK_BM, TICK_BN, 2, 2, 16, 4, true>(g, m0, n0, lds);
        else tile_f32<TICK_BM, TICK_BN, 2, 2, 16, 4, false>(g, m0, n0, lds);
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

template <int BM, int BN, int PF, bool FAST>
__global__ __launch_bounds__(256) void k_gemm_f32(GemmArgs g) {
    __shared__ __align__(16) float lds[2 * 16 * (BM + 4 + BN + 4)];
    const int tn = (g.N + BN - 1) / BN, tiles = ((g.M + BM - 1) / BM) * tn;
    for (int blk = blockIdx.x; blk < tiles; blk += gridDim.x) {
        tile_f32<BM, BN, 2, 2, 16, PF, FAST>(g, (blk / tn) * BM, (blk % tn) * BN, lds);
        __syncthreads();
    }
}

template <int BM, int BN>
static hipError_t launch_gemm_tile(const GemmArgs &g, hipStream_t stream) {
    const long long tiles = (long long)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
    const int grid = (int)(tiles < 8192 ? tiles : 8192);
    if (grid == 0) return hipSuccess;
    if (gemm_fast(g, 16)) k_gemm_f32<BM, BN, 4, true><<<grid, 256, 0, stream>>>(g);
    else k_gemm_f32<BM, BN, 4, false><<<grid, 256, 0, stream>>>(g);
    return hipGetLastError();
}

hipError_t launch_gemm(const GemmArgs &g, int tile, hipStream_t stream) {
    return tile == 128 ? launch_gemm_tile<128, 64>(g, stream) : launch_gemm_tile<64, 64>(g, stream);
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
// per 16 x 16 x 32 block and half pairing), PF slices in flight in register
// slots and two LDS buffers, as in tile_f32. Rows past M or N are clamped to the
// last one: they only reach outputs that are never stored.
struct NtArgs {
    const __bf16 *ah = nullptr, *al = nullptr, *bh = nullptr, *bl = nullptr;
    int lda = 0, ldb = 0;
    float *C = nullptr;
    const float *scale = nullptr;
    int ldc = 0, M = 0, N = 0, K = 0;
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int BM, int BN, int PF>
__global__ __launch_bounds__(256) void k_gemm_nt_bf16x3(NtArgs g) {
    constexpr int GK = 32, P = GK + 8;   // LDS row pitch in bf16 (80 bytes: 16-byte reads stay aligned)
    constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
    constexpr int CA = BM * (GK / 8) / 256, CB = BN * (GK / 8) / 256;   // 16-byte chunks per thread per half
    constexpr int BUF = 2 * (BM + BN) * P;                              // one LDS buffer, in bf16
    static_assert(CA >= 1 && CB >= 1, "tile too small for 256 threads");
    __shared__ __align__(16) __bf16 lds[2 * BUF];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
    const int l16 = lane & 15, kq = lane >> 4;
    const int tn = (g.N + BN - 1) / BN, tiles = ((g.M + BM - 1) / BM) * tn;
    const int steps = g.K / GK, last = (steps - 1) * GK;
    for (int blk = blockIdx.x; blk < tiles; blk += gridDim.x) {
        const int m0 = (blk / tn) * BM, n0 = (blk % tn) * BN;
        floatx4 acc[TM][TN];
#pragma unroll
        for (int x = 0; x < TM; x++)
#pragma unroll
            for (int y = 0; y < TN; y++) acc[x][y] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        size_t offa[CA], offb[CB];
        int lda_[CA], ldb_[CB];   // LDS offsets of each chunk
#pragma unroll
        for (int v = 0; v < CA; v++) {
            const int c = tid + 256 * v, row = c / (GK / 8), part = c % (GK / 8);
            offa[v] = (size_t)min(m0 + row, g.M - 1) * g.lda + part * 8;
            lda_[v] = row * P + part * 8;
        }
#pragma unroll
        for (int v = 0; v < CB; v++) {
            const int c = tid + 256 * v, row = c / (GK / 8), part = c % (GK / 8);
            offb[v] = (size_t)min(n0 + row, g.N - 1) * g.ldb + part * 8;
            ldb_[v] = row * P + part * 8;
        }
        u32x4 rah[PF][CA], ral[PF][CA], rbh[PF][CB], rbl[PF][CB];
        auto load = [&](auto slot, int k0) {
#pragma unroll
            for (int v = 0; v < CA; v++) {
                rah[slot][v] = *reinterpret_cast<const u32x4 *>(g.ah + offa[v] + k0);
                ral[slot][v] = *reinterpret_cast<const u32x4 *>(g.al + offa[v] + k0);
            }
#pragma unroll
            for (int v = 0; v < CB; v++) {
                rbh[slot][v] = *reinterpret_cast<const u32x4 *>(g.bh + offb[v] + k0);
                rbl[slot][v] = *reinterpret_cast<const u32x4 *>(g.bl + offb[v] + k0);
            }
        };
        auto store = [&](auto slot, int buf) {
            __bf16 *Ah = lds + buf * BUF, *Al = Ah + BM * P, *Bh = Al + BM * P, *Bl = Bh + BN * P;
#pragma unroll
            for (int v = 0; v < CA; v++) {
                *reinterpret_cast<u32x4 *>(Ah + lda_[v]) = rah[slot][v];
                *reinterpret_cast<u32x4 *>(Al + lda_[v]) = ral[slot][v];
            }
#pragma unroll
            for (int v = 0; v < CB; v++) {
                *reinterpret_cast<u32x4 *>(Bh + ldb_[v]) = rbh[slot][v];
                *reinterpret_cast<u32x4 *>(Bl + ldb_[v]) = rbl[slot][v];
            }
        };
        auto multiply = [&](int buf) {
            const __bf16 *Ah = lds + buf * BUF, *Al = Ah + BM * P, *Bh = Al + BM * P, *Bl = Bh + BN * P;
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
        };
        load(std::integral_constant<int, 0>{}, 0);
        store(std::integral_constant<int, 0>{}, 0);
        static_for<PF>([&](auto u) { load(u, min((1 + u) * GK, last)); });
        __syncthreads();
        int s0 = 0;
        for (; s0 + PF <= steps; s0 += PF) {
            static_for<PF>([&](auto u) {
                const int st = s0 + u;
                multiply(st & 1);
                store(u, (st + 1) & 1);
                load(u, min((st + 1 + PF) * GK, last));
                __syncthreads();
            });
        }
        static_for<PF>([&](auto u) {
            const int st = s0 + u;
            if (st < steps) multiply(st & 1);
            store(u, (st + 1) & 1);
            __syncthreads();
        });
        if (g.scale) {
#pragma unroll
            for (int x = 0; x < TM; x++)
#pragma unroll
                for (int y = 0; y < TN; y++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int row = min(m0 + wm + x * 16 + kq * 4 + i, g.M - 1);
                        const int col = min(n0 + wn + y * 16 + l16, g.N - 1);
                        acc[x][y][i] *= g.scale[(size_t)row * g.ldc + col];
                    }
        }
#pragma unroll
        for (int x = 0; x < TM; x++)
#pragma unroll
            for (int y = 0; y < TN; y++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int row = m0 + wm + x * 16 + kq * 4 + i, col = n0 + wn + y * 16 + l16;
                    if (row < g.M && col < g.N) g.C[(size_t)row * g.ldc + col] = acc[x][y][i];
                }
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
    if (grid == 0 || K == 0) return hipSuccess;
    if (K % 32 != 0 || lda % 8 != 0 || ldb % 8 != 0) return hipErrorInvalidValue;
    if (t == 128) k_gemm_nt_bf16x3<128, 128, 2><<<grid, 256, 0, stream>>>(g);
    else k_gemm_nt_bf16x3<64, 64, 4><<<grid, 256, 0, stream>>>(g);
    return hipGetLastError();
}

}  // namespace emb
