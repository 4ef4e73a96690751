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
// S[k][N-1-n] = (-1)^k S[k][n], so S x splits into two half-size products:
// rows 2i of the result are se . p and rows 2i+1 are so . q, with the folded
// input p_n = x_n + x_{N-1-n}, q_n = x_n - x_{N-1-n} (n < N/2; p keeps the
// middle x_n alone when N is odd) and se, so the even and odd rows of S over
// the first half of its columns (sine_halves). The FP32 products fold while
// loading their data operand and write every other row or column: half the
// flops and half the K slices of the whole product.
//
// Two tile kernels compute C = A . B (row-major, optional elementwise scale):
//   * FP32: v_mfma_f32_16x16x4f32, FP32 in and out; what the tick runs, folded.
//     Four K slices are in flight per block, which is what the tick's small K
//     needs.
//   * BF16x3: each operand split as x = hi + lo (hi = bf16(x), lo = bf16(x - hi)),
//     A.B ~ lo.hi + hi.lo + hi.hi with v_mfma_f32_16x16x32_bf16 and FP32
//     accumulation (the lo.lo term is dropped). Offered for the isolated solve;
//     its error is measured against FP32 and FP64. The operands are split (and
//     transposed where the contraction runs over rows) by an O(n^2) pass before
//     each product, so the O(n^3) kernel only loads bf16 and multiplies.
// Both stage operands through LDS in K slices with the next slices prefetched
// into registers; the tick's products run 32 x 32 tiles on two waves when the
// batch has few tiles, else 64 x 64 tiles (the isolated solve's are larger) on
// a 2 x 2 grid of waves.

#include "device.hpp"
#include "gemm.hpp"

#include <algorithm>
#include <type_traits>
#include <utility>

namespace emb {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(1))) const float gfloat;
typedef __attribute__((address_space(1))) const floatx4 gfloatx4;

// Operand loads through global-memory pointers: global_load instructions, which
// count in vmcnt alone. Through a generic pointer they are flat_loads, which
// also count in lgkmcnt, so the wait before the next LDS read of the slice being
// multiplied would also wait for every prefetch still in flight.
__device__ __forceinline__ floatx4 gload4(const float *p) { return *(gfloatx4 *)p; }
__device__ __forceinline__ float gload1(const float *p) { return *(gfloat *)p; }

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

// MVEC (FAST + FOLD 2): the mirror columns of A are loaded as a float4 (fold_n
// a multiple of 4, so the mirror is aligned) or as four floats.
template <int BM, int BN, int GM = 2, int GN = 2, int GK = 16, int PF = 1, bool FAST = false, int FOLD = 0,
          bool MVEC = true>
__device__ __forceinline__ void tile_f32(const GemmArgs &g, int m0, int n0, float *lds) {
    constexpr int NT = 64 * GM * GN;
    constexpr int LA = BM + 4, LB = BN + 4, KV = GK / 4;
    constexpr int WM = BM / GM, WN = BN / GN, TM = WM / 16, TN = WN / 16;
    constexpr int AV = BM * GK / 4 / NT, BV = BN * GK / 4 / NT;
    static_assert(AV >= 1 && BV >= 1 && AV * NT * 4 == BM * GK && BV * NT * 4 == BN * GK, "tile and threads mismatch");
    // FAST + FOLD: the mirror of the folded operand is loaded into slots of its
    // own and added when the slice is stored (adding at load time would wait).
    constexpr int A2 = (FAST && FOLD == 2) ? AV : 1, B2 = (FAST && FOLD == 1) ? BV : 1;
    float *As = lds, *Bs = lds + 2 * GK * LA;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / GN) * WM, wn = (wave % GN) * WN;
    const int l16 = lane & 15, kq = lane >> 4;
    const bool avec = (g.lda & 3) == 0 && (reinterpret_cast<uintptr_t>(g.A) & 15) == 0;
    const bool bvec = (g.ldb & 3) == 0 && (reinterpret_cast<uintptr_t>(g.B) & 15) == 0;
    const float sgn = g.odd ? -1.0f : 1.0f;
    const int fn = g.fold_n;
    floatx4 acc[TM][TN];
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
        for (int y = 0; y < TN; y++) acc[x][y] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // Per-thread operand addresses at slice 0 (FAST: clamped, so every load is valid).
    const float *pa[AV], *pa2[A2];
    const float *pb[BV], *pb2[B2];
    int ra_row[AV], ra_k[AV], rb_k[BV], rb_col[BV];
#pragma unroll
    for (int v = 0; v < AV; v++) {
        const int idx = tid + NT * v;
        ra_row[v] = idx / KV;
        ra_k[v] = (idx % KV) * 4;
        const int gr = FAST ? min(m0 + ra_row[v], g.M - 1) : m0 + ra_row[v];
        pa[v] = g.A + (size_t)gr * g.lda + ra_k[v];
        if (FAST && FOLD == 2) pa2[v] = g.A + (size_t)gr * g.lda + (fn - 4 - ra_k[v]);   // moves left as k grows
    }
#pragma unroll
    for (int v = 0; v < BV; v++) {
        const int idx = tid + NT * v;
        rb_k[v] = idx / (BN / 4);
        rb_col[v] = (idx % (BN / 4)) * 4;
        const int gc = FAST ? min(n0 + rb_col[v], (g.Np ? g.Np : g.N) - 4) : n0 + rb_col[v];
        pb[v] = g.B + (size_t)rb_k[v] * g.ldb + gc;
        if (FAST && FOLD == 1) pb2[v] = g.B + (size_t)(fn - 1 - rb_k[v]) * g.ldb + gc;   // moves up as k grows
    }
    // Native vectors, not float4 structs: a struct copy between address spaces
    // goes through private memory.
    floatx4 ra[PF][AV], rb[PF][BV], ra2[PF][A2], rb2[PF][B2];
    auto vec4 = [](float4 f) { return floatx4{f.x, f.y, f.z, f.w}; };
    // Element e of a folded A row at column c (< K): A[r][c] + s A[r][fn - 1 - c].
    auto fold_a = [&](const float *row, int c) {
        const int c2 = fn - 1 - c;
        return row[c] + (c2 == c ? 0.0f : sgn * row[c2]);
    };
    auto load = [&](auto slot, int k0) {
#pragma unroll
        for (int v = 0; v < AV; v++) {
            if (FAST) {
                ra[slot][v] = gload4(pa[v] + k0);
                if (FOLD == 2) {
                    if (MVEC) {
                        ra2[slot][v] = gload4(pa2[v] - k0);
                    } else {
                        const float *m = pa2[v] - k0;
                        ra2[slot][v] = floatx4{gload1(m), gload1(m + 1), gload1(m + 2), gload1(m + 3)};
                    }
                }
            } else {
                const int gr = m0 + ra_row[v], gk = k0 + ra_k[v];
                if (FOLD == 2) {
                    floatx4 r = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                    if (gr < g.M) {
                        const float *row = g.A + (size_t)gr * g.lda;
#pragma unroll
                        for (int e = 0; e < 4; e++)
                            if (gk + e < g.K) r[e] = fold_a(row, gk + e);
                    }
                    ra[slot][v] = r;
                } else {
                    ra[slot][v] = gr < g.M ? vec4(load4(pa[v] + k0, g.K - gk, avec)) : floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                }
            }
        }
#pragma unroll
        for (int v = 0; v < BV; v++) {
            if (FAST) {
                rb[slot][v] = gload4(pb[v] + (size_t)k0 * g.ldb);
                if (FOLD == 1) rb2[slot][v] = gload4(pb2[v] - (size_t)k0 * g.ldb);
            } else {
                const int gk = k0 + rb_k[v], gc = n0 + rb_col[v];
                floatx4 r = gk < g.K ? vec4(load4(pb[v] + (size_t)k0 * g.ldb, g.N - gc, bvec))
                                     : floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                if (FOLD == 1 && gk < g.K && fn - 1 - gk != gk)
                    r += sgn * vec4(load4(g.B + (size_t)(fn - 1 - gk) * g.ldb + gc, g.N - gc, bvec));
                rb[slot][v] = r;
            }
        }
    };
    auto store = [&](auto slot, int buf, int k0) {
        float *a = As + buf * GK * LA, *b = Bs + buf * GK * LB;
#pragma unroll
        for (int v = 0; v < AV; v++) {
            const int row = ra_row[v], kk = ra_k[v];
            floatx4 x = ra[slot][v];
            if (FAST && FOLD == 2) {   // the mirror float4 runs backwards
                const floatx4 y = ra2[slot][v];
                if (MVEC) {   // fold_n even: no column is its own mirror
                    x = x + sgn * floatx4{y[3], y[2], y[1], y[0]};
                } else {      // the middle column of an odd fold_n is taken once
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int c = k0 + kk + e;
                        x[e] += (fn - 1 - c == c) ? 0.0f : sgn * y[3 - e];
                    }
                }
            }
            a[(kk + 0) * LA + row] = x[0];
            a[(kk + 1) * LA + row] = x[1];
            a[(kk + 2) * LA + row] = x[2];
            a[(kk + 3) * LA + row] = x[3];
        }
#pragma unroll
        for (int v = 0; v < BV; v++) {
            floatx4 x = rb[slot][v];
            if (FAST && FOLD == 1) {
                const int r = k0 + rb_k[v];
                x = x + ((fn - 1 - r == r) ? 0.0f : sgn) * rb2[slot][v];
            }
            *reinterpret_cast<floatx4 *>(b + rb_k[v] * LB + rb_col[v]) = x;
        }
    };
    auto multiply = [&](int buf) {
        const float *a = As + buf * GK * LA, *b = Bs + buf * GK * LB;
        // every fragment of the slice read before the first MFMA: one wait for
        // the LDS, not one per step of 4
        float af[GK / 4][TM], bf[GK / 4][TN];
#pragma unroll
        for (int q = 0; q < GK / 4; q++) {
#pragma unroll
            for (int x = 0; x < TM; x++) af[q][x] = a[(4 * q + kq) * LA + wm + x * 16 + l16];
#pragma unroll
            for (int y = 0; y < TN; y++) bf[q][y] = b[(4 * q + kq) * LB + wn + y * 16 + l16];
        }
#pragma unroll
        for (int q = 0; q < GK / 4; q++)
#pragma unroll
            for (int x = 0; x < TM; x++)
#pragma unroll
                for (int y = 0; y < TN; y++)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[q][x], bf[q][y], acc[x][y], 0, 0, 0);
    };
    // Slices past the last are clamped to it, so every load and store runs
    // unconditionally in the unrolled body (a repeated slice is never multiplied).
    const int steps = (g.K + GK - 1) / GK, last = (steps - 1) * GK;
    load(std::integral_constant<int, 0>{}, 0);
    store(std::integral_constant<int, 0>{}, 0, 0);
    static_for<PF>([&](auto u) { load(u, min((1 + u) * GK, last)); });
    __syncthreads();
    int s0 = 0;
    for (; s0 + PF <= steps; s0 += PF) {
        static_for<PF>([&](auto u) {
            const int st = s0 + u;
            multiply(st & 1);
            store(u, (st + 1) & 1, min((st + 1) * GK, last));
            load(u, min((st + 1 + PF) * GK, last));
            __syncthreads();
        });
    }
    // Tail: only the multiply is conditional (a branch around the register
    // slots would make their index dynamic); stores past the end are unread.
    static_for<PF>([&](auto u) {
        const int st = s0 + u;
        if (st < steps) multiply(st & 1);
        store(u, (st + 1) & 1, min((st + 1) * GK, last));
        __syncthreads();
    });
    // D[row][col]: lane holds rows 4 * (lane / 16) + i of column lane % 16; a
    // folded half writes every other row (FOLD 1) or column (FOLD 2) of C.
    auto crow = [&](int row) { return FOLD == 1 ? 2 * row + g.odd : row; };
    auto ccol = [&](int col) { return FOLD == 2 ? 2 * col + g.odd : col; };
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
                    acc[x][y][i] *= g.scale[(size_t)crow(row) * g.ldc + ccol(col)];
                }
    }
#pragma unroll
    for (int x = 0; x < TM; x++)
#pragma unroll
        for (int y = 0; y < TN; y++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = m0 + wm + x * 16 + kq * 4 + i, col = n0 + wn + y * 16 + l16;
                if (row < g.M && col < g.N) g.C[(size_t)crow(row) * g.ldc + ccol(col)] = acc[x][y][i];
            }
}

// Whether a product can take tile_f32's FAST path for slices of GK: 0 no,
// 1 yes, 2 yes with the mirror columns of a column fold loaded as single
// floats (fold_n not a multiple of 4, so the mirror float4 is not aligned).
__host__ __device__ inline int gemm_fast(const GemmArgs &g, int gk) {
    const int np = g.Np ? g.Np : g.N;
    const bool fast = g.K % gk == 0 && np % 4 == 0 && np >= 4 && g.lda % 4 == 0 && g.ldb % 4 == 0 &&
                      (reinterpret_cast<uintptr_t>(g.A) & 15) == 0 && (reinterpret_cast<uintptr_t>(g.B) & 15) == 0;
    if (!fast) return 0;
    return (g.fold != 2 || g.fold_n % 4 == 0) ? 1 : 2;
}


// ---------------------------------------------------------------------------------------
// The tick's solve over a batch of scenes: blocks walk (scene, channel, half,
// tile). Each product runs as its two folded halves (sine_halves): half the
// flops and half the K slices of the whole product.

constexpr int TICK_BM = 64, TICK_BN = 64;

// Output tiles of the even and odd halves of one stage's product, BM x BN each.
__host__ __device__ inline void dst_half_tiles(bool left, int nx, int ny, int bm, int bn, int &te, int &to) {
    auto cdiv = [](int a, int b) { return (a + b - 1) / b; };
    if (left) {   // rows of Sx: ceil(nx/2) even, floor(nx/2) odd
        te = cdiv((nx + 1) / 2, bm) * cdiv(ny, bn);
        to = cdiv(nx / 2, bm) * cdiv(ny, bn);
    } else {      // columns of Sy
        te = cdiv(nx, bm) * cdiv((ny + 1) / 2, bn);
        to = cdiv(nx, bm) * cdiv(ny / 2, bn);
    }
}

__host__ __device__ inline int dst_blocks(int nx, int ny, int bm, int bn) {
    int le, lo, re, ro;
    dst_half_tiles(true, nx, ny, bm, bn, le, lo);
    dst_half_tiles(false, nx, ny, bm, bn, re, ro);
    return 2 * (le + lo > re + ro ? le + lo : re + ro);
}

// A batch whose stages have at most DST_SMALL_TILES output tiles of 64 x 64 (both
// channels, over all its scenes) does not fill the chip's 256 CUs with them:
// its products run SMALL_BM x SMALL_BN tiles on blocks of SMALL_GM x SMALL_GN
// waves instead (KC_GEMM_SMALL). Every element's sum over K runs in the same
// order either way, so the results have the same bits. At 544 nodes (one
// scene) the four products took 51 us against 54.5 us with 64 x 64 tiles;
// 32 x 64 tiles (69 us) and 64 x 64 tiles with K split over the block's four
// waves and the partial tiles summed in LDS (72 us) were slower. A full batch
// (the 256-request serving stream) keeps the 64 x 64 tiles: small tiles for its
// small scenes, in a launch of their own, took its throughput from 2.87 to 2.52
// G cell-steps/s.
constexpr int DST_SMALL_TILES = 512;
constexpr int SMALL_BM = 32, SMALL_BN = 32, SMALL_GM = 2, SMALL_GN = 1;

int dst_gemm_blocks(int nx, int ny) { return dst_blocks(nx, ny, TICK_BM, TICK_BN); }
int dst_gemm_small_blocks(int nx, int ny) { return dst_blocks(nx, ny, SMALL_BM, SMALL_BN); }

// Block `local` of a scene's stage STAGE: its product's arguments and output
// tile origin, or false when the block has nothing to do.
template <int STAGE, int BM, int BN>
__device__ __forceinline__ bool dst_tile(const DevScene &S, int local, int chans, GemmArgs &g, int &m0, int &n0) {
    constexpr bool LEFT = (STAGE == 0 || STAGE == 2);
    const int nx = S.nx, ny = S.ny;
    int te, to;
    dst_half_tiles(LEFT, nx, ny, BM, BN, te, to);
    const int ch = local / (te + to);
    if (ch >= 2) return false;   // the scene's block count covers the larger stage
    // Channel 0 (the charge's potential) only where it is wanted: every
    // tick in a scene with battery stacks (the trim reads it), otherwise on
    // demand (DST_POT) when the potential is read.
    if (chans == DST_POT ? ch != 0 : (chans == DST_TICK && ch == 0 && S.n_stacks == 0)) return false;
    const int rem0 = local - ch * (te + to);
    const int odd = rem0 >= te ? 1 : 0;
    const int rem = rem0 - odd * te;
    // The Poisson arrays' rows are padded to pny (a multiple of 4) and the
    // sine halves to whole slices, so every scene takes the FAST tiles.
    const int pny = S.pny;
    const size_t pnodes = (size_t)nx * pny;
    const float *in = (LEFT ? S.w1 : S.w2) + ch * pnodes;
    float *out = (STAGE == 3 ? S.phi : (STAGE == 1 ? S.w1 : S.w2)) + ch * pnodes;
    g = GemmArgs{};
    g.C = out;
    g.ldc = pny;
    g.scale = STAGE == 1 ? S.inv_lap : nullptr;
    g.odd = odd;
    if (LEFT) {   // out = Sx . in: rows 2 i + odd from the half of Sx and the folded rows of in
        const int mh = odd ? nx / 2 : (nx + 1) / 2;
        g.A = odd ? S.sx_o : S.sx_e;
        g.lda = sine_half_pitch(mh);
        g.B = in;
        g.ldb = pny;
        g.M = mh;
        g.N = ny;
        g.Np = pny;
        g.K = fold_k(mh, nx);
        g.fold = 1;
        g.fold_n = nx;
    } else {      // out = in . Sy: columns 2 j + odd from the folded columns of in
        const int mh = odd ? ny / 2 : (ny + 1) / 2;
        g.A = in;
        g.lda = pny;
        g.B = odd ? S.sy_o : S.sy_et;
        g.ldb = sine_half_pitch(mh);
        g.M = nx;
        g.N = mh;
        g.Np = sine_half_pitch(mh);
        g.K = fold_k(mh, ny);
        g.fold = 2;
        g.fold_n = ny;
    }
    const int tn = (g.N + BN - 1) / BN;
    m0 = (rem / tn) * BM;
    n0 = (rem % tn) * BN;
    return true;
}

template <int STAGE>
__global__ __launch_bounds__(256) void k_dst_gemm(Batch b, int chans) {
    constexpr int F = (STAGE == 0 || STAGE == 2) ? 1 : 2;
    __shared__ __align__(16) float lds[2 * 16 * (TICK_BM + 4 + TICK_BN + 4)];
    for (int blk = blockIdx.x; blk < b.total[KC_GEMM]; blk += gridDim.x) {
        const int s = scene_of(b, KC_GEMM, blk);
        GemmArgs g;
        int m0, n0;
        if (!dst_tile<STAGE, TICK_BM, TICK_BN>(b.scenes[s], block_in_scene(b, KC_GEMM, s, blk), chans, g, m0, n0)) continue;
        const int fast = gemm_fast(g, 16);
        if (fast == 1) tile_f32<TICK_BM, TICK_BN, 2, 2, 16, 4, true, F>(g, m0, n0, lds);
        else if (F == 2 && fast == 2) tile_f32<TICK_BM, TICK_BN, 2, 2, 16, 4, true, F, false>(g, m0, n0, lds);
        else tile_f32<TICK_BM, TICK_BN, 2, 2, 16, 4, false, F>(g, m0, n0, lds);
        __syncthreads();
    }
}

// A batch with few 64 x 64 tiles: SMALL_BM x SMALL_BN tiles.
constexpr int SMALL_THREADS = 64 * SMALL_GM * SMALL_GN;
template <int STAGE>
__global__ __launch_bounds__(SMALL_THREADS) void k_dst_gemm_small(Batch b, int chans) {
    constexpr int F = (STAGE == 0 || STAGE == 2) ? 1 : 2;
    constexpr int GM = SMALL_GM, GN = SMALL_GN;
    __shared__ __align__(16) float lds[2 * 16 * (SMALL_BM + 4 + SMALL_BN + 4)];
    for (int blk = blockIdx.x; blk < b.total[KC_GEMM_SMALL]; blk += gridDim.x) {
        const int s = scene_of(b, KC_GEMM_SMALL, blk);
        GemmArgs g;
        int m0, n0;
        if (!dst_tile<STAGE, SMALL_BM, SMALL_BN>(b.scenes[s], block_in_scene(b, KC_GEMM_SMALL, s, blk), chans, g, m0,
                                                 n0))
            continue;
        const int fast = gemm_fast(g, 16);
        if (fast == 1) tile_f32<SMALL_BM, SMALL_BN, GM, GN, 16, 4, true, F>(g, m0, n0, lds);
        else if (F == 2 && fast == 2) tile_f32<SMALL_BM, SMALL_BN, GM, GN, 16, 4, true, F, false>(g, m0, n0, lds);
        else tile_f32<SMALL_BM, SMALL_BN, GM, GN, 16, 4, false, F>(g, m0, n0, lds);
        __syncthreads();
    }
}

template __global__ void k_dst_gemm<0>(Batch, int);
template __global__ void k_dst_gemm<1>(Batch, int);
template __global__ void k_dst_gemm<2>(Batch, int);
template __global__ void k_dst_gemm<3>(Batch, int);
template __global__ void k_dst_gemm_small<0>(Batch, int);
template __global__ void k_dst_gemm_small<1>(Batch, int);
template __global__ void k_dst_gemm_small<2>(Batch, int);
template __global__ void k_dst_gemm_small<3>(Batch, int);

// The four products of one Poisson solve for every scene of the batch, one
// launch per stage, in stage order: small tiles when the batch's 64 x 64 tiles
// would not fill the chip.
hipError_t launch_dst_gemms(const Batch &b, hipStream_t stream, int chans) {
    if (b.total[KC_GEMM] == 0) return hipSuccess;
    const bool small = b.total[KC_GEMM] <= DST_SMALL_TILES;
    const int whole = small ? 0 : std::min(b.total[KC_GEMM], 8192);
    const int split = small ? std::min(b.total[KC_GEMM_SMALL], 8192) : 0;
#define EMB_DST_STAGE(st)                                                                         \
    if (whole) k_dst_gemm<st><<<whole, 256, 0, stream>>>(b, chans);                               \
    if (split) k_dst_gemm_small<st><<<split, SMALL_THREADS, 0, stream>>>(b, chans);
    EMB_DST_STAGE(0)
    EMB_DST_STAGE(1)
    EMB_DST_STAGE(2)
    EMB_DST_STAGE(3)
#undef EMB_DST_STAGE
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// One product on plain pointers (the isolated solve).

template <int BM, int BN, int PF, bool FAST, int FOLD, bool MVEC = true>
__global__ __launch_bounds__(256) void k_gemm_f32(GemmArgs g0, GemmArgs g1) {
    __shared__ __align__(16) float lds[2 * 16 * (BM + 4 + BN + 4)];
    const int tn0 = (g0.N + BN - 1) / BN, tiles0 = ((g0.M + BM - 1) / BM) * tn0;
    const int tn1 = (g1.N + BN - 1) / BN, tiles1 = ((g1.M + BM - 1) / BM) * tn1;
    for (int blk = blockIdx.x; blk < tiles0 + tiles1; blk += gridDim.x) {
        const bool second = blk >= tiles0;
        const GemmArgs &g = second ? g1 : g0;
        const int t = second ? blk - tiles0 : blk, tn = second ? tn1 : tn0;
        tile_f32<BM, BN, 2, 2, 16, PF, FAST, FOLD, MVEC>(g, (t / tn) * BM, (t % tn) * BN, lds);
        __syncthreads();
    }
}

// Products g0 and g1 (g1.M = 0: none) in one launch, all FAST or none.
template <int BM, int BN, int FOLD>
static hipError_t launch_gemm_tiles(const GemmArgs &g0, const GemmArgs &g1, hipStream_t stream) {
    auto tiles = [](const GemmArgs &g) {
        return g.M > 0 && g.N > 0 ? (long long)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN) : 0LL;
    };
    const long long total = tiles(g0) + tiles(g1);
    const int grid = (int)(total < 8192 ? total : 8192);
    if (grid == 0) return hipSuccess;
    const int f0 = gemm_fast(g0, 16), f1 = g1.M == 0 ? 1 : gemm_fast(g1, 16);
    if (f0 == 0 || f1 == 0) k_gemm_f32<BM, BN, 4, false, FOLD><<<grid, 256, 0, stream>>>(g0, g1);
    else if (FOLD == 2 && (f0 == 2 || f1 == 2)) k_gemm_f32<BM, BN, 4, true, FOLD, false><<<grid, 256, 0, stream>>>(g0, g1);
    else k_gemm_f32<BM, BN, 4, true, FOLD><<<grid, 256, 0, stream>>>(g0, g1);
    return hipGetLastError();
}

hipError_t launch_gemm(const GemmArgs &g, int tile, hipStream_t stream) {
    GemmArgs none;
    return tile == 128 ? launch_gemm_tiles<128, 64, 0>(g, none, stream) : launch_gemm_tiles<64, 64, 0>(g, none, stream);
}

hipError_t launch_gemm_fold(const GemmArgs &even, const GemmArgs &odd, int tile, hipStream_t stream) {
    if (even.fold != odd.fold || (even.fold != 1 && even.fold != 2)) return hipErrorInvalidValue;
    if (even.fold == 1)
        return tile == 128 ? launch_gemm_tiles<128, 64, 1>(even, odd, stream) : launch_gemm_tiles<64, 64, 1>(even, odd, stream);
    return tile == 128 ? launch_gemm_tiles<128, 64, 2>(even, odd, stream) : launch_gemm_tiles<64, 64, 2>(even, odd, stream);
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
                rah[slot][v] = *(__attribute__((address_space(1))) const u32x4 *)(g.ah + offa[v] + k0);
                ral[slot][v] = *(__attribute__((address_space(1))) const u32x4 *)(g.al + offa[v] + k0);
            }
#pragma unroll
            for (int v = 0; v < CB; v++) {
                rbh[slot][v] = *(__attribute__((address_space(1))) const u32x4 *)(g.bh + offb[v] + k0);
                rbl[slot][v] = *(__attribute__((address_space(1))) const u32x4 *)(g.bl + offb[v] + k0);
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
