// tables.hip - the close-range tables built on the device, in FP64
// (close_range::tables, crates/embloks-core/src/close_range.rs:67-242).
//
// The dominant cost is grid_force (close_range.rs:73-108): 8 x 8 source
// positions, each a unit charge deposited, smoothed, solved by the DST-I
// Poisson product phi = S ((S rhs S) * L) S on a small node grid, and its field
// read back at 16 directions x (cutoff / 0.01 + 1) radii. One block per source
// position does all of that with its matrices in global memory (they stay in
// L2); the radial tables, the square table (midpoint-rule box sums through the
// summed-area table, whose recurrence runs as an anti-diagonal wavefront), the
// ion stencil and the summed ion table follow as small kernels. Every sum runs in
// the reference's order and this file is compiled without FP contraction, so
// the tables agree with an FP64 restatement of the reference to the last bits
// of the transcendental functions (sin, cos, exp).

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <memory>
#include <vector>

#include "runtime.hpp"

namespace emb {

// Kernels and helpers of this file only (k_tab_*).
constexpr double TPI = 3.14159265358979323846;
constexpr int T_SQUARE = 32;       // SQUARE_SAMPLES, close_range.rs:32
constexpr int T_SOURCES = 8;       // SOURCE_SAMPLES, close_range.rs:35
constexpr int T_DIRS = 16;         // DIRECTIONS, close_range.rs:36
constexpr double T_STEP = 0.01;    // RADIAL_STEP, close_range.rs:37
constexpr int T_MARGIN = 16;       // Settings::default().absorbing_margin
constexpr int T_THREADS = 256;
constexpr double MAX_CUTOFF = 32.0;

struct TabGeom {
    int n, N, P, passes, count;   // grid cells, nodes with margin, padded smoothing grid, passes, radii
    double reach, softening;
    const double *S, *L;          // N x N DST-I matrix and scaled inverse Laplacian
    double *work;                 // per source block: pad, nxt (P^2), A, B (N^2), ex ((N+1) N), ey (N (N+1))
    size_t per_block;
    double *part;                 // T_SOURCES^2 x count: each source's radial force / DIRECTIONS
};

static __device__ double compact_d(double r2, double eps) {   // close_range.rs:41-49
    const double x = r2 / (eps * eps);
    const double v = x < 1e-4 ? 2.0 * (1.0 - x / 2.0 + x * x / 6.0) : -2.0 * expm1(-x) / x;
    return v / (eps * eps);
}

// numpy.interp over ascending xs (close_range.rs:56-65).
static __device__ double interp_d(double x, const double *xs, const double *ys, int len) {
    if (x <= xs[0]) return ys[0];
    if (x >= xs[len - 1]) return ys[len - 1];
    int lo = 0, hi = len;   // first index with xs > x
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xs[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    const int i = lo - 1;
    return ys[i] + (ys[i + 1] - ys[i]) * (x - xs[i]) / (xs[i + 1] - xs[i]);
}

static __device__ double radial_d(const double *table, double r2, double cutoff) {   // close_range.rs:174-181
    if (r2 >= cutoff * cutoff) return 0.0;
    const double x = fmin(r2 * ((double)PAIR_SAMPLES / (cutoff * cutoff)), (double)PAIR_SAMPLES);
    const int i = (int)x;
    return table[i] + (table[i + 1] - table[i]) * (x - (double)i);
}

__global__ void k_tab_matrices(int N, double *S, double *L) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N * N) return;
    const int k = q / N, j = q % N;
    // sin(pi (k+1)(j+1) / (N+1)) with the argument reduced modulo 2 (N + 1) first (maxwell.rs:56-67).
    S[q] = sin(TPI * (double)(((long long)(k + 1) * (j + 1)) % (2 * (N + 1))) / (double)(N + 1));
    const double li = 2.0 * cos(TPI * (double)(k + 1) / (double)(N + 1)) - 2.0;
    const double lj = 2.0 * cos(TPI * (double)(j + 1) / (double)(N + 1)) - 2.0;
    L[q] = (4.0 / (double)((N + 1) * (N + 1))) / (li + lj);
}

// Z = X Y for N x N row-major matrices, each entry summed over k in order.
static __device__ void block_matmul(int N, const double *X, const double *Y, double *Z) {
    for (int q = threadIdx.x; q < N * N; q += T_THREADS) {
        const int i = q / N, j = q % N;
        double s = 0.0;
        for (int k = 0; k < N; k++) s += X[i * N + k] * Y[k * N + j];
        Z[q] = s;
    }
    __syncthreads();
}

static __device__ double bilin_d(const double *a, int w, int h, double gx, double gy) {   // Maxwell::gather's lerp
    const long long i = min(max((long long)floor(gx), 0LL), (long long)w - 2);
    const long long j = min(max((long long)floor(gy), 0LL), (long long)h - 2);
    const double fx = gx - (double)i, fy = gy - (double)j;
    return a[i * h + j] * (1.0 - fx) * (1.0 - fy) + a[(i + 1) * h + j] * fx * (1.0 - fy) +
           a[i * h + j + 1] * (1.0 - fx) * fy + a[(i + 1) * h + j + 1] * fx * fy;
}

// One source position per block: deposit, smooth, solve, gather (close_range.rs:88-104).
__global__ __launch_bounds__(T_THREADS) void k_tab_source(TabGeom g) {
    const int src = blockIdx.x, si = src / T_SOURCES, sj = src % T_SOURCES;
    const int n = g.n, N = g.N, P = g.P, kp = g.passes, m = T_MARGIN;
    double *pad = g.work + (size_t)src * g.per_block;
    double *nxt = pad + (size_t)P * P;
    double *A = nxt + (size_t)P * P;
    double *B = A + (size_t)N * N;
    double *ex = B + (size_t)N * N;
    double *ey = ex + (size_t)(N + 1) * N;
    const double sx = (double)(n / 2) + ((double)si + 0.5) / T_SOURCES;
    const double sy = (double)(n / 2) + ((double)sj + 0.5) / T_SOURCES;
    for (int q = threadIdx.x; q < P * P; q += T_THREADS) pad[q] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {   // CloudInCell of the unit charge (fields.rs:27-51)
        const double u = sx - 0.5, v = sy - 0.5;
        const long long c0 = (long long)floor(u), r0 = (long long)floor(v);
        const double fu = u - (double)c0, fv = v - (double)r0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                const long long c = min(max(c0 + a, 0LL), (long long)n - 1), r = min(max(r0 + b, 0LL), (long long)n - 1);
                pad[(c + kp) * P + (r + kp)] += (a ? fu : 1.0 - fu) * (b ? fv : 1.0 - fv);
            }
    }
    __syncthreads();
    // `passes` passes of 1-2-1, x then y, zero beyond the padding (maxwell.rs:397-414).
    for (int p = 0; p < kp; p++)
        for (int axis = 0; axis < 2; axis++) {
            const int di = axis == 0 ? 1 : 0, dj = 1 - di;
            for (int q = threadIdx.x; q < P * P; q += T_THREADS) {
                const int i = q / P, j = q % P, along = axis == 0 ? i : j;
                double s = 0.5 * pad[q];
                if (along > 0) s += 0.25 * pad[(i - di) * P + (j - dj)];
                if (along + 1 < P) s += 0.25 * pad[(i + di) * P + (j + dj)];
                nxt[q] = s;
            }
            __syncthreads();
            double *t = pad;
            pad = nxt;
            nxt = t;
        }
    // The right-hand side -4 pi rho on the N x N nodes, the smoothed grid at offset m - passes.
    for (int q = threadIdx.x; q < N * N; q += T_THREADS) {
        const int i = q / N - (m - kp), j = q % N - (m - kp);
        A[q] = (i >= 0 && i < P && j >= 0 && j < P) ? -4.0 * TPI * pad[i * P + j] : 0.0;
    }
    __syncthreads();
    block_matmul(N, g.S, A, B);
    block_matmul(N, B, g.S, A);
    for (int q = threadIdx.x; q < N * N; q += T_THREADS) A[q] *= g.L[q];
    __syncthreads();
    block_matmul(N, g.S, A, B);
    block_matmul(N, B, g.S, A);
    const double *phi = A;
    // E = -grad phi, phi = 0 outside the node grid (maxwell.rs:181-204).
    for (int q = threadIdx.x; q < (N + 1) * N; q += T_THREADS) {
        const int i = q / N, j = q % N;
        ex[q] = -((i < N ? phi[i * N + j] : 0.0) - (i > 0 ? phi[(i - 1) * N + j] : 0.0));
    }
    for (int q = threadIdx.x; q < N * (N + 1); q += T_THREADS) {
        const int i = q / (N + 1), j = q % (N + 1);
        ey[q] = -((j < N ? phi[i * N + j] : 0.0) - (j > 0 ? phi[i * N + j - 1] : 0.0));
    }
    __syncthreads();
    for (int k = threadIdx.x; k < g.count; k += T_THREADS) {
        const double r = 0.0 + (g.reach - 0.0) * (double)k / (double)(g.count - 1);
        double along = 0.0;
        for (int d = 0; d < T_DIRS; d++) {
            const double ang = 2.0 * TPI * ((double)d + 0.5) / (double)T_DIRS;
            const double ux = cos(ang), uy = sin(ang);
            const double x = sx + r * ux, y = sy + r * uy;
            const double fx = bilin_d(ex, N + 1, N, x + m, y - 0.5 + m);
            const double fy = bilin_d(ey, N, N + 1, x - 0.5 + m, y + m);
            along += fx * ux + fy * uy;
        }
        g.part[(size_t)src * g.count + k] = along / T_DIRS;
    }
}

struct RadialArgs {
    int count;
    double reach, softening, cutoff;
    const double *part;
    double *r, *f, *near, *pot;   // count each
    double *g, *potential;        // PAIR_SAMPLES + 2 each
};

// The source average and the radial tables (close_range.rs:108, 151-171); one block.
__global__ __launch_bounds__(T_THREADS) void k_tab_radial(RadialArgs a) {
    const int count = a.count;
    for (int k = threadIdx.x; k < count; k += T_THREADS) {
        double t = 0.0;
        for (int s = 0; s < T_SOURCES * T_SOURCES; s++) t += a.part[(size_t)s * count + k];
        a.r[k] = 0.0 + (a.reach - 0.0) * (double)k / (double)(count - 1);
        a.f[k] = t / (double)(T_SOURCES * T_SOURCES);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < count; k += T_THREADS) a.near[k] = a.r[k] * compact_d(a.r[k] * a.r[k], a.softening) - a.f[k];
    __syncthreads();
    if (threadIdx.x == 0) {   // the potential: integral of the correction inwards from the cutoff
        a.pot[count - 1] = 0.0;
        for (int i = count - 2; i >= 0; i--)
            a.pot[i] = a.pot[i + 1] + 0.5 * (a.near[i + 1] + a.near[i]) * (a.r[i + 1] - a.r[i]);
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= PAIR_SAMPLES + 1; i += T_THREADS) {
        if (i == PAIR_SAMPLES + 1) {
            a.g[i] = 0.0;
            a.potential[i] = 0.0;
            continue;
        }
        const double r2 = 0.0 + (a.cutoff * a.cutoff - 0.0) * (double)i / (double)PAIR_SAMPLES;
        const double rs = sqrt(r2);
        const double grid_g = i == 0 ? a.f[1] / a.r[1] : interp_d(rs, a.r, a.f, count) / rs;
        a.g[i] = compact_d(r2, a.softening) - grid_g;
        a.potential[i] = interp_d(rs, a.r, a.pot, count);
    }
}

struct SquareArgs {
    int cnt, n;
    double half, cutoff;
    const double *g, *potential;
    double *pts;      // cnt^2 x 3 point corrections
    double *sat;      // (cnt + 1)^2 x 3, leading row and column zero
    double *square;   // n^2 x 3
};

static __device__ double z_at(const SquareArgs &a, int k) {
    return -a.half - 0.5 + ((double)k + 0.5) * (1.0 / T_SQUARE);
}

// point_correction at every midpoint sample (close_range.rs:184-189, 201-211).
__global__ void k_tab_points(SquareArgs a) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)a.cnt * a.cnt) return;
    const int ia = (int)(q / a.cnt), ib = (int)(q % a.cnt);
    const double za = z_at(a, ia), zb = z_at(a, ib);
    const double r2 = za * za + zb * zb;
    const double gr = radial_d(a.g, r2, a.cutoff);
    a.pts[q * 3 + 0] = radial_d(a.potential, r2, a.cutoff);
    a.pts[q * 3 + 1] = gr * za;
    a.pts[q * 3 + 2] = gr * zb;
}

// The summed-area recurrence of close_range.rs:205-213 as an anti-diagonal
// wavefront: sample (ia, ib) waits only for (ia - 1, ib), (ia, ib - 1) and
// (ia - 1, ib - 1), all on earlier diagonals. One block; the barrier orders the
// diagonals.
__global__ __launch_bounds__(1024) void k_tab_sat(SquareArgs a) {
    const int cnt = a.cnt, W = cnt + 1;
    for (int s = 0; s <= 2 * (cnt - 1); s++) {
        for (int ia = threadIdx.x; ia < cnt; ia += 1024) {
            const int ib = s - ia;
            if (ib < 0 || ib >= cnt) continue;
            const double *pt = a.pts + ((size_t)ia * cnt + ib) * 3;
            for (int k = 0; k < 3; k++)
                a.sat[((size_t)(ia + 1) * W + ib + 1) * 3 + k] =
                    pt[k] + a.sat[((size_t)ia * W + ib + 1) * 3 + k] + a.sat[((size_t)(ia + 1) * W + ib) * 3 + k] -
                    a.sat[((size_t)ia * W + ib) * 3 + k];
        }
        __syncthreads();
    }
}

__global__ void k_tab_box(SquareArgs a) {   // close_range.rs:214-218
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.n, W = a.cnt + 1, M = T_SQUARE;
    if (q >= (long long)n * n * 3) return;
    const int k = (int)(q % 3), ij = (int)(q / 3), i = ij / n, j = ij % n;
    const double h = 1.0 / M;
    a.square[q] = (a.sat[((size_t)(i + M) * W + j + M) * 3 + k] - a.sat[((size_t)i * W + j + M) * 3 + k] -
                   a.sat[((size_t)(i + M) * W + j) * 3 + k] + a.sat[((size_t)i * W + j) * 3 + k]) * h * h;
}

struct IonArgs {
    int n, reach, a0;
    double half;
    const double *square;
    double *stencil;   // (2 reach + 1)^2
    double *ion_sum;   // ION_SUM^2 x 2
};

// ion_stencil (close_range.rs:223-242) and the force channels summed over the
// reach window at each fractional offset (the summed ion table).
__global__ void k_tab_ion(IonArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int R = a.reach, side = 2 * R + 1, n = a.n, M = T_SQUARE;
    if (q < side * side) {
        const int i = q / side, j = q % side;
        const long long ia = llround(((double)(i - R) - 0.5 + a.half) * M);
        const long long ib = llround(((double)(j - R) - 0.5 + a.half) * M);
        double sum = 0.0;
        for (int x = 0; x <= M; x++)
            for (int y = 0; y <= M; y++) {
                const double wx = (x == 0 || x == M) ? 0.5 / M : 1.0 / M;
                const double wy = (y == 0 || y == M) ? 0.5 / M : 1.0 / M;
                sum += wx * a.square[((ia + x) * n + (ib + y)) * 3] * wy;
            }
        a.stencil[q] = sum;
        return;
    }
    const int r = q - side * side;
    if (r >= ION_SUM * ION_SUM * 2) return;
    const int ch = r % 2, ij = r / 2, i = ij / ION_SUM, j = ij % ION_SUM;
    double sum = 0.0;
    for (int ox = -R; ox <= R; ox++)
        for (int oy = -R; oy <= R; oy++) {
            const long long ia = a.a0 + i + (long long)M * ox, ib = a.a0 + j + (long long)M * oy;
            if (ia < 0 || ib < 0 || ia >= n || ib >= n) continue;
            sum += a.square[((size_t)ia * n + ib) * 3 + 1 + ch];
        }
    a.ion_sum[r] = sum;
}

static int cdiv_i(long long a, int b) { return (int)((a + b - 1) / b); }

// Device buffers of one build, freed on every exit.
struct TabMem {
    std::vector<void *> ptrs;
    ~TabMem() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    hipError_t get(double **p, size_t count) {
        void *v = nullptr;
        const hipError_t e = hipMalloc(&v, count * sizeof(double));
        if (e == hipSuccess) ptrs.push_back(v);
        *p = (double *)v;
        return e;
    }
};

// Build the tables for (softening, passes, cutoff) on the device.
static int build_on_device(EmbDevice *dev, double softening, int passes, double cutoff, std::shared_ptr<HostTables> &out) {
    if (!(softening > 0.0) || !std::isfinite(softening)) return fail(EMB_E_INVALID, "electron_softening must be positive");
    // The smoothed charge lands inside the margin only while passes <= margin.
    if (passes < 0 || passes > T_MARGIN) return fail(EMB_E_INVALID, "smoothing_passes must be in [0, 16]");
    // At least three radii (grid_force reads the second), and a bounded build.
    if (!(cutoff >= 0.02) || cutoff > MAX_CUTOFF) return fail(EMB_E_INVALID, "electron_cutoff must be in [0.02, 32] cells");
    EMB_HIP(hipSetDevice(dev->ordinal));
    auto t = std::make_shared<HostTables>();
    t->softening = softening;
    t->passes = passes;
    t->cutoff = cutoff;
    t->reach = (int)std::ceil(cutoff);
    TabGeom g{};
    g.n = 2 * t->reach + 24;
    g.N = g.n + 2 * T_MARGIN;
    g.P = g.n + 2 * passes;
    g.passes = passes;
    g.count = (int)std::llround(cutoff / T_STEP) + 1;
    g.reach = cutoff;
    g.softening = softening;
    g.per_block = 2 * (size_t)g.P * g.P + 2 * (size_t)g.N * g.N + 2 * (size_t)(g.N + 1) * g.N;
    const int count = g.count, NS = T_SOURCES * T_SOURCES;
    t->half = std::ceil(cutoff) + 1.0;
    const int cnt = (int)std::llround((2.0 * t->half + 1.0) * T_SQUARE);
    const int n = cnt - T_SQUARE + 1;
    t->square_n = n;
    t->ion_a0 = (int)std::llround((t->half - 0.5) * T_SQUARE);
    const int side = 2 * t->reach + 1;

    TabMem mem;
    double *S, *L, *rr, *ff, *near, *pot, *gt, *pt, *pts, *sat, *sq, *st, *is;
    EMB_HIP(mem.get(&S, (size_t)g.N * g.N));
    EMB_HIP(mem.get(&L, (size_t)g.N * g.N));
    EMB_HIP(mem.get(&g.work, g.per_block * NS));
    EMB_HIP(mem.get(&g.part, (size_t)count * NS));
    EMB_HIP(mem.get(&rr, count));
    EMB_HIP(mem.get(&ff, count));
    EMB_HIP(mem.get(&near, count));
    EMB_HIP(mem.get(&pot, count));
    EMB_HIP(mem.get(&gt, PAIR_SAMPLES + 2));
    EMB_HIP(mem.get(&pt, PAIR_SAMPLES + 2));
    EMB_HIP(mem.get(&pts, (size_t)cnt * cnt * 3));
    EMB_HIP(mem.get(&sat, (size_t)(cnt + 1) * (cnt + 1) * 3));
    EMB_HIP(mem.get(&sq, (size_t)n * n * 3));
    EMB_HIP(mem.get(&st, (size_t)side * side));
    EMB_HIP(mem.get(&is, (size_t)ION_SUM * ION_SUM * 2));
    g.S = S;
    g.L = L;
    hipStream_t st_ = nullptr;
    EMB_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st_};
    hipLaunchKernelGGL(k_tab_matrices, dim3(cdiv_i((long long)g.N * g.N, 256)), dim3(256), 0, st_, g.N, S, L);
    hipLaunchKernelGGL(k_tab_source, dim3(NS), dim3(T_THREADS), 0, st_, g);
    RadialArgs ra{count, cutoff, softening, cutoff, g.part, rr, ff, near, pot, gt, pt};
    hipLaunchKernelGGL(k_tab_radial, dim3(1), dim3(T_THREADS), 0, st_, ra);
    SquareArgs sa{cnt, n, t->half, cutoff, gt, pt, pts, sat, sq};
    EMB_HIP(hipMemsetAsync(sat, 0, (size_t)(cnt + 1) * (cnt + 1) * 3 * sizeof(double), st_));
    hipLaunchKernelGGL(k_tab_points, dim3(cdiv_i((long long)cnt * cnt, 256)), dim3(256), 0, st_, sa);
    hipLaunchKernelGGL(k_tab_sat, dim3(1), dim3(1024), 0, st_, sa);
    hipLaunchKernelGGL(k_tab_box, dim3(cdiv_i((long long)n * n * 3, 256)), dim3(256), 0, st_, sa);
    IonArgs ia{n, t->reach, t->ion_a0, t->half, sq, st, is};
    hipLaunchKernelGGL(k_tab_ion, dim3(cdiv_i((long long)side * side + ION_SUM * ION_SUM * 2, 256)), dim3(256), 0, st_, ia);
    EMB_HIP(hipGetLastError());
    auto back = [&](std::vector<double> &v, const double *p, size_t len) -> hipError_t {
        v.resize(len);
        return hipMemcpyAsync(v.data(), p, len * sizeof(double), hipMemcpyDeviceToHost, st_);
    };
    EMB_HIP(back(t->grid_r, rr, count));
    EMB_HIP(back(t->grid_f, ff, count));
    EMB_HIP(back(t->g, gt, PAIR_SAMPLES + 2));
    EMB_HIP(back(t->potential, pt, PAIR_SAMPLES + 2));
    EMB_HIP(back(t->square, sq, (size_t)n * n * 3));
    EMB_HIP(back(t->ion_stencil, st, (size_t)side * side));
    EMB_HIP(back(t->ion_sum, is, (size_t)ION_SUM * ION_SUM * 2));
    EMB_HIP(hipStreamSynchronize(st_));
    out = t;
    return EMB_OK;
}

int ensure_tables(EmbDevice *dev, double softening, int passes, double cutoff) {
    if (!(cutoff > 0.0)) return EMB_OK;   // no close-range forces, no tables
    std::lock_guard<std::mutex> guard(dev->lock);
    if (find_tables(softening, passes, cutoff)) return EMB_OK;
    std::shared_ptr<HostTables> t;
    const int rc = build_on_device(dev, softening, passes, cutoff, t);
    if (rc != EMB_OK) return rc;
    store_tables(t);
    return EMB_OK;
}

}  // namespace emb

extern "C" int emb_close_tables(EmbDevice *dev, double softening, int32_t passes, double cutoff, double *g,
                                double *potential, double *square, int32_t *square_n, double *half,
                                double *ion_stencil, int32_t *reach, double *build_ms) {
    using namespace emb;
    EMB_GUARD({
        if (!dev) return fail(EMB_E_INVALID, "null device");
        std::shared_ptr<const HostTables> t;
        if (build_ms) {   // a fresh build, timed on the host clock (launches, kernels and the copy back)
            std::lock_guard<std::mutex> guard(dev->lock);
            std::shared_ptr<HostTables> fresh;
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = build_on_device(dev, softening, passes, cutoff, fresh);
            if (rc != EMB_OK) return rc;
            *build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (!find_tables(softening, passes, cutoff)) store_tables(fresh);
            t = fresh;
        } else {
            const int rc = ensure_tables(dev, softening, passes, cutoff);
            if (rc != EMB_OK) return rc;
            t = find_tables(softening, passes, cutoff);
            if (!t) return fail(EMB_E_INVALID, "electron_cutoff must be positive");
        }
        const int side = 2 * t->reach + 1;
        if (square_n) *square_n = t->square_n;
        if (half) *half = t->half;
        if (reach) *reach = t->reach;
        if (g) std::copy(t->g.begin(), t->g.end(), g);
        if (potential) std::copy(t->potential.begin(), t->potential.end(), potential);
        if (square) std::copy(t->square.begin(), t->square.end(), square);
        if (ion_stencil) std::copy(t->ion_stencil.begin(), t->ion_stencil.begin() + (size_t)side * side, ion_stencil);
        return EMB_OK;
    })
}
