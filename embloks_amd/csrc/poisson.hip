// poisson.hip - the DST-I Poisson solve on its own, for one nx x ny node grid.
//
// Maxwell::poisson (crates/embloks-core/src/maxwell.rs:139-168): with Dirichlet
// boundaries (phi = 0 outside the grid), phi = P(r) solves the 5-point
// Laplacian lap(phi) = r exactly, through
//   P(r) = Sx ((Sx r Sy) * inv_lap) Sy,
//   S[k][n] = sin(pi (k+1)(n+1) / (N+1)),
//   inv_lap[i][j] = [4 / ((nx+1)(ny+1))] / (lx_i + ly_j),  l_k = 2 cos(pi (k+1) / (N+1)) - 2
// (maxwell.rs:56-85). The tick runs the same products (gemm.hip) on its
// scenes; this entry point runs them on a caller's right-hand side, in FP32 or
// BF16x3 MFMA, and times them with HIP events on its own stream.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gemm.hpp"
#include "runtime.hpp"

struct EmbPoisson {
    EmbDevice *dev = nullptr;
    int nx = 0, ny = 0;
    int sine_x = 0, sine_y = 0;   // sine lengths referenced (0: none)
    hipStream_t stream = nullptr;
    float *rhs = nullptr, *w1 = nullptr, *w2 = nullptr, *phi = nullptr, *inv_lap = nullptr;
    const float *sx = nullptr, *sy = nullptr;                                   // whole DST-I matrices
    const float *sx_e = nullptr, *sx_o = nullptr, *sy_et = nullptr, *sy_o = nullptr;   // their halves
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // BF16x3: halves of Sx, Sy and of each product's data operand, rows padded to 32
    int kx = 0, ky = 0;
    __bf16 *sxh = nullptr, *sxl = nullptr, *syh = nullptr, *syl = nullptr, *dh = nullptr, *dl = nullptr;
};

namespace emb {
namespace {

// out = Sx . in (rows of out from the halves of Sx and the folded rows of in).
hipError_t left_product(const EmbPoisson &p, const float *in, float *out, int tile) {
    GemmArgs h[2];
    for (int odd = 0; odd < 2; odd++) {
        GemmArgs &g = h[odd];
        const int mh = odd ? p.nx / 2 : (p.nx + 1) / 2;
        g.A = odd ? p.sx_o : p.sx_e;
        g.lda = sine_half_pitch(mh);
        g.B = in;
        g.ldb = p.ny;
        g.C = out;
        g.ldc = p.ny;
        g.M = mh;
        g.K = fold_k(mh, p.nx);
        g.N = p.ny;
        g.fold = 1;
        g.fold_n = p.nx;
        g.odd = odd;
    }
    return launch_gemm_fold(h[0], h[1], tile, p.stream);
}

// out = (in . Sy) .* scale (columns from the folded columns of in and the halves of Sy).
hipError_t right_product(const EmbPoisson &p, const float *in, float *out, const float *scale, int tile) {
    GemmArgs h[2];
    for (int odd = 0; odd < 2; odd++) {
        GemmArgs &g = h[odd];
        const int mh = odd ? p.ny / 2 : (p.ny + 1) / 2;
        g.A = in;
        g.lda = p.ny;
        g.B = odd ? p.sy_o : p.sy_et;
        g.ldb = sine_half_pitch(mh);
        g.C = out;
        g.ldc = p.ny;
        g.scale = scale;
        g.M = p.nx;
        g.N = mh;
        g.Np = sine_half_pitch(mh);
        g.K = fold_k(mh, p.ny);
        g.fold = 2;
        g.fold_n = p.ny;
        g.odd = odd;
    }
    return launch_gemm_fold(h[0], h[1], tile, p.stream);
}

// The four products, each as its two folded halves (half the flops of S . x).
hipError_t enqueue_solve_fp32(const EmbPoisson &p, int tile) {
    hipError_t e;
    if ((e = left_product(p, p.rhs, p.w1, tile)) != hipSuccess) return e;            // w1 = Sx . rhs
    if ((e = right_product(p, p.w1, p.w2, p.inv_lap, tile)) != hipSuccess) return e;  // w2 = (w1 . Sy) * inv_lap
    if ((e = left_product(p, p.w2, p.w1, tile)) != hipSuccess) return e;             // w1 = Sx . w2
    return right_product(p, p.w1, p.phi, nullptr, tile);                             // phi = w1 . Sy
}

// The same four products as C = A . B^T on split operands: Sx is A as it is
// (rows K-contiguous), Sy is B^T as it is (symmetric); the data operand is
// split before each product, transposed where the contraction runs over its
// rows (stages 0 and 2).
hipError_t enqueue_solve_bf16x3(const EmbPoisson &p, int tile) {
    const int nx = p.nx, ny = p.ny;
    hipError_t e;
    // stage 0: w1 = Sx . rhs = Sx . (rhs^T)^T
    if ((e = launch_split(p.rhs, nx, ny, ny, true, p.dh, p.dl, p.kx, p.stream)) != hipSuccess) return e;
    if ((e = launch_gemm_nt_bf16x3(p.sxh, p.sxl, p.kx, p.dh, p.dl, p.kx, p.w1, nullptr, ny, nx, ny, p.kx, tile,
                                   p.stream)) != hipSuccess)
        return e;
    // stage 1: w2 = (w1 . Sy) * inv_lap
    if ((e = launch_split(p.w1, nx, ny, ny, false, p.dh, p.dl, p.ky, p.stream)) != hipSuccess) return e;
    if ((e = launch_gemm_nt_bf16x3(p.dh, p.dl, p.ky, p.syh, p.syl, p.ky, p.w2, p.inv_lap, ny, nx, ny, p.ky, tile,
                                   p.stream)) != hipSuccess)
        return e;
    // stage 2: w1 = Sx . w2
    if ((e = launch_split(p.w2, nx, ny, ny, true, p.dh, p.dl, p.kx, p.stream)) != hipSuccess) return e;
    if ((e = launch_gemm_nt_bf16x3(p.sxh, p.sxl, p.kx, p.dh, p.dl, p.kx, p.w1, nullptr, ny, nx, ny, p.kx, tile,
                                   p.stream)) != hipSuccess)
        return e;
    // stage 3: phi = w1 . Sy
    if ((e = launch_split(p.w1, nx, ny, ny, false, p.dh, p.dl, p.ky, p.stream)) != hipSuccess) return e;
    return launch_gemm_nt_bf16x3(p.dh, p.dl, p.ky, p.syh, p.syl, p.ky, p.phi, nullptr, ny, nx, ny, p.ky, tile, p.stream);
}

}  // namespace
}  // namespace emb

using emb::fail;

extern "C" {

int emb_poisson_create(EmbDevice *dev, int32_t nx, int32_t ny, EmbPoisson **out) {
    EMB_GUARD({
        if (!dev || !out) return fail(EMB_E_INVALID, "null argument");
        if (nx < 1 || ny < 1 || nx > 32768 || ny > 32768) return fail(EMB_E_INVALID, "grid must be 1..32768 nodes per side");
        EMB_HIP(hipSetDevice(dev->ordinal));
        auto *p = new EmbPoisson();
        p->dev = dev;
        p->nx = nx;
        p->ny = ny;
        auto bail = [&](int rc) {
            (void)hipFree(p->rhs);
            (void)hipFree(p->w1);
            (void)hipFree(p->w2);
            (void)hipFree(p->phi);
            (void)hipFree(p->inv_lap);
            for (__bf16 *a : {p->sxh, p->sxl, p->syh, p->syl, p->dh, p->dl}) (void)hipFree(a);
            if (p->e0) (void)hipEventDestroy(p->e0);
            if (p->e1) (void)hipEventDestroy(p->e1);
            if (p->stream) (void)hipStreamDestroy(p->stream);
            if (p->sine_x) emb::release_sine(p->dev, p->sine_x);
            if (p->sine_y) emb::release_sine(p->dev, p->sine_y);
            delete p;
            return rc;
        };
        const size_t nodes = (size_t)nx * ny;
        for (float **a : {&p->rhs, &p->w1, &p->w2, &p->phi, &p->inv_lap})
            if (hipMalloc(a, nodes * sizeof(float)) != hipSuccess) return bail(fail(EMB_E_NOMEM, "device memory"));
        if (hipStreamCreate(&p->stream) != hipSuccess || hipEventCreate(&p->e0) != hipSuccess ||
            hipEventCreate(&p->e1) != hipSuccess)
            return bail(fail(EMB_E_DEVICE, "stream or event creation failed"));
        const emb::SineSet *set_x = nullptr, *set_y = nullptr;
        int rc = emb::device_sine(dev, nx, &set_x);
        if (rc != EMB_OK) return bail(rc);
        p->sine_x = nx;   // the solver holds both lengths until destroyed
        rc = emb::device_sine(dev, ny, &set_y);
        if (rc != EMB_OK) return bail(rc);
        p->sine_y = ny;
        p->sx = set_x->full;
        p->sy = set_y->full;
        p->sx_e = set_x->e;
        p->sx_o = set_x->o;
        p->sy_et = set_y->e_t;
        p->sy_o = set_y->o;
        p->kx = (nx + 31) / 32 * 32;
        p->ky = (ny + 31) / 32 * 32;
        const size_t halves[6] = {(size_t)nx * p->kx, (size_t)nx * p->kx, (size_t)ny * p->ky, (size_t)ny * p->ky,
                                  std::max((size_t)ny * p->kx, (size_t)nx * p->ky), std::max((size_t)ny * p->kx, (size_t)nx * p->ky)};
        __bf16 **hp[6] = {&p->sxh, &p->sxl, &p->syh, &p->syl, &p->dh, &p->dl};
        for (int q = 0; q < 6; q++)
            if (hipMalloc(hp[q], halves[q] * sizeof(__bf16)) != hipSuccess) return bail(fail(EMB_E_NOMEM, "device memory"));
        if (emb::launch_split(p->sx, nx, nx, nx, false, p->sxh, p->sxl, p->kx, p->stream) != hipSuccess ||
            emb::launch_split(p->sy, ny, ny, ny, false, p->syh, p->syl, p->ky, p->stream) != hipSuccess ||
            hipStreamSynchronize(p->stream) != hipSuccess)
            return bail(fail(EMB_E_DEVICE, "splitting the sine matrices failed"));
        // Inverse Laplacian with the transforms' normalisation (maxwell.rs:60-67), FP64 -> FP32.
        std::vector<double> kx(nx), ky(ny);
        const double pi = 3.14159265358979323846;
        for (int k = 1; k <= nx; k++) kx[k - 1] = 2.0 * std::cos(pi * (double)k / (double)(nx + 1)) - 2.0;
        for (int k = 1; k <= ny; k++) ky[k - 1] = 2.0 * std::cos(pi * (double)k / (double)(ny + 1)) - 2.0;
        const double norm = 4.0 / ((double)(nx + 1) * (double)(ny + 1));
        std::vector<float> il(nodes);
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) il[(size_t)i * ny + j] = (float)(norm / (kx[i] + ky[j]));
        if (hipMemcpy(p->inv_lap, il.data(), nodes * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(p->rhs, 0, nodes * sizeof(float)) != hipSuccess)
            return bail(fail(EMB_E_DEVICE, "upload failed"));
        *out = p;
        return EMB_OK;
    })
}

int emb_poisson_set_rhs(EmbPoisson *p, const double *rhs) {
    EMB_GUARD({
        if (!p || !rhs) return fail(EMB_E_INVALID, "null argument");
        const size_t nodes = (size_t)p->nx * p->ny;
        std::vector<float> h(nodes);
        for (size_t i = 0; i < nodes; i++) h[i] = (float)rhs[i];
        EMB_HIP(hipSetDevice(p->dev->ordinal));
        EMB_HIP(hipMemcpyAsync(p->rhs, h.data(), nodes * sizeof(float), hipMemcpyHostToDevice, p->stream));
        EMB_HIP(hipStreamSynchronize(p->stream));
        return EMB_OK;
    })
}

int emb_poisson_solve(EmbPoisson *p, int32_t precision, int32_t tile, int32_t repeats, double *avg_ms) {
    EMB_GUARD({
        if (!p) return fail(EMB_E_INVALID, "null argument");
        if (precision != EMB_DST_FP32 && precision != EMB_DST_BF16X3) return fail(EMB_E_INVALID, "unknown precision");
        if (tile != 0 && tile != 64 && tile != 128) return fail(EMB_E_INVALID, "tile must be 0, 64 or 128");
        if (repeats < 1) return fail(EMB_E_INVALID, "repeats must be >= 1");
        if (tile == 0) tile = (long long)p->nx * p->ny >= (1LL << 20) ? 128 : 64;
        EMB_HIP(hipSetDevice(p->dev->ordinal));
        EMB_HIP(hipEventRecord(p->e0, p->stream));
        for (int r = 0; r < repeats; r++)
            EMB_HIP(precision == EMB_DST_BF16X3 ? emb::enqueue_solve_bf16x3(*p, tile) : emb::enqueue_solve_fp32(*p, tile));
        EMB_HIP(hipEventRecord(p->e1, p->stream));
        EMB_HIP(hipEventSynchronize(p->e1));
        if (avg_ms) {
            float ms = 0.0f;
            EMB_HIP(hipEventElapsedTime(&ms, p->e0, p->e1));
            *avg_ms = (double)ms / repeats;
        }
        return EMB_OK;
    })
}

int emb_poisson_read(EmbPoisson *p, double *phi) {
    EMB_GUARD({
        if (!p || !phi) return fail(EMB_E_INVALID, "null argument");
        const size_t nodes = (size_t)p->nx * p->ny;
        std::vector<float> h(nodes);
        EMB_HIP(hipSetDevice(p->dev->ordinal));
        EMB_HIP(hipMemcpyAsync(h.data(), p->phi, nodes * sizeof(float), hipMemcpyDeviceToHost, p->stream));
        EMB_HIP(hipStreamSynchronize(p->stream));
        for (size_t i = 0; i < nodes; i++) phi[i] = (double)h[i];
        return EMB_OK;
    })
}

int emb_poisson_destroy(EmbPoisson *p) {
    EMB_GUARD({
        if (!p) return EMB_OK;
        (void)hipSetDevice(p->dev->ordinal);
        if (p->stream) (void)hipStreamSynchronize(p->stream);
        for (float *a : {p->rhs, p->w1, p->w2, p->phi, p->inv_lap}) (void)hipFree(a);
        for (__bf16 *a : {p->sxh, p->sxl, p->syh, p->syl, p->dh, p->dl}) (void)hipFree(a);
        if (p->e0) (void)hipEventDestroy(p->e0);
        if (p->e1) (void)hipEventDestroy(p->e1);
        if (p->stream) (void)hipStreamDestroy(p->stream);
        if (p->sine_x) emb::release_sine(p->dev, p->sine_x);
        if (p->sine_y) emb::release_sine(p->dev, p->sine_y);
        delete p;
        return EMB_OK;
    })
}

}  // extern "C"
