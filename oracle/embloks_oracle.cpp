// embloks_oracle.cpp - FP64 CPU restatement of embloks-core's per-tick physics.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP backend. Only the tests, the
// smoke check and the benchmark's CPU-baseline leg load it; the product library
// never links it. Built with -ffp-contract=off so that every product and sum is
// rounded as in the reference (Rust does not contract a * b + c into an FMA).
//
// Each function cites the lines of crates/embloks-core/src it restates. Arrays
// are x-major (index i * h + j) like the reference's ndarray Array2 (w, h).

#include "embloks_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double PI = 3.14159265358979323846;
int g_threads = 1;

// --- Small containers -----------------------------------------------------------------

struct Grid {
    int w = 0, h = 0;
    std::vector<double> v;
    Grid() = default;
    Grid(int w_, int h_, double fill = 0.0) : w(w_), h(h_), v((size_t)w_ * h_, fill) {}
    double &at(int i, int j) { return v[(size_t)i * h + j]; }
    double at(int i, int j) const { return v[(size_t)i * h + j]; }
};

// (w, h, d) array, index (i * h + j) * d + k, like ndarray Array3 in standard layout.
struct Grid3 {
    int w = 0, h = 0, d = 0;
    std::vector<double> v;
    Grid3() = default;
    Grid3(int w_, int h_, int d_) : w(w_), h(h_), d(d_), v((size_t)w_ * h_ * d_, 0.0) {}
    double &at(int i, int j, int k) { return v[((size_t)i * h + j) * d + k]; }
    double at(int i, int j, int k) const { return v[((size_t)i * h + j) * d + k]; }
};

// --- DST-I (the arithmetic rustdct 0.7.1's Dst1 computes, maxwell.rs:28, 139-159) ----------
// Unnormalised type-I sine transform: y[k] = sum_n x[n] sin(pi (k+1)(n+1) / (N+1)).
// Small lengths by the definition; larger ones through an FFT of the odd extension
// of length 2 (N + 1) (Bluestein's chirp-z for lengths that are not powers of two).

using cplx = std::complex<double>;

// In-place radix-2 FFT of a power-of-two length; `tw` holds exp(-2 pi i k / n), k < n / 2.
void fft_pow2(std::vector<cplx> &a, const std::vector<cplx> &tw, bool inverse) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t stride = n / len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t k = 0; k < len / 2; k++) {
                const cplx w = inverse ? std::conj(tw[k * stride]) : tw[k * stride];
                const cplx u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
        }
    }
    if (inverse)
        for (auto &x : a) x /= (double)n;
}

struct DstPlan {
    int n = 0;           // transform length N
    int m = 0;           // FFT length 2 (N + 1)
    bool direct = true;
    std::vector<double> sine;   // direct: N x N table
    // Bluestein for length m
    int mp = 0;
    std::vector<cplx> chirp;    // exp(-i pi k^2 / m), k < m
    std::vector<cplx> kernel_f; // FFT of conj chirp, length mp
    std::vector<cplx> tw;       // twiddles of the length-mp FFT

    explicit DstPlan(int n_) : n(n_), m(2 * (n_ + 1)) {
        direct = n <= 96;
        if (direct) {
            sine.resize((size_t)n * n);
            for (int k = 0; k < n; k++)
                for (int j = 0; j < n; j++)
                    sine[(size_t)k * n + j] = std::sin(PI * (double)(((long long)(k + 1) * (j + 1)) % (2 * (n + 1))) / (double)(n + 1));
            return;
        }
        mp = 1;
        while (mp < 2 * m - 1) mp <<= 1;
        tw.resize(mp / 2);
        for (int k = 0; k < mp / 2; k++) {
            const double ang = -2.0 * PI * (double)k / (double)mp;
            tw[k] = cplx(std::cos(ang), std::sin(ang));
        }
        chirp.resize(m);
        for (int k = 0; k < m; k++) {
            const long long q = ((long long)k * k) % (2LL * m);
            const double ang = -PI * (double)q / (double)m;
            chirp[k] = cplx(std::cos(ang), std::sin(ang));
        }
        kernel_f.assign(mp, cplx(0, 0));
        kernel_f[0] = std::conj(chirp[0]);
        for (int k = 1; k < m; k++) kernel_f[k] = kernel_f[mp - k] = std::conj(chirp[k]);
        fft_pow2(kernel_f, tw, false);
    }

    void apply(double *x, std::vector<cplx> &work) const {
        if (direct) {
            double tmp[96];
            for (int k = 0; k < n; k++) {
                double acc = 0.0;
                const double *row = &sine[(size_t)k * n];
                for (int j = 0; j < n; j++) acc += x[j] * row[j];
                tmp[k] = acc;
            }
            std::memcpy(x, tmp, sizeof(double) * n);
            return;
        }
        // Odd extension y of length m: y[0] = 0, y[1..N] = x, y[N+1] = 0, y[m-k] = -y[k].
        // FFT(y)[k] = -2 i DST(x)[k-1] for k = 1..N.
        work.assign(mp, cplx(0, 0));
        for (int j = 0; j < n; j++) {
            work[j + 1] = x[j] * chirp[j + 1];
            work[m - 1 - j] = -x[j] * chirp[m - 1 - j];
        }
        fft_pow2(work, tw, false);
        for (int k = 0; k < mp; k++) work[k] *= kernel_f[k];
        fft_pow2(work, tw, true);
        for (int k = 1; k <= n; k++) {
            const cplx yk = work[k] * chirp[k];
            x[k - 1] = -0.5 * yk.imag();
        }
    }
};

// --- Settings helpers -------------------------------------------------------------------

// maxwell.rs:376-378
int fdtd_substeps(const EmbSettings &s, double dt) {
    const double v = std::ceil(s.speed_of_light * dt * std::sqrt(2.0) / 0.9);
    return std::max(1, (int)v);
}

// --- Maxwell (maxwell.rs:34-414) -----------------------------------------------------------

// maxwell.rs:381-390
double bilinear(const Grid &a, double gx, double gy) {
    const long long i = std::clamp((long long)std::floor(gx), 0LL, (long long)a.w - 2);
    const long long j = std::clamp((long long)std::floor(gy), 0LL, (long long)a.h - 2);
    const double fx = gx - (double)i, fy = gy - (double)j;
    return a.at(i, j) * (1.0 - fx) * (1.0 - fy) + a.at(i + 1, j) * fx * (1.0 - fy) +
           a.at(i, j + 1) * (1.0 - fx) * fy + a.at(i + 1, j + 1) * fx * fy;
}

// One pass of the 1-2-1 filter along x and y, zero beyond the edges (maxwell.rs:397-414).
Grid3 smooth(const Grid3 &input) {
    Grid3 array = input;
    for (int axis = 0; axis < 2; axis++) {
        Grid3 out = array;
        for (auto &x : out.v) x = 0.5 * x;
        const int n = axis == 0 ? array.w : array.h;
        const int other = axis == 0 ? array.h : array.w;
        for (int i = 0; i < n; i++) {
            for (int o = 0; o < other; o++) {
                for (int k = 0; k < array.d; k++) {
                    const double quarter = 0.25 * (axis == 0 ? array.at(i, o, k) : array.at(o, i, k));
                    if (i > 0) (axis == 0 ? out.at(i - 1, o, k) : out.at(o, i - 1, k)) += quarter;
                    if (i + 1 < n) (axis == 0 ? out.at(i + 1, o, k) : out.at(o, i + 1, k)) += quarter;
                }
            }
        }
        array = std::move(out);
    }
    return array;
}

struct Maxwell {
    int cols, rows, margin, nx, ny;
    EmbSettings s;
    Grid ex, ey, bz, potential, inv_laplacian;
    std::shared_ptr<DstPlan> dst_x, dst_y;
    bool has_damping = false;
    double damping_key = 0.0;
    Grid damping[3];

    // maxwell.rs:56-85
    Maxwell(int cols_, int rows_, const EmbSettings &settings) : cols(cols_), rows(rows_), s(settings) {
        margin = s.absorbing_margin;
        nx = cols + 2 * margin;
        ny = rows + 2 * margin;
        auto eigen = [](int n) {
            std::vector<double> out;
            for (int k = 1; k <= n; k++) out.push_back(2.0 * std::cos(PI * (double)k / (double)(n + 1)) - 2.0);
            return out;
        };
        const auto kx = eigen(nx), ky = eigen(ny);
        const double norm = 4.0 / (double)((nx + 1) * (ny + 1));
        inv_laplacian = Grid(nx, ny);
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) inv_laplacian.at(i, j) = norm / (kx[i] + ky[j]);
        ex = Grid(nx + 1, ny);
        ey = Grid(nx, ny + 1);
        bz = Grid(nx + 1, ny + 1);
        potential = Grid(nx, ny);
        dst_x = std::make_shared<DstPlan>(nx);
        dst_y = ny == nx ? dst_x : std::make_shared<DstPlan>(ny);
    }

    // maxwell.rs:111-134
    const Grid *damping_factors(double dt) {
        if (!has_damping || damping_key != dt) {
            const double m = (double)margin;
            auto depth = [m](double coord, int n) {
                const double outside = std::max(std::max(m - coord, coord - ((double)n - 1.0 - m)), 0.0);
                return std::min(outside / m, 1.0);
            };
            const double strength = s.absorbing_strength;
            auto factor = [&](int w, int h, double xo, double yo) {
                Grid g(w, h);
                for (int i = 0; i < w; i++)
                    for (int j = 0; j < h; j++) {
                        const double d = std::max(depth((double)i + xo, nx), depth((double)j + yo, ny));
                        g.at(i, j) = std::exp(-strength * (d * d * d) * dt);
                    }
                return g;
            };
            damping[0] = factor(nx + 1, ny, -0.5, 0.0);
            damping[1] = factor(nx, ny + 1, 0.0, -0.5);
            damping[2] = factor(nx + 1, ny + 1, -0.5, -0.5);
            has_damping = true;
            damping_key = dt;
        }
        return damping;
    }

    // maxwell.rs:139-159: DST-I along axis 0 (each column j), then along axis 1.
    void dst2(Grid &a) const {
        {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
            for (int j = 0; j < ny; j++) {
                std::vector<double> lane(nx);
                std::vector<cplx> work;
                for (int i = 0; i < nx; i++) lane[i] = a.at(i, j);
                dst_x->apply(lane.data(), work);
                for (int i = 0; i < nx; i++) a.at(i, j) = lane[i];
            }
        }
        {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
            for (int i = 0; i < nx; i++) {
                std::vector<cplx> work;
                dst_y->apply(&a.v[(size_t)i * ny], work);
            }
        }
    }

    // maxwell.rs:162-168
    Grid poisson(const Grid &rhs) const {
        Grid phi = rhs;
        dst2(phi);
        for (size_t k = 0; k < phi.v.size(); k++) phi.v[k] *= inv_laplacian.v[k];
        dst2(phi);
        return phi;
    }

    // maxwell.rs:170-173
    Grid divergence() const {
        Grid d(nx, ny);
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                d.at(i, j) = (ex.at(i + 1, j) - ex.at(i, j)) + (ey.at(i, j + 1) - ey.at(i, j));
        return d;
    }

    // maxwell.rs:176-193
    void gradient_subtract(const Grid &phi) {
        for (int i = 0; i <= nx; i++)
            for (int j = 0; j < ny; j++) {
                const double high = i < nx ? phi.at(i, j) : 0.0;
                const double low = i > 0 ? phi.at(i - 1, j) : 0.0;
                ex.at(i, j) -= high - low;
            }
        for (int i = 0; i < nx; i++)
            for (int j = 0; j <= ny; j++) {
                const double high = j < ny ? phi.at(i, j) : 0.0;
                const double low = j > 0 ? phi.at(i, j - 1) : 0.0;
                ey.at(i, j) -= high - low;
            }
    }

    // maxwell.rs:200-204
    void enforce_gauss(const Grid &rho) {
        Grid rhs(nx, ny);
        for (size_t k = 0; k < rhs.v.size(); k++) rhs.v[k] = -4.0 * PI * s.coulomb_k * rho.v[k];
        potential = poisson(rhs);
        Grid correction = poisson(divergence());
        for (size_t k = 0; k < correction.v.size(); k++) correction.v[k] += potential.v[k];
        gradient_subtract(correction);
    }

    // maxwell.rs:207-227
    void curl_e(Grid &curl) const {
        for (int i = 0; i <= nx; i++)
            for (int j = 0; j <= ny; j++) {
                double v = 0.0;
                if (i < nx) v += ey.at(i, j);
                if (i > 0) v -= ey.at(i - 1, j);
                if (j < ny) v -= ex.at(i, j);
                if (j > 0) v += ex.at(i, j - 1);
                curl.at(i, j) = v;
            }
    }

    // maxwell.rs:235-272
    void step(const Grid &rho, const Grid3 &current, double dt) {
        const int substeps = fdtd_substeps(s, dt);
        const double h = dt / (double)substeps;
        const double c2 = s.speed_of_light * s.speed_of_light;
        const double scale = 4.0 * PI * s.coulomb_k * h;
        Grid jx(nx + 1, ny), jy(nx, ny + 1);
        for (int i = 0; i <= nx; i++)
            for (int j = 0; j < ny; j++)
                jx.at(i, j) = (i == 0 || i == nx) ? 0.0 : 0.5 * scale * (current.at(i - 1, j, 0) + current.at(i, j, 0));
        for (int i = 0; i < nx; i++)
            for (int j = 0; j <= ny; j++)
                jy.at(i, j) = (j == 0 || j == ny) ? 0.0 : 0.5 * scale * (current.at(i, j - 1, 1) + current.at(i, j, 1));
        const Grid *damp = damping_factors(h);
        Grid curl(nx + 1, ny + 1);
        for (int sub = 0; sub < substeps; sub++) {
            curl_e(curl);
            for (size_t k = 0; k < bz.v.size(); k++) bz.v[k] = bz.v[k] + (-0.5 * h) * curl.v[k];
            for (int i = 0; i <= nx; i++)
                for (int j = 0; j < ny; j++) ex.at(i, j) += h * c2 * (bz.at(i, j + 1) - bz.at(i, j)) - jx.at(i, j);
            for (int i = 0; i < nx; i++)
                for (int j = 0; j <= ny; j++) ey.at(i, j) -= h * c2 * (bz.at(i + 1, j) - bz.at(i, j)) + jy.at(i, j);
            curl_e(curl);
            for (size_t k = 0; k < bz.v.size(); k++) bz.v[k] = bz.v[k] + (-0.5 * h) * curl.v[k];
            for (size_t k = 0; k < ex.v.size(); k++) ex.v[k] *= damp[0].v[k];
            for (size_t k = 0; k < ey.v.size(); k++) ey.v[k] *= damp[1].v[k];
            for (size_t k = 0; k < bz.v.size(); k++) bz.v[k] *= damp[2].v[k];
        }
        enforce_gauss(rho);
    }

    // maxwell.rs:280-292
    Grid3 smoothed(const Grid3 &values) const {
        const int k = s.smoothing_passes;
        Grid3 padded(values.w + 2 * k, values.h + 2 * k, values.d);
        for (int i = 0; i < values.w; i++)
            for (int j = 0; j < values.h; j++)
                for (int c = 0; c < values.d; c++) padded.at(i + k, j + k, c) = values.at(i, j, c);
        for (int p = 0; p < k; p++) padded = smooth(padded);
        Grid3 out(nx, ny, values.d);
        const int m = margin - k;
        for (int i = 0; i < padded.w; i++)
            for (int j = 0; j < padded.h; j++)
                for (int c = 0; c < values.d; c++) out.at(m + i, m + j, c) = padded.at(i, j, c);
        return out;
    }

    // maxwell.rs:336-350
    void gather(int64_t n, const double *pos, double *out) const {
        const double m = (double)margin;
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
        for (int64_t p = 0; p < n; p++) {
            const double x = pos[2 * p], y = pos[2 * p + 1];
            out[3 * p + 0] = bilinear(ex, x + m, y - 0.5 + m);
            out[3 * p + 1] = bilinear(ey, x - 0.5 + m, y + m);
            out[3 * p + 2] = bilinear(bz, x + m, y + m);
        }
    }

    // maxwell.rs:353-368
    void node_fields(double *e, double *b) const {
        for (int c = 0; c < cols; c++)
            for (int r = 0; r < rows; r++) {
                const int i = c + margin, j = r + margin;
                e[((size_t)c * rows + r) * 2 + 0] = 0.5 * (ex.at(i, j) + ex.at(i + 1, j));
                e[((size_t)c * rows + r) * 2 + 1] = 0.5 * (ey.at(i, j) + ey.at(i, j + 1));
                if (b) b[(size_t)c * rows + r] = 0.25 * (bz.at(i, j) + bz.at(i + 1, j) + bz.at(i, j + 1) + bz.at(i + 1, j + 1));
            }
    }
};

// --- Cloud in cell (fields.rs:18-112) ------------------------------------------------------

struct CloudInCell {
    int cols, rows;
    std::vector<std::array<size_t, 4>> index;
    std::vector<std::array<double, 4>> weight;

    // fields.rs:27-40
    CloudInCell(int64_t n, const double *pos, int cols_, int rows_) : cols(cols_), rows(rows_) {
        auto clip = [](long long v, int n) { return (size_t)std::clamp(v, 0LL, (long long)n - 1); };
        index.resize(n);
        weight.resize(n);
        for (int64_t p = 0; p < n; p++) {
            const double u = pos[2 * p] - 0.5, v = pos[2 * p + 1] - 0.5;
            const long long c0 = (long long)std::floor(u), r0 = (long long)std::floor(v);
            const double fu = u - (double)c0, fv = v - (double)r0;
            const size_t ca = clip(c0, cols), cb = clip(c0 + 1, cols);
            const size_t ra = clip(r0, rows), rb = clip(r0 + 1, rows);
            index[p] = {ca * rows + ra, ca * rows + rb, cb * rows + ra, cb * rows + rb};
            weight[p] = {(1.0 - fu) * (1.0 - fv), (1.0 - fu) * fv, fu * (1.0 - fv), fu * fv};
        }
    }

    // fields.rs:43-51
    void deposit(const double *values, double *flat) const {
        for (size_t p = 0; p < index.size(); p++)
            for (int k = 0; k < 4; k++) flat[index[p][k]] += weight[p][k] * values[p];
    }

    // fields.rs:54-63
    void deposit_vectors(const double *values, double *flat) const {
        for (size_t p = 0; p < index.size(); p++)
            for (int k = 0; k < 4; k++) {
                flat[2 * index[p][k]] += weight[p][k] * values[2 * p];
                flat[2 * index[p][k] + 1] += weight[p][k] * values[2 * p + 1];
            }
    }
};

// fields.rs:91-107
Grid spread_over_cells(const Grid &input) {
    Grid values = input;
    for (int axis = 0; axis < 2; axis++) {
        const int n = axis == 0 ? values.w : values.h;
        const int other = axis == 0 ? values.h : values.w;
        Grid out = values;
        for (auto &x : out.v) x = 0.75 * x;
        for (int i = 0; i < n; i++) {
            const int below = i > 0 ? i - 1 : i;
            const int above = i + 1 < n ? i + 1 : i;
            for (int o = 0; o < other; o++) {
                const double here = (axis == 0 ? values.at(i, o) : values.at(o, i)) / 8.0;
                (axis == 0 ? out.at(below, o) : out.at(o, below)) += here;
                (axis == 0 ? out.at(above, o) : out.at(o, above)) += here;
            }
        }
        values = std::move(out);
    }
    return values;
}

// --- Close-range tables (close_range.rs:29-242) --------------------------------------------

const int PAIR_SAMPLES = 4096;
const int SQUARE_SAMPLES = 32;
const int SOURCE_SAMPLES = 8;
const int DIRECTIONS = 16;
const double RADIAL_STEP = 0.01;
const double SERIES_BELOW = 1e-4;

// close_range.rs:45-49
double compact_g(double r2, double softening) {
    const double x = r2 / (softening * softening);
    const double value = x < SERIES_BELOW ? 2.0 * (1.0 - x / 2.0 + x * x / 6.0) : -2.0 * std::expm1(-x) / x;
    return value / (softening * softening);
}

// close_range.rs:51-53
std::vector<double> linspace(double start, double end, size_t n) {
    std::vector<double> out(n);
    for (size_t i = 0; i < n; i++) out[i] = start + (end - start) * (double)i / (double)(n - 1);
    return out;
}

// close_range.rs:56-65 (numpy.interp)
double interp(double x, const std::vector<double> &xs, const std::vector<double> &ys) {
    if (x <= xs[0]) return ys[0];
    if (x >= xs.back()) return ys.back();
    const size_t i = (size_t)(std::upper_bound(xs.begin(), xs.end(), x) - xs.begin()) - 1;
    return ys[i] + (ys[i + 1] - ys[i]) * (x - xs[i]) / (xs[i + 1] - xs[i]);
}

// The physics defaults of Settings::default (settings.rs:164-236).
EmbSettings default_settings() {
    EmbSettings s;
    std::memset(&s, 0, sizeof(s));
    s.dt = 1.0 / 30.0;
    s.cell_charge = 4.0;
    s.carriers_per_cell = 8;
    s.battery_density = 4;
    s.carrier_charge_to_mass = 1.0;
    s.drag_per_resistance = 1.0;
    s.battery_resistance = 0.1;
    s.battery_trim_time = 0.25;
    s.battery_gain_low = 0.5;
    s.battery_gain_high = 4.0;
    s.current_smoothing_time = 2.0;
    s.wall_restitution = 0.3;
    s.max_step = 0.4;
    s.max_substeps = 16;
    s.smoothing_passes = 2;
    s.coulomb_k = 1.0;
    s.speed_of_light = 8.0;
    s.absorbing_margin = 16;
    s.absorbing_strength = 3.0;
    s.electron_softening = 0.25;
    s.electron_cutoff = 3.0;
    s.display_smoothing_time = 0.5;
    return s;
}

// close_range.rs:73-108
void grid_force(int passes, double reach, std::vector<double> &radii, std::vector<double> &force) {
    const int n = 2 * (int)std::ceil(reach) + 24;
    EmbSettings settings = default_settings();
    settings.smoothing_passes = passes;
    settings.coulomb_k = 1.0;
    Maxwell mx(n, n, settings);
    radii = linspace(0.0, reach, (size_t)std::llround(reach / RADIAL_STEP) + 1);
    std::vector<std::array<double, 2>> dirs(DIRECTIONS);
    for (int k = 0; k < DIRECTIONS; k++) {
        const double angle = 2.0 * PI * ((double)k + 0.5) / (double)DIRECTIONS;
        dirs[k] = {std::cos(angle), std::sin(angle)};
    }
    std::vector<double> total(radii.size(), 0.0);
    std::vector<double> points(radii.size() * DIRECTIONS * 2), field(radii.size() * DIRECTIONS * 3);
    for (int i = 0; i < SOURCE_SAMPLES; i++) {
        for (int j = 0; j < SOURCE_SAMPLES; j++) {
            const double source[2] = {(double)(n / 2) + ((double)i + 0.5) / (double)SOURCE_SAMPLES,
                                      (double)(n / 2) + ((double)j + 0.5) / (double)SOURCE_SAMPLES};
            std::fill(mx.ex.v.begin(), mx.ex.v.end(), 0.0);
            std::fill(mx.ey.v.begin(), mx.ey.v.end(), 0.0);
            Grid3 dep(n, n, 1);
            const double one = 1.0;
            CloudInCell(1, source, n, n).deposit(&one, dep.v.data());
            Grid3 sm = mx.smoothed(dep);
            Grid rho(mx.nx, mx.ny);
            rho.v = sm.v;
            mx.enforce_gauss(rho);
            size_t p = 0;
            for (double r : radii)
                for (auto &u : dirs) {
                    points[2 * p] = source[0] + r * u[0];
                    points[2 * p + 1] = source[1] + r * u[1];
                    p++;
                }
            mx.gather((int64_t)p, points.data(), field.data());
            for (size_t k = 0; k < radii.size(); k++) {
                double along = 0.0;
                for (int d = 0; d < DIRECTIONS; d++) {
                    const double *f = &field[3 * (k * DIRECTIONS + d)];
                    along += f[0] * dirs[d][0] + f[1] * dirs[d][1];
                }
                total[k] += along / (double)DIRECTIONS;
            }
        }
    }
    const double samples = (double)(SOURCE_SAMPLES * SOURCE_SAMPLES);
    force.resize(total.size());
    for (size_t k = 0; k < total.size(); k++) force[k] = total[k] / samples;
}

// close_range.rs:174-181
double radial(const std::vector<double> &table, double r2, double cutoff) {
    if (r2 >= cutoff * cutoff) return 0.0;
    const double x = std::min(r2 * ((double)PAIR_SAMPLES / (cutoff * cutoff)), (double)PAIR_SAMPLES);
    const size_t i = (size_t)x;
    return table[i] + (table[i + 1] - table[i]) * (x - (double)i);
}

struct Tables {
    std::vector<double> g, potential;
    Grid3 square;
    double half = 0.0;
    Grid ion_stencil;
    int reach = 0;
};

// close_range.rs:151-171
void radial_tables(double softening, int passes, double cutoff, std::vector<double> &g, std::vector<double> &pot) {
    std::vector<double> r, f_grid;
    grid_force(passes, cutoff, r, f_grid);
    std::vector<double> f_near(r.size());
    for (size_t i = 0; i < r.size(); i++) f_near[i] = r[i] * compact_g(r[i] * r[i], softening) - f_grid[i];
    std::vector<double> potential(r.size(), 0.0);
    for (size_t i = r.size() - 1; i-- > 0;)
        potential[i] = potential[i + 1] + 0.5 * (f_near[i + 1] + f_near[i]) * (r[i + 1] - r[i]);
    const auto r2s = linspace(0.0, cutoff * cutoff, PAIR_SAMPLES + 1);
    g.clear();
    pot.clear();
    for (size_t i = 0; i < r2s.size(); i++) {
        const double r2 = r2s[i];
        const double rs = std::sqrt(r2);
        const double grid_g = i == 0 ? f_grid[1] / r[1] : interp(rs, r, f_grid) / rs;
        g.push_back(compact_g(r2, softening) - grid_g);
        pot.push_back(interp(rs, r, potential));
    }
    g.push_back(0.0);
    pot.push_back(0.0);
}

// close_range.rs:184-188
void point_correction(const double d[2], const std::vector<double> &g, const std::vector<double> &pot, double cutoff,
                      double out[3]) {
    const double r2 = d[0] * d[0] + d[1] * d[1];
    const double gr = radial(g, r2, cutoff);
    out[0] = radial(pot, r2, cutoff);
    out[1] = gr * d[0];
    out[2] = gr * d[1];
}

// close_range.rs:195-219
void square_table(const std::vector<double> &g, const std::vector<double> &pot, double cutoff, Grid3 &table,
                  double &half) {
    const int m = SQUARE_SAMPLES;
    const double h = 1.0 / (double)m;
    half = std::ceil(cutoff) + 1.0;
    const int count = (int)std::llround((2.0 * half + 1.0) * (double)m);
    std::vector<double> z(count);
    for (int k = 0; k < count; k++) z[k] = -half - 0.5 + ((double)k + 0.5) * h;
    Grid3 total(count + 1, count + 1, 3);
    for (int a = 0; a < count; a++)
        for (int b = 0; b < count; b++) {
            const double d[2] = {z[a], z[b]};
            double point[3];
            point_correction(d, g, pot, cutoff, point);
            for (int k = 0; k < 3; k++)
                total.at(a + 1, b + 1, k) = point[k] + total.at(a, b + 1, k) + total.at(a + 1, b, k) - total.at(a, b, k);
        }
    const int n = count - m + 1;
    table = Grid3(n, n, 3);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < 3; k++)
                table.at(i, j, k) = (total.at(i + m, j + m, k) - total.at(i, j + m, k) - total.at(i + m, j, k) + total.at(i, j, k)) * h * h;
}

// close_range.rs:223-242
Grid ion_stencil(const Grid3 &square, double half, double cutoff) {
    const int m = SQUARE_SAMPLES;
    const int reach = (int)std::ceil(cutoff);
    std::vector<double> weights(m + 1, 1.0 / (double)m);
    weights[0] *= 0.5;
    weights[m] *= 0.5;
    const int size = 2 * reach + 1;
    Grid out(size, size);
    for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) {
            const int dc = i - reach, dr = j - reach;
            const size_t a = (size_t)std::llround(((double)dc - 0.5 + half) * (double)m);
            const size_t b = (size_t)std::llround(((double)dr - 0.5 + half) * (double)m);
            double sum = 0.0;
            for (int x = 0; x <= m; x++)
                for (int y = 0; y <= m; y++) sum += weights[x] * square.at(a + x, b + y, 0) * weights[y];
            out.at(i, j) = sum;
        }
    return out;
}

// close_range.rs:129-146 (cached per process, like the reference)
std::shared_ptr<const Tables> tables(double softening, int passes, double cutoff) {
    static std::mutex lock;
    static std::map<std::tuple<double, int, double>, std::shared_ptr<const Tables>> cache;
    const auto key = std::make_tuple(softening, passes, cutoff);
    {
        std::lock_guard<std::mutex> guard(lock);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    auto t = std::make_shared<Tables>();
    radial_tables(softening, passes, cutoff, t->g, t->potential);
    square_table(t->g, t->potential, cutoff, t->square, t->half);
    t->ion_stencil = ion_stencil(t->square, t->half, cutoff);
    t->reach = (int)std::ceil(cutoff);
    std::lock_guard<std::mutex> guard(lock);
    cache[key] = t;
    return t;
}

// close_range.rs:245-254
double table_bilinear(const Grid3 &table, int k, double u, double v) {
    const int n = table.w;
    const long long a = std::clamp((long long)std::floor(u), 0LL, (long long)n - 2);
    const long long b = std::clamp((long long)std::floor(v), 0LL, (long long)n - 2);
    const double fu = u - (double)a, fv = v - (double)b;
    return table.at(a, b, k) * (1.0 - fu) * (1.0 - fv) + table.at(a + 1, b, k) * fu * (1.0 - fv) +
           table.at(a, b + 1, k) * (1.0 - fu) * fv + table.at(a + 1, b + 1, k) * fu * fv;
}

// --- World (world.rs, materials.rs, connectivity.rs, battery.rs) -----------------------------

enum { VACUUM = 0, CONDUCTOR = 1, BATTERY = 2 };
const double ORIENT[4][2] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0}};  // materials.rs:25

struct Stack {
    std::vector<std::pair<int, int>> cells;
    long long low[2], high[2];
    double emf, internal_resistance;
    std::vector<std::pair<double, double>> sources;

    double amplitude() const {  // battery.rs:81-83
        double a = 0.0;
        for (auto &s : sources) a += std::fabs(s.first);
        return a;
    }
    double emf_at(double time) const {  // battery.rs:86-96
        if (sources.empty()) return emf;
        double total = 0.0;
        for (auto &s : sources) total += s.second == 0.0 ? s.first : s.first * std::sin(2.0 * PI * s.second * time);
        return total;
    }
};

struct World {
    EmbSettings s;
    int cols, rows;
    std::vector<uint8_t> kind, orientation;
    std::vector<double> resistance, density, emf, frequency;
    std::vector<uint8_t> open_x, open_y;   // (cols-1, rows), (cols, rows-1)
    std::vector<long long> labels;
    std::vector<Stack> stacks;
    std::vector<int> has_gain;
    std::vector<double> gain;
    std::vector<double> pos, vel, charge, mass;
    std::vector<double> fixed_pos, fixed_charge;

    size_t n() const { return charge.size(); }
    bool in_bounds(long long c, long long r) const { return c >= 0 && r >= 0 && c < cols && r < rows; }
    int material(int c, int r) const {
        const int k = kind[(size_t)c * rows + r];
        return (k == 1 || k == 2) ? k : 0;
    }
    bool conductive(int c, int r) const { return material(c, r) != VACUUM; }

    // connectivity.rs:20-62
    void compute_connectivity() {
        open_x.assign((size_t)std::max(cols - 1, 0) * rows, 0);
        open_y.assign((size_t)cols * std::max(rows - 1, 0), 0);
        for (int c = 0; c + 1 < cols; c++)
            for (int r = 0; r < rows; r++) open_x[(size_t)c * rows + r] = conductive(c, r) && conductive(c + 1, r);
        for (int c = 0; c < cols; c++)
            for (int r = 0; r + 1 < rows; r++) open_y[(size_t)c * (rows - 1) + r] = conductive(c, r) && conductive(c, r + 1);
        labels.assign((size_t)cols * rows, -1);
        long long count = 0;
        std::vector<std::pair<int, int>> stack;
        for (int c = 0; c < cols; c++)
            for (int r = 0; r < rows; r++) {
                if (!conductive(c, r) || labels[(size_t)c * rows + r] >= 0) continue;
                labels[(size_t)c * rows + r] = count;
                stack.push_back({c, r});
                while (!stack.empty()) {
                    auto [x, y] = stack.back();
                    stack.pop_back();
                    auto visit = [&](int cx, int cy, bool open) {
                        if (open && labels[(size_t)cx * rows + cy] < 0) {
                            labels[(size_t)cx * rows + cy] = count;
                            stack.push_back({cx, cy});
                        }
                    };
                    if (x > 0) visit(x - 1, y, open_x[(size_t)(x - 1) * rows + y]);
                    if (x + 1 < cols) visit(x + 1, y, open_x[(size_t)x * rows + y]);
                    if (y > 0) visit(x, y - 1, open_y[(size_t)x * (rows - 1) + y - 1]);
                    if (y + 1 < rows) visit(x, y + 1, open_y[(size_t)x * (rows - 1) + y]);
                }
                count++;
            }
    }

    // battery.rs:100-147
    void find_stacks() {
        stacks.clear();
        auto horizontal = [&](int c, int r) {
            const int o = orientation[(size_t)c * rows + r] % 4;
            return o == 0 || o == 2;
        };
        auto in_run = [&](long long c, long long r, bool axis_h) {
            return in_bounds(c, r) && material((int)c, (int)r) == BATTERY && horizontal((int)c, (int)r) == axis_h;
        };
        std::vector<uint8_t> seen((size_t)cols * rows, 0);
        for (int c = 0; c < cols; c++)
            for (int r = 0; r < rows; r++) {
                if (kind[(size_t)c * rows + r] != BATTERY || seen[(size_t)c * rows + r]) continue;
                const bool axis_h = horizontal(c, r);
                const long long sx = axis_h ? 1 : 0, sy = axis_h ? 0 : 1;
                long long first[2] = {c, r};
                while (in_run(first[0] - sx, first[1] - sy, axis_h)) {
                    first[0] -= sx;
                    first[1] -= sy;
                }
                Stack st;
                long long cell[2] = {first[0], first[1]};
                while (in_run(cell[0], cell[1], axis_h)) {
                    st.cells.push_back({(int)cell[0], (int)cell[1]});
                    cell[0] += sx;
                    cell[1] += sy;
                }
                for (auto &x : st.cells) seen[(size_t)x.first * rows + x.second] = 1;
                double total = 0.0, resist = 0.0;
                for (auto &x : st.cells) {
                    const size_t k = (size_t)x.first * rows + x.second;
                    const int o = orientation[k] % 4;
                    const double sign = (o == 0 || o == 1) ? 1.0 : -1.0;
                    st.sources.push_back({sign * emf[k], frequency[k]});
                }
                for (auto &src : st.sources) total += src.first;
                for (auto &x : st.cells) resist += resistance[(size_t)x.first * rows + x.second];
                st.low[0] = first[0] - sx;
                st.low[1] = first[1] - sy;
                st.high[0] = cell[0];
                st.high[1] = cell[1];
                st.emf = total;
                st.internal_resistance = resist;
                stacks.push_back(st);
            }
    }

    // battery.rs:150-152
    bool wired(const long long t[2]) const {
        return in_bounds(t[0], t[1]) && material((int)t[0], (int)t[1]) == CONDUCTOR;
    }

    // battery.rs:155-158: 3x3 around a terminal, clipped, columns outer.
    template <class F> void terminal_block(const long long t[2], F f) const {
        for (long long x = std::max(t[0] - 1, 0LL); x < std::min(t[0] + 2, (long long)cols); x++)
            for (long long y = std::max(t[1] - 1, 0LL); y < std::min(t[1] + 2, (long long)rows); y++) f((int)x, (int)y);
    }

    // battery.rs:165-176
    bool terminal_potential(const Grid &potential, const long long t[2], double &out) const {
        if (!wired(t)) return false;
        const long long island = labels[(size_t)t[0] * rows + t[1]];
        double sum = 0.0;
        int count = 0;
        terminal_block(t, [&](int x, int y) {
            if (labels[(size_t)x * rows + y] == island && material(x, y) == CONDUCTOR) {
                sum += potential.at(x, y);
                count++;
            }
        });
        out = sum / (double)count;
        return true;
    }

    // battery.rs:179-190
    std::vector<uint8_t> probe_cells() const {
        std::vector<uint8_t> mask((size_t)cols * rows, 0);
        for (auto &st : stacks) {
            terminal_block(st.low, [&](int x, int y) { mask[(size_t)x * rows + y] = 1; });
            terminal_block(st.high, [&](int x, int y) { mask[(size_t)x * rows + y] = 1; });
        }
        return mask;
    }

    // battery.rs:193-228
    Grid3 impressed_field_grid(double time) const {
        Grid gain_grid(cols, rows, 1.0);
        for (size_t s = 0; s < stacks.size(); s++)
            if (has_gain[s])
                for (auto &x : stacks[s].cells) gain_grid.at(x.first, x.second) = gain[s];
        Grid3 field(cols, rows, 2);
        for (int c = 0; c < cols; c++)
            for (int r = 0; r < rows; r++) {
                const size_t k = (size_t)c * rows + r;
                if (emf[k] == 0.0) continue;
                double value = emf[k] * gain_grid.at(c, r);
                if (frequency[k] > 0.0) value *= std::sin(2.0 * PI * frequency[k] * time);
                const int o = orientation[k];
                field.at(c, r, 0) = value * ORIENT[o][0];
                field.at(c, r, 1) = value * ORIENT[o][1];
            }
        return field;
    }

    // battery.rs:240-252
    double stack_current(const Stack &st, const Grid3 &current) const {
        const double dx = (double)((st.high[0] > st.low[0]) - (st.high[0] < st.low[0]));
        const double dy = (double)((st.high[1] > st.low[1]) - (st.high[1] < st.low[1]));
        std::vector<std::pair<int, int>> cells = st.cells;
        if (wired(st.low)) cells.push_back({(int)st.low[0], (int)st.low[1]});
        if (wired(st.high)) cells.push_back({(int)st.high[0], (int)st.high[1]});
        double total = 0.0;
        for (auto &x : cells) total += current.at(x.first, x.second, 0) * dx + current.at(x.first, x.second, 1) * dy;
        return total / (double)cells.size();
    }

    // physics.rs:17-21
    std::vector<double> damping_rates() const {
        const double per_ohm = s.drag_per_resistance * s.cell_charge * s.carrier_charge_to_mass;
        std::vector<double> out(resistance.size());
        for (size_t k = 0; k < out.size(); k++) out[k] = resistance[k] * density[k] * per_ohm;
        return out;
    }
};

// physics.rs:24-29
inline void grid_cell(const double *p, int cols, int rows, int &c, int &r) {
    c = (int)std::clamp((long long)std::floor(p[0]), 0LL, (long long)cols - 1);
    r = (int)std::clamp((long long)std::floor(p[1]), 0LL, (long long)rows - 1);
}

// physics.rs:34-43
Grid3 current_density(const World &w) {
    Grid3 current(w.cols, w.rows, 2);
    for (size_t i = 0; i < w.n(); i++) {
        int c, r;
        grid_cell(&w.pos[2 * i], w.cols, w.rows, c, r);
        current.at(c, r, 0) += w.charge[i] * w.vel[2 * i];
        current.at(c, r, 1) += w.charge[i] * w.vel[2 * i + 1];
    }
    return current;
}

const double WALL_GAP = 1e-6;  // physics.rs:10

// physics.rs:83-115
void resolve(int axis, long long start, long long other, double pos[2], double vel[2], const World &w, double restitution) {
    const long long end = (long long)std::floor(pos[axis]);
    if (end == start) return;
    const bool forward = end > start;
    const long long face = forward ? start : start - 1;
    const long long n = axis == 0 ? w.cols : w.rows;
    const long long n_other = axis == 0 ? w.rows : w.cols;
    const bool on_grid = face >= 0 && face < n - 1 && other >= 0 && other < n_other;
    bool open = false;
    if (on_grid) {
        open = axis == 0 ? w.open_x[(size_t)face * w.rows + other] : w.open_y[(size_t)other * (w.rows - 1) + face];
    }
    if (open) return;
    const double wall = (double)(forward ? start + 1 : start);
    const double bounced = wall - restitution * (pos[axis] - wall);
    pos[axis] = std::clamp(bounced, (double)start + WALL_GAP, (double)start + 1.0 - WALL_GAP);
    vel[axis] *= -restitution;
}

// physics.rs:66-79
void confine(const double old_pos[2], double pos[2], double vel[2], const World &w, double restitution) {
    const long long old_cell[2] = {(long long)std::floor(old_pos[0]), (long long)std::floor(old_pos[1])};
    resolve(0, old_cell[0], old_cell[1], pos, vel, w, restitution);
    const long long column = (long long)std::floor(pos[0]);
    resolve(1, old_cell[1], column, pos, vel, w, restitution);
}

// --- Close range (close_range.rs:256-465) --------------------------------------------------

struct CellList {  // close_range.rs:258-286
    std::vector<size_t> cell, start, order;
    CellList(const World &w) {
        const size_t n = w.n(), cells = (size_t)w.cols * w.rows;
        cell.resize(n);
        start.assign(cells + 1, 0);
        for (size_t i = 0; i < n; i++) {
            int c, r;
            grid_cell(&w.pos[2 * i], w.cols, w.rows, c, r);
            cell[i] = (size_t)c * w.rows + r;
            start[cell[i] + 1]++;
        }
        for (size_t i = 0; i < cells; i++) start[i + 1] += start[i];
        std::vector<size_t> fill(start.begin(), start.end() - 1);
        order.resize(n);
        for (size_t i = 0; i < n; i++) order[fill[cell[i]]++] = i;
    }
};

struct CloseRange {
    EmbSettings s;
    bool enabled() const { return s.electron_cutoff > 0.0; }
    std::shared_ptr<const Tables> t() const { return tables(s.electron_softening, s.smoothing_passes, s.electron_cutoff); }

    // close_range.rs:314-376
    std::vector<double> force(const World &w) const {
        auto tab = t();
        const double k = s.coulomb_k, k_ion = s.coulomb_k * s.cell_charge;
        const double cutoff = s.electron_cutoff;
        const int reach = (int)std::ceil(cutoff);
        const double cutoff2 = cutoff * cutoff;
        const double scale = (double)(tab->g.size() - 2) / cutoff2;
        const int m = tab->square.w;
        const double samples = (double)(m - 1) / (2.0 * tab->half);
        const int cols = w.cols, rows = w.rows;
        CellList list(w);
        const size_t n = w.n();
        std::vector<double> force(2 * n, 0.0);
        if (g_threads > 1) {
            // Timing only (the CPU baseline on all cores): each electron's force
            // from its own side, equal to the loop below up to FP64 rounding.
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
            for (long long i = 0; i < (long long)n; i++) force_on(w, list, (size_t)i, &force[2 * (size_t)i]);
            return force;
        }
        for (size_t i = 0; i < n; i++) {
            const double x = w.pos[2 * i], y = w.pos[2 * i + 1];
            const int ci = (int)(list.cell[i] / rows), ri = (int)(list.cell[i] % rows);
            for (int c = std::max(ci - reach, 0); c < std::min(ci + reach + 1, cols); c++) {
                for (int r = std::max(ri - reach, 0); r < std::min(ri + reach + 1, rows); r++) {
                    const double ion = w.density[(size_t)c * rows + r];
                    if (ion > 0.0) {
                        const double u = (x - (double)c - 0.5 + tab->half) * samples;
                        const double v = (y - (double)r - 0.5 + tab->half) * samples;
                        const long long a = std::clamp((long long)u, 0LL, (long long)m - 2);
                        const long long b = std::clamp((long long)v, 0LL, (long long)m - 2);
                        const double fu = u - (double)a, fv = v - (double)b;
                        for (int axis = 0; axis < 2; axis++) {
                            const Grid3 &sq = tab->square;
                            const double value = sq.at(a, b, axis + 1) * (1.0 - fu) * (1.0 - fv) +
                                                 sq.at(a + 1, b, axis + 1) * fu * (1.0 - fv) +
                                                 sq.at(a, b + 1, axis + 1) * (1.0 - fu) * fv +
                                                 sq.at(a + 1, b + 1, axis + 1) * fu * fv;
                            force[2 * i + axis] += k_ion * ion * w.charge[i] * value;
                        }
                    }
                    const size_t other = (size_t)c * rows + r;
                    if (other < list.cell[i]) continue;
                    for (size_t s = list.start[other]; s < list.start[other + 1]; s++) {
                        const size_t j = list.order[s];
                        if (other == list.cell[i] && j <= i) continue;
                        const double dx = x - w.pos[2 * j], dy = y - w.pos[2 * j + 1];
                        const double d2 = dx * dx + dy * dy;
                        if (d2 >= cutoff2) continue;
                        const double tt = d2 * scale;
                        const size_t si = (size_t)tt;
                        const double f = k * w.charge[i] * w.charge[j] *
                                         (tab->g[si] + (tab->g[si + 1] - tab->g[si]) * (tt - (double)si));
                        force[2 * i] += f * dx;
                        force[2 * i + 1] += f * dy;
                        force[2 * j] -= f * dx;
                        force[2 * j + 1] -= f * dy;
                    }
                }
            }
        }
        return force;
    }

    // close_range.rs:314-376 for electron i alone: every pair seen from i's side
    // (the force the Newton-3 loop above accumulates on i), plus the ions.
    void force_on(const World &w, const CellList &list, size_t i, double out[2]) const {
        auto tab = t();
        const double k = s.coulomb_k, k_ion = s.coulomb_k * s.cell_charge;
        const double cutoff = s.electron_cutoff;
        const int reach = (int)std::ceil(cutoff);
        const double cutoff2 = cutoff * cutoff;
        const double scale = (double)(tab->g.size() - 2) / cutoff2;
        const int m = tab->square.w;
        const double samples = (double)(m - 1) / (2.0 * tab->half);
        const int cols = w.cols, rows = w.rows;
        const double x = w.pos[2 * i], y = w.pos[2 * i + 1];
        const int ci = (int)(list.cell[i] / rows), ri = (int)(list.cell[i] % rows);
        out[0] = out[1] = 0.0;
        for (int c = std::max(ci - reach, 0); c < std::min(ci + reach + 1, cols); c++) {
            for (int r = std::max(ri - reach, 0); r < std::min(ri + reach + 1, rows); r++) {
                const double ion = w.density[(size_t)c * rows + r];
                if (ion > 0.0) {
                    const double u = (x - (double)c - 0.5 + tab->half) * samples;
                    const double v = (y - (double)r - 0.5 + tab->half) * samples;
                    const long long a = std::clamp((long long)u, 0LL, (long long)m - 2);
                    const long long b = std::clamp((long long)v, 0LL, (long long)m - 2);
                    const double fu = u - (double)a, fv = v - (double)b;
                    for (int axis = 0; axis < 2; axis++) {
                        const Grid3 &sq = tab->square;
                        const double value = sq.at(a, b, axis + 1) * (1.0 - fu) * (1.0 - fv) +
                                             sq.at(a + 1, b, axis + 1) * fu * (1.0 - fv) +
                                             sq.at(a, b + 1, axis + 1) * (1.0 - fu) * fv +
                                             sq.at(a + 1, b + 1, axis + 1) * fu * fv;
                        out[axis] += k_ion * ion * w.charge[i] * value;
                    }
                }
                const size_t other = (size_t)c * rows + r;
                for (size_t s2 = list.start[other]; s2 < list.start[other + 1]; s2++) {
                    const size_t j = list.order[s2];
                    if (j == i) continue;
                    const double dx = x - w.pos[2 * j], dy = y - w.pos[2 * j + 1];
                    const double d2 = dx * dx + dy * dy;
                    if (d2 >= cutoff2) continue;
                    const double tt = d2 * scale;
                    const size_t si = (size_t)tt;
                    const double f = k * w.charge[i] * w.charge[j] *
                                     (tab->g[si] + (tab->g[si + 1] - tab->g[si]) * (tt - (double)si));
                    out[0] += f * dx;
                    out[1] += f * dy;
                }
            }
        }
    }

    // close_range.rs:379-403
    Grid ion_potential(const World &w) const {
        auto tab = t();
        const Grid &stencil = tab->ion_stencil;
        const int reach = stencil.w / 2;
        const double scale = s.coulomb_k * s.cell_charge;
        Grid ions(w.cols, w.rows);
        for (int c = 0; c < w.cols; c++)
            for (int r = 0; r < w.rows; r++) {
                double sum = 0.0;
                for (int dc = -reach; dc <= reach; dc++)
                    for (int dr = -reach; dr <= reach; dr++) {
                        const long long x = c + dc, y = r + dr;
                        if (w.in_bounds(x, y)) sum += w.density[(size_t)x * w.rows + y] * stencil.at(dc + reach, dr + reach);
                    }
                ions.at(c, r) = scale * sum;
            }
        return ions;
    }

    // close_range.rs:427-464
    Grid charges_potential(const World &w, const std::vector<uint8_t> *where) const {
        auto tab = t();
        const int cols = w.cols, rows = w.rows;
        const long long reach = (long long)std::ceil(s.electron_cutoff);
        std::vector<uint8_t> near;
        if (where) {  // dilate (close_range.rs:468-477)
            near.assign((size_t)cols * rows, 0);
            for (int c = 0; c < cols; c++)
                for (int r = 0; r < rows; r++) {
                    if (!(*where)[(size_t)c * rows + r]) continue;
                    for (long long x = std::max(c - reach, 0LL); x < std::min(c + reach + 1, (long long)cols); x++)
                        for (long long y = std::max(r - reach, 0LL); y < std::min(r + reach + 1, (long long)rows); y++)
                            near[(size_t)x * rows + y] = 1;
                }
        }
        Grid total(cols, rows);
        for (size_t p = 0; p < w.n(); p++) {
            const double px = w.pos[2 * p], py = w.pos[2 * p + 1], q = w.charge[p];
            const long long pc = (long long)std::floor(px), pr = (long long)std::floor(py);
            if (where) {
                const long long cc = std::clamp(pc, 0LL, (long long)cols - 1), rr = std::clamp(pr, 0LL, (long long)rows - 1);
                if (!near[(size_t)cc * rows + rr]) continue;
            }
            for (long long dc = -reach; dc <= reach; dc++)
                for (long long dr = -reach; dr <= reach; dr++) {
                    const long long c = pc + dc, r = pr + dr;
                    if (!w.in_bounds(c, r)) continue;
                    const double d0 = px - ((double)c + 0.5), d1 = py - ((double)r + 0.5);
                    const double u = (d0 + tab->half) * (double)SQUARE_SAMPLES;
                    const double v = (d1 + tab->half) * (double)SQUARE_SAMPLES;
                    total.at(c, r) += table_bilinear(tab->square, 0, u, v) * q;
                }
        }
        for (auto &x : total.v) x *= s.coulomb_k;
        return total;
    }

    // close_range.rs:415-423
    Grid potential(const World &w, const std::vector<uint8_t> *where) const {
        Grid out = ion_potential(w);
        Grid charges = charges_potential(w, where);
        for (size_t k = 0; k < out.v.size(); k++) {
            out.v[k] += charges.v[k];
            if (where && !(*where)[k]) out.v[k] = 0.0;
        }
        return out;
    }
};

}  // namespace

// --- Simulation (simulation.rs:15-81, physics.rs:124-162) ---------------------------------

struct OrcSim {
    World world;
    std::unique_ptr<Maxwell> maxwell;
    CloseRange close;
    double time = 0.0;
    int last_substeps = 0;

    // maxwell.rs:303-316
    void sources(Grid &rho, Grid3 &current) const {
        const World &w = world;
        CloudInCell cic((int64_t)w.n(), w.pos.data(), w.cols, w.rows);
        Grid3 values(w.cols, w.rows, 3);
        Grid dens(w.cols, w.rows);
        for (size_t k = 0; k < dens.v.size(); k++) dens.v[k] = w.density[k] * w.s.cell_charge;
        Grid rho_v = spread_over_cells(dens);
        std::vector<double> dep((size_t)w.cols * w.rows, 0.0);
        cic.deposit(w.charge.data(), dep.data());
        for (size_t k = 0; k < dep.size(); k++) rho_v.v[k] += dep[k];
        if (!w.fixed_charge.empty()) {
            std::vector<double> fixed((size_t)w.cols * w.rows, 0.0);
            CloudInCell((int64_t)w.fixed_charge.size(), w.fixed_pos.data(), w.cols, w.rows).deposit(w.fixed_charge.data(), fixed.data());
            for (size_t k = 0; k < fixed.size(); k++) rho_v.v[k] += fixed[k];
        }
        std::vector<double> qv(2 * w.n());
        for (size_t i = 0; i < w.n(); i++) {
            qv[2 * i] = w.charge[i] * w.vel[2 * i];
            qv[2 * i + 1] = w.charge[i] * w.vel[2 * i + 1];
        }
        std::vector<double> vec((size_t)w.cols * w.rows * 2, 0.0);
        cic.deposit_vectors(qv.data(), vec.data());
        for (size_t k = 0; k < (size_t)w.cols * w.rows; k++) {
            values.v[3 * k] = rho_v.v[k];
            values.v[3 * k + 1] = vec[2 * k];
            values.v[3 * k + 2] = vec[2 * k + 1];
        }
        Grid3 full = maxwell->smoothed(values);
        rho = Grid(maxwell->nx, maxwell->ny);
        current = Grid3(maxwell->nx, maxwell->ny, 2);
        for (size_t k = 0; k < rho.v.size(); k++) {
            rho.v[k] = full.v[3 * k];
            current.v[2 * k] = full.v[3 * k + 1];
            current.v[2 * k + 1] = full.v[3 * k + 2];
        }
    }

    // maxwell.rs:326-329
    void advance(double dt) {
        Grid rho;
        Grid3 current;
        sources(rho, current);
        maxwell->step(rho, current, dt);
    }

    // simulation.rs:48-55 (sync is a no-op: the scene never changes inside the oracle)
    Grid potential(const std::vector<uint8_t> *where) const {
        Grid out(world.cols, world.rows);
        for (int c = 0; c < world.cols; c++)
            for (int r = 0; r < world.rows; r++) out.at(c, r) = maxwell->potential.at(c + maxwell->margin, r + maxwell->margin);
        if (close.enabled()) {
            Grid extra = close.potential(world, where);
            for (size_t k = 0; k < out.v.size(); k++) out.v[k] += extra.v[k];
        }
        return out;
    }

    // battery.rs:260-281
    void trim(const Grid &potential, const Grid3 &current, double dt) {
        World &w = world;
        for (size_t si = 0; si < w.stacks.size(); si++) {
            const Stack &st = w.stacks[si];
            double v_low = 0.0, v_high = 0.0;
            const bool ok_low = w.terminal_potential(potential, st.low, v_low);
            const bool ok_high = w.terminal_potential(potential, st.high, v_high);
            const double amplitude = st.amplitude();
            if (!ok_low || !ok_high || amplitude == 0.0) {
                w.has_gain[si] = 0;
                continue;
            }
            const double target = st.emf_at(time);
            const double delivered = v_high - v_low + w.stack_current(st, current) * st.internal_resistance;
            double gain = w.has_gain[si] ? w.gain[si] : 1.0;
            gain += dt / w.s.battery_trim_time * (target - delivered) * target / (amplitude * amplitude);
            w.gain[si] = std::clamp(gain, w.s.battery_gain_low, w.s.battery_gain_high);
            w.has_gain[si] = 1;
        }
    }

    // physics.rs:124-162
    int physics_step(double dt) {
        World &w = world;
        const EmbSettings &s = w.s;
        if (w.n() == 0) {
            advance(dt);
            return 0;
        }
        double speed = 0.0;
        for (size_t i = 0; i < w.n(); i++) speed = std::max(speed, std::hypot(w.vel[2 * i], w.vel[2 * i + 1]));
        const int substeps = (int)std::clamp((long long)std::ceil(speed * dt / s.max_step), 1LL, (long long)s.max_substeps);
        const double h = dt / (double)substeps;
        const std::vector<double> damping = w.damping_rates();
        std::vector<double> fields(3 * w.n());
        for (int k = 0; k < substeps; k++) {
            const Grid3 impressed = w.impressed_field_grid(time + (double)k * h);
            maxwell->gather((int64_t)w.n(), w.pos.data(), fields.data());
            std::vector<double> closef;
            if (close.enabled()) closef = close.force(w);
            // Electrons are independent here: threads change the timing, not the result.
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
            for (size_t i = 0; i < w.n(); i++) {
                int c, r;
                grid_cell(&w.pos[2 * i], w.cols, w.rows, c, r);
                const double e0 = fields[3 * i] + impressed.at(c, r, 0), e1 = fields[3 * i + 1] + impressed.at(c, r, 1);
                const double qm = w.charge[i] / w.mass[i];
                const double v0 = w.vel[2 * i], v1 = w.vel[2 * i + 1], bz = fields[3 * i + 2];
                const double l0 = e0 + v1 * bz, l1 = e1 - v0 * bz;  // physics.rs:49-51
                double acc0 = qm * l0, acc1 = qm * l1;
                if (!closef.empty()) {
                    acc0 += closef[2 * i] / w.mass[i];
                    acc1 += closef[2 * i + 1] / w.mass[i];
                }
                const double sc = 1.0 / (1.0 + damping[(size_t)c * w.rows + r] * h);  // physics.rs:57-60
                double vel[2] = {(v0 + acc0 * h) * sc, (v1 + acc1 * h) * sc};
                double moved[2] = {vel[0] * h, vel[1] * h};
                const double length = std::hypot(moved[0], moved[1]);
                const double limit = std::min(s.max_step / std::max(length, 1e-300), 1.0);
                moved[0] *= limit;
                moved[1] *= limit;
                const double old_pos[2] = {w.pos[2 * i], w.pos[2 * i + 1]};
                double target[2] = {old_pos[0] + moved[0], old_pos[1] + moved[1]};
                confine(old_pos, target, vel, w, s.wall_restitution);
                w.pos[2 * i] = target[0];
                w.pos[2 * i + 1] = target[1];
                w.vel[2 * i] = vel[0];
                w.vel[2 * i + 1] = vel[1];
            }
        }
        advance(dt);
        return substeps;
    }

    // simulation.rs:63-72
    int step() {
        const double dt = world.s.dt;
        const std::vector<uint8_t> probe = world.probe_cells();
        const Grid potential_now = potential(&probe);
        const Grid3 current = current_density(world);
        trim(potential_now, current, dt);
        const int substeps = physics_step(dt);
        time += dt;
        return substeps;
    }
};

// --- C ABI ---------------------------------------------------------------------------------

extern "C" {

void orc_set_threads(int threads) { g_threads = std::max(1, threads); }

struct OrcMaxwell {
    Maxwell mx;
};

OrcMaxwell *orc_mx_new(const EmbSettings *s, int cols, int rows) { return new OrcMaxwell{Maxwell(cols, rows, *s)}; }
void orc_mx_free(OrcMaxwell *m) { delete m; }
void orc_mx_dims(const OrcMaxwell *m, int *nx, int *ny) {
    *nx = m->mx.nx;
    *ny = m->mx.ny;
}
void orc_mx_set(OrcMaxwell *m, const double *ex, const double *ey, const double *bz) {
    if (ex) std::copy(ex, ex + m->mx.ex.v.size(), m->mx.ex.v.begin());
    if (ey) std::copy(ey, ey + m->mx.ey.v.size(), m->mx.ey.v.begin());
    if (bz) std::copy(bz, bz + m->mx.bz.v.size(), m->mx.bz.v.begin());
}
void orc_mx_get(const OrcMaxwell *m, double *ex, double *ey, double *bz, double *potential) {
    if (ex) std::copy(m->mx.ex.v.begin(), m->mx.ex.v.end(), ex);
    if (ey) std::copy(m->mx.ey.v.begin(), m->mx.ey.v.end(), ey);
    if (bz) std::copy(m->mx.bz.v.begin(), m->mx.bz.v.end(), bz);
    if (potential) std::copy(m->mx.potential.v.begin(), m->mx.potential.v.end(), potential);
}
void orc_mx_enforce_gauss(OrcMaxwell *m, const double *rho) {
    Grid g(m->mx.nx, m->mx.ny);
    std::copy(rho, rho + g.v.size(), g.v.begin());
    m->mx.enforce_gauss(g);
}
void orc_mx_step(OrcMaxwell *m, const double *rho, const double *current, double dt) {
    Grid g(m->mx.nx, m->mx.ny);
    std::copy(rho, rho + g.v.size(), g.v.begin());
    Grid3 c(m->mx.nx, m->mx.ny, 2);
    std::copy(current, current + c.v.size(), c.v.begin());
    m->mx.step(g, c, dt);
}
void orc_mx_poisson(const OrcMaxwell *m, const double *rhs, double *phi) {
    Grid g(m->mx.nx, m->mx.ny);
    std::copy(rhs, rhs + g.v.size(), g.v.begin());
    Grid out = m->mx.poisson(g);
    std::copy(out.v.begin(), out.v.end(), phi);
}
void orc_mx_dst2(const OrcMaxwell *m, double *array) {
    Grid g(m->mx.nx, m->mx.ny);
    std::copy(array, array + g.v.size(), g.v.begin());
    m->mx.dst2(g);
    std::copy(g.v.begin(), g.v.end(), array);
}
void orc_mx_smoothed(const OrcMaxwell *m, const double *values, int d, double *out) {
    Grid3 v(m->mx.cols, m->mx.rows, d);
    std::copy(values, values + v.v.size(), v.v.begin());
    Grid3 o = m->mx.smoothed(v);
    std::copy(o.v.begin(), o.v.end(), out);
}
void orc_mx_gather(const OrcMaxwell *m, int64_t n, const double *pos, double *out) { m->mx.gather(n, pos, out); }
void orc_mx_node_fields(const OrcMaxwell *m, double *e, double *b) { m->mx.node_fields(e, b); }
void orc_mx_divergence(const OrcMaxwell *m, double *div) {
    Grid d = m->mx.divergence();
    std::copy(d.v.begin(), d.v.end(), div);
}
void orc_mx_damping(OrcMaxwell *m, double h, double *ex, double *ey, double *bz) {
    const Grid *d = m->mx.damping_factors(h);
    if (ex) std::copy(d[0].v.begin(), d[0].v.end(), ex);
    if (ey) std::copy(d[1].v.begin(), d[1].v.end(), ey);
    if (bz) std::copy(d[2].v.begin(), d[2].v.end(), bz);
}

void orc_cic_deposit(int64_t n, const double *pos, const double *q, int cols, int rows, double *out) {
    std::fill(out, out + (size_t)cols * rows, 0.0);
    CloudInCell(n, pos, cols, rows).deposit(q, out);
}
void orc_cic_deposit_vectors(int64_t n, const double *pos, const double *v, int cols, int rows, double *out) {
    std::fill(out, out + (size_t)cols * rows * 2, 0.0);
    CloudInCell(n, pos, cols, rows).deposit_vectors(v, out);
}
void orc_spread_over_cells(int cols, int rows, const double *in, double *out) {
    Grid g(cols, rows);
    std::copy(in, in + g.v.size(), g.v.begin());
    Grid o = spread_over_cells(g);
    std::copy(o.v.begin(), o.v.end(), out);
}

int orc_grid_force(int passes, double reach, double *radii, double *force, int capacity) {
    std::vector<double> r, f;
    grid_force(passes, reach, r, f);
    const int n = (int)r.size();
    for (int i = 0; i < std::min(n, capacity); i++) {
        radii[i] = r[i];
        force[i] = f[i];
    }
    return n;
}

void orc_tables(double softening, int passes, double cutoff, double *g, double *potential, double *square,
                int *square_n, double *half, double *stencil, int *reach) {
    auto t = tables(softening, passes, cutoff);
    if (square_n) *square_n = t->square.w;
    if (half) *half = t->half;
    if (reach) *reach = t->reach;
    if (g) std::copy(t->g.begin(), t->g.end(), g);
    if (potential) std::copy(t->potential.begin(), t->potential.end(), potential);
    if (square) std::copy(t->square.v.begin(), t->square.v.end(), square);
    if (stencil) std::copy(t->ion_stencil.v.begin(), t->ion_stencil.v.end(), stencil);
}

double orc_compact_g(double r2, double softening) { return compact_g(r2, softening); }

OrcSim *orc_sim_new(const EmbSettings *s, const EmbScene *sc) {
    auto *sim = new OrcSim();
    World &w = sim->world;
    w.s = *s;
    w.cols = sc->cols;
    w.rows = sc->rows;
    const size_t cells = (size_t)w.cols * w.rows;
    w.kind.assign(sc->kind, sc->kind + cells);
    w.orientation.assign(sc->orientation, sc->orientation + cells);
    w.resistance.assign(sc->resistance, sc->resistance + cells);
    w.density.assign(sc->density, sc->density + cells);
    w.emf.assign(sc->emf, sc->emf + cells);
    w.frequency.assign(sc->frequency, sc->frequency + cells);
    w.compute_connectivity();
    w.find_stacks();
    w.has_gain.assign(w.stacks.size(), 0);
    w.gain.assign(w.stacks.size(), 1.0);
    for (int k = 0; k < sc->n_gains && k < (int)w.stacks.size(); k++)
        if (!std::isnan(sc->gains[k])) {
            w.has_gain[k] = 1;
            w.gain[k] = sc->gains[k];
        }
    const size_t n = (size_t)sc->n_electrons;
    w.pos.assign(sc->pos, sc->pos + 2 * n);
    w.vel.assign(sc->vel, sc->vel + 2 * n);
    const double q = -s->cell_charge / (double)s->carriers_per_cell;  // world.rs:59-66
    w.charge.assign(n, q);
    w.mass.assign(n, -q / s->carrier_charge_to_mass);
    if (sc->n_fixed > 0) {
        w.fixed_pos.assign(sc->fixed_pos, sc->fixed_pos + 2 * sc->n_fixed);
        w.fixed_charge.assign(sc->fixed_charge, sc->fixed_charge + sc->n_fixed);
    }
    sim->close.s = *s;
    sim->time = sc->time;
    // Maxwell::for_world (maxwell.rs:88-94)
    sim->maxwell = std::make_unique<Maxwell>(w.cols, w.rows, *s);
    Grid rho;
    Grid3 current;
    sim->sources(rho, current);
    sim->maxwell->enforce_gauss(rho);
    return sim;
}

void orc_sim_free(OrcSim *sim) { delete sim; }

void orc_sim_dims(const OrcSim *sim, EmbDims *out) {
    std::memset(out, 0, sizeof(*out));
    out->cols = sim->world.cols;
    out->rows = sim->world.rows;
    out->margin = sim->maxwell->margin;
    out->nx = sim->maxwell->nx;
    out->ny = sim->maxwell->ny;
    out->n_stacks = (int)sim->world.stacks.size();
    out->n_electrons = (int64_t)sim->world.n();
    out->fdtd_substeps = fdtd_substeps(sim->world.s, sim->world.s.dt);
    out->last_substeps = sim->last_substeps;
    out->time = sim->time;
}

int orc_sim_step(OrcSim *sim, int ticks) {
    for (int t = 0; t < ticks; t++) sim->last_substeps = sim->step();
    return sim->last_substeps;
}

void orc_sim_read_state(const OrcSim *sim, double *pos, double *vel, double *gains) {
    const World &w = sim->world;
    if (pos) std::copy(w.pos.begin(), w.pos.end(), pos);
    if (vel) std::copy(w.vel.begin(), w.vel.end(), vel);
    if (gains)
        for (size_t k = 0; k < w.stacks.size(); k++) gains[k] = w.has_gain[k] ? w.gain[k] : std::nan("");
}

void orc_sim_read_yee(const OrcSim *sim, double *ex, double *ey, double *bz) {
    const Maxwell &m = *sim->maxwell;
    if (ex) std::copy(m.ex.v.begin(), m.ex.v.end(), ex);
    if (ey) std::copy(m.ey.v.begin(), m.ey.v.end(), ey);
    if (bz) std::copy(m.bz.v.begin(), m.bz.v.end(), bz);
}

void orc_sim_read_potential(const OrcSim *sim, double *potential) {
    std::copy(sim->maxwell->potential.v.begin(), sim->maxwell->potential.v.end(), potential);
}

void orc_sim_write_state(OrcSim *sim, const double *pos, const double *vel, const double *ex, const double *ey,
                         const double *bz) {
    World &w = sim->world;
    if (pos) std::copy(pos, pos + w.pos.size(), w.pos.begin());
    if (vel) std::copy(vel, vel + w.vel.size(), w.vel.begin());
    Maxwell &m = *sim->maxwell;
    if (ex) std::copy(ex, ex + m.ex.v.size(), m.ex.v.begin());
    if (ey) std::copy(ey, ey + m.ey.v.size(), m.ey.v.begin());
    if (bz) std::copy(bz, bz + m.bz.v.size(), m.bz.v.begin());
}

void orc_sim_current(const OrcSim *sim, double *current) {
    Grid3 c = current_density(sim->world);
    std::copy(c.v.begin(), c.v.end(), current);
}

void orc_sim_node_fields(const OrcSim *sim, double *e, double *b) { sim->maxwell->node_fields(e, b); }

void orc_sim_potential(OrcSim *sim, double *potential) {
    Grid p = sim->potential(nullptr);
    std::copy(p.v.begin(), p.v.end(), potential);
}

void orc_sim_close_force(const OrcSim *sim, double *force) {
    std::vector<double> f = sim->close.force(sim->world);
    std::copy(f.begin(), f.end(), force);
}

int orc_sim_stacks(const OrcSim *sim, int *low_high, int *n_cells) {
    const auto &st = sim->world.stacks;
    for (size_t k = 0; k < st.size(); k++) {
        if (low_high) {
            low_high[4 * k + 0] = (int)st[k].low[0];
            low_high[4 * k + 1] = (int)st[k].low[1];
            low_high[4 * k + 2] = (int)st[k].high[0];
            low_high[4 * k + 3] = (int)st[k].high[1];
        }
        if (n_cells) n_cells[k] = (int)st[k].cells.size();
    }
    return (int)st.size();
}

void orc_close_force_subset(const EmbSettings *s, const EmbScene *sc, int64_t m, const int64_t *idx, double *out) {
    World w;
    w.s = *s;
    w.cols = sc->cols;
    w.rows = sc->rows;
    const size_t cells = (size_t)w.cols * w.rows;
    w.kind.assign(sc->kind, sc->kind + cells);
    w.density.assign(sc->density, sc->density + cells);
    const size_t n = (size_t)sc->n_electrons;
    w.pos.assign(sc->pos, sc->pos + 2 * n);
    w.charge.assign(n, -s->cell_charge / (double)s->carriers_per_cell);   // world.rs:59-66
    CloseRange close;
    close.s = *s;
    const CellList list(w);
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads)
    for (int64_t q = 0; q < m; q++) close.force_on(w, list, (size_t)idx[q], out + 2 * q);
}

void orc_sim_confine(const OrcSim *sim, const double *old_pos, const double *new_pos, const double *vel,
                     double restitution, double *out_pos, double *out_vel) {
    double pos[2] = {new_pos[0], new_pos[1]}, v[2] = {vel[0], vel[1]};
    confine(old_pos, pos, v, sim->world, restitution);
    out_pos[0] = pos[0];
    out_pos[1] = pos[1];
    out_vel[0] = v[0];
    out_vel[1] = v[1];
}

void orc_kick(const double *vel, const double *acc, double damping, double dt, double *out) {
    const double scale = 1.0 / (1.0 + damping * dt);  // physics.rs:57-60
    out[0] = (vel[0] + acc[0] * dt) * scale;
    out[1] = (vel[1] + acc[1] * dt) * scale;
}

void orc_sim_labels(const OrcSim *sim, int64_t *labels) {
    std::copy(sim->world.labels.begin(), sim->world.labels.end(), labels);
}

}  // extern "C"

extern "C" void orc_settings_default(EmbSettings *out) { *out = default_settings(); }
