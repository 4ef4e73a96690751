// setup.cpp - host-side FP64 preparation of a scene (see setup.hpp).

#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

namespace emb {

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr int SQUARE_SAMPLES = 32;     // close_range.rs:32
constexpr int SOURCE_SAMPLES = 8;      // close_range.rs:35
constexpr int DIRECTIONS = 16;         // close_range.rs:36
constexpr double RADIAL_STEP = 0.01;   // close_range.rs:37

// g of a unit line charge smeared as a Gaussian of width `eps`: field g d,
// 2 (1 - exp(-r^2/eps^2)) / r^2 (close_range.rs:41-49).
double compact(double r2, double eps) {
    const double x = r2 / (eps * eps);
    const double v = x < 1e-4 ? 2.0 * (1.0 - x / 2.0 + x * x / 6.0) : -2.0 * std::expm1(-x) / x;
    return v / (eps * eps);
}

// Piecewise-linear interpolation, clamped at the ends (numpy.interp).
double lerp_table(double x, const std::vector<double> &xs, const std::vector<double> &ys) {
    if (x <= xs.front()) return ys.front();
    if (x >= xs.back()) return ys.back();
    const size_t i = (size_t)(std::upper_bound(xs.begin(), xs.end(), x) - xs.begin()) - 1;
    return ys[i] + (ys[i + 1] - ys[i]) * (x - xs[i]) / (xs[i + 1] - xs[i]);
}

// Dense row-major n x n product C = A B.
void matmul(int n, const std::vector<double> &A, const std::vector<double> &B, std::vector<double> &C) {
    C.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
            const double a = A[(size_t)i * n + k];
            if (a == 0.0) continue;
            const double *b = &B[(size_t)k * n];
            double *c = &C[(size_t)i * n];
            for (int j = 0; j < n; j++) c[j] += a * b[j];
        }
}

// The radial force between two unit charges through a square grid of the
// default margin, measured as close_range::grid_force does (close_range.rs:73-108):
// one charge deposited (CIC), smoothed, its static field solved (DST-I Poisson
// and E = -grad phi) and read back bilinearly around it; averaged over 8 x 8
// source positions in the cell and 16 directions.
void grid_force(int passes, double reach, std::vector<double> &radii, std::vector<double> &force) {
    const int n = 2 * (int)std::ceil(reach) + 24;
    const int m = 16;          // Settings::default().absorbing_margin
    const int N = n + 2 * m;
    // Sine matrix and inverse Laplacian of the N x N node grid (maxwell.rs:56-67).
    std::vector<double> S((size_t)N * N), L((size_t)N * N);
    for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
            S[(size_t)k * N + j] = std::sin(PI * (double)(((long long)(k + 1) * (j + 1)) % (2 * (N + 1))) / (double)(N + 1));
    std::vector<double> lam(N);
    for (int k = 0; k < N; k++) lam[k] = 2.0 * std::cos(PI * (double)(k + 1) / (double)(N + 1)) - 2.0;
    const double norm = 4.0 / (double)((N + 1) * (N + 1));
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) L[(size_t)i * N + j] = norm / (lam[i] + lam[j]);

    const size_t count = (size_t)std::llround(reach / RADIAL_STEP) + 1;
    radii.resize(count);
    for (size_t i = 0; i < count; i++) radii[i] = 0.0 + (reach - 0.0) * (double)i / (double)(count - 1);
    double dir[DIRECTIONS][2];
    for (int k = 0; k < DIRECTIONS; k++) {
        const double a = 2.0 * PI * ((double)k + 0.5) / (double)DIRECTIONS;
        dir[k][0] = std::cos(a);
        dir[k][1] = std::sin(a);
    }
    std::vector<double> total(count, 0.0), rhs((size_t)N * N), t1, t2;
    std::vector<double> ex((size_t)(N + 1) * N), ey((size_t)N * (N + 1));
    const int k_pass = passes;
    const int P = n + 2 * k_pass;
    std::vector<double> pad((size_t)P * P), nxt((size_t)P * P);
    auto bilin = [](const std::vector<double> &a, int w, int h, double gx, double gy) {
        const long long i = std::clamp((long long)std::floor(gx), 0LL, (long long)w - 2);
        const long long j = std::clamp((long long)std::floor(gy), 0LL, (long long)h - 2);
        const double fx = gx - (double)i, fy = gy - (double)j;
        return a[(size_t)i * h + j] * (1.0 - fx) * (1.0 - fy) + a[(size_t)(i + 1) * h + j] * fx * (1.0 - fy) +
               a[(size_t)i * h + j + 1] * (1.0 - fx) * fy + a[(size_t)(i + 1) * h + j + 1] * fx * fy;
    };
    for (int si = 0; si < SOURCE_SAMPLES; si++)
        for (int sj = 0; sj < SOURCE_SAMPLES; sj++) {
            const double src[2] = {(double)(n / 2) + ((double)si + 0.5) / SOURCE_SAMPLES,
                                   (double)(n / 2) + ((double)sj + 0.5) / SOURCE_SAMPLES};
            // CIC of the unit charge on the n x n grid (fields.rs:27-51), padded by `passes`.
            std::fill(pad.begin(), pad.end(), 0.0);
            const double u = src[0] - 0.5, v = src[1] - 0.5;
            const long long c0 = (long long)std::floor(u), r0 = (long long)std::floor(v);
            const double fu = u - (double)c0, fv = v - (double)r0;
            for (int a = 0; a < 2; a++)
                for (int bb = 0; bb < 2; bb++) {
                    const long long c = std::clamp(c0 + a, 0LL, (long long)n - 1), r = std::clamp(r0 + bb, 0LL, (long long)n - 1);
                    pad[(size_t)(c + k_pass) * P + (r + k_pass)] += (a ? fu : 1.0 - fu) * (bb ? fv : 1.0 - fv);
                }
            // `passes` passes of the 1-2-1 filter, x then y, zero beyond the padding (maxwell.rs:397-414).
            for (int p = 0; p < k_pass; p++)
                for (int axis = 0; axis < 2; axis++) {
                    for (int i = 0; i < P; i++)
                        for (int j = 0; j < P; j++) {
                            const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
                            double s = 0.5 * pad[(size_t)i * P + j];
                            if ((axis == 0 ? i : j) > 0) s += 0.25 * pad[(size_t)(i - di) * P + (j - dj)];
                            if ((axis == 0 ? i : j) + 1 < P) s += 0.25 * pad[(size_t)(i + di) * P + (j + dj)];
                            nxt[(size_t)i * P + j] = s;
                        }
                    pad.swap(nxt);
                }
            std::fill(rhs.begin(), rhs.end(), 0.0);
            for (int i = 0; i < P; i++)
                for (int j = 0; j < P; j++) rhs[(size_t)(i + m - k_pass) * N + (j + m - k_pass)] = -4.0 * PI * pad[(size_t)i * P + j];
            // phi = S ((S rhs S) * L) S; the field of a static charge is -grad phi.
            matmul(N, S, rhs, t1);
            matmul(N, t1, S, t2);
            for (size_t q = 0; q < t2.size(); q++) t2[q] *= L[q];
            matmul(N, S, t2, t1);
            matmul(N, t1, S, t2);
            const std::vector<double> &phi = t2;
            for (int i = 0; i <= N; i++)
                for (int j = 0; j < N; j++)
                    ex[(size_t)i * N + j] = -((i < N ? phi[(size_t)i * N + j] : 0.0) - (i > 0 ? phi[(size_t)(i - 1) * N + j] : 0.0));
            for (int i = 0; i < N; i++)
                for (int j = 0; j <= N; j++)
                    ey[(size_t)i * (N + 1) + j] = -((j < N ? phi[(size_t)i * N + j] : 0.0) - (j > 0 ? phi[(size_t)i * N + j - 1] : 0.0));
            for (size_t k = 0; k < count; k++) {
                double along = 0.0;
                for (int d = 0; d < DIRECTIONS; d++) {
                    const double x = src[0] + radii[k] * dir[d][0], y = src[1] + radii[k] * dir[d][1];
                    const double fx = bilin(ex, N + 1, N, x + m, y - 0.5 + m);
                    const double fy = bilin(ey, N, N + 1, x - 0.5 + m, y + m);
                    along += fx * dir[d][0] + fy * dir[d][1];
                }
                total[k] += along / DIRECTIONS;
            }
        }
    force.resize(count);
    for (size_t k = 0; k < count; k++) force[k] = total[k] / (double)(SOURCE_SAMPLES * SOURCE_SAMPLES);
}

double radial_at(const std::vector<double> &table, double r2, double cutoff) {  // close_range.rs:174-181
    if (r2 >= cutoff * cutoff) return 0.0;
    const double x = std::min(r2 * ((double)PAIR_SAMPLES / (cutoff * cutoff)), (double)PAIR_SAMPLES);
    const size_t i = (size_t)x;
    return table[i] + (table[i + 1] - table[i]) * (x - (double)i);
}

std::shared_ptr<HostTables> build_tables(double softening, int passes, double cutoff) {
    auto t = std::make_shared<HostTables>();
    t->softening = softening;
    t->passes = passes;
    t->cutoff = cutoff;
    t->reach = (int)std::ceil(cutoff);
    if (cutoff <= 0.0) return t;
    grid_force(passes, cutoff, t->grid_r, t->grid_f);
    const auto &r = t->grid_r, &fg = t->grid_f;
    // The correction's potential: the integral of (compact - grid) force inwards
    // from the cutoff (close_range.rs:151-171).
    std::vector<double> near(r.size()), pot(r.size(), 0.0);
    for (size_t i = 0; i < r.size(); i++) near[i] = r[i] * compact(r[i] * r[i], softening) - fg[i];
    for (size_t i = r.size() - 1; i-- > 0;) pot[i] = pot[i + 1] + 0.5 * (near[i + 1] + near[i]) * (r[i + 1] - r[i]);
    for (int i = 0; i <= PAIR_SAMPLES; i++) {
        const double r2 = 0.0 + (cutoff * cutoff - 0.0) * (double)i / (double)PAIR_SAMPLES;
        const double rs = std::sqrt(r2);
        const double grid_g = i == 0 ? fg[1] / r[1] : lerp_table(rs, r, fg) / rs;
        t->g.push_back(compact(r2, softening) - grid_g);
        t->potential.push_back(lerp_table(rs, r, pot));
    }
    t->g.push_back(0.0);
    t->potential.push_back(0.0);
    // The correction of a unit charge spread evenly over a cell: midpoint-rule
    // box sums through a summed-area table (close_range.rs:195-219).
    const int M = SQUARE_SAMPLES;
    const double h = 1.0 / M;
    t->half = std::ceil(cutoff) + 1.0;
    const int cnt = (int)std::llround((2.0 * t->half + 1.0) * M);
    std::vector<double> z(cnt);
    for (int k = 0; k < cnt; k++) z[k] = -t->half - 0.5 + ((double)k + 0.5) * h;
    const int W = cnt + 1;
    std::vector<double> sat((size_t)W * W * 3, 0.0);
    for (int a = 0; a < cnt; a++)
        for (int b = 0; b < cnt; b++) {
            const double r2 = z[a] * z[a] + z[b] * z[b];
            const double gr = radial_at(t->g, r2, cutoff);
            const double pt[3] = {radial_at(t->potential, r2, cutoff), gr * z[a], gr * z[b]};
            for (int k = 0; k < 3; k++)
                sat[((size_t)(a + 1) * W + b + 1) * 3 + k] =
                    pt[k] + sat[((size_t)a * W + b + 1) * 3 + k] + sat[((size_t)(a + 1) * W + b) * 3 + k] - sat[((size_t)a * W + b) * 3 + k];
        }
    const int n = cnt - M + 1;
    t->square_n = n;
    t->square.resize((size_t)n * n * 3);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < 3; k++)
                t->square[((size_t)i * n + j) * 3 + k] =
                    (sat[((size_t)(i + M) * W + j + M) * 3 + k] - sat[((size_t)i * W + j + M) * 3 + k] -
                     sat[((size_t)(i + M) * W + j) * 3 + k] + sat[((size_t)i * W + j) * 3 + k]) * h * h;
    // The square's potential averaged over each cell at offset (dc, dr): trapezoid
    // rule on the table (close_range.rs:223-242).
    const int R = t->reach, side = 2 * R + 1;
    t->ion_stencil.assign((size_t)side * side, 0.0);
    for (int i = 0; i < side; i++)
        for (int j = 0; j < side; j++) {
            const size_t a = (size_t)std::llround(((double)(i - R) - 0.5 + t->half) * M);
            const size_t b = (size_t)std::llround(((double)(j - R) - 0.5 + t->half) * M);
            double sum = 0.0;
            for (int x = 0; x <= M; x++)
                for (int y = 0; y <= M; y++) {
                    const double wx = (x == 0 || x == M) ? 0.5 / M : 1.0 / M;
                    const double wy = (y == 0 || y == M) ? 0.5 / M : 1.0 / M;
                    sum += wx * t->square[((a + x) * n + (b + y)) * 3] * wy;
                }
            t->ion_stencil[(size_t)i * side + j] = sum;
        }
    // The force channels summed over every cell offset of the reach window: the
    // close-range ion force of a uniform-density neighbourhood is one bilinear
    // lookup in this table (the offsets shift the table index by whole multiples
    // of SQUARE_SAMPLES, so every cell shares the electron's fractional offset).
    t->ion_a0 = (int)std::llround((t->half - 0.5) * M);
    t->ion_sum.assign((size_t)ION_SUM * ION_SUM * 2, 0.0);
    for (int i = 0; i < ION_SUM; i++)
        for (int j = 0; j < ION_SUM; j++)
            for (int ox = -R; ox <= R; ox++)
                for (int oy = -R; oy <= R; oy++) {
                    const long long a = t->ion_a0 + i + (long long)M * ox, b = t->ion_a0 + j + (long long)M * oy;
                    if (a < 0 || b < 0 || a >= n || b >= n) continue;
                    for (int ch = 0; ch < 2; ch++)
                        t->ion_sum[((size_t)i * ION_SUM + j) * 2 + ch] += t->square[((size_t)a * n + b) * 3 + 1 + ch];
                }
    return t;
}

}  // namespace

std::shared_ptr<const HostTables> host_tables(double softening, int passes, double cutoff) {
    static std::mutex lock;
    static std::map<std::tuple<double, int, double>, std::shared_ptr<const HostTables>> cache;
    const auto key = std::make_tuple(softening, passes, cutoff);
    {
        std::lock_guard<std::mutex> g(lock);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    auto t = build_tables(softening, passes, cutoff);
    std::lock_guard<std::mutex> g(lock);
    cache[key] = t;
    return t;
}

double HostStack::emf_at(double t) const {
    if (sources.empty()) return emf;
    double total = 0.0;
    for (auto &s : sources) total += s.second == 0.0 ? s.first : s.first * std::sin(2.0 * PI * s.second * t);
    return total;
}

int fdtd_substeps_of(const EmbSettings &s) {
    return std::max(1, (int)std::ceil(s.speed_of_light * s.dt * std::sqrt(2.0) / 0.9));
}

std::vector<float> sine_matrix(int n) {
    std::vector<float> out((size_t)n * n);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            out[(size_t)k * n + j] = (float)std::sin(PI * (double)(((long long)(k + 1) * (j + 1)) % (2 * (n + 1))) / (double)(n + 1));
    return out;
}

SineHalves sine_halves(int n) {
    const int m = (n + 1) / 2, h = n / 2;
    auto S = [n](long long k, long long j) {
        return (float)std::sin(PI * (double)(((k + 1) * (j + 1)) % (2 * (n + 1))) / (double)(n + 1));
    };
    SineHalves o;
    o.se.resize((size_t)m * m);
    o.se_t.resize((size_t)m * m);
    o.so.resize((size_t)h * h);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
            o.se[(size_t)i * m + j] = S(2 * i, j);
            o.se_t[(size_t)j * m + i] = S(2 * i, j);
        }
    for (int i = 0; i < h; i++)
        for (int j = 0; j < h; j++) o.so[(size_t)i * h + j] = S(2 * i + 1, j);
    return o;
}

std::string prepare_scene(const EmbSettings &s, const EmbScene &sc, HostScene &o) {
    if (sc.cols <= 0 || sc.rows <= 0) return "scene has no cells";
    if (!sc.kind || !sc.resistance || !sc.density || !sc.emf || !sc.orientation || !sc.frequency)
        return "scene lacks a material array";
    if (sc.n_electrons < 0 || (sc.n_electrons > 0 && (!sc.pos || !sc.vel))) return "electron arrays missing";
    if (sc.n_electrons > INT32_MAX / 2) return "too many electrons";
    if (s.smoothing_passes < 0 || s.smoothing_passes > 4) return "smoothing_passes must be 0..4";
    if (s.absorbing_margin < s.smoothing_passes || s.absorbing_margin < 1)
        return "absorbing_margin must be at least smoothing_passes";   // maxwell.rs:58
    if (s.carriers_per_cell <= 0 || s.cell_charge <= 0.0) return "carriers_per_cell and cell_charge must be positive";
    if (s.max_substeps < 1 || s.dt <= 0.0 || s.max_step <= 0.0) return "bad time stepping settings";
    if (s.electron_cutoff > 4.0) return "electron_cutoff above 4 cells is not supported";
    const int cols = sc.cols, rows = sc.rows;
    const long long cells = (long long)cols * rows;
    if (cells > (1LL << 26) || cols > 32767 || rows > 32767) return "grid too large";
    o.cols = cols;
    o.rows = rows;
    o.margin = s.absorbing_margin;
    o.nx = cols + 2 * o.margin;
    o.ny = rows + 2 * o.margin;
    o.fdtd_substeps = fdtd_substeps_of(s);
    o.h_fdtd = s.dt / (double)o.fdtd_substeps;
    o.n = sc.n_electrons;
    o.q = -s.cell_charge / (double)s.carriers_per_cell;   // world.rs:59-66
    o.mass = -o.q / s.carrier_charge_to_mass;
    auto kind = [&](long long c, long long r) -> int {
        const int k = sc.kind[c * rows + r];
        return (k == 1 || k == 2) ? k : 0;
    };
    auto in_bounds = [&](long long c, long long r) { return c >= 0 && r >= 0 && c < cols && r < rows; };
    auto conductive = [&](long long c, long long r) { return kind(c, r) != 0; };

    // Islands: flood fill in scan order through shared faces (connectivity.rs:20-62).
    o.labels.assign((size_t)cells, -1);
    {
        long long count = 0;
        std::vector<long long> stack;
        for (long long c = 0; c < cols; c++)
            for (long long r = 0; r < rows; r++) {
                if (!conductive(c, r) || o.labels[c * rows + r] >= 0) continue;
                o.labels[c * rows + r] = count;
                stack.push_back(c * rows + r);
                while (!stack.empty()) {
                    const long long k = stack.back();
                    stack.pop_back();
                    const long long x = k / rows, y = k % rows;
                    const long long nb[4][2] = {{x - 1, y}, {x + 1, y}, {x, y - 1}, {x, y + 1}};
                    for (auto &p : nb)
                        if (in_bounds(p[0], p[1]) && conductive(p[0], p[1]) && o.labels[p[0] * rows + p[1]] < 0) {
                            o.labels[p[0] * rows + p[1]] = count;
                            stack.push_back(p[0] * rows + p[1]);
                        }
                }
                count++;
            }
    }

    // Battery stacks: runs of battery cells along their own axis (battery.rs:100-147).
    o.stacks.clear();
    {
        auto horizontal = [&](long long c, long long r) {
            const int ori = sc.orientation[c * rows + r] % 4;
            return ori == 0 || ori == 2;
        };
        auto in_run = [&](long long c, long long r, bool h) {
            return in_bounds(c, r) && kind(c, r) == 2 && horizontal(c, r) == h;
        };
        std::vector<uint8_t> seen((size_t)cells, 0);
        for (long long c = 0; c < cols; c++)
            for (long long r = 0; r < rows; r++) {
                if (sc.kind[c * rows + r] != 2 || seen[c * rows + r]) continue;
                const bool h = horizontal(c, r);
                const long long sx = h ? 1 : 0, sy = h ? 0 : 1;
                long long fc = c, fr = r;
                while (in_run(fc - sx, fr - sy, h)) {
                    fc -= sx;
                    fr -= sy;
                }
                HostStack st;
                long long xc = fc, xr = fr;
                while (in_run(xc, xr, h)) {
                    st.cells.push_back({(int)xc, (int)xr});
                    seen[xc * rows + xr] = 1;
                    xc += sx;
                    xr += sy;
                }
                st.emf = 0.0;
                st.internal_resistance = 0.0;
                st.amplitude = 0.0;
                for (auto &x : st.cells) {
                    const long long k = (long long)x.first * rows + x.second;
                    const int ori = sc.orientation[k] % 4;
                    const double sign = (ori == 0 || ori == 1) ? 1.0 : -1.0;
                    st.sources.push_back({sign * sc.emf[k], sc.frequency[k]});
                }
                for (auto &src : st.sources) st.emf += src.first;
                for (auto &x : st.cells) st.internal_resistance += sc.resistance[(long long)x.first * rows + x.second];
                for (auto &src : st.sources) st.amplitude += std::fabs(src.first);
                st.low[0] = (int)(fc - sx);
                st.low[1] = (int)(fr - sy);
                st.high[0] = (int)xc;
                st.high[1] = (int)xr;
                o.stacks.push_back(st);
            }
    }
    // Per stack: the terminal blocks a trim reads and the cells whose current it
    // averages (battery.rs:150-190, 240-252; sim.rs:403-436).
    o.stack_info.clear();
    o.stack_cells.clear();
    std::vector<int32_t> stack_of((size_t)cells, -1);
    for (size_t k = 0; k < o.stacks.size(); k++) {
        const HostStack &st = o.stacks[k];
        for (auto &x : st.cells) stack_of[(size_t)x.first * rows + x.second] = (int32_t)k;
        StackInfo I;
        std::memset(&I, 0, sizeof(I));
        auto wired = [&](const int t[2]) { return in_bounds(t[0], t[1]) && kind(t[0], t[1]) == 1; };
        const int *terms[2] = {st.low, st.high};
        for (int side = 0; side < 2; side++) {
            const int *t = terms[side];
            const int off = (int)o.stack_cells.size();
            if (wired(t)) {
                const long long island = o.labels[(long long)t[0] * rows + t[1]];
                for (long long x = std::max(t[0] - 1, 0); x < std::min(t[0] + 2, cols); x++)
                    for (long long y = std::max(t[1] - 1, 0); y < std::min(t[1] + 2, rows); y++)
                        if (o.labels[x * rows + y] == island && kind(x, y) == 1) o.stack_cells.push_back((int32_t)(x * rows + y));
            }
            const int cnt = (int)o.stack_cells.size() - off;
            if (side == 0) {
                I.low_off = off;
                I.low_n = cnt;
            } else {
                I.high_off = off;
                I.high_n = cnt;
            }
        }
        I.cur_off = (int)o.stack_cells.size();
        for (auto &x : st.cells) o.stack_cells.push_back((int32_t)((long long)x.first * rows + x.second));
        for (int side = 0; side < 2; side++)
            if (wired(terms[side])) o.stack_cells.push_back((int32_t)((long long)terms[side][0] * rows + terms[side][1]));
        I.cur_n = (int)o.stack_cells.size() - I.cur_off;
        I.dir_x = (float)((st.high[0] > st.low[0]) - (st.high[0] < st.low[0]));
        I.dir_y = (float)((st.high[1] > st.low[1]) - (st.high[1] < st.low[1]));
        I.amplitude = (float)st.amplitude;
        I.internal_resistance = (float)st.internal_resistance;
        I.both_wired = wired(st.low) && wired(st.high);
        o.stack_info.push_back(I);
    }

    // Per cell: drag, flags, battery cells (physics.rs:17-21, battery.rs:210-228).
    const double per_ohm = s.drag_per_resistance * s.cell_charge * s.carrier_charge_to_mass;
    o.info.resize((size_t)cells);
    o.density.resize((size_t)cells);
    o.bats.clear();
    o.ac_freq.clear();
    for (long long k = 0; k < cells; k++) {
        const int kd = kind(k / rows, k % rows);
        CellInfo ci;
        ci.damping = (float)(sc.resistance[k] * sc.density[k] * per_ohm);
        ci.flags = (kd != 0 ? F_CONDUCTIVE : 0u) | (kd == 1 ? F_WIRE : 0u);
        ci.flags |= (uint32_t)(sc.orientation[k] & 3) << F_ORI_SHIFT;
        if (sc.emf[k] != 0.0) {
            BatteryCell b;
            b.emf = (float)sc.emf[k];
            b.frequency = (float)sc.frequency[k];
            b.stack = stack_of[(size_t)k];
            b.ac_slot = -1;
            if (sc.frequency[k] > 0.0) {
                b.ac_slot = (int32_t)o.ac_freq.size();
                o.ac_freq.push_back(sc.frequency[k]);
            }
            o.bats.push_back(b);
            ci.flags |= F_BATTERY | ((uint32_t)o.bats.size() << F_BAT_SHIFT);
        }
        o.info[(size_t)k] = ci;
        o.density[(size_t)k] = (float)sc.density[k];
    }
    if (o.bats.size() >= (1u << (32 - F_BAT_SHIFT)) - 1) return "too many battery cells";

    // Cells whose whole reach window lies on the grid with one ion density.
    if (s.electron_cutoff > 0.0) {
        const int R = (int)std::ceil(s.electron_cutoff);
        for (long long c = R; c + R < cols; c++)
            for (long long r = R; r + R < rows; r++) {
                const double d = sc.density[c * rows + r];
                if (!(d > 0.0)) continue;
                bool same = true;
                for (long long dc = -R; dc <= R && same; dc++)
                    for (long long dr = -R; dr <= R && same; dr++) same = sc.density[(c + dc) * rows + (r + dr)] == d;
                if (same) o.info[(size_t)(c * rows + r)].flags |= F_ION_UNIFORM;
            }
    }

    // Static charge: ions spread over their cells plus the fixed charges' CIC
    // (fields.rs:91-112; maxwell.rs:307-310), in units of 2^-40 electron charges.
    {
        std::vector<double> v((size_t)cells), out((size_t)cells);
        for (long long k = 0; k < cells; k++) v[(size_t)k] = sc.density[k] * s.cell_charge;
        for (int axis = 0; axis < 2; axis++) {
            const int n = axis == 0 ? cols : rows;
            for (long long k = 0; k < cells; k++) out[(size_t)k] = 0.75 * v[(size_t)k];
            for (long long c = 0; c < cols; c++)
                for (long long r = 0; r < rows; r++) {
                    const long long i = axis == 0 ? c : r;
                    const long long lo = i > 0 ? i - 1 : i, hi = i + 1 < n ? i + 1 : i;
                    const double here = v[(size_t)(c * rows + r)] / 8.0;
                    out[(size_t)(axis == 0 ? lo * rows + r : c * rows + lo)] += here;
                    out[(size_t)(axis == 0 ? hi * rows + r : c * rows + hi)] += here;
                }
            v.swap(out);
        }
        for (int64_t f = 0; f < sc.n_fixed; f++) {
            const double u = sc.fixed_pos[2 * f] - 0.5, w = sc.fixed_pos[2 * f + 1] - 0.5;
            const long long c0 = (long long)std::floor(u), r0 = (long long)std::floor(w);
            const double fu = u - (double)c0, fv = w - (double)r0;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) {
                    const long long c = std::clamp(c0 + a, 0LL, (long long)cols - 1), r = std::clamp(r0 + b, 0LL, (long long)rows - 1);
                    v[(size_t)(c * rows + r)] += (a ? fu : 1.0 - fu) * (b ? fv : 1.0 - fv) * sc.fixed_charge[f];
                }
        }
        const double unit = s.cell_charge / (double)s.carriers_per_cell;
        o.static_rho.resize((size_t)cells);
        for (long long k = 0; k < cells; k++) o.static_rho[(size_t)k] = std::llround(v[(size_t)k] / unit * RHO_SCALE);
    }

    // Close-range potential of all ions per cell (close_range.rs:379-403).
    o.ion_pot.assign((size_t)cells, 0.0f);
    if (s.electron_cutoff > 0.0) {
        auto t = host_tables(s.electron_softening, s.smoothing_passes, s.electron_cutoff);
        const int R = t->reach, side = 2 * R + 1;
        const double scale = s.coulomb_k * s.cell_charge;
        for (long long c = 0; c < cols; c++)
            for (long long r = 0; r < rows; r++) {
                double sum = 0.0;
                for (int dc = -R; dc <= R; dc++)
                    for (int dr = -R; dr <= R; dr++)
                        if (in_bounds(c + dc, r + dr))
                            sum += sc.density[(c + dc) * rows + (r + dr)] * t->ion_stencil[(size_t)(dc + R) * side + (dr + R)];
                o.ion_pot[(size_t)(c * rows + r)] = (float)(scale * sum);
            }
    }

    // Absorbing margin: exp(-strength d^3 h) with d the larger of the two axis
    // depths (maxwell.rs:111-134), stored per axis since the factor is the smaller
    // of the two per-axis factors.
    {
        const double m = (double)o.margin;
        auto factor = [&](double coord, int n) {
            const double outside = std::max(std::max(m - coord, coord - ((double)n - 1.0 - m)), 0.0);
            const double d = std::min(outside / m, 1.0);
            return (float)std::exp(-s.absorbing_strength * (d * d * d) * o.h_fdtd);
        };
        auto axis = [&](int len, double offset, int n) {
            std::vector<float> f((size_t)len);
            for (int i = 0; i < len; i++) f[(size_t)i] = factor((double)i + offset, n);
            return f;
        };
        o.damp[0] = axis(o.nx + 1, -0.5, o.nx);
        o.damp[1] = axis(o.ny, 0.0, o.ny);
        o.damp[2] = axis(o.nx, 0.0, o.nx);
        o.damp[3] = axis(o.ny + 1, -0.5, o.ny);
        o.damp[4] = axis(o.nx + 1, -0.5, o.nx);
        o.damp[5] = axis(o.ny + 1, -0.5, o.ny);
    }
    // Inverse Laplacian with the transforms' normalisation (maxwell.rs:60-67).
    {
        std::vector<double> kx(o.nx), ky(o.ny);
        for (int k = 1; k <= o.nx; k++) kx[k - 1] = 2.0 * std::cos(PI * (double)k / (double)(o.nx + 1)) - 2.0;
        for (int k = 1; k <= o.ny; k++) ky[k - 1] = 2.0 * std::cos(PI * (double)k / (double)(o.ny + 1)) - 2.0;
        const double norm = 4.0 / (double)((o.nx + 1) * (o.ny + 1));
        o.inv_lap.resize((size_t)o.nx * o.ny);
        for (int i = 0; i < o.nx; i++)
            for (int j = 0; j < o.ny; j++) o.inv_lap[(size_t)i * o.ny + j] = (float)(norm / (kx[i] + ky[j]));
    }
    // Gains (NaN = no entry) and electrons.
    o.gains.assign(o.stacks.size(), 1.0f);
    o.has_gain.assign(o.stacks.size(), 0);
    for (int k = 0; k < sc.n_gains && k < (int)o.stacks.size(); k++)
        if (sc.gains && !std::isnan(sc.gains[k])) {
            o.gains[(size_t)k] = (float)sc.gains[k];
            o.has_gain[(size_t)k] = 1;
        }
    o.cxy.resize((size_t)o.n);
    o.fx.resize((size_t)o.n);
    o.fy.resize((size_t)o.n);
    o.vx.resize((size_t)o.n);
    o.vy.resize((size_t)o.n);
    for (int64_t i = 0; i < o.n; i++) {
        const double x = sc.pos[2 * i], y = sc.pos[2 * i + 1];
        if (!std::isfinite(x) || !std::isfinite(y)) return "electron position not finite";
        const double cx = std::floor(x), cy = std::floor(y);
        if (cx < 0 || cy < 0 || cx >= cols || cy >= rows) return "electron outside the grid";
        float fx = (float)(x - cx), fy = (float)(y - cy);
        if (fx >= 1.0f) fx = 0.99999994f;   // x just below a face: stay in its cell
        if (fy >= 1.0f) fy = 0.99999994f;
        o.cxy[(size_t)i] = pack_cell((int)cx, (int)cy);
        o.fx[(size_t)i] = fx;
        o.fy[(size_t)i] = fy;
        o.vx[(size_t)i] = (float)sc.vel[2 * i];
        o.vy[(size_t)i] = (float)sc.vel[2 * i + 1];
    }
    return "";
}

}  // namespace emb
