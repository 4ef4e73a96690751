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

std::mutex g_tables_lock;
std::map<std::tuple<double, int, double>, std::shared_ptr<const HostTables>> g_tables;
}  // namespace

std::shared_ptr<const HostTables> find_tables(double softening, int passes, double cutoff) {
    std::lock_guard<std::mutex> g(g_tables_lock);
    auto it = g_tables.find(std::make_tuple(softening, passes, cutoff));
    return it == g_tables.end() ? nullptr : it->second;
}

void store_tables(std::shared_ptr<const HostTables> t) {
    std::lock_guard<std::mutex> g(g_tables_lock);
    g_tables[std::make_tuple(t->softening, t->passes, t->cutoff)] = std::move(t);
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
    const int pm = sine_half_pitch(m), ph = sine_half_pitch(h);
    auto S = [n](long long k, long long j) {
        return (float)std::sin(PI * (double)(((k + 1) * (j + 1)) % (2 * (n + 1))) / (double)(n + 1));
    };
    // Each half square, zero-padded to a side of sine_half_pitch.
    SineHalves o;
    o.se.assign((size_t)pm * pm, 0.0f);
    o.se_t.assign((size_t)pm * pm, 0.0f);
    o.so.assign((size_t)ph * ph, 0.0f);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
            o.se[(size_t)i * pm + j] = S(2 * i, j);
            o.se_t[(size_t)j * pm + i] = S(2 * i, j);
        }
    for (int i = 0; i < h; i++)
        for (int j = 0; j < h; j++) o.so[(size_t)i * ph + j] = S(2 * i + 1, j);
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
        auto t = find_tables(s.electron_softening, s.smoothing_passes, s.electron_cutoff);
        if (!t) return "close-range tables not built for these settings";
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
        const int pny = poisson_pitch(o.ny);
        o.inv_lap.assign((size_t)o.nx * pny, 0.0f);
        for (int i = 0; i < o.nx; i++)
            for (int j = 0; j < o.ny; j++) o.inv_lap[(size_t)i * pny + j] = (float)(norm / (kx[i] + ky[j]));
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
