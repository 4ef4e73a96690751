// setup.hpp - host-side FP64 preparation of a scene for the device.
//
// Everything here runs once per scene (or once per table parameter set) on the
// host, in FP64, and is uploaded as FP32 / int64 arrays:
//   * islands and battery stacks (connectivity.rs:20-62, battery.rs:100-190);
//   * per-cell drag, static charge (ions + fixed charges) in fixed point, the
//     ions' close-range potential (physics.rs:17-21, fields.rs:91-112,
//     close_range.rs:379-403);
//   * absorbing-margin damping profiles, DST-I matrices and the inverse
//     Laplacian (maxwell.rs:56-134).
#pragma once

#include <stdint.h>

#include <memory>
#include <vector>

#include "../../include/embloks_hip.h"
#include "common.hpp"

namespace emb {

struct HostTables {
    double softening, cutoff;
    int passes;
    std::vector<double> grid_r, grid_f;   // grid_force
    std::vector<double> g, potential;     // PAIR_SAMPLES + 2
    std::vector<double> square;           // n * n * 3
    int square_n = 0;
    double half = 0.0;
    std::vector<double> ion_stencil;      // (2 reach + 1)^2
    int reach = 0;
    std::vector<double> ion_sum;          // ION_SUM^2 * 2: force channels summed over cell offsets
    int ion_a0 = 0;
};

// close_range::tables per parameter set, cached for the process. They are built
// on the device (tables.hip: ensure_tables) before a scene that needs them is
// prepared; find_tables returns null for a set not built yet.
std::shared_ptr<const HostTables> find_tables(double softening, int passes, double cutoff);
void store_tables(std::shared_ptr<const HostTables> t);

struct HostStack {
    std::vector<std::pair<int, int>> cells;
    int low[2], high[2];
    double emf, internal_resistance, amplitude;
    std::vector<std::pair<double, double>> sources;   // (signed EMF, frequency) per cell
    double emf_at(double t) const;                     // battery.rs:86-96
};

// Everything about one scene that the device needs, prepared in FP64.
struct HostScene {
    int cols = 0, rows = 0, margin = 0, nx = 0, ny = 0, fdtd_substeps = 1;
    int64_t n = 0;
    double h_fdtd = 0.0;
    double q = 0.0, mass = 0.0;
    std::vector<int64_t> labels;
    std::vector<HostStack> stacks;
    std::vector<StackInfo> stack_info;
    std::vector<int32_t> stack_cells;
    std::vector<CellInfo> info;
    std::vector<BatteryCell> bats;
    std::vector<double> ac_freq;              // per AC slot
    std::vector<float> density, ion_pot;
    std::vector<long long> static_rho;
    std::vector<float> damp[6];               // ex_x, ex_y, ey_x, ey_y, bz_x, bz_y
    std::vector<float> inv_lap;
    std::vector<float> gains;
    std::vector<int32_t> has_gain;
    std::vector<int32_t> cxy;                 // electrons: cell + fraction, velocity
    std::vector<float> fx, fy, vx, vy;
};

// Returns an empty string on success, else what is wrong with the scene.
std::string prepare_scene(const EmbSettings &s, const EmbScene &scene, HostScene &out);

// The DST-I matrix sin(pi (k+1)(j+1) / (n+1)), FP64 rounded to FP32.
std::vector<float> sine_matrix(int n);

// What the symmetry S[k][n-1-j] = (-1)^k S[k][j] of that matrix S leaves of it:
// the even rows se[i][j] = S[2i][j] over j < m (m x m, m = ceil(n/2)), their
// transpose, and the odd rows so[i][j] = S[2i+1][j] over j < h (h x h,
// h = floor(n/2); so is symmetric). S x is then se p (rows 2i) and so q (rows
// 2i+1) with p_j = x_j + x_{n-1-j}, q_j = x_j - x_{n-1-j} (j < h) and, for odd n,
// p_h = x_h. Each is stored square and zero-padded to sine_half_pitch of its
// side (row-major, that pitch), so a product's K can run in whole slices.
struct SineHalves {
    std::vector<float> se, se_t, so;
};
SineHalves sine_halves(int n);

// Electron substeps of the FDTD step (maxwell.rs:376-378).
int fdtd_substeps_of(const EmbSettings &s);

}  // namespace emb
