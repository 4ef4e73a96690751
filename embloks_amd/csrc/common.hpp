// common.hpp - device data model shared by the kernels and the host runtime.
//
// One `DevScene` describes one resident scene: its dimensions, physics
// constants (FP32) and pointers into its device arena. Kernels receive an
// array of them plus a per-kernel block prefix (blocks of scene s are
// prefix[s] .. prefix[s+1]), so the same kernel advances one scene (the plain
// path) or every resident scene in one launch (the batched path).
//
// Layouts follow the reference (crates/embloks-core/src/maxwell.rs:12-15,
// fields.rs:21-22): x-major arrays, index = i * height + j.
//   Ex (nx+1, ny)   Ey (nx, ny+1)   Bz (nx+1, ny+1)   node arrays (nx, ny)
//   cell arrays (cols, rows), cell = c * rows + r.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
typedef float2 lerp2;
typedef float2 force2;
typedef float4 snap4;
#else
#define __host__
#define __device__
struct lerp2 {
    float x, y;
};
struct force2 {
    float x, y;
};
struct snap4 {
    float x, y, z, w;
};
#endif

namespace emb {

// Fixed-point scales of the deterministic deposit: electron charge in units of
// 2^-44 of one electron (CIC weights carry 22 fractional bits per axis), q*v in
// units of 2^-28 (charge velocity of one electron). The x pass of the smoothing
// multiplies by at most 4^passes and the y pass by as much again: with two
// passes (x 256) a node holds at most 2^(63-44-8) = 2048 electron charges of
// either sign. Three or four passes run the y pass in FP64 instead (the int64
// grid itself stays exact); a node past the int64 headroom of the passes done
// in integers sets SceneCounters::charge_overflow (EMB_E_NUMERIC at the next
// sync) instead of wrapping silently.
constexpr int W_BITS = 22;
constexpr double RHO_SCALE = 17592186044416.0;  // 2^44
constexpr double J_SCALE = 268435456.0;         // 2^28
constexpr float J_FROM_W = 1.0f / 65536.0f;     // 2^28 / 2^44

// Electron positions are kept as an integer cell (two int16 packed in an int32,
// x in the low half) plus FP32 fractions in [0, 1): 2^-24 cells of resolution
// anywhere on the grid, where a plain FP32 coordinate near x = 1000 resolves
// only 6e-5 cells.
__host__ __device__ inline int32_t pack_cell(int cx, int cy) { return (int32_t)(((uint32_t)(uint16_t)cx) | ((uint32_t)cy << 16)); }
__host__ __device__ inline int cell_x(int32_t p) { return (int)(int16_t)(p & 0xffff); }
__host__ __device__ inline int cell_y(int32_t p) { return (int)(p >> 16); }

constexpr int MAX_STACKS_PER_BLOCK = 64;
constexpr int SCAN_TILE = 1024;   // cells per block of the cell-list scan
constexpr int TICK_TABLE = 64;    // ticks of host-computed AC phases / EMF targets
constexpr int PUSH_TILE = 8;      // cells per side of a push tile
constexpr int PUSH_BLOCKS_MAX = 2048;   // tiled-push blocks per scene: four rounds of the chip's 512 resident blocks
constexpr int DEP_TILE = 24;      // cells per side of a deposit + smoothing tile
constexpr int CELL_MODE_MIN = 6;  // a push tile with a cell this crowded runs the cell mode
constexpr int HEAVY_CELL = 64;    // a cell this crowded is pushed by k_push_heavy, spread over blocks
constexpr int TRIM_BLOCKS = 4;    // blocks per battery stack in the trim (trim_block, 1024 threads each)
constexpr int REST_PHASES = 7;    // work phases of one extra electron substep in k_rest
constexpr int REST_THREADS = 1024; // threads per k_rest block (the push tiles' layout)
constexpr int FDTD_STRIP = 60;    // corner columns per FDTD wave (64 lanes: two halo lanes each side)
constexpr int FDTD_WAVES = 4;     // FDTD waves per block (adjacent strips)
constexpr int FDTD_AHEAD = 4;     // corner rows an FDTD wave has loads in flight for
constexpr int FDTD_MIN_BLOCKS = 4096;   // FDTD blocks a scene keeps while its rows per wave double
// Corner rows one FDTD wave walks: 32 where the grid still gives FDTD_MIN_BLOCKS
// blocks (each wave re-reads 3 rows of its neighbour above: 9% at 32 rows), fewer,
// down to 4, on small grids, whose launch needs the waves more.
__host__ __device__ inline int fdtd_rows(int nx, int ny) {
    const long long sblocks = ((ny + 1 + FDTD_STRIP - 1) / FDTD_STRIP + FDTD_WAVES - 1) / FDTD_WAVES;
    int rows = 32;
    while (rows > 4 && sblocks * ((nx + 1 + rows - 1) / rows) < FDTD_MIN_BLOCKS) rows >>= 1;
    return rows;
}
__host__ __device__ inline int rest_sync_len(int max_substeps) {
    return 1 + REST_PHASES * (max_substeps > 1 ? max_substeps - 1 : 0);
}
constexpr int ION_SUM = 33;       // samples per axis of the summed ion table
// Row pitch of the Poisson arrays of a scene with ny nodes per row: a multiple of
// four floats, so every row starts 16-byte aligned (the padding stays zero).
__host__ __device__ inline int poisson_pitch(int ny) { return (ny + 3) & ~3; }
// Side of a stored half of the DST-I matrix (sine_halves) of m rows: zero-padded
// to a multiple of 16, the products' K slice.
__host__ __device__ inline int sine_half_pitch(int m) { return (m + 15) & ~15; }
constexpr int PAIR_SAMPLES = 4096;   // samples of the radial pair table (close_range.rs:30)

// Cell flags (CellInfo.flags).
constexpr uint32_t F_CONDUCTIVE = 1u;
constexpr uint32_t F_WIRE = 2u;
constexpr uint32_t F_BATTERY = 4u;
constexpr int F_ORI_SHIFT = 3;      // 2 bits of orientation
constexpr uint32_t F_ION_UNIFORM = 32u;   // the (2 reach + 1)^2 cells around share one ion density
constexpr int F_BAT_SHIFT = 8;      // battery-cell index + 1 in the upper bits

struct CellInfo {
    float damping;      // drag rate R * density * cell_charge * q/m (physics.rs:17-21)
    uint32_t flags;
};

struct BatteryCell {
    float emf;          // |EMF| of the cell, along its orientation (battery.rs:210-228)
    float frequency;    // 0 = DC
    int32_t stack;      // index of its stack, -1 if none
    int32_t ac_slot;    // column in the per-tick phase table, -1 for DC
};

struct StackInfo {
    int32_t low_off, low_n, high_off, high_n;   // terminal-block cell lists
    int32_t cur_off, cur_n;                     // cells whose current is averaged
    float dir_x, dir_y;
    float amplitude, internal_resistance;
    int32_t both_wired;
    int32_t pad;
};

// Per-scene device counters and reductions.
struct SceneCounters {
    uint32_t tick;            // ticks completed (parity selects ping-pong buffers)
    uint32_t vmax_bits[2];    // max |v| (float bits) of the electrons, by tick parity
    int32_t heavy_n;          // cells listed in heavy_cells by the last sort
    int32_t substeps;         // electron substeps of the current tick
    int32_t rest_error;       // set when a k_rest wait ran past its limit (the tick's later substeps were dropped)
    int32_t charge_overflow;  // set when a node's fixed-point charge came near the int64 range in the smoothing
    int32_t dense_n;          // tiles listed in dense_tiles by the last sort
};

struct DevScene {
    // --- shape
    int32_t cols, rows, nx, ny, margin;
    int32_t n;                // electrons
    int32_t n_stacks;
    int32_t fdtd_substeps;
    int32_t max_substeps;
    int32_t reach;            // ceil(cutoff)
    int32_t square_n;
    int32_t close_on;
    int32_t pny;              // row pitch of the Poisson arrays (w1, w2, phi, inv_lap): ny rounded up to 4
    int32_t pad0;
    // --- constants (FP32)
    float dt, max_step, restitution;
    float qm;                 // q / m
    float inv_mass;
    float q;                  // electron charge
    float k_pair;             // k q q (close_range.rs:366)
    float k_pair_fx;          // k q q 2^-24: the pair sum over 24-bit fixed-point offsets back to a force
    float k_ion_q;            // k * cell_charge * q (close_range.rs:348)
    float k_q;                // k * q (charges_potential, close_range.rs:463)
    float cutoff2, g_scale, half, square_samples;
    float c2h;                // h * c^2 of one FDTD substep (maxwell.rs:257)
    float half_h;             // 0.5 * h (maxwell.rs:254)
    float jscale;             // 0.5 * 4 pi k h (maxwell.rs:240-248)
    float four_pi_k;
    float trim_rate, gain_low, gain_high;
    float rho_from_fixed;     // |q_e| / 2^44 / 4^(2 passes): int64 charge -> FP32
    float j_from_fixed;       // q_e / 2^28 / 4^(2 passes)
    float i_from_fixed;       // q_e / 2^28 (cell-binned current)
    float pair_s2;            // g_scale 2^-48: a squared 24-bit fixed-point offset to a table coordinate
    float alpha_current;      // min(dt / current_smoothing_time, 1) (sim.rs:256)
    float alpha_display;      // 1 - exp(-dt / display_smoothing_time), 1 when that time is 0 (sim.rs:257-261)
    int32_t passes;
    int32_t pad1;
    // --- fields (FP32)
    float *ex, *ey, *bz;
    float *ex2, *ey2, *bz2;   // second buffers of the FDTD step (written, then read back by k_correct)
    float *rho, *jx, *jy;     // smoothed sources on nodes
    float *w1, *w2, *phi;     // Poisson work: 2 channels each, channel stride nodes
    const float *damp_ex_x, *damp_ex_y, *damp_ey_x, *damp_ey_y, *damp_bz_x, *damp_bz_y;
    // halves of the DST-I matrices (sine_halves): even rows of Sx (ceil(nx/2)^2),
    // odd rows of Sx (floor(nx/2)^2), the even rows of Sy transposed, odd rows of Sy
    const float *sx_e, *sx_o, *sy_et, *sy_o;
    const float *inv_lap;             // norm / (lambda_x + lambda_y), (nx, ny) (maxwell.rs:60-67)
    // --- electrons (SoA, upload order): cell + fraction, velocity
    int32_t *cxy;
    float *fx, *fy, *vx, *vy;
    // --- cell list
    int32_t *scell;           // per sorted slot: flat cell index (c * rows + r)
    snap4 *snap;                  // electrons in sorted order (snapshot of the substep start): x, y, vx, vy
    force2 *pforce;           // per sorted slot: close-range + ion force of the cell-mode electrons (k_push_tiled)
    int32_t *order;           // electron index per sorted slot
    int32_t *ecell, *erank;   // per electron: cell and rank within it
    int32_t *count;           // per cell (zero between sorts)
    int32_t *lstart;          // per cell: start within its scan tile
    int32_t *toff;            // per scan tile: its electron count (k_scan), then its offset (k_rest's queue)
    int32_t *toff_pre;        // per scan tile: offset of the tile (the main sort's top-level scan)
    int32_t *cstart;          // per cell + 1: absolute start in the sorted order
    uint8_t *tile_dense;      // per push tile: a cell holds >= CELL_MODE_MIN electrons
    int32_t *dense_tiles;     // the tiles with tile_dense set (ctr->dense_n of them), in no particular order
    int32_t *heavy_cells;     // cells holding >= HEAVY_CELL electrons (ctr->heavy_n of them)
    // --- deposit
    long long *cur;           // [2 * cells]: cell-binned sum of v, units 2^-28 (x q_e)
    float *cur_smooth;        // [2 * cells]: exponential average of the cell current (channel-planar)
    float *display;           // [3 * cells]: exponential average of E (x, y) and B at cell centres
    const long long *static_rho;   // ions + fixed charges per cell, units 2^-44 electron charges
    long long *dep_dump;      // optional [3 * cells]: unsmoothed electron deposit (diagnostics)
    // --- materials, tables, batteries
    const CellInfo *info;
    const float *density;     // per cell (ion density)
    const float *ion_pot;     // per cell: close-range potential of all ions
    const float *g_table;     // PAIR_SAMPLES + 4: the reference's PAIR_SAMPLES + 2 samples, then two zeros
    const lerp2 *g_lerp;      // PAIR_SAMPLES + 4: (g[i] - i d, d), d = g[i+1] - g[i], below the cutoff, then (0, 0)
    const float *square;      // square_n^2 * 3
    const lerp2 *square_f2;   // square_n^2: the force channels (1, 2) of the square table as pairs
    const float *ion_sum;     // ION_SUM^2 * 2: the square table's force channels summed over
                              // the (2 reach + 1)^2 cell offsets (uniform-density neighbourhoods)
    int32_t ion_a0;           // square-table index of the home cell's low corner offset
    const BatteryCell *bats;
    const StackInfo *stacks;
    const int32_t *stack_cells;
    float *gains;             // per stack
    int32_t *has_gain;        // per stack
    float *trim_part;         // per stack and trim block: terminal potential sums (low, high)
    float *stack_read;        // per stack: the last trim's v_low, v_high, and 1 if it trimmed (else 0)
    int32_t *rest_sync;       // [rest_sync_len(max_substeps)]: k_rest's work ticket, then items done per
                              // (substep, phase); zeroed by the trim at the start of every tick
    const float *phase_tab;   // [TICK_TABLE][n_ac]: frac(f t) per AC cell and tick
    const float *target_tab;  // [TICK_TABLE][n_stacks]: EMF target per tick
    int32_t n_ac;
    int32_t pad2;
    SceneCounters *ctr;
};

// Kernel classes with their own block prefix over the batch.
enum KernelClass {
    KC_ELECTRONS = 0,   // 256 electrons per block
    KC_SCAN,            // SCAN_TILE cells per block
    KC_CELLS,           // 256 cells per block
    KC_SMOOTH,          // 32 x 32 tiles of the smoothing region
    KC_FDTD,            // FDTD_WAVES strips x fdtd_rows() rows of the (nx+1) x (ny+1) corner space
    KC_GEMM,            // 2 channels x 64 x 64 output tiles of nx x ny
    KC_GEMM_SMALL,      // the same with small tiles (a batch with few 64 x 64 tiles runs these)
    KC_NODES,           // 256 corners of (nx+1) x (ny+1) per block
    KC_STACKS,          // TRIM_BLOCKS blocks per battery stack (at least one stack per scene)
    KC_SCENE,           // one block per scene
    KC_PUSH,            // PUSH_TILE x PUSH_TILE cell tiles
    KC_HEAVY,           // blocks sharing the scene's heavy cells (one per 16k electrons, at most 256)
    KC_REST,            // k_rest work-queue blocks (one per 1024 electrons, at most 256: launching more costs
                        // ticks that need no second substep more than it gains)
    KC_COUNT
};

struct Batch {
    const DevScene *scenes;     // device array
    const int32_t *prefix;      // device: KC_COUNT * (n_scenes + 1)
    int32_t n_scenes;
    int32_t total[KC_COUNT];    // blocks per class over the batch
    int32_t even_fdtd;          // some scene takes an even number of FDTD substeps (its tick ends in k_finish)
    int32_t top_small;          // every scene has at most 256 scan tiles (k_scatter scans their totals itself)
};

}  // namespace emb
