/*
 * embloks_hip.h - C ABI of the MI355X (gfx950) physics backend for embloks scenes.
 *
 * This library replaces the reference's GPU backend `GpuSim`
 * (crates/embloks-gpu/src/sim.rs:180-850) behind a plain C interface: upload a
 * scene, step it, read back fields and electrons. All structs are POD; host
 * arrays are owned by the caller and copied in or out; device memory is owned
 * by the library. Every call returns 0 on success or a negative EMB_E* code,
 * with a message in emb_last_error() (thread-local).
 *
 * Array conventions follow the reference: per-cell arrays are x-major,
 * index = c * rows + r (crates/embloks-core/src/fields.rs:21-22). Yee arrays are
 * shaped like `Maxwell`'s (crates/embloks-core/src/maxwell.rs:76-79):
 *   Ex (nx + 1, ny), Ey (nx, ny + 1), Bz (nx + 1, ny + 1), potential (nx, ny),
 * with nx = cols + 2 * margin, ny = rows + 2 * margin, index = i * width_y + j.
 *
 * Handles are thread-compatible, not thread-safe: one host thread drives one
 * simulation at a time (the reference's GpuSim is used the same way,
 * crates/embloks-web/src/app.rs:168-175).
 */
#ifndef EMBLOKS_HIP_H
#define EMBLOKS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    EMB_OK = 0,
    EMB_E_INVALID = -1,   /* bad argument or scene */
    EMB_E_DEVICE = -2,    /* a HIP call failed */
    EMB_E_NOMEM = -3,     /* device memory or the admission budget is exhausted */
    EMB_E_STATE = -4,     /* call not valid in the handle's state */
    EMB_E_NUMERIC = -5,   /* a device-side check failed (see emb_last_error) */
};

/* Physics subset of `Settings` (crates/embloks-core/src/settings.rs:164-236).
 * emb_settings_default fills the reference's defaults. */
typedef struct EmbSettings {
    double dt;
    double cell_charge;
    int32_t carriers_per_cell;
    int32_t battery_density;
    double carrier_charge_to_mass;
    double drag_per_resistance;
    double battery_resistance;
    double battery_trim_time;
    double battery_gain_low;
    double battery_gain_high;
    double current_smoothing_time;
    double wall_restitution;
    double max_step;
    int32_t max_substeps;
    int32_t smoothing_passes;
    double coulomb_k;
    double speed_of_light;
    int32_t absorbing_margin;
    int32_t reserved0;
    double absorbing_strength;
    double electron_softening;
    double electron_cutoff;
    double display_smoothing_time;
} EmbSettings;

/* A scene: the reference's `World` (crates/embloks-core/src/world.rs:24-38) as
 * flat arrays. Electrons carry the World's uniform charge and mass
 * (world.rs:59-66). `gains` lists the battery gain per stack in find_stacks
 * order (battery.rs:100-147); NaN marks a stack without a gain entry (gain 1).
 * `time` is the scene's simulated time (Simulation::time). */
typedef struct EmbScene {
    int32_t cols, rows;
    const uint8_t *kind;        /* 0 vacuum, 1 conductor, 2 battery (materials.rs:8-12) */
    const double *resistance;
    const double *density;
    const double *emf;
    const uint8_t *orientation; /* 0 right, 1 down, 2 left, 3 up (materials.rs:30-35) */
    const double *frequency;
    int64_t n_electrons;
    const double *pos;          /* [n_electrons][2] */
    const double *vel;          /* [n_electrons][2] */
    int64_t n_fixed;
    const double *fixed_pos;    /* [n_fixed][2] */
    const double *fixed_charge; /* [n_fixed] */
    int32_t n_gains;
    int32_t reserved0;
    const double *gains;        /* [n_gains], NaN = no entry */
    double time;
} EmbScene;

typedef struct EmbDims {
    int32_t cols, rows, margin, nx, ny;
    int32_t n_stacks;
    int64_t n_electrons;
    int32_t fdtd_substeps;
    int32_t last_substeps;      /* electron substeps of the most recent tick */
    double time;
} EmbDims;

/* The info panel's readings (Probe, crates/embloks-gpu/src/sim.rs:82-101). */
typedef struct EmbProbe {
    int32_t valid;              /* a visible cell was probed (Probe::cell is Some) */
    int32_t electrons;          /* electrons in the cell */
    double potential;           /* Simulation::potential at the cell, close range included */
    double e[2], b, s[2];       /* displayed E and B, S = (E x B) / mu0 */
    double current[2];          /* smoothed cell current */
    double wire_current;        /* Probe::wire_current: flow summed across the wire ... */
    int32_t wire_width;         /* ... and the wire's width in cells */
    int32_t stack_state;        /* 0 not a stacked battery cell, 1 not connected, 2 connected */
    double v_low, v_high;       /* the stack's terminal potentials at its last trim */
    double stack_current, gain;
    int32_t has_mean_e;         /* the scene has wire cells */
    int32_t pad;
    double mean_e_in_wires;
    uint32_t histograms[3][256];  /* |E|, |B| c^2, |S| (arrow cells): counts per log10 bin of width 18/256 from 1e-12 */
} EmbProbe;

typedef struct EmbDevice EmbDevice;
typedef struct EmbSim EmbSim;

const char *emb_last_error(void);
void emb_settings_default(EmbSettings *out);

/* Device lifecycle; the device handle is shared by every simulation on it, as
 * `Kernels` is in the reference (crates/embloks-gpu/src/sim.rs:164-178). */
int emb_device_open(int ordinal, EmbDevice **out);
int emb_device_close(EmbDevice *dev);

/* GpuSim::new / with_kernels + upload (sim.rs:214-370): builds the tables,
 * uploads the scene and brings the fields to the static field of its charges
 * (Simulation::new, crates/embloks-core/src/simulation.rs:25-30). */
int emb_sim_create(EmbDevice *dev, const EmbSettings *settings, const EmbScene *scene, EmbSim **out);
/* GpuSim::upload (sim.rs:395-529): replace electrons, materials, gains and
 * static charges of an existing simulation of the same grid; the fields carry
 * on, corrected for Gauss's law with the new charges (Simulation::sync). */
int emb_sim_upload(EmbSim *sim, const EmbScene *scene, int reset_display);
/* GpuSim::encode_ticks + submit (sim.rs:607-661, 784-791): enqueue `ticks`
 * physics ticks (Simulation::step, simulation.rs:63-72) on the sim's stream.
 * Asynchronous; emb_sim_sync waits. */
int emb_sim_step(EmbSim *sim, int ticks);
/* Waits for the queued ticks. EMB_E_NUMERIC if a tick's later electron substeps
 * were dropped (the substep queue's wait limit, a fault never expected). */
int emb_sim_sync(EmbSim *sim);
int emb_sim_dims(EmbSim *sim, EmbDims *out);
/* read_state (sim.rs:719-768): electrons in upload order and the gains per
 * stack (NaN where the stack has no gain entry). */
int emb_sim_read_state(EmbSim *sim, double *pos, double *vel, double *gains);
/* read_yee (sim.rs:818-823). */
int emb_sim_read_yee(EmbSim *sim, double *ex, double *ey, double *bz);
/* Overwrite the Yee fields (NULL leaves a component as is). The CPU backend
 * exposes them the same way (`Simulation::maxwell` is public, simulation.rs:15-20;
 * crates/embloks-gpu/tests/sim.rs:208 assigns cpu.maxwell.ex). */
int emb_sim_write_yee(EmbSim *sim, const double *ex, const double *ey, const double *bz);
/* read_potential (sim.rs:826-830): the potential of the charges (nx, ny). */
int emb_sim_read_potential(EmbSim *sim, double *potential);
/* read_current (sim.rs:843-849): physics::current_density, (cols, rows, 2). */
int emb_sim_read_current(EmbSim *sim, double *current);
/* GpuSim::read_display (crates/embloks-gpu/src/sim.rs:831-840): the displayed E
 * (cols x rows x 2) and B (cols x rows) at cell centres, exponential averages
 * over display_smoothing_time updated at the end of every tick (finish_tick,
 * fields.wgsl:270-281) and reset to the exact fields on create and on an upload
 * with reset_display (finish_sync). */
int emb_sim_read_display(EmbSim *sim, double *e, double *b);
/* The smoothed (ammeter) cell current, cols x rows x 2, averaged over
 * current_smoothing_time: what GpuSim::read_state / apply_state return
 * (sim.rs:739-760). */
int emb_sim_read_smoothed_current(EmbSim *sim, double *current);
/* GpuSim::set_current (sim.rs:531-539): replace the smoothed current (from a save). */
int emb_sim_set_current(EmbSim *sim, const double *current);
/* The charge deposit as the int64 fixed-point grid the kernels hold: electron
 * charge in units of 2^-44 of one electron (cols * rows, unsmoothed, from the
 * electrons' current positions). No counterpart in the reference, which
 * deposits in floating point (fields.rs:43-63); exported so that the deposit
 * can be checked bit for bit. */
int emb_sim_read_charge_fixed(EmbSim *sim, int64_t *rho);
/* FP64 device reductions over the current state: max |div E - 4 pi k rho|,
 * total deposited electron charge (exact, in electron charges), field energy
 * sum(Ex^2) + sum(Ey^2) + c^2 sum(Bz^2). No counterpart in the reference. */
int emb_sim_diagnostics(EmbSim *sim, double *gauss_residual, double *electron_charge, double *field_energy);
/* emb_sim_diagnostics' Gauss residual and field energy (the same values, summed
 * in the same order) without allocating anything: no hipMalloc, hipFree or
 * allocating library call. No counterpart in the reference: for callers that
 * cannot allocate (code under HIP graph capture, serving loops that check
 * Gauss's law every N ticks, harnesses that forbid allocation). The per-block
 * partials live in a file-scope device array of 2 x 1024 + 2 doubles (16 KiB)
 * per GPU, held until the process ends; the reduction is grid-stride, so every
 * scene fits it. The exact electron charge is not offered (it needs a deposit
 * dump as large as the grid). Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (several threads, or simulations on different
 * streams) race. */
int emb_sim_diagnostics_static(EmbSim *sim, double *gauss_residual, double *field_energy);
/* Run `ticks` more ticks with a HIP event between the kernel phases of each tick
 * on the sim's stream and write each phase's average duration (ms) to avg_ms.
 * Returns the number of phases (avg_ms NULL: just the count). Names from
 * emb_tick_phase_name. No counterpart in the reference (which times nothing
 * but the wall clock, crates/embloks-gpu/examples/bench.rs:47-54). */
int emb_sim_kernel_times(EmbSim *sim, int ticks, double *avg_ms, int capacity);
/* CloseRange::force (crates/embloks-core/src/close_range.rs:314-376): the
 * close-range pair and ion force on every electron at its current position
 * (n_electrons x 2 doubles, world order), computed `repeats` times on the sim's
 * stream; *avg_ms (if not NULL) gets the mean time per evaluation. */
int emb_sim_close_force(EmbSim *sim, int32_t repeats, double *force, double *avg_ms);
const char *emb_tick_phase_name(int phase);
/* GpuSim::read_probe (sim.rs:663-715; probe_cell and probe_grid,
 * fields.wgsl:368-466): the readings at visible cell (c, r) (c < 0: none) and
 * the grid's mean |E| in wires and magnitude histograms, the S histogram over
 * the cells (c, r) with c % arrow_stride == r % arrow_stride == arrow_stride / 2. */
int emb_sim_probe(EmbSim *sim, int32_t c, int32_t r, int32_t arrow_stride, EmbProbe *out);
int emb_sim_destroy(EmbSim *sim);

/* --- Continuous-batching serving loop (no counterpart in the reference) ----------
 * Scene requests of different sizes are submitted with a step count; the server
 * admits them while their device footprint fits `hbm_budget_bytes`, keeps their
 * state resident, and advances every admitted scene together, one batched
 * kernel per phase of the tick, replayed as a captured HIP graph. `batched` = 0
 * runs the same scenes one scene per launch instead (the plain comparison). */
typedef struct EmbServerConfig {
    int64_t hbm_budget_bytes;
    int32_t max_slots;          /* scenes resident at once */
    int32_t batched;            /* 1: batched kernels + graphs, 0: one scene per launch */
    int32_t ticks_per_round;    /* ticks advanced between admission rounds (<= 64) */
    int32_t use_graph;          /* batched mode: replay each tick as a captured HIP graph */
    int32_t keep_results;       /* 1: keep each finished request's state until read */
    int32_t reserved0;
} EmbServerConfig;

typedef struct EmbServerStats {
    int64_t completed;
    int64_t cell_steps;         /* nx * ny * ticks summed over completed work */
    double elapsed_s;
    double cell_steps_per_s;
    double step_latency_p50_ms;   /* one tick of a resident scene: the batch tick it took part in */
    double step_latency_p99_ms;
    int64_t step_samples;
    int32_t peak_resident;
    int32_t reserved0;
    double request_latency_p50_ms;   /* per completed request: submit -> finished (host clock) */
    double request_latency_p99_ms;
    double queue_wait_p50_ms;        /* per admitted request: submit -> admitted */
    double queue_wait_p99_ms;
} EmbServerStats;

typedef struct EmbServer EmbServer;

int emb_server_create(EmbDevice *dev, const EmbSettings *settings, const EmbServerConfig *config, EmbServer **out);
/* Queue a scene for `steps` ticks; returns its id in *id. Requests whose
 * footprint alone exceeds the budget are refused with EMB_E_NOMEM. */
int emb_server_submit(EmbServer *srv, const EmbScene *scene, int64_t steps, int64_t *id);
/* Cancel a queued or running request. */
int emb_server_cancel(EmbServer *srv, int64_t id);
/* Run one admission round: admit what fits, advance every resident scene by up
 * to ticks_per_round ticks, retire finished ones. *active gets the number of
 * scenes resident or queued afterwards. */
int emb_server_poll(EmbServer *srv, int32_t *active);
/* 1 if request `id` finished (its result can be read), 0 if not yet, <0 error. */
int emb_server_done(EmbServer *srv, int64_t id);
/* Result of a finished request (kept when keep_results is set): electrons and
 * fields as emb_sim_read_state / emb_sim_read_yee would give them. Any pointer
 * may be NULL. Reading releases the result. EMB_E_NUMERIC if the run dropped
 * electron substeps (the substep queue's wait limit, a fault never expected). */
int emb_server_result(EmbServer *srv, int64_t id, double *pos, double *vel, double *ex, double *ey, double *bz);
int emb_server_stats(EmbServer *srv, EmbServerStats *out);
int emb_server_destroy(EmbServer *srv);

/* --- The DST-I Poisson solve on its own ----------------------------------------
 * Replaces Maxwell::poisson (crates/embloks-core/src/maxwell.rs:139-168) for a
 * caller's right-hand side on an nx x ny node grid (x-major, index i * ny + j,
 * phi = 0 outside): phi = Sx ((Sx r Sy) .* inv_lap) Sy, as four dense products
 * on MFMA, in FP32 or in BF16 with a 3-term split (hi.hi + hi.lo + lo.hi, FP32
 * accumulation). The tick itself always solves in FP32. */
enum { EMB_DST_FP32 = 0, EMB_DST_BF16X3 = 1 };
typedef struct EmbPoisson EmbPoisson;
int emb_poisson_create(EmbDevice *dev, int32_t nx, int32_t ny, EmbPoisson **out);
/* Upload the right-hand side (nx * ny doubles, rounded to FP32 on the device). */
int emb_poisson_set_rhs(EmbPoisson *p, const double *rhs);
/* Run `repeats` solves back to back on the handle's stream and wait; *avg_ms
 * (if not NULL) gets the mean time per solve from HIP events around them.
 * tile: 64 (64 x 64 output tiles) or 128 (FP32: 128 x 64, BF16x3: 128 x 128),
 * 0 = by grid size. */
int emb_poisson_solve(EmbPoisson *p, int32_t precision, int32_t tile, int32_t repeats, double *avg_ms);
/* The last solve's potential (nx * ny doubles). */
int emb_poisson_read(EmbPoisson *p, double *phi);
int emb_poisson_destroy(EmbPoisson *p);

/* --- The close-range tables -------------------------------------------------
 * Replaces close_range::tables (crates/embloks-core/src/close_range.rs:120-138,
 * built by build_tables, 140-146): the radial correction g and its potential
 * (4098 samples each, uniform in r^2 up to the cutoff plus a zero sample), the
 * square table (square_n x square_n x 3: potential, field x, field y over
 * offsets in [-half, half]^2) and the ion stencil ((2 reach + 1)^2). Built on
 * the device in FP64 (grid_force's 64 Poisson solves and gathers, then the
 * radial, square and ion tables) and cached for the process; emb_sim_create
 * and emb_server_submit build what their settings need. NULL arrays are
 * skipped, so a first call with NULLs reads the sizes. If build_ms is not
 * NULL the tables are built afresh and *build_ms gets the build's host-clock
 * time (launches, kernels, copy back). cutoff must be in [0.02, 32] cells and
 * passes in [0, 16] (the smoothed charge must stay inside the 16-node margin). */
int emb_close_tables(EmbDevice *dev, double softening, int32_t passes, double cutoff, double *g, double *potential,
                     double *square, int32_t *square_n, double *half, double *ion_stencil, int32_t *reach,
                     double *build_ms);

/* Footprint in bytes that a scene of this shape would occupy on the device. */
int64_t emb_scene_footprint(const EmbSettings *settings, int32_t cols, int32_t rows, int64_t n_electrons);

#ifdef __cplusplus
}
#endif

#endif
