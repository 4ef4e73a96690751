/*
 * embloks_oracle.h - FP64 CPU restatement of embloks-core's per-tick physics.
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle the HIP backend is checked
 * against: only the tests, the smoke check and the benchmark's CPU-baseline leg
 * load it. The product library never links it.
 *
 * It follows crates/embloks-core/src/{maxwell,fields,physics,close_range,
 * battery,simulation,connectivity}.rs function by function (each function cites
 * the reference lines it restates) and is pinned by the reference's golden
 * fixtures (the JSON files in crates/embloks-core/tests/fixtures), closed forms and
 * invariants; see tests/test_oracle_*.py.
 */
#ifndef EMBLOKS_ORACLE_H
#define EMBLOKS_ORACLE_H

#include <stdint.h>
#include "../include/embloks_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Threads used by the lane/particle loops (1 = the reference's "one core"). */
void orc_set_threads(int threads);
/* Settings::default (settings.rs:164-236), physics subset. */
void orc_settings_default(EmbSettings *out);

/* --- Maxwell (crates/embloks-core/src/maxwell.rs) -------------------------------- */
typedef struct OrcMaxwell OrcMaxwell;
OrcMaxwell *orc_mx_new(const EmbSettings *s, int cols, int rows);
void orc_mx_free(OrcMaxwell *mx);
void orc_mx_dims(const OrcMaxwell *mx, int *nx, int *ny);
void orc_mx_set(OrcMaxwell *mx, const double *ex, const double *ey, const double *bz);
void orc_mx_get(const OrcMaxwell *mx, double *ex, double *ey, double *bz, double *potential);
void orc_mx_enforce_gauss(OrcMaxwell *mx, const double *rho);
void orc_mx_step(OrcMaxwell *mx, const double *rho, const double *current, double dt);
void orc_mx_poisson(const OrcMaxwell *mx, const double *rhs, double *phi);
void orc_mx_dst2(const OrcMaxwell *mx, double *array);
void orc_mx_smoothed(const OrcMaxwell *mx, const double *values, int d, double *out);
void orc_mx_gather(const OrcMaxwell *mx, int64_t n, const double *pos, double *out);
void orc_mx_node_fields(const OrcMaxwell *mx, double *e, double *b);
void orc_mx_divergence(const OrcMaxwell *mx, double *div);
void orc_mx_damping(OrcMaxwell *mx, double h, double *ex, double *ey, double *bz);

/* --- Deposit (crates/embloks-core/src/fields.rs) ----------------------------------- */
void orc_cic_deposit(int64_t n, const double *pos, const double *q, int cols, int rows, double *out);
void orc_cic_deposit_vectors(int64_t n, const double *pos, const double *v, int cols, int rows, double *out);
void orc_spread_over_cells(int cols, int rows, const double *in, double *out);

/* --- Close-range tables (crates/embloks-core/src/close_range.rs) -------------------- */
int orc_grid_force(int passes, double reach, double *radii, double *force, int capacity);
/* Sizes: g and potential have PAIR_SAMPLES + 2 entries; square is n*n*3; the ion
 * stencil is (2 reach + 1)^2. Pass NULL outputs to query n and reach. */
void orc_tables(double softening, int passes, double cutoff, double *g, double *potential,
                double *square, int *square_n, double *half, double *ion_stencil, int *reach);
double orc_compact_g(double r2, double softening);

/* --- Whole simulation (crates/embloks-core/src/simulation.rs) ------------------------ */
typedef struct OrcSim OrcSim;
OrcSim *orc_sim_new(const EmbSettings *s, const EmbScene *scene);
void orc_sim_free(OrcSim *sim);
void orc_sim_dims(const OrcSim *sim, EmbDims *out);
int orc_sim_step(OrcSim *sim, int ticks);
void orc_sim_read_state(const OrcSim *sim, double *pos, double *vel, double *gains);
void orc_sim_read_yee(const OrcSim *sim, double *ex, double *ey, double *bz);
void orc_sim_read_potential(const OrcSim *sim, double *potential);
void orc_sim_write_state(OrcSim *sim, const double *pos, const double *vel, const double *ex,
                         const double *ey, const double *bz);
void orc_sim_current(const OrcSim *sim, double *current);
void orc_sim_node_fields(const OrcSim *sim, double *e, double *b);
/* Simulation::potential(None): visible potential plus the close-range part. */
void orc_sim_potential(OrcSim *sim, double *potential);
void orc_sim_close_force(const OrcSim *sim, double *force);
/* CloseRange::force for the electrons listed in idx only (m of them), on a
 * scene built from `scene` without a field solve: the one-sided sum over every
 * neighbour within the cutoff plus the ions, in FP64 (OpenMP over the list). */
void orc_close_force_subset(const EmbSettings *s, const EmbScene *scene, int64_t m, const int64_t *idx, double *out);
/* Stacks found in the scene: low/high terminals (c, r) and cell counts. */
int orc_sim_stacks(const OrcSim *sim, int *low_high, int *n_cells);
void orc_sim_labels(const OrcSim *sim, int64_t *labels);
/* physics::confine (physics.rs:66-115) against the sim's islands. */
void orc_sim_confine(const OrcSim *sim, const double *old_pos, const double *new_pos, const double *vel,
                     double restitution, double *out_pos, double *out_vel);
/* physics::kick (physics.rs:57-60). */
void orc_kick(const double *vel, const double *acc, double damping, double dt, double *out);

#ifdef __cplusplus
}
#endif

#endif
