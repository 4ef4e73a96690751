"""The HIP backend against the FP64 oracle, through the C ABI.

Mirrors crates/embloks-gpu/tests/sim.rs (static fields, ticks following the CPU)
with the reference's own tolerances, and adds the stricter checks this backend
is built to: per-tick reset-and-compare at 1e-5 relative (max|a - e| / max|e|,
fixtures.rs:39-46), a bit-exact integer deposit, Gauss's law, determinism.
"""
import numpy as np
import pytest

from embloks_amd import world as W
from oracle import orc
from tests.golden.fixtures import SCENES, array, load, relative_error, world_from

gpu = pytest.mark.gpu


def gpusim():
    from embloks_amd.gpusim import GpuSim
    return GpuSim


def fixture_world(name):
    return world_from(load(name))


@gpu
@pytest.mark.parametrize("name", ["scene_f1", "scene_f2", "scene_f4", "scene_f7"])
def test_static_fields_match_the_oracle(name):
    # crates/embloks-gpu/tests/sim.rs:45-60 (potential 1e-4, E 1e-3); held here to 1e-5.
    world = fixture_world(name)
    cpu = orc.Sim(world)
    sim = gpusim()(world)
    ex, ey, bz = sim.read_yee()
    cex, cey, cbz = cpu.yee()
    assert relative_error(sim.read_potential(), cpu.potential_nodes()) < 1e-5
    assert relative_error(ex, cex) < 1e-5
    assert relative_error(ey, cey) < 1e-5
    assert np.all(bz == 0.0)
    assert relative_error(sim.read_current(), cpu.current()) < 1e-6


@gpu
def test_potential_solved_on_read_is_the_same_whenever_it_is_read():
    # A scene without battery stacks solves the charge's potential only when it
    # is read (read_potential, the probe), from the right-hand side the last
    # tick left: reading it after every tick, twice in a row or after a probe
    # gives the same bits. (The one-tick comparisons hold it to the oracle.)
    world = W.random_wire(40, 32)
    assert not np.any(world.kind == W.BATTERY)
    a, b = gpusim()(world), gpusim()(world)
    for _ in range(3):
        a.run_ticks(1)
        a.read_potential()
    b.run_ticks(3)
    want = b.read_potential()
    np.testing.assert_array_equal(a.read_potential(), want)
    np.testing.assert_array_equal(b.read_potential(), want)
    p = b.read_probe((10, 10))["potential"]
    np.testing.assert_array_equal(b.read_potential(), want)
    assert p == a.read_probe((10, 10))["potential"]
    a.destroy()
    b.destroy()


def reset_and_compare(world, ticks_before=0):
    """Both backends from the same state; compare after one tick."""
    cpu = orc.Sim(world)
    if ticks_before:
        cpu.step(ticks_before)
        pos, vel, gains = cpu.state()
        ex, ey, bz = cpu.yee()
        w = world.copy()
        w.pos, w.vel, w.gains, w.time = pos, vel, gains, cpu.time
        sim = gpusim()(w)
        sim.write_yee(ex, ey, bz)
    else:
        sim = gpusim()(world)
    cpu.step(1)
    sim.run_ticks(1)
    return cpu, sim


@gpu
@pytest.mark.parametrize("name", SCENES)
@pytest.mark.parametrize("before", [0, 20])
def test_one_tick_matches_the_oracle(name, before):
    cpu, sim = reset_and_compare(fixture_world(name), before)
    ex, ey, bz = sim.read_yee()
    cex, cey, cbz = cpu.yee()
    err = {
        "ex": relative_error(ex, cex),
        "ey": relative_error(ey, cey),
        "bz": relative_error(bz, cbz),
        "phi": relative_error(sim.read_potential(), cpu.potential_nodes()),
    }
    pos, vel, gains = sim.read_state()
    cpos, cvel, cgains = cpu.state()
    err["pos"] = float(np.max(np.abs(pos - cpos)))
    print(name, before, err)
    for k in ("ex", "ey", "phi"):
        assert err[k] < 1e-5, err
    assert err["bz"] < 1e-4, err
    assert err["pos"] < 1e-4, err
    for g, c in zip(gains, cgains):
        assert (g is None) == (c is None)
        if g is not None:
            assert abs(g - c) < 1e-5


def odd_world(cols=31, rows=32):
    """A conductor ring with one battery on a cols x rows grid. 31 x 32 gives 63
    x 64 Yee nodes, so the DST's row fold has a middle row (odd length); 32 x 31
    gives a middle column (the column fold of an odd row length, whose rows are
    padded and whose mirror columns are not 16-byte aligned); 33 x 34 gives row
    lengths of 66, even but not a multiple of 4."""
    w = W.World(cols, rows, seed=7)
    ring = [(c, r) for c in range(4, cols - 4) for r in range(6, rows - 8)
            if c in (4, cols - 5) or r in (6, rows - 9)]
    w.paint(ring, W.CONDUCTOR, resistance=1.0)
    w.paint([(15, 6)], W.BATTERY, emf=1.5, orientation=W.RIGHT)
    return w


@gpu
@pytest.mark.parametrize("shape", [(31, 32), (32, 31), (33, 34)])
@pytest.mark.parametrize("before", [0, 5])
def test_odd_node_count_one_tick_matches_the_oracle(shape, before):
    cpu, sim = reset_and_compare(odd_world(*shape), before)
    ex, ey, bz = sim.read_yee()
    cex, cey, cbz = cpu.yee()
    nx, ny = shape[0] + 32, shape[1] + 32
    assert ex.shape == (nx + 1, ny) and ey.shape == (nx, ny + 1)
    err = {"ex": relative_error(ex, cex), "ey": relative_error(ey, cey),
           "phi": relative_error(sim.read_potential(), cpu.potential_nodes())}
    print(shape, before, err)
    for k in ("ex", "ey", "phi"):
        assert err[k] < 1e-5, err


def fast_world(world, speed, seed):
    """`world` with its electrons thrown in random directions at up to `speed`
    cells per time unit along each axis: vmax dt / max_step > 1, so a tick takes
    several electron substeps, each after a fresh cell list (physics.rs:124-135,
    sim.rs:643-651)."""
    w = world.copy()
    w.vel = np.random.default_rng(seed).uniform(-speed, speed, size=w.vel.shape)
    return w


def expected_substeps(world):
    s = world.settings
    vmax = float(np.max(np.hypot(world.vel[:, 0], world.vel[:, 1])))
    return int(min(max(np.ceil(vmax * s["dt"] / s["max_step"]), 1), s["max_substeps"]))


@gpu
@pytest.mark.parametrize("name,speed", [("scene_f1", 20.0), ("scene_f1", 60.0), ("scene_f6", 200.0), ("wire128", 30.0)])
def test_fast_electrons_take_several_substeps(name, speed):
    base = W.random_wire(128, 128, carriers_per_cell=8, seed=9) if name == "wire128" else fixture_world(name)
    world = fast_world(base, speed, 3)
    want = expected_substeps(world)
    assert want > 1
    cpu, sim = reset_and_compare(world)
    assert sim.dims().last_substeps == want
    pos, vel, _ = sim.read_state()
    cpos, cvel, _ = cpu.state()
    ex, ey, _ = sim.read_yee()
    cex, cey, _ = cpu.yee()
    err = {"pos": float(np.max(np.abs(pos - cpos))), "vel": relative_error(vel, cvel),
           "ex": relative_error(ex, cex), "ey": relative_error(ey, cey),
           "phi": relative_error(sim.read_potential(), cpu.potential_nodes())}
    print(name, speed, want, world.n_electrons, err)
    assert err["pos"] < 1e-4, err
    assert err["vel"] < 1e-5, err
    for k in ("ex", "ey", "phi"):
        assert err[k] < 1e-5, err


@gpu
def test_fast_electrons_are_deterministic():
    world = fast_world(W.random_wire(128, 128, carriers_per_cell=8, seed=9), 30.0, 4)
    runs = []
    for _ in range(2):
        sim = gpusim()(world)
        sim.run_ticks(3)
        runs.append(sim.read_state()[:2] + sim.read_yee())
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


def pulse_fields(nx, ny, seed):
    """A smooth initial field: a Gaussian Bz pulse and a swirl of E, in FP32 values."""
    i = np.arange(nx + 1)[:, None]
    j = np.arange(ny + 1)[None, :]
    g = np.exp(-((i - 0.45 * nx) ** 2 + (j - 0.55 * ny) ** 2) / 30.0)
    bz = (0.8 * g).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(seed)
    a, b = rng.normal(size=2)
    ex = (a * np.sin(2 * np.pi * np.arange(nx + 1)[:, None] / nx) * np.cos(np.pi * np.arange(ny)[None, :] / ny))
    ey = (b * np.cos(np.pi * np.arange(nx)[:, None] / nx) * np.sin(2 * np.pi * np.arange(ny + 1)[None, :] / ny))
    return ex.astype(np.float32).astype(np.float64), ey.astype(np.float32).astype(np.float64), bz


@gpu
def test_field_only_linear_run_follows_the_oracle():
    # The field path alone over a 2000-tick run (leapfrog, absorbing damping,
    # Gauss cleaning against fixed charges): no electrons, so no current, and
    # the run is linear in the fields. Both backends start from the same FP32
    # values; the reference's metric max|a - e| / max|e| at checkpoints.
    w = W.World(40, 32, seed=3)
    w.place_charges([(12, 10), (27, 20)], 3.0)
    w.place_charges([(20, 16)], -2.0)
    assert w.n_electrons == 0
    cpu = orc.Sim(w)
    sim = gpusim()(w)
    ex0, ey0, bz0 = pulse_fields(sim.nx, sim.ny, 5)
    # start both from the static field plus the pulse
    cex, cey, _ = cpu.yee()
    ex0, ey0 = (cex + ex0).astype(np.float32).astype(np.float64), (cey + ey0).astype(np.float32).astype(np.float64)
    cpu.write_state(ex=ex0, ey=ey0, bz=bz0)
    sim.write_yee(ex0, ey0, bz0)
    done, errs = 0, {}
    for upto in (1, 10, 100, 500, 2000):
        cpu.step(upto - done)
        sim.run_ticks(upto - done)
        done = upto
        gex, gey, gbz = sim.read_yee()
        cex, cey, cbz = cpu.yee()
        errs[upto] = (relative_error(gex, cex), relative_error(gey, cey), relative_error(gbz, cbz),
                      float(np.abs(cbz).max()))
    print("field-only run, relative errors (Ex, Ey, Bz) and max |Bz| by tick:", errs)
    for upto, (eex, eey, ebz, _) in errs.items():
        assert eex < 1e-5 and eey < 1e-5, (upto, errs[upto])
        # Bz is normalised by its own maximum, which falls from 0.79 to 0.008 by
        # tick 500 as the pulse leaves through the absorbing margin, while the
        # FP32 rounding of the charges' static E (differenced in curl E) feeds it
        # at a steady rate: measured 1.1e-5 there (9e-8 absolute, 1e-7 of the
        # starting pulse), held here at 2e-5.
        assert ebz < 2e-5, (upto, errs[upto])


@gpu
@pytest.mark.parametrize("before", [0, 20])
def test_fast_light_takes_three_fdtd_substeps(before):
    # c = 40: c dt sqrt 2 / 0.9 > 2, so three FDTD substeps per tick
    # (maxwell.rs:376-378; the oracle's substepping is pinned by the cavity
    # eigenmode and wave-front tests). Reset to the oracle's state and compare
    # one coupled tick at the per-tick bars.
    world = W.scene(0, speed_of_light=40.0)
    cpu, sim = reset_and_compare(world, before)
    assert sim.dims().fdtd_substeps == 3
    ex, ey, bz = sim.read_yee()
    cex, cey, cbz = cpu.yee()
    err = {"ex": relative_error(ex, cex), "ey": relative_error(ey, cey), "bz": relative_error(bz, cbz),
           "phi": relative_error(sim.read_potential(), cpu.potential_nodes()),
           "pos": float(np.max(np.abs(sim.read_state()[0] - cpu.state()[0])))}
    print("c = 40, before", before, err)
    for k in ("ex", "ey", "phi"):
        assert err[k] < 1e-5, err
    assert err["bz"] < 1e-4, err
    assert err["pos"] < 1e-4, err


@gpu
@pytest.mark.parametrize("index", [0, 1, 5])
def test_ticks_follow_the_oracle(index):
    # crates/embloks-gpu/tests/sim.rs:62-90 with its tolerances.
    world = W.scene(index)
    cpu = orc.Sim(world)
    sim = gpusim()(world)
    ticks = 10
    sim.run_ticks(ticks)
    cpu.step(ticks)
    pos, vel, gains = sim.read_state()
    cpos, cvel, cgains = cpu.state()
    worst = float(np.max(np.abs(pos - cpos)))
    assert worst < 2e-3, worst
    ex, ey, bz = sim.read_yee()
    cex, cey, cbz = cpu.yee()
    assert relative_error(ex, cex) < 2e-2
    assert relative_error(ey, cey) < 2e-2
    assert relative_error(bz, cbz) < 5e-2
    assert relative_error(sim.read_current(), cpu.current()) < 5e-2
    for g, c in zip(gains, cgains):
        assert abs((g if g is not None else 1.0) - (c if c is not None else 1.0)) < 1e-3


def fixed_point_deposit(pos, cols, rows):
    """The integer CIC deposit (fields.rs:27-51 with 22-bit weights per axis),
    restated in numpy from the same cell + FP32-fraction positions: electron
    charge in units of 2^-44."""
    cell = np.floor(pos).astype(np.int64)
    frac = (pos - cell).astype(np.float32)   # exact: the backend stores these fractions
    half = np.float32(0.5)
    low = frac >= half
    c0 = np.where(low, cell, cell - 1)
    u = np.where(low, frac - half, frac + half).astype(np.float32)
    f = np.rint(u * np.float32(4194304.0)).astype(np.int64)
    out = np.zeros(cols * rows, dtype=np.int64)
    for a in (0, 1):
        wx = f[:, 0] if a else (1 << 22) - f[:, 0]
        cc = np.clip(c0[:, 0] + a, 0, cols - 1)
        for b in (0, 1):
            wy = f[:, 1] if b else (1 << 22) - f[:, 1]
            rr = np.clip(c0[:, 1] + b, 0, rows - 1)
            np.add.at(out, cc * rows + rr, -(wx * wy))
    return out.reshape(cols, rows)


@gpu
@pytest.mark.parametrize("index", [0, 2, 5])
def test_deposit_is_bit_exact_and_conserves_charge(index):
    world = W.scene(index)
    sim = gpusim()(world)
    sim.run_ticks(7)
    pos, _, _ = sim.read_state()
    got = sim.read_charge_fixed()
    want = fixed_point_deposit(pos, world.cols, world.rows)
    np.testing.assert_array_equal(got, want)
    assert int(got.sum()) == -world.n_electrons * (1 << 44)
    residual, charge, energy = sim.diagnostics()
    q0 = world.settings["cell_charge"] / world.settings["carriers_per_cell"]
    assert charge == -world.n_electrons * q0
    assert np.isfinite(energy)


@gpu
def test_total_charge_is_exact_past_the_int64_range():
    # 1.01M electrons: their deposit sums to -N 2^44, beyond int64 from 2^19 electrons on
    world = W.rectangles_circuit(512, carriers_per_cell=16, seed=2)
    sim = gpusim()(world)
    assert sim.n > (1 << 19)
    sim.run_ticks(2)
    _, charge, _ = sim.diagnostics()
    q0 = world.settings["cell_charge"] / world.settings["carriers_per_cell"]
    assert charge == -sim.n * q0
    wrapped = int(sim.read_charge_fixed().view(np.uint64).sum(dtype=np.uint64))
    assert wrapped == (-sim.n * (1 << 44)) % (1 << 64)


@gpu
def test_gauss_law_holds_after_ticks():
    world = W.scene(0)
    sim = gpusim()(world)
    sim.run_ticks(30)
    residual, _, _ = sim.diagnostics()
    # div E - 4 pi k rho in FP32 fields: round-off of the O(10) values involved.
    assert residual < 1e-4, residual


@gpu
@pytest.mark.parametrize("index", [0, 1, 5])
def test_runs_are_bitwise_deterministic(index):
    world = W.scene(index)
    a, b = gpusim()(world), gpusim()(world)
    a.run_ticks(40)
    b.run_ticks(40)
    for x, y in zip(a.read_yee() + a.read_state()[:2], b.read_yee() + b.read_state()[:2]):
        np.testing.assert_array_equal(x, y)


@gpu
def test_fixture_scene_thirty_ticks():
    # The golden 30-tick runs (fixtures.rs:128-186) in FP32: positions and fields
    # at the reference GPU's own 10-tick bar.
    for name in ("scene_f1", "scene_f7"):
        f = load(name)
        sim = gpusim()(world_from(f))
        sim.run_ticks(f["ticks"])
        pos, _, gains = sim.read_state()
        assert np.max(np.abs(pos - array(f["final_pos"]))) < 2e-3
        e, b = sim.node_fields()
        assert relative_error(e, array(f["final_e"])) < 2e-2
        for g, want in zip(gains, f["gains"]):
            assert abs(g - want) < 1e-3


@gpu
def test_ring_circuit_two_thousand_ticks():
    # The 40 x 32 ring (one DC battery, one carrier per cell) run 2000 ticks on
    # both backends from the same scene: a coupled run, so the FP32/FP64
    # difference is reported as it grows rather than held to the one-tick bar.
    world = W.ring_circuit()
    cpu = orc.Sim(world)
    sim = gpusim()(world)
    cpu.step(2000)
    sim.run_ticks(2000)
    pos, vel, gains = sim.read_state()
    cpos, cvel, cgains = cpu.state()
    ex, ey, bz = sim.read_yee()
    cex, cey, cbz = cpu.yee()
    err = {"pos": float(np.max(np.abs(pos - cpos))), "ex": relative_error(ex, cex), "ey": relative_error(ey, cey),
           "bz": relative_error(bz, cbz), "phi": relative_error(sim.read_potential(), cpu.potential_nodes()),
           "gain": abs(gains[0] - cgains[0])}
    residual, _, _ = sim.diagnostics()
    # The current along the top and bottom edges, averaged over 20 samples, beside
    # Ohm's EMF / (71 wire cells x 1 ohm + the battery's 0.1). At one carrier per
    # cell the ring does not reach it in the FP64 oracle either (1.1e-4 there,
    # 0.0126 = 89% at 8 carriers per cell), so it is reported, not asserted;
    # Ohm's law is held at 8 carriers by test_loop_current_follows_ohms_law.
    samples = []
    for _ in range(20):
        sim.run_ticks(15)
        cur = sim.read_current()
        samples.append(0.5 * (abs(cur[10:30, 8, 0].mean()) + abs(cur[10:30, 23, 0].mean())))
    ohm = 1.0 / (71 * 1.0 + world.settings["battery_resistance"])
    print("ring 2000 ticks", err, "gauss residual", residual, "current", np.mean(samples), "Ohm", ohm)
    assert err["pos"] < 1e-2, err
    assert err["ex"] < 1e-2 and err["ey"] < 1e-2 and err["phi"] < 1e-2, err
    assert err["gain"] < 1e-3, err
    assert residual < 1e-4, residual


def crowded_world(carriers=80):
    # A conductor block at 80 carriers per cell: every cell takes the crowded-cell
    # path (k_push_heavy), its neighbours the cell mode, the vacuum edge nothing.
    s = dict(W.abi.DEFAULTS)
    s["carriers_per_cell"] = carriers
    w = W.World(24, 20, s, seed=3)
    w.paint(W.block(8, 16, 6, 14), W.CONDUCTOR, resistance=1.0)
    w.paint([(8, 10)], W.BATTERY, emf=1.0, orientation=W.RIGHT)
    return w


@gpu
def test_crowded_cells_one_tick_match_the_oracle():
    world = crowded_world()
    assert world.n_electrons > 64 * 60
    cpu, sim = reset_and_compare(world)
    pos, _, _ = sim.read_state()
    cpos, _, _ = cpu.state()
    ex, ey, _ = sim.read_yee()
    cex, cey, _ = cpu.yee()
    err = {"pos": float(np.max(np.abs(pos - cpos))), "ex": relative_error(ex, cex), "ey": relative_error(ey, cey),
           "phi": relative_error(sim.read_potential(), cpu.potential_nodes())}
    print("crowded cells one tick", err)
    assert err["pos"] < 1e-4 and err["ex"] < 1e-5 and err["ey"] < 1e-5 and err["phi"] < 1e-5, err


@gpu
def test_crowded_cells_are_deterministic():
    world = crowded_world()
    a, b = gpusim()(world), gpusim()(world)
    a.run_ticks(5)
    b.run_ticks(5)
    for x, y in zip(a.read_yee() + a.read_state()[:2], b.read_yee() + b.read_state()[:2]):
        np.testing.assert_array_equal(x, y)


@gpu
@pytest.mark.parametrize("c,fdtd_substeps", [(None, 1), (24.0, 2)])
def test_display_and_smoothed_current_follow_the_oracle(c, fdtd_substeps):
    # finish_tick (fields.wgsl:270-281): per tick, display += a_d (node fields -
    # display) and smoothed current += a_c (cell current - smoothed current),
    # with a_d = 1 - exp(-dt / display_smoothing_time) and
    # a_c = min(dt / current_smoothing_time, 1) (sim.rs:256-261); the display
    # starts at the exact fields (finish_sync) and the current average at zero.
    # An odd FDTD substep count ends the tick inside the Gauss correction, an
    # even one in a kernel of its own: both are checked.
    import math
    world = W.scene(5) if c is None else W.scene(5, speed_of_light=c)
    s = world.settings
    a_d = 1.0 - math.exp(-s["dt"] / s["display_smoothing_time"])
    a_c = min(s["dt"] / s["current_smoothing_time"], 1.0)
    cpu = orc.Sim(world)
    sim = gpusim()(world)
    assert sim.dims().fdtd_substeps == fdtd_substeps
    e_avg, b_avg = cpu.node_fields()
    cur_avg = np.zeros_like(cpu.current())
    e0, b0 = sim.read_display()
    assert relative_error(e0, e_avg) < 1e-5
    for _ in range(12):
        cpu.step(1)
        sim.run_ticks(1)
        e, b = cpu.node_fields()
        e_avg = e_avg + a_d * (e - e_avg)
        b_avg = b_avg + a_d * (b - b_avg)
        cur_avg = cur_avg + a_c * (cpu.current() - cur_avg)
    e_gpu, b_gpu = sim.read_display()
    err = {"e": relative_error(e_gpu, e_avg), "b": relative_error(b_gpu, b_avg),
           "current": relative_error(sim.read_smoothed_current(), cur_avg)}
    print("display and smoothed current after 12 ticks", err)
    # coupled ticks: the reference GPU's 10-tick bars are 2e-2 / 5e-2 / 5e-2
    # (tests/sim.rs:62-90); measured here 2-4e-6 at one and at two substeps
    assert err["e"] < 2e-5 and err["b"] < 2e-5 and err["current"] < 2e-5, err
    # set_current replaces the average exactly (to FP32), and upload keeps it
    want = np.random.default_rng(0).normal(size=cur_avg.shape)
    sim.set_current(want)
    assert np.max(np.abs(sim.read_smoothed_current() - want)) < 1e-6 * np.abs(want).max()
    sim.upload(world, reset_display=False)
    assert np.max(np.abs(sim.read_smoothed_current() - want)) < 1e-6 * np.abs(want).max()
    assert np.array_equal(sim.read_display()[1], b_gpu)


def away_from_the_cutoff(world, tol=1e-6, shift=1e-3):
    """Move each electron with a neighbour within tol * cutoff of the cutoff
    distance by up to `shift` cells in a direction of its own (seeded, so a
    pair's two electrons part whatever its orientation). The reference's pair force jumps by a whole
    term at r = cutoff (g(cutoff^2) != 0, close_range.rs:359-372), so such a
    pair may count on one backend and not on the other: FP32 resolves d^2 to
    about 1e-7 relative. The 512^2 scene's 4 x 4 carrier lattice holds 232
    ordered pairs within 1e-7 of the cutoff (none in the 40 x 32 scenes)."""
    from scipy.spatial import cKDTree
    c = world.settings["electron_cutoff"]
    rng = np.random.default_rng(11)
    for _ in range(8):
        tree = cKDTree(world.pos)
        near = (tree.query_ball_point(world.pos, c * (1 + tol), return_length=True, workers=-1) -
                tree.query_ball_point(world.pos, c * (1 - tol), return_length=True, workers=-1))
        moved = np.flatnonzero(near > 0)
        if len(moved) == 0:
            return
        angle = rng.uniform(0.0, 2.0 * np.pi, len(moved))
        step = rng.uniform(0.25 * shift, shift, len(moved))
        world.pos[moved, 0] += step * np.cos(angle)
        world.pos[moved, 1] += step * np.sin(angle)
    raise AssertionError("electrons left at the cutoff distance")


@gpu
@pytest.mark.parametrize("before", [0, 3])
def test_full_size_scene_one_tick_matches_the_oracle(before):
    # The 512 x 512 circuit of 64 rectangles and 16 batteries at 16 carriers per
    # conductor cell (1,011,088 electrons), both backends from the same state,
    # one tick, at the per-tick bars above. From rest (before = 0) the first
    # tick's velocities are the close-range forces alone, so the few pairs at
    # the cutoff distance would set the current's error: those electrons are
    # moved off it first.
    orc.set_threads(16)
    try:
        world = W.rectangles_circuit(512, carriers_per_cell=16, seed=2)
        assert world.n_electrons > 1_000_000
        if before == 0:
            away_from_the_cutoff(world)
        cpu, sim = reset_and_compare(world, before)
        ex, ey, bz = sim.read_yee()
        cex, cey, cbz = cpu.yee()
        pos, _, gains = sim.read_state()
        cpos, _, cgains = cpu.state()
        err = {"ex": relative_error(ex, cex), "ey": relative_error(ey, cey), "bz": relative_error(bz, cbz),
               "phi": relative_error(sim.read_potential(), cpu.potential_nodes()),
               "pos": float(np.max(np.abs(pos - cpos)))}
        print("512^2", before, err)
        for k in ("ex", "ey", "phi"):
            assert err[k] < 1e-5, err
        assert err["bz"] < 1e-4 and err["pos"] < 1e-4, err
        # From rest the first trim moves a gain by up to 0.4 in one step (1.0 ->
        # 0.586 here), driven by terminal potentials summed in FP32 over the whole
        # grid: 1.8e-5 of gain is 4e-5 of that step.
        for g, c in zip(gains, cgains):
            assert (g is None) == (c is None)
            if g is not None:
                assert abs(g - c) < 5e-5, (g, c)
    finally:
        orc.set_threads(1)


@gpu
@pytest.mark.parametrize("index", [0, 4])
def test_scenes_without_close_range_forces_follow_the_oracle(index):
    # electron_cutoff = 0 turns the close-range forces off (CloseRange::enabled,
    # close_range.rs:300-302): every electron takes the field-only push, no
    # tables are built, and the battery terminals read no pair potentials.
    world = W.scene(index, electron_cutoff=0.0)
    cpu, sim = reset_and_compare(world, 0)
    ex, ey, bz = sim.read_yee()
    cex, cey, cbz = cpu.yee()
    pos, _, _ = sim.read_state()
    err = {"ex": relative_error(ex, cex), "ey": relative_error(ey, cey),
           "phi": relative_error(sim.read_potential(), cpu.potential_nodes()),
           "pos": float(np.max(np.abs(pos - cpu.state()[0])))}
    sim.run_ticks(9)
    cpu.step(9)
    gex, gey, _ = sim.read_yee()
    cex, cey, _ = cpu.yee()
    err["ex_10"], err["ey_10"] = relative_error(gex, cex), relative_error(gey, cey)
    print("cutoff 0, scene", index, err)
    for k in ("ex", "ey", "phi"):
        assert err[k] < 1e-5, err
    assert err["pos"] < 1e-4, err
    assert err["ex_10"] < 1e-4 and err["ey_10"] < 1e-4, err
