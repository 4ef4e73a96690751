"""The edit path: GpuSim::upload of an edited world while the simulation runs
(crates/embloks-gpu/src/sim.rs:395-529) through the C ABI.

The first test is the reference's own (crates/embloks-gpu/tests/sim.rs:196-214).
The second paints a new conductor with its carriers (more electrons than the
scene had), uploads, and checks the upload's invariants (Gauss's law over the
new charge, the electron count, the fields carried), then one tick against the
FP64 oracle started from the uploaded state.
"""
import numpy as np
import pytest

from embloks_amd import world as W
from oracle import orc
from tests.golden.fixtures import relative_error

gpu = pytest.mark.gpu


def gpusim():
    from embloks_amd.gpusim import GpuSim
    return GpuSim


def take_state(sim, world):
    """World::apply of GpuSim::read_state: the device's electrons and gains."""
    pos, vel, gains = sim.read_state()
    world.pos, world.vel = pos.copy(), vel.copy()
    world.gains = list(gains)
    world.time = sim.dims().time


@gpu
def test_edits_keep_the_fields_going_and_restart_the_display():
    world = W.scene(0)
    sim = gpusim()(world)
    sim.run_ticks(60)
    take_state(sim, world)
    gains_before = list(world.gains)
    # Paint a test charge: Gauss's law takes it in at once, without a tick.
    world.place_charges([(5, 5)], 5.0)
    sim.upload(world, reset_display=True)
    e, _ = sim.read_display()
    assert e[6, 5, 0] > 0.5, e[6, 5, 0]
    assert sim.read_state()[2] == gains_before
    residual, _, _ = sim.diagnostics()
    assert residual < 1e-4, residual


@gpu
def test_painting_a_wire_with_carriers_then_ticking_follows_the_oracle():
    world = W.scene(4)
    sim = gpusim()(world)
    sim.run_ticks(40)
    take_state(sim, world)
    n0 = world.n_electrons
    ex0, ey0, bz0 = sim.read_yee()
    world.paint(W.line(2, 2, 2, 25), W.CONDUCTOR, resistance=0.5)   # a new wire and its electrons
    assert world.n_electrons > n0
    sim.upload(world, reset_display=False)
    assert sim.dims().n_electrons == world.n_electrons
    pos, vel, _ = sim.read_state()
    # the device keeps a cell and an FP32 fraction: the new electrons round to 2^-24 of a cell
    assert np.max(np.abs(pos - world.pos)) < 1e-7
    residual, charge, _ = sim.diagnostics()
    assert residual < 1e-4, residual
    q0 = world.settings["cell_charge"] / world.settings["carriers_per_cell"]
    assert charge == -world.n_electrons * q0
    # The upload corrects E to the new charges with a gradient and leaves Bz as it was.
    ex, ey, bz = sim.read_yee()
    np.testing.assert_array_equal(bz, bz0)
    assert relative_error(ex, ex0) > 0.0
    # One tick from the uploaded state, on both backends.
    cpu = orc.Sim(world)
    cpu.write_state(pos, vel, ex, ey, bz)
    cpu.step(1)
    sim.run_ticks(1)
    gex, gey, _ = sim.read_yee()
    cex, cey, _ = cpu.yee()
    gpos = sim.read_state()[0]
    err = {"ex": relative_error(gex, cex), "ey": relative_error(gey, cey),
           "pos": float(np.max(np.abs(gpos - cpu.state()[0])))}
    print(err)
    assert err["ex"] < 1e-5 and err["ey"] < 1e-5 and err["pos"] < 1e-4, err


@gpu
def test_erasing_part_of_a_wire_then_ticking_follows_the_oracle():
    # World::paint with vacuum over a running circuit (world.rs:180-238: the
    # erased cells' electrons leave by island quota, stranded ones move), then
    # GpuSim::upload and one tick against the oracle from the uploaded state.
    world = W.scene(0)
    sim = gpusim()(world)
    sim.run_ticks(30)
    take_state(sim, world)
    n0 = world.n_electrons
    world.paint(W.line(20, 22, 24, 24), W.VACUUM)   # cut the loop's bottom wire
    assert world.n_electrons == n0 - 3 * world.settings["carriers_per_cell"]
    sim.upload(world, reset_display=False)
    assert sim.dims().n_electrons == world.n_electrons
    pos, vel, _ = sim.read_state()
    cell = np.floor(pos).astype(np.int64)
    assert np.all(world.kind[cell[:, 0], cell[:, 1]] > 0)
    ex, ey, bz = sim.read_yee()
    cpu = orc.Sim(world)
    cpu.write_state(pos, vel, ex, ey, bz)
    cpu.step(1)
    sim.run_ticks(1)
    gex, gey, _ = sim.read_yee()
    cex, cey, _ = cpu.yee()
    err = {"ex": relative_error(gex, cex), "ey": relative_error(gey, cey),
           "pos": float(np.max(np.abs(sim.read_state()[0] - cpu.state()[0])))}
    print(err)
    assert err["ex"] < 1e-5 and err["ey"] < 1e-5 and err["pos"] < 1e-4, err


@gpu
def test_close_force_after_an_upload_that_adds_electrons():
    # The close-force output grows with the electron count an upload brings.
    from tests.test_gpu_p3m import FIXTURE_FORCE_TOL, check_forces

    world = W.scene(4)
    sim = gpusim()(world)
    sim.close_force()
    n0 = world.n_electrons
    world.paint(W.block(2, 4, 2, 26), W.CONDUCTOR, resistance=0.5)
    assert world.n_electrons > n0
    sim.upload(world)
    got, _ = sim.close_force()
    assert got.shape == (world.n_electrons, 2)
    pos = sim.read_state()[0]
    w = world.copy()
    w.pos = pos
    want = orc.Sim(w).close_force()
    err, edge = check_forces(w, got, want, tol=FIXTURE_FORCE_TOL)
    print("force error", err, "boundary pairs", edge)
    assert err < FIXTURE_FORCE_TOL, err
