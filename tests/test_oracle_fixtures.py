"""The FP64 oracle against the reference's golden fixtures.

Mirrors crates/embloks-core/tests/fixtures.rs:48-186 with its tolerances: the
oracle must reproduce what the Python original computed (Poisson/Gauss/FDTD at
1e-12, deposit/spread/smoothing at 1e-12, close-range tables at 1e-9, and six
example scenes over 30 ticks).
"""
import numpy as np
import pytest

from oracle import orc
from tests.golden.fixtures import SCENES, array, load, relative_error, world_from


def test_poisson_and_gauss_match_scipy():
    f = load("maxwell")
    s = orc.settings(absorbing_margin=4)
    mx = orc.Maxwell(10, 8, s)
    assert (mx.nx, mx.ny) == (f["nx"], f["ny"])
    mx.set(array(f["ex0"]), array(f["ey0"]), array(f["bz0"]))
    rho = array(f["rho"])
    mx.enforce_gauss(rho)
    ex, ey, bz, phi = mx.get()
    g = f["after_gauss"]
    assert relative_error(phi, array(g["potential"])) < 1e-12
    assert relative_error(ex, array(g["ex"])) < 1e-12
    assert relative_error(ey, array(g["ey"])) < 1e-12
    mx.step(rho, array(f["current"]), s.dt)
    ex, ey, bz, phi = mx.get()
    after = f["after_step"]
    for ours, name in ((ex, "ex"), (ey, "ey"), (bz, "bz"), (phi, "potential")):
        assert relative_error(ours, array(after[name])) < 1e-12, name


def test_deposit_spread_and_smoothing_match_numpy():
    f = load("deposit")
    pos = array(f["pos"])
    q = array(f["q"])
    assert relative_error(orc.cic_deposit(pos, q, 10, 8), array(f["deposit"])) < 1e-12
    values = array(f["spread_in"])
    assert relative_error(orc.spread_over_cells(values), array(f["spread_out"])) < 1e-12
    mx = orc.Maxwell(10, 8, orc.settings(absorbing_margin=4))
    assert relative_error(mx.smoothed(values), array(f["smoothed"])) < 1e-12


def test_close_range_tables_match_python():
    f = load("close_range_tables")
    r, force = orc.grid_force(2, 3.0)
    assert relative_error(r, array(f["grid_force_r"])) < 1e-12
    assert relative_error(force, array(f["grid_force_f"])) < 1e-9
    g, pot, square, half, stencil = orc.tables(0.25, 2, 3.0)
    assert relative_error(g, array(f["g"])) < 1e-9
    assert relative_error(pot, array(f["potential"])) < 1e-9
    assert half == f["half"]
    assert list(square.shape) == f["square_shape"]
    assert relative_error(square[::8, ::8, :], array(f["square_sampled"])) < 1e-9
    assert relative_error(stencil, array(f["ion_stencil"])) < 1e-9


@pytest.mark.parametrize("name", SCENES)
def test_scene_matches_python(name):
    f = load(name)
    world = world_from(f)
    sim = orc.Sim(world)
    np.testing.assert_array_equal(sim.labels(), array(f["initial"]["labels"]), "island labels")
    assert relative_error(sim.potential(), array(f["initial_potential"])) < 1e-9, "initial potential"
    assert relative_error(sim.close_force(), array(f["initial_close_range_force"])) < 1e-9, "close-range force"
    sim.step(f["ticks"])
    pos, vel, gains = sim.state()
    assert np.max(np.abs(pos - array(f["final_pos"]))) < 1e-6, "positions"
    e, b = sim.node_fields()
    assert relative_error(e, array(f["final_e"])) < 1e-6, "E"
    assert relative_error(b, array(f["final_b"])) < 1e-6, "B"
    phi = sim.potential_nodes()
    m = world.settings["absorbing_margin"]
    visible = phi[m:m + world.cols, m:m + world.rows]
    assert relative_error(visible, array(f["final_potential"])) < 1e-6, "V"
    assert relative_error(sim.current(), array(f["final_current"])) < 1e-5, "I"
    expected = f["gains"]
    assert len(gains) == len(expected)
    for g, e in zip(gains, expected):
        assert (g is None and e is None) or abs(g - e) < 1e-6, (gains, expected)
