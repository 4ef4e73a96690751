"""Numeric edge cases of the HIP backend through the C ABI, against the FP64
oracle: smoothing at three and four passes (the int64 deposit's headroom and
the FP64 second pass), the overflow flag past that headroom, and electrons at
the last FP32 fraction below a cell face (the summed ion table's last sample).
"""
import numpy as np
import pytest

from embloks_amd import world as W
from oracle import orc
from tests.golden.fixtures import relative_error
from tests.test_gpu_p3m import check_forces
from tests.test_gpu_parity import fixed_point_deposit, reset_and_compare

gpu = pytest.mark.gpu

LAST_BELOW_ONE = float(np.nextafter(np.float32(1.0), np.float32(0.0)))   # 0.99999994


def gpusim():
    from embloks_amd.gpusim import GpuSim
    return GpuSim


@gpu
@pytest.mark.parametrize("passes", [3, 4])
@pytest.mark.parametrize("before", [0, 10])
def test_more_smoothing_passes_follow_the_oracle(passes, before):
    # maxwell.rs:280-297 with `passes` 1-2-1 passes per axis (the reference takes
    # any count): the backend's x pass stays in int64 and the y pass runs in
    # FP64 beyond two passes. The battery cell holds 32 electrons on a dense ion.
    world = W.scene(0, smoothing_passes=passes)
    cpu, sim = reset_and_compare(world, before)
    ex, ey, _ = sim.read_yee()
    cex, cey, _ = cpu.yee()
    err = {"ex": relative_error(ex, cex), "ey": relative_error(ey, cey),
           "phi": relative_error(sim.read_potential(), cpu.potential_nodes()),
           "pos": float(np.max(np.abs(sim.read_state()[0] - cpu.state()[0])))}
    print("passes", passes, "before", before, err)
    for k in ("ex", "ey", "phi"):
        assert err[k] < 1e-5, err
    assert err["pos"] < 1e-4, err
    sim.sync()   # no overflow flag


@gpu
def test_deposit_stays_exact_at_four_passes():
    world = W.scene(5, smoothing_passes=4)
    sim = gpusim()(world)
    sim.run_ticks(5)
    pos, _, _ = sim.read_state()
    np.testing.assert_array_equal(sim.read_charge_fixed(), fixed_point_deposit(pos, world.cols, world.rows))


@gpu
def test_charge_past_the_fixed_point_headroom_raises():
    # Four passes weigh a node by up to 4^4 in the x pass: 4000 electrons on one
    # node (a cell's centre puts all of its CIC weight there) pass int64's
    # headroom. The backend flags it (EMB_E_NUMERIC at the next sync) instead of
    # wrapping silently.
    from embloks_amd.gpusim import EmbError

    w = W.World(16, 12, dict(W.abi.DEFAULTS, smoothing_passes=4, electron_cutoff=0.0))
    w.paint(W.block(5, 9, 4, 8), W.CONDUCTOR, resistance=1.0)
    crowd = np.tile([[6.5, 5.5]], (4000, 1))
    w.pos = np.concatenate([w.pos, crowd])
    w.vel = np.zeros_like(w.pos)
    sim = gpusim()(w)
    with pytest.raises(EmbError, match="error -5"):
        sim.sync()
    # At two passes a node may weigh 2^63 / 16 / 2^44 = 32768 electrons going
    # into either pass: 1000 on one node (6 x 1000 after the x pass) fit.
    w2 = w.copy()
    w2.settings = dict(w.settings, smoothing_passes=2)
    w2.pos = np.concatenate([w.pos[:-4000], np.tile([[6.5, 5.5]], (1000, 1))])
    w2.vel = np.zeros_like(w2.pos)
    sim2 = gpusim()(w2)
    sim2.sync()


@gpu
@pytest.mark.parametrize("cutoff", [1.0, 2.0, 3.0, 4.0])
def test_electrons_at_the_last_fraction_below_a_face(cutoff):
    # A fraction of 1 - 2^-24 rounds the ion table coordinate up to its last
    # sample; the bulk of a wide conductor takes the summed (uniform) ion table.
    w = W.World(24, 24, dict(W.abi.DEFAULTS, electron_cutoff=cutoff), seed=3)
    w.paint(W.block(3, 21, 3, 21), W.CONDUCTOR, resistance=1.0)
    rng = np.random.default_rng(int(cutoff * 10))
    pick = rng.choice(w.n_electrons, 300, replace=False)
    cell = np.floor(w.pos[pick])
    corner = rng.integers(0, 3, size=len(pick))
    w.pos[pick, 0] = cell[:, 0] + np.where(corner != 1, LAST_BELOW_ONE, w.pos[pick, 0] - cell[:, 0])
    w.pos[pick, 1] = cell[:, 1] + np.where(corner != 0, LAST_BELOW_ONE, w.pos[pick, 1] - cell[:, 1])
    want = orc.Sim(w).close_force()
    sim = gpusim()(w)
    pos = sim.read_state()[0]
    assert np.max(np.abs(pos - w.pos)) < 1e-7   # the fraction survives as FP32
    got, _ = sim.close_force()
    # Electrons moved to the same fraction in neighbouring cells sit exactly one
    # cell apart: at cutoff 1 many pairs lie on the cutoff itself, where FP32 and
    # FP64 may decide differently; each such electron must own one.
    err, edge = check_forces(w, got, want, max_edge=len(pick))
    print("cutoff", cutoff, "force error", err, "boundary pairs", edge)
    assert np.all(np.isfinite(got))
    assert err < 1e-5, err
