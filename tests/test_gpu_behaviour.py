"""Circuit behaviour over long runs, the reference's own GPU tests
(crates/embloks-gpu/tests/sim.rs:96-152) through the C ABI, with its bars:
Ohm's law around a loop, screening of an external charge by a conductor, and a
series battery stack driving three times the current of a single cell.
"""
import numpy as np
import pytest

from embloks_amd import world as W

gpu = pytest.mark.gpu


def gpusim():
    from embloks_amd.gpusim import GpuSim
    return GpuSim


@gpu
def test_loop_current_follows_ohms_law():
    # sim.rs:96-115
    s = dict(W.abi.DEFAULTS)
    w = W.World(30, 20, s, seed=0)
    loop = W.outline(4, 23, 3, 12)
    w.paint(loop, W.CONDUCTOR, resistance=0.5)
    w.paint([(14, 12)], W.BATTERY, emf=2.0, orientation=W.RIGHT)
    expected = 2.0 / ((len(loop) - 1) * 0.5 + s["battery_resistance"])
    sim = gpusim()(w)
    sim.run_ticks(int(30.0 / s["dt"]))
    samples = []
    for _ in range(20):
        sim.run_ticks(int(0.5 / s["dt"]))
        samples.append(sim.read_current()[6:22, 3, 0].mean())
    mean = float(np.mean(samples))
    print("loop current", mean, "Ohm's law", expected)
    assert abs(-mean / expected - 1.0) < 0.25, (mean, expected)


@gpu
def test_conductor_screens_external_charge():
    # sim.rs:117-134
    w = W.World(40, 32, dict(W.abi.DEFAULTS), seed=0)
    w.paint([(c, r) for c in range(14, 26) for r in range(13, 19)], W.CONDUCTOR, resistance=1.0)
    w.place_charges([(8, 16)], 5.0)

    def mean_e(e):
        inner = e[15:25, 14:18, :]
        n = inner.size / 2
        return float(np.hypot(inner[..., 0].sum() / n, inner[..., 1].sum() / n))

    sim = gpusim()(w)
    unscreened = mean_e(sim.read_display()[0])
    sim.run_ticks(int(10.0 / w.settings["dt"]))
    screened = mean_e(sim.read_display()[0])
    print("mean E inside the conductor", unscreened, "->", screened)
    assert screened < 0.15 * unscreened, (screened, unscreened)


@gpu
def test_series_scene_drives_three_times_the_current():
    # sim.rs:136-152
    world = W.scene(2)
    sim = gpusim()(world)
    sim.run_ticks(int(30.0 / world.settings["dt"]))
    left = right = 0.0
    for _ in range(60):
        sim.run_ticks(15)
        current = sim.read_current()
        left += current[5:16, 8, 0].mean()
        right += current[24:35, 8, 0].mean()
    print("series / single current", right / left)
    assert abs(right / left / 3.0 - 1.0) < 0.2, right / left
