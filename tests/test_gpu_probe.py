"""The info-panel probe (GpuSim::read_probe, crates/embloks-gpu/src/sim.rs:663-715;
probe_cell and probe_grid, fields.wgsl:368-466) through the C ABI.

The two behavioural tests are the reference's own (crates/embloks-gpu/tests/sim.rs:
153-173, 175-194). The others pin every reading: the potential against the FP64
oracle's Simulation::potential, and the display, current, wire-current, mean-|E|
and histogram readings against the same quantities formed in numpy from the
backend's own readbacks (which the parity tests compare with the oracle).
"""
import numpy as np
import pytest

from embloks_amd import world as W
from oracle import orc
from tests.golden.fixtures import load, world_from

gpu = pytest.mark.gpu


def gpusim():
    from embloks_amd.gpusim import GpuSim
    return GpuSim


def hist_bins(m):
    """probe_grid's log bins (fields.wgsl:422-427) in FP32: -1 for m <= 0."""
    m = np.asarray(m, dtype=np.float32)
    out = np.full(m.shape, -1, dtype=np.int64)
    pos = m > 0
    x = (np.log2(m[pos]) * np.float32(0.30103) + np.float32(12.0)) / np.float32(18.0) * np.float32(256.0)
    out[pos] = np.clip(x, 0.0, 255.0).astype(np.int64)
    return out


def histogram(m):
    b = hist_bins(m)
    return np.bincount(b[b >= 0], minlength=256)


@gpu
def test_probe_counts_electrons_and_reads_the_wire_current():
    # crates/embloks-gpu/tests/sim.rs:175-194 (parallel paths)
    world = W.scene(4)
    sim = gpusim()(world)
    probe = sim.read_probe((5, 15), 2)
    assert probe["cell"] is not None and probe["electrons"] > 0
    assert probe["mean_e_in_wires"] is not None
    assert probe["histograms"][0].sum() > 0 and np.all(probe["histograms"][1] == 0)
    sim.run_ticks(int(50.0 / world.settings["dt"]))
    thin = wide = 0.0
    for _ in range(20):
        sim.run_ticks(30)
        thin += sim.read_probe((20, 10), 2)["wire_current"][0]
        wide += sim.read_probe((20, 22), 2)["wire_current"][0]
    assert sim.read_probe((20, 22), 2)["wire_current"][1] == 4
    assert abs(wide / thin / 2.0 - 1.0) < 0.3, wide / thin
    assert sim.read_probe(None, 2)["cell"] is None


@gpu
def test_capacitor_reaches_the_emf_and_the_probe_reads_it():
    # crates/embloks-gpu/tests/sim.rs:153-173
    world = W.scene(3)
    emf = 8.0
    sim = gpusim()(world)
    sim.run_ticks(int(60.0 / world.settings["dt"]))

    def plate(col):
        return np.mean([sim.read_probe((col, r), 1)["potential"] for r in range(10, 22)])

    difference = plate(22) - plate(16)
    assert abs(difference / emf - 1.0) < 0.1, difference
    battery = sim.read_probe((19, 28), 1)
    assert battery["stack_state"] == 2
    assert abs((battery["v_high"] - battery["v_low"]) / emf - 1.0) < 0.1


@gpu
@pytest.mark.parametrize("name", ["scene_f1", "scene_f2", "scene_f6"])
def test_probe_potential_matches_the_oracle(name):
    world = world_from(load(name))
    cpu = orc.Sim(world)
    sim = gpusim()(world)
    want = cpu.potential()   # Simulation::potential at the cell centres (simulation.rs:48-55)
    scale = np.max(np.abs(want))
    rng = np.random.default_rng(1)
    cells = [(int(c), int(r)) for c, r in zip(rng.integers(0, world.cols, 12), rng.integers(0, world.rows, 12))]
    for c, r in cells + [(0, 0), (world.cols - 1, world.rows - 1)]:
        got = sim.read_probe((c, r), 2)["potential"]
        assert abs(got - want[c, r]) < 1e-5 * scale, (c, r, got, want[c, r])


def wire_current(kind, current, c, r):
    cols, rows = kind.shape
    ix, iy = current[c, r]
    along = 0 if abs(ix) >= abs(iy) else 1
    total, width = current[c, r, along], 1
    dc, dr = (1, 0) if along == 1 else (0, 1)
    for side in (-1, 1):
        ac, ar = c + side * dc, r + side * dr
        while 0 <= ac < cols and 0 <= ar < rows and kind[ac, ar] != 0:
            total += current[ac, ar, along]
            width += 1
            ac, ar = ac + side * dc, ar + side * dr
    return total, width


@gpu
@pytest.mark.parametrize("index", [0, 4, 5])
def test_probe_reads_the_tick_state(index):
    world = W.scene(index)
    sim = gpusim()(world)
    sim.run_ticks(90)
    e, b = sim.read_display()
    cur = sim.read_smoothed_current()
    pos, _, gains = sim.read_state()
    s = world.settings
    c2 = s["speed_of_light"] ** 2
    inv_mu0 = np.float32(c2 / (4.0 * np.pi * s["coulomb_k"]))
    kind = world.kind.reshape(world.cols, world.rows)
    wire = kind == W.CONDUCTOR
    counts = np.zeros((world.cols, world.rows), dtype=np.int64)
    cc = np.clip(np.floor(pos[:, 0]).astype(np.int64), 0, world.cols - 1)
    rr = np.clip(np.floor(pos[:, 1]).astype(np.int64), 0, world.rows - 1)
    np.add.at(counts, (cc, rr), 1)
    conductive = np.argwhere(kind != 0)
    for stride in (1, 2, 3):
        probe = sim.read_probe(tuple(conductive[0]), stride)
        # probe_grid
        ex32, ey32 = e[..., 0].astype(np.float32), e[..., 1].astype(np.float32)
        emag = np.sqrt(ex32 * ex32 + ey32 * ey32)
        mean = float(emag[wire].astype(np.float64).mean())
        assert abs(probe["mean_e_in_wires"] - mean) <= 1e-5 * max(mean, 1e-30)
        arrows = np.zeros_like(wire)
        arrows[stride // 2::stride, stride // 2::stride] = True
        hb = np.abs(b).astype(np.float32) * np.float32(c2)
        hs = inv_mu0 * emag * np.abs(b).astype(np.float32)
        for got, want in zip(probe["histograms"], (histogram(emag[arrows]), histogram(hb), histogram(hs[arrows]))):
            assert got.sum() == want.sum()
            assert np.abs(got.astype(np.int64) - want).sum() <= 4   # log2 rounding at a bin edge
    rng = np.random.default_rng(index)
    picks = [tuple(conductive[k]) for k in rng.integers(0, len(conductive), 10)]
    for c, r in picks:
        p = sim.read_probe((c, r), 2)
        assert p["cell"] == (c, r)
        assert p["electrons"] == counts[c, r]
        assert p["e"] == (e[c, r, 0], e[c, r, 1]) and p["b"] == b[c, r]
        assert p["current"] == (cur[c, r, 0], cur[c, r, 1])
        s_want = (inv_mu0 * np.float32(e[c, r, 1]) * np.float32(b[c, r]),
                  -inv_mu0 * np.float32(e[c, r, 0]) * np.float32(b[c, r]))
        np.testing.assert_allclose(p["s"], s_want, rtol=1e-6, atol=1e-30)
        total, width = wire_current(kind, cur, c, r)
        assert p["wire_current"][1] == width
        assert abs(p["wire_current"][0] - total) <= 1e-5 * np.abs(cur).max() * width
    # A stacked battery cell reads its stack; the gain is the one read_state gives.
    bat = np.argwhere(kind == W.BATTERY)
    if len(bat) and gains:
        p = sim.read_probe(tuple(bat[0]), 2)
        assert p["stack_state"] in (1, 2)
        assert any(abs(p["gain"] - (1.0 if g is None else g)) < 1e-6 for g in gains)
    assert sim.read_probe((world.cols, 0), 2)["cell"] is None
