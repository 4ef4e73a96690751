"""The oracle against the reference's unit tests, closed forms and invariants.

These pin what the golden fixtures leave open: the DST-I beyond the fixture's
18 x 16 grid (the FFT path of the oracle, lengths > 96), FDTD with several
substeps, confinement, the cell-list close-range force against a brute-force
sum, and the physics checks of crates/embloks-core (Gauss's law, light-speed
fronts, absorbing margin, solenoid, screening, Ohm's law). Each test cites the
reference test it restates.
"""
import math

import numpy as np
import pytest

from embloks_amd import world as W
from oracle import orc

PI = math.pi


def test_dst_matches_its_definition_on_the_fft_path():
    # rustdct's unnormalised DST-I (maxwell.rs:64-66): y_k = sum_n x_n sin(pi (k+1)(n+1)/(N+1)).
    # cols x rows chosen so nx, ny > 96 take the oracle's FFT path.
    mx = orc.Maxwell(101, 77, orc.settings(absorbing_margin=4))
    rng = np.random.default_rng(0)
    a = rng.normal(size=(mx.nx, mx.ny))
    s = lambda n: np.sin(np.pi * np.outer(np.arange(1, n + 1), np.arange(1, n + 1)) / (n + 1))
    want = s(mx.nx) @ a @ s(mx.ny)
    got = mx.dst2(a)
    assert np.max(np.abs(got - want)) < 1e-10 * np.max(np.abs(want))
    # Applying it twice multiplies by (n + 1) / 2 per axis.
    twice = mx.dst2(got)
    np.testing.assert_allclose(twice, a * (mx.nx + 1) * (mx.ny + 1) / 4, rtol=0, atol=1e-9 * np.abs(a).max() * mx.nx * mx.ny)


@pytest.mark.parametrize("cols,rows,margin", [(10, 8, 4), (10, 8, 16), (40, 32, 4), (40, 32, 16), (120, 90, 16)])
def test_poisson_inverts_the_laplacian(cols, rows, margin):
    # maxwell.rs:594-610, plus one size on the FFT path.
    mx = orc.Maxwell(cols, rows, orc.settings(absorbing_margin=margin))
    i, j = np.meshgrid(np.arange(mx.nx), np.arange(mx.ny), indexing="ij")
    rhs = ((i * 7 + j * 3) % 11).astype(float) - 5.0
    phi = np.pad(mx.poisson(rhs), 1)
    lap = phi[2:, 1:-1] + phi[:-2, 1:-1] + phi[1:-1, 2:] + phi[1:-1, :-2] - 4 * phi[1:-1, 1:-1]
    assert np.max(np.abs(lap - rhs)) < 1e-9


def test_gauss_law_holds_exactly_and_static_field_is_2d_coulomb():
    # maxwell.rs:439-453
    s = orc.settings()
    mx = orc.Maxwell(40, 32, s)
    rho = np.zeros((mx.nx, mx.ny))
    rho[20 + 16, 16 + 16] = 1.0
    mx.enforce_gauss(rho)
    assert np.max(np.abs(mx.divergence() - 4 * PI * rho)) < 1e-12
    e, _ = mx.node_fields()
    for r in (4, 8):
        assert abs(math.hypot(*e[20 + r, 16]) / (2.0 / r) - 1) < 0.05


def gaussian_pulse(mx, col, row, width2):
    m = mx.s.absorbing_margin
    i, j = np.meshgrid(np.arange(mx.nx + 1), np.arange(mx.ny + 1), indexing="ij")
    x, y = i - 0.5 - col - m, j - 0.5 - row - m
    mx.set(bz=np.exp(-(x * x + y * y) / width2))


def test_waves_travel_at_the_speed_of_light():
    # maxwell.rs:467-487
    s = orc.settings(absorbing_margin=40, speed_of_light=2.0)
    mx = orc.Maxwell(80, 80, s)
    gaussian_pulse(mx, 40.0, 40.0, 0.5)
    rho, cur = np.zeros((mx.nx, mx.ny)), np.zeros((mx.nx, mx.ny, 2))
    fronts, t = [], 0.0
    for until in (5.0, 10.0, 15.0):
        while t < until - 1e-9:
            mx.step(rho, cur, s.dt)
            t += s.dt
        start = 40 + 40
        profile = mx.get()[2][start:, start]
        fronts.append(float(np.nonzero(np.abs(profile) > 1e-3)[0][-1]))
    for a, b in zip(fronts, fronts[1:]):
        assert abs((b - a) / 5.0 - s.speed_of_light) <= 0.25


def test_absorbing_border_swallows_outgoing_waves():
    # maxwell.rs:489-506
    s = orc.settings()
    mx = orc.Maxwell(40, 32, s)
    gaussian_pulse(mx, 20.0, 16.0, 2.0)
    rho, cur = np.zeros((mx.nx, mx.ny)), np.zeros((mx.nx, mx.ny, 2))
    energy = lambda: sum(float((a * a).sum()) * w for a, w in zip(mx.get()[:3], (1, 1, s.speed_of_light ** 2)))
    for _ in range(int(5.0 / s.dt)):
        mx.step(rho, cur, s.dt)
    before = energy()
    for _ in range(int(40.0 / s.dt)):
        mx.step(rho, cur, s.dt)
    assert energy() < 1e-3 * before


def test_fast_light_is_substepped_and_stays_stable():
    # maxwell.rs:508-519 (c dt sqrt 2 ~ 1.9: three FDTD substeps)
    s = orc.settings(speed_of_light=40.0)
    mx = orc.Maxwell(30, 30, s)
    gaussian_pulse(mx, 15.0, 15.0, 2.0)
    rho, cur = np.zeros((mx.nx, mx.ny)), np.zeros((mx.nx, mx.ny, 2))
    for _ in range(300):
        mx.step(rho, cur, s.dt)
    bz = mx.get()[2]
    assert np.all(np.isfinite(bz)) and np.max(np.abs(bz)) < 1.0


def test_staggered_gather_reads_uniform_fields_exactly():
    # maxwell.rs:521-530
    mx = orc.Maxwell(10, 8)
    mx.set(np.full((mx.nx + 1, mx.ny), 0.3), np.full((mx.nx, mx.ny + 1), -0.2), np.full((mx.nx + 1, mx.ny + 1), 0.7))
    for f in mx.gather([[2.3, 4.6], [7.9, 1.1]]):
        assert abs(f[0] - 0.3) < 1e-12 and abs(f[1] + 0.2) < 1e-12 and abs(f[2] - 0.7) < 1e-12


def test_smoothing_conserves_charge_and_visible_equals_whole():
    # maxwell.rs:532-562
    rng = np.random.default_rng(1)
    mx = orc.Maxwell(9, 7)
    values = rng.normal(size=(9, 7, 3))
    out = mx.smoothed(values)
    np.testing.assert_allclose(out.sum(axis=(0, 1)), values.sum(axis=(0, 1)), rtol=1e-12)
    # The same as the whole (otherwise zero) grid smoothed with the separable
    # binomial [1 4 6 4 1] / 16 per axis (two 1-2-1 passes).
    m = mx.s.absorbing_margin
    whole = np.zeros((mx.nx, mx.ny, 3))
    whole[m:m + 9, m:m + 7] = values
    b = np.array([1, 4, 6, 4, 1]) / 16.0
    conv = np.apply_along_axis(lambda v: np.convolve(v, b, mode="same"), 0, whole)
    conv = np.apply_along_axis(lambda v: np.convolve(v, b, mode="same"), 1, conv)
    assert np.max(np.abs(out - conv)) < 1e-14


def test_current_loop_is_a_solenoid_following_the_right_hand_rule():
    # maxwell.rs:564-586: B = 4 pi k I / c^2 inside, none outside.
    s = orc.settings()
    cur = np.zeros((30, 30, 2))
    for i in range(8, 22):
        cur[i, 8, 0] = 1.0
        cur[22, i, 1] = 1.0
        cur[i, 22, 0] = -1.0
        cur[8, i, 1] = -1.0
    mx = orc.Maxwell(30, 30, s)
    full = mx.smoothed(cur)
    rho = np.zeros((mx.nx, mx.ny))
    for _ in range(int(40.0 / s.dt)):
        mx.step(rho, full, s.dt)
    _, bz = mx.node_fields()
    inside, outside = bz[12:19, 12:19], bz[1:5, 1:5]
    mean = inside.mean()
    assert np.all(inside > 0) and inside.std() < 0.1 * mean
    assert np.all(np.abs(outside) < 0.1 * mean)
    assert abs(mean / (4 * PI / s.speed_of_light ** 2) - 1) < 0.1


def test_deposit_conserves_charge_and_is_exact_on_centres():
    # fields.rs:139-148, 222-233
    rng = np.random.default_rng(3)
    pos, q = rng.uniform(0, 5, size=(50, 2)), rng.normal(size=50)
    assert abs(orc.cic_deposit(pos, q, 5, 5).sum() / q.sum() - 1) < 1e-9
    on = orc.cic_deposit([[2.5, 1.5]], [3.0], 5, 5)
    assert on[2, 1] == 3.0 and on.sum() == 3.0
    v = np.zeros((5, 4))
    v[2, 1], v[0, 3] = 64.0, 8.0
    sp = orc.spread_over_cells(v)
    assert abs(sp.sum() - 72.0) < 1e-12 and abs(sp[2, 1] - 36.0) < 1e-12 and abs(sp[1, 1] - 6.0) < 1e-12
    assert abs(sp[1, 0] - 1.0) < 1e-12 and abs(sp[0, 3] - 8.0 * (7 / 8) ** 2) < 1e-12


def test_kick_terminal_velocity_and_huge_damping():
    # physics.rs:180-194
    v = np.zeros(2)
    for _ in range(2000):
        v = orc.kick(v, [2.0, -1.0], 4.0, 0.01)
    assert abs(v[0] - 0.5) < 1e-8 and abs(v[1] + 0.25) < 1e-8
    v = orc.kick([5.0, 0.0], [0.0, 0.0], 1e6, 0.1)
    assert 0.0 <= v[0] < 1e-4


def island_world(cells, battery=None):
    w = W.World(6, 5)
    w.paint(cells, W.CONDUCTOR, resistance=1.0)
    if battery:
        w.paint([battery[0]], W.BATTERY, emf=1.0, orientation=battery[1])
    return orc.Sim(w)


def test_confine_cases():
    # physics.rs:205-245
    sim = island_world([(1, 1), (2, 1)])
    p, v = sim.confine([1.9, 1.5], [2.1, 1.5], [1.0, 0.0], 0.5)
    assert list(p) == [2.1, 1.5] and list(v) == [1.0, 0.0]
    for old, new, vel in (([1.9, 1.5], [2.1, 1.5], [1.0, 0.0]), ([1.5, 1.1], [1.5, 0.9], [0.0, -1.0]),
                          ([1.1, 1.9], [0.9, 2.1], [-1.0, 1.0])):
        sim = island_world([(1, 1)])
        p, v = sim.confine(old, new, vel, 0.5)
        assert (math.floor(p[0]), math.floor(p[1])) == (1, 1)
        for a in range(2):
            if vel[a] != 0.0:
                assert abs(v[a] + 0.5 * vel[a]) < 1e-12
    sim = island_world([(0, 0)])
    p, _ = sim.confine([0.1, 0.5], [-0.1, 0.5], [-1.0, 0.0], 0.0)
    assert 0.0 <= p[0] < 1.0
    sim = island_world([(1, 1), (3, 1)], ((2, 1), W.RIGHT))
    p, _ = sim.confine([1.9, 1.5], [2.1, 1.5], [1.0, 0.0], 0.0)
    assert (math.floor(p[0]), math.floor(p[1])) == (2, 1)


def test_random_walk_never_leaves_its_island():
    # physics.rs:247-264
    rng = np.random.default_rng(5)
    w = W.World(12, 10)
    cells = [(c, r) for c in range(12) for r in range(10) if rng.random() < 0.55]
    w.paint(cells, W.CONDUCTOR, resistance=1.0)
    sim = orc.Sim(w)
    labels = sim.labels()
    pos = np.array([[c + 0.5, r + 0.5] for c, r in cells])
    start = np.array([labels[c, r] for c, r in cells])
    for _ in range(300):
        rest = rng.random()
        for k in range(len(pos)):
            moved = rng.uniform(-0.45, 0.45, size=2)
            pos[k], _ = sim.confine(pos[k], pos[k] + moved, moved, rest)
            assert labels[int(math.floor(pos[k][0])), int(math.floor(pos[k][1]))] == start[k]


def brute_force(world, sim):
    """CloseRange::force the slow way (close_range.rs:493-525): all pairs, every ion."""
    g, pot, square, half, _ = orc.tables()
    s = world.settings
    cutoff = s["electron_cutoff"]
    q = world.electron_charge
    pos = world.pos

    def radial(table, r2):
        if r2 >= cutoff * cutoff:
            return 0.0
        x = min(r2 * (4096 / (cutoff * cutoff)), 4096.0)
        i = int(x)
        return table[i] + (table[i + 1] - table[i]) * (x - i)

    def bil(k, u, v):
        n = square.shape[0]
        a = min(max(int(math.floor(u)), 0), n - 2)
        b = min(max(int(math.floor(v)), 0), n - 2)
        fu, fv = u - a, v - b
        return (square[a, b, k] * (1 - fu) * (1 - fv) + square[a + 1, b, k] * fu * (1 - fv)
                + square[a, b + 1, k] * (1 - fu) * fv + square[a + 1, b + 1, k] * fu * fv)

    f = np.zeros_like(pos)
    for i in range(len(pos)):
        for j in range(i + 1, len(pos)):
            d = pos[i] - pos[j]
            if math.hypot(*d) >= cutoff:
                continue
            gr = radial(g, d @ d)
            f[i] += s["coulomb_k"] * q * q * gr * d
            f[j] -= s["coulomb_k"] * q * q * gr * d
    for c, r in np.argwhere(world.kind > 0):
        for i in range(len(pos)):
            d = pos[i] - (c + 0.5, r + 0.5)
            if abs(d[0]) >= half or abs(d[1]) >= half:
                continue
            u, v = (d[0] + half) * 32, (d[1] + half) * 32
            f[i] += s["coulomb_k"] * s["cell_charge"] * q * np.array([bil(1, u, v), bil(2, u, v)])
    return f


def test_cell_list_force_matches_brute_force():
    # close_range.rs:527-544
    rng = np.random.default_rng(2)
    w = W.World(12, 9, seed=1)
    cells = [(c, r) for c in range(12) for r in range(9) if rng.random() < 0.5]
    w.paint(cells, W.CONDUCTOR, resistance=1.0)
    w.pos = np.clip(w.pos + rng.uniform(-0.3, 0.3, size=w.pos.shape), [0.001, 0.001], [11.999, 8.999])
    sim = orc.Sim(w)
    force = sim.close_force()
    ref = brute_force(w, sim)
    assert np.max(np.abs(force - ref)) < 1e-12 * np.max(np.abs(ref))


def test_stacked_pair_repels_and_momentum_balances():
    # close_range.rs:570-584
    w = W.World(10, 10)
    w.paint([(4, 4), (5, 4)], W.CONDUCTOR, resistance=1.0)
    w.pos[:] = [5.0, 4.5]
    w.pos[0, 0] -= 0.05
    w.pos[1, 0] += 0.05
    f = orc.Sim(w).close_force()
    assert f[0, 0] < -0.5 and f[1, 0] > 0.5
    assert np.all(np.abs(f.sum(axis=0)) < 1e-9)


def test_grid_plus_close_range_gives_the_compact_force():
    # close_range.rs:546-568
    w = W.World(24, 24)
    rng = np.random.default_rng(0)
    for r in (0.25, 0.5, 1.0, 2.0, 2.8):
        totals = []
        for _ in range(40):
            a = rng.uniform(10.0, 13.0, size=2)
            ang = rng.uniform(0, 2 * PI)
            u = np.array([math.cos(ang), math.sin(ang)])
            w.pos = np.array([a, a + r * u])
            w.vel = np.zeros((2, 2))
            w.settings = dict(w.settings, carriers_per_cell=4, cell_charge=4.0)  # charge -1, mass 1
            sim = orc.Sim(w)
            mx = orc.Maxwell(24, 24)
            ex, ey, bz = sim.yee()
            mx.set(ex, ey, bz)
            grid = mx.gather(w.pos)[1]
            close = sim.close_force()[1]
            totals.append((-grid[0] + close[0]) * u[0] + (-grid[1] + close[1]) * u[1])
        assert abs(np.mean(totals) / (r * orc.compact_g(r * r, 0.25)) - 1) < 0.05


def block_world():
    w = W.World(40, 32)
    w.paint(W.block(14, 26, 13, 19), W.CONDUCTOR, resistance=1.0)
    return w


def test_neutral_conductor_stays_quiet():
    # physics.rs:273-280
    sim = orc.Sim(block_world())
    sim.step(int(round(5.0 / (1 / 30))))
    _, vel, _ = sim.state()
    assert np.max(np.hypot(vel[:, 0], vel[:, 1])) < 0.02


def test_conductor_screens_external_charge():
    # physics.rs:282-306
    w = block_world()
    w.place_charges([(8, 16)], 5.0)
    inner = (slice(15, 25), slice(14, 18))
    sim = orc.Sim(w)

    def mean_e():
        e, _ = sim.node_fields()
        return math.hypot(e[inner][..., 0].mean(), e[inner][..., 1].mean())

    unscreened = mean_e()
    sim.step(300)
    assert mean_e() < 0.15 * unscreened
    pos, _, _ = sim.state()
    assert pos[:, 0].mean() < 20.0 - 0.01


def test_loop_current_follows_ohms_law():
    # crates/embloks-gpu/tests/sim.rs:96-115 (and battery.rs:404-415), on the oracle.
    w = W.World(30, 20)
    lp = W.outline(4, 23, 3, 12)
    w.paint(lp, W.CONDUCTOR, resistance=0.5)
    w.paint([(14, 12)], W.BATTERY, emf=2.0, orientation=W.RIGHT)
    expected = 2.0 / ((len(lp) - 1) * 0.5 + 0.1)
    sim = orc.Sim(w)
    sim.step(900)
    samples = []
    for _ in range(20):
        sim.step(15)
        samples.append(sim.current()[6:22, 3, 0].mean())
    assert abs(-np.mean(samples) / expected - 1) < 0.25


@pytest.mark.parametrize("index", [0, 5])
def test_subset_force_equals_the_full_force(index):
    # The sampled one-sided sum (used against the 4M-electron GPU run) against
    # CloseRange::force's Newton-3 loop over the whole scene.
    w = W.scene(index)
    full = orc.Sim(w).close_force()
    sub = orc.close_force_subset(w, np.arange(w.n_electrons))
    assert np.abs(full - sub).max() < 1e-12 * np.abs(full).max()
    pick = np.array([3, 0, 17])
    assert np.array_equal(orc.close_force_subset(w, pick), sub[pick])


def test_clustered_scene_is_neutral_per_cell():
    w = W.clustered_charges(side=128, n_electrons=20000, n_clusters=5, seed=2)
    cell = np.floor(w.pos).astype(np.int64)
    counts = np.zeros((128, 128))
    np.add.at(counts, (cell[:, 0], cell[:, 1]), 1.0)
    assert np.array_equal(counts, w.density)
    assert np.array_equal(w.kind > 0, counts > 0)
