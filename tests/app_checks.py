"""The reference's whole-app physics checks, written once for any backend.

crates/embloks-core/tests/app.rs runs four checks of the physics through the
app (`App::step` = `Simulation::step` + the smoothed current, app.rs:61-65 /
controller.rs:547-552): the transformer's induction (app.rs:739-769), the
antenna's travelling wave (771-813), the energy flowing into the resistor
(571-610) and the steady B next to the battery (627-651). Here each is a
function of an `App`: the FP64 oracle (`OracleApp`, CPU tests) or the HIP
backend through its C ABI (`GpuApp`, GPU tests). The app-side helpers they
need (display averaging, Poynting vector, sine fit) are restated from
controller.rs:569-590, fields.rs:113-127 and app.rs:682-711.
"""
from __future__ import annotations

import math

import numpy as np

from embloks_amd import world as W

PI = math.pi


class App:
    """What the checks need of a backend: step, the cell current
    (physics::current_density), the node fields (Maxwell::node_fields) and the
    app's display average of them (App::display_fields)."""

    def __init__(self, world):
        self.world = world
        self.settings = world.settings
        self.dt = world.settings["dt"]
        self.time = world.time
        self._display = None
        self._display_time = 0.0

    def run(self, units: float):
        # app.rs:61-65: round(time_units / dt) ticks.
        ticks = int(round(units / self.dt))
        self._ticks(ticks)
        for _ in range(ticks):
            self.time += self.dt

    def step(self):
        self._ticks(1)
        self.time += self.dt

    def display_fields(self):
        # controller.rs:569-590: an average over display_smoothing_time of
        # simulated time, started from the instantaneous fields on first use.
        e, b = self.node_fields()
        tau = self.settings["display_smoothing_time"]
        if self._display is not None and tau > 0.0:
            alpha = 1.0 - math.exp(-(self.time - self._display_time) / tau)
            de, db = self._display
            de += alpha * (e - de)
            db += alpha * (b - db)
        else:
            self._display = (e.copy(), b.copy())
        self._display_time = self.time
        return self._display[0].copy(), self._display[1].copy()


class OracleApp(App):
    def __init__(self, world):
        from oracle import orc

        super().__init__(world)
        self.sim = orc.Sim(world)

    def _ticks(self, n):
        self.sim.step(n)

    def current(self):
        return self.sim.current()

    def node_fields(self):
        return self.sim.node_fields()


class GpuApp(App):
    def __init__(self, world):
        from embloks_amd.gpusim import GpuSim

        super().__init__(world)
        self.sim = GpuSim(world)

    def _ticks(self, n):
        self.sim.run_ticks(n)

    def current(self):
        return self.sim.read_current()

    def node_fields(self):
        return self.sim.node_fields()


def sine_fit(t, x, frequency):
    """Amplitude and phase (degrees) of x's component at `frequency`: least
    squares on sin, cos and a constant (app.rs:682-711)."""
    w = 2.0 * PI * frequency * np.asarray(t)
    basis = np.stack([np.sin(w), np.cos(w), np.ones_like(w)], axis=1)
    coef, *_ = np.linalg.lstsq(basis, np.asarray(x), rcond=None)
    return math.hypot(coef[0], coef[1]), math.degrees(math.atan2(coef[1], coef[0]))


def poynting(e, bz, settings):
    """S = E x B / mu0 per cell, 1/mu0 = c^2 / (4 pi k) (fields.rs:113-127)."""
    inv_mu0 = settings["speed_of_light"] ** 2 / (4.0 * PI * settings["coulomb_k"])
    return np.stack([inv_mu0 * e[:, :, 1] * bz, -inv_mu0 * e[:, :, 0] * bz], axis=2)


def transformer(make_app):
    """app.rs:739-769: the inner loop's load current is induced. Returns
    [(load / primary amplitude, phase difference)] at c = 8 and c = 40."""
    def primary(j):   # mean tangential current over three sides
        return (j[10:30, 4, 0].mean() + j[35, 8:24, 1].mean() - j[10:30, 27, 0].mean()) / 3.0

    def load(j):      # the inner resistor
        return j[27, 13:19, 1].mean()

    results = []
    for light in (8.0, 40.0):
        app = make_app(W.scene(5, speed_of_light=light))
        app.run(60.0)
        t, p, q = [], [], []
        for _ in range(int(100.0 / app.dt)):
            app.step()
            t.append(app.time)
            j = app.current()
            p.append(primary(j))
            q.append(load(j))
        fp, fq = sine_fit(t, p, 0.02), sine_fit(t, q, 0.02)
        results.append((fq[0] / fp[0], fq[1] - fp[1]))
    return results


def check_transformer(results):
    assert results[0][0] > 5.0 * results[1][0], results
    # Faraday + Lenz with a resistive load: I_load ~ -dI_primary/dt, 90 degrees behind.
    lag = (results[0][1] + 180.0) % 360.0 - 180.0
    assert abs(lag + 90.0) <= 40.0, (lag, results)


def antenna(make_app):
    """app.rs:771-813: (phase slope along the row, receiver amplitude), driven
    and with the source painted over with wire."""
    frequency, columns = 0.1, list(range(12, 27, 2))
    results = []
    for driven in (True, False):
        w = W.scene(6)
        if not driven:
            w.paint([(9, 16)], W.CONDUCTOR, resistance=0.1)   # nothing radiates
        app = make_app(w)
        app.run(30.0)
        t, fields, receiver = [], [[] for _ in columns], []
        for _ in range(int(30.0 / app.dt)):
            app.step()
            t.append(app.time)
            e, _ = app.display_fields()
            for f, x in zip(fields, columns):
                f.append(e[x, 16, 1])
            receiver.append(app.current()[29, 15:18, 1].mean())
        phases = [math.radians(sine_fit(t, f, frequency)[1]) for f in fields]
        for i in range(1, len(phases)):
            while phases[i] - phases[i - 1] > PI:
                phases[i] -= 2.0 * PI
            while phases[i] - phases[i - 1] < -PI:
                phases[i] += 2.0 * PI
        xs = np.array(columns, float)
        ps = np.array(phases)
        slope = float(((xs - xs.mean()) * (ps - ps.mean())).sum() / ((xs - xs.mean()) ** 2).sum())
        results.append((slope, sine_fit(t, receiver, frequency)[0]))
    return results


def check_antenna(results):
    wave = 2.0 * PI * 0.1 / 2.0   # the antenna scene's speed of light is 2
    assert abs(results[0][0] + wave) <= 0.35 * wave, results
    assert results[0][1] > 5.0 * results[1][1], results


def energy_into_resistor(make_app):
    """app.rs:571-610: time-averaged Poynting flux into the resistor, and the
    loop current against Ohm's law."""
    w = W.scene(0)
    resistor = [tuple(c) for c in np.argwhere((w.kind == W.CONDUCTOR) & (w.resistance == 5.0))]
    col = min(c for c, _ in resistor)
    rows = [r for c, r in resistor if c == col]
    app = make_app(w)
    app.run(40.0)
    flow = np.zeros((w.cols, w.rows, 2))
    current = []
    for i in range(int(20.0 / app.dt)):
        app.step()
        if i % 3 == 0:
            e, b = app.node_fields()
            flow += poynting(e, b, w.settings)
        current.append(app.current()[13:27, 7, 0].mean())
    wire = float(w.resistance[(w.kind == W.CONDUCTOR) & (w.resistance != 5.0)].sum())
    total = wire + 5.0 * len(resistor) + w.settings["battery_resistance"]
    emf = 6.0   # the scene's one stack
    return dict(into=[flow[col - 1, r, 0] for r in rows], outside=sum(abs(flow[col + 3, r, 0]) for r in rows),
                inside=sum(flow[col - 1, r, 0] for r in rows), current=float(np.mean(current)), ohm=emf / total)


def check_energy_into_resistor(r):
    # S points into the resistor from inside the loop; in 2D the loop is a
    # solenoid, so S is nearly zero outside it.
    assert all(v > 0.0 for v in r["into"]), r
    assert r["outside"] < 0.1 * r["inside"], r
    # Ohm's law around the loop (thin wires cut corners: 25%).
    assert abs(abs(r["current"]) - r["ohm"]) <= 0.25 * r["ohm"], r


def steady_b(make_app):
    """app.rs:627-651: the displayed B next to the battery holds steady.
    Returns the spread / level ratio."""
    app = make_app(W.scene(0))
    app.run(40.0)
    samples = []
    for i in range(int(20.0 / app.dt)):
        app.step()
        if i % 4 == 0:   # one displayed frame at 8x speed
            _, b = app.display_fields()
            samples.append([b[c, r] for c in range(11, 13) for r in range(13, 18)])
    s = np.array(samples)
    spread = float(np.sqrt(((s - s.mean(axis=0)) ** 2).mean(axis=0)).sum())
    level = float(np.abs(s.mean(axis=0)).sum())
    return spread / level


def check_steady_b(flicker):
    assert flicker < 0.3, flicker
