"""Scene construction: the inputs the physics backend is given.

A `World` mirrors the reference's world model as flat arrays
(crates/embloks-core/src/world.rs:24-38, materials.rs:80-146): per-cell
material kind, resistance, carrier density, EMF, orientation and AC frequency
(x-major, [col, row]), plus electrons (positions and velocities in cell units),
fixed test charges and per-stack battery gains.

Painting follows World::paint (world.rs:74-93): conductors get density 1,
batteries `battery_density` and `battery_resistance`, and every cell that
gains density spawns `density * carriers_per_cell` electrons on the rank-1
lattice of world.rs:266-280 with a small jitter. The jitter comes from numpy's
PCG64, not the reference's rand_pcg stream, so scenes built here match the
reference's scenes in layout and count, not in the last digits of electron
positions (the golden fixtures carry their own positions for that).

Also here: the example scenes F1-F7 (scenes.rs:55-134), the 30%-wire world of
the reference's bench (crates/embloks-gpu/examples/bench.rs:16-22) and seeded
generators for the benchmark workloads (a DC ring, random-rectangle circuits
and the serving stream).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

VACUUM, CONDUCTOR, BATTERY = 0, 1, 2
RIGHT, DOWN, LEFT, UP = 0, 1, 2, 3
SPAWN_JITTER = 0.02  # world.rs:15


def fill_pattern(n: int) -> np.ndarray:
    """n points spread evenly over a unit cell (world.rs:266-280)."""
    target = 0.382 * n
    best, best_d = 1, math.inf
    for g in range(1, n + 1):
        if math.gcd(g, n) == 1:
            d = abs(g - target)
            if d < best_d:  # the first minimum, like Iterator::min_by
                best, best_d = g, d
    i = np.arange(n)
    return np.stack([(i + 0.5) / n, ((i * best) % n) / n + 0.5 / n], axis=1)


@dataclass
class World:
    cols: int
    rows: int
    settings: dict = field(default_factory=lambda: dict(abi.DEFAULTS))
    seed: int = 0

    def __post_init__(self):
        shape = (self.cols, self.rows)
        self.kind = np.zeros(shape, np.uint8)
        self.resistance = np.zeros(shape)
        self.density = np.zeros(shape)
        self.emf = np.zeros(shape)
        self.orientation = np.zeros(shape, np.uint8)
        self.frequency = np.zeros(shape)
        self.pos = np.zeros((0, 2))
        self.vel = np.zeros((0, 2))
        self.fixed_pos = np.zeros((0, 2))
        self.fixed_charge = np.zeros(0)
        self.gains = None  # per stack, None = no entry
        self.time = 0.0
        self.rng = np.random.Generator(np.random.PCG64(self.seed))

    # --- painting (world.rs:74-93, materials.rs:134-146) -----------------------------

    def paint(self, cells, material, resistance=0.0, emf=0.0, orientation=RIGHT, frequency=0.0, jitter=SPAWN_JITTER):
        cells = np.asarray(list(cells), dtype=np.int64).reshape(-1, 2)
        s = self.settings
        density = 1.0
        if material == BATTERY:
            resistance = s["battery_resistance"]
            density = float(s["battery_density"])
        old = self.density.copy()
        c, r = cells[:, 0], cells[:, 1]
        conductive = material in (CONDUCTOR, BATTERY)
        new_density = density if conductive else 0.0
        # The islands before the paint decide where removed carriers come from
        # (world.rs:80-91); only needed when some cell loses density.
        old_labels = self.labels() if (old[c, r] > new_density).any() else None
        self.kind[c, r] = material
        self.resistance[c, r] = resistance if conductive else 0.0
        self.density[c, r] = new_density
        self.emf[c, r] = emf if material == BATTERY else 0.0
        self.orientation[c, r] = orientation if material == BATTERY else 0
        self.frequency[c, r] = frequency if material == BATTERY else 0.0
        n = s["carriers_per_cell"]
        change = np.rint((self.density - old) * n).astype(np.int64)
        if (change < 0).any():
            self._remove_carriers(np.maximum(-change, 0), old_labels)
        if (change > 0).any():
            self.spawn(np.maximum(change, 0), jitter)

    def labels(self) -> np.ndarray:
        """Island id per cell (-1 off conductors): conductive cells joined through
        shared edges, never corners, numbered in scan order of their first cell
        (columns outer, connectivity.rs:20-62)."""
        from scipy import ndimage

        lab, _ = ndimage.label(self.kind > 0)   # 4-connected; features numbered in C (scan) order
        return lab.astype(np.int64) - 1

    def spawn(self, counts: np.ndarray, jitter=SPAWN_JITTER):
        """Add counts[c, r] electrons per cell, spread over it (world.rs:148-164)."""
        new = []
        for count in sorted(set(int(v) for v in counts.reshape(-1) if v > 0)):
            pattern = fill_pattern(count)
            cells = np.argwhere(counts == count)  # row-major: columns outer, like indexed_iter
            base = np.repeat(cells.astype(np.float64), count, axis=0) + np.tile(pattern, (len(cells), 1))
            j = self.rng.uniform(-jitter, jitter, size=base.shape) if jitter > 0 else 0.0
            new.append(base + j)
        if new:
            added = np.concatenate(new)
            self.pos = np.concatenate([self.pos, added])
            self.vel = np.concatenate([self.vel, np.zeros_like(added)])

    def _remove_carriers(self, counts: np.ndarray, old_labels: np.ndarray):
        """Take away counts[c, r] electrons for the cells that lost density
        (world.rs:180-238). Per island of the old materials: the electrons in the
        cells that lost density leave first, those in erased (now non-conductive)
        cells before the others; if they fall short of the island's quota, the
        electrons nearest to those cells' centres follow. Electrons still in
        erased cells then move to the nearest remaining conductive cell of their
        old island (with the spawn jitter and zero velocity), or vanish if nothing
        of it is left."""
        n = len(self.pos)
        conductive = self.kind > 0
        cell = np.clip(np.floor(self.pos).astype(np.int64), 0, [self.cols - 1, self.rows - 1])
        island = old_labels[cell[:, 0], cell[:, 1]]
        lost = counts > 0
        in_lost = lost[cell[:, 0], cell[:, 1]]
        stranded = ~conductive[cell[:, 0], cell[:, 1]]
        keep = np.ones(n, bool)
        for isl in np.unique(old_labels[lost]):
            lost_here = np.argwhere(lost & (old_labels == isl))   # scan order
            quota = int(counts[lost_here[:, 0], lost_here[:, 1]].sum())
            members = island == isl
            inside = np.nonzero(members & in_lost)[0]
            inside = inside[np.argsort(~stranded[inside], kind="stable")]   # erased cells first
            keep[inside[:quota]] = False
            if len(inside) < quota:
                others = np.nonzero(members & ~in_lost)[0]
                centres = lost_here + 0.5
                d2 = ((self.pos[others, None, :] - centres[None, :, :]) ** 2).sum(axis=2)
                dist = np.sqrt(d2.min(axis=1))
                nearest = others[np.argsort(dist, kind="stable")]
                keep[nearest[: quota - len(inside)]] = False
            leftover = np.nonzero(keep & members & stranded)[0]
            if len(leftover):
                targets = np.argwhere(conductive & (old_labels == isl))
                if len(targets) == 0:
                    keep[leftover] = False
                else:
                    centres = targets + 0.5
                    for i in leftover:
                        d2 = ((centres - self.pos[i]) ** 2).sum(axis=1)
                        j = self.rng.uniform(-SPAWN_JITTER, SPAWN_JITTER, size=2)
                        self.pos[i] = centres[int(np.argmin(d2))] + j
                        self.vel[i] = 0.0
        self.pos, self.vel = self.pos[keep], self.vel[keep]

    def place_charges(self, cells, charge):
        """A fixed test charge at the centre of each cell (world.rs:113-123)."""
        cells = sorted(set(tuple(c) for c in cells))
        pos = np.array([[c + 0.5, r + 0.5] for c, r in cells])
        self.fixed_pos = np.concatenate([self.fixed_pos, pos.reshape(-1, 2)])
        self.fixed_charge = np.concatenate([self.fixed_charge, np.full(len(cells), float(charge))])

    # --- views ----------------------------------------------------------------------

    @property
    def n_electrons(self) -> int:
        return len(self.pos)

    @property
    def electron_charge(self) -> float:  # world.rs:59-61
        return -self.settings["cell_charge"] / self.settings["carriers_per_cell"]

    def emb_settings(self) -> abi.EmbSettings:
        return abi.settings(**self.settings)

    def scene(self) -> abi.SceneHolder:
        return abi.SceneHolder(self)

    def copy(self) -> "World":
        w = World(self.cols, self.rows, dict(self.settings), self.seed)
        for name in ("kind", "resistance", "density", "emf", "orientation", "frequency", "pos", "vel",
                     "fixed_pos", "fixed_charge"):
            setattr(w, name, getattr(self, name).copy())
        w.gains = None if self.gains is None else list(self.gains)
        w.time = self.time
        return w


# --- cell sets (scenes.rs:28-44) -------------------------------------------------------

def outline(c0, c1, r0, r1):
    cells = {(c, r0) for c in range(c0, c1 + 1)} | {(c, r1) for c in range(c0, c1 + 1)}
    cells |= {(c0, r) for r in range(r0, r1 + 1)} | {(c1, r) for r in range(r0, r1 + 1)}
    return sorted(cells)


def block(c0, c1, r0, r1):
    return [(c, r) for c in range(c0, c1) for r in range(r0, r1)]


def line(c0, c1, r0, r1):
    return block(c0, c1 + 1, r0, r1 + 1)


# --- example scenes (scenes.rs:55-222) ------------------------------------------------------

def battery_and_resistor(w):
    w.paint(outline(10, 29, 7, 24), CONDUCTOR, resistance=0.2)
    w.paint(line(29, 29, 12, 19), CONDUCTOR, resistance=5.0)
    w.paint(line(10, 10, 15, 15), BATTERY, emf=6.0, orientation=UP)


def screening(w):
    w.paint(block(15, 25, 11, 21), CONDUCTOR, resistance=1.0)
    w.place_charges([(8, 16)], 5.0)
    w.place_charges([(31, 16)], -5.0)


def series_batteries(w):
    for c0, cells in ((3, 1), (22, 3)):
        w.paint(outline(c0, c0 + 14, 8, 23), CONDUCTOR, resistance=0.5)
        middle = c0 + 7
        w.paint(line(middle - cells // 2, middle + cells // 2, 23, 23), BATTERY, emf=2.0, orientation=RIGHT)


def capacitor(w):
    w.paint(line(16, 16, 8, 23), CONDUCTOR, resistance=0.1)
    w.paint(line(22, 22, 8, 23), CONDUCTOR, resistance=0.1)
    w.paint(line(19, 19, 28, 28), BATTERY, emf=8.0, orientation=RIGHT)
    for cells in (line(8, 18, 28, 28), line(8, 8, 15, 27), line(9, 15, 15, 15),
                  line(20, 30, 28, 28), line(30, 30, 15, 27), line(23, 29, 15, 15)):
        w.paint(cells, CONDUCTOR, resistance=0.1)


def parallel_paths(w):
    for cols, rows in (((4, 7), (9, 30)), ((33, 36), (9, 30)), ((4, 36), (28, 30))):
        w.paint(block(cols[0], cols[1], rows[0], rows[1]), CONDUCTOR, resistance=0.05)
    w.paint(line(19, 19, 28, 29), BATTERY, emf=6.0, orientation=LEFT)
    for r0, r1 in ((10, 11), (16, 17), (21, 24)):
        w.paint(line(7, 32, r0, r1), CONDUCTOR, resistance=0.05)
        w.paint(line(12, 27, r0, r1), CONDUCTOR, resistance=2.0)


def transformer(w):
    w.paint(outline(4, 35, 4, 27), CONDUCTOR, resistance=0.1)
    w.paint(line(4, 4, 15, 15), BATTERY, emf=5.0, orientation=UP, frequency=0.02)
    w.paint(line(35, 35, 16, 16), BATTERY, emf=5.0, orientation=DOWN, frequency=0.02)
    w.paint(outline(12, 27, 10, 21), CONDUCTOR, resistance=0.1)
    w.paint(line(27, 27, 13, 18), CONDUCTOR, resistance=2.0)


def antenna(w):
    w.paint(line(9, 9, 11, 21), CONDUCTOR, resistance=0.1)
    w.paint(line(9, 9, 16, 16), BATTERY, emf=10.0, orientation=UP, frequency=0.1)
    w.paint(line(29, 29, 11, 21), CONDUCTOR, resistance=0.1)
    w.paint(line(29, 29, 16, 16), CONDUCTOR, resistance=2.0)


SCENES = [
    ("battery_and_resistor", battery_and_resistor, {}),
    ("screening", screening, {}),
    ("series_batteries", series_batteries, {}),
    ("capacitor", capacitor, {}),
    ("parallel_paths", parallel_paths, {}),
    ("transformer", transformer, {}),
    ("antenna", antenna, {"speed_of_light": 2.0}),
]


def scene(index: int, seed: int = 0, **settings) -> World:
    """Example scene F(index + 1) on its 40x32 grid (scenes.rs:138-222)."""
    _, build, extra = SCENES[index]
    s = dict(abi.DEFAULTS)
    s.update(extra)
    s.update(settings)
    w = World(40, 32, s, seed)
    build(w)
    return w


# --- benchmark workloads ----------------------------------------------------------------------

def random_wire(cols: int, rows: int, fraction: float = 0.3, seed: int = 1, **settings) -> World:
    """The reference bench's world: each cell a 1-ohm wire with probability 0.3
    (crates/embloks-gpu/examples/bench.rs:16-22), 8 electrons per cell."""
    s = dict(abi.DEFAULTS)
    s.update(settings)
    w = World(cols, rows, s, seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    mask = rng.random((cols, rows)) < fraction
    w.paint([tuple(x) for x in np.argwhere(mask)], CONDUCTOR, resistance=1.0)
    return w


def _place_electrons(w: World, jitter: float, rng) -> None:
    """One electron per carrier at the cell centre plus uniform jitter when a cell
    holds one; the rank-1 lattice (plus the same jitter, scaled) when it holds more."""
    n = w.settings["carriers_per_cell"]
    counts = np.rint(w.density * n).astype(np.int64)
    parts = []
    for count in sorted(set(int(v) for v in counts.reshape(-1) if v > 0)):
        cells = np.argwhere(counts == count).astype(np.float64)
        if count == 1:
            base = cells + 0.5
            spread = jitter
        else:
            base = np.repeat(cells, count, axis=0) + np.tile(fill_pattern(count), (len(cells), 1))
            spread = jitter / count
        parts.append(base + rng.uniform(-spread, spread, size=base.shape))
    w.pos = np.concatenate(parts) if parts else np.zeros((0, 2))
    w.vel = np.zeros_like(w.pos)


def ring_circuit(seed: int = 1, steps_note: str = "") -> World:
    """A 40x32 grid with a rectangular ring conductor one cell wide (inner 20x14),
    1 ohm per cell, one DC battery cell of EMF 1 on its top edge pointing right
    (clockwise on screen), one carrier per conductor cell jittered by up to
    +-0.3 cells."""
    s = dict(abi.DEFAULTS)
    s["carriers_per_cell"] = 1
    w = World(40, 32, s, seed)
    for c, r in outline(9, 30, 8, 23):
        w.kind[c, r] = CONDUCTOR
        w.resistance[c, r] = 1.0
        w.density[c, r] = 1.0
    bc, br = 19, 8
    w.kind[bc, br] = BATTERY
    w.resistance[bc, br] = s["battery_resistance"]
    w.density[bc, br] = float(s["battery_density"])
    w.emf[bc, br] = 1.0
    w.orientation[bc, br] = RIGHT
    _place_electrons(w, 0.3, np.random.Generator(np.random.PCG64(seed)))
    return w


def rectangles_circuit(side: int = 512, n_rects: int = 64, n_batteries: int = 16, min_side: int = 4,
                       max_side: int = 64, carriers_per_cell: int = 1, seed: int = 2, **settings) -> World:
    """Random axis-aligned filled conductor rectangles (sides uniform in
    [min_side, max_side], resistance log-uniform in [0.1, 10]) on a side x side
    grid, with battery cells on rectangle edges (both terminals wired), half DC
    and half AC (frequency uniform 0.02-0.2), EMF uniform in [-2, 2] with its
    sign giving the orientation. One electron per carrier, jittered by up to
    +-0.3 cells."""
    s = dict(abi.DEFAULTS)
    s["carriers_per_cell"] = carriers_per_cell
    s.update(settings)
    w = World(side, side, s, seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    max_side = min(max_side, side - 2)
    min_side = min(min_side, max_side)
    rects = []
    for _ in range(n_rects):
        a, b = rng.integers(min_side, max_side + 1, size=2)
        c0 = int(rng.integers(0, side - a + 1))
        r0 = int(rng.integers(0, side - b + 1))
        R = float(math.exp(rng.uniform(math.log(0.1), math.log(10.0))))
        w.kind[c0:c0 + a, r0:r0 + b] = CONDUCTOR
        w.resistance[c0:c0 + a, r0:r0 + b] = R
        w.density[c0:c0 + a, r0:r0 + b] = 1.0
        rects.append((c0, r0, int(a), int(b)))
    placed = 0
    tries = 0
    while placed < n_batteries and tries < 100 * n_batteries:
        tries += 1
        c0, r0, a, b = rects[int(rng.integers(len(rects)))]
        horizontal = bool(rng.integers(2))
        if horizontal:  # on the top or bottom edge, not at a corner
            if a < 3:
                continue
            c = c0 + int(rng.integers(1, a - 1))
            r = r0 if rng.integers(2) == 0 else r0 + b - 1
            lo, hi = (c - 1, r), (c + 1, r)
        else:
            if b < 3:
                continue
            r = r0 + int(rng.integers(1, b - 1))
            c = c0 if rng.integers(2) == 0 else c0 + a - 1
            lo, hi = (c, r - 1), (c, r + 1)
        if w.kind[c, r] != CONDUCTOR or w.kind[lo] != CONDUCTOR or w.kind[hi] != CONDUCTOR:
            continue
        emf = float(rng.uniform(-2.0, 2.0))
        ac = placed % 2 == 1
        w.kind[c, r] = BATTERY
        w.resistance[c, r] = s["battery_resistance"]
        w.density[c, r] = float(s["battery_density"])
        w.emf[c, r] = abs(emf)
        w.orientation[c, r] = (RIGHT if emf >= 0 else LEFT) if horizontal else (DOWN if emf >= 0 else UP)
        w.frequency[c, r] = float(rng.uniform(0.02, 0.2)) if ac else 0.0
        placed += 1
    _place_electrons(w, 0.3, rng)
    return w


def serving_stream(n_requests: int = 256, mean_gap_s: float = 0.002, min_side: int = 64, max_side: int = 1024,
                   min_steps: int = 1000, max_steps: int = 20000, seed: int = 5, carriers_per_cell: int = 1):
    """The serving workload: requests with exponential inter-arrival times, square
    grids log-uniform in [min_side, max_side] built like rectangles_circuit at the
    conductor density of its 512 x 512 case, each asking for a uniform number of
    steps. Yields (arrival_s, steps, world)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for k in range(n_requests):
        t += float(rng.exponential(mean_gap_s))
        side = int(round(math.exp(rng.uniform(math.log(min_side), math.log(max_side)))))
        steps = int(rng.integers(min_steps, max_steps + 1))
        area = (side / 512.0) ** 2
        world = rectangles_circuit(side, n_rects=max(1, int(round(64 * area))),
                                   n_batteries=max(1, int(round(16 * area))),
                                   max_side=max(8, min(64, side // 4)), carriers_per_cell=carriers_per_cell,
                                   seed=seed * 1000003 + k)
        out.append((t, steps, world))
    return out


def gaussian_blobs(n: int = 2048, count: int = 200, min_sigma: float = 2.0, max_sigma: float = 40.0,
                   seed: int = 3) -> np.ndarray:
    """A charge density on an n x n node grid: `count` Gaussian blobs with sigma
    uniform in [min_sigma, max_sigma] cells, amplitude normal(0, 1) and centres
    uniform over the grid, summed (each blob is separable, so it is built as an
    outer product)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    axis = np.arange(n, dtype=np.float64)
    rho = np.zeros((n, n))
    for _ in range(count):
        sigma = float(rng.uniform(min_sigma, max_sigma))
        amp = float(rng.normal(0.0, 1.0))
        cx, cy = rng.uniform(0.0, n, size=2)
        gx = np.exp(-0.5 * ((axis - cx) / sigma) ** 2)
        gy = np.exp(-0.5 * ((axis - cy) / sigma) ** 2)
        rho += amp * np.outer(gx, gy)
    return rho


def clustered_charges(side: int = 4096, n_electrons: int = 4_000_000, cluster_frac: float = 0.7,
                      n_clusters: int = 1000, min_spread: float = 4.0, max_spread: float = 24.0,
                      cutoff: float = 2.0, seed: int = 4) -> World:
    """Electrons for the close-range force alone: `cluster_frac` of them in
    Gaussian clusters (cluster sizes log-normal, centres uniform, spread
    uniform in [min_spread, max_spread] cells), the rest uniform. Every cell
    that holds electrons is a conductor whose ion density equals its electron
    count (one carrier per cell), so each cell is neutral and the reference's
    per-cell ion squares (close_range.rs:335-350) come into play everywhere."""
    s = dict(abi.DEFAULTS)
    s["carriers_per_cell"] = 1
    s["electron_cutoff"] = cutoff
    w = World(side, side, s, seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    n_cl = int(round(cluster_frac * n_electrons))
    weights = rng.lognormal(0.0, 1.0, n_clusters)
    counts = rng.multinomial(n_cl, weights / weights.sum())
    centres = rng.uniform(0.0, side, size=(n_clusters, 2))
    spreads = rng.uniform(min_spread, max_spread, n_clusters)
    clustered = np.repeat(centres, counts, axis=0) + rng.normal(size=(n_cl, 2)) * np.repeat(spreads, counts)[:, None]
    background = rng.uniform(0.0, side, size=(n_electrons - n_cl, 2))
    pos = np.clip(np.concatenate([clustered, background]), 1e-3, side - 1e-3)
    cell = np.floor(pos).astype(np.int64)
    per_cell = np.bincount(cell[:, 0] * side + cell[:, 1], minlength=side * side).reshape(side, side)
    occupied = per_cell > 0
    w.kind[occupied] = CONDUCTOR
    w.resistance[occupied] = 1.0
    w.density[occupied] = per_cell[occupied].astype(np.float64)
    w.pos = pos
    w.vel = np.zeros_like(pos)
    return w
