"""Python mirror of the reference's GPU backend interface over the C ABI.

`GpuSim` follows crates/embloks-gpu/src/sim.rs (`GpuSim::new`, `upload`,
`run_ticks`, `read_state`, `read_yee`, `read_potential`, `read_current`); the
work runs in libembloks_hip.so (HIP kernels for gfx950). There is no fallback:
if the library or a gfx950 device is missing, construction raises.
"""
from __future__ import annotations

import ctypes as C
import math
import pathlib

import numpy as np

from . import abi

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libembloks_hip.so"

D = C.POINTER(C.c_double)
I64 = C.POINTER(C.c_int64)
VP = C.c_void_p

# Every entry point of include/embloks_hip.h with its ctypes signature.
SIGNATURES = {
    "emb_last_error": (C.c_char_p, []),
    "emb_settings_default": (None, [C.POINTER(abi.EmbSettings)]),
    "emb_device_open": (C.c_int, [C.c_int, C.POINTER(VP)]),
    "emb_device_close": (C.c_int, [VP]),
    "emb_sim_create": (C.c_int, [VP, C.POINTER(abi.EmbSettings), C.POINTER(abi.EmbScene), C.POINTER(VP)]),
    "emb_sim_upload": (C.c_int, [VP, C.POINTER(abi.EmbScene), C.c_int]),
    "emb_sim_step": (C.c_int, [VP, C.c_int]),
    "emb_sim_sync": (C.c_int, [VP]),
    "emb_sim_dims": (C.c_int, [VP, C.POINTER(abi.EmbDims)]),
    "emb_sim_read_state": (C.c_int, [VP, D, D, D]),
    "emb_sim_read_yee": (C.c_int, [VP, D, D, D]),
    "emb_sim_write_yee": (C.c_int, [VP, D, D, D]),
    "emb_sim_read_potential": (C.c_int, [VP, D]),
    "emb_sim_read_current": (C.c_int, [VP, D]),
    "emb_sim_read_display": (C.c_int, [VP, D, D]),
    "emb_sim_read_smoothed_current": (C.c_int, [VP, D]),
    "emb_sim_set_current": (C.c_int, [VP, D]),
    "emb_sim_read_charge_fixed": (C.c_int, [VP, I64]),
    "emb_sim_diagnostics": (C.c_int, [VP, D, D, D]),
    "emb_sim_diagnostics_static": (C.c_int, [VP, D, D]),
    "emb_sim_kernel_times": (C.c_int, [VP, C.c_int, D, C.c_int]),
    "emb_sim_close_force": (C.c_int, [VP, C.c_int32, D, D]),
    "emb_tick_phase_name": (C.c_char_p, [C.c_int]),
    "emb_sim_probe": (C.c_int, [VP, C.c_int32, C.c_int32, C.c_int32, C.POINTER(abi.EmbProbe)]),
    "emb_sim_destroy": (C.c_int, [VP]),
    "emb_server_create": (C.c_int, [VP, C.POINTER(abi.EmbSettings), C.POINTER(abi.EmbServerConfig), C.POINTER(VP)]),
    "emb_server_submit": (C.c_int, [VP, C.POINTER(abi.EmbScene), C.c_int64, C.POINTER(C.c_int64)]),
    "emb_server_cancel": (C.c_int, [VP, C.c_int64]),
    "emb_server_poll": (C.c_int, [VP, C.POINTER(C.c_int32)]),
    "emb_server_done": (C.c_int, [VP, C.c_int64]),
    "emb_server_result": (C.c_int, [VP, C.c_int64, D, D, D, D, D]),
    "emb_server_stats": (C.c_int, [VP, C.POINTER(abi.EmbServerStats)]),
    "emb_server_destroy": (C.c_int, [VP]),
    "emb_scene_footprint": (C.c_int64, [C.POINTER(abi.EmbSettings), C.c_int32, C.c_int32, C.c_int64]),
    "emb_close_tables": (C.c_int, [VP, C.c_double, C.c_int32, C.c_double, D, D, D, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double), D, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "emb_poisson_create": (C.c_int, [VP, C.c_int32, C.c_int32, C.POINTER(VP)]),
    "emb_poisson_set_rhs": (C.c_int, [VP, D]),
    "emb_poisson_solve": (C.c_int, [VP, C.c_int32, C.c_int32, C.c_int32, D]),
    "emb_poisson_read": (C.c_int, [VP, D]),
    "emb_poisson_destroy": (C.c_int, [VP]),
}

_lib = None


class EmbError(RuntimeError):
    pass


def lib():
    """The loaded HIP backend; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise EmbError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        raise EmbError(f"embloks_hip error {rc}: {lib().emb_last_error().decode()}")
    return rc


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(D)


class Device:
    """A gfx950 device (the reference's `Gpu` plus shared `Kernels`, sim.rs:164-178)."""

    _default = None

    def __init__(self, ordinal: int = 0):
        h = VP()
        check(lib().emb_device_open(ordinal, C.byref(h)))
        self.h = h

    @classmethod
    def default(cls) -> "Device":
        if cls._default is None:
            cls._default = Device(0)
        return cls._default

    def close(self):
        if self.h:
            lib().emb_device_close(self.h)
            self.h = None


class GpuSim:
    """The physics of a World on the device (GpuSim, crates/embloks-gpu/src/sim.rs:180-850)."""

    def __init__(self, world, device: Device | None = None):
        self.device = device or Device.default()
        self.settings = world.emb_settings()
        self._scene = world.scene()
        h = VP()
        check(lib().emb_sim_create(self.device.h, C.byref(self.settings), self._scene.ref, C.byref(h)))
        self.h = h
        d = self.dims()
        self.cols, self.rows, self.nx, self.ny = d.cols, d.rows, d.nx, d.ny
        self.n, self.n_stacks, self.margin = d.n_electrons, d.n_stacks, d.margin

    def __del__(self):
        self.destroy()

    def destroy(self):
        if getattr(self, "h", None):
            lib().emb_sim_destroy(self.h)
            self.h = None

    def dims(self) -> abi.EmbDims:
        d = abi.EmbDims()
        check(lib().emb_sim_dims(self.h, C.byref(d)))
        return d

    def upload(self, world, reset_display: bool = True):
        self._scene = world.scene()
        check(lib().emb_sim_upload(self.h, self._scene.ref, int(reset_display)))
        self.n = self.dims().n_electrons

    def step(self, ticks: int):
        """Enqueue ticks (encode_ticks + submit); asynchronous."""
        check(lib().emb_sim_step(self.h, int(ticks)))

    def sync(self):
        check(lib().emb_sim_sync(self.h))

    def run_ticks(self, ticks: int):
        """Run ticks and wait (GpuSim::run_ticks, sim.rs:784-791)."""
        self.step(ticks)
        self.sync()

    @property
    def time(self) -> float:
        return self.dims().time

    def read_state(self):
        """(pos (n, 2), vel (n, 2), gains per stack with None for no entry)."""
        pos, vel = np.zeros((self.n, 2)), np.zeros((self.n, 2))
        gains = np.zeros(max(self.n_stacks, 1))
        check(lib().emb_sim_read_state(self.h, _p(pos), _p(vel), _p(gains)))
        return pos, vel, [None if math.isnan(g) else float(g) for g in gains[: self.n_stacks]]

    def read_yee(self):
        nx, ny = self.nx, self.ny
        ex, ey, bz = np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), np.zeros((nx + 1, ny + 1))
        check(lib().emb_sim_read_yee(self.h, _p(ex), _p(ey), _p(bz)))
        return ex, ey, bz

    def write_yee(self, ex=None, ey=None, bz=None):
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (ex, ey, bz)]
        check(lib().emb_sim_write_yee(self.h, *(D() if a is None else _p(a) for a in keep)))

    def read_potential(self):
        out = np.zeros((self.nx, self.ny))
        check(lib().emb_sim_read_potential(self.h, _p(out)))
        return out

    def read_current(self):
        out = np.zeros((self.cols, self.rows, 2))
        check(lib().emb_sim_read_current(self.h, _p(out)))
        return out

    def read_display(self):
        """GpuSim::read_display (sim.rs:831-840): averaged E (cols, rows, 2), B (cols, rows)."""
        e, b = np.zeros((self.cols, self.rows, 2)), np.zeros((self.cols, self.rows))
        check(lib().emb_sim_read_display(self.h, _p(e), _p(b)))
        return e, b

    def read_smoothed_current(self):
        """The smoothed cell current (cols, rows, 2) that GpuSim::read_state returns."""
        out = np.zeros((self.cols, self.rows, 2))
        check(lib().emb_sim_read_smoothed_current(self.h, _p(out)))
        return out

    def set_current(self, current):
        """GpuSim::set_current (sim.rs:531-539)."""
        current = np.ascontiguousarray(current, dtype=np.float64)
        assert current.shape == (self.cols, self.rows, 2)
        check(lib().emb_sim_set_current(self.h, _p(current)))

    def read_charge_fixed(self):
        out = np.zeros((self.cols, self.rows), dtype=np.int64)
        check(lib().emb_sim_read_charge_fixed(self.h, out.ctypes.data_as(I64)))
        return out

    def read_probe(self, cell=None, arrow_stride: int = 2) -> dict:
        """GpuSim::read_probe (sim.rs:663-715): the info panel's readings at
        visible cell `cell` = (c, r) or None, as Probe's fields."""
        p = abi.EmbProbe()
        c, r = cell if cell is not None else (-1, -1)
        check(lib().emb_sim_probe(self.h, int(c), int(r), int(arrow_stride), C.byref(p)))
        return {
            "cell": tuple(cell) if (cell is not None and p.valid) else None,
            "electrons": p.electrons, "potential": p.potential,
            "e": (p.e[0], p.e[1]), "b": p.b, "s": (p.s[0], p.s[1]), "current": (p.current[0], p.current[1]),
            "wire_current": (p.wire_current, p.wire_width), "stack_state": p.stack_state,
            "v_low": p.v_low, "v_high": p.v_high, "stack_current": p.stack_current, "gain": p.gain,
            "mean_e_in_wires": p.mean_e_in_wires if p.has_mean_e else None,
            "histograms": [np.array(p.histograms[k][:], dtype=np.uint32) for k in range(3)],
        }

    def diagnostics(self):
        """(Gauss residual max |div E - 4 pi k rho|, total electron charge, field energy), FP64."""
        r, q, e = C.c_double(), C.c_double(), C.c_double()
        check(lib().emb_sim_diagnostics(self.h, C.byref(r), C.byref(q), C.byref(e)))
        return r.value, q.value, e.value

    def diagnostics_static(self):
        """(Gauss residual, field energy) as diagnostics() gives them, through the
        entry point that allocates nothing (a per-GPU static scratch array)."""
        r, e = C.c_double(), C.c_double()
        check(lib().emb_sim_diagnostics_static(self.h, C.byref(r), C.byref(e)))
        return r.value, e.value

    def close_force(self, repeats: int = 1):
        """CloseRange::force (close_range.rs:314-376) on every electron at its
        current position: (forces n x 2, mean ms per evaluation)."""
        out = np.zeros((self.n, 2))
        ms = C.c_double()
        check(lib().emb_sim_close_force(self.h, repeats, _p(out), C.byref(ms)))
        return out, ms.value

    def kernel_times(self, ticks: int) -> dict:
        """Run `ticks` more ticks with HIP events between the kernel phases; average ms per phase."""
        n = check(lib().emb_sim_kernel_times(self.h, 1, None, 0))
        out = np.zeros(n)
        check(lib().emb_sim_kernel_times(self.h, int(ticks), _p(out), n))
        return {lib().emb_tick_phase_name(p).decode(): float(out[p]) for p in range(n)}

    def node_fields(self):
        """E (cols, rows, 2) and Bz (cols, rows) at the visible cell centres (maxwell.rs:353-368)."""
        ex, ey, bz = self.read_yee()
        m, c, r = self.margin, self.cols, self.rows
        e = np.zeros((c, r, 2))
        e[:, :, 0] = 0.5 * (ex[m:m + c, m:m + r] + ex[m + 1:m + c + 1, m:m + r])
        e[:, :, 1] = 0.5 * (ey[m:m + c, m:m + r] + ey[m:m + c, m + 1:m + r + 1])
        b = 0.25 * (bz[m:m + c, m:m + r] + bz[m + 1:m + c + 1, m:m + r] + bz[m:m + c, m + 1:m + r + 1]
                    + bz[m + 1:m + c + 1, m + 1:m + r + 1])
        return e, b


class Server:
    """The continuous-batching serving loop (no counterpart in the reference)."""

    def __init__(self, settings: abi.EmbSettings, hbm_budget_bytes: int, max_slots: int = 256, batched: bool = True,
                 ticks_per_round: int = 16, use_graph: bool = True, keep_results: bool = False,
                 device: Device | None = None):
        self.device = device or Device.default()
        self.settings = settings
        cfg = abi.EmbServerConfig()
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.max_slots = int(max_slots)
        cfg.batched = int(batched)
        cfg.ticks_per_round = int(ticks_per_round)
        cfg.use_graph = int(use_graph)
        cfg.keep_results = int(keep_results)
        self.cfg = cfg
        h = VP()
        check(lib().emb_server_create(self.device.h, C.byref(settings), C.byref(cfg), C.byref(h)))
        self.h = h
        self.shapes = {}

    def __del__(self):
        self.destroy()

    def destroy(self):
        if getattr(self, "h", None):
            lib().emb_server_destroy(self.h)
            self.h = None

    def submit(self, world, steps: int) -> int:
        holder = world.scene()
        rid = C.c_int64()
        check(lib().emb_server_submit(self.h, holder.ref, int(steps), C.byref(rid)))
        m = world.settings["absorbing_margin"]
        self.shapes[rid.value] = (world.cols + 2 * m, world.rows + 2 * m, world.n_electrons)
        return rid.value

    def cancel(self, rid: int):
        check(lib().emb_server_cancel(self.h, rid))

    def poll(self) -> int:
        active = C.c_int32()
        check(lib().emb_server_poll(self.h, C.byref(active)))
        return active.value

    def done(self, rid: int) -> bool:
        return check(lib().emb_server_done(self.h, rid)) == 1

    def result(self, rid: int):
        nx, ny, n = self.shapes[rid]
        pos, vel = np.zeros((n, 2)), np.zeros((n, 2))
        ex, ey, bz = np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), np.zeros((nx + 1, ny + 1))
        check(lib().emb_server_result(self.h, rid, _p(pos), _p(vel), _p(ex), _p(ey), _p(bz)))
        return pos, vel, ex, ey, bz

    def stats(self) -> dict:
        st = abi.EmbServerStats()
        check(lib().emb_server_stats(self.h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in abi.EmbServerStats._fields_ if name != "reserved0"}


def footprint(settings: abi.EmbSettings, cols: int, rows: int, n_electrons: int) -> int:
    return int(lib().emb_scene_footprint(C.byref(settings), cols, rows, n_electrons))


class Poisson:
    """The DST-I Poisson solve on its own (Maxwell::poisson,
    crates/embloks-core/src/maxwell.rs:139-168) on an nx x ny node grid."""

    FP32, BF16X3 = 0, 1

    def __init__(self, nx: int, ny: int, device: "Device | None" = None):
        self.device = device or Device.default()
        self.nx, self.ny = nx, ny
        h = VP()
        check(lib().emb_poisson_create(self.device.h, nx, ny, C.byref(h)))
        self.h = h

    def set_rhs(self, rhs: np.ndarray) -> None:
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        assert rhs.shape == (self.nx, self.ny)
        check(lib().emb_poisson_set_rhs(self.h, _p(rhs)))

    def solve(self, precision: int = 0, tile: int = 0, repeats: int = 1) -> float:
        """Run `repeats` solves; returns the mean milliseconds per solve."""
        ms = C.c_double()
        check(lib().emb_poisson_solve(self.h, precision, tile, repeats, C.byref(ms)))
        return ms.value

    def read(self) -> np.ndarray:
        out = np.zeros((self.nx, self.ny))
        check(lib().emb_poisson_read(self.h, _p(out)))
        return out

    def destroy(self) -> None:
        if getattr(self, "h", None):
            check(lib().emb_poisson_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass



def close_tables(softening: float, passes: int, cutoff: float, timed: bool = False, device: "Device | None" = None):
    """close_range::tables (crates/embloks-core/src/close_range.rs:120-138) built on
    the device: (g, potential, square (n, n, 3), half, ion_stencil), plus the
    build's milliseconds when `timed` (a fresh build)."""
    dev = (device or Device.default()).h
    n, reach, half = C.c_int32(), C.c_int32(), C.c_double()
    ms = C.c_double()
    check(lib().emb_close_tables(dev, softening, passes, cutoff, D(), D(), D(), C.byref(n), C.byref(half), D(),
                                 C.byref(reach), C.byref(ms) if timed else C.POINTER(C.c_double)()))
    g, pot = np.zeros(4098), np.zeros(4098)   # PAIR_SAMPLES + 1 samples and the zero past the cutoff
    square = np.zeros((n.value, n.value, 3))
    size = 2 * reach.value + 1
    stencil = np.zeros((size, size))
    check(lib().emb_close_tables(dev, softening, passes, cutoff, _p(g), _p(pot), _p(square), C.byref(n),
                                 C.byref(half), _p(stencil), C.byref(reach), C.POINTER(C.c_double)()))
    out = (g, pot, square, half.value, stencil)
    return out + (ms.value,) if timed else out
