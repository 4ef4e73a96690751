"""ctypes mirror of the POD structs in include/embloks_hip.h.

`Settings` mirrors the physics subset of the reference's `Settings`
(crates/embloks-core/src/settings.rs:164-236); `scene_struct` flattens a
`World` (crates/embloks-core/src/world.rs:24-38) into an `EmbScene` whose
arrays stay alive as long as the returned holder.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np


class EmbSettings(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("cell_charge", C.c_double),
        ("carriers_per_cell", C.c_int32),
        ("battery_density", C.c_int32),
        ("carrier_charge_to_mass", C.c_double),
        ("drag_per_resistance", C.c_double),
        ("battery_resistance", C.c_double),
        ("battery_trim_time", C.c_double),
        ("battery_gain_low", C.c_double),
        ("battery_gain_high", C.c_double),
        ("current_smoothing_time", C.c_double),
        ("wall_restitution", C.c_double),
        ("max_step", C.c_double),
        ("max_substeps", C.c_int32),
        ("smoothing_passes", C.c_int32),
        ("coulomb_k", C.c_double),
        ("speed_of_light", C.c_double),
        ("absorbing_margin", C.c_int32),
        ("reserved0", C.c_int32),
        ("absorbing_strength", C.c_double),
        ("electron_softening", C.c_double),
        ("electron_cutoff", C.c_double),
        ("display_smoothing_time", C.c_double),
    ]


class EmbScene(C.Structure):
    _fields_ = [
        ("cols", C.c_int32),
        ("rows", C.c_int32),
        ("kind", C.POINTER(C.c_uint8)),
        ("resistance", C.POINTER(C.c_double)),
        ("density", C.POINTER(C.c_double)),
        ("emf", C.POINTER(C.c_double)),
        ("orientation", C.POINTER(C.c_uint8)),
        ("frequency", C.POINTER(C.c_double)),
        ("n_electrons", C.c_int64),
        ("pos", C.POINTER(C.c_double)),
        ("vel", C.POINTER(C.c_double)),
        ("n_fixed", C.c_int64),
        ("fixed_pos", C.POINTER(C.c_double)),
        ("fixed_charge", C.POINTER(C.c_double)),
        ("n_gains", C.c_int32),
        ("reserved0", C.c_int32),
        ("gains", C.POINTER(C.c_double)),
        ("time", C.c_double),
    ]


class EmbProbe(C.Structure):
    """The info panel's readings (Probe, crates/embloks-gpu/src/sim.rs:82-101)."""
    _fields_ = [
        ("valid", C.c_int32),
        ("electrons", C.c_int32),
        ("potential", C.c_double),
        ("e", C.c_double * 2),
        ("b", C.c_double),
        ("s", C.c_double * 2),
        ("current", C.c_double * 2),
        ("wire_current", C.c_double),
        ("wire_width", C.c_int32),
        ("stack_state", C.c_int32),
        ("v_low", C.c_double),
        ("v_high", C.c_double),
        ("stack_current", C.c_double),
        ("gain", C.c_double),
        ("has_mean_e", C.c_int32),
        ("pad", C.c_int32),
        ("mean_e_in_wires", C.c_double),
        ("histograms", (C.c_uint32 * 256) * 3),
    ]


class EmbDims(C.Structure):
    _fields_ = [
        ("cols", C.c_int32),
        ("rows", C.c_int32),
        ("margin", C.c_int32),
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("n_stacks", C.c_int32),
        ("n_electrons", C.c_int64),
        ("fdtd_substeps", C.c_int32),
        ("last_substeps", C.c_int32),
        ("time", C.c_double),
    ]


class EmbServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_int64),
        ("max_slots", C.c_int32),
        ("batched", C.c_int32),
        ("ticks_per_round", C.c_int32),
        ("use_graph", C.c_int32),
        ("keep_results", C.c_int32),
        ("reserved0", C.c_int32),
    ]


class EmbServerStats(C.Structure):
    _fields_ = [
        ("completed", C.c_int64),
        ("cell_steps", C.c_int64),
        ("elapsed_s", C.c_double),
        ("cell_steps_per_s", C.c_double),
        ("step_latency_p50_ms", C.c_double),
        ("step_latency_p99_ms", C.c_double),
        ("step_samples", C.c_int64),
        ("peak_resident", C.c_int32),
        ("reserved0", C.c_int32),
        ("request_latency_p50_ms", C.c_double),
        ("request_latency_p99_ms", C.c_double),
        ("queue_wait_p50_ms", C.c_double),
        ("queue_wait_p99_ms", C.c_double),
    ]


# Settings::default, physics subset (settings.rs:164-236).
DEFAULTS = dict(
    dt=1.0 / 30.0,
    cell_charge=4.0,
    carriers_per_cell=8,
    battery_density=4,
    carrier_charge_to_mass=1.0,
    drag_per_resistance=1.0,
    battery_resistance=0.1,
    battery_trim_time=0.25,
    battery_gain_low=0.5,
    battery_gain_high=4.0,
    current_smoothing_time=2.0,
    wall_restitution=0.3,
    max_step=0.4,
    max_substeps=16,
    smoothing_passes=2,
    coulomb_k=1.0,
    speed_of_light=8.0,
    absorbing_margin=16,
    absorbing_strength=3.0,
    electron_softening=0.25,
    electron_cutoff=3.0,
    display_smoothing_time=0.5,
)


def settings(**overrides) -> EmbSettings:
    """An EmbSettings with the reference defaults, then `overrides`."""
    values = dict(DEFAULTS)
    unknown = set(overrides) - set(values)
    if unknown:
        raise KeyError(f"unknown settings: {sorted(unknown)}")
    values.update(overrides)
    s = EmbSettings()
    for k, v in values.items():
        setattr(s, k, v)
    return s


def settings_dict(s: EmbSettings) -> dict:
    return {name: getattr(s, name) for name, _ in EmbSettings._fields_ if name != "reserved0"}


def _ptr(array: np.ndarray | None, ctype):
    if array is None or array.size == 0:
        return C.POINTER(ctype)()
    return array.ctypes.data_as(C.POINTER(ctype))


class SceneHolder:
    """An EmbScene plus the numpy arrays it points into (kept alive here)."""

    def __init__(self, world):
        cols, rows = world.cols, world.rows
        self.kind = np.ascontiguousarray(world.kind, dtype=np.uint8).reshape(-1)
        self.resistance = np.ascontiguousarray(world.resistance, dtype=np.float64).reshape(-1)
        self.density = np.ascontiguousarray(world.density, dtype=np.float64).reshape(-1)
        self.emf = np.ascontiguousarray(world.emf, dtype=np.float64).reshape(-1)
        self.orientation = np.ascontiguousarray(world.orientation, dtype=np.uint8).reshape(-1)
        self.frequency = np.ascontiguousarray(world.frequency, dtype=np.float64).reshape(-1)
        self.pos = np.ascontiguousarray(world.pos, dtype=np.float64).reshape(-1)
        self.vel = np.ascontiguousarray(world.vel, dtype=np.float64).reshape(-1)
        self.fixed_pos = np.ascontiguousarray(world.fixed_pos, dtype=np.float64).reshape(-1)
        self.fixed_charge = np.ascontiguousarray(world.fixed_charge, dtype=np.float64).reshape(-1)
        gains = world.gains if world.gains is not None else []
        self.gains = np.ascontiguousarray([math.nan if g is None else g for g in gains], dtype=np.float64)
        s = EmbScene()
        s.cols, s.rows = cols, rows
        s.kind = _ptr(self.kind, C.c_uint8)
        s.resistance = _ptr(self.resistance, C.c_double)
        s.density = _ptr(self.density, C.c_double)
        s.emf = _ptr(self.emf, C.c_double)
        s.orientation = _ptr(self.orientation, C.c_uint8)
        s.frequency = _ptr(self.frequency, C.c_double)
        s.n_electrons = self.pos.size // 2
        s.pos = _ptr(self.pos, C.c_double)
        s.vel = _ptr(self.vel, C.c_double)
        s.n_fixed = self.fixed_charge.size
        s.fixed_pos = _ptr(self.fixed_pos, C.c_double)
        s.fixed_charge = _ptr(self.fixed_charge, C.c_double)
        s.n_gains = self.gains.size
        s.gains = _ptr(self.gains, C.c_double)
        s.time = float(getattr(world, "time", 0.0))
        self.struct = s

    @property
    def ref(self):
        return C.byref(self.struct)
