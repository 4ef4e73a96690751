"""ctypes wrapper of the FP64 oracle (liborc.so, built from embloks_oracle.cpp).

TEST INFRASTRUCTURE ONLY: the checker for the HIP backend. Only the tests, the
smoke check and the benchmark's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import pathlib
import subprocess

import numpy as np

from embloks_amd import abi

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"

_lib = None

D = C.POINTER(C.c_double)
I64 = C.POINTER(C.c_int64)
I32 = C.POINTER(C.c_int)


def build(force: bool = False) -> pathlib.Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-B" if force else "-s", "-C", str(HERE)], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        sig = {
            "orc_set_threads": (None, [C.c_int]),
            "orc_settings_default": (None, [C.POINTER(abi.EmbSettings)]),
            "orc_mx_new": (vp, [C.POINTER(abi.EmbSettings), C.c_int, C.c_int]),
            "orc_mx_free": (None, [vp]),
            "orc_mx_dims": (None, [vp, I32, I32]),
            "orc_mx_set": (None, [vp, D, D, D]),
            "orc_mx_get": (None, [vp, D, D, D, D]),
            "orc_mx_enforce_gauss": (None, [vp, D]),
            "orc_mx_step": (None, [vp, D, D, C.c_double]),
            "orc_mx_poisson": (None, [vp, D, D]),
            "orc_mx_dst2": (None, [vp, D]),
            "orc_mx_smoothed": (None, [vp, D, C.c_int, D]),
            "orc_mx_gather": (None, [vp, C.c_int64, D, D]),
            "orc_mx_node_fields": (None, [vp, D, D]),
            "orc_mx_divergence": (None, [vp, D]),
            "orc_mx_damping": (None, [vp, C.c_double, D, D, D]),
            "orc_cic_deposit": (None, [C.c_int64, D, D, C.c_int, C.c_int, D]),
            "orc_cic_deposit_vectors": (None, [C.c_int64, D, D, C.c_int, C.c_int, D]),
            "orc_spread_over_cells": (None, [C.c_int, C.c_int, D, D]),
            "orc_grid_force": (C.c_int, [C.c_int, C.c_double, D, D, C.c_int]),
            "orc_tables": (None, [C.c_double, C.c_int, C.c_double, D, D, D, I32, D, D, I32]),
            "orc_compact_g": (C.c_double, [C.c_double, C.c_double]),
            "orc_sim_new": (vp, [C.POINTER(abi.EmbSettings), C.POINTER(abi.EmbScene)]),
            "orc_sim_free": (None, [vp]),
            "orc_sim_dims": (None, [vp, C.POINTER(abi.EmbDims)]),
            "orc_sim_step": (C.c_int, [vp, C.c_int]),
            "orc_sim_read_state": (None, [vp, D, D, D]),
            "orc_sim_read_yee": (None, [vp, D, D, D]),
            "orc_sim_read_potential": (None, [vp, D]),
            "orc_sim_write_state": (None, [vp, D, D, D, D, D]),
            "orc_sim_current": (None, [vp, D]),
            "orc_sim_node_fields": (None, [vp, D, D]),
            "orc_sim_potential": (None, [vp, D]),
            "orc_sim_close_force": (None, [vp, D]),
            "orc_close_force_subset": (None, [C.POINTER(abi.EmbSettings), C.POINTER(abi.EmbScene), C.c_int64, I64, D]),
            "orc_sim_stacks": (C.c_int, [vp, I32, I32]),
            "orc_sim_labels": (None, [vp, I64]),
            "orc_sim_confine": (None, [vp, D, D, D, C.c_double, D, D]),
            "orc_kick": (None, [D, D, C.c_double, C.c_double, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray | None):
    if a is None:
        return D()
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(D)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def settings(**overrides) -> abi.EmbSettings:
    return abi.settings(**overrides)


class Maxwell:
    """Maxwell (crates/embloks-core/src/maxwell.rs) on a cols x rows grid."""

    def __init__(self, cols: int, rows: int, s: abi.EmbSettings | None = None):
        self.s = s if s is not None else settings()
        self.h = lib().orc_mx_new(C.byref(self.s), cols, rows)
        nx, ny = C.c_int(), C.c_int()
        lib().orc_mx_dims(self.h, C.byref(nx), C.byref(ny))
        self.cols, self.rows, self.nx, self.ny = cols, rows, nx.value, ny.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mx_free(self.h)
            self.h = None

    def set(self, ex=None, ey=None, bz=None):
        keep = [None if a is None else f64(a) for a in (ex, ey, bz)]
        lib().orc_mx_set(self.h, *(ptr(a) for a in keep))

    def get(self):
        nx, ny = self.nx, self.ny
        ex, ey, bz, phi = np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), np.zeros((nx + 1, ny + 1)), np.zeros((nx, ny))
        lib().orc_mx_get(self.h, ptr(ex), ptr(ey), ptr(bz), ptr(phi))
        return ex, ey, bz, phi

    def enforce_gauss(self, rho):
        lib().orc_mx_enforce_gauss(self.h, ptr(f64(rho)))

    def step(self, rho, current, dt):
        lib().orc_mx_step(self.h, ptr(f64(rho)), ptr(f64(current)), float(dt))

    def poisson(self, rhs):
        out = np.zeros((self.nx, self.ny))
        lib().orc_mx_poisson(self.h, ptr(f64(rhs)), ptr(out))
        return out

    def dst2(self, a):
        a = f64(a).copy()
        lib().orc_mx_dst2(self.h, ptr(a))
        return a

    def smoothed(self, values):
        values = f64(values)
        d = values.shape[2] if values.ndim == 3 else 1
        out = np.zeros((self.nx, self.ny, d))
        lib().orc_mx_smoothed(self.h, ptr(values), d, ptr(out))
        return out if values.ndim == 3 else out[:, :, 0]

    def gather(self, pos):
        pos = f64(pos).reshape(-1, 2)
        out = np.zeros((len(pos), 3))
        lib().orc_mx_gather(self.h, len(pos), ptr(pos), ptr(out))
        return out

    def node_fields(self):
        e, b = np.zeros((self.cols, self.rows, 2)), np.zeros((self.cols, self.rows))
        lib().orc_mx_node_fields(self.h, ptr(e), ptr(b))
        return e, b

    def divergence(self):
        out = np.zeros((self.nx, self.ny))
        lib().orc_mx_divergence(self.h, ptr(out))
        return out

    def damping(self, h):
        nx, ny = self.nx, self.ny
        ex, ey, bz = np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), np.zeros((nx + 1, ny + 1))
        lib().orc_mx_damping(self.h, float(h), ptr(ex), ptr(ey), ptr(bz))
        return ex, ey, bz


def cic_deposit(pos, q, cols, rows):
    pos, q = f64(pos).reshape(-1, 2), f64(q)
    out = np.zeros((cols, rows))
    lib().orc_cic_deposit(len(q), ptr(pos), ptr(q), cols, rows, ptr(out))
    return out


def cic_deposit_vectors(pos, v, cols, rows):
    pos, v = f64(pos).reshape(-1, 2), f64(v).reshape(-1, 2)
    out = np.zeros((cols, rows, 2))
    lib().orc_cic_deposit_vectors(len(pos), ptr(pos), ptr(v), cols, rows, ptr(out))
    return out


def spread_over_cells(values):
    values = f64(values)
    out = np.zeros_like(values)
    lib().orc_spread_over_cells(values.shape[0], values.shape[1], ptr(values), ptr(out))
    return out


def grid_force(passes=2, reach=3.0):
    cap = 100000
    r, f = np.zeros(cap), np.zeros(cap)
    n = lib().orc_grid_force(passes, float(reach), ptr(r), ptr(f), cap)
    return r[:n].copy(), f[:n].copy()


def tables(softening=0.25, passes=2, cutoff=3.0):
    """(g, potential, square (n, n, 3), half, ion_stencil) of close_range::tables."""
    n, reach, half = C.c_int(), C.c_int(), C.c_double()
    lib().orc_tables(softening, passes, cutoff, D(), D(), D(), C.byref(n), C.byref(half), D(), C.byref(reach))
    g, pot = np.zeros(4098), np.zeros(4098)
    square = np.zeros((n.value, n.value, 3))
    size = 2 * reach.value + 1
    stencil = np.zeros((size, size))
    lib().orc_tables(softening, passes, cutoff, ptr(g), ptr(pot), ptr(square), C.byref(n), C.byref(half),
                     ptr(stencil), C.byref(reach))
    return g, pot, square, half.value, stencil


def kick(vel, acc, damping, dt):
    out = np.zeros(2)
    lib().orc_kick(ptr(f64(vel)), ptr(f64(acc)), float(damping), float(dt), ptr(out))
    return out


def compact_g(r2, softening=0.25):
    return lib().orc_compact_g(float(r2), float(softening))


def close_force_subset(world, idx) -> np.ndarray:
    """CloseRange::force (close_range.rs:314-376) on the electrons `idx` of
    `world` only, FP64, without building the field solver."""
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    s = world.emb_settings()
    holder = world.scene()
    out = np.zeros((len(idx), 2))
    lib().orc_close_force_subset(C.byref(s), holder.ref, len(idx), idx.ctypes.data_as(I64), ptr(out))
    return out


class Sim:
    """Simulation (crates/embloks-core/src/simulation.rs) of a World."""

    def __init__(self, world):
        self.world = world
        self.s = world.emb_settings()
        holder = world.scene()
        self.h = lib().orc_sim_new(C.byref(self.s), holder.ref)
        d = abi.EmbDims()
        lib().orc_sim_dims(self.h, C.byref(d))
        self.dims = d
        self.cols, self.rows, self.nx, self.ny = d.cols, d.rows, d.nx, d.ny
        self.n = d.n_electrons
        self.n_stacks = d.n_stacks

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_free(self.h)
            self.h = None

    def step(self, ticks: int = 1) -> int:
        return lib().orc_sim_step(self.h, int(ticks))

    @property
    def time(self) -> float:
        d = abi.EmbDims()
        lib().orc_sim_dims(self.h, C.byref(d))
        return d.time

    def state(self):
        pos, vel, gains = np.zeros((self.n, 2)), np.zeros((self.n, 2)), np.zeros(max(self.n_stacks, 1))
        lib().orc_sim_read_state(self.h, ptr(pos), ptr(vel), ptr(gains))
        return pos, vel, [None if math.isnan(g) else g for g in gains[: self.n_stacks]]

    def yee(self):
        nx, ny = self.nx, self.ny
        ex, ey, bz = np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1)), np.zeros((nx + 1, ny + 1))
        lib().orc_sim_read_yee(self.h, ptr(ex), ptr(ey), ptr(bz))
        return ex, ey, bz

    def potential_nodes(self):
        out = np.zeros((self.nx, self.ny))
        lib().orc_sim_read_potential(self.h, ptr(out))
        return out

    def write_state(self, pos=None, vel=None, ex=None, ey=None, bz=None):
        keep = [None if a is None else f64(a) for a in (pos, vel, ex, ey, bz)]
        lib().orc_sim_write_state(self.h, *(ptr(a) for a in keep))

    def current(self):
        out = np.zeros((self.cols, self.rows, 2))
        lib().orc_sim_current(self.h, ptr(out))
        return out

    def node_fields(self):
        e, b = np.zeros((self.cols, self.rows, 2)), np.zeros((self.cols, self.rows))
        lib().orc_sim_node_fields(self.h, ptr(e), ptr(b))
        return e, b

    def potential(self):
        out = np.zeros((self.cols, self.rows))
        lib().orc_sim_potential(self.h, ptr(out))
        return out

    def close_force(self):
        out = np.zeros((self.n, 2))
        lib().orc_sim_close_force(self.h, ptr(out))
        return out

    def stacks(self):
        lh = np.zeros(4 * max(self.n_stacks, 1), dtype=np.int32)
        nc = np.zeros(max(self.n_stacks, 1), dtype=np.int32)
        lib().orc_sim_stacks(self.h, lh.ctypes.data_as(I32), nc.ctypes.data_as(I32))
        return lh[: 4 * self.n_stacks].reshape(-1, 4), nc[: self.n_stacks]

    def confine(self, old, new, vel, restitution):
        old, new, vel = f64(old), f64(new), f64(vel)
        p, v = np.zeros(2), np.zeros(2)
        lib().orc_sim_confine(self.h, ptr(old), ptr(new), ptr(vel), float(restitution), ptr(p), ptr(v))
        return p, v

    def labels(self):
        out = np.zeros((self.cols, self.rows), dtype=np.int64)
        lib().orc_sim_labels(self.h, out.ctypes.data_as(I64))
        return out
