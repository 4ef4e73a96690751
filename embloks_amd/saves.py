"""The `.embloks` save format (crates/embloks-core/src/storage.rs:1-186), so that
served requests and simulations can start from real user scenes.

A save is the 8-byte magic `EMBLOKS\\0` followed by a raw deflate stream
(miniz_oxide 0.8.9/0.9.1 `compress_to_vec` at level 6, RFC 1951, no zlib header)
of the postcard 1.1.3 encoding of `SaveFile` (storage.rs:75-98). Postcard writes
a struct as its fields in order with no names or tags: unsigned integers
(`u32`, `usize`) as LEB128 varints, `u8` and `bool` as one byte, `f64` as 8
little-endian bytes, a `Vec` or `String` as a varint length and its elements,
an `Option` as a 0/1 byte and the value, fixed arrays (`[f64; 2]`,
`[Layer; 4]`) and newtype structs (`Layers`) as their elements alone.

`load` returns a `Saved` with a `World` (materials, electrons, test charges,
gains, time and the saved physics settings over this build's defaults), the
speed, the four display layers and the smoothed current; `save` writes the
same. The deflate stream is Python's zlib (the same RFC 1951 format; its bytes
may differ from miniz_oxide's, both inflate to the same postcard body). Errors
are the reference's load errors (storage.rs:138-185), raised as `SaveError`.
"""
from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .world import World

MAGIC = b"EMBLOKS\x00"     # storage.rs:17
FORMAT_VERSION = 3         # storage.rs:16

# PhysicsSettings (storage.rs:21-42), in field order.
PHYSICS = (("dt", "f64"), ("cell_charge", "f64"), ("carriers_per_cell", "u32"), ("carrier_charge_to_mass", "f64"),
           ("drag_per_resistance", "f64"), ("battery_resistance", "f64"), ("battery_density", "u32"),
           ("current_smoothing_time", "f64"), ("wall_restitution", "f64"), ("max_step", "f64"),
           ("max_substeps", "u32"), ("coulomb_k", "f64"), ("speed_of_light", "f64"), ("electron_softening", "f64"),
           ("electron_cutoff", "f64"), ("battery_trim_time", "f64"), ("smoothing_passes", "usize"),
           ("absorbing_margin", "usize"), ("absorbing_strength", "f64"))
# Layer (layers.rs:10-22), in field order.
LAYER = (("label", "str"), ("visible", "bool"), ("gain", "f64"), ("opacity", "f64"), ("log_scale", "bool"),
         ("auto", "bool"), ("adjustable", "bool"))
U32_MAX = (1 << 32) - 1
U64_MAX = (1 << 64) - 1


class SaveError(ValueError):
    pass


def default_layers() -> list:
    """Layers::default (layers.rs:150-166)."""
    def layer(label, **kw):
        d = {"label": label, "visible": True, "gain": 1.0, "opacity": 1.0, "log_scale": True, "auto": False,
             "adjustable": True}
        d.update(kw)
        return d
    return [layer("E field"), layer("B field", auto=True, log_scale=False),
            layer("S energy flow", visible=False, auto=True), layer("Particles", adjustable=False)]


@dataclass
class Saved:
    world: World
    sim_speed: float = 1.0
    layers: list = field(default_factory=default_layers)
    current: np.ndarray | None = None   # (cols, rows, 2) smoothed current


# --- postcard ---------------------------------------------------------------------------

class Writer:
    def __init__(self):
        self.out = bytearray()

    def varint(self, v: int, limit: int = U64_MAX):
        v = int(v)
        if v < 0 or v > limit:
            raise SaveError(f"{v} does not fit the field")
        while True:
            b = v & 0x7F
            v >>= 7
            if v:
                self.out.append(b | 0x80)
            else:
                self.out.append(b)
                return

    def f64(self, v):
        self.out += struct.pack("<d", float(v))

    def byte(self, v):
        self.out.append(int(v) & 0xFF)

    def value(self, kind: str, v):
        if kind == "f64":
            self.f64(v)
        elif kind == "u32":
            self.varint(v, U32_MAX)
        elif kind == "usize":
            self.varint(v)
        elif kind == "bool":
            self.byte(1 if v else 0)
        elif kind == "str":
            data = str(v).encode("utf-8")
            self.varint(len(data))
            self.out += data
        else:
            raise AssertionError(kind)

    def f64_vec(self, a):
        a = np.ascontiguousarray(a, dtype="<f8").ravel()
        self.varint(a.size)
        self.out += a.tobytes()

    def pair_vec(self, a):   # Vec<[f64; 2]>
        a = np.ascontiguousarray(a, dtype="<f8").reshape(-1, 2)
        self.varint(len(a))
        self.out += a.tobytes()

    def u8_vec(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint8).ravel()
        self.varint(a.size)
        self.out += a.tobytes()


class Reader:
    def __init__(self, data: bytes):
        self.data = memoryview(data)
        self.at = 0

    def need(self, n: int):
        if n < 0 or self.at + n > len(self.data):
            raise SaveError("corrupt save (unexpected end of the data)")

    def varint(self, limit: int = U64_MAX) -> int:
        v, shift = 0, 0
        while True:
            self.need(1)
            b = self.data[self.at]
            self.at += 1
            v |= (b & 0x7F) << shift
            if not b & 0x80:
                break
            shift += 7
            if shift > 63:
                raise SaveError("corrupt save (varint too long)")
        if v > limit:
            raise SaveError("corrupt save (integer out of range)")
        return v

    def f64(self) -> float:
        self.need(8)
        v = struct.unpack_from("<d", self.data, self.at)[0]
        self.at += 8
        return v

    def byte(self) -> int:
        self.need(1)
        v = self.data[self.at]
        self.at += 1
        return v

    def value(self, kind: str):
        if kind == "f64":
            return self.f64()
        if kind == "u32":
            return self.varint(U32_MAX)
        if kind == "usize":
            return self.varint()
        if kind == "bool":
            b = self.byte()
            if b > 1:
                raise SaveError("corrupt save (bad bool)")
            return bool(b)
        if kind == "str":
            n = self.varint()
            self.need(n)
            s = bytes(self.data[self.at:self.at + n])
            self.at += n
            try:
                return s.decode("utf-8")
            except UnicodeDecodeError:
                raise SaveError("corrupt save (bad string)") from None
        raise AssertionError(kind)

    def array(self, count: int, width: int, dtype) -> np.ndarray:
        n = count * width
        self.need(n * np.dtype(dtype).itemsize)
        a = np.frombuffer(self.data, dtype=dtype, count=n, offset=self.at).copy()
        self.at += n * np.dtype(dtype).itemsize
        return a

    def f64_vec(self) -> np.ndarray:
        return self.array(self.varint(), 1, "<f8").astype(np.float64)

    def pair_vec(self) -> np.ndarray:
        n = self.varint()
        return self.array(n, 2, "<f8").astype(np.float64).reshape(n, 2)

    def u8_vec(self) -> np.ndarray:
        return self.array(self.varint(), 1, np.uint8)


# --- SaveFile ---------------------------------------------------------------------------

def save(world: World, sim_speed: float = 1.0, layers: list | None = None, current: np.ndarray | None = None,
         level: int = 6) -> bytes:
    """storage::save (storage.rs:106-135) of a world, its view and its smoothed current."""
    w = Writer()
    cols, rows = world.cols, world.rows
    w.varint(FORMAT_VERSION, U32_MAX)
    w.varint(cols)
    w.varint(rows)
    w.f64(world.time)
    w.f64(sim_speed)
    for name, kind in PHYSICS:
        w.value(kind, world.settings[name])
    gains = list(world.gains or [])
    w.varint(len(gains))
    for g in gains:
        if g is None:
            w.byte(0)
        else:
            w.byte(1)
            w.f64(g)
    layers = default_layers() if layers is None else layers
    if len(layers) != 4:
        raise SaveError("there are four layers")
    for layer in layers:
        for name, kind in LAYER:
            w.value(kind, layer[name])
    shape = (cols, rows)
    for name in ("kind", "resistance", "density", "emf", "orientation", "frequency"):
        a = np.asarray(getattr(world, name)).reshape(shape)
        (w.u8_vec if name in ("kind", "orientation") else w.f64_vec)(a)
    n = len(world.pos)
    charge = world.electron_charge
    w.pair_vec(world.pos)
    w.pair_vec(world.vel)
    w.f64_vec(np.full(n, charge))
    w.f64_vec(np.full(n, -charge / world.settings["carrier_charge_to_mass"]))   # World::electron_mass
    w.pair_vec(world.fixed_pos)
    w.f64_vec(world.fixed_charge)
    cur = np.zeros((cols, rows, 2)) if current is None else np.asarray(current, dtype=np.float64)
    if cur.shape != (cols, rows, 2):
        raise SaveError("current has the wrong size")
    w.f64_vec(cur)
    comp = zlib.compressobj(level, zlib.DEFLATED, -15)
    return MAGIC + comp.compress(bytes(w.out)) + comp.flush()


def load(data: bytes, seed: int = 0) -> Saved:
    """storage::load (storage.rs:142-185): a World with the saved physics settings
    over this build's defaults, and the saved view and smoothed current."""
    if not bytes(data[:len(MAGIC)]) == MAGIC:
        raise SaveError("not an Embloks save")
    try:
        d = zlib.decompressobj(-15)
        raw = d.decompress(bytes(data[len(MAGIC):]))
        if not d.eof:
            raise zlib.error("truncated")
    except zlib.error:
        raise SaveError("corrupt save (compression)") from None
    r = Reader(raw)
    try:
        fmt = r.varint(U32_MAX)
    except SaveError:
        raise SaveError("corrupt save") from None
    if fmt != FORMAT_VERSION:
        raise SaveError(f"unsupported save format {fmt}")
    cols, rows = r.varint(), r.varint()
    sim_time, sim_speed = r.f64(), r.f64()
    settings = dict(abi.DEFAULTS)
    for name, kind in PHYSICS:
        settings[name] = r.value(kind)
    gains = []
    for _ in range(r.varint()):
        tag = r.byte()
        if tag > 1:
            raise SaveError("corrupt save (bad option)")
        gains.append(r.f64() if tag else None)
    layers = [{name: r.value(kind) for name, kind in LAYER} for _ in range(4)]
    arrays = {}
    for name in ("kind", "resistance", "density", "emf", "orientation", "frequency"):
        arrays[name] = r.u8_vec() if name in ("kind", "orientation") else r.f64_vec()
    pos, vel, charge, mass = r.pair_vec(), r.pair_vec(), r.f64_vec(), r.f64_vec()
    fixed_pos, fixed_charge = r.pair_vec(), r.f64_vec()
    current = r.f64_vec()
    if r.at != len(raw):
        raise SaveError("corrupt save (trailing bytes)")
    if cols <= 0 or rows <= 0:
        raise SaveError("corrupt save (empty grid)")

    world = World(cols, rows, settings, seed)
    for name, a in arrays.items():
        if a.size != cols * rows:
            raise SaveError(f"{name} has the wrong size")
        setattr(world, name, a.reshape(cols, rows).astype(getattr(world, name).dtype))
    n = len(pos)
    if any(len(a) != n for a in (vel, charge, mass)):
        raise SaveError("electron arrays differ in length")
    if len(fixed_charge) != len(fixed_pos):
        raise SaveError("test charge arrays differ in length")
    # The device stores one electron charge and charge-to-mass ratio per scene
    # (World::electron_charge / electron_mass, world.rs:59-66).
    q = world.electron_charge
    if n and (np.any(np.abs(charge - q) > 1e-12 * abs(q)) or np.any(np.abs(charge / mass + settings["carrier_charge_to_mass"]) > 1e-12 * settings["carrier_charge_to_mass"])):
        raise SaveError("electrons of another charge or mass than the settings give are not supported")
    world.pos, world.vel = pos, vel
    world.fixed_pos, world.fixed_charge = fixed_pos, fixed_charge
    world.gains = gains
    world.time = sim_time
    if current.size != cols * rows * 2:
        raise SaveError("current has the wrong size")
    return Saved(world, sim_speed, layers, current.reshape(cols, rows, 2))
