"""Loader for the reference's golden fixtures (copied by import_fixtures.py).

The fixtures hold inputs and expected outputs dumped from the Python original of
embloks-core; crates/embloks-core/tests/fixtures.rs reads them the same way: each
array is {"shape": [...], "data": [...]} in row-major order.
"""
import functools
import gzip
import json
import pathlib

import numpy as np

from embloks_amd import abi
from embloks_amd.world import World

DIR = pathlib.Path(__file__).resolve().parent / "fixtures"
SCENES = ["scene_f1", "scene_f2", "scene_f3", "scene_f4", "scene_f6", "scene_f7"]


@functools.lru_cache(maxsize=None)
def load(name: str) -> dict:
    with gzip.open(DIR / f"{name}.json.gz", "rt") as f:
        return json.load(f)


def array(value) -> np.ndarray:
    return np.asarray(value["data"], dtype=np.float64).reshape(value["shape"])


def relative_error(actual, expected) -> float:
    """max |a - e| / max |e| (crates/embloks-core/tests/fixtures.rs:39-46)."""
    a = np.asarray(actual, dtype=np.float64).reshape(-1)
    e = np.asarray(expected, dtype=np.float64).reshape(-1)
    return float(np.max(np.abs(a - e), initial=0.0) / max(float(np.max(np.abs(e), initial=0.0)), 1e-300))


def world_from(f: dict) -> World:
    """The world exactly as the Python original built it (fixtures.rs:102-126)."""
    state = f["initial"]
    s = dict(abi.DEFAULTS)
    s["speed_of_light"] = f["settings"]["speed_of_light"]
    w = World(state["cols"], state["rows"], s, 0)
    w.kind = array(state["kind"]).astype(np.uint8)
    w.resistance = array(state["resistance"])
    w.density = array(state["density"])
    w.emf = array(state["emf"])
    w.orientation = array(state["orientation"]).astype(np.uint8)
    w.frequency = array(state["frequency"])
    w.pos = array(state["pos"]).reshape(-1, 2)
    w.vel = array(state["vel"]).reshape(-1, 2)
    if len(state["fixed_charge"]["data"]):
        w.fixed_pos = array(state["fixed_pos"]).reshape(-1, 2)
        w.fixed_charge = array(state["fixed_charge"])
    return w
