"""The `.embloks` save format (crates/embloks-core/src/storage.rs) on the host.

No save written by the reference exists in its tree, so the format is pinned by
the published encodings it is built from (postcard 1.1.3's wire format: LEB128
varints, little-endian f64, one-byte bools and option tags; RFC 1951 raw
deflate, which miniz_oxide writes and zlib reads), by the reference's own
save/load tests (crates/embloks-core/tests/app.rs:428-541) restated on a World,
and by its load errors (storage.rs:138-185). The field order is read from
storage.rs:21-98 and layers.rs:10-22; a save from the reference itself would be
the remaining check (parity of the layout is otherwise unpinned).
"""
import struct
import zlib

import numpy as np
import pytest

from embloks_amd import saves
from embloks_amd import world as W


@pytest.mark.parametrize("value,encoded", [(0, "00"), (1, "01"), (127, "7f"), (128, "8001"), (300, "ac02"),
                                           (16384, "808001"), (2 ** 32 - 1, "ffffffff0f")])
def test_varints_follow_the_postcard_wire_format(value, encoded):
    # postcard's spec: unsigned integers as LEB128, low 7 bits first, high bit = more.
    w = saves.Writer()
    w.varint(value)
    assert bytes(w.out).hex() == encoded
    assert saves.Reader(bytes.fromhex(encoded)).varint() == value


def test_scalars_strings_and_options_follow_the_postcard_wire_format():
    w = saves.Writer()
    w.value("f64", 1.0)
    w.value("bool", True)
    w.value("bool", False)
    w.value("str", "E field")
    assert bytes(w.out) == struct.pack("<d", 1.0) + b"\x01\x00\x07E field"
    r = saves.Reader(bytes(w.out))
    assert (r.value("f64"), r.value("bool"), r.value("bool"), r.value("str")) == (1.0, True, False, "E field")
    with pytest.raises(saves.SaveError):
        saves.Reader(b"\x02").value("bool")
    with pytest.raises(saves.SaveError):
        saves.Writer().varint(2 ** 32, saves.U32_MAX)


def test_a_save_is_the_magic_and_a_raw_deflate_stream_starting_with_the_format():
    data = saves.save(W.scene(0))
    assert data[:8] == b"EMBLOKS\x00"
    body = zlib.decompressobj(-15).decompress(data[8:])
    assert body[:3] == bytes([3, 40, 32])   # format 3, cols 40, rows 32 as varints
    assert struct.unpack_from("<d", body, 3)[0] == 0.0   # sim_time


def assert_same_world(a, b):
    for name in ("kind", "resistance", "density", "emf", "orientation", "frequency", "pos", "vel", "fixed_pos",
                 "fixed_charge"):
        x, y = getattr(a, name), getattr(b, name)
        assert x.dtype == y.dtype and x.shape == y.shape, name
        np.testing.assert_array_equal(x, y, err_msg=name)
    assert a.time == b.time and list(a.gains or []) == list(b.gains or [])
    for k in saves.PHYSICS:
        assert a.settings[k[0]] == b.settings[k[0]], k


def test_round_trip_restores_everything():
    # app.rs:428-458
    world = W.scene(4)
    world.place_charges([(5, 5)], 2.5)
    world.time = 12.375
    world.gains = [1.25, None]
    world.settings["speed_of_light"] = 3.0
    layers = saves.default_layers()
    layers[2]["visible"] = True
    layers[0]["opacity"] = 0.5
    layers[1]["log_scale"] = False
    current = np.random.default_rng(0).normal(size=(40, 32, 2))
    loaded = saves.load(saves.save(world, sim_speed=16.0, layers=layers, current=current))
    assert_same_world(loaded.world, world)
    assert loaded.sim_speed == 16.0 and loaded.layers == layers
    np.testing.assert_array_equal(loaded.current, current)
    assert loaded.world.settings["speed_of_light"] == 3.0


def test_ac_frequency_and_light_speed_survive_save_and_load():
    # app.rs:521-541: the antenna scene (AC source, slow light)
    world = W.scene(6)
    loaded = saves.load(saves.save(world)).world
    assert loaded.settings["speed_of_light"] == 2.0
    np.testing.assert_array_equal(loaded.frequency, world.frequency)
    assert np.any(loaded.frequency > 0)


def test_loading_restores_the_saved_grid_shape():
    # app.rs:460-470
    world = W.World(23, 17)
    world.paint(W.outline(2, 20, 2, 14), W.CONDUCTOR, resistance=1.0)
    loaded = saves.load(saves.save(world)).world
    assert (loaded.cols, loaded.rows) == (23, 17) and loaded.kind.shape == (23, 17)
    assert_same_world(loaded, world)


def body_of(world, **kw):
    return zlib.decompressobj(-15).decompress(saves.save(world, **kw)[8:])


def pack(body):
    c = zlib.compressobj(6, zlib.DEFLATED, -15)
    return saves.MAGIC + c.compress(body) + c.flush()


@pytest.mark.parametrize("data,message", [
    (b"not a save at all", "not an Embloks save"),
    (b"EMBLOKS\x00not a save", "corrupt save (compression)"),   # app.rs:510-518
    (pack(b"\x02"), "unsupported save format 2"),
    (pack(b"\x03\x28"), "corrupt save"),
])
def test_bad_saves_give_the_reference_messages(data, message):
    with pytest.raises(saves.SaveError, match=message.replace("(", r"\(").replace(")", r"\)")):
        saves.load(data)


def test_truncated_and_inconsistent_saves_are_refused():
    world = W.scene(0)
    data = saves.save(world)
    with pytest.raises(saves.SaveError):
        saves.load(data[:len(data) // 2])
    body = bytearray(body_of(world))
    with pytest.raises(saves.SaveError, match="trailing"):
        saves.load(pack(bytes(body) + b"\x00"))
    # cols 40 -> 41 keeps the stream well formed but the arrays no longer fit the grid
    body[1] = 41
    with pytest.raises(saves.SaveError, match="kind has the wrong size"):
        saves.load(pack(bytes(body)))
