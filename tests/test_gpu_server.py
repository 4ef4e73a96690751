"""The continuous-batching serving loop on the GPU.

Each request's result must be the same alone and in the mix: the batched
kernels, the one-scene-per-launch path and a standalone simulation give
bitwise-identical electrons and fields (the kernels are deterministic and one
scene's work never reads another's). Admission by footprint, cancellation and
slot reuse are exercised too.
"""
import numpy as np
import pytest

from embloks_amd import abi
from embloks_amd import world as W
from embloks_amd.gpusim import EmbError, GpuSim, Server

gpu = pytest.mark.gpu


def mix():
    # One settings for every request (the server's), so carriers_per_cell is the default 8.
    scenes = [W.rectangles_circuit(48, n_rects=6, n_batteries=2, max_side=16, carriers_per_cell=8, seed=11),
              W.rectangles_circuit(96, n_rects=12, n_batteries=4, max_side=24, carriers_per_cell=8, seed=12),
              W.scene(0),
              W.scene(5),
              W.rectangles_circuit(160, n_rects=30, n_batteries=6, max_side=32, carriers_per_cell=8, seed=13)]
    steps = [7, 23, 16, 31, 12]
    return scenes, steps


def run_server(batched, use_graph, scenes, steps, max_slots=8, ticks_per_round=5):
    srv = Server(abi.settings(), hbm_budget_bytes=8 << 30, max_slots=max_slots, batched=batched,
                 ticks_per_round=ticks_per_round, use_graph=use_graph, keep_results=True)
    ids = [srv.submit(w, n) for w, n in zip(scenes, steps)]
    while srv.poll():
        pass
    assert all(srv.done(i) for i in ids)
    out = [srv.result(i) for i in ids]
    st = srv.stats()
    srv.destroy()
    return out, st


@gpu
def test_batched_plain_and_alone_agree_bitwise():
    scenes, steps = mix()
    alone = []
    for w, n in zip(scenes, steps):
        sim = GpuSim(w)
        sim.run_ticks(n)
        pos, vel, _ = sim.read_state()
        alone.append((pos, vel) + sim.read_yee())
        sim.destroy()
    batched, st = run_server(True, True, scenes, steps)
    plain, _ = run_server(False, False, scenes, steps)
    eager, _ = run_server(True, False, scenes, steps, ticks_per_round=3)
    for a, b, p, e in zip(alone, batched, plain, eager):
        for x, y, z, u in zip(a, b, p, e):
            np.testing.assert_array_equal(x, y)
            np.testing.assert_array_equal(x, z)
            np.testing.assert_array_equal(x, u)
    assert st["completed"] == len(scenes)
    want_cells = sum((w.cols + 32) * (w.rows + 32) * n for w, n in zip(scenes, steps))
    assert st["cell_steps"] == want_cells


@gpu
def test_admission_refuses_what_cannot_fit_and_reuses_slots():
    scenes, steps = mix()
    small = Server(abi.settings(), hbm_budget_bytes=1 << 20, max_slots=4)
    with pytest.raises(EmbError, match="budget"):
        small.submit(scenes[-1], 5)
    small.destroy()
    # Two slots for five requests: every request still completes, never more than two resident.
    out, st = run_server(True, True, scenes, steps, max_slots=2)
    assert st["completed"] == len(scenes) and st["peak_resident"] <= 2


@gpu
def test_cancel_a_running_request():
    scenes, steps = mix()
    srv = Server(abi.settings(), hbm_budget_bytes=8 << 30, max_slots=8, ticks_per_round=2, keep_results=True)
    ids = [srv.submit(w, 40) for w in scenes[:3]]
    srv.poll()
    srv.cancel(ids[1])
    while srv.poll():
        pass
    assert srv.done(ids[0]) and srv.done(ids[2])
    with pytest.raises(EmbError):
        srv.done(ids[1])
    srv.destroy()


def drain(srv, limit=200000):
    for _ in range(limit):
        if not srv.poll():
            return True
    return False


@gpu
def test_admission_evicts_idle_sine_matrices_to_fit_the_budget():
    # The budget holds the largest scene with its two DST-I sets and little
    # more: every later size needs the idle sets of the earlier ones freed
    # (least recently used first) to be admitted.
    from embloks_amd.gpusim import footprint

    scenes, steps = mix()
    order = [0, 2, 1, 3, 4]   # sizes 48, 40 x 32, 96, 40 x 32, 160
    s = abi.settings()

    def sine(n):
        return (n * n + 2 * ((n + 1) // 2) ** 2 + (n // 2) ** 2) * 4

    need = [footprint(s, w.cols, w.rows, w.n_electrons) + sine(w.cols + 32) + sine(w.rows + 32) for w in scenes]
    srv = Server(s, hbm_budget_bytes=max(need) + (1 << 20), max_slots=4, ticks_per_round=4, keep_results=True)
    ids = [srv.submit(scenes[k], steps[k]) for k in order]
    assert drain(srv), "a request never fit the budget"
    assert all(srv.done(i) for i in ids)
    st = srv.stats()
    assert st["completed"] == len(ids)
    assert st["queue_wait_p99_ms"] >= st["queue_wait_p50_ms"] >= 0.0
    assert st["request_latency_p99_ms"] >= st["request_latency_p50_ms"] > 0.0
    # Each still equals its standalone run.
    for k, i in zip(order, ids):
        sim = GpuSim(scenes[k])
        sim.run_ticks(steps[k])
        pos, vel, _ = sim.read_state()
        got = srv.result(i)
        for x, y in zip((pos, vel) + sim.read_yee(), got):
            np.testing.assert_array_equal(x, y)
        sim.destroy()
    srv.destroy()


@gpu
def test_serving_stream_of_256_requests_each_equal_to_its_standalone_run():
    # The serving workload: 256 requests, sides log-uniform 64-1024, conductor
    # density of the 512^2 circuit, one carrier per cell, Poisson arrivals
    # (all submitted at once here), steps cut to 1% (10-200 per request).
    stream = W.serving_stream(n_requests=256, seed=5, min_steps=10, max_steps=200)
    s = abi.settings(carriers_per_cell=1)
    srv = Server(s, hbm_budget_bytes=64 << 30, max_slots=256, batched=True, ticks_per_round=16, use_graph=True,
                 keep_results=True)
    ids = [srv.submit(w, n) for _, n, w in stream]
    print("submitted 256 requests", flush=True)
    assert drain(srv)
    st = srv.stats()
    print("stream of 256:", {k: st[k] for k in ("completed", "cell_steps_per_s", "step_latency_p50_ms",
                                                 "step_latency_p99_ms", "request_latency_p50_ms",
                                                 "request_latency_p99_ms", "queue_wait_p99_ms", "peak_resident")})
    assert st["completed"] == 256
    assert st["cell_steps"] == sum((w.cols + 32) * (w.rows + 32) * n for _, n, w in stream)
    sides = np.array([w.cols for _, _, w in stream])
    pick = {int(np.argmin(sides)), int(np.argmax(sides))}
    pick |= set(int(k) for k in np.random.default_rng(1).choice(256, 6, replace=False))
    for k in sorted(pick):
        _, n, w = stream[k]
        sim = GpuSim(w)
        sim.run_ticks(n)
        pos, vel, _ = sim.read_state()
        got = srv.result(ids[k])
        for x, y in zip((pos, vel) + sim.read_yee(), got):
            np.testing.assert_array_equal(x, y)
        print("request", k, "side", w.cols, "steps", n, "electrons", w.n_electrons, "equal to its standalone run",
              flush=True)
        sim.destroy()
    srv.destroy()
