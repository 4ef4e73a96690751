"""A scene read from an `.embloks` save (storage::load, crates/embloks-core/src/
storage.rs:142-185) runs on the device exactly as the world it was saved from:
through GpuSim and as a served request, bit for bit."""
import numpy as np
import pytest

from embloks_amd import abi, saves
from embloks_amd import world as W

gpu = pytest.mark.gpu


@gpu
def test_a_loaded_save_runs_like_the_saved_world():
    from embloks_amd.gpusim import GpuSim, Server
    world = W.scene(2)   # series batteries: two stacks with gains
    sim = GpuSim(world)
    sim.run_ticks(40)
    pos, vel, gains = sim.read_state()
    world.pos, world.vel, world.gains, world.time = pos.copy(), vel.copy(), list(gains), sim.dims().time
    sim.destroy()
    loaded = saves.load(saves.save(world)).world

    def run(w):
        s = GpuSim(w)
        s.run_ticks(25)
        out = s.read_state()[:2] + s.read_yee()
        s.destroy()
        return out

    a, b = run(world), run(loaded)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    srv = Server(abi.settings(**loaded.settings), hbm_budget_bytes=1 << 30, max_slots=2, keep_results=True)
    rid = srv.submit(loaded, 25)
    while srv.poll():
        pass
    served = srv.result(rid)
    srv.destroy()
    for x, y in zip(a, served):
        np.testing.assert_array_equal(x, y)
