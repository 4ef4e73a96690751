"""emb_sim_diagnostics_static, the entry point that allocates nothing, against
emb_sim_diagnostics through the C ABI: the same Gauss residual and field
energy, bit for bit (same partials, same summation order), on two calls in a
row on one GPU, the second with a smaller scene."""
import pytest

from embloks_amd import world as W

gpu = pytest.mark.gpu


@gpu
def test_static_diagnostics_equal_the_allocating_ones():
    from embloks_amd.gpusim import GpuSim

    big = GpuSim(W.random_wire(256, 192, carriers_per_cell=4, seed=5))
    small = GpuSim(W.scene(5))
    for sim, ticks in ((big, 3), (small, 7)):
        sim.run_ticks(ticks)
        res, _, energy = sim.diagnostics()
        got = sim.diagnostics_static()
        print("nodes", sim.nx * sim.ny, "residual", res, got[0], "energy", energy, got[1])
        assert got == (res, energy)
        assert energy > 0.0 and res < 1e-4
    # The small scene's call left the scratch with its partials; the big
    # scene's next call rewrites every one of them.
    big.run_ticks(1)
    res, _, energy = big.diagnostics()
    assert big.diagnostics_static() == (res, energy)
