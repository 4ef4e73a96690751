"""The close-range tables built on the device (close_range::tables,
crates/embloks-core/src/close_range.rs:67-242) against the FP64 oracle's build
of the same tables.

Both run the reference's sums in its order without fused multiply-adds, so they
differ only where the device's and the host's sin, cos and exp round their last
bit differently: the bar is 1e-12 relative to each table's largest entry. The
oracle's tables are pinned by the reference's close-range table fixture
(tests/test_oracle_fixtures.py::test_close_range_tables_match_python) and by the
compact-force identity (tests/test_oracle_physics.py::
test_grid_plus_close_range_gives_the_compact_force).
"""
import numpy as np
import pytest

from oracle import orc

gpu = pytest.mark.gpu


def gpusim():
    from embloks_amd import gpusim as G
    return G


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@gpu
@pytest.mark.parametrize("softening,passes,cutoff", [(0.25, 2, 3.0), (0.25, 2, 2.0), (0.5, 1, 2.5), (0.2, 3, 4.0),
                                                      (0.3, 0, 1.0)])
def test_device_tables_match_the_oracle(softening, passes, cutoff):
    g, pot, square, half, stencil = gpusim().close_tables(softening, passes, cutoff)
    og, opot, osquare, ohalf, ostencil = orc.tables(softening, passes, cutoff)
    assert half == ohalf and square.shape == osquare.shape and stencil.shape == ostencil.shape
    err = {"g": rel(g, og), "potential": rel(pot, opot), "square": rel(square, osquare),
           "stencil": rel(stencil, ostencil)}
    print(softening, passes, cutoff, err)
    assert max(err.values()) < 1e-12, err
    assert g[-1] == 0.0 and pot[-1] == 0.0 and pot[-2] == 0.0   # zero at and past the cutoff


@gpu
def test_a_timed_build_gives_the_cached_tables_again():
    G = gpusim()
    cached = G.close_tables(0.25, 2, 3.0)
    *fresh, ms = G.close_tables(0.25, 2, 3.0, timed=True)
    print("device table build", round(ms, 2), "ms")
    assert ms > 0.0
    for a, b in zip(cached, fresh):
        assert np.array_equal(np.asarray(a), np.asarray(b))


@gpu
@pytest.mark.parametrize("softening,passes,cutoff", [(0.25, 2, 0.0), (0.25, 2, 40.0), (0.0, 2, 3.0), (0.25, -1, 3.0),
                                                      (0.25, 17, 3.0)])
def test_bad_table_parameters_are_refused(softening, passes, cutoff):
    G = gpusim()
    with pytest.raises(Exception):
        G.close_tables(softening, passes, cutoff)


@gpu
def test_device_tables_match_the_reference_fixture():
    # tests/fixtures/close_range_tables.json, checked at the reference's 1e-9 (fixtures.rs).
    from tests.golden.fixtures import array, load, relative_error
    f = load("close_range_tables")
    g, pot, square, half, stencil = gpusim().close_tables(0.25, 2, 3.0)
    assert relative_error(g, array(f["g"])) < 1e-9
    assert relative_error(pot, array(f["potential"])) < 1e-9
    assert half == f["half"] and list(square.shape) == f["square_shape"]
    assert relative_error(square[::8, ::8, :], array(f["square_sampled"])) < 1e-9
    assert relative_error(stencil, array(f["ion_stencil"])) < 1e-9
