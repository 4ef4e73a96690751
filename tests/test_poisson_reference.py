"""The FP64 scipy restatement of Maxwell::poisson that bench.py and the GPU
Poisson tests use as their error reference, pinned against the oracle (itself
pinned by maxwell.json) and by the residual of the Laplacian it inverts."""
import numpy as np
import pytest

from bench import fp64_poisson, laplacian_residual
from embloks_amd import world as W
from oracle import orc


@pytest.mark.parametrize("nx,ny", [(72, 64), (130, 97), (18, 16)])
def test_scipy_restatement_equals_the_oracle(nx, ny):
    rhs = np.random.default_rng(nx + ny).normal(size=(nx, ny))
    want = orc.Maxwell(nx - 32, ny - 32).poisson(rhs) if nx > 32 else None
    got = fp64_poisson(rhs)
    if want is not None:
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    assert laplacian_residual(got, rhs) < 1e-12


def test_residual_sees_a_wrong_solution():
    # A dropped normalisation or transposed axis must fail the residual check.
    rhs = np.random.default_rng(1).normal(size=(40, 30))
    phi = fp64_poisson(rhs)
    assert laplacian_residual(2.0 * phi, rhs) > 0.5
    assert laplacian_residual(fp64_poisson(rhs.T).T * 1.0 + 1e-3, rhs) > 1e-4


def test_blobs_are_seeded_and_bounded():
    a, b = W.gaussian_blobs(128, count=20), W.gaussian_blobs(128, count=20)
    assert np.array_equal(a, b) and np.isfinite(a).all() and np.abs(a).max() > 0.1
