"""The DST-I Poisson solve on its own (Maxwell::poisson,
crates/embloks-core/src/maxwell.rs:139-168) in FP32 and BF16x3 MFMA, against
the FP64 oracle (small grids) and an FP64 scipy restatement that the oracle
pins (2048 x 2048, the Gaussian-blob charge density)."""
import numpy as np
import pytest

from bench import fp64_poisson, laplacian_residual
from embloks_amd import world as W
from oracle import orc
from tests.golden.fixtures import relative_error

gpu = pytest.mark.gpu


def solver():
    from embloks_amd.gpusim import Poisson
    return Poisson


@gpu
@pytest.mark.parametrize("nx,ny", [(72, 64), (130, 97), (97, 130), (33, 31), (256, 256), (300, 200)])
def test_small_grids_match_the_oracle(nx, ny):
    rng = np.random.default_rng(nx * 1000 + ny)
    rhs = rng.normal(size=(nx, ny))
    want = orc.Maxwell(nx - 32, ny - 32).poisson(rhs)
    p = solver()(nx, ny)
    p.set_rhs(rhs)
    errs = {}
    for prec, name in ((0, "fp32"), (1, "bf16x3")):
        for tile in (64, 128):
            p.solve(prec, tile)
            errs[(name, tile)] = relative_error(p.read(), want)
    p.destroy()
    print(nx, ny, errs)
    # FP32: a few units of FP32 rounding through four products of length <= 300.
    assert errs[("fp32", 64)] < 1e-5 and errs[("fp32", 128)] < 1e-5, errs
    # BF16x3 keeps ~16 bits of each operand: it must be far better than BF16
    # (~4e-3) yet is measurably worse than FP32.
    assert errs[("bf16x3", 64)] < 2e-4 and errs[("bf16x3", 128)] < 2e-4, errs


@gpu
def test_blobs_2048_fp32_and_bf16x3():
    n = 2048
    rho = W.gaussian_blobs(n)
    want = fp64_poisson(rho)
    p = solver()(n, n)
    p.set_rhs(rho)
    out = {}
    for prec, name in ((0, "fp32"), (1, "bf16x3")):
        p.solve(prec, 0)
        phi = p.read()
        out[name] = (relative_error(phi, want), laplacian_residual(phi, rho))
    p.destroy()
    # The 5-point Laplacian of phi cancels ~|phi| / |rho| ~ 5e3 digits here, so
    # even the exact solution rounded to FP32 leaves a visible residual: report
    # each precision's residual beside that floor.
    floor = laplacian_residual(want.astype(np.float32).astype(np.float64), rho)
    print("2048 blobs (max rel potential error, relative residual):", out, "fp32-rounding floor", floor)
    assert out["fp32"][0] < 1e-4, out
    assert out["bf16x3"][0] < 1e-3, out
    assert out["bf16x3"][0] > out["fp32"][0], out
    assert floor <= out["fp32"][1] < 100 * floor and out["bf16x3"][1] < 1000 * floor, (out, floor)
