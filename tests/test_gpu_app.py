"""The reference's whole-app physics checks (crates/embloks-core/tests/app.rs:
induction at c = 8 and c = 40, the antenna's travelling wave, energy into the
resistor, the steady B) on the HIP backend through its C ABI, with the
reference's own bars. The same checks pass on the FP64 oracle
(test_oracle_app.py)."""
import pytest

from tests import app_checks as A

gpu = pytest.mark.gpu


@gpu
def test_transformer_drives_the_inner_loop_by_induction_only():
    r = A.transformer(A.GpuApp)
    print("transformer (load/primary, phase) at c = 8 and 40:", r)
    A.check_transformer(r)


@gpu
def test_antenna_radiates_waves_that_the_receiver_picks_up():
    r = A.antenna(A.GpuApp)
    print("antenna (phase slope, receiver amplitude) driven and not:", r)
    A.check_antenna(r)


@gpu
def test_textbook_circuit_sends_energy_into_the_resistor():
    r = A.energy_into_resistor(A.GpuApp)
    print("energy into the resistor:", r)
    A.check_energy_into_resistor(r)


@gpu
def test_magnetic_field_next_to_the_battery_is_steady():
    r = A.steady_b(A.GpuApp)
    print("steady B flicker:", r)
    A.check_steady_b(r)
