"""The reference's whole-app physics checks (crates/embloks-core/tests/app.rs)
on the FP64 oracle: induction, the antenna's travelling wave, energy into the
resistor and the steady B. The same checks run on the HIP backend in
test_gpu_app.py."""
import pytest

from tests import app_checks as A


@pytest.fixture(scope="module", autouse=True)
def threads():
    from oracle import orc
    orc.set_threads(8)
    yield
    orc.set_threads(1)


def test_transformer_drives_the_inner_loop_by_induction_only():
    A.check_transformer(A.transformer(A.OracleApp))


def test_antenna_radiates_waves_that_the_receiver_picks_up():
    A.check_antenna(A.antenna(A.OracleApp))


def test_textbook_circuit_sends_energy_into_the_resistor():
    A.check_energy_into_resistor(A.energy_into_resistor(A.OracleApp))


def test_magnetic_field_next_to_the_battery_is_steady():
    A.check_steady_b(A.steady_b(A.OracleApp))
