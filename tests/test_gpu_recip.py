"""The pair passes divide floats through a double reciprocal
(device_math.h: recip, div_recip): the hardware estimate refined by two
Newton steps instead of the division sequence.  It must equal the correctly
rounded 1.0 / double(b) for every float b, which is what the oracle computes
(and what keeps float(double(a) * recip(b)) the correctly rounded a / b,
tests/test_division_identity.py).  The library's self-check compares the two
on the device over all 2^32 bit patterns.  The fluid density pass takes its
square roots the same way (hardware estimate, then a correction), which must
equal the correctly rounded sqrtf of the oracle for every float."""
import pytest

pytestmark = pytest.mark.gpu


def test_device_reciprocal_is_correctly_rounded_for_every_float(hip_device):
    from sph import bindings
    assert bindings.check_recip() == 0


def test_device_square_root_is_correctly_rounded_for_every_float(hip_device):
    from sph import bindings
    assert bindings.check_sqrt() == 0
