"""Device numerics the bit-exact parity rests on.  Needs an MI355X.

On slices whose coordinates are proven in range (0 or magnitude in
[2^-50, 2^20]), the Jacobi gather divides with a shared refined reciprocal
and Markstein's FMA correction, and takes square roots with the hardware
root plus Tuckerman's neighbour test, instead of the compiler's general
expansions.  Those are correctly rounded iff the refined reciprocal equals
RN(1/b) (checked for EVERY significand, both signs, over the exponents used),
the quotient step is (checked against '/' on random operands over the
numerator / denominator exponent ranges of the gather), and the root equals
sqrtf (checked for EVERY positive normal input with exponent in [-100, 127]
and for 0).  Below about 2^-101 the hardware root is not within one ulp;
the gather takes squared lengths under 2^-100 as 0, which its 1e-8 length
test skips exactly as it skips the true root.
"""
import pytest

pytestmark = pytest.mark.gpu


def test_refined_reciprocal_and_quotient_are_correctly_rounded(X):
    print("case division selftest", flush=True)
    r = X.selftest_division(-60, 60, pairs=1 << 28, seed=7)
    assert r["rcp_mismatches"] == 0, r
    assert r["div_mismatches"] == 0, r


def test_fast_sqrt_is_correctly_rounded_on_every_input_used(X):
    print("case sqrt selftest", flush=True)
    r = X.selftest_sqrt(-100, 127)
    assert r["sqrt_mismatches"] == 0, r
