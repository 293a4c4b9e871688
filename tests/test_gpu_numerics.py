"""Device numerics the bit-exact parity rests on.  Needs an MI355X.

The Jacobi gather divides with a shared refined reciprocal and Markstein's
FMA correction instead of the compiler's general f32 division expansion.
That is correctly rounded iff the refined reciprocal equals RN(1/b); this
checks it for EVERY significand (both signs) over the exponent range the
kernel admits, and the quotient against '/' on random pairs.
"""
import pytest

pytestmark = pytest.mark.gpu


def test_refined_reciprocal_and_quotient_are_correctly_rounded(X):
    print("case division selftest", flush=True)
    r = X.selftest_division(-60, 60, pairs=1 << 28, seed=7)
    assert r["rcp_mismatches"] == 0, r
    assert r["div_mismatches"] == 0, r
