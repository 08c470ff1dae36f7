"""The oracle reproduces the reference's manufactured-solution errors
(results/convergence2d.csv:14-15, P = 4 at N = 4 and 8): the forced RHS, RK4
to t = 0.1 and the L2 error of h, as apps/convergence2d.cpp:92-148 runs them.
The printed values carry 7 significant digits (a rounding of at most 4e-7
relative); the run differs from the reference's only in the summation order of
the error's quadrature and in numpy's sin/cos, so the bar is 1e-6 relative
(observed: under 1e-7)."""
import numpy as np
import pytest

import mms_lib
import oracle_lib as OL


@pytest.mark.parametrize("N", [4, 8])
def test_oracle_mms_errors_match_reference(N):
    from sgnx import cases
    c = cases.mms(N)
    o = OL.Oracle.from_case(c)
    qw = np.empty(5)
    x = np.empty(5); D = np.empty(25); lL = np.empty(5); lR = np.empty(5)
    gL = np.empty(5); gR = np.empty(5)
    OL.oracle().orc_fr_operators(4, 0, OL.dp(x), OL.dp(D), OL.dp(lL), OL.dp(lR), OL.dp(gL),
                                 OL.dp(gR), OL.dp(qw))
    err = mms_lib.run(c, o.rhs, qw)
    o.close()
    want = mms_lib.golden_rows()[N]["hL2err"]
    assert err == pytest.approx(want, rel=1e-6), (N, err, want)
