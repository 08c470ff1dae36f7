"""Accuracy of the GPU right-hand side pinned by the reference's own
manufactured-solution errors (results/convergence2d.csv:14-17, P = 4 at
N = 4, 8, 16, 32): sgnx_rhs (host buffers, through the C-ABI) plus the
manufactured forcing on the host, RK4 to t = 0.1 and the L2 error of h, as
apps/convergence2d.cpp:92-148 runs them. The GPU RHS is bitwise the CPU
rhs(), so the errors match the printed 7 digits (bar 1e-6 relative, the
printed rounding plus margin), and the observed orders are the reference's
(4.0 between consecutive meshes)."""
import math

import numpy as np
import pytest

import mms_lib

pytestmark = pytest.mark.gpu


def test_gpu_mms_errors_match_reference():
    import sgnx
    from sgnx import cases
    want = mms_lib.golden_rows()
    got = {}
    for N in (4, 8, 16, 32):
        c = cases.mms(N)
        s = sgnx.SGN2D(**c.solver_kwargs())
        qw = s.ops.weights()
        print(f"[mms] N={N}: {c.meta['steps']} RK4 steps", flush=True)
        got[N] = mms_lib.run(c, lambda U, t: s.rhs(U, t), qw)
        s.close()
        assert got[N] == pytest.approx(want[N]["hL2err"], rel=1e-6), (N, got[N], want[N])
    for N in (8, 16, 32):
        order = math.log2(got[N // 2] / got[N])
        assert order == pytest.approx(want[N]["order"], abs=2e-3), (N, order)
