"""sgnx_gauges_noalloc (the allocation-free gauge read) against sgnx_gauges,
bit for bit, through the C-ABI: two calls in a row, the second with fewer
nodes, then the size limit. The static scratch array keeps the first call's
contents, so a wrong offset or a stale value would show in the second call.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _read(lib, h, fn, nodes):
    nodes = np.ascontiguousarray(nodes, dtype=np.int64)
    out = np.full(nodes.size, np.nan)
    rc = fn(h, nodes.ctypes.data_as(C.POINTER(C.c_longlong)), int(nodes.size),
            out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_gauges_noalloc_matches_gauges():
    import sgnx
    from sgnx import cases
    from sgnx import _lib as L
    c = cases.solitary()
    s = sgnx.SGN2D(**c.solver_kwargs())
    s.upload(c.U0)
    # a state that is not the initial one: a few adaptive steps
    st = s.integrateAdaptive(0.0, 1e9, s.cflTimeStep(None, 0.5), sgnx.RkOpts(max_accepted=5))
    assert st.accepted == 5
    lib, h = s.lib, s.h
    nd = c.nd
    rng = np.random.default_rng(3)
    first = rng.choice(nd, size=L_MAX(), replace=False)
    second = rng.choice(nd, size=37, replace=False)
    for nodes in (first, second):
        rc0, want = _read(lib, h, lib.sgnx_gauges, nodes)
        rc1, got = _read(lib, h, lib.sgnx_gauges_noalloc, nodes)
        assert rc0 == L.SGNX_OK and rc1 == L.SGNX_OK
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), nodes.size
        # and eta = h - H from the downloaded state, as the gauges define it
        U = s.download()
        assert np.array_equal(got, U[nodes] - c.depth["Hnod"][nodes])
    # beyond the scratch array: refused before anything is copied
    rc, out = _read(lib, h, lib.sgnx_gauges_noalloc, np.arange(L_MAX() + 1))
    assert rc == L.SGNX_EINVAL and np.isnan(out).all()
    assert b"at most" in lib.sgnx_last_error(h)
    # an empty call and a node out of range
    assert _read(lib, h, lib.sgnx_gauges_noalloc, np.zeros(0))[0] == L.SGNX_OK
    assert _read(lib, h, lib.sgnx_gauges_noalloc, np.array([nd]))[0] == L.SGNX_EINVAL
    # the solver wrapper takes the same path
    assert np.array_equal(s.gauges(second, noalloc=True), s.gauges(second))
    s.close()


def L_MAX():
    """SGNX_GAUGES_NOALLOC_MAX as include/sgnx.h defines it."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "include", "sgnx.h")) as f:
        return int(re.search(r"#define SGNX_GAUGES_NOALLOC_MAX (\d+)", f.read()).group(1))
