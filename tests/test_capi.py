"""CPU checks of the product library: it loads, exports every symbol that
include/sgnx.h declares, its host setup helpers reproduce the reference tables
and coordinates bit for bit, and its descriptor validation reports errors
(no GPU needed; no kernel is launched)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as OL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "sgnx.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(sgnx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(sgnx_mod):
    lib = sgnx_mod.load()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers exactly the declared set
    from sgnx import _lib
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == names


@pytest.mark.parametrize("P", range(1, 7))
def test_fr_tables_bitwise_vs_oracle(sgnx_mod, P):
    ops = sgnx_mod.FROperators1D(P)
    n = P + 1
    arrs = [np.zeros(n), np.zeros(n * n)] + [np.zeros(n) for _ in range(5)]
    OL.oracle().orc_fr_operators(P, 0, *[OL.dp(a) for a in arrs])
    got = [ops.nodes(), ops.D(), ops.lL(), ops.lR(), ops.gLprime(), ops.gRprime(), ops.weights()]
    for g, w in zip(got, arrs):
        assert np.array_equal(g.view(np.uint64), w.view(np.uint64))


@pytest.mark.parametrize("Nx,Ny,Lx,Ly", [(8, 6, 10.0, 8.0), (13, 7, 3.3, 1.9), (2, 2, 1.0, 1.0)])
def test_coordinates_bitwise_vs_oracle(sgnx_mod, Nx, Ny, Lx, Ly):
    x, y = sgnx_mod.node_coords(4, Nx, Ny, Lx, Ly)
    nd = Nx * Ny * 25
    ox, oy = np.zeros(nd), np.zeros(nd)
    OL.oracle().orc_node_coords(4, Nx, Ny, Lx, Ly, OL.dp(ox), OL.dp(oy))
    assert np.array_equal(x, ox) and np.array_equal(y, oy)
    for wx, wy in ((1, 1), (0, 0), (1, 0)):
        got = sgnx_mod.face_coords(4, Nx, Ny, Lx, Ly, wx, wy)
        nfx, nfy = (Nx + 1) * Ny * 5, Nx * (Ny + 1) * 5
        want = [np.zeros(nfx), np.zeros(nfx), np.zeros(nfy), np.zeros(nfy)]
        OL.oracle().orc_face_coords(4, Nx, Ny, Lx, Ly, wx, wy, *[OL.dp(a) for a in want])
        for g, w in zip(got, want):
            assert np.array_equal(g, w)


def _desc(sgnx_mod, P=4, Nx=4, Ny=4, lL=None, bc=(0, 0, 0, 0)):
    from sgnx import _lib as L
    ops = sgnx_mod.FROperators1D(4)
    keep = [ops.D(), ops.lL() if lL is None else lL, ops.lR(), ops.gLprime(), ops.gRprime(),
            ops.weights()]
    d = L.Desc()
    d.P, d.Nx, d.Ny, d.Lx, d.Ly = P, Nx, Ny, 1.0, 1.0
    d.par = L.Params(9.81, 1.0, 1.5, 0.0)
    d.D, d.lL, d.lR, d.gLp, d.gRp, d.qw = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in keep]
    for k in range(4):
        d.bc[k] = bc[k]
    return d, keep


@pytest.mark.parametrize("change,code", [
    (dict(P=3), 4), (dict(Nx=1), 1), (dict(bc=(0, 1, 0, 0)), 1), (dict(bc=(1, 1, 2, 1)), 4),
    (dict(bc=(2, 1, 1, 1)), 1), (dict(lL=np.array([1.0, 0.0, 0.0, 0.0, 1e-300])), 4),
    (dict(Nx=4096, Ny=2622), 4)])   # 268,492,800 nodes per variable: over the 2^28 the kernels index
def test_descriptor_validation_reports_errors(sgnx_mod, change, code):
    from sgnx import _lib as L
    d, keep = _desc(sgnx_mod, **change)
    h = C.c_void_p()
    rc = sgnx_mod.load().sgnx_create(C.byref(d), C.byref(h))
    assert rc == code and not h.value
    assert sgnx_mod.load().sgnx_last_error(None)
    del keep, L


def test_python_mirror_raises_on_bad_arguments(sgnx_mod):
    with pytest.raises(ValueError):
        sgnx_mod.SGN2D(4, 1, 4, 1.0, 1.0, sgnx_mod.Params())
    with pytest.raises(sgnx_mod.SgnxError) as e:
        sgnx_mod.SGN2D(3, 4, 4, 1.0, 1.0, sgnx_mod.Params())
    assert e.value.code == 4


def test_no_device_is_an_error_not_an_abort():
    """Creating a solver where no GPU is visible returns SGNX_EDEVICE (run in a
    child process with the GPUs hidden, so it holds on a machine with a GPU too)."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import sgnx\n"
        "from sgnx import cases\n"
        "try:\n"
        "    sgnx.SGN2D(**cases.gate().solver_kwargs())\n"
        "    print('CREATED')\n"
        "except sgnx.SgnxError as e:\n"
        "    print('ERR', e.code)\n" % os.path.join(ROOT, "sgn-hyperbolic-gpu_amd"))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1",
               CUDA_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert "ERR 2" in out.stdout or "ERR 3" in out.stdout, out.stdout + out.stderr


def _revolve_cost(L, c, memo={}):
    """Forward steps an optimal schedule needs to reverse L steps from a held
    state with c free state slots (dynamic programme over the first checkpoint)."""
    if L <= 1:
        return 0
    if c <= 0:
        return L * (L - 1) // 2
    key = (L, c)
    if key not in memo:
        memo[key] = min(m + _revolve_cost(L - m, c - 1) + _revolve_cost(m, c) for m in range(1, L))
    return memo[key]


@pytest.mark.parametrize("steps,states", [(13, 3), (13, 4), (13, 6), (13, 13), (40, 5), (60, 7),
                                          (97, 9), (2, 1), (1, 3), (30, 1)])
def test_planned_snapshots_give_the_binomial_optimum(sgnx_mod, steps, states):
    """sgnx_tape_plan_positions (host only): snapshot steps from 0, ascending,
    at most `states` of them; reversing segment j with states - j - 1 free
    slots, the last step's input held from the taped run, recomputes exactly
    F(L, states - 1) - (L - 1) steps, F the dynamic programme's optimum."""
    lib = sgnx_mod.load()
    at = (C.c_long * 64)()
    n = C.c_long()
    assert lib.sgnx_tape_plan_positions(steps, states, at, 64, C.byref(n)) == 0
    p = list(at[:n.value])
    assert p[0] == 0 and all(a < b for a, b in zip(p, p[1:])) and p[-1] < steps
    assert len(p) <= states
    c = states - 1
    cost = 0
    for j, a in enumerate(p):
        b = p[j + 1] if j + 1 < len(p) else steps
        if j == len(p) - 1 and b - a >= 2:
            b -= 1                       # the last step's input is held: reversed without recomputation
        cost += _revolve_cost(b - a, c - j)
    assert cost == _revolve_cost(steps, c) - (steps - 1), (p, cost)


def test_plan_positions_rejects_bad_arguments(sgnx_mod):
    lib = sgnx_mod.load()
    n = C.c_long()
    assert lib.sgnx_tape_plan_positions(0, 3, None, 0, C.byref(n)) != 0
    assert lib.sgnx_tape_plan_positions(10, 0, None, 0, C.byref(n)) != 0
    assert lib.sgnx_tape_plan_positions(10, 3, None, 0, C.byref(n)) == 0 and n.value >= 1
    assert lib.sgnx_tape_plan_positions((1 << 20) + 1, 40, None, 0, C.byref(n)) != 0
    at = (C.c_long * 64)()
    assert lib.sgnx_tape_plan_positions(1 << 20, 40, at, 64, C.byref(n)) == 0
    assert 1 <= n.value <= 40 and at[0] == 0
