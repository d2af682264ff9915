"""d2f_step_scene_noalloc, the scene step with the library's static work arrays, against
d2f_step_scene with the solver's own work arrays, through the C ABI.

Two calls in a row on one device: first a grip scene of two environments (affine block, soft
pads, kinematic pad backs, IPC contact with friction), then a smaller grip scene of one
environment whose kinematic nodes sit at other indices, so that the second call starts on what
the first left in the static array. The same bar as every scene-path test: 1e-9 of the body
size in float64, 1e-5 in mixed precision (the two runs differ only in the order of float64
atomics). The solver's own work arrays stay untouched, and a scene larger than the static
array is refused with D2F_EINVAL.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from diff_2d_fem_amd import ImplicitIntegrator, _lib
from tests.scenes import close_jaws, grip

pytestmark = pytest.mark.gpu

KW = dict(max_iter=60, eps=0.0, conv_check=0)
WORK = ("xhat", "grad", "grad_prev", "p", "diag", "diag_el", "x0", "bc_q")


def _pair(make, precision, steps=8):
    (ma, pa), (mb, pb) = make(), make()
    ga = ImplicitIntegrator(path="scene", precision=precision, **KW)
    gb = ImplicitIntegrator(path="scene", precision=precision, **KW)
    for k in range(1, steps + 1):
        close_jaws(ma, pa, k, n_close=6)
        close_jaws(mb, pb, k, n_close=6)
        if k == 1:
            ga.attach(ma, 1e-3)
            gb.attach(mb, 1e-3)
            mb._solver.noalloc = True
        ga.simulate(ma, 1e-3)
        gb.simulate(mb, 1e-3)
    torch.cuda.synchronize()
    return ma, mb


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_noalloc_step_equals_the_allocating_step(precision):
    bar = 1e-9 if precision == "double" else 1e-5
    for label, make, size in (("grip, 2 environments", lambda: grip(n_env=2), 4e-3),
                              ("grip, 1 smaller environment", lambda: grip(n_env=1, block=6), 3e-3)):
        ma, mb = _pair(make, precision)
        sa, sb = ma._solver, mb._solver
        assert sb.N < _lib.NOALLOC_MAX_NODES
        err = np.abs(mb.get_x() - ma.get_x()).max() / size
        errv = np.abs(sb.get_v() - sa.get_v()).max() / max(np.abs(sa.get_v()).max(), 1e-12)
        print(f"{label}, {precision}: N={sb.N}, max |dx| / size = {err:.3e}, |dv| rel = {errv:.3e}",
              flush=True)
        assert np.isfinite(mb.get_x()).all()
        assert err < bar and errv < 1e3 * bar
        assert np.abs(sb.get_x() - sb.me.rest_pos).max() > 1e-5      # the pads moved the scene
        for k in WORK:
            assert not torch.any(getattr(sb, k) != 0), f"the caller's {k} was written"


def test_noalloc_refuses_a_scene_beyond_its_static_arrays():
    L = _lib.lib()
    s, q = _lib.Scene(), _lib.Params()
    s.B, s.N = 1, _lib.NOALLOC_MAX_NODES + 1
    rc = L.d2f_step_scene_noalloc(C.byref(s), C.byref(q), None)
    assert rc == -1
    assert b"static work arrays hold at most" in L.d2f_last_error()
