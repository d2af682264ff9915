"""The reference's behavioural acceptance tests on both GPU paths, float64, at the reference's
own sizes, step counts and float64 bars (scenes in tests/behaviour.py; the oracle runs the same
scenes in test_oracle_behaviour.py):

* reference tests/test_ipc.py:128-159  slide at the Coulomb speed (four cases)
* reference tests/test_ipc.py:163-181  a stack carries its weight
* reference tests/test_ipc.py:207-212  a grip needs friction
* reference tests/test_ipc.py:245-253  nothing passes through; momentum through an impact
* reference tests/test_ipc.py:280-295  a prescribed body is swept
* reference tests/test_objects.py:84-110  sliding friction = mu x normal, every shape + ground
* reference tests/test_rigid.py:162-189  a block on a slope (rigid and soft)
* reference tests/test_rigid.py:252-274  momentum through a rigid-soft impact
"""
import numpy as np
import pytest

from tests import behaviour as BH

pytestmark = pytest.mark.gpu
PATHS = ["scene", "plain"]


@pytest.mark.parametrize("path", PATHS)
def test_a_prescribed_body_is_swept(path):
    lag_px, squash = BH.check_swept(lambda: BH.GPUStepper(path))
    print(f"{path}: against the wall {lag_px:.1f} px short, smallest area {squash:.3f}")


@pytest.mark.parametrize("path", PATHS)
def test_nothing_passes_through(path):
    for other, diagonal in (("soft", False), ("rigid", False), ("held", True), ("held", False)):
        gap, inside, _ = BH.throw(BH.GPUStepper(path), other, 5.0, diagonal)
        assert gap > 0.0 and not inside, \
            f"thrown at a {other} block at 5 px/step: gap {gap:.2e} dhat, inside {inside}"
    _, _, dp = BH.throw(BH.GPUStepper(path), "soft", 0.5)
    print(f"{path}: momentum change through the impact {dp:.1e}")
    assert dp < 2e-3


@pytest.mark.parametrize("path", PATHS)
def test_sliding_friction_is_mu_times_normal(path):
    r = BH.friction_ratio(lambda: BH.GPUStepper(path))
    print(path, {k: round(v, 5) for k, v in r.items()})
    for shape, ratio in r.items():
        assert abs(ratio - BH.MU_OBJ) < 0.03 * BH.MU_OBJ, f"{shape}: {ratio:.4f}"


CASES = [("rigid", 0.0, 40.0), ("rigid", 0.3, 30.0), ("rigid", 0.5, 15.0), ("soft", 0.3, 30.0)]


@pytest.mark.parametrize("path,case", [("scene", c) for c in CASES] + [("plain", CASES[1]), ("plain", CASES[3])])
def test_slides_at_the_coulomb_speed(path, case):
    kind, mu, th = case
    v, expect, gap = BH.slide(BH.GPUStepper(path), kind, mu, th)
    print(f"{path}: {kind} block, mu {mu}, {th:.0f} deg: {v:.5e} m/s, closed form {expect:.5e}; "
          f"resting gap {gap:.2f} dhat")
    assert 0.5 < gap < 1.0
    assert abs(v / expect - 1.0) < (0.01 if kind == "rigid" else 0.02)


@pytest.mark.parametrize("path", PATHS)
def test_a_stack_carries_its_weight(path):
    gap, vmax, ratio = BH.stack(BH.GPUStepper(path))
    print(f"{path}: gap {gap:.3f} dhat, upper block speed {vmax:.1e}, ground carries {ratio:.5f}")
    assert 0.0 < gap < 1.0
    assert vmax < 2e-3
    assert abs(ratio - 1.0) < 1e-3


@pytest.mark.parametrize("path", ["scene"])
def test_a_grip_needs_friction(path):
    rose = BH.pinch_and_lift(BH.GPUStepper(path), 0.0)
    assert abs(rose) < 0.05, f"with no friction the block came up {rose:.3f} of the lift"
    rose = BH.pinch_and_lift(BH.GPUStepper(path), 0.8)
    print(f"{path}: with friction the block rose {rose:.3f} of the lift")
    assert rose > 0.95


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("kind", ["rigid", "soft"])
def test_block_on_a_slope(path, kind):
    for th, mu in ((30.0, 0.30), (15.0, 0.60)):             # slides; holds and creeps
        NW, v, expect = BH.slope(BH.GPUStepper(path, self_collision=False), kind, th, mu)
        print(f"{path}: {kind} block, {th:.0f} deg, mu {mu}: ground carries {NW:.4f}; "
              f"{v:.4e} m/s, closed form {expect:.4e}")
        assert abs(NW - 1.0) < 5e-3
        assert abs(v / expect - 1.0) < 0.01


@pytest.mark.parametrize("path", PATHS)
def test_momentum_through_a_rigid_soft_impact(path):
    vs, dp, gap = BH.impact(BH.GPUStepper(path, friction=False))
    print(f"{path}: soft block pushed to {vs:.3f} m/s, momentum change {dp:.1e}, gap {gap:.2e} m")
    assert vs > 0.05 and dp < 1e-4 and gap > 0.0
