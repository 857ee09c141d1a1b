"""User perturbations in the oracle, against the reference's own outputs.

The golden cases grid21_perturb and grid5_perturb (tests/golden/make_golden.py)
drive the compiled reference's cBaraffCloth with tPerturb records before
given steps: forces through cBaseCloth::ApplyPerturb (BaseCloth.cpp:76-87)
and dragged vertices through cBaraffCloth::ApplyUserPerturbForceOnce
(BaraffCloth.cpp:165-178), plus a y drag force on the top and bottom UV
edges every step (mYDragForce, ApplyYDragForce BaseCloth.cpp:592-623;
grid21_perturb only).  In reference-solve mode the oracle reproduces those
trajectories bit for bit.  Closed forms besides: a perturbation force F
changes the rhs by exactly dt^2 F / 3 on each of the triangle's vertices,
and a dragged vertex ends its step at the goal."""
import numpy as np
import pytest

import oracle_ffi as of
from acousticsim_amd import workloads as wl
from conftest import load_golden


def apply(oc, g, step):
    for st, kind, tri, a, b, c in g["perturb"]:
        if int(st) == step:
            if int(kind) == 0:
                oc.apply_perturb(int(tri), [a, b, c])
            else:
                oc.apply_user_perturb_force_once(int(tri), [a, b, c])


@pytest.mark.parametrize("name", ["grid21_perturb", "grid5_perturb"])
def test_perturbed_trajectory_bitexact(name):
    g = load_golden(name)
    oc = of.OracleCloth(g["config"])
    oc.set_state(g["x0"], g["x0"])
    oc.set_y_drag_force(float(g["ydrag"]))
    steps = [int(s) for s in g["steps"]]
    for s in range(1, max(steps) + 1):
        apply(oc, g, s)
        oc.step(of.SOLVE_REFERENCE)
        if s == 1:
            assert np.array_equal(oc.dx(), g["s1_dx"])
        if s in steps:
            assert np.array_equal(oc.pos(), g["x_step%d" % s]), s


def test_fixture_sees_every_input():
    """each input changes the reference's trajectory: dropping the y drag
    force, a perturbation force or a drag leaves the fixture unmatched"""
    g = load_golden("grid21_perturb")

    def run(ydrag=True, skip=None):
        oc = of.OracleCloth(g["config"])
        oc.set_state(g["x0"], g["x0"])
        if ydrag:
            oc.set_y_drag_force(float(g["ydrag"]))
        for s in range(1, 4):
            for k, (st, kind, tri, a, b, c) in enumerate(g["perturb"]):
                if int(st) == s and k != skip:
                    if int(kind) == 0:
                        oc.apply_perturb(int(tri), [a, b, c])
                    else:
                        oc.apply_user_perturb_force_once(int(tri), [a, b, c])
            oc.step(of.SOLVE_REFERENCE)
        return oc.pos()

    ref = g["x_step3"]
    assert np.array_equal(run(), ref)
    assert not np.array_equal(run(ydrag=False), ref)
    assert not np.array_equal(run(skip=0), ref)  # a force at step 1
    assert not np.array_equal(run(skip=3), ref)  # the drag at step 2


def test_force_enters_rhs_as_dt2_third():
    c = wl.grid_config(9)
    x0 = wl.jittered_positions(c, 3)
    oc = of.OracleCloth(c)
    oc.set_state(x0, x0)
    oc.assemble()
    b0 = oc.rhs()
    F = np.array([0.3, -0.2, 0.7])
    oc.apply_perturb(17, F)
    oc.assemble()
    b1 = oc.rhs()
    tri = oc.triangles()[17]
    dt = c["default_timestep"]
    expect = np.zeros_like(b0).reshape(-1, 3)
    expect[tri] = dt * dt * (F / 3)
    assert np.allclose((b1 - b0).reshape(-1, 3), expect, rtol=1e-12, atol=1e-18)
    # the user force lasts one step
    oc.step(of.SOLVE_TOLERANCE, 1e-10)
    oc.set_state(x0, x0)
    oc.assemble()
    assert np.array_equal(oc.rhs(), b0)


def test_drag_vertex_ends_at_goal_and_is_released():
    c = wl.grid_config(9)
    x0 = wl.jittered_positions(c, 4)
    oc = of.OracleCloth(c)
    oc.set_state(x0, x0)
    tri = 40
    v0 = int(oc.triangles()[tri][0])
    goal = x0.reshape(-1, 3)[v0] + np.array([0.0, 0.01, 0.02])
    oc.apply_user_perturb_force_once(tri, goal)
    oc.step(of.SOLVE_TOLERANCE, 1e-10)
    assert np.array_equal(oc.pos().reshape(-1, 3)[v0], goal)
    assert np.array_equal(oc.pre_pos().reshape(-1, 3)[v0], goal)
    # the next step moves it again (the fix lasts one step)
    oc.step(of.SOLVE_TOLERANCE, 1e-10)
    assert not np.array_equal(oc.pos().reshape(-1, 3)[v0], goal)
    with pytest.raises(ValueError):
        oc.apply_perturb(oc.nt, [1.0, 0.0, 0.0])
