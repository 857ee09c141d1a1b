"""User perturbations on the GPU, through the C ABI: bw_apply_perturb
(cBaseCloth::ApplyPerturb), bw_apply_user_perturb_force_once
(cBaraffCloth::ApplyUserPerturbForceOnce) and bw_set_y_drag_force
(cBaseCloth::mYDragForce).

Bounds:
  * reference semantics (max_iters-1 FP64 PCG iterations) against the
    compiled reference's golden trajectories: the first step within 1e-9 of
    the displacement, later steps of the 21 x 21 cloth within 1e-4 (its
    trajectory amplifies rounding-level differences, as in the unperturbed
    golden tests);
  * tolerance mode (FP64 and mixed, 1e-6) against the oracle solved to
    1e-10 from the same inputs: within 1e-3 of the displacement over 12
    steps, every step's residual <= 1e-6;
  * a perturbed member of a batch equals the same cloth alone, bit for bit;
  * the dragged vertex ends its step exactly at the goal.
"""
import copy
import sys

import numpy as np
import pytest

import oracle_ffi as of
from acousticsim_amd import workloads as wl
from conftest import load_golden

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def capi():
    from acousticsim_amd import capi as c
    c.lib()
    return c


def _inputs(g, step):
    return [(int(kind), int(tri), np.array([a, b, c]))
            for st, kind, tri, a, b, c in g["perturb"] if int(st) == step]


def _apply_cloth(cl, g, step):
    for kind, tri, v in _inputs(g, step):
        if kind == 0:
            cl.ApplyPerturb(tri, v)
        else:
            cl.ApplyUserPerturbForceOnce(tri, v)


def _apply_oracle(oc, g, step):
    for kind, tri, v in _inputs(g, step):
        if kind == 0:
            oc.apply_perturb(tri, v)
        else:
            oc.apply_user_perturb_force_once(tri, v)


@pytest.mark.parametrize("name", ["grid21_perturb", "grid5_perturb"])
def test_reference_semantics_with_perturbations(capi, name):
    g = load_golden(name)
    conf = dict(g["config"], precision="fp64", solver_mode="reference")
    say("case perturbed reference-mode", name)
    cl = capi.BaraffClothHip(conf, g["x0"])
    cl.SetYDragForce(float(g["ydrag"]))
    steps = [int(s) for s in g["steps"]]
    for s in range(1, max(steps) + 1):
        _apply_cloth(cl, g, s)
        cl.UpdatePos()
        if s in steps:
            x = cl.GetPos()
            xr = g["x_step%d" % s]
            disp = np.abs(xr - g["x0"]).max()
            err = np.abs(x - xr).max()
            say("  step %d err/disp %.3e" % (s, err / disp))
            assert err <= (1e-9 if s == 1 or name != "grid21_perturb" else 1e-4) * disp


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tolerance_mode_with_perturbations(capi, precision):
    g = load_golden("grid21_perturb")
    conf = dict(g["config"], precision=precision)
    say("case perturbed tolerance-mode", precision)
    cl = capi.BaraffClothHip(conf, g["x0"])
    cl.SetYDragForce(float(g["ydrag"]))
    oc = of.OracleCloth(g["config"])
    oc.set_state(g["x0"], g["x0"])
    oc.set_y_drag_force(float(g["ydrag"]))
    worst = 0.0
    for s in range(1, 13):
        _apply_cloth(cl, g, s)
        _apply_oracle(oc, g, s)
        cl.UpdatePos()
        oc.step(of.SOLVE_TOLERANCE, 1e-10)
        assert cl.last_stats()["rel_residual"] <= 1e-6
        xr = oc.pos()
        err = np.abs(cl.GetPos() - xr).max() / np.abs(xr - g["x0"]).max()
        worst = max(worst, err)
        for kind, tri, v in _inputs(g, s):
            if kind == 1:
                v0 = oc.triangles()[tri][0]
                assert np.array_equal(cl.GetPos().reshape(-1, 3)[v0], v)
    say("  worst err/disp %.3e" % worst)
    assert worst <= 1e-3


def test_perturbed_member_alone_equals_in_batch(capi, monkeypatch):
    monkeypatch.setenv("BW_SPMV_LPR", "1")
    monkeypatch.setenv("BW_SPMV_STAGE", "0")
    g = load_golden("grid21_perturb")
    conf = dict(g["config"], precision="mixed")
    others = []
    for side, seed in ((17, 5), (33, 6)):
        c = wl.grid_config(side)
        c["precision"] = "mixed"
        others.append((c, wl.jittered_positions(c, seed)))
    say("case perturbed member in a batch")
    b = capi.ClothBatch([others[0][0], conf, others[1][0]], [others[0][1], g["x0"], others[1][1]])
    b.set_y_drag_force(1, float(g["ydrag"]))
    a = capi.ClothBatch([conf], [g["x0"]])
    a.set_y_drag_force(0, float(g["ydrag"]))
    for s in range(1, 6):
        for kind, tri, v in _inputs(g, s):
            if kind == 0:
                b.apply_perturb(1, tri, v)
                a.apply_perturb(0, tri, v)
            else:
                b.apply_user_perturb_force_once(1, tri, v)
                a.apply_user_perturb_force_once(0, tri, v)
        b.step(1)
        a.step(1)
    assert np.array_equal(a.get_pos(0), b.get_pos(1))
    # the unperturbed members are unaffected
    for k, i in ((0, 0), (1, 2)):
        c, x0 = others[k]
        alone = capi.ClothBatch([c], [x0])
        alone.step(5)
        assert np.array_equal(alone.get_pos(0), b.get_pos(i))


def test_reset_clears_user_force_and_errors(capi, monkeypatch):
    # one SpMV variant for both cloths (the autotuner's pick may differ
    # between them, and variants differ at the rounding level)
    monkeypatch.setenv("BW_SPMV_LPR", "1")
    monkeypatch.setenv("BW_SPMV_STAGE", "0")
    c = wl.grid_config(9)
    c["precision"] = "fp64"
    x0 = wl.jittered_positions(c, 8)
    say("case perturb reset and errors")
    cl = capi.BaraffClothHip(c, x0)
    cl.ApplyPerturb(10, [0.5, 0.0, 0.0])
    cl.Reset()
    cl.SetPos(x0)
    cl.SetPrePos(x0)
    cl.UpdatePos()
    fresh = capi.BaraffClothHip(c, x0)
    fresh.UpdatePos()
    assert np.array_equal(cl.GetPos(), fresh.GetPos())
    with pytest.raises(capi.BwError) as e:
        cl.ApplyPerturb(cl.GetNumOfTriangles(), [1.0, 0.0, 0.0])
    assert e.value.code == -1
    with pytest.raises(capi.BwError) as e:
        cl.ApplyUserPerturbForceOnce(-1, [0.0, 0.0, 0.0])
    assert e.value.code == -1
    with pytest.raises(capi.BwError) as e:
        cl.ApplyPerturb(0, [np.nan, 0.0, 0.0])
    assert e.value.code == -1
    gpu_form = capi.BaraffClothHip(dict(c, cloth_type="fem_gpu"), x0)
    with pytest.raises(capi.BwError) as e:
        gpu_form.ApplyPerturb(0, [1.0, 0.0, 0.0])
    assert e.value.code == -4
