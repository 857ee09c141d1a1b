"""Collision detection on the GPU (bw_detect_contacts) against the oracle's
restatement of cBVHCollisionDetecter::PerformCD (pinned in
tests/test_collide_cpu.py).  The device applies the same float/double
arithmetic with contraction off, so contact lists are compared exactly:
ids, order, normals, barycentric coordinates and depths bit for bit.
Closed loops (detection before every step) are compared to 1e-6 of the
displacement in reference-semantics FP64 mode.
"""
import sys

import numpy as np
import pytest

import oracle_ffi as of
from acousticsim_amd import workloads as wl

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def capi():
    from acousticsim_amd import capi as c
    c.lib()
    return c


def floor_body(y, half=2.0):
    xyz = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]])
    tri = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    return xyz, tri


def sphere_body(center, radius, n_lat=12, n_lon=24):
    return wl.uv_sphere(center, radius, n_lat, n_lon)


def flat_cloth(side=9, height=0.0, precision="fp64", mode="reference"):
    conf = wl.grid_config(side)
    conf["cloth_init_orientation"] = [-90.0 * np.pi / 180.0, 0.0, 0.0, 0.0]
    conf["cloth_init_pos"] = [0.0, height, 0.0]
    conf["precision"] = precision
    conf["solver_mode"] = mode
    return conf


def assert_same_contacts(a, b):
    for k in ("pt_ids", "pt_normal", "pt_bary", "pt_depth", "ee_ids", "ee_normal", "ee_bary",
              "ee_depth"):
        assert a[k].shape == b[k].shape, (k, a[k].shape, b[k].shape)
        assert np.array_equal(a[k], b[k]), k


def _pair(capi, conf, x, xyz, tri):
    cl = capi.BaraffClothHip(conf, x)
    oc = of.OracleCloth(conf)
    oc.set_state(x, x)
    return cl, oc, capi.KinematicBody(xyz, tri), of.OracleBody(xyz, tri)


@pytest.mark.parametrize("cloth_first", [True, False])
def test_floor_contacts_match_oracle(capi, cloth_first):
    conf = flat_cloth(21)
    oc0 = of.OracleCloth(conf)
    x = oc0.rest_pos().reshape(-1, 3).copy()
    x[:, 1] = np.random.default_rng(5).uniform(-2e-3, 6e-3, len(x))
    xyz, tri = floor_body(0.0)
    cl, oc, body, obody = _pair(capi, conf, x.ravel(), xyz, tri)
    say("case floor contacts, cloth_first", cloth_first)
    assert np.array_equal(body.edges(), obody.edges())
    mine = cl.detect_contacts(body, cloth_first)
    ref = oc.detect_contacts(obody, cloth_first)
    say("  point-triangle %d, edge-edge %d" % (len(ref["pt_ids"]), len(ref["ee_ids"])))
    assert len(ref["pt_ids"]) > 50
    assert_same_contacts(mine, ref)


@pytest.mark.parametrize("cloth_first", [True, False])
def test_sphere_contacts_match_oracle(capi, cloth_first):
    """a cloth draped (by hand) over a sphere: vertex, body-vertex and
    edge-edge contacts all occur"""
    conf = flat_cloth(33)
    oc0 = of.OracleCloth(conf)
    x = oc0.rest_pos().reshape(-1, 3).copy()
    r = 0.15
    rho = np.hypot(x[:, 0], x[:, 2])
    top = np.sqrt(np.maximum(r * r - rho * rho, 0.0))
    rng = np.random.default_rng(6)
    x[:, 1] = np.where(rho < r, top + rng.uniform(-1.5e-3, 2.5e-3, len(x)), -0.05)
    xyz, tri = sphere_body([0.0, 0.0, 0.0], r, n_lat=16, n_lon=32)
    cl, oc, body, obody = _pair(capi, conf, x.ravel(), xyz, tri)
    say("case sphere contacts, cloth_first", cloth_first)
    mine = cl.detect_contacts(body, cloth_first)
    ref = oc.detect_contacts(obody, cloth_first)
    kinds = ref["pt_ids"][:, 0]
    say("  cloth-vertex %d, body-vertex %d, edge-edge %d" % (
        int((kinds == 1).sum()), int((kinds == 0).sum()), len(ref["ee_ids"])))
    assert (kinds == 1).sum() > 20 and (kinds == 0).sum() > 0 and len(ref["ee_ids"]) > 10
    assert_same_contacts(mine, ref)


def test_closed_loop_against_oracle(capi):
    """a cloth falling onto a floor, detection before every step,
    reference-semantics FP64: same contact counts every step, positions
    within 1e-6 of the displacement (rounding-level differences of the
    device's solve, accumulated over 12 steps)"""
    conf = flat_cloth(15, height=2.5e-3)
    conf["constraints"] = {"static_point": []}
    oc0 = of.OracleCloth(conf)
    x = oc0.rest_pos()
    xyz, tri = floor_body(0.0)
    cl, oc, body, obody = _pair(capi, conf, x, xyz, tri)
    say("case closed loop")
    total = 0
    for s in range(12):
        n_dev = cl.detect_contacts(body, True)
        n_ref = oc.detect_contacts(obody, True)
        assert len(n_dev["pt_ids"]) == len(n_ref["pt_ids"]), s
        assert len(n_dev["ee_ids"]) == len(n_ref["ee_ids"]), s
        total += len(n_ref["pt_ids"])
        cl.UpdatePos()
        oc.step(of.SOLVE_REFERENCE)
    disp = np.abs(oc.pos() - x).max()
    err = np.abs(cl.GetPos() - oc.pos()).max()
    say("  contacts %d, err/disp %.3e, lowest %.3e" % (total, err / disp,
                                                      oc.pos().reshape(-1, 3)[:, 1].min()))
    assert total > 0
    assert err <= 1e-6 * disp


def test_batch_member_detection(capi, monkeypatch):
    """detection on one batch member sets only that member's contacts; the
    member then steps as the same cloth alone"""
    monkeypatch.setenv("BW_SPMV_LPR", "2")
    monkeypatch.setenv("BW_SPMV_STAGE", "0")
    confs = [flat_cloth(9, 2e-3, "mixed", "tolerance"), flat_cloth(13, 2e-3, "mixed", "tolerance")]
    xs = [of.OracleCloth(c).rest_pos() for c in confs]
    xyz, tri = floor_body(0.0)
    body = capi.KinematicBody(xyz, tri)
    b = capi.ClothBatch(confs, xs)
    n_pt, n_ee = b.detect_contacts(1, body, True)
    assert n_pt > 0
    b.step(1)
    a = capi.ClothBatch([confs[1]], [xs[1]])
    assert a.detect_contacts(0, body, True) == (n_pt, n_ee)
    a.step(1)
    assert np.array_equal(a.get_pos(0), b.get_pos(1))
    z = capi.ClothBatch([confs[0]], [xs[0]])
    z.step(1)
    assert np.array_equal(z.get_pos(0), b.get_pos(0))


def test_body_errors(capi):
    xyz, tri = floor_body(0.0)
    with pytest.raises(capi.BwError):
        capi.KinematicBody(xyz, np.array([[0, 1, 7]], np.int32))
    fan = np.array([[0, 1, 2], [0, 1, 3], [1, 0, 2]], np.int32)  # edge {0,1} in three triangles
    with pytest.raises(capi.BwError):
        capi.KinematicBody(xyz, fan)
    body = capi.KinematicBody(xyz, tri)
    with pytest.raises(ValueError):
        body.set_pos(np.zeros(3))
    with pytest.raises(capi.BwError):
        body.set_pos(np.full(12, np.nan))
