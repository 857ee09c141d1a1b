"""Collision response on the GPU through the C ABI: contacts handed in
before a step (bw_set_point_triangle_contacts / bw_set_edge_edge_contacts,
the counterparts of cBaseObject::SetPointTriangleCollisionInfo and
SetEdgeEdgeCollisionInfo) against the oracle, whose collision response is
pinned bit for bit to the reference's own (tests/test_contacts_cpu.py).

Bounds (per test):
  * W with contacts: as the system without them, 1e-12 of the largest entry
    (the collision Hessian is summed on the host in the reference's order)
  * the response itself, step with contacts minus step without, on the
    device and in the oracle: within 1e-9 of the response (a wrong constant,
    order or vertex would be off by O(1))
  * reference-semantics trajectories with contacts on every step: step 1
    within 1e-9 of the displacement, later steps 1e-4 (the 21x21 cloth's own
    sensitivity, as without contacts)
  * tolerance solves (mixed and FP64): one step within 1e-4 of the step
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


def _mode(conf, precision, mode="tolerance", tol=1e-6):
    c = copy.deepcopy(conf)
    c["precision"] = precision
    c["solver_mode"] = mode
    c["solver_tolerance"] = tol
    return c


def contacts_of(g):
    return {k[len("contacts_"):]: g[k] for k in g if k.startswith("contacts_")}


def _oracle_step(conf, x, xp, ct, mode=of.SOLVE_REFERENCE):
    """one oracle step; tolerance mode solves to 1e-13"""
    oc = of.OracleCloth(conf)
    oc.set_state(x, xp)
    if ct is not None:
        oc.set_contacts(ct)
    oc.step(mode, tol=1e-13, max_iters=100000)
    return oc.pos()


@pytest.mark.parametrize("name", ["grid5_contacts", "grid21_contacts"])
def test_contact_hessian_in_W(capi, name):
    say("case contact W", name)
    g = load_golden(name)
    ct = contacts_of(g)
    conf = _mode(g["config"], "fp64")
    cl = capi.BaraffClothHip(conf, g["x0"])
    W0, _ = cl.assemble_system()
    cl.set_contacts(ct)
    W1, _ = cl.assemble_system()
    oc = of.OracleCloth(g["config"])
    oc.set_state(g["x0"], g["x0"])
    oc.assemble()
    Wo0 = oc.W()
    oc.set_contacts(ct)
    oc.assemble()
    Wo1 = oc.W()
    rp, cols = oc.pattern()
    pinned = np.zeros(oc.nv, bool)
    pinned[oc.pins()] = True
    rows = np.repeat(np.arange(oc.nv), np.diff(rp))
    free = ~pinned[rows] & ~pinned[cols]
    scale = np.abs(Wo1[free]).max()
    err = np.abs(W1[free] - Wo1[free]).max()
    derr = np.abs((W1 - W0)[free] - (Wo1 - Wo0)[free]).max()
    say("  W err %.3e  dW err %.3e  (scale %.3e, |dW| %.3e)" % (
        err, derr, scale, np.abs((Wo1 - Wo0)[free]).max()))
    assert np.abs((Wo1 - Wo0)[free]).max() > 0
    assert err <= 1e-12 * scale
    assert derr <= 1e-12 * scale
    assert np.all(W1[~free] == 0)


@pytest.mark.parametrize("precision,mode", [("fp64", "reference"), ("fp64", "tolerance"),
                                            ("mixed", "tolerance")])
def test_response_matches_oracle(capi, precision, mode):
    """one step of the 21x21 cloth from the golden x0, with and without the
    golden contacts: the device's response (difference of the two) against
    the oracle's"""
    g = load_golden("grid21_contacts")
    ct = contacts_of(g)
    conf = _mode(g["config"], precision, mode, tol=1e-10 if precision == "fp64" else 1e-8)
    conf["solver_max_iters"] = 20000
    x0 = g["x0"]
    say("case response", precision, mode)
    a = capi.BaraffClothHip(conf, x0)
    a.set_contacts(ct)
    a.UpdatePos()
    b = capi.BaraffClothHip(conf, x0)
    b.UpdatePos()
    om = of.SOLVE_REFERENCE if mode == "reference" else of.SOLVE_TOLERANCE
    oa = _oracle_step(g["config"], x0, x0, ct, om)
    ob = _oracle_step(g["config"], x0, x0, None, om)
    resp = a.GetPos() - b.GetPos()
    oresp = oa - ob
    scale = np.abs(oresp).max()
    err = np.abs(resp - oresp).max()
    say("  response err/|response| %.3e (|response| %.3e)" % (err / scale, scale))
    assert scale > 1e-6
    bound = 1e-9 if mode == "reference" else (1e-6 if precision == "fp64" else 1e-4)
    assert err <= bound * scale


def test_reference_trajectory_with_contacts(capi):
    """contacts on every step, reference semantics, against the reference's
    own trajectory (golden grid21_contacts, 30 steps)"""
    g = load_golden("grid21_contacts")
    ct = contacts_of(g)
    conf = _mode(g["config"], "fp64", "reference")
    cl = capi.BaraffClothHip(conf, g["x0"])
    steps = [int(s) for s in g["steps"]]
    say("case reference trajectory with contacts")
    for s in range(1, max(steps) + 1):
        cl.set_contacts(ct)
        cl.UpdatePos()
        if s in steps:
            xr = g["x_step%d" % s]
            disp = np.abs(xr - g["x0"]).max()
            err = np.abs(cl.GetPos() - xr).max()
            say("  step %d err/disp %.3e" % (s, err / disp))
            assert err <= (1e-9 if s == 1 else 1e-4) * disp


def test_grid5_trajectory_with_contacts(capi):
    """5x5 with damping: reference semantics, 4 steps with contacts each"""
    g = load_golden("grid5_contacts")
    ct = contacts_of(g)
    cl = capi.BaraffClothHip(_mode(g["config"], "fp64", "reference"), g["x0"])
    say("case grid5 contacts")
    for s in range(1, 5):
        cl.set_contacts(ct)
        cl.UpdatePos()
        if s in list(g["steps"]):
            xr = g["x_step%d" % s]
            disp = np.abs(xr - g["x0"]).max()
            err = np.abs(cl.GetPos() - xr).max()
            say("  step %d err/disp %.3e" % (s, err / disp))
            assert err <= 1e-9 * disp


def test_chained_impulses_follow_contact_order(capi):
    """many contacts on few vertices (deep levels in the parallel schedule):
    the response equals the oracle's serial one"""
    conf = wl.grid_config(21)
    x0 = wl.jittered_positions(conf, 3)
    ct = wl.random_contacts(441, 800, 1240, seed=9, n_point=120, n_tri=60, n_edge=60)
    rng = np.random.default_rng(1)
    hot = np.array([200, 201, 222])
    ct["pt_ids"][:120, 1] = rng.choice(hot, 120)
    ct["pt_ids"][120:, 2] = rng.choice([380, 381, 382, 420], 60)
    ct["ee_ids"][:, 0] = 1
    ct["ee_ids"][:, 1] = rng.choice([560, 561, 600], 60)
    say("case chained impulses")
    c = _mode(conf, "fp64", "reference")
    a = capi.BaraffClothHip(c, x0)
    a.set_contacts(ct)
    a.UpdatePos()
    b = capi.BaraffClothHip(c, x0)
    b.UpdatePos()
    oresp = _oracle_step(conf, x0, x0, ct) - _oracle_step(conf, x0, x0, None)
    resp = a.GetPos() - b.GetPos()
    scale = np.abs(oresp).max()
    err = np.abs(resp - oresp).max()
    say("  err/|response| %.3e" % (err / scale))
    assert err <= 1e-9 * scale


def test_contacts_last_one_step_and_batch_members(capi, monkeypatch):
    """contacts apply to the next step only; in a batch they reach only
    their member, which then equals the same cloth alone"""
    monkeypatch.setenv("BW_SPMV_LPR", "2")
    monkeypatch.setenv("BW_SPMV_STAGE", "0")
    g = load_golden("grid21_contacts")
    ct = contacts_of(g)
    confs, x0s = [], []
    for side, seed in ((21, 1), (17, 2), (21, 3)):
        c = wl.grid_config(side)
        c["precision"] = "mixed"
        confs.append(c)
        x0s.append(wl.jittered_positions(c, seed))
    say("case batch contacts")
    b = capi.ClothBatch(confs, x0s)
    b.set_contacts(0, ct)
    b.set_contacts(2, ct)
    b.step(2)
    for i in range(3):
        a = capi.ClothBatch([confs[i]], [x0s[i]])
        if i != 1:
            a.set_contacts(0, ct)
        a.step(2)
        assert np.array_equal(a.get_pos(0), b.get_pos(i)), i
    # member 0 with contacts differs from member 0 without
    n = capi.ClothBatch([confs[0]], [x0s[0]])
    n.step(2)
    assert not np.array_equal(n.get_pos(0), b.get_pos(0))


def test_contact_errors(capi):
    conf = wl.grid_config(5)
    cl = capi.BaraffClothHip(conf)
    ids = np.array([[1, 25, 0]], np.int32)
    nrm = np.array([[0.0, 1.0, 0.0]])
    bary = np.array([[1.0, 0.0, 0.0]])
    with pytest.raises(capi.BwError):  # vertex out of range
        cl.SetPointTriangleCollisionInfo(ids, nrm, bary, np.array([0.0], np.float32))
    with pytest.raises(capi.BwError):  # depth not below 3e-3
        cl.SetPointTriangleCollisionInfo(np.array([[1, 3, 0]], np.int32), nrm, bary,
                                         np.array([3e-3], np.float32))
    with pytest.raises(capi.BwError):  # cloth triangle out of range
        cl.SetPointTriangleCollisionInfo(np.array([[0, 3, 32]], np.int32), nrm, bary,
                                         np.array([0.0], np.float32))
    with pytest.raises(capi.BwError):  # cloth edge out of range
        cl.SetEdgeEdgeCollisionInfo(np.array([[1, 56, 0]], np.int32), nrm,
                                    np.array([[0.5, 0.5]], np.float32),
                                    np.array([0.0], np.float32))
    with pytest.raises(capi.BwError):  # non-finite normal
        cl.SetEdgeEdgeCollisionInfo(np.array([[1, 3, 0]], np.int32), nrm * np.nan,
                                    np.array([[0.5, 0.5]], np.float32),
                                    np.array([0.0], np.float32))
    # valid ones are accepted and the step runs
    cl.SetPointTriangleCollisionInfo(np.array([[1, 3, 0]], np.int32), nrm, bary,
                                     np.array([0.0], np.float32))
    cl.UpdatePos()
    assert np.isfinite(cl.GetPos()).all()
