"""Irregular meshes on the GPU, through the C ABI: Delaunay cloths
(geometry_type "delaunay") and caller-given meshes (bw_create_mesh) against
the oracle's mesh constructor on the same triangles.

Tolerances (stated per test):
  * topology, edges (first-appearance numbering), pins, ELL structure,
    rest positions: identical
  * reference-semantics solve (max_iters-1 FP64 PCG iterations): first step
    within 1e-9 of the step displacement
  * 1e-6 tolerance solve, 20 steps, top side pinned: within 1e-4 relative to
    the positions, every step's FP64 true residual <= 1e-6
  * the same mesh through bw_create and bw_create_mesh: bit-identical
"""
import copy
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


def _mode(conf, precision, mode="tolerance", tol=1e-6):
    c = copy.deepcopy(conf)
    c["precision"] = precision
    c["solver_mode"] = mode
    c["solver_tolerance"] = tol
    return c


def _case(points=300, seed=4, **kw):
    conf = wl.delaunay_config(points, seed=seed, **kw)
    pos, tri = wl.delaunay_rest_positions(conf)
    return conf, pos, tri, wl.jitter(pos, seed + 100)


@pytest.mark.parametrize("points,seed,rot", [(300, 4, 0), (1200, 5, 45)])
def test_delaunay_topology_matches_oracle(capi, points, seed, rot):
    say("case delaunay topology", points, seed)
    conf, pos, tri, x0 = _case(points, seed, cloth_uv_rotation=rot,
                               cloth_init_orientation=[0.3, 0.0, 0.1])
    cl = capi.BaraffClothHip(conf)
    oc = of.OracleCloth(conf, mesh=(pos, tri))
    assert np.array_equal(cl.triangles(), tri)
    assert np.array_equal(cl.edges(), oc.edges())
    assert np.array_equal(np.sort(cl.GetConstraintStaticVertices()), np.sort(oc.pins()))
    rp, cols = cl.ell_map()
    orp, ocols = oc.ref_ell()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    assert np.array_equal(cl.GetInitPos(), oc.rest_pos())
    width = np.diff(rp)
    say("  ELL row widths %d..%d" % (width.min(), width.max()))


def test_delaunay_reference_semantics_first_steps(capi):
    conf, pos, tri, x0 = _case(300, 4, top_edge=False)
    say("case delaunay reference-mode")
    cl = capi.BaraffClothHip(_mode(conf, "fp64", "reference"), x0)
    oc = of.OracleCloth(conf, mesh=(pos, tri))
    oc.set_state(x0, x0)
    for s in range(3):
        xs = cl.GetPos()
        cl.UpdatePos()
        oc.step(of.SOLVE_REFERENCE)
        d = np.abs(oc.pos() - xs).max()
        err = np.abs(cl.GetPos() - oc.pos()).max()
        say("  step %d err/|dx| %.3e" % (s + 1, err / d))
        assert err <= (1e-9 if s == 0 else 1e-6) * d


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("stage", ["0", "1", "1/192"])
def test_delaunay_twenty_tolerance_steps(capi, precision, stage, monkeypatch):
    """20 tolerance-mode steps against the oracle, with the SpMV's directions
    gathered per slot (stage 0) or staged in LDS per slice column (stage 1);
    1/192: LDS for 192 list entries, so that the slices with longer lists
    take the gathered product within the staged kernel."""
    monkeypatch.setenv("BW_SPMV_STAGE", stage[0])
    if "/" in stage:
        monkeypatch.setenv("BW_STAGE_LDS", stage.split("/")[1])
    conf, pos, tri, x0 = _case(1200, 5)
    say("case delaunay 20 steps", precision, "stage", stage)
    cl = capi.BaraffClothHip(_mode(conf, precision), x0)
    oc = of.OracleCloth(conf, mesh=(pos, tri))
    oc.set_state(x0, x0)
    worst = 0.0
    for _ in range(20):
        cl.UpdatePos()
        oc.step(of.SOLVE_TOLERANCE, 1e-12)
        worst = max(worst, cl.last_stats()["rel_residual"])
    x, xr = cl.GetPos(), oc.pos()
    err = np.abs(x - xr).max()
    say("  err/|x| %.3e err/disp %.3e worst residual %.2e" % (
        err / np.abs(xr).max(), err / np.abs(xr - x0).max(), worst))
    assert worst <= 1e-6
    assert err <= 1e-4 * np.abs(xr).max()


def test_mesh_entry_equals_delaunay_geometry(capi, monkeypatch):
    monkeypatch.setenv("BW_SPMV_LPR", "1")  # the two objects autotune apart
    monkeypatch.setenv("BW_SPMV_STAGE", "0")
    conf, pos, tri, x0 = _case(500, 6)
    say("case bw_create_mesh vs delaunay")
    a = capi.BaraffClothHip(conf, x0)
    mconf = dict(conf, geometry_type="mesh")
    del mconf["cloth_size"], mconf["subdivision"]
    b = capi.BaraffClothHip(mconf, x0, mesh=(pos, tri))
    assert np.array_equal(a.edges(), b.edges())
    for _ in range(3):
        a.UpdatePos()
        b.UpdatePos()
    assert np.array_equal(a.GetPos(), b.GetPos())


@pytest.mark.parametrize("stage", ["0", "1", "1/192"])
def test_delaunay_in_batch_equals_alone(capi, stage, monkeypatch):
    monkeypatch.setenv("BW_SPMV_LPR", "2")  # the two batches autotune apart
    monkeypatch.setenv("BW_SPMV_STAGE", stage[0])
    if "/" in stage:
        monkeypatch.setenv("BW_STAGE_LDS", stage.split("/")[1])
    conf, pos, tri, x0 = _case(800, 7)
    g = wl.grid_config(33)
    gx = wl.jittered_positions(g, 8)
    say("case delaunay in batch, stage", stage)
    alone = capi.ClothBatch([conf], [x0])
    alone.step(3)
    mix = capi.ClothBatch([g, conf], [gx, x0])
    mix.step(3)
    assert np.array_equal(alone.get_pos(0), mix.get_pos(1))


@pytest.mark.parametrize("kind", ["grid", "delaunay"])
def test_normals_match_oracle(capi, kind, monkeypatch):
    """Normals read back after 4 steps (cBaseCloth::Update's refresh) equal
    the oracle's at the same positions to 1e-14, and a batch member's equal
    the single cloth's (one SpMV variant fixed, as the two autotune apart)."""
    monkeypatch.setenv("BW_SPMV_LPR", "1")
    monkeypatch.setenv("BW_SPMV_STAGE", "0")
    say("case normals", kind)
    if kind == "grid":
        conf = wl.grid_config(21)
        x0 = wl.jittered_positions(conf, 3)
        oc = of.OracleCloth(conf)
    else:
        conf, pos, tri, x0 = _case(700, 9)
        oc = of.OracleCloth(conf, mesh=(pos, tri))
    cl = capi.BaraffClothHip(conf, x0)
    for _ in range(4):
        cl.UpdatePos()
    x = cl.GetPos()
    oc.set_state(x, x)
    vn, tn = cl.normals()
    ovn, otn = oc.normals()
    say("  max diff vertex %.2e triangle %.2e" % (np.abs(vn - ovn).max(), np.abs(tn - otn).max()))
    assert np.abs(tn - otn).max() <= 1e-14
    assert np.abs(vn - ovn).max() <= 1e-14
    b = capi.ClothBatch([conf], [x0])
    b.step(4)
    bvn, btn = b.normals(0)
    assert np.array_equal(bvn, vn) and np.array_equal(btn, tn)


@pytest.mark.parametrize("relocate", [False, True])
def test_obj_cloth_matches_oracle(capi, tmp_path, relocate):
    """enable_cloth_from_obj: the library's OBJ reader and the reference's
    OBJ-cloth transforms and BuildEdge numbering against the oracle's
    restatement; 5 tolerance steps within 1e-4 of the displacement."""
    import obj_fixture as fx

    say("case OBJ cloth, relocate", relocate)
    path = tmp_path / "sheet.obj"
    fx.write_obj(path)
    conf = fx.obj_config(path, relocate=relocate)
    oc = of.OracleCloth(conf)
    cl = capi.BaraffClothHip(conf)
    assert np.array_equal(cl.triangles(), oc.triangles())
    assert np.array_equal(cl.edges(), oc.edges())
    assert np.array_equal(np.sort(cl.GetConstraintStaticVertices()), np.sort(oc.pins()))
    rp, cols = cl.ell_map()
    orp, ocols = oc.ref_ell()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    x0 = cl.GetInitPos()
    assert np.array_equal(x0, oc.rest_pos())
    oc.set_state(x0, x0)
    for _ in range(5):
        cl.UpdatePos()
        oc.step(of.SOLVE_TOLERANCE, 1e-12)
        assert cl.last_stats()["rel_residual"] <= 1e-6
    d = np.abs(oc.pos() - x0).max()
    err = np.abs(cl.GetPos() - oc.pos()).max()
    say("  err/disp %.3e" % (err / d))
    assert err <= 1e-4 * d


def test_obj_errors_are_reported(capi, tmp_path):
    import obj_fixture as fx

    say("case OBJ errors")
    quad = tmp_path / "quad.obj"
    quad.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    for path in (quad, tmp_path / "missing.obj"):
        with pytest.raises(capi.BwError) as e:
            capi.BaraffClothHip(fx.obj_config(path))
        assert e.value.code == -1
