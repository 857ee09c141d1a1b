"""Oracle pieces pinned by dumps of the compiled reference
(tests/golden/make_golden.py runs oracle/_ref/ref_cloth with --qbend and
--normals):

  * the reference GPU form's constant bending matrix
    (cBaraffClothGpu::InitBendingMatrix / UpdateBendingMatrix,
    BaraffClothGpu.cpp:573-680): per interior edge the quadratic-bending
    element of cQBendingMaterial (QBendingMaterial.cpp:47-118, dumped at
    unit stiffness: cotangent coefficients times 3 / (A0 + A1)) scaled by
    K_e = Ku cos^2(theta_e) + Kv sin^2(theta_e), theta_e the angle of the
    edge's UV direction from the u axis folded into [0, pi/2]; the oracle's
    matrix against that sum with warp and weft stiffness different, theta
    taken from the reference's own UV, and on the unrotated 21 x 21 grid
    against the edge classes' known angles (u edges 0, v edges pi/2,
    diagonals pi/4), so a wrong prefactor or swapped axes fail;
  * vertex normals (cBaseObject::UpdateVertexNormalFromTriangleNormal,
    BaseObject.cpp:138-160): the sum of incident triangle normals weighted
    by the triangles' initial areas, normalised; on an OBJ cloth of jittered
    vertices (triangle areas 0.3x..2x the mean) after 6 steps, where
    unweighted or current-area weighting differ from the reference's dump.
"""
import json

import numpy as np
import pytest

import oracle_ffi as of
from conftest import load_golden


def _qbend_expect(oc, g, ku, kv, theta):
    """per pattern block: sum over interior edges of K_e(theta) * element"""
    rp, cols = oc.pattern()
    out = np.zeros(oc.nnzb)
    ele = g["qbend_ele"].reshape(-1, 12, 12)
    st = oc.stencils()
    for k, e in enumerate(g["qbend_edges"]):
        ke = ku * np.cos(theta[e]) ** 2 + kv * np.sin(theta[e]) ** 2
        s4 = st[e]
        for i in range(4):
            for j in range(4):
                r, c = s4[i], s4[j]
                slot = rp[r] + np.searchsorted(cols[rp[r]:rp[r + 1]], c)
                out[slot] += ke * ele[k][3 * i, 3 * j]
    return out


def _uv_theta(g):
    uv = g["uv"].reshape(-1, 2).astype(np.float64)
    e = g["edges"].reshape(-1, 5)
    d = uv[e[:, 1]] - uv[e[:, 0]]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    th = np.arccos(np.clip(d[:, 0], -1.0, 1.0))
    return np.where(th > np.pi / 2, np.pi - th, th)


def test_qbend_elements_assemble_to_the_reference_matrix():
    """the dumped elements (edge order, float) sum to the dumped assembled
    matrix (double triplets), which GetStiffnessMatrix returns negated
    (BaseBendingMaterial.cpp: the material's K is minus the energy
    Hessian)"""
    g = load_golden("grid21")
    oc = of.OracleCloth(g["config"])
    n = 3 * oc.nv
    K = np.zeros((n, n))
    K[g["qbend_K_row"], g["qbend_K_col"]] = g["qbend_K_val"]
    S = np.zeros((n, n))
    ele = g["qbend_ele"].reshape(-1, 12, 12).astype(np.float64)
    st = oc.stencils()
    for k, e in enumerate(g["qbend_edges"]):
        idx = np.concatenate([3 * st[e][i] + np.arange(3) for i in range(4)])
        S[np.ix_(idx, idx)] += ele[k]
    assert np.abs(S + K).max() <= 1e-6 * np.abs(K).max()


@pytest.mark.parametrize("name", ["grid21", "rect7x12", "grid5_rot45"])
def test_gpu_form_bending_matches_reference_elements(name):
    g = load_golden(name)
    ku, kv = 3.0e-4, 0.7e-4
    conf = dict(g["config"], cloth_bending_stiffness=[ku, kv])
    oc = of.OracleCloth(conf)
    mine = oc.gpu_form_bending()
    ref = _qbend_expect(oc, g, float(np.float32(ku)), float(np.float32(kv)), _uv_theta(g))
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.abs(mine - ref).max() <= 2e-6 * scale, name
    # the same with the axes swapped is far off (the test sees the axis rule)
    swapped = _qbend_expect(oc, g, float(np.float32(kv)), float(np.float32(ku)), _uv_theta(g))
    if name != "grid5_rot45":  # at 45 degrees both axis classes sit at pi/4
        assert np.abs(mine - swapped).max() > 0.1 * scale


def test_gpu_form_bending_edge_classes_on_the_unrotated_grid():
    """21 x 21 grid, UV = rest xy: vertex i*(hd+1)+j sits at column i, row
    j, so an edge between ids differing by 1 runs along v (theta = pi/2,
    K = Kv), by hd+1 along u (theta = 0, K = Ku), otherwise it is a cell
    diagonal of a square cell (theta = pi/4, K = (Ku + Kv)/2)"""
    g = load_golden("grid21")
    ku, kv = 3.0e-4, 0.7e-4
    conf = dict(g["config"], cloth_bending_stiffness=[ku, kv])
    oc = of.OracleCloth(conf)
    hd = conf["subdivision"][0]
    e = g["edges"].reshape(-1, 5)
    diff = np.abs(e[:, 1] - e[:, 0])
    theta = np.where(diff == 1, np.pi / 2, np.where(diff == hd + 1, 0.0, np.pi / 4))
    ref = _qbend_expect(oc, g, float(np.float32(ku)), float(np.float32(kv)), theta)
    mine = oc.gpu_form_bending()
    assert np.abs(mine - ref).max() <= 2e-6 * np.abs(ref).max()


def _obj_oracle(g, tmp_path):
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import write_sheet_obj

    path = tmp_path / "sheet.obj"
    write_sheet_obj(str(path), g["obj_pos"], g["obj_uv"], g["obj_tri"])
    conf = json.loads(str(g["config_json"]))
    conf["cloth_obj_path"] = str(path)
    return of.OracleCloth(conf)


def test_obj_cloth_trajectory_matches_reference(tmp_path):
    g = load_golden("obj_irregular")
    oc = _obj_oracle(g, tmp_path)
    assert np.array_equal(oc.triangles().ravel(), g["triangles"])
    assert np.array_equal(oc.edges().ravel(), g["edges"])
    oc.set_state(g["x0"], g["x0"])
    for s in range(1, 7):
        oc.step(of.SOLVE_REFERENCE)
        if s == 1:
            assert np.array_equal(oc.dx(), g["s1_dx"])
        if s in (1, 6):
            assert np.array_equal(oc.pos(), g["x_step%d" % s]), s


def test_vertex_normals_weight_by_initial_area(tmp_path):
    g = load_golden("obj_irregular")
    oc = _obj_oracle(g, tmp_path)
    x = g["x_step6"]
    oc.set_state(x, x)
    vn, tn = oc.normals()
    rvn = g["vertex_normals"].reshape(-1, 3)
    rtn = g["triangle_normals"].reshape(-1, 3)
    assert np.abs(tn - rtn).max() <= 1e-15
    assert np.abs(vn - rvn).max() <= 1e-15
    # the alternatives the reference's dump rules out
    tri = g["triangles"].reshape(-1, 3)
    xv = x.reshape(-1, 3)
    cur = 0.5 * np.linalg.norm(np.cross(xv[tri[:, 1]] - xv[tri[:, 0]],
                                        xv[tri[:, 2]] - xv[tri[:, 0]]), axis=1)
    for w in (np.ones(len(tri)), cur):
        alt = np.zeros_like(rvn)
        for k in range(3):
            np.add.at(alt, tri[:, k], rtn * w[:, None])
        alt /= np.linalg.norm(alt, axis=1, keepdims=True)
        assert np.abs(alt - rvn).max() > 1e-6
