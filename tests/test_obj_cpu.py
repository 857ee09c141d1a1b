"""OBJ cloth ingest in the oracle (the library's reader is compared with it
on the GPU): positions and UV as float values, the reference's OBJ-cloth
transforms, and the reference's BuildEdge numbering, pinned against a
literal transcription of cObjUtil::BuildEdge (ObjUtil.cpp:141-224)."""
import numpy as np
import pytest

import obj_fixture as fx
import oracle_ffi as of


def _build_edge(tri):
    """cObjUtil::BuildEdge as written: a map keyed by (min, max) to
    (first triangle, second triangle), iterated in key order"""
    info = {}
    for t, tr in enumerate(tri.tolist()):
        for i in range(3):
            a, b = tr[i], tr[(i + 1) % 3]
            key = (min(a, b), max(a, b))
            if key not in info:
                info[key] = [t, -1]
            else:
                assert info[key][0] != -1 and info[key][1] == -1
                info[key][1] = t
    return np.array([[k[0], k[1], v[0], v[1], int(v[1] == -1)]
                     for k, v in sorted(info.items())], dtype=np.int32)


def test_obj_reader_takes_float_values_and_corner_forms(tmp_path):
    path = tmp_path / "sheet.obj"
    pos, uv, tri = fx.write_obj(path)
    p, u, t = of.read_obj_triangles(str(path))
    assert np.array_equal(t, tri)
    assert np.array_equal(p, pos.astype(np.float32).astype(np.float64))
    assert np.array_equal(u, uv.astype(np.float32))


def test_obj_cloth_edges_transforms_and_pins(tmp_path):
    path = tmp_path / "sheet.obj"
    pos, uv, tri = fx.write_obj(path)
    conf = fx.obj_config(path)
    oc = of.OracleCloth(conf)
    assert np.array_equal(oc.triangles(), tri)
    assert np.array_equal(oc.edges(), _build_edge(tri))
    rest = oc.rest_pos().reshape(-1, 3)
    p32 = pos.astype(np.float32).astype(np.float64)
    assert np.array_equal(rest, p32 + [0.0, 0.2, 0.0] + np.array([0.01, -0.02, 0.03]))
    assert np.array_equal(oc.uv(), (uv.astype(np.float32) * np.float32(10 / 3.0)).astype(np.float64))
    # UV targets (0, 1) and (1, 1) are the top corners of the sheet
    assert sorted(oc.pins().tolist()) == [(fx.NY - 1) * fx.NX, fx.NY * fx.NX - 1]


def test_obj_relocation_puts_top_at_027_and_centres_it(tmp_path):
    path = tmp_path / "sheet.obj"
    fx.write_obj(path)
    oc = of.OracleCloth(fx.obj_config(path, relocate=True))
    rest = oc.rest_pos().reshape(-1, 3)
    assert abs(rest[:, 1].max() - 0.27) < 1e-7
    top = (0.27 - rest[:, 1]) < 0.01
    assert top.sum() == fx.NX
    assert abs(rest[top, 0].mean()) < 1e-12 and abs(rest[top, 2].mean()) < 1e-12


def test_obj_reader_rejects_polygons(tmp_path):
    path = tmp_path / "quad.obj"
    path.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    with pytest.raises(ValueError):
        of.read_obj_triangles(str(path))
