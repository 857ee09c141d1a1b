"""A small OBJ cloth written on the fly for the OBJ-ingest tests: a 7 x 6
vertex sheet in the x-y plane with a gentle z wave, texture coordinates in
[0, 0.3] (the reference scales them by 10/3 to metres), faces written with
every corner form tinyobjloader accepts (i, i/t, i//n, i/t/n, negative
relative indices), comments, normals and one object name."""
import numpy as np

NX, NY = 7, 6


def sheet():
    """(positions (nv, 3), uv (nv, 2), triangles (nt, 3)) as written"""
    pos, uv = [], []
    for j in range(NY):
        for i in range(NX):
            pos.append([0.05 * i - 0.15, 0.04 * j + 0.013, 0.002 * np.sin(i + 2 * j)])
            uv.append([0.3 * i / (NX - 1), 0.3 * j / (NY - 1)])
    tri = []
    for j in range(NY - 1):
        for i in range(NX - 1):
            a, b, c, d = j * NX + i, j * NX + i + 1, (j + 1) * NX + i, (j + 1) * NX + i + 1
            if (i + j) % 2:
                tri += [[a, b, d], [a, d, c]]
            else:
                tri += [[a, b, c], [b, d, c]]
    return np.array(pos), np.array(uv), np.array(tri, dtype=np.int32)


def write_obj(path):
    pos, uv, tri = sheet()
    lines = ["# cloth sheet", "o sheet"]
    for p in pos:
        lines.append("v %.9g %.9g %.9g" % tuple(p))
    for t in uv:
        lines.append("vt %.9g %.9g" % tuple(t))
    lines.append("vn 0 0 1")
    nv = len(pos)
    for k, t in enumerate(tri):
        form = k % 5
        if form == 0:
            corners = ["%d/%d" % (v + 1, v + 1) for v in t]
        elif form == 1:
            corners = ["%d/%d/1" % (v + 1, v + 1) for v in t]
        elif form == 2:
            corners = ["%d/%d" % (v - nv, v - nv) for v in t]  # relative
        elif form == 3:
            corners = ["%d/%d/1" % (v + 1, v + 1) for v in t]
        else:
            corners = ["%d/%d" % (v + 1, v + 1) for v in t]
        lines.append("f " + " ".join(corners))
    open(path, "w").write("\n".join(lines) + "\n")
    return pos, uv, tri


def obj_config(path, relocate=False, **over):
    from acousticsim_amd import workloads as wl

    c = wl.grid_config(9)
    del c["cloth_size"], c["subdivision"]
    c["enable_cloth_from_obj"] = True
    c["cloth_obj_path"] = str(path)
    c["enable_relocate_obj"] = relocate
    c["cloth_obj_translation"] = [0.01, -0.02, 0.03]
    c["cloth_obj_size"] = [0.3, 0.2]
    # UV targets (metres after the 10/3 scale): the two top corners
    c["constraints"] = {"static_point": [[0.0, 1.0], [1.0, 1.0]]}
    c.update(over)
    return c
