"""Generate the golden fixtures from the reference's own CPU cloth class.

Runs oracle/_ref/ref_cloth (the reference's cBaraffCloth compiled from its
sources by oracle/Makefile.ref, driven by oracle/ref_harness.cpp) on seeded
inputs and stores inputs and outputs as small .npz fixtures in this folder.
Run in the build container, where the reference tree exists:

    make -f oracle/Makefile.ref -j8
    python tests/golden/make_golden.py            # all cases
    python tests/golden/make_golden.py grid21     # one case

The fixtures are data: configuration, initial positions, pins, and the
reference's positions after given steps (and, for small meshes, its
per-stage K, forces and dx of the first step).
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from acousticsim_amd import workloads  # noqa: E402
from refdump import read_dump  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_cloth")


def case_defs():
    cases = {}
    c = workloads.grid_config(3, dt=0.01)
    cases["tiny3"] = dict(conf=c, seed=11, steps=5, dump=[1, 2, 5], stages=True, full_k=True)

    c = workloads.grid_config(5, dt=0.005)
    c.update({"cloth_uv_rotation": 45, "rayleigh_damping_a": 0.1, "rayleigh_damping_b": -2e-3,
              "cloth_stretch_stiffness": [500, 800, 50], "cloth_bending_stiffness": [3e-4, 1e-4],
              "cloth_density": 0.3, "max_iters": 200})
    cases["grid5_rot45"] = dict(conf=c, seed=12, steps=5, dump=[1, 5], stages=True, full_k=True,
                                qbend=True)

    c = workloads.grid_config(7)
    c.update({"subdivision": [6, 11], "cloth_size": [0.3, 0.5], "cloth_init_pos": [0, 0.4, 0],
              "constraints": {"static_point": [[-0.25, 0.15], [0.25, 0.15], [0.0, 0.15]]},
              "cloth_uv_rotation": 90, "max_iters": 300})
    cases["rect7x12"] = dict(conf=c, seed=13, steps=8, dump=[1, 8], stages=True, full_k=True,
                             qbend=True, normals=True)

    c = workloads.grid_config(21, dt=0.01)
    cases["grid21"] = dict(conf=c, seed=workloads.SINGLE_SMALL["seed"], steps=100,
                           dump=[1, 10, 50, 100], stages=True, full_k=False, qbend=True,
                           normals=True)

    # contacts handed to the cloth before every step (collision Hessian in W,
    # repulsion impulses on dx); stiffness values whose float mean depends on
    # the summation order
    c = workloads.grid_config(21, dt=0.01)
    c.update({"cloth_stretch_stiffness": [440.6, 974.4, 898.7]})
    cases["grid21_contacts"] = dict(conf=c, seed=21, steps=30, dump=[1, 10, 30], stages=True,
                                    full_k=False, contacts=dict(seed=5))
    c = workloads.grid_config(5, dt=0.005)
    c.update({"rayleigh_damping_a": 0.1, "max_iters": 200})
    cases["grid5_contacts"] = dict(conf=c, seed=22, steps=4, dump=[1, 2, 4], stages=True,
                                   full_k=True, contacts=dict(seed=6, n_point=10, n_tri=6,
                                                              n_edge=8, repeat=3))

    # user inputs before given steps: forces split over a triangle's vertices
    # (cBaseCloth::ApplyPerturb), a dragged vertex moved to a goal and fixed
    # for the step (cBaraffCloth::ApplyUserPerturbForceOnce), and the y drag
    # force on the top and bottom UV edges every step (mYDragForce)
    c = workloads.grid_config(21, dt=0.01)
    cases["grid21_perturb"] = dict(conf=c, seed=23, steps=12, dump=[1, 2, 3, 5, 8, 12],
                                   stages=True, full_k=False, perturb=perturb_records(c),
                                   ydrag=0.35)
    c = workloads.grid_config(5, dt=0.005)
    c.update({"rayleigh_damping_a": 0.1, "max_iters": 200, "cloth_uv_rotation": 45})
    cases["grid5_perturb"] = dict(conf=c, seed=24, steps=4, dump=[1, 2, 3, 4], stages=True,
                                  full_k=True, perturb=[(1, 0, 3, (0.01, -0.02, 0.004)),
                                                        (1, 1, 5, (0.0, 0.0, 0.003)),
                                                        (2, 0, 5, (0.0, 0.0, 0.02)),
                                                        (2, 0, 5, (0.001, 0.0, 0.0)),
                                                        (3, 1, 0, (-0.125, -0.125, -0.002))],
                                  ydrag=0.0)

    # an OBJ cloth of jittered vertices (triangle areas from 0.3x to 2x the
    # mean), so that area-weighted vertex normals differ from unweighted ones
    cases["obj_irregular"] = dict(conf=None, seed=None, steps=6, dump=[1, 6], stages=True,
                                  full_k=False, normals=True, obj=dict(seed=7))

    c = workloads.grid_config(257, dt=0.005)
    cases["grid257"] = dict(conf=c, seed=workloads.SINGLE_LARGE["seed"], steps=5,
                            dump=[1, 5], stages=False, full_k=False, stride=17)
    return cases


def perturb_records(conf):
    """[(step, kind, triangle, vector)] for the 21 x 21 cloth: kind 0 is a
    force on the triangle (ApplyPerturb), kind 1 the goal of the triangle's
    first vertex (ApplyUserPerturbForceOnce); goals are set relative to the
    vertex's rest position"""
    rest = workloads.grid_rest_positions(conf).reshape(-1, 3)
    hd, wd = conf["subdivision"]
    # first vertex of triangle t is the cell's lower-left or upper-left corner;
    # the harness reads the actual vertex from the reference's triangle array
    # and these goals only need to be near the cloth
    goal_a = rest[(wd // 2) * (hd + 1) + hd // 2] + (0.0, 0.0, 0.02)
    goal_b = rest[(wd // 3) * (hd + 1) + hd // 4] + (0.01, -0.01, -0.015)
    return [(1, 0, 100, (0.02, -0.01, 0.05)), (1, 0, 101, (0.02, -0.01, 0.05)),
            (1, 0, 100, (0.0, 0.0, -0.01)),
            (2, 1, 400, tuple(goal_a)), (3, 1, 400, tuple(goal_a + (0.0, 0.0, 0.005))),
            (5, 0, 250, (0.0, 0.1, 0.0)), (5, 1, 150, tuple(goal_b)),
            (8, 1, 799, tuple(rest[-1] + (0.0, 0.0, 0.01)))]


def irregular_sheet(seed):
    """(positions, uv, triangles) of the OBJ fixture sheet with its interior
    vertices jittered in the plane (positions and texture coordinates alike)"""
    import obj_fixture as ofx

    rng = np.random.default_rng(seed)
    pos, uv, tri = ofx.sheet()
    inner = [k for k in range(len(pos))
             if 0 < k % ofx.NX < ofx.NX - 1 and 0 < k // ofx.NX < ofx.NY - 1]
    d = rng.uniform(-0.4, 0.4, (len(inner), 2))
    pos[inner, 0] += d[:, 0] * 0.05
    pos[inner, 1] += d[:, 1] * 0.04
    uv[inner, 0] += d[:, 0] * 0.05
    uv[inner, 1] += d[:, 1] * 0.05
    return pos, uv, tri


def write_sheet_obj(path, pos, uv, tri):
    lines = ["o sheet"] + ["v %.17g %.17g %.17g" % tuple(p) for p in pos]
    lines += ["vt %.17g %.17g" % tuple(t) for t in uv]
    lines += ["f " + " ".join("%d/%d" % (v + 1, v + 1) for v in t) for t in tri]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def initial_positions(conf, seed):
    x = workloads.jittered_positions(conf, seed).reshape(-1, 3)
    x += np.array((list(conf.get("cloth_init_pos", [0, 0, 0])) + [0, 0, 0])[:3])
    return x.reshape(-1)


def run_case(name, d):
    conf = d["conf"]
    with tempfile.TemporaryDirectory() as td:
        sheet = None
        if "obj" in d:
            import obj_fixture as ofx

            sheet = irregular_sheet(d["obj"]["seed"])
            objpath = os.path.join(td, "sheet.obj")
            write_sheet_obj(objpath, *sheet)
            conf = ofx.obj_config(objpath)
            x0 = None
        else:
            x0 = initial_positions(conf, d["seed"])
        cpath = os.path.join(td, "conf.json")
        xpath = os.path.join(td, "x0.bin")
        opath = os.path.join(td, "out.bin")
        with open(cpath, "w") as f:
            json.dump(conf, f)
        if x0 is not None:
            x0.astype("<f8").tofile(xpath)
        cmd = [REF_BIN, cpath, xpath if x0 is not None else "-", str(d["steps"]), opath,
               "--dump-steps", ",".join(str(s) for s in d["dump"]), "--time"]
        if d["stages"]:
            cmd.append("--stages")
        if d.get("qbend"):
            cmd.append("--qbend")
        if d.get("normals"):
            cmd.append("--normals")
        ct = None
        if "contacts" in d:
            side = conf["subdivision"]
            nv = (side[0] + 1) * (side[1] + 1)
            nt = 2 * side[0] * side[1]
            ne = side[0] * (side[1] + 1) + side[1] * (side[0] + 1) + side[0] * side[1]
            ct = workloads.random_contacts(nv, nt, ne, **d["contacts"])
            ctpath = os.path.join(td, "contacts.bin")
            workloads.contacts_file_records(ct).astype("<f8").tofile(ctpath)
            cmd += ["--contacts", ctpath]
        if "perturb" in d:
            ppath = os.path.join(td, "perturb.bin")
            np.array([[st, kind, tri, v[0], v[1], v[2]] for st, kind, tri, v in d["perturb"]],
                     dtype="<f8").tofile(ppath)
            cmd += ["--perturb", ppath, "--ydrag", repr(d["ydrag"])]
        subprocess.run(cmd, check=True)
        dump = read_dump(opath)
    xi = dump["x_init"]
    if x0 is None:
        x0 = dump["x0"]
    else:
        # the injected x0 must be a jitter of the reference's own rest positions
        assert np.max(np.abs((x0 - xi).reshape(-1, 3)[:, :2])) == 0.0, name
    out = {"config_json": np.array(json.dumps(conf)),
           "seed": np.array(d["seed"] if d["seed"] is not None else -1),
           "steps": np.array(d["dump"], dtype=np.int32)}
    if sheet is not None:
        # the OBJ's content (the test writes it again; the config's path is
        # a placeholder)
        out["obj_pos"], out["obj_uv"], out["obj_tri"] = sheet
    stride = d.get("stride", 1)
    sel = np.arange(0, len(xi) // 3, stride)
    out["vertex_subset"] = sel.astype(np.int32)
    if stride == 1:
        out["x0"] = x0
    else:
        # large meshes: x0 is regenerated from (config, seed); keep the
        # subset as a check that the regeneration matches
        out["x0_subset"] = x0.reshape(-1, 3)[sel].reshape(-1)
    if ct is not None:
        for k, v in ct.items():
            out["contacts_" + k] = v
    if "perturb" in d:
        out["perturb"] = np.array([[st, kind, tri, v[0], v[1], v[2]]
                                   for st, kind, tri, v in d["perturb"]], dtype=np.float64)
        out["ydrag"] = np.array(d["ydrag"], dtype=np.float32)
    out["pins"] = dump["pins"]
    out["mass"] = dump["mass_diag"][0::3][sel]
    for s in d["dump"]:
        x = dump["x_step%d" % s].reshape(-1, 3)
        out["x_step%d" % s] = x[sel].reshape(-1)
        out["norm_step%d" % s] = np.array(np.linalg.norm(x))
        out["sum_step%d" % s] = x.sum(axis=0)
    out["step_seconds"] = dump["step_seconds"]
    if d.get("qbend"):
        # the reference's quadratic bending matrix at unit stiffness
        # (assembled COO and per-interior-edge 12x12 element matrices)
        for k in ("qbend_K_row", "qbend_K_col", "qbend_K_val", "qbend_ele", "qbend_edges"):
            out[k] = dump[k]
    if d.get("normals"):
        # normals after the last step (cBaseCloth::Update's refresh)
        out["vertex_normals"] = dump["vertex_normals"]
        out["triangle_normals"] = dump["triangle_normals"]
    if d["stages"]:
        out["triangles"] = dump["triangles"]
        out["edges"] = dump["edges"]
        out["stencils"] = dump["stencils"]
        out["uv"] = dump["uv"]
        for k in ("s1_fint", "s1_dx", "s1_f_stretch", "s1_f_bend"):
            out[k] = dump[k]
        if d["full_k"]:
            for k in ("s1_K", "s1_K_stretch", "s1_K_bend"):
                for part in ("row", "col", "val"):
                    out[k + "_" + part] = dump[k + "_" + part]
        else:
            # the block pattern only (rows/cols of the scalar CSR, one per block)
            r, c = dump["s1_K_row"], dump["s1_K_col"]
            keep = (r % 3 == 0) & (c % 3 == 0)
            out["s1_K_block_row"] = (r[keep] // 3).astype(np.int32)
            out["s1_K_block_col"] = (c[keep] // 3).astype(np.int32)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("%-12s V=%d steps=%d  ref %.3f s/step  -> %s (%d KB)" % (
        name, len(xi) // 3, d["steps"], float(np.mean(dump["step_seconds"])), path,
        os.path.getsize(path) // 1024))


def main():
    if not os.path.exists(REF_BIN):
        sys.exit("build the reference first: make -f oracle/Makefile.ref -j8")
    cases = case_defs()
    names = sys.argv[1:] or list(cases)
    for n in names:
        run_case(n, cases[n])


if __name__ == "__main__":
    main()
