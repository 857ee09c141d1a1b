"""Python binding of the C ABI (include/bw_cloth.h) in libbwcloth.so.

`BaraffClothHip` mirrors the reference's cloth-object interface
(cBaseCloth, src/sim/cloth/BaseCloth.h:19-80, as implemented by cBaraffCloth):
Init from the cloth JSON, UpdatePos, GetPos / SetPos / SetPrePos, Reset,
constraint vertices, plus readback of mesh, ELL structure and step counters.
`ClothBatch` steps many cloths together.  The library is loaded from the
package folder; if it is missing the import of this module fails loudly
(there is no fallback path).
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbwcloth.so")

BW_OK = 0


class BwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("bw error %d: %s" % (code, msg))
        self.code = code


class StepStats(ctypes.Structure):
    _fields_ = [
        ("pcg_iterations", ctypes.c_int),
        ("outer_iterations", ctypes.c_int),
        ("rel_residual", ctypes.c_double),
        ("step_us", ctypes.c_double),
        ("matrix_fp64", ctypes.c_int),
        ("reserved", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("done", ctypes.c_int),
        ("num_vertices", ctypes.c_int),
        ("steps", ctypes.c_int),
        ("steps_done", ctypes.c_int),
        ("pcg_iterations", ctypes.c_longlong),
        ("max_rel_residual", ctypes.c_double),
        ("t_submit", ctypes.c_double),
        ("t_admit", ctypes.c_double),
        ("t_first_step", ctypes.c_double),
        ("t_done", ctypes.c_double),
        ("device_bytes", ctypes.c_double),
        ("admitted_bytes", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerCounters(ctypes.Structure):
    _fields_ = [
        ("queued", ctypes.c_int),
        ("resident", ctypes.c_int),
        ("completed", ctypes.c_int),
        ("max_resident_seen", ctypes.c_int),
        ("bytes_in_use", ctypes.c_double),
        ("chunks", ctypes.c_longlong),
        ("launches", ctypes.c_longlong),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
EXPORTS = [
    "bw_create", "bw_create_mesh", "bw_delaunay_mesh", "bw_destroy", "bw_update_pos", "bw_reset", "bw_get_pos", "bw_get_pre_pos",
    "bw_get_velocity", "bw_set_pos", "bw_set_pre_pos", "bw_get_init_pos",
    "bw_get_default_timestep", "bw_num_vertices", "bw_num_triangles", "bw_num_edges",
    "bw_get_constraint_static_vertices", "bw_add_constraint_static_vertices",
    "bw_clear_constraint_static_vertices", "bw_get_triangles", "bw_get_edges",
    "bw_get_ell_map", "bw_get_normals", "bw_batch_get_normals",
    "bw_apply_perturb", "bw_apply_user_perturb_force_once", "bw_set_y_drag_force",
    "bw_batch_apply_perturb", "bw_batch_apply_user_perturb_force_once",
    "bw_batch_set_y_drag_force", "bw_assemble_system", "bw_get_last_stats", "bw_last_error", "bw_batch_create",
    "bw_batch_destroy", "bw_batch_size", "bw_batch_step", "bw_batch_step_noalloc", "bw_batch_num_vertices",
    "bw_batch_num_triangles", "bw_batch_num_blocks", "bw_set_point_triangle_contacts", "bw_set_edge_edge_contacts",
    "bw_batch_set_point_triangle_contacts", "bw_batch_set_edge_edge_contacts",
    "bw_get_point_triangle_contacts", "bw_get_edge_edge_contacts", "bw_body_create",
    "bw_body_destroy", "bw_body_set_pos", "bw_body_num_edges", "bw_body_get_edges",
    "bw_detect_contacts", "bw_batch_detect_contacts",
    "bw_batch_get_pos", "bw_batch_set_state", "bw_batch_get_stats", "bw_batch_set_graphs",
    "bw_batch_time_solver_kernel", "bw_batch_time_launch_floor", "bw_batch_autotune", "bw_server_create", "bw_server_destroy", "bw_server_submit",
    "bw_server_wait", "bw_server_get_info", "bw_server_get_pos", "bw_server_release",
    "bw_server_get_counters", "bw_server_now",
]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libbwcloth.so is not built (run python -m acousticsim_amd.build): "
                          + LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp, dp, ip, cp = ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_int), ctypes.c_char_p
    i32 = ctypes.c_int
    fp = P(ctypes.c_float)
    sig = {
        "bw_create": (i32, [cp, dp, P(vp)]),
        "bw_create_mesh": (i32, [cp, dp, i32, ip, i32, dp, P(vp)]),
        "bw_delaunay_mesh": (i32, [ctypes.c_double, ctypes.c_double, i32, ctypes.c_ulonglong,
                                   ctypes.c_double, dp, i32, ip, i32, ip, ip]),
        "bw_destroy": (None, [vp]),
        "bw_update_pos": (i32, [vp, ctypes.c_double]),
        "bw_reset": (i32, [vp]),
        "bw_get_pos": (i32, [vp, dp]),
        "bw_get_pre_pos": (i32, [vp, dp]),
        "bw_get_velocity": (i32, [vp, dp]),
        "bw_set_pos": (i32, [vp, dp]),
        "bw_set_pre_pos": (i32, [vp, dp]),
        "bw_get_init_pos": (i32, [vp, dp]),
        "bw_get_default_timestep": (ctypes.c_double, [vp]),
        "bw_num_vertices": (i32, [vp]),
        "bw_num_triangles": (i32, [vp]),
        "bw_num_edges": (i32, [vp]),
        "bw_get_constraint_static_vertices": (i32, [vp, ip, i32]),
        "bw_add_constraint_static_vertices": (i32, [vp, ip, i32]),
        "bw_clear_constraint_static_vertices": (i32, [vp]),
        "bw_get_triangles": (i32, [vp, ip]),
        "bw_get_edges": (i32, [vp, ip]),
        "bw_get_ell_map": (i32, [vp, ip, ip]),
        "bw_get_normals": (i32, [vp, dp, dp]),
        "bw_apply_perturb": (i32, [vp, i32, dp]),
        "bw_apply_user_perturb_force_once": (i32, [vp, i32, dp]),
        "bw_set_y_drag_force": (i32, [vp, ctypes.c_float]),
        "bw_batch_apply_perturb": (i32, [vp, i32, i32, dp]),
        "bw_batch_apply_user_perturb_force_once": (i32, [vp, i32, i32, dp]),
        "bw_batch_set_y_drag_force": (i32, [vp, i32, ctypes.c_float]),
        "bw_batch_get_normals": (i32, [vp, i32, dp, dp]),
        "bw_assemble_system": (i32, [vp, dp, dp]),
        "bw_get_last_stats": (i32, [vp, P(StepStats)]),
        "bw_last_error": (cp, []),
        "bw_batch_create": (i32, [P(cp), P(dp), i32, P(vp)]),
        "bw_batch_destroy": (None, [vp]),
        "bw_batch_size": (i32, [vp]),
        "bw_batch_step": (i32, [vp, i32]),
        "bw_batch_step_noalloc": (i32, [vp, i32]),
        "bw_batch_num_vertices": (i32, [vp, i32]),
        "bw_batch_num_triangles": (i32, [vp, i32]),
        "bw_batch_num_blocks": (i32, [vp, i32]),
        "bw_set_point_triangle_contacts": (i32, [vp, i32, ip, dp, dp, fp]),
        "bw_set_edge_edge_contacts": (i32, [vp, i32, ip, dp, fp, fp]),
        "bw_batch_set_point_triangle_contacts": (i32, [vp, i32, i32, ip, dp, dp, fp]),
        "bw_batch_set_edge_edge_contacts": (i32, [vp, i32, i32, ip, dp, fp, fp]),
        "bw_get_point_triangle_contacts": (i32, [vp, i32, ip, dp, dp, fp]),
        "bw_get_edge_edge_contacts": (i32, [vp, i32, ip, dp, fp, fp]),
        "bw_body_create": (i32, [dp, i32, ip, i32, P(vp)]),
        "bw_body_destroy": (None, [vp]),
        "bw_body_set_pos": (i32, [vp, dp]),
        "bw_body_num_edges": (i32, [vp]),
        "bw_body_get_edges": (i32, [vp, ip]),
        "bw_detect_contacts": (i32, [vp, vp, i32, ip, ip]),
        "bw_batch_detect_contacts": (i32, [vp, i32, vp, i32, ip, ip]),
        "bw_batch_get_pos": (i32, [vp, i32, dp]),
        "bw_batch_set_state": (i32, [vp, i32, dp, dp]),
        "bw_batch_get_stats": (i32, [vp, i32, P(StepStats)]),
        "bw_batch_set_graphs": (i32, [vp, i32]),
        "bw_batch_time_solver_kernel": (i32, [vp, i32, dp, dp]),
        "bw_batch_time_launch_floor": (i32, [vp, i32, dp]),
        "bw_batch_autotune": (i32, [vp, ip, ip, ip, dp]),
        "bw_server_create": (i32, [ctypes.c_double, i32, cp, P(vp)]),
        "bw_server_destroy": (None, [vp]),
        "bw_server_submit": (i32, [vp, cp, dp, i32, i32, P(ctypes.c_longlong)]),
        "bw_server_wait": (i32, [vp, ctypes.c_longlong, ctypes.c_double]),
        "bw_server_get_info": (i32, [vp, ctypes.c_longlong, P(RequestInfo)]),
        "bw_server_get_pos": (i32, [vp, ctypes.c_longlong, dp]),
        "bw_server_release": (i32, [vp, ctypes.c_longlong]),
        "bw_server_get_counters": (i32, [vp, P(ServerCounters)]),
        "bw_server_now": (ctypes.c_double, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _contact_arrays(ct):
    """contiguous arrays of a contact dict (keys as
    acousticsim_amd.workloads.random_contacts returns)"""
    def a(key, dtype):
        return np.ascontiguousarray(ct.get(key, np.zeros(0)), dtype=dtype).reshape(-1)
    return {"pt_ids": a("pt_ids", np.int32), "pt_normal": a("pt_normal", np.float64),
            "pt_bary": a("pt_bary", np.float64), "pt_depth": a("pt_depth", np.float32),
            "ee_ids": a("ee_ids", np.int32), "ee_normal": a("ee_normal", np.float64),
            "ee_bary": a("ee_bary", np.float32), "ee_depth": a("ee_depth", np.float32)}


def _check(rc):
    if rc < 0:
        raise BwError(rc, lib().bw_last_error().decode())
    return rc


def _json(conf):
    return (conf if isinstance(conf, str) else json.dumps(conf)).encode()


def delaunay_mesh(cloth_size, n, seed=0, min_spacing=0.0):
    """The mesh of geometry_type "delaunay" (cloth_size = [height, width],
    n points, delaunay_seed, delaunay_min_spacing), before the init
    transform: rest positions (n, 3) and triangles (T, 3) in the cloth's
    winding.  Host only."""
    nv, nt = ctypes.c_int(), ctypes.c_int()
    args = (float(cloth_size[0]), float(cloth_size[1]), int(n), int(seed), float(min_spacing))
    cap = max(int(n), 4)
    pos = np.zeros(3 * cap, dtype=np.float64)
    tri = np.zeros(6 * cap, dtype=np.int32)
    rc = lib().bw_delaunay_mesh(*args, _dp(pos), cap, _ip(tri), 2 * cap, ctypes.byref(nv),
                                ctypes.byref(nt))
    if rc == -5:  # more boundary points than n: ask again with room for them
        pos = np.zeros(3 * nv.value, dtype=np.float64)
        tri = np.zeros(3 * nt.value, dtype=np.int32)
        rc = lib().bw_delaunay_mesh(*args, _dp(pos), nv.value, _ip(tri), nt.value,
                                    ctypes.byref(nv), ctypes.byref(nt))
    _check(rc)
    return (pos[:3 * nv.value].reshape(-1, 3).copy(),
            tri[:3 * nt.value].reshape(-1, 3).copy())


class BaraffClothHip:
    """The plain implicit cloth step on the GPU, cBaseCloth-style interface."""

    def __init__(self, conf, x0=None, mesh=None):
        """conf: the cloth JSON (dict); x0: initial positions (3V) applied as
        SetPos + SetPrePos; mesh: optional (rest_pos (V x 3), triangles
        (T x 3)) for geometry_type "mesh"."""
        self._h = ctypes.c_void_p()
        if mesh is None:
            _check(lib().bw_create(_json(conf), None, ctypes.byref(self._h)))
        else:
            pos = np.ascontiguousarray(mesh[0], dtype=np.float64).reshape(-1)
            tri = np.ascontiguousarray(mesh[1], dtype=np.int32).reshape(-1)
            _check(lib().bw_create_mesh(_json(conf), _dp(pos), pos.size // 3, _ip(tri),
                                        tri.size // 3, None, ctypes.byref(self._h)))
        self.nv = lib().bw_num_vertices(self._h)
        if x0 is not None:
            x = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1)
            if x.size != 3 * self.nv:
                self.close()
                raise ValueError("x0 has %d values, expected %d" % (x.size, 3 * self.nv))
            self.SetPos(x)
            self.SetPrePos(x)

    def close(self):
        if self._h:
            lib().bw_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- cBaseCloth names -------------------------------------------------
    def UpdatePos(self, dt=0.0):
        _check(lib().bw_update_pos(self._h, dt))

    def Reset(self):
        _check(lib().bw_reset(self._h))

    def GetPos(self):
        x = np.empty(3 * self.nv)
        _check(lib().bw_get_pos(self._h, _dp(x)))
        return x

    def GetPrePos(self):
        x = np.empty(3 * self.nv)
        _check(lib().bw_get_pre_pos(self._h, _dp(x)))
        return x

    def GetVelocity(self):
        v = np.empty(3 * self.nv)
        _check(lib().bw_get_velocity(self._h, _dp(v)))
        return v

    def SetPos(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == 3 * self.nv
        _check(lib().bw_set_pos(self._h, _dp(x)))

    def SetPrePos(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == 3 * self.nv
        _check(lib().bw_set_pre_pos(self._h, _dp(x)))

    def GetInitPos(self):
        x = np.empty(3 * self.nv)
        _check(lib().bw_get_init_pos(self._h, _dp(x)))
        return x

    def normals(self):
        """(vertex normals (V, 3), triangle normals (T, 3)) at the current
        positions, as cBaseCloth::Update refreshes them."""
        nt = lib().bw_num_triangles(self._h)
        vn = np.empty(3 * self.nv)
        tn = np.empty(3 * nt)
        _check(lib().bw_get_normals(self._h, _dp(vn), _dp(tn)))
        return vn.reshape(-1, 3), tn.reshape(-1, 3)

    def ApplyPerturb(self, triangle, force):
        """cBaseCloth::ApplyPerturb: force / 3 on each vertex of the triangle,
        as user force of the next substep"""
        f = np.ascontiguousarray(force, dtype=np.float64)
        _check(lib().bw_apply_perturb(self._h, int(triangle), _dp(f)))

    def ApplyUserPerturbForceOnce(self, triangle, goal):
        """cBaraffCloth::ApplyUserPerturbForceOnce: the triangle's first vertex
        moves to goal and is fixed in the next substep"""
        g = np.ascontiguousarray(goal, dtype=np.float64)
        _check(lib().bw_apply_user_perturb_force_once(self._h, int(triangle), _dp(g)))

    def SetYDragForce(self, force):
        """cBaseCloth::mYDragForce, applied by every substep"""
        _check(lib().bw_set_y_drag_force(self._h, float(force)))

    def SetPointTriangleCollisionInfo(self, ids, normal, bary, depth):
        """contacts for the next substep (cBaseObject::SetPointTriangleCollisionInfo):
        ids (n, 3) {cloth_is_point, vertex, triangle}, normal (n, 3), bary (n, 3),
        depth (n,)"""
        a = _contact_arrays({"pt_ids": ids, "pt_normal": normal, "pt_bary": bary,
                             "pt_depth": depth})
        _check(lib().bw_set_point_triangle_contacts(self._h, a["pt_depth"].size, _ip(a["pt_ids"]),
                                                    _dp(a["pt_normal"]), _dp(a["pt_bary"]),
                                                    _fp(a["pt_depth"])))

    def SetEdgeEdgeCollisionInfo(self, ids, normal, bary, depth):
        """contacts for the next substep (cBaseObject::SetEdgeEdgeCollisionInfo):
        ids (n, 3) {cloth_is_edge0, edge0, edge1}, normal (n, 3), bary (n, 2),
        depth (n,)"""
        a = _contact_arrays({"ee_ids": ids, "ee_normal": normal, "ee_bary": bary,
                             "ee_depth": depth})
        _check(lib().bw_set_edge_edge_contacts(self._h, a["ee_depth"].size, _ip(a["ee_ids"]),
                                               _dp(a["ee_normal"]), _fp(a["ee_bary"]),
                                               _fp(a["ee_depth"])))

    def contacts(self):
        """the contacts set for the next substep, as a contact dict"""
        L = lib()
        n = L.bw_get_point_triangle_contacts(self._h, 0, None, None, None, None)
        _check(n)
        m = L.bw_get_edge_edge_contacts(self._h, 0, None, None, None, None)
        _check(m)
        ct = {"pt_ids": np.zeros((n, 3), np.int32), "pt_normal": np.zeros((n, 3)),
              "pt_bary": np.zeros((n, 3)), "pt_depth": np.zeros(n, np.float32),
              "ee_ids": np.zeros((m, 3), np.int32), "ee_normal": np.zeros((m, 3)),
              "ee_bary": np.zeros((m, 2), np.float32), "ee_depth": np.zeros(m, np.float32)}
        _check(L.bw_get_point_triangle_contacts(self._h, n, _ip(ct["pt_ids"]), _dp(ct["pt_normal"]),
                                                _dp(ct["pt_bary"]), _fp(ct["pt_depth"])))
        _check(L.bw_get_edge_edge_contacts(self._h, m, _ip(ct["ee_ids"]), _dp(ct["ee_normal"]),
                                           _fp(ct["ee_bary"]), _fp(ct["ee_depth"])))
        return ct

    def detect_contacts(self, body, cloth_first=True):
        """collision detection against a KinematicBody; the contacts
        replace this cloth's for its next substep; returns them"""
        _check(lib().bw_detect_contacts(self._h, body._h, 1 if cloth_first else 0, None, None))
        return self.contacts()

    def set_contacts(self, ct):
        """both kinds from one contact dict"""
        self.SetPointTriangleCollisionInfo(ct["pt_ids"], ct["pt_normal"], ct["pt_bary"],
                                           ct["pt_depth"])
        self.SetEdgeEdgeCollisionInfo(ct["ee_ids"], ct["ee_normal"], ct["ee_bary"],
                                      ct["ee_depth"])

    def GetDefaultTimestep(self):
        return lib().bw_get_default_timestep(self._h)

    def GetNumOfVertices(self):
        return self.nv

    def GetNumOfTriangles(self):
        return _check(lib().bw_num_triangles(self._h))

    def GetNumOfEdges(self):
        return _check(lib().bw_num_edges(self._h))

    def GetConstraintStaticVertices(self):
        n = _check(lib().bw_get_constraint_static_vertices(self._h, None, 0))
        out = np.zeros(n, dtype=np.int32)
        lib().bw_get_constraint_static_vertices(self._h, _ip(out), n)
        return out

    def AddConstraintStaticVertices(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        _check(lib().bw_add_constraint_static_vertices(self._h, _ip(ids), ids.size))

    def ClearConstraintStaticVertices(self):
        _check(lib().bw_clear_constraint_static_vertices(self._h))

    # --- readback ----------------------------------------------------------
    def triangles(self):
        t = np.zeros(3 * self.GetNumOfTriangles(), dtype=np.int32)
        _check(lib().bw_get_triangles(self._h, _ip(t)))
        return t.reshape(-1, 3)

    def edges(self):
        e = np.zeros(5 * self.GetNumOfEdges(), dtype=np.int32)
        _check(lib().bw_get_edges(self._h, _ip(e)))
        return e.reshape(-1, 5)

    def ell_map(self):
        n = _check(lib().bw_get_ell_map(self._h, None, None))
        rp = np.zeros(self.nv + 1, dtype=np.int32)
        cols = np.zeros(n, dtype=np.int32)
        _check(lib().bw_get_ell_map(self._h, _ip(rp), _ip(cols)))
        return rp, cols

    def assemble_system(self):
        """(W blocks in ell_map order, rhs) assembled at the current state."""
        rp, cols = self.ell_map()
        W = np.zeros(9 * cols.size)
        b = np.zeros(3 * self.nv)
        _check(lib().bw_assemble_system(self._h, _dp(W), _dp(b)))
        return W.reshape(-1, 3, 3), b

    def last_stats(self):
        st = StepStats()
        _check(lib().bw_get_last_stats(self._h, ctypes.byref(st)))
        return st.as_dict()


class ClothBatch:
    """Many independent cloths stepped together (one launch chain per phase)."""

    def __init__(self, confs, x0s=None):
        """confs: cloth JSON dicts; x0s: per cloth, initial positions (3V)
        applied as SetPos + SetPrePos, or None for the rest state."""
        n = len(confs)
        if x0s is not None and len(x0s) != n:
            raise ValueError("%d configs but %d initial states" % (n, len(x0s)))
        texts = [_json(c) for c in confs]
        arr = (ctypes.c_char_p * n)(*texts)
        ptrs = (ctypes.POINTER(ctypes.c_double) * n)()
        self._h = ctypes.c_void_p()
        _check(lib().bw_batch_create(arr, ptrs, n, ctypes.byref(self._h)))
        self.n = n
        self.nv = [lib().bw_batch_num_vertices(self._h, i) for i in range(n)]
        for i, x in enumerate(x0s or []):
            if x is None:
                continue
            x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
            if x.size != 3 * self.nv[i]:
                self.close()
                raise ValueError("x0 of cloth %d has %d values, expected %d"
                                 % (i, x.size, 3 * self.nv[i]))
            _check(lib().bw_batch_set_state(self._h, i, _dp(x), _dp(x)))

    def close(self):
        if self._h:
            lib().bw_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps=1):
        _check(lib().bw_batch_step(self._h, nsteps))

    def num_blocks(self, i):
        """true 3x3 blocks of member i's matrix"""
        n = lib().bw_batch_num_blocks(self._h, i)
        if n < 0:
            _check(n)
        return n

    def step_noalloc(self, nsteps=1):
        """bw_batch_step_noalloc: the same steps with the solver scratch in the
        library's static device array (no allocation)"""
        _check(lib().bw_batch_step_noalloc(self._h, nsteps))

    def set_graphs(self, on):
        _check(lib().bw_batch_set_graphs(self._h, 1 if on else 0))

    def get_pos(self, i):
        x = np.empty(3 * self.nv[i])
        _check(lib().bw_batch_get_pos(self._h, i, _dp(x)))
        return x

    def time_solver_kernel(self, reps=50):
        """(mean us per launch of the dominant solver kernel, algorithmic
        bytes per launch), on the system assembled at the current state."""
        us = np.zeros(1)
        nbytes = np.zeros(1)
        _check(lib().bw_batch_time_solver_kernel(self._h, reps, _dp(us), _dp(nbytes)))
        return float(us[0]), float(nbytes[0])

    def time_launch_floor(self, reps=200):
        """mean us per launch of an empty kernel with the grid and workgroup
        of the solver kernel's launch (bw_batch_time_launch_floor)"""
        us = np.zeros(1)
        _check(lib().bw_batch_time_launch_floor(self._h, reps, _dp(us)))
        return float(us[0])

    def autotune(self):
        """(lanes per row chosen, {lanes: mean SpMV launch us}) for the
        direction form and load hint chosen; autotune_all() has every form"""
        lanes, staged, nt, us = self.autotune_all()
        return lanes, us[("staged" if staged else "gathered") + ("_nt" if nt else "")]

    def autotune_all(self):
        """(lanes per row, staged (bool), nontemporal (bool),
        {"gathered"|"staged"|"gathered_nt"|"staged_nt": {lanes: us}}); a
        variant not timed has 0 us"""
        lanes = ctypes.c_int()
        staged = ctypes.c_int()
        nt = ctypes.c_int()
        us = np.zeros(12)
        _check(lib().bw_batch_autotune(self._h, ctypes.byref(lanes), ctypes.byref(staged),
                                       ctypes.byref(nt), _dp(us)))
        forms = ("gathered", "staged", "gathered_nt", "staged_nt")
        return (lanes.value, bool(staged.value), bool(nt.value),
                {f: {1: float(us[3 * i]), 2: float(us[3 * i + 1]), 4: float(us[3 * i + 2])}
                 for i, f in enumerate(forms)})

    def set_state(self, i, x, xpre):
        x = np.ascontiguousarray(x, dtype=np.float64)
        xpre = np.ascontiguousarray(xpre, dtype=np.float64)
        _check(lib().bw_batch_set_state(self._h, i, _dp(x), _dp(xpre)))

    def normals(self, i):
        """(vertex normals (V, 3), triangle normals (T, 3)) of member i"""
        vn = np.empty(3 * self.nv[i])
        tn = np.empty(3 * lib().bw_batch_num_triangles(self._h, i))
        _check(lib().bw_batch_get_normals(self._h, i, _dp(vn), _dp(tn)))
        return vn.reshape(-1, 3), tn.reshape(-1, 3)

    def apply_perturb(self, i, triangle, force):
        f = np.ascontiguousarray(force, dtype=np.float64)
        _check(lib().bw_batch_apply_perturb(self._h, i, int(triangle), _dp(f)))

    def apply_user_perturb_force_once(self, i, triangle, goal):
        g = np.ascontiguousarray(goal, dtype=np.float64)
        _check(lib().bw_batch_apply_user_perturb_force_once(self._h, i, int(triangle), _dp(g)))

    def set_y_drag_force(self, i, force):
        _check(lib().bw_batch_set_y_drag_force(self._h, i, float(force)))

    def set_contacts(self, i, ct):
        """contacts (a dict as workloads.random_contacts returns) for member
        i's next step"""
        a = _contact_arrays(ct)
        L = lib()
        _check(L.bw_batch_set_point_triangle_contacts(
            self._h, i, a["pt_depth"].size, _ip(a["pt_ids"]), _dp(a["pt_normal"]),
            _dp(a["pt_bary"]), _fp(a["pt_depth"])))
        _check(L.bw_batch_set_edge_edge_contacts(
            self._h, i, a["ee_depth"].size, _ip(a["ee_ids"]), _dp(a["ee_normal"]),
            _fp(a["ee_bary"]), _fp(a["ee_depth"])))

    def detect_contacts(self, i, body, cloth_first=True):
        """collision detection of member i against a KinematicBody;
        returns (point-triangle, edge-edge) counts"""
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().bw_batch_detect_contacts(self._h, i, body._h, 1 if cloth_first else 0,
                                              ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def stats(self, i):
        st = StepStats()
        _check(lib().bw_batch_get_stats(self._h, i, ctypes.byref(st)))
        return st.as_dict()


class KinematicBody:
    """a kinematic body (triangle mesh) for collision detection
    (bw_body_create)"""

    def __init__(self, pos, tri):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1)
        self.nv, self.nt = pos.size // 3, tri.size // 3
        h = ctypes.c_void_p()
        _check(lib().bw_body_create(_dp(pos), self.nv, _ip(tri), self.nt, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().bw_body_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pos(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        if pos.size != 3 * self.nv:
            raise ValueError("expected %d positions" % (3 * self.nv))
        _check(lib().bw_body_set_pos(self._h, _dp(pos)))

    def edges(self):
        ne = lib().bw_body_num_edges(self._h)
        _check(ne)
        e = np.zeros(4 * ne, np.int32)
        _check(lib().bw_body_get_edges(self._h, _ip(e)))
        return e.reshape(-1, 4)


class Server:
    """Serving runtime (bw_server_*): submit step requests of any cloth at
    any time; they are admitted against a device-memory budget, solved
    together with everything resident, and retired with their final
    positions."""

    def __init__(self, budget_bytes=200e9, max_resident=256, precision="mixed"):
        h = ctypes.c_void_p()
        _check(lib().bw_server_create(float(budget_bytes), int(max_resident),
                                      precision.encode(), ctypes.byref(h)))
        self._h = h

    def submit(self, config, x0=None, nsteps=10):
        text = json.dumps(config).encode()
        rid = ctypes.c_longlong()
        if x0 is None:
            _check(lib().bw_server_submit(self._h, text, None, 0, int(nsteps), ctypes.byref(rid)))
        else:
            x0 = np.ascontiguousarray(x0, dtype=np.float64)
            _check(lib().bw_server_submit(self._h, text, _dp(x0), x0.size, int(nsteps),
                                          ctypes.byref(rid)))
        return rid.value

    def wait(self, rid, timeout=600.0):
        """True when the request is done (or failed) within timeout seconds."""
        rc = lib().bw_server_wait(self._h, rid, float(timeout))
        if rc == -4:
            return False
        _check(rc)
        return True

    def info(self, rid):
        st = RequestInfo()
        _check(lib().bw_server_get_info(self._h, rid, ctypes.byref(st)))
        d = st.as_dict()
        d["error"] = lib().bw_last_error().decode() if st.status != 0 else ""
        return d

    def get_pos(self, rid):
        nv = self.info(rid)["num_vertices"]
        x = np.empty(3 * nv)
        _check(lib().bw_server_get_pos(self._h, rid, _dp(x)))
        return x

    def release(self, rid):
        _check(lib().bw_server_release(self._h, rid))

    def counters(self):
        c = ServerCounters()
        _check(lib().bw_server_get_counters(self._h, ctypes.byref(c)))
        return c.as_dict()

    def now(self):
        return lib().bw_server_now(self._h)

    def close(self):
        if self._h:
            lib().bw_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
