"""The reference's GPU form (cloth_type "fem_gpu") in the oracle, pinned by
mathematics and by the plain method rather than by retyping it:

  * the quadratic-bending matrix is symmetric positive semi-definite,
    annihilates the flat rest positions and rigid translations, scales with
    the stiffness, and rotating the UV frame by 90 degrees swaps the warp
    and weft stiffness;
  * with bending off, its FP32 stretch, assembly and 100-iteration
    block-Jacobi PCG agree with the plain method's FP64 solve to 1e-3 of the
    step (both start with the pinned vertices at their targets);
  * pinned vertices are put back on their targets every step.
"""
import numpy as np

import oracle_ffi as of
from acousticsim_amd import workloads as wl


def _dense_bending(oc):
    qb = oc.gpu_form_bending()
    rp, cols = oc.pattern()
    K = np.zeros((oc.nv, oc.nv))
    K[np.repeat(np.arange(oc.nv), np.diff(rp)), cols] = qb
    return K


def test_quadratic_bending_matrix_properties():
    c = wl.grid_config(9)
    c["cloth_init_orientation"] = [0.4, -0.2, 0.3]  # a tilted flat cloth
    oc = of.OracleCloth(c)
    K = _dense_bending(oc)
    assert np.array_equal(K, K.T)
    ev = np.linalg.eigvalsh(K)
    assert ev[0] >= -1e-6 * ev[-1] and ev[-1] > 0
    x = oc.rest_pos().reshape(-1, 3)
    assert np.abs(K @ x).max() <= 1e-5 * np.abs(K).max() * np.abs(x).max()
    assert np.abs(K.sum(axis=1)).max() <= 1e-5 * np.abs(K).max()


def test_bending_scales_with_stiffness_and_uv_rotation_swaps_axes():
    def qb(ku, kv, rot):
        c = wl.grid_config(9, cloth_uv_rotation=rot)
        c["cloth_bending_stiffness"] = [ku, kv]
        return of.OracleCloth(c).gpu_form_bending()
    a = qb(1e-4, 1e-4, 0)
    b = qb(2e-4, 2e-4, 0)
    assert np.abs(b - 2 * a).max() <= 1e-6 * np.abs(a).max()
    s = qb(3e-4, 1e-4, 90)
    t = qb(1e-4, 3e-4, 0)
    assert np.abs(s - t).max() <= 1e-5 * np.abs(t).max()
    assert np.abs(qb(3e-4, 1e-4, 0) - t).max() > 1e-2 * np.abs(t).max()  # anisotropic


def _start(conf, oc):
    x0 = oc.rest_pos().copy()
    x0[2::3] += 1e-3 * np.sin(np.arange(oc.nv))
    for p in oc.pins():
        x0[3 * p:3 * p + 3] = np.float32(oc.rest_pos()[3 * p:3 * p + 3])
    return x0


def test_without_bending_matches_the_plain_method():
    c = wl.grid_config(11)
    c["cloth_bending_stiffness"] = [0.0, 0.0]
    a, b = of.OracleCloth(c), of.OracleCloth(c)
    x0 = _start(c, a)
    a.set_state(x0, x0)
    b.set_state(x0, x0)
    for _ in range(3):
        st = a.gpu_form_step()
        b.step(of.SOLVE_TOLERANCE, 1e-10)
        assert st["iterations"] == 100
        d = np.abs(b.pos() - x0).max()
        assert np.abs(a.pos() - b.pos()).max() <= 1e-3 * d


def test_pins_return_to_their_targets():
    c = wl.grid_config(9)
    oc = of.OracleCloth(c)
    x0 = wl.jittered_positions(c, 4)  # pins off target by the jitter
    oc.set_state(x0, x0)
    oc.gpu_form_step()
    x = oc.pos().reshape(-1, 3)
    rest = oc.rest_pos().reshape(-1, 3)
    for p in oc.pins():
        assert np.array_equal(x[p], rest[p].astype(np.float32).astype(np.float64))
