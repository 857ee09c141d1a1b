"""Closed-form, finite-difference and structural checks of the oracle that do
not rely on the reference's outputs (a second, independent pin)."""
import math

import numpy as np
import pytest

import oracle_ffi as of
from acousticsim_amd import workloads as wl


def test_dot_reduction_matches_exact_sum():
    rng = np.random.default_rng(0)
    for n in (0, 1, 2, 3, 4, 5, 7, 8, 9, 1001):
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        exact = math.fsum((a * b).tolist())
        got = of.lib().oc_dot(of._dp(a), of._dp(b), n)
        assert abs(got - exact) <= 1e-13 * max(1.0, np.abs(a * b).sum())


def test_dbeta_dx_matches_central_differences():
    rng = np.random.default_rng(1)
    for _ in range(20):
        v = [rng.standard_normal(3) for _ in range(4)]
        beta, g = of.dbeta_dx(*v)
        num = np.zeros(12)
        h = 1e-6
        for k in range(12):
            vp = [x.copy() for x in v]
            vm = [x.copy() for x in v]
            vp[k // 3][k % 3] += h
            vm[k // 3][k % 3] -= h
            num[k] = (of.dbeta_dx(*vp)[0] - of.dbeta_dx(*vm)[0]) / (2 * h)
        assert np.allclose(g, num, rtol=1e-6, atol=1e-8)
        # beta is invariant to translation: gradient sums to zero
        assert np.abs(g.reshape(4, 3).sum(0)).max() < 1e-10 * np.abs(g).max()


def _stretch_energy(x, coef, area, k):
    e = 0.0
    for s, (Ku, Kv) in enumerate(((k[0], k[1]), (k[2], k[2]))):
        cu, cv = coef[6 * s:6 * s + 3], coef[6 * s + 3:6 * s + 6]
        Fu = cu[0] * x[0] + cu[1] * x[1] + cu[2] * x[2]
        Fv = cv[0] * x[0] + cv[1] * x[1] + cv[2] * x[2]
        Cu, Cv = np.linalg.norm(Fu) - 1, np.linalg.norm(Fv) - 1
        e += area / 2 * (Ku * Cu * Cu + Kv * Cv * Cv)
    return e


def test_stretch_force_is_minus_energy_gradient():
    """f = -dE/dx with E = A/2 (Ku Cu^2 + Kv Cv^2) summed over both direction
    sets (BaraffMaterial.cpp:97-125)."""
    oc = of.OracleCloth(wl.grid_config(5))
    coef, area = oc.stretch_coef()
    rng = np.random.default_rng(2)
    k = np.array([1e3, 7e2, 1e2])
    tri = oc.triangles()
    xr = oc.rest_pos().reshape(-1, 3)
    for t in range(0, len(tri), 5):
        x = xr[tri[t]] * (1 + 0.2 * rng.random()) + 0.01 * rng.standard_normal((3, 3))
        f, _ = of.stretch_element(x[0], x[1], x[2], coef[t], area[t], k)
        num = np.zeros((3, 3))
        h = 1e-7
        for i in range(3):
            for a in range(3):
                xp, xm = x.copy(), x.copy()
                xp[i, a] += h
                xm[i, a] -= h
                num[i, a] = -(_stretch_energy(xp, coef[t], area[t], k) -
                              _stretch_energy(xm, coef[t], area[t], k)) / (2 * h)
        assert np.allclose(f, num, rtol=1e-5, atol=1e-9 * np.abs(f).max())


def test_stretch_hessian_exact_when_stretched_isotropically():
    """With |Fu| = |Fv| > 1 the reference's Dv scaling (BaraffMaterial.cpp:142)
    coincides with the true one, so K = df/dx exactly: check against central
    differences of the force under a uniform 10% stretch."""
    oc = of.OracleCloth(wl.grid_config(5))
    coef, area = oc.stretch_coef()
    tri = oc.triangles()
    xr = oc.rest_pos().reshape(-1, 3)
    k = np.array([1e3, 1e3, 0.0])
    t = 3
    x = xr[tri[t]] * 1.1
    f, H = of.stretch_element(x[0], x[1], x[2], coef[t], area[t], k)
    num = np.zeros((9, 9))
    h = 1e-7
    for j in range(3):
        for b in range(3):
            xp, xm = x.copy(), x.copy()
            xp[j, b] += h
            xm[j, b] -= h
            fp, _ = of.stretch_element(xp[0], xp[1], xp[2], coef[t], area[t], k)
            fm, _ = of.stretch_element(xm[0], xm[1], xm[2], coef[t], area[t], k)
            num[:, 3 * j + b] = ((fp - fm) / (2 * h)).ravel()
    Hfull = np.zeros((9, 9))
    for i in range(3):
        for j in range(3):
            Hfull[3 * i:3 * i + 3, 3 * j:3 * j + 3] = H[3 * i + j]
    assert np.allclose(Hfull, num, rtol=1e-5, atol=1e-6 * np.abs(Hfull).max())


def test_rest_state_is_force_free_and_K_translation_invariant():
    conf = wl.grid_config(9)
    conf["enable_gravity"] = False
    oc = of.OracleCloth(conf)
    x = oc.rest_pos()
    oc.set_state(x, x)
    oc.assemble()
    # at rest |Fu| = 1 up to the float rounding of the stretch coefficients
    # (BaraffMaterial.cpp:21-24), so the residual force is at float level
    coef, area = oc.stretch_coef()
    scale = 1e3 * area.max() * np.abs(coef).max()
    assert np.abs(oc.force()).max() < 1e-6 * scale
    # a flat, unstretched cloth stays put (to that same level)
    oc.step(of.SOLVE_TOLERANCE, 1e-12)
    assert np.abs(oc.pos() - x).max() < 1e-9
    # K * (1 (x) e_k) = 0 up to the float rounding of the coefficients
    rng = np.random.default_rng(3)
    xj = x + 1e-3 * rng.standard_normal(x.size)
    oc.set_state(xj, xj)
    oc.assemble()
    K = oc.K()
    rp, cols = oc.pattern()
    for v in range(oc.nv):
        blk = K[rp[v]:rp[v + 1]].sum(0)
        scale = np.abs(K[rp[v]:rp[v + 1]]).max()
        assert np.abs(blk).max() <= 1e-5 * scale


def _dense(oc, W):
    rp, cols = oc.pattern()
    n = 3 * oc.nv
    A = np.zeros((n, n))
    for v in range(oc.nv):
        for s in range(rp[v], rp[v + 1]):
            c = cols[s]
            A[3 * v:3 * v + 3, 3 * c:3 * c + 3] = W[s]
    return A


def test_W_symmetric_positive_definite_and_solve_matches_dense():
    conf = wl.grid_config(7)
    oc = of.OracleCloth(conf)
    x0 = wl.jittered_positions(conf, 5)
    oc.set_state(x0, x0)
    oc.assemble()
    A = _dense(oc, oc.W())
    assert np.abs(A - A.T).max() <= 1e-7 * np.abs(A).max()  # float-stored bending H
    free = np.ones(3 * oc.nv, bool)
    for p in oc.pins():
        free[3 * p:3 * p + 3] = False
    Af = A[np.ix_(free, free)]
    assert np.linalg.eigvalsh(0.5 * (Af + Af.T)).min() > 0
    b = oc.rhs()
    dx_dense = np.zeros(3 * oc.nv)
    dx_dense[free] = np.linalg.solve(Af, b[free])
    oc.set_state(x0, x0)
    oc.step(of.SOLVE_TOLERANCE, 1e-13)
    assert np.allclose(oc.dx(), dx_dense, rtol=0, atol=1e-10 * np.abs(dx_dense).max())


# ELL widths: an independent recount of the reference GPU rule
# (BaraffClothGpu.cpp:438-500) on the reference grid topology.
WIDTHS = {
    21: (4521, {6: 44, 9: 216, 11: 4, 12: 32, 13: 145}),
    257: (722433, {6: 516, 9: 33020, 11: 4, 12: 504, 13: 32005}),
    17: (2913, None),
    129: (180993, None),
}


@pytest.mark.parametrize("side", sorted(WIDTHS))
def test_reference_ell_rule_widths(side):
    oc = of.OracleCloth(wl.grid_config(side))
    rp, cols = oc.ref_ell()
    nnzb, hist = WIDTHS[side]
    assert rp[-1] == nnzb
    w = np.diff(rp)
    if hist is not None:
        u, c = np.unique(w, return_counts=True)
        assert dict(zip(u.tolist(), c.tolist())) == hist
    # on grids the rule equals the full assembly pattern (every triangle has
    # an interior edge, so triangle pairs are covered by some stencil)
    prp, pcols = oc.pattern()
    assert np.array_equal(rp, prp) and np.array_equal(cols, pcols)
    # rows sorted ascending, self included
    for v in range(0, oc.nv, max(1, oc.nv // 50)):
        row = cols[rp[v]:rp[v + 1]]
        assert np.all(np.diff(row) > 0) and v in row
