"""Batched stepping: every cloth's result is the same alone and in the mix
(bit for bit: per-cloth reductions run in a fixed order that does not depend
on the batch), and the batched result agrees with the oracle."""
import sys

import numpy as np
import pytest

import oracle_ffi as of
from acousticsim_amd import workloads as wl

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


MEMBERS = [(9, 1.0, 21), (30, 3.0, 22), (17, 0.2, 23), (45, 1.0, 24), (21, 7.0, 25)]


def _confs(precision):
    out = []
    for side, scale, seed in MEMBERS:
        c = wl.grid_config(side, stiffness_scale=scale)
        c["precision"] = precision
        out.append((c, wl.jittered_positions(c, seed)))
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("graphs", [True, False])
def test_alone_equals_in_mix(precision, graphs, monkeypatch):
    """Bit-identical for one SpMV variant (the autotuner may pick different
    variants for different batches, and variants sum a row's blocks in
    different orders), so the variant is fixed here."""
    from acousticsim_amd import capi

    monkeypatch.setenv("BW_SPMV_LPR", "2")
    cs = _confs(precision)
    say("case batch", precision, "graphs", graphs)
    b = capi.ClothBatch([c for c, _ in cs], [x for _, x in cs])
    b.set_graphs(graphs)
    b.step(5)
    mixed = [b.get_pos(i) for i in range(len(cs))]
    for i, (c, x0) in enumerate(cs):
        say("  alone", i)
        a = capi.ClothBatch([c], [x0])
        a.set_graphs(graphs)
        a.step(5)
        assert np.array_equal(a.get_pos(0), mixed[i]), i
        assert a.stats(0)["pcg_iterations"] == b.stats(i)["pcg_iterations"]
    # and the batch member equals the single-cloth C ABI object
    cl = capi.BaraffClothHip(cs[1][0], cs[1][1])
    for _ in range(5):
        cl.UpdatePos()
    assert np.array_equal(cl.GetPos(), mixed[1])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batched_step_matches_oracle(precision):
    from acousticsim_amd import capi

    cs = _confs(precision)
    say("case batch vs oracle", precision)
    b = capi.ClothBatch([c for c, _ in cs], [x for _, x in cs])
    b.step(1)
    for i, (c, x0) in enumerate(cs):
        oc = of.OracleCloth(c)
        oc.set_state(x0, x0)
        oc.step(of.SOLVE_TOLERANCE, 1e-12)
        xr = oc.pos()
        d = np.abs(xr - x0).max()
        err = np.abs(b.get_pos(i) - xr).max()
        st = b.stats(i)
        say("  member %d err/|dx| %.3e iters %d residual %.2e" % (i, err / d, st["pcg_iterations"],
                                                                 st["rel_residual"]))
        assert st["rel_residual"] <= 1e-6
        assert err <= 1e-4 * d


def test_autotuner_variants_agree(monkeypatch):
    """The autotuner times the 1-, 2- and 4-lane SpMV variants and keeps the
    fastest; every variant gives the same step within the solve tolerance."""
    from acousticsim_amd import capi

    cs = _confs("mixed")
    b = capi.ClothBatch([c for c, _ in cs], [x for _, x in cs])
    lanes, us = b.autotune()
    say("case autotune: picked %d, us %s" % (lanes, us))
    assert lanes in (1, 2, 4) and us[lanes] == min(us.values())
    out = {}
    for lpr in ("1", "2", "4"):
        monkeypatch.setenv("BW_SPMV_LPR", lpr)
        b = capi.ClothBatch([c for c, _ in cs], [x for _, x in cs])
        b.step(1)
        out[lpr] = [b.get_pos(i) for i in range(len(cs))]
        assert max(b.stats(i)["rel_residual"] for i in range(len(cs))) <= 1e-6
    for i, (c, x0) in enumerate(cs):
        d = np.abs(out["4"][i] - x0).max()
        for lpr in ("1", "2"):
            assert np.abs(out[lpr][i] - out["4"][i]).max() <= 1e-4 * d


@pytest.mark.parametrize("graphs", [True, False])
def test_group_sums_alone_equals_mix_and_oracle(graphs, monkeypatch):
    """Cloths above 512 slices reduce their solver sums through group sums
    (one extra small launch per solver launch).  Forced on for every cloth
    here: each result is still the same alone and in the mix (bit for bit),
    and one step agrees with the oracle within 1e-4 of the displacement."""
    from acousticsim_amd import capi

    monkeypatch.setenv("BW_SPMV_LPR", "2")
    monkeypatch.setenv("BW_GROUP_MIN", "0")
    cs = _confs("mixed")
    say("case group sums, graphs", graphs)
    b = capi.ClothBatch([c for c, _ in cs], [x for _, x in cs])
    b.set_graphs(graphs)
    b.step(3)
    for i, (c, x0) in enumerate(cs):
        a = capi.ClothBatch([c], [x0])
        a.set_graphs(graphs)
        a.step(3)
        assert np.array_equal(a.get_pos(0), b.get_pos(i)), i
    one = capi.ClothBatch([c for c, _ in cs], [x for _, x in cs])
    one.step(1)
    for i, (c, x0) in enumerate(cs):
        oc = of.OracleCloth(c)
        oc.set_state(x0, x0)
        oc.step(of.SOLVE_TOLERANCE, 1e-12)
        d = np.abs(oc.pos() - x0).max()
        err = np.abs(one.get_pos(i) - oc.pos()).max()
        say("  member %d err/|dx| %.3e" % (i, err / d))
        assert one.stats(i)["rel_residual"] <= 1e-6
        assert err <= 1e-4 * d


def test_stalled_solves_move_to_fp64_matrix(monkeypatch):
    """Mixed form: a true-residual check forced by the iteration count that
    finds the residual stalled moves the cloth's SpMV to the FP64 matrix for
    the rest of its solve.  Forced here (a check every 20 iterations, any
    progress counted as a stall): every cloth switches, still converges to
    1e-6, matches the oracle within 1e-4 of the step, and is the same alone
    and in the mix."""
    from acousticsim_amd import capi

    monkeypatch.setenv("BW_SPMV_LPR", "2")
    monkeypatch.setenv("BW_MAX_SINCE", "20")
    monkeypatch.setenv("BW_STALL", "1e-30")
    cs = _confs("mixed")
    say("case forced FP64-matrix switch")
    b = capi.ClothBatch([c for c, _ in cs], [x for _, x in cs])
    b.step(1)
    for i, (c, x0) in enumerate(cs):
        st = b.stats(i)
        assert st["matrix_fp64"] == 1 and st["rel_residual"] <= 1e-6
        oc = of.OracleCloth(c)
        oc.set_state(x0, x0)
        oc.step(of.SOLVE_TOLERANCE, 1e-12)
        d = np.abs(oc.pos() - x0).max()
        err = np.abs(b.get_pos(i) - oc.pos()).max()
        say("  member %d err/|dx| %.3e iters %d" % (i, err / d, st["pcg_iterations"]))
        assert err <= 1e-4 * d
        a = capi.ClothBatch([c], [x0])
        a.step(1)
        assert np.array_equal(a.get_pos(0), b.get_pos(i))
    monkeypatch.setenv("BW_STALL", "0")
    n = capi.ClothBatch([c for c, _ in cs], [x for _, x in cs])
    n.step(1)
    assert all(n.stats(i)["matrix_fp64"] == 0 for i in range(len(cs)))
