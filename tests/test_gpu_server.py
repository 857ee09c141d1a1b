"""Serving runtime: requests admitted while others run give the same result
as alone (bit for bit), the same as a batch stepping the cloth by itself,
agree with the oracle, respect the device-memory budget, and report bad
requests through their status."""
import sys
import time

import numpy as np
import pytest

import oracle_ffi as of
from acousticsim_amd import workloads as wl

pytestmark = pytest.mark.gpu

STEPS = 4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _req(side, scale, seed, precision="mixed"):
    c = wl.grid_config(side, stiffness_scale=scale)
    c["precision"] = precision
    return c, wl.jittered_positions(c, seed)


REQS = [_req(33, 1.0, 31), _req(65, 4.0, 32), _req(17, 0.3, 33), _req(48, 1.0, 34)]


def _served(reqs, stagger=0.0, **kw):
    from acousticsim_amd import capi

    s = capi.Server(**kw)
    ids = []
    for c, x0 in reqs:
        ids.append(s.submit(c, x0, STEPS))
        if stagger:
            time.sleep(stagger)
    out = []
    for rid in ids:
        assert s.wait(rid, 120.0), "request did not finish"
        info = s.info(rid)
        out.append((info, s.get_pos(rid) if info["status"] == 0 else None))
    counters = s.counters()
    s.close()
    return out, counters


@pytest.mark.parametrize("stage", ["0", "1"])
def test_served_alone_equals_batch_alone(stage, monkeypatch):
    """The server runs the 1-lane SpMV variant, the directions staged in LDS
    (its default) or gathered; the batch is held to the same variant (the
    nontemporal hint changes no result bit)."""
    from acousticsim_amd import capi

    monkeypatch.setenv("BW_SPMV_LPR", "1")
    monkeypatch.setenv("BW_SPMV_STAGE", stage)
    monkeypatch.setenv("BW_SPMV_NT", "0")
    c, x0 = REQS[0]
    # the server stages the lists of 99% of a request's slices; the batch too
    monkeypatch.setenv("BW_STAGE_LDS", "p99")
    say("case served alone vs batch alone, stage", stage)
    info, x = _served([(c, x0)])[0][0]
    assert info["status"] == 0 and info["steps_done"] == STEPS
    b = capi.ClothBatch([c], [x0])
    b.step(STEPS)
    assert np.array_equal(b.get_pos(0), x)


def test_served_in_mix_equals_alone():
    say("case served mix")
    alone = [_served([r])[0][0] for r in REQS]
    mix, counters = _served(REQS, stagger=0.01)
    say("  counters", counters)
    assert counters["max_resident_seen"] >= 2, "requests never overlapped"
    for i, ((ia, xa), (im, xm)) in enumerate(zip(alone, mix)):
        say("  request %d iters alone %d mix %d" % (i, ia["pcg_iterations"], im["pcg_iterations"]))
        assert ia["status"] == 0 and im["status"] == 0
        assert np.array_equal(xa, xm), i
        assert ia["pcg_iterations"] == im["pcg_iterations"]
        assert im["max_rel_residual"] <= 1e-6


def test_served_matches_oracle():
    say("case served vs oracle")
    c, x0 = REQS[2]
    (info, x), = _served([(c, x0)])[0]
    oc = of.OracleCloth(c)
    oc.set_state(x0, x0)
    for _ in range(STEPS):
        oc.step(of.SOLVE_TOLERANCE, 1e-12)
    xr = oc.pos()
    d = np.abs(xr - x0).max()
    err = np.abs(x - xr).max()
    say("  err/disp %.3e" % (err / d))
    assert info["max_rel_residual"] <= 1e-6
    assert err <= 1e-4 * d


def test_budget_limits_residency():
    from acousticsim_amd import capi

    say("case budget")
    reqs = [_req(40, 1.0, 40 + k) for k in range(3)]
    # the budget of one request: what admission reserves for one (an upper
    # bound of what it allocates), then allow 1.5x
    (info, _), = _served([reqs[0]])[0]
    assert 0 < info["device_bytes"] <= info["admitted_bytes"]
    budget = 1.5 * info["admitted_bytes"]
    out, counters = _served(reqs, budget_bytes=budget)
    say("  counters", counters)
    assert counters["max_resident_seen"] == 1
    assert all(i["status"] == 0 for i, _ in out)
    # a request larger than the whole budget is refused, not queued forever
    s = capi.Server(budget_bytes=1e6)
    rid = s.submit(reqs[0][0], reqs[0][1], 1)
    assert s.wait(rid, 60.0)
    assert s.info(rid)["status"] == -5
    s.close()


def test_bad_requests_report_errors():
    from acousticsim_amd import capi

    say("case bad requests")
    s = capi.Server()
    c = wl.grid_config(9)
    bad_type = dict(c, cloth_type="pbd")
    bad_prec = dict(c, precision="fp64")
    bad_mode = dict(c, solver_mode="reference")
    # a grid whose vertex count would overflow: refused while the request's
    # configuration is read, the builder thread and the server unharmed
    bad_size = dict(c, subdivision=[65536, 65536])
    ids = [s.submit(bad_type, None, 1), s.submit(bad_prec, None, 1), s.submit(bad_mode, None, 1),
           s.submit(bad_size, None, 1)]
    good = s.submit(c, None, 1)
    for rid in ids:
        assert s.wait(rid, 60.0)
        assert s.info(rid)["status"] == -1
        with pytest.raises(capi.BwError):
            s.get_pos(rid)
    assert s.wait(good, 60.0) and s.info(good)["status"] == 0
    with pytest.raises(capi.BwError):
        s.submit(c, None, 0)
    s.close()


@pytest.mark.parametrize("stage", ["0", "1"])
def test_served_with_group_sums_equals_batch(stage, monkeypatch):
    """Requests whose cloths reduce through group sums (forced for every
    cloth here) give, in a mix, the same result as a batch stepping each one
    alone."""
    from acousticsim_amd import capi

    monkeypatch.setenv("BW_SPMV_LPR", "1")
    monkeypatch.setenv("BW_SPMV_STAGE", stage)
    monkeypatch.setenv("BW_SPMV_NT", "0")
    monkeypatch.setenv("BW_GROUP_MIN", "0")
    # the server stages the lists of 99% of a request's slices; the batch too
    monkeypatch.setenv("BW_STAGE_LDS", "p99")
    say("case served group sums, stage", stage)
    mix, counters = _served(REQS, stagger=0.01)
    for i, (c, x0) in enumerate(REQS):
        info, x = mix[i]
        assert info["status"] == 0
        b = capi.ClothBatch([c], [x0])
        b.step(STEPS)
        assert np.array_equal(b.get_pos(0), x), i
