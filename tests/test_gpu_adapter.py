"""The call sequence of the documented reference-side adapter
(cBaraffClothHip::UpdatePos: bw_update_pos, then bw_get_pos, and only the
host-side cBaseCloth::SetPos) keeps the device solver's state across steps:
it equals stepping the cloth without host calls in between, bit for bit,
while re-uploading the positions after every step (bw_set_pos, the virtual
SetPos) restarts each solve on the FP32 matrix and changes the result."""
import sys

import numpy as np
import pytest

from acousticsim_amd import workloads as wl

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_adapter_sequence_keeps_the_fp64_matrix_choice(monkeypatch):
    from acousticsim_amd import capi

    # every solve's first forced residual check counts as stalled, so each
    # cloth moves to the FP64 matrix in step 1 and starts the next steps on
    # it -- unless its state was uploaded in between, which clears that
    # history
    monkeypatch.setenv("BW_SPMV_LPR", "1")
    monkeypatch.setenv("BW_SPMV_STAGE", "0")
    monkeypatch.setenv("BW_MAX_SINCE", "20")
    monkeypatch.setenv("BW_STALL", "1e-30")
    c = wl.grid_config(33, stiffness_scale=2.0)
    c["precision"] = "mixed"
    x0 = wl.jittered_positions(c, 9)
    say("case adapter call sequence")
    plain = capi.BaraffClothHip(c, x0)
    adapter = capi.BaraffClothHip(c, x0)
    upload = capi.BaraffClothHip(c, x0)
    iters = {"plain": [], "adapter": [], "upload": []}
    for _ in range(4):
        plain.UpdatePos()
        iters["plain"].append(plain.last_stats()["pcg_iterations"])
        adapter.UpdatePos()
        x = adapter.GetPos()  # then cBaseCloth::SetPos(x) on the host only
        iters["adapter"].append(adapter.last_stats()["pcg_iterations"])
        upload.UpdatePos()
        upload.SetPos(upload.GetPos())  # the virtual SetPos: uploads again
        iters["upload"].append(upload.last_stats()["pcg_iterations"])
        assert plain.last_stats()["matrix_fp64"] == 1
    say("  iterations", iters)
    assert np.array_equal(adapter.GetPos(), plain.GetPos())
    assert iters["adapter"] == iters["plain"]
    assert not np.array_equal(upload.GetPos(), plain.GetPos())
    assert iters["upload"][1:] != iters["plain"][1:]
    assert np.array_equal(x, adapter.GetPos())


def test_fp64_matrix_carries_over_after_the_set_count(monkeypatch):
    """BW_STALL_CARRY=2: a cloth starts a solve on the FP64 matrix only after
    two solves in a row switched to it.  With every forced check counted as a
    stall, steps 1 and 2 both start on the FP32 matrix (bit-identical to a
    cloth whose history is cleared after every step) and step 3 starts on
    the FP64 matrix (different from that cloth)."""
    from acousticsim_amd import capi

    monkeypatch.setenv("BW_SPMV_LPR", "1")
    monkeypatch.setenv("BW_SPMV_STAGE", "0")
    monkeypatch.setenv("BW_MAX_SINCE", "20")
    monkeypatch.setenv("BW_STALL", "1e-30")
    monkeypatch.setenv("BW_STALL_CARRY", "2")
    c = wl.grid_config(33, stiffness_scale=2.0)
    c["precision"] = "mixed"
    x0 = wl.jittered_positions(c, 9)
    say("case FP64-matrix carry-over after two stalled solves")
    carry = capi.BaraffClothHip(c, x0)
    cleared = capi.BaraffClothHip(c, x0)
    for step in range(1, 4):
        carry.UpdatePos()
        cleared.UpdatePos()
        cleared.SetPos(cleared.GetPos())  # an upload clears the solver history
        assert carry.last_stats()["matrix_fp64"] == 1
        same = np.array_equal(carry.GetPos(), cleared.GetPos())
        say("  step %d: iterations %d / %d, same %s" % (step, carry.last_stats()["pcg_iterations"],
                                                       cleared.last_stats()["pcg_iterations"], same))
        assert same == (step < 3), step
