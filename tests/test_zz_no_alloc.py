"""bw_batch_step_noalloc against bw_batch_step through the C ABI.

The entry point that allocates nothing runs the same kernels with the
solver's scratch in the library's static device array; its results must be
bit-identical to bw_batch_step's with the same SpMV variant (fixed here
through BW_SPMV_*), on two calls in a row, the second on a smaller batch
(the array then still holds the first call's contents), and again on the
first batch after the second.  Batches whose scratch exceeds the array, and
batches the entry point does not cover, are refused with an error."""
import sys

import numpy as np
import pytest

from acousticsim_amd import workloads as wl

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _members(spec, precision):
    out = []
    for side, scale, seed in spec:
        c = wl.grid_config(side, stiffness_scale=scale)
        c["precision"] = precision
        out.append((c, wl.jittered_positions(c, seed)))
    return out


BIG = [(9, 1.0, 61), (45, 3.0, 62), (17, 0.2, 63), (64, 1.0, 64), (30, 7.0, 65)]
SMALL = [(21, 2.0, 66), (12, 0.5, 67)]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_step_equals_batch_step(precision, monkeypatch):
    from acousticsim_amd import capi

    monkeypatch.setenv("BW_SPMV_LPR", "1")
    monkeypatch.setenv("BW_SPMV_STAGE", "0")
    monkeypatch.setenv("BW_SPMV_NT", "0")
    big, small = _members(BIG, precision), _members(SMALL, precision)

    def pair(ms):
        return (capi.ClothBatch([c for c, _ in ms], [x for _, x in ms]),
                capi.ClothBatch([c for c, _ in ms], [x for _, x in ms]))

    ref_big, na_big = pair(big)
    ref_small, na_small = pair(small)
    for label, ref, na, ms, steps in (("big", ref_big, na_big, big, 3),
                                      ("small", ref_small, na_small, small, 2),
                                      ("big again", ref_big, na_big, big, 2)):
        say("case noalloc %s %s: %d cloths, %d steps" % (precision, label, len(ms), steps))
        ref.step(steps)
        na.step_noalloc(steps)
        for i in range(len(ms)):
            a, b = ref.get_pos(i), na.get_pos(i)
            sa, sb = ref.stats(i), na.stats(i)
            assert np.array_equal(a, b), (label, i)
            assert sa["pcg_iterations"] == sb["pcg_iterations"], (label, i)
            assert sa["rel_residual"] == sb["rel_residual"], (label, i)
            assert sb["rel_residual"] <= 1e-6
            assert np.abs(b - ms[i][1]).max() > 0


def test_noalloc_refusals():
    from acousticsim_amd import capi

    say("case noalloc refusals")
    # reference-semantics batches are not covered
    c = wl.grid_config(9)
    c["solver_mode"] = "reference"
    c["precision"] = "fp64"
    b = capi.ClothBatch([c], [wl.jittered_positions(c, 1)])
    with pytest.raises(capi.BwError) as e:
        b.step_noalloc(1)
    assert e.value.code == -4
    # pending contacts or perturbations need bw_batch_step
    c = wl.grid_config(9)
    b = capi.ClothBatch([c], [wl.jittered_positions(c, 2)])
    b.apply_perturb(0, 3, [0.1, 0.0, 0.0])
    with pytest.raises(capi.BwError) as e:
        b.step_noalloc(1)
    assert e.value.code == -4
    b.step(1)  # consumes the perturbation
    b.step_noalloc(1)
    # 9,000 small cloths need 4.6 MB of solver scratch, more than the 4 MiB
    c = wl.grid_config(3)
    many = capi.ClothBatch([c] * 9000, [None] * 9000)
    with pytest.raises(capi.BwError) as e:
        many.step_noalloc(1)
    assert e.value.code == -5
    assert "exceeds" in str(e.value)
