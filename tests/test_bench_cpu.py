"""Host-side pieces of bench.py: the slice count that picks the SpMV
instance, the instance name in the bench line, the traffic entries keyed by
the autotuner's pick, and that every traffic entry names a profile that
holds its figure."""
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_batch_slices_pad_each_cloth_to_four_slices():
    # 257 x 257 = 66,049 vertices: 1,033 slices of 64 rows, padded to 1,036
    assert bench.batch_slices(bench.members("single257", "mixed")) == 1036
    ms = bench.members("batch64", "mixed")
    n = bench.batch_slices(ms)
    assert n % 4 == 0 and n == 6248
    # one cloth of 64 vertices is one slice, padded to 4
    assert bench.batch_slices([(None, [0.0] * (3 * 64))]) == 4
    assert bench.batch_slices([(None, [0.0] * (3 * 65))]) == 4
    assert bench.batch_slices([(None, [0.0] * (3 * 64 * 5))]) == 8


@pytest.mark.parametrize("lanes,staged,slices,chunk", [
    (1, False, 1036, 8), (2, False, 1036, 4), (4, False, 1036, 2),
    (1, False, 6248, 4), (2, False, 6248, 2), (1, True, 1036, 4), (2, True, 6248, 2)])
def test_kernel_name_names_the_launched_instance(lanes, staged, slices, chunk, monkeypatch):
    monkeypatch.delenv("BW_ILP_SLICES", raising=False)
    name = bench.kernel_name(lanes, staged, False, slices)
    assert name.startswith("k_cg_spmv<float, %d, %s, false, %d>"
                           % (lanes, "true" if staged else "false", chunk))


def test_traffic_only_for_profiled_picks():
    b, src = bench.traffic_for("batch64", (1, False, False, {}))
    assert b > 0 and "k_cg_spmv_batch64.md" in src
    b, why = bench.traffic_for("batch64", (4, True, False, {}))
    assert b is None and "no counter profile" in why
    b, _ = bench.traffic_for("irregular", (1, True, True, {}))
    assert b is None


def test_traffic_entries_appear_in_their_profiles():
    for workload, entries in bench.TRAFFIC.items():
        for key, (nbytes, profile) in entries.items():
            text = open(os.path.join(ROOT, profile)).read()
            figures = {int(x.replace(",", "")) for x in re.findall(r"\*\*([\d,]+) B per launch", text)}
            assert int(nbytes) in figures, (workload, key, nbytes, profile)
