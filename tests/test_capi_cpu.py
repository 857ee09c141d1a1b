"""CPU-only checks of the product library: it builds, loads, exports every
entry point include/bw_cloth.h declares, and its host-side workload helpers
behave.  No device computation here."""
import ctypes
import os
import re

import numpy as np
import pytest

from acousticsim_amd import capi
from acousticsim_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bw_cloth.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bw_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libbwcloth.so not built")
    lib = ctypes.CDLL(capi.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(capi.EXPORTS) == syms


def test_last_error_is_callable_without_device():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libbwcloth.so not built")
    assert isinstance(capi.lib().bw_last_error(), bytes)


def test_splitmix_stream_known_values():
    # splitmix64 reference outputs for seed 0 (first three draws)
    z = wl.splitmix64_stream(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_jitter_is_bounded_and_seeded():
    c = wl.grid_config(21)
    a = wl.jittered_positions(c, 1)
    b = wl.jittered_positions(c, 1)
    d = wl.jittered_positions(c, 2)
    assert np.array_equal(a, b) and not np.array_equal(a, d)
    z = a.reshape(-1, 3)[:, 2]
    assert np.abs(z).max() <= 1e-3 and np.abs(z).max() > 9e-4
    assert np.array_equal(a.reshape(-1, 3)[:, :2], wl.grid_rest_positions(c).reshape(-1, 3)[:, :2])


def test_batch64_members_in_range():
    m = wl.batch64_members()
    assert len(m) == 64
    sides = [s for s, _, _ in m]
    scales = [sc for _, sc, _ in m]
    assert min(sides) >= 17 and max(sides) <= 129
    assert min(scales) >= 0.1 and max(scales) <= 10
    total_v = sum(s * s for s in sides)
    assert 2e5 < total_v < 7e5


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_device_fails_loudly():
    """No CPU fallback: without a GPU the library refuses with BW_ERR_DEVICE."""
    with pytest.raises(capi.BwError) as e:
        capi.BaraffClothHip(wl.grid_config(5), None)
    assert e.value.code == -2


@pytest.mark.parametrize("sub", [[65536, 65536], [1e300, 2], [-1, 5], [0.5, 4], [16385, 16385]])
def test_out_of_range_grid_is_a_config_error(sub):
    """Grid sizes whose vertex, triangle or edge counts would overflow are
    rejected while the configuration is read, before any device call, with
    BW_ERR_ARG (the host arrays are never sized from a wrapped count)."""
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libbwcloth.so not built")
    c = wl.grid_config(5)
    c["subdivision"] = sub
    with pytest.raises(capi.BwError) as e:
        capi.BaraffClothHip(c, None)
    assert e.value.code == -1, e.value
    assert "subdivision" in str(e.value)
