"""The C-ABI library loads and exports every entry point declared in include/tlrgp.h.  CPU only: no
computation is launched."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tlrgp.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tlr_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "tlr_loglik" in names and "tlr_ctx_create" in names and "tlr_loglik_batch" in names


def test_library_exports_every_declared_symbol():
    import tlrgp
    lib = ctypes.CDLL(tlrgp._lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    # the Python binding covers the whole header
    assert set(_declared()) == set(tlrgp._lib.SIGNATURES)


def test_no_device_is_reported_not_faked():
    import tlrgp
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    S = np.random.default_rng(0).random((100, 2))
    y = np.zeros(100)
    with pytest.raises(tlrgp.TLRError, match="no HIP device"):
        tlrgp.TLRContext(S, y, 50)


def test_facade_rejects_nonuniform_bounds():
    import tlrgp
    S = np.random.default_rng(0).random((100, 2))
    with pytest.raises(ValueError):
        tlrgp.tlr_loglik(S, np.zeros(100), (1, 0.3, 0.05), [(0, 30), (30, 100)], "adaptive", 0.9)


def test_host_helpers_match_oracle():
    import tlrgp
    import tlr_oracle as O
    xy = np.random.default_rng(4).random((1000, 2))
    assert np.array_equal(tlrgp.morton_order(xy), O.morton_order(xy))
    assert tlrgp.tile_bounds(1001, 250) == O.tile_bounds(1001, 250)
    rs = np.tril(np.random.default_rng(1).integers(0, 9, (4, 4)), -1)
    info = tlrgp.rank_info(rs, rs, 1000, 250, True)
    assert info["stored_sigma"] == 4 * 250 ** 2 + 2 * 250 * rs.sum()
    assert info["stored_factor"] == 4 * 250 ** 2 + 2 * 250 * rs.sum()
