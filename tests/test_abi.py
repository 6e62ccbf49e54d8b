"""The C ABI library loads and exports every entry point include/simkit_amd.h declares (no GPU needed)."""

import ctypes
import os
import re

import pytest

from simkit_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "simkit_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(sk_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_symbols():
    syms = declared_symbols()
    assert len(syms) >= 20
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} is not built (python -m simkit_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_host_only_calls():
    lib = _lib.load()
    assert lib.sk_version() == 1
    assert lib.sk_status_string(_lib.SK_ERR_LS_FAIL) == b"line search failed"
    # argument validation happens before any device work
    assert lib.sk_mesh_create(None, None, 0, 0, ctypes.byref(ctypes.c_void_p())) == _lib.SK_ERR_INVALID
    assert b"empty mesh" in lib.sk_last_error()
    assert lib.sk_psd_project(None, 0, 9, 7, None, None) == _lib.SK_ERR_INVALID


def _coarse_order(X, per_agg):
    import numpy as np
    order = np.empty(len(X), np.int32)
    Xc = np.ascontiguousarray(X, dtype=np.float64)
    n_agg = ctypes.c_int()
    ptr = np.empty(len(X) + 1, np.int32)
    _lib.call("sk_coarse_order", Xc.ctypes.data_as(ctypes.c_void_p), len(X), per_agg,
              order.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n_agg), ptr.ctypes.data_as(ctypes.c_void_p))
    _coarse_order.n_agg = n_agg.value
    _coarse_order.ptr = ptr[:n_agg.value + 1].copy()
    return order


def test_coarse_order_groups_spatial_bins():
    """sk_coarse_order (host only): a permutation whose runs are the two-level preconditioner's
    aggregates, checked against an independent numpy binning of the same rule (each axis of the
    bounding box cut evenly into round(extent / s) bins, s the edge of a cube holding per_agg
    vertices at the mean density), on a cube, a jittered cube and a slender beam."""
    import numpy as np
    from simkit_amd.mesh import beam, jitter, unit_cube

    def bins(X, per_agg):
        lo, hi = X.min(0), X.max(0)
        ext = hi - lo
        box = np.prod(np.maximum(ext, 1e-6 * ext.max()))
        s = np.cbrt(box / max(1.0, len(X) / per_agg))
        nb = np.maximum(1, np.rint(ext / s)).astype(np.int64)
        sa = np.where(ext > 0, ext / nb, 1.0)
        ix = np.minimum(nb - 1, np.maximum(0, ((X - lo) / sa).astype(np.int64)))
        return (ix[:, 0] * nb[1] + ix[:, 1]) * nb[2] + ix[:, 2]

    X12, _ = unit_cube(12)
    Xb, _ = beam(30, 4, 0.025)
    cases = [(X12, 27), (jitter(X12, 1 / 12, 0.2, seed=3), 64), (Xb, 50), (X12, 1), (X12, 10 ** 6)]
    for X, per_agg in cases:
        order = _coarse_order(X, per_agg)
        assert np.array_equal(np.sort(order), np.arange(len(X)))
        key = bins(X, per_agg)
        k = key[order]
        assert np.all(np.diff(k) >= 0)                           # grouped by bin, bins ascending
        same = np.diff(k) == 0
        assert np.all(np.diff(order)[same] > 0)                   # ascending inside a bin
        assert len(np.unique(key)) == len(np.unique(k)) == _coarse_order.n_agg
        ptr = _coarse_order.ptr                                   # the runs' starts: one bin each
        assert ptr[0] == 0 and ptr[-1] == len(X) and np.all(np.diff(ptr) > 0)
        assert all(len(np.unique(k[ptr[a]:ptr[a + 1]])) == 1 for a in range(len(ptr) - 1))
        assert np.all(np.diff(k[ptr[:-1]]) > 0)
    # 13^3 vertices / 27 = 81.4 bins -> edge 0.231 -> round(4.33) = 4 per axis
    assert len(np.unique(bins(X12, 27))) == 64
    # 775 vertices / 50 = 15.5 bins in a 0.75 x 0.1 x 0.1 box -> edge 0.0785 -> 10 x 1 x 1
    assert len(np.unique(bins(Xb, 50))) == 10


def test_coarse_order_rejects_bad_input():
    import numpy as np
    X = np.zeros((4, 3))
    X[1, 0] = np.nan
    with pytest.raises(_lib.SimkitAmdError, match="non-finite"):
        _coarse_order(X, 2)
    with pytest.raises(_lib.SimkitAmdError, match="vertices per aggregate"):
        _coarse_order(np.zeros((4, 3)), 0)
