"""The C-ABI library loads, exports every entry point include/*.h declares,
and rejects bad arguments with an error message (host-side checks only; no
compute calls without a GPU)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import liar_amd as la
from liar_amd import _native as N

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(lr_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "lr_trace_closest" in names and "lr_trace_any" in names and len(names) >= 10


@pytest.mark.parametrize("name", declared_symbols())
def test_library_exports(name):
    lib = C.CDLL(str(N.LIB_PATH))
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_bindings_cover_header():
    assert set(declared_symbols()) <= set(N.SIGNATURES)


def test_version():
    assert b"gfx950" in N.lib().lr_version()


def test_null_arguments_are_errors():
    L = N.lib()
    assert L.lr_context_create(0, None) == -1
    assert L.lr_trace_closest(None, None, 0, None, None) == -1
    assert b"NULL" in L.lr_last_error()
    assert L.lr_qbvh_get_info(None, None) == -1


def test_host_only_qbvh_cannot_be_traced():
    v, t = la.random_triangles(10, seed=1)
    b = la.Qbvh(None, v, t)
    rs = N.RaysSoa()
    hs = N.HitsSoa()
    rc = N.lib().lr_trace_closest(b.handle, C.byref(rs), 1, C.byref(hs), None)
    assert rc == -1 and b"host-only" in N.lib().lr_last_error()


def test_bad_geometry_rejected():
    v = np.zeros((3, 3), np.float32)
    with pytest.raises(la.LiarError, match="references vertex"):
        la.Qbvh(None, v, [[0, 1, 3]])
    with pytest.raises(la.LiarError, match="radius"):
        la.Qbvh(None, v[:0], np.zeros((0, 3)), [[0, 0, 0, -1.0]])
    with pytest.raises(la.LiarError, match="non-finite"):
        la.Qbvh(None, np.array([[0, 0, np.nan], [1, 0, 0], [0, 1, 0]], np.float32), [[0, 1, 2]])


def test_empty_scene_info():
    b = la.Qbvh(None, np.zeros((0, 3)), np.zeros((0, 3)))
    assert b.info()["n_nodes"] == 0


def test_generators_deterministic():
    a = la.random_rays(1000, seed=3)
    b = la.random_rays(1000, seed=3)
    np.testing.assert_array_equal(a, b)
    d = a[3:6]
    np.testing.assert_allclose(np.linalg.norm(d, axis=0), 1.0, atol=1e-6)
    assert (a[0:3] >= 0).all() and (a[0:3] < 1).all() and (a[6] == 0).all() and np.isinf(a[7]).all()
    v, t = la.random_triangles(1000, seed=2)
    c = v.reshape(-1, 3, 3).mean(1)
    assert (np.abs(v.reshape(-1, 3, 3) - c[:, None, :]) <= 0.02).all()
