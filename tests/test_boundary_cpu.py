"""C-ABI boundary and host logic without a GPU: the library loads, exports every
symbol declared in include/*.h, reports errors as status codes, and the
Python host mirrors the reference's engine surface."""

import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(cx_\w+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    from covertreex_amd import _ffi

    lib = _ffi.load()
    declared = _declared_symbols()
    assert len(declared) >= 15
    for s in sorted(declared):
        assert hasattr(lib, s), f"{s} declared in include/ but not exported"
    assert declared == set(_ffi.SIGNATURES)


def test_status_codes_without_device():
    from covertreex_amd import _ffi

    lib = _ffi.load()
    if _ffi.device_count() > 0:
        pytest.skip("a GPU is visible; this checks the no-device path")
    v = np.zeros((4, 2), np.float32)
    p = np.ones(4, np.float32)
    x = np.zeros((4, 3), np.float32)
    out = C.c_void_p()
    st = lib.cx_tree_build_residual(_ffi.ptr(v), _ffi.ptr(p), _ffi.ptr(x), 4, 2, 3, 1.0, None, 0,
                                    1, 512, 0, 1.0, C.byref(out))
    assert st == _ffi.CX_ERR_NO_DEVICE
    assert b"no HIP device" in lib.cx_last_error()
    # argument errors are reported before any device work
    st = lib.cx_tree_build_residual(None, None, None, 4, 2, 3, 1.0, None, 0, 1, 512, 0, 1.0,
                                    C.byref(out))
    assert st == _ffi.CX_ERR_INVALID
    with pytest.raises(_ffi.HipError):
        _ffi.check(st)
    st = lib.cx_tree_knn_residual(None, None, 0, 5, 0, 1, None, None, None)
    assert st == _ffi.CX_ERR_INVALID
    assert lib.cx_tree_set_screen_segment(None, 1024) == _ffi.CX_ERR_INVALID
    assert lib.cx_tree_get_screen_segment(None, None) == _ffi.CX_ERR_INVALID


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from covertreex_amd import _ffi

    monkeypatch.setattr(_ffi, "_lib", None)
    with pytest.raises(ImportError, match="no CPU fallback"):
        _ffi.load(str(tmp_path / "absent.so"))


def test_format_knn_output_shapes():
    from covertreex_amd.engine import _format_knn_output

    idx = np.arange(6).reshape(2, 3)
    dst = np.arange(6, dtype=np.float32).reshape(2, 3)
    assert _format_knn_output(idx, dst, return_distances=False) is idx
    i1, d1 = _format_knn_output(idx[:1], dst[:1], return_distances=True)
    assert i1.shape == (3,) and d1.shape == (3,)
    single = _format_knn_output(idx[:1, :1], dst[:1, :1], return_distances=False)
    assert single == 0


def test_engine_registry_and_query_payload_check():
    import covertreex_amd as cx
    from covertreex_amd.engine import CoverTree, HipEngine

    assert cx.get_engine("hip").name == "hip"
    with pytest.raises(ValueError, match="Unknown engine"):
        cx.get_engine("python-numba")
    tree = CoverTree(engine=HipEngine(), handle=None, metric="residual_correlation")
    with pytest.raises(ValueError, match="integer query payloads"):
        tree.knn(np.zeros((3, 1), np.float32), k=2)


def test_kernels_closed_forms():
    from covertreex_amd.kernels import Matern52, RBF

    a = np.array([[0.0, 0.0]])
    b = np.array([[3.0, 4.0]])
    assert RBF(lengthscale=5.0)(a, b)[0, 0] == pytest.approx(np.exp(-0.5))
    r = np.sqrt(5.0)
    assert Matern52(lengthscale=5.0, variance=2.0)(a, b)[0, 0] == pytest.approx(
        2.0 * (1 + r + r * r / 3) * np.exp(-r))
    assert RBF().kernel_type == 0 and Matern52().kernel_type == 1


def test_backend_container_validation():
    from covertreex_amd.residual import ResidualBackend, inducing_indices

    with pytest.raises(ValueError):
        ResidualBackend(v_matrix=np.zeros(3), p_diag=np.zeros(3), kernel_diag=np.zeros(3),
                        kernel_points_f32=np.zeros((3, 2)))
    b = ResidualBackend(v_matrix=np.zeros((3, 2)), p_diag=np.ones(3), kernel_diag=np.ones(3),
                        kernel_points_f32=np.zeros((3, 2)))
    assert b.v_matrix.dtype == np.float32 and b.rank == 2
    # host_backend.py:191-199: sorted default_rng(seed).choice
    idx = inducing_indices(100, 10, 42)
    assert np.array_equal(idx, np.sort(np.random.default_rng(42).choice(100, 10, replace=False)))
    assert np.array_equal(inducing_indices(5, 10, 0), np.arange(5))


def test_bench_rejects_multi_gpu():
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
