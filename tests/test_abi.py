"""The C ABI library loads and exports every function include/uwmesh.h declares; host-only entry
points agree with the oracle. No device compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "uwmesh.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(uw_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["uw_ctx_create", "uw_ctx_destroy", "uw_build_chunks_plain", "uw_sample_density",
                 "uw_chunk_build_host", "uw_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from underwaterworld_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from underwaterworld_amd import _lib
    assert set(declared_functions()) == set(_lib.SIGNATURES)


@pytest.mark.parametrize("seed", [0, 1, 42, 1234, 0xFFFFFFFF])
def test_engine_permutation_equals_oracle(seed):
    from underwaterworld_amd import permutation_table
    assert np.array_equal(permutation_table(seed), O.perm(seed))


def test_size_bounds():
    from underwaterworld_amd import _lib
    L = _lib.lib()
    assert L.uw_max_verts(12) == 5 * 12 * 13 * 13
    assert L.uw_max_inds(12) == 15 * 12 ** 3
    assert L.uw_max_verts(0) == 0


def test_reference_params():
    from underwaterworld_amd import _lib, params
    s = _lib.UwParams()
    _lib.lib().uw_params_reference(ctypes.byref(s), 77)
    ref = _lib.make_params(params.preset("reference", seed=77))
    assert bytes(s) == bytes(ref)


def test_ctx_create_rejects_bad_params_without_device():
    from underwaterworld_amd import _lib, params
    s = _lib.make_params(params.preset("reference", octaves=0))
    ctx = ctypes.c_void_p()
    assert _lib.lib().uw_ctx_create(ctypes.byref(s), 0, ctypes.byref(ctx)) == _lib.UW_ERR_INVALID_ARG
    assert not ctx.value
