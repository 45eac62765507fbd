"""The C ABI on the CPU: every function include/ccs.h declares is exported by
libccs.so and bound by the ctypes table with the same arity, and the calls that
touch no device (version, strides, sizes, argument checks) behave as documented."""
import ctypes
import os
import re

import pytest

from cnn_crf_stereo_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ccs.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^(?:const char\s*\*|int|size_t)\s+(ccs_\w+)\s*\(([^;]*)\)\s*;", text, flags=re.M):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_header_parses():
    d = declared()
    assert len(d) >= 25
    assert "ccs_slackprop_forward" in d and "ccs_correlation_forward" in d


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    d = declared()
    assert set(d) == set(capi.SIGNATURES)
    for name, nargs in d.items():
        assert len(capi.SIGNATURES[name][1]) == nargs, name


def test_host_only_calls():
    L = capi.lib()
    assert L.ccs_version().decode()
    assert [capi.feature_stride(c) for c in (1, 3, 8, 100, 101)] == [8, 8, 8, 104, 104]
    assert L.ccs_conv_packed_bytes(100, 100, 3, 3, capi.BF16) > 0
    assert L.ccs_conv_packed_bytes(100, 100, 3, 3, capi.FP32) >= L.ccs_conv_packed_bytes(100, 100, 3, 3, capi.BF16)


@pytest.mark.parametrize("name,args", [
    ("ccs_slackprop", (None, None, None, None, 4, 4, 8, 0.5, 2.0, 2, 1, None, None, None, None)),
    ("ccs_slackprop_decode_gap", (None, None, None, 4, 4, 8, 0.5, 2.0, 2, None, None)),
    ("ccs_correlation_volume", (None, None, None, 1, 4, 4, 8, 0, 1, 8, 1, None)),
    ("ccs_argmin", (None, None, 4, 4, 8, 0.0, 1.0, None)),
])
def test_null_arguments_are_rejected(name, args):
    rc = getattr(capi.lib(), name)(*args)
    assert rc == -1  # CCS_ERR_ARG, before any device work
    assert capi.last_error()
