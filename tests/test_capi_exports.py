"""The C-ABI libraries load and export every symbol their headers declare
(no GPU call is made)."""
import ctypes
import os
import re

from tests.conftest import ROOT


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(feat_[a-z0-9_]+)\s*\(", txt)))


def test_kernel_library_exports(feat):
    k, _ = feat.libs()
    names = declared("feat3amd.h")
    assert len(names) > 40
    missing = [n for n in names if not hasattr(k, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's entry points
    assert sorted(n for n, _, _ in feat.KERNEL_API) == names


def test_host_library_exports(feat):
    _, h = feat.libs()
    names = declared("feat3amd_host.h")
    missing = [n for n in names if not hasattr(h, n)]
    assert not missing, missing
    assert sorted(n for n, _, _ in feat.HOST_API) == names


def test_library_is_gfx950_code_object(feat):
    # the kernel library carries a gfx950 offload bundle
    data = open(feat.KERNEL_LIB, "rb").read()
    assert b"gfx950" in data


def test_no_oracle_linkage(feat):
    # the product libraries never link the oracle or the reference
    for path in (feat.KERNEL_LIB, feat.HOST_LIB):
        data = open(path, "rb").read()
        assert b"liboracle" not in data and b"libfeatref" not in data and b"orc_" not in data
