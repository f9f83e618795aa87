"""The C-ABI library: it loads without a GPU and exports every entry point
include/ipgpu.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ipgpu_[a-z_]+)\s*\(", txt)))


def test_ipgpu_exports_all_declared_symbols():
    path = os.path.join(ROOT, "ipole_amd", "lib", "libipgpu.so")
    if not os.path.exists(path):
        pytest.skip("libipgpu.so not built")
    lib = ctypes.CDLL(path)
    names = declared("ipgpu.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_no_device_no_context():
    """Without a visible GPU the library refuses a context instead of falling
    back to anything."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from ipole_amd import ipgpu as G
    with pytest.raises(G.IpgpuError):
        G.Context(0)
