"""The C ABI library loads and exports every entry point declared in include/vcl.h; host-only
entry points and argument validation work without a GPU (no compute is launched here)."""
import ctypes as C
import os
import re

import pytest

from vcl_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_][a-z0-9_]*\s*\**\s*\b(vcl_[a-z0-9_]+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libvcl_amd.so not built")
    return L.load()


def test_header_functions_exported(lib):
    names = declared_functions()
    assert len(names) >= 40, names
    raw = C.CDLL(L.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(raw, n)]
    assert not missing, f"declared but not exported: {missing}"
    # and the Python binding knows every declared entry point
    assert names <= set(L.SIGNATURES), sorted(names - set(L.SIGNATURES))


def test_host_only_entry_points(lib):
    assert lib.vcl_texture_mip_levels(2048, 2048, 12) == 12
    assert lib.vcl_texture_mip_levels(512, 512, 100) == 10
    assert lib.vcl_texture_mip_levels(96, 64, 16) == 6
    assert lib.vcl_texture_mip_levels(64, 64, 0) == 1
    n = lib.vcl_rasterize_workspace_size(1024, 1024, 1, 49922, 99840)
    assert 0 < n < 1 << 31
    assert lib.vcl_rasterize_workspace_size(0, 10, 1, 3, 1) == 0
    assert lib.vcl_spherical_gaussian_workspace_size(1024, 1024, 1, 4, 128) > 0
    assert lib.vcl_loss_workspace_size(1 << 20) > 0
    assert lib.vcl_status_string(L.VCL_ERR_INVALID) == b"invalid argument"
    assert lib.vcl_build_info().startswith(b"vcl_amd")


def test_argument_validation_without_gpu(lib):
    # NULL / inconsistent parameters are rejected before any HIP call
    assert lib.vcl_project_forward(None, None) == L.VCL_ERR_INVALID
    p = L.RasterizeParams()
    p.width, p.height, p.depth = 0, 16, 1
    assert lib.vcl_rasterize_forward(C.byref(p), None) == L.VCL_ERR_INVALID
    t = L.TexturemapParams()
    t.width = t.height = t.depth = 1
    t.miplevel = 17
    assert lib.vcl_texturemap_forward(C.byref(t), None) == L.VCL_ERR_INVALID
    s = L.SphericalGaussianParams()
    s.channel = 9
    assert lib.vcl_spherical_gaussian_forward(C.byref(s), None) == L.VCL_ERR_INVALID
    d = L.FitDesc()
    h = C.c_void_p()
    assert lib.vcl_fit_create(C.byref(d), C.byref(h)) == L.VCL_ERR_INVALID
    assert lib.vcl_fit_step(None, None) == L.VCL_ERR_INVALID
    # a workspace smaller than required is refused
    r = L.RasterizeParams(64, 64, 1, 0, 2 / 64, 2 / 64, 1, 1, 1, None, 3, 1, 1, 16)
    assert lib.vcl_rasterize_forward(C.byref(r), None) == L.VCL_ERR_OOM
