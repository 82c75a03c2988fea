"""On-disk formats through the host-only C ABI (no GPU): OBJ meshes (Attribute::loadOBJ,
buffer.cu:203-306), BMP textures (Texture::loadBMP, buffer.cu:486-541) and their writer, PFM
output, and the SG lobe text format (SGBuffer::loadTXT, buffer.cu:706-732) on the reference's
own data file data/envmap2.txt (copied as tests/golden/envmap2.txt)."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from vcl_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F32 = np.float32


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libvcl_amd.so not built")
    return L.load()


def load_obj(lib, path):
    m = L.ObjMesh()
    st = lib.vcl_load_obj(str(path).encode(), C.byref(m))
    if st:
        return st, None
    arr = lambda p, n, k, dt: np.ctypeslib.as_array(p, shape=(n * k,)).astype(dt).reshape(n, k).copy() if n and p else None
    out = {"pos": arr(m.pos, m.posNum, 3, F32), "texel": arr(m.texel, m.texelNum, 2, F32), "normal": arr(m.normal, m.normalNum, 3, F32),
           "pi": arr(m.posIndex, m.indexNum, 3, np.int64), "ti": arr(m.texelIndex, m.indexNum, 3, np.int64),
           "ni": arr(m.normalIndex, m.indexNum, 3, np.int64)}
    lib.vcl_obj_free(C.byref(m))
    return 0, out


CUBE_V = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
CUBE_F = [(1, 2, 3), (1, 3, 4), (5, 8, 7), (5, 7, 6), (1, 5, 6), (1, 6, 2), (2, 6, 7), (2, 7, 3), (3, 7, 8), (3, 8, 4), (5, 1, 4), (5, 4, 8)]


def write_obj(path, form):
    lines = ["# cube written by the test", "mtllib cube.mtl", "o Cube"]
    lines += [f"v {x:.6f} {y:.6f} {z:.6f}" for x, y, z in CUBE_V]
    if form in ("vt", "vtn"):
        lines += [f"vt {u:.3f} {v:.3f}" for u, v in ((0, 0), (1, 0), (1, 1), (0, 1))]
    if form in ("vn", "vtn"):
        lines += ["vn 0 0 -1", "vn 0 0 1", "vn 0 -1 0", "vn 1 0 0", "vn 0 1 0", "vn -1 0 0"]
    lines += ["usemtl Material", "s off"]
    for k, f in enumerate(CUBE_F):
        t = [(k + j) % 4 + 1 for j in range(3)]
        n = k // 2 + 1
        if form == "v":
            c = [str(a) for a in f]
        elif form == "vt":
            c = [f"{a}/{b}" for a, b in zip(f, t)]
        elif form == "vn":
            c = [f"{a}//{n}" for a in f]
        else:
            c = [f"{a}/{b}/{n}" for a, b in zip(f, t)]
        lines.append("f " + " ".join(c))
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("form", ["v", "vt", "vn", "vtn"])
def test_obj_forms(lib, tmp_path, form):
    p = tmp_path / f"cube_{form}.obj"
    write_obj(p, form)
    st, m = load_obj(lib, p)
    assert st == 0
    np.testing.assert_array_equal(m["pos"], np.array(CUBE_V, F32))
    np.testing.assert_array_equal(m["pi"], np.array(CUBE_F) - 1)
    if "t" in form:
        assert m["texel"].shape == (4, 2)
        np.testing.assert_array_equal(m["ti"], np.array([[(k + j) % 4 for j in range(3)] for k in range(12)]))
    else:
        assert m["texel"] is None and m["ti"] is None
    if "n" in form:
        assert m["normal"].shape == (6, 3)
        np.testing.assert_array_equal(m["ni"], np.repeat(np.arange(6), 2)[:, None].repeat(3, 1))
    else:
        assert m["normal"] is None and m["ni"] is None


def test_obj_errors(lib, tmp_path):
    assert lib.vcl_load_obj(str(tmp_path / "missing.obj").encode(), C.byref(L.ObjMesh())) == 1
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1 2 3\n")  # texels exist: v/vt form required
    assert load_obj(lib, p)[0] == 1
    p.write_text("f 1 2 3\nv 0 0 0\nv 1 0 0\nv 0 1 0\n")  # a face before any position is skipped
    st, m = load_obj(lib, p)
    assert st == 0 and m["pi"] is None and m["pos"].shape == (3, 3)


def test_bmp_roundtrip_and_layout(lib, tmp_path):
    rng = np.random.default_rng(0)
    for w, h, c in ((8, 5, 3), (16, 4, 4), (12, 7, 1)):
        img = (rng.integers(0, 256, (h, w, c)) / 255.0).astype(F32)
        p = tmp_path / f"t{c}.bmp"
        assert lib.vcl_save_bmp(str(p).encode(), img.ctypes.data, w, h, c) == 0
        raw = p.read_bytes()
        # header fields the reference reads (buffer.cu:510-518), little endian
        assert raw[:2] == b"BM"
        off, = struct.unpack_from("<I", raw, 0x0A)
        W, H = struct.unpack_from("<ii", raw, 0x12)
        bpp, = struct.unpack_from("<H", raw, 0x1C)
        size, = struct.unpack_from("<I", raw, 0x22)
        assert (W, H, bpp, size) == (w, h, 8 * c, w * h * c)
        # pixel bytes: rows bottom first, each pixel's channels in reverse order (BGR)
        pix = np.frombuffer(raw[off: off + size], np.uint8).reshape(h, w, c)
        np.testing.assert_array_equal(pix[..., ::-1], np.round(img * 255).astype(np.uint8))
        im = L.Image()
        assert lib.vcl_load_bmp(str(p).encode(), C.byref(im)) == 0
        got = np.ctypeslib.as_array(im.data, shape=(w * h * c,)).reshape(h, w, c).copy()
        lib.vcl_image_free(C.byref(im))
        assert (im.width, im.height) == (0, 0)
        np.testing.assert_array_equal(got, (np.round(img * 255) / F32(255)).astype(F32))
    # not a BMP
    p = tmp_path / "x.bmp"
    p.write_bytes(b"PNG" + bytes(100))
    assert lib.vcl_load_bmp(str(p).encode(), C.byref(L.Image())) == 1


def test_pfm(lib, tmp_path):
    img = np.random.default_rng(1).normal(size=(6, 9, 4)).astype(F32)
    p = tmp_path / "x.pfm"
    assert lib.vcl_save_pfm(str(p).encode(), img.ctypes.data, 9, 6, 4) == 0
    raw = p.read_bytes()
    head = b"PF\n9 6\n-1.0\n"
    assert raw.startswith(head)
    np.testing.assert_array_equal(np.frombuffer(raw[len(head):], "<f4").reshape(6, 9, 3), img[..., :3])
    g = img[..., :1].copy()
    assert lib.vcl_save_pfm(str(p).encode(), g.ctypes.data, 9, 6, 1) == 0
    assert p.read_bytes().startswith(b"Pf\n9 6\n-1.0\n")


def test_sg_txt_reference_data(lib):
    """The reference's envmap2.txt (256 lobes), parsed independently as text."""
    path = os.path.join(ROOT, "tests", "golden", "envmap2.txt")
    rows = np.array([[float(t) for t in line.split()] for line in open(path) if line.strip()], F32)
    assert rows.shape == (256, 7)
    ax, sh, am = np.zeros(768, F32), np.zeros(256, F32), np.zeros(768, F32)
    assert lib.vcl_load_sg_txt(path.encode(), 256, ax.ctypes.data, sh.ctypes.data, am.ctypes.data) == 0
    np.testing.assert_array_equal(ax.reshape(256, 3), rows[:, :3])
    np.testing.assert_array_equal(sh, rows[:, 3])
    np.testing.assert_array_equal(am.reshape(256, 3), rows[:, 4:])
    # more lobes than the file holds: an error, as the reference's "Impossible to read"
    big = np.zeros(3 * 300, F32)
    assert lib.vcl_load_sg_txt(path.encode(), 300, big.ctypes.data, np.zeros(300, F32).ctypes.data, big.ctypes.data) == 1
