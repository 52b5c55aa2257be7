"""The .cube reader behind GailHipColor.lut3d (an app 3D LUT file): the Adobe Cube LUT
Specification 1.0 layout -- keywords, then size^3 lines of three numbers, red fastest --
with comments, blank lines, CRLF line ends and keywords of other tools, and its refusals."""
import ctypes as C

import numpy as np
import pytest

import libgail_amd as g


def _read(path):
    h = g.host()
    h.gail_read_cube_lut.restype = C.c_int64
    h.gail_read_cube_lut.argtypes = [C.c_char_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_int), C.c_char_p, C.c_int]
    out = np.zeros(3 * 17 ** 3, np.float32)
    size = C.c_int(0)
    msg = C.create_string_buffer(512)
    n = h.gail_read_cube_lut(str(path).encode(), out.ctypes.data, out.size, C.byref(size), msg, 512)
    if n < 0:
        raise ValueError(msg.value.decode())
    return size.value, out[:n].reshape(-1, 3)


def _table(n):
    """The identity LUT of size n plus a small twist, red fastest."""
    idx = np.arange(n)
    b, gg, r = np.meshgrid(idx, idx, idx, indexing="ij")
    t = np.stack([r, gg, b], -1).reshape(-1, 3) / (n - 1)
    return (t * 0.9 + 0.05 * np.roll(t, 1, axis=1)).astype(np.float32)


def test_reads_a_table(tmp_path):
    n = 5
    t = _table(n)
    lines = ["# made for the test", "TITLE \"twist\"", "", "LUT_3D_SIZE 5", "DOMAIN_MIN 0 0 0",
             "DOMAIN_MAX 1.0 1.0 1.0", "LUT_3D_INPUT_RANGE 0 1   # a keyword of another tool"]
    lines += ["%.9g %.9g %.9g" % tuple(v) + ("  # first" if i == 0 else "") for i, v in enumerate(t)]
    p = tmp_path / "twist.cube"
    p.write_bytes(("\r\n".join(lines) + "\r\n").encode())
    size, got = _read(p)
    assert size == n
    assert np.array_equal(got, t)


@pytest.mark.parametrize("text,why", [
    ("LUT_1D_SIZE 4\n0 0 0\n", "1D LUTs"),
    ("LUT_3D_SIZE 2\n" + "0 0 0\n" * 7, "7 table lines, 8 expected"),
    ("LUT_3D_SIZE 2\n" + "0 0 0\n" * 8 + "TITLE late\n", "after the table"),
    ("LUT_3D_SIZE 2\nDOMAIN_MAX 2 2 2\n" + "0 0 0\n" * 8, "domain 0..1"),
    ("LUT_3D_SIZE 2\n" + "0 0 0\n" * 7 + "0 0 x\n", "exactly three numbers"),
    ("LUT_3D_SIZE 2\n" + "0 0 0 0\n" * 8, "exactly three numbers"),
    ("LUT_3D_SIZE 1\n0 0 0\n", "2 to 256"),
    ("0 0 0\n" * 8, "no LUT_3D_SIZE"),
])
def test_refusals(tmp_path, text, why):
    p = tmp_path / "bad.cube"
    p.write_text(text)
    with pytest.raises(ValueError, match=why):
        _read(p)


def test_missing_file(tmp_path):
    with pytest.raises(ValueError, match="cannot be opened"):
        _read(tmp_path / "absent.cube")
