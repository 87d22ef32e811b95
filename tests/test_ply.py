"""The .ply codec (src/ply.rs): round trip (port of src/ply.rs:192-238), the byte layout
against an independently assembled file, reading files whose properties are reordered or
carry extra elements, and the reference's error cases (host code, no GPU)."""
import struct

import numpy as np
import pytest

from splatrs_amd import abi


def five_gaussians():
    """The records of src/ply.rs:194-209."""
    g = np.zeros((5, 59), np.float32)
    for i in range(5):
        f = float(i)
        for k in range(16):
            for c in range(3):
                g[i, 11 + 3 * k + c] = f + k * 0.1 + c * 0.01
        g[i, 0:3] = [f, f + 0.5, f - 0.5]
        g[i, 3:6] = [-f, -f - 1.0, -f - 2.0]
        g[i, 6:10] = [1.0, 0.1 * f, 0.2 * f, 0.3 * f]
        g[i, 10] = f * 0.3
    return g


def test_roundtrip(tmp_path):
    g = five_gaussians()
    p = tmp_path / "m.ply"
    abi.ply_write(p, g, 3)
    back, deg = abi.ply_read(p)
    assert deg == 3 and back.shape == (5, 59)
    assert np.array_equal(back, g)


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_layout_matches_independent_writer(tmp_path, degree):
    g = five_gaussians()
    p = tmp_path / "m.ply"
    abi.ply_write(p, g, degree)
    k = (degree + 1) ** 2
    names = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    names += [f"f_rest_{i}" for i in range(3 * (k - 1))]
    names += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    want = b"ply\nformat binary_little_endian 1.0\nelement vertex 5\n"
    want += b"".join(f"property float {n}\n".encode() for n in names) + b"end_header\n"
    for r in g:
        vals = list(r[0:3]) + [0.0, 0.0, 0.0] + list(r[11:14])
        vals += [r[11 + 3 * kk + c] for c in range(3) for kk in range(1, k)]
        vals += [r[10]] + list(r[3:6]) + list(r[6:10])
        want += struct.pack(f"<{len(vals)}f", *vals)
    assert p.read_bytes() == want
    back, deg = abi.ply_read(p)
    assert deg == degree
    mask = np.zeros(59, bool)
    mask[:11 + 3 * k] = True
    assert np.array_equal(back[:, mask], g[:, mask]) and not back[:, ~mask].any()


def test_reads_reordered_properties_and_other_elements(tmp_path):
    names = ["rot_3", "opacity", "scale_2", "x", "f_dc_1", "y", "rot_0", "z", "scale_0", "f_dc_0",
             "rot_1", "scale_1", "f_dc_2", "rot_2", "extra"]
    vals = {n: float(i) + 0.25 for i, n in enumerate(names)}
    hdr = "ply\nformat binary_little_endian 1.0\ncomment made elsewhere\nelement vertex 2\n"
    hdr += "".join(f"property float32 {n}\n" for n in names)
    hdr += "element face 0\nproperty list uchar int vertex_indices\nend_header\n"
    body = struct.pack(f"<{2 * len(names)}f", *([vals[n] for n in names] * 2))
    p = tmp_path / "r.ply"
    p.write_bytes(hdr.encode() + body)
    g, deg = abi.ply_read(p)
    assert deg == 0 and g.shape == (2, 59)
    assert list(g[1, 0:3]) == [vals["x"], vals["y"], vals["z"]]
    assert list(g[1, 3:6]) == [vals["scale_0"], vals["scale_1"], vals["scale_2"]]
    assert list(g[1, 6:10]) == [vals[f"rot_{i}"] for i in range(4)]
    assert g[1, 10] == vals["opacity"] and list(g[1, 11:14]) == [vals[f"f_dc_{i}"] for i in range(3)]


def _header(fmt="binary_little_endian 1.0", count="1", names=None, ty="float"):
    names = names or ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0",
                      "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    h = f"ply\nformat {fmt}\nelement vertex {count}\n"
    h += "".join(f"property {ty} {n}\n" for n in names) + "end_header\n"
    return h.encode(), len(names)


@pytest.mark.parametrize("case,msg", [
    ("ascii", "only binary_little_endian"),
    ("no_end", "missing end_header"),
    ("double", "only float is supported"),
    ("missing", "missing property opacity"),
    ("rest", "unsupported number of f_rest"),
    ("short", "shorter than the header"),
    ("count", "bad vertex count"),
])
def test_format_errors(tmp_path, case, msg):
    p = tmp_path / "bad.ply"
    if case == "ascii":
        h, m = _header(fmt="ascii 1.0")
    elif case == "double":
        h, m = _header(ty="double")
    elif case == "missing":
        h, m = _header(names=["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2"])
    elif case == "rest":
        h, m = _header(names=["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0",
                              "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3",
                              "f_rest_0"])
    elif case == "count":
        h, m = _header(count="-3")
    else:
        h, m = _header()
    body = struct.pack(f"<{m}f", *([0.0] * m))
    if case == "no_end":
        data = h.replace(b"end_header\n", b"") + body
    elif case == "short":
        data = h + body[:-4]
    else:
        data = h + body
    p.write_bytes(data)
    with pytest.raises(abi.PlyError) as ei:
        abi.ply_read(p)
    assert ei.value.status == abi.ERR_FORMAT and msg in str(ei.value)


def test_io_errors(tmp_path):
    with pytest.raises(abi.PlyError) as ei:
        abi.ply_read(tmp_path / "absent.ply")
    assert ei.value.status == abi.ERR_IO
    with pytest.raises(abi.PlyError) as ei:
        abi.ply_write(tmp_path / "no_dir" / "m.ply", five_gaussians(), 3)
    assert ei.value.status == abi.ERR_IO
