"""COLMAP sparse models (src/colmap/*.rs) and their conversion to training cameras
(src/dataset.rs:83-129, 213-237), host code:

- binary and text round trips of a synthetic model (ring cameras, the reference's
  tests/formats.rs:54-85 with every camera model's parameter count), exact;
- binary files packed here from COLMAP's layout, with 2D observations and point tracks
  that must be skipped;
- the text reader's header detection (observation lines present, absent, or numeric in
  their tenth token), comments and blank lines;
- camera_from_colmap against closed forms (identity, a quarter turn about z, an
  unnormalised and a zero quaternion) and against the cameras the model was built from;
- match_image_size: ports of src/dataset.rs:271-305;
- malformed inputs and their error codes.
"""
import struct

import numpy as np
import pytest

from splatrs_amd import abi, scenes
from splatrs_amd import records as rec

MODEL_ID = {name: k for k, (name, _) in enumerate(rec.COLMAP_MODELS)}


def rotation_to_qvec(m):
    """Shepperd's method (as src/colmap/mod.rs:241-272), test-side: builds the model."""
    tr = m[0][0] + m[1][1] + m[2][2]
    if tr > 0:
        s = np.sqrt(tr + 1.0) * 2.0
        q = [0.25 * s, (m[2][1] - m[1][2]) / s, (m[0][2] - m[2][0]) / s, (m[1][0] - m[0][1]) / s]
    elif m[0][0] > m[1][1] and m[0][0] > m[2][2]:
        s = np.sqrt(1.0 + m[0][0] - m[1][1] - m[2][2]) * 2.0
        q = [(m[2][1] - m[1][2]) / s, 0.25 * s, (m[0][1] + m[1][0]) / s, (m[0][2] + m[2][0]) / s]
    elif m[1][1] > m[2][2]:
        s = np.sqrt(1.0 + m[1][1] - m[0][0] - m[2][2]) * 2.0
        q = [(m[0][2] - m[2][0]) / s, (m[0][1] + m[1][0]) / s, 0.25 * s, (m[1][2] + m[2][1]) / s]
    else:
        s = np.sqrt(1.0 + m[2][2] - m[0][0] - m[1][1]) * 2.0
        q = [(m[1][0] - m[0][1]) / s, (m[0][2] + m[2][0]) / s, (m[1][2] + m[2][1]) / s, 0.25 * s]
    q = np.array(q)
    return q / np.linalg.norm(q)


def synthetic_model(cams):
    cameras = np.zeros(len(cams) + 2, rec.COLMAP_CAMERA)
    images = np.zeros(len(cams), rec.COLMAP_IMAGE)
    for k, c in enumerate(cams):
        cameras[k]["id"] = 10 + k
        cameras[k]["model"] = MODEL_ID["PINHOLE"]
        cameras[k]["width"], cameras[k]["height"] = int(c["width"]), int(c["height"])
        cameras[k]["num_params"] = 4
        cameras[k]["params"][:4] = [c["fx"], c["fy"], c["cx"], c["cy"]]
        rot = np.asarray(c["rot"], np.float64).reshape(3, 3)
        images[k]["id"] = 100 + k
        images[k]["camera_id"] = 10 + k
        images[k]["qvec"] = rotation_to_qvec(rot)
        images[k]["tvec"] = np.asarray(c["trans"], np.float64)
        images[k]["name"] = f"view_{k:03d}.png".encode()
    # two more cameras: a distortion model and a simple pinhole with awkward numbers
    cameras[-2]["id"], cameras[-2]["model"] = 3, MODEL_ID["OPENCV"]
    cameras[-2]["width"], cameras[-2]["height"], cameras[-2]["num_params"] = 640, 480, 8
    cameras[-2]["params"][:8] = [500.25, 501.5, 320.0, 240.0, -0.1, 0.01, 1e-7, -2.5e-9]
    cameras[-1]["id"], cameras[-1]["model"] = 1, MODEL_ID["SIMPLE_PINHOLE"]
    cameras[-1]["width"], cameras[-1]["height"], cameras[-1]["num_params"] = 33, 17, 3
    cameras[-1]["params"][:3] = [0.1 + 0.2, 1e300, -0.0]
    points = np.zeros(3, rec.COLMAP_POINT)
    points["id"] = [7, 2, 2**40]
    points["xyz"] = [[0.1, 0.2, 0.3], [-0.5, 0.0, 0.25], [1e-12, -3.3333333333333335, 5e20]]
    points["rgb"] = [[1, 2, 3], [200, 100, 50], [255, 0, 128]]
    points["error"] = [0.5, 1.25, 0.0]
    return cameras, images, points


def sorted_by_id(c):
    return c[np.argsort(c["id"], kind="stable")]


@pytest.mark.parametrize("binary", [True, False])
def test_round_trip_is_exact(tmp_path, binary):
    cams = scenes.small_scene_cameras(5)
    cameras, images, points = synthetic_model(cams)
    abi.colmap_write(tmp_path, cameras, images, points, binary=binary)
    for stem in ("cameras", "images", "points3D"):
        assert (tmp_path / f"{stem}.{'bin' if binary else 'txt'}").is_file()
    c, i, p = abi.colmap_read(tmp_path)
    # the writers order cameras by id; images and points keep their order
    assert c.tobytes() == sorted_by_id(cameras).tobytes()
    assert i.tobytes() == images.tobytes()
    assert p.tobytes() == points.tobytes()
    for k, cam in enumerate(cams):
        back = abi.camera_from_colmap(c, i[k])
        assert int(back["width"]) == int(cam["width"]) and back["fx"] == cam["fx"]
        assert back["cx"] == cam["cx"] and back["cy"] == cam["cy"]
        assert np.array_equal(back["trans"], cam["trans"])
        assert np.abs(back["rot"] - cam["rot"]).max() < 1e-6


def test_text_numbers_are_shortest_without_exponent(tmp_path):
    cameras, images, points = synthetic_model(scenes.small_scene_cameras(1))
    abi.colmap_write(tmp_path, cameras, images, points, binary=False)
    lines = (tmp_path / "cameras.txt").read_text().splitlines()
    assert lines[2] == "1 SIMPLE_PINHOLE 33 17 0.30000000000000004 1" + "0" * 300 + " -0"
    assert lines[3].startswith("3 OPENCV 640 480 500.25 501.5 320 240 -0.1 0.01 0.0000001 -0.0000000025")
    pts = (tmp_path / "points3D.txt").read_text().splitlines()
    assert pts[4] == ("1099511627776 0.000000000001 -3.3333333333333335 500000000000000000000 "
                      "255 0 128 0")
    imgs = (tmp_path / "images.txt").read_text()
    assert imgs.count("\n\n") == 1  # one image, each followed by an empty observation line


def test_binary_layout_with_observations_and_tracks(tmp_path):
    cam = struct.pack("<Q", 1) + struct.pack("<IiQQ", 5, 0, 800, 600) + struct.pack("<3d", 700.0, 400.0, 300.0)
    img = struct.pack("<Q", 1) + struct.pack("<I4d3dI", 9, 1.0, 0.0, 0.0, 0.0, 0.5, -1.0, 2.0, 5)
    img += b"dir/a b.jpg\0" + struct.pack("<Q", 2) + struct.pack("<ddQ", 1.5, 2.5, 7) * 2
    pts = struct.pack("<Q", 2)
    pts += struct.pack("<Q3d3Bd", 11, 1.0, 2.0, 3.0, 10, 20, 30, 0.75) + struct.pack("<Q", 3)
    pts += struct.pack("<II", 9, 0) * 3
    pts += struct.pack("<Q3d3Bd", 12, -1.0, -2.0, -3.0, 1, 2, 3, 0.25) + struct.pack("<Q", 0)
    (tmp_path / "cameras.bin").write_bytes(cam)
    (tmp_path / "images.bin").write_bytes(img)
    (tmp_path / "points3D.bin").write_bytes(pts)
    (tmp_path / "cameras.txt").write_text("not read: the .bin files win\n")
    c, i, p = abi.colmap_read(tmp_path)
    assert len(c) == 1 and int(c[0]["id"]) == 5 and int(c[0]["model"]) == 0
    assert (int(c[0]["width"]), int(c[0]["height"]), int(c[0]["num_params"])) == (800, 600, 3)
    assert list(c[0]["params"][:3]) == [700.0, 400.0, 300.0]
    assert int(i[0]["id"]) == 9 and int(i[0]["camera_id"]) == 5 and i[0]["name"] == b"dir/a b.jpg"
    assert list(i[0]["tvec"]) == [0.5, -1.0, 2.0]
    assert list(p["id"]) == [11, 12] and p[0]["rgb"].tolist() == [10, 20, 30]
    assert p[1]["error"] == 0.25 and list(p[1]["xyz"]) == [-1.0, -2.0, -3.0]
    cam0 = abi.camera_from_colmap(c, i[0])
    assert (cam0["fx"], cam0["fy"], cam0["cx"], cam0["cy"]) == (700.0, 700.0, 400.0, 300.0)
    assert np.array_equal(cam0["rot"], np.eye(3, dtype=np.float32).ravel())


def test_text_header_detection(tmp_path):
    (tmp_path / "cameras.txt").write_text(
        "# comment\n\n  2 PINHOLE 64 48 60 61 32 24  \n1 SIMPLE_PINHOLE 10 10 +5 5 5\n")
    (tmp_path / "images.txt").write_text(
        "# IMAGE_ID, QW, QX, QY, QZ, TX, TY, TZ, CAMERA_ID, NAME\n"
        "1 1 0 0 0 0 0 0 2 first.png\n"
        "1.5 2.5 -1 3.5 4.5 7\n"                       # observations
        "2 0.7071067811865476 0 0 0.7071067811865476 1 2 3 1 second.png\n"
        # no observation line after the second header
        "3 1 0 0 0 0 0 0 2 0.5 1 2 3 4 5 6 7 8 9\n"    # tenth token numeric: not a header
        "4 1 0 0 0 -0 inf 0 2 fourth.png\n")
    (tmp_path / "points3D.txt").write_text("# none\n")
    c, i, p = abi.colmap_read(tmp_path)
    assert list(c["id"]) == [2, 1] and c[1]["params"][0] == 5.0
    assert [n.decode() for n in i["name"]] == ["first.png", "second.png", "fourth.png"]
    assert list(i["id"]) == [1, 2, 4] and np.isinf(i[2]["tvec"][1])
    assert len(p) == 0
    quarter = abi.camera_from_colmap(c, i[1])  # a quarter turn about z
    want = np.array([0, -1, 0, 1, 0, 0, 0, 0, 1], np.float32)
    assert np.abs(quarter["rot"] - want).max() < 1e-6
    assert list(quarter["trans"]) == [1.0, 2.0, 3.0]


def test_camera_conversion_closed_forms():
    cams = np.zeros(2, rec.COLMAP_CAMERA)
    cams[0]["id"], cams[0]["model"], cams[0]["num_params"] = 4, MODEL_ID["PINHOLE"], 4
    cams[0]["width"], cams[0]["height"] = 100, 50
    cams[0]["params"][:4] = [90.0, 91.0, 49.5, 24.5]
    cams[1]["id"], cams[1]["model"], cams[1]["num_params"] = 8, MODEL_ID["RADIAL"], 5
    img = np.zeros((), rec.COLMAP_IMAGE)
    img["camera_id"] = 4
    img["name"] = b"x.png"
    eye = np.eye(3, dtype=np.float32).ravel()
    for q in ([1, 0, 0, 0], [2, 0, 0, 0], [0, 0, 0, 0]):  # identity, unnormalised, zero
        img["qvec"] = q
        c = abi.camera_from_colmap(cams, img)
        assert np.array_equal(c["rot"], eye), q
        assert (int(c["width"]), int(c["height"]), c["fx"], c["fy"]) == (100, 50, 90.0, 91.0)
    img["qvec"] = [0.0, 1.0, 0.0, 0.0]  # half turn about x
    c = abi.camera_from_colmap(cams, img)
    assert np.array_equal(c["rot"], np.diag([1, -1, -1]).astype(np.float32).ravel())
    img["camera_id"] = 8
    with pytest.raises(abi.ColmapError) as e:
        abi.camera_from_colmap(cams, img)
    assert e.value.status == abi.ERR_INVALID and "RADIAL" in str(e.value)
    img["camera_id"] = 99
    with pytest.raises(abi.ColmapError) as e:
        abi.camera_from_colmap(cams, img)
    assert e.value.status == abi.ERR_INVALID and "unknown camera 99" in str(e.value)


def truck_camera():
    c = np.zeros((), rec.CAMERA)
    c["width"], c["height"] = 1957, 1091
    c["fx"], c["fy"], c["cx"], c["cy"] = 1163.2547, 1156.2804, 978.5, 545.5
    c["rot"] = np.eye(3, dtype=np.float32).ravel()
    return c


def test_match_image_size():
    c = truck_camera()
    assert abi.match_image_size(c, 1957, 1091).tobytes() == c.tobytes()
    h = abi.match_image_size(c, 979, 546)
    assert (int(h["width"]), int(h["height"])) == (979, 546)
    sx, sy = 979.0 / 1957.0, 546.0 / 1091.0
    assert abs(float(h["fx"]) - 1163.2547 * sx) < 1e-3 and abs(float(h["fy"]) - 1156.2804 * sy) < 1e-3
    assert abs(float(h["cx"]) - 978.5 * sx) < 1e-3 and abs(float(h["cy"]) - 545.5 * sy) < 1e-3
    # Camera::resized in binary32: fx * (979 / 1957)
    assert h["fx"] == np.float32(c["fx"]) * (np.float32(979) / np.float32(1957))
    for w, hh in ((979, 600), (1091, 1957), (0, 0)):
        with pytest.raises(abi.ColmapError):
            abi.match_image_size(c, w, hh)


def write(tmp_path, name, data):
    p = tmp_path / name
    if isinstance(data, str):
        p.write_text(data)
    else:
        p.write_bytes(data)


@pytest.mark.parametrize("case", ["truncated", "count", "model", "params", "name", "token",
                                  "missing", "rgb"])
def test_malformed_models(tmp_path, case):
    good_cam = struct.pack("<Q", 1) + struct.pack("<IiQQ3d", 1, 0, 8, 8, 1.0, 4.0, 4.0)
    write(tmp_path, "images.txt", "")
    write(tmp_path, "points3D.txt", "")
    want = abi.ERR_FORMAT
    if case == "truncated":
        write(tmp_path, "cameras.bin", good_cam[:-3])
    elif case == "count":
        write(tmp_path, "cameras.bin", struct.pack("<Q", 10**12) + good_cam[8:])
    elif case == "model":
        write(tmp_path, "cameras.bin", struct.pack("<Q", 1) + struct.pack("<IiQQ", 1, 12, 8, 8))
    elif case == "params":
        write(tmp_path, "cameras.txt", "1 PINHOLE 8 8 1 2 3\n")
    elif case == "name":
        write(tmp_path, "cameras.txt", "1 PINHOLE 8 8 1 2 3 4\n")
        write(tmp_path, "images.txt", "1 1 0 0 0 0 0 0 1 " + "n" * 256 + "\n")
    elif case == "token":
        write(tmp_path, "cameras.txt", "1 PINHOLE 8 eight 1 2 3 4\n")
    elif case == "rgb":
        write(tmp_path, "cameras.txt", "1 PINHOLE 8 8 1 2 3 4\n")
        write(tmp_path, "points3D.txt", "1 0 0 0 256 0 0 0.5\n")
    else:
        want = abi.ERR_INVALID  # no cameras file at all
    with pytest.raises(abi.ColmapError) as e:
        abi.colmap_read(tmp_path)
    assert e.value.status == want, str(e.value)


def test_repeated_camera_id_keeps_the_last_record(tmp_path):
    # read_model collects cameras into a map by id (src/colmap/mod.rs:234): the record read
    # last wins; it stays at the place of the id's first occurrence.
    (tmp_path / "cameras.txt").write_text(
        "# Camera list\n"
        "1 PINHOLE 100 80 50 50 40 30\n"
        "2 SIMPLE_PINHOLE 10 10 5 5 5\n"
        "1 PINHOLE 200 160 60 61 80 60\n")
    (tmp_path / "images.txt").write_text("")
    (tmp_path / "points3D.txt").write_text("")
    c, i, p = abi.colmap_read(tmp_path)
    assert len(i) == 0 and len(p) == 0
    assert list(c["id"]) == [1, 2]
    assert int(c[0]["width"]) == 200 and int(c[0]["height"]) == 160
    assert list(c[0]["params"][:4]) == [60.0, 61.0, 80.0, 60.0]
    assert int(c[1]["model"]) == MODEL_ID["SIMPLE_PINHOLE"]
