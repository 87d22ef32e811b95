"""Dataset loading on the GPU (needs a gfx950 GPU): dataset::load's downscale
(src/dataset.rs:131-205) through the trainer.

- splat_resize_rgb8 equals the oracle's restatement of the `image` crate's Lanczos3
  resampler bitwise: integer downscales of a 1957x1091 image (the tandt_db camera), odd
  sizes, an upscale, a copy and a 1x1 result;
- a synthetic COLMAP project whose images are stored at half the cameras' resolution
  loads in name order with downscale 2: every view's camera is the model camera matched
  to its image and resized (src/dataset.rs:159-184), and the pixels the trainer holds are
  the oracle's resize of the image; split_eval holds out every 8th view.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from splatrs_amd import abi, scenes
from splatrs_amd import records as rec
from test_dataset import photo

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


def test_resize_matches_oracle(ctx):
    rng = np.random.default_rng(5)
    big = photo(rng, 1091, 1957)
    for d in (2, 3, 4, 8):
        w, h = abi.downscaled_size(1957, 1091, d)
        say(f"1957x1091 downscale {d} -> {w}x{h}")
        assert np.array_equal(ctx.resize_rgb8(big, w, h), O.resize_rgb8(big, w, h))
    small = photo(rng, 61, 97)
    for ow, oh in ((32, 20), (97, 30), (40, 61), (150, 90), (97, 61), (1, 1)):
        say(f"97x61 -> {ow}x{oh}")
        assert np.array_equal(ctx.resize_rgb8(small, ow, oh), O.resize_rgb8(small, ow, oh))


def test_load_dataset_downscales_on_the_gpu(ctx, tmp_path):
    from test_colmap import rotation_to_qvec
    full_w, full_h, n = 400, 300, 10
    cams = scenes.small_scene_cameras(n, size=256)
    cameras = np.zeros(1, rec.COLMAP_CAMERA)
    cameras[0]["id"], cameras[0]["model"], cameras[0]["num_params"] = 1, 1, 4  # PINHOLE
    cameras[0]["width"], cameras[0]["height"] = full_w, full_h
    cameras[0]["params"][:4] = [350.0, 351.0, 200.0, 150.0]
    images = np.zeros(n, rec.COLMAP_IMAGE)
    rng = np.random.default_rng(8)
    files = {}
    for k in range(n):
        images[k]["id"] = 100 + k
        images[k]["camera_id"] = 1
        images[k]["qvec"] = rotation_to_qvec(np.asarray(cams[k]["rot"], np.float64).reshape(3, 3))
        images[k]["tvec"] = np.asarray(cams[k]["trans"], np.float64)
        name = f"img_{(7 * k) % n:02d}.npy"  # file order differs from name order
        images[k]["name"] = name.encode()
        files[name] = photo(rng, full_h // 2, full_w // 2)  # stored at half resolution
    points = np.zeros(3, rec.COLMAP_POINT)
    points["xyz"] = [[0.1, 0.2, 0.3], [-0.5, 0.0, 0.25], [0.0, 0.4, -0.2]]
    points["rgb"] = [[1, 2, 3], [200, 100, 50], [255, 0, 128]]
    sparse = tmp_path / "sparse" / "0"
    sparse.mkdir(parents=True)
    (tmp_path / "images").mkdir()
    for name, img in files.items():
        np.save(tmp_path / "images" / name, img)
    abi.colmap_write(sparse, cameras, images, points)
    sp, im = abi.locate(str(tmp_path))
    tr = abi.Trainer(ctx, scenes.small_scene(0, 3000))
    say("load_dataset with downscale 2")
    used, xyz, rgb, extent = abi.load_dataset(tr, sp, im, 2, lambda p: np.load(p))
    assert len(used) == n and xyz.shape == (3, 3) and (rgb == points["rgb"]).all()
    order = np.argsort([bytes(x) for x in images["name"]], kind="stable")
    for v, k in enumerate(order):
        base = abi.camera_from_colmap(cameras, images[k])
        matched = abi.match_image_size(base, full_w // 2, full_h // 2)
        w, h = abi.downscaled_size(full_w // 2, full_h // 2, 2)
        want = abi.match_image_size(matched, w, h)  # Camera::resized, same aspect
        assert used[v].tobytes() == want.tobytes(), v
        name = images[k]["name"].decode()
        cam, pixels = tr.view(v)
        assert cam.tobytes() == want.tobytes()
        assert np.array_equal(pixels, O.resize_rgb8(files[name], w, h)), v
    assert abs(extent - scenes.scene_extent(used)) == 0
    train, test = abi.split_eval(n, 8)
    assert list(test) == [0, 8]
    tr.close()
