"""Resumable checkpoints (splat_checkpoint_write / splat_checkpoint_read), host code: the
byte layout, a bitwise round trip (including non-finite values and an empty model), sizing
calls, zero-filled missing arrays, and the errors for a foreign, truncated, padded or
future-version file."""
import struct

import numpy as np
import pytest

from splatrs_amd import abi
from splatrs_amd import records as rec


def state(n, seed=0):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, rec.NUM_PARAMS)).astype(np.float32)
    return {"params": g,
            "m": rng.standard_normal(g.shape).astype(np.float32),
            "v": rng.random(g.shape).astype(np.float32),
            "grad_accum": rng.random(n).astype(np.float32),
            "denom": rng.integers(0, 100, n).astype(np.uint32),
            "max_radii": rng.integers(0, 500, n).astype(np.uint32)}


def test_round_trip_is_bitwise(tmp_path):
    s = state(257)
    s["params"][3, 7] = np.inf
    s["params"][5, 2] = np.nan
    s["m"][0, 0] = -0.0
    p = tmp_path / "a.ckpt"
    abi.checkpoint_write(p, s, 1234567, 3)
    back, step, deg = abi.checkpoint_read(p)
    assert step == 1234567 and deg == 3
    for k, a in s.items():
        assert back[k].dtype == a.dtype
        assert np.array_equal(back[k].view(np.uint32), a.view(np.uint32)), k


def test_layout(tmp_path):
    s = state(3, 1)
    p = tmp_path / "b.ckpt"
    abi.checkpoint_write(p, s, 9, 2)
    raw = p.read_bytes()
    assert raw[:8] == b"SPLATCKP"
    assert struct.unpack("<IiQQ", raw[8:32]) == (1, 2, 3, 9)
    body = raw[32:]
    assert len(body) == 3 * (3 * 59 * 4 + 12)
    f = np.frombuffer(body[:3 * 3 * 59 * 4], np.float32).reshape(3, 3, 59)
    assert np.array_equal(f[0], s["params"]) and np.array_equal(f[1], s["m"])
    assert np.array_equal(f[2], s["v"])
    tail = body[3 * 3 * 59 * 4:]
    assert np.array_equal(np.frombuffer(tail[:12], np.float32), s["grad_accum"])
    assert np.array_equal(np.frombuffer(tail[12:24], np.uint32), s["denom"])
    assert np.array_equal(np.frombuffer(tail[24:], np.uint32), s["max_radii"])


def test_missing_arrays_are_zeros_and_empty_model(tmp_path):
    s = state(10, 2)
    p = tmp_path / "c.ckpt"
    abi.checkpoint_write(p, {"params": s["params"]}, 5, 0)
    back, step, deg = abi.checkpoint_read(p)
    assert step == 5 and deg == 0
    assert np.array_equal(back["params"], s["params"])
    for k in ("m", "v", "grad_accum", "denom", "max_radii"):
        assert not back[k].any(), k
    e = tmp_path / "e.ckpt"
    abi.checkpoint_write(e, {"params": np.zeros((0, 59), np.float32)}, 0, 3)
    back, step, _ = abi.checkpoint_read(e)
    assert back["params"].shape == (0, 59) and step == 0


@pytest.mark.parametrize("damage", ["magic", "version", "truncated", "padded", "degree",
                                    "wrapping_count"])
def test_malformed_files_are_errors(tmp_path, damage):
    s = state(4, 3)
    p = tmp_path / "d.ckpt"
    abi.checkpoint_write(p, s, 1, 3)
    raw = bytearray(p.read_bytes())
    if damage == "magic":
        raw[0:8] = b"NOTACKPT"
    elif damage == "version":
        raw[8:12] = struct.pack("<I", 2)
    elif damage == "truncated":
        raw = raw[:-1]
    elif damage == "padded":
        raw += b"\0"
    elif damage == "wrapping_count":
        # 4 + 2^60 rows of 720 bytes: the product wraps to this file's true body size
        raw[16:24] = struct.pack("<Q", 4 + (1 << 60))
    else:
        raw[12:16] = struct.pack("<i", 4)
    p.write_bytes(bytes(raw))
    with pytest.raises(abi.CheckpointError) as e:
        abi.checkpoint_read(p)
    assert e.value.status == abi.ERR_FORMAT


def test_invalid_arguments(tmp_path):
    with pytest.raises(abi.CheckpointError) as e:
        abi.checkpoint_write(tmp_path / "x.ckpt", state(2), 0, 4)
    assert e.value.status == abi.ERR_INVALID
    with pytest.raises(abi.CheckpointError) as e:
        abi.checkpoint_read(tmp_path / "missing.ckpt")
    assert e.value.status == abi.ERR_IO
