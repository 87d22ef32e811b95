"""The C-ABI library loads and exports every entry point include/splat.h declares; its
host-only helpers (to_rgb8, psnr) match the reference; context creation without a gfx950
device fails with SPLAT_ERR_NO_DEVICE instead of falling back to anything."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from splatrs_amd import abi


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    syms = abi.declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_no_device_is_an_error_not_a_fallback():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert abi.lib().splat_ctx_create(0, ctypes.byref(h)) == abi.ERR_NO_DEVICE
    assert not h.value


def test_to_rgb8_and_psnr_match_reference():
    img = np.array([-0.2, 0.0, 0.5, 0.998, 1.0, 1.7, np.nan, 0.0019608, 0.5 / 255],
                   np.float32)
    got = abi.to_rgb8(img)
    # src/render.rs:26-31: round(clamp(v) * 255), NaN saturates to 0
    want = [0, 0, 128, 254, 255, 255, 0, 1, 1]
    assert list(got) == want
    rng = np.random.default_rng(0)
    a = rng.uniform(-0.1, 1.1, 999).astype(np.float32)
    b = rng.uniform(0, 1, 999).astype(np.float32)
    assert abi.psnr(a, b) == O.psnr(a, b)
    assert abi.psnr(a, a) == np.inf
