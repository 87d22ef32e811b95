"""splat_photometric_loss_device_noalloc (needs a gfx950 GPU): the loss on device buffers
with its scratch in a static device array, against splat_photometric_loss (the host-tier
entry point, which allocates its buffers) on the same inputs.  Two calls in a row, the
second on a smaller image, so the second reads scratch the first has written: gradients
bitwise equal, loss values within one binary32 rounding of each other (both sum the
per-workgroup partials in binary64, in different orders).  An image above the maximum is
refused before anything is launched."""
import sys

import numpy as np
import pytest

from helpers import DeviceArrays
from splatrs_amd import abi

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_noalloc_loss_equals_allocating_entry_point():
    ctx = abi.Context(0)
    mem = DeviceArrays()
    try:
        for k, (w, h) in enumerate([(1920, 1080), (333, 217)]):
            say(f"no-allocation loss call {k + 1}: {w}x{h}")
            rng = np.random.default_rng(40 + k)
            pred = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
            target = np.clip(pred + rng.normal(0, 0.1, pred.shape), 0, 1).astype(np.float32)
            want_loss, want_grad = ctx.photometric_loss(pred, target, w, h, 0.2)
            dp, dt = mem.upload(pred), mem.upload(target)
            dg = mem.upload(np.full(pred.shape, np.nan, np.float32))
            dl = mem.upload(np.full(1, np.nan, np.float32))
            ctx.photometric_loss_device_noalloc(dp, dt, w, h, 0.2, dl, dg)
            ctx.sync()
            got_grad = mem.download(dg, pred.shape, np.float32)
            got_loss = float(mem.download(dl, (1,), np.float32)[0])
            say(f"loss {got_loss!r} vs {want_loss!r}")
            assert np.array_equal(got_grad.view(np.uint32), want_grad.view(np.uint32))
            assert abs(got_loss - want_loss) <= np.spacing(np.float32(want_loss))
        say("image above the maximum")
        with pytest.raises(abi.SplatError) as e:
            ctx.photometric_loss_device_noalloc(dp, dt, 4096, 4096, 0.2, 0, dg)
        assert e.value.status == abi.ERR_INVALID
    finally:
        mem.free()
        ctx.close()
