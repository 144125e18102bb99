"""Allocation-free batched render (fpt_hip_render_batch_noalloc), called
through the C ABI: its images equal fpt_hip_batch_create +
fpt_hip_batch_render's bit for bit, on two calls in a row, the second with a
smaller batch (the device array keeps the first call's contents), and it
refuses sizes beyond its fixed maximum."""
import ctypes as C

import numpy as np
import pytest

from fractalpathtracer_amd import api, ffi, scenes

pytestmark = pytest.mark.gpu


def _batch_reference(cfgs, stream):
    with api.Batch(cfgs) as b:
        b.tune(0, 0, use_graph=False)
        buf = api.DeviceBuffer(b.pixels * 4)
        b.render_into(buf.address, stream.address)
        stream.synchronize()
        return buf.numpy().reshape(-1, 4), b.pixels


def _noalloc(lib, cfgs, buf, floats, stream):
    arr = (ffi.FptRenderConfig * len(cfgs))(*cfgs)
    err = ffi.error_buffer()
    rc = lib.fpt_hip_render_batch_noalloc(arr, len(cfgs), C.c_void_p(buf.address), floats,
                                          C.c_void_p(stream.address), err, len(err))
    return rc, err.value.decode()


def test_noalloc_batch_matches_batch_render_on_two_calls():
    lib = ffi.load_library()
    st = api.Stream()
    big = scenes.batch_scenes(count=7, seed=11, min_tile=24, max_tile=80, max_spp=3)
    # A lens scene exercises the focus prepass, which writes into the array.
    big[1].camera_dof = 0.02
    big[1].focus_distance = 0.0
    small = scenes.batch_scenes(count=3, seed=12, min_tile=16, max_tile=40, max_spp=2)
    ref_big, px_big = _batch_reference(big, st)
    ref_small, px_small = _batch_reference(small, st)
    assert px_small < px_big
    out = api.DeviceBuffer(px_big * 4)
    for name, cfgs, ref, px in (("first", big, ref_big, px_big), ("second", small, ref_small, px_small)):
        print(f"no-alloc call {name}: {len(cfgs)} scenes, {px} pixels", flush=True)
        out.fill(-1.0, st)
        rc, msg = _noalloc(lib, cfgs, out, px * 4, st)
        assert rc == 0, msg
        st.synchronize()
        got = out.numpy().reshape(-1, 4)
        assert np.array_equal(got[:px], ref), f"{name} call differs from fpt_hip_batch_render"
        # Nothing is written past the batch's own pixels.
        assert np.all(got[px:] == -1.0)


def test_noalloc_batch_rejects_oversized_requests():
    lib = ffi.load_library()
    st = api.Stream()
    cfgs = scenes.batch_scenes(count=2, seed=13, min_tile=16, max_tile=32, max_spp=1)
    px = sum(c.width * c.height for c in cfgs)
    out = api.DeviceBuffer(px * 4)
    rc, msg = _noalloc(lib, cfgs, out, px * 4 - 1, st)
    assert rc == 1 and "too small" in msg
    many = [cfgs[0]] * (ffi.NOALLOC_MAX_SCENES + 1)
    rc, msg = _noalloc(lib, many, out, px * 4, st)
    assert rc == 1 and "1..64 scenes" in msg
