"""The allocation-free stitch (kmig_batch_stitch_noalloc, through the C ABI)
against kmig_batch_stitch on the same layers: bit for bit, on two calls in a
row, the second with a smaller input (three layers, a 20-sample angle window,
other lateral weights), so the second call meets the first call's scratch
contents.  Output buffers are device memory allocated by the test."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def layers():
    from kirch_datuming_amd import configs, kmig
    print("no-alloc stitch: five overlapping layers, one batch", flush=True)
    ws = configs.layers(threads=8, scale=0.04, overlap=12)
    ms = []
    for w in ws:
        ax = w.data_axes
        p = kmig.parse_params(dict(w.kv, tile=9), int(ax[1][0]), ax[1][1], ax[1][2])
        m = kmig.Migrator(p)
        m.set_tables(0, w.stab, w.sder, w.s_axes)
        m.set_tables(1, w.rtab, w.rder, w.r_axes)
        m.set_data(w.data, ax)
        ms.append(m)
    b = kmig.Batch(ms)
    b.run()
    b.sync()
    yield ms, b, ws
    b.close()
    for m in ms:
        m.close()


def _hip():
    """The HIP runtime libkmig itself links (torch ships another one, and two
    runtimes in one process do not share devices)."""
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return H


def _device_stitch(b, nz, nx, ng, **opts):
    """Stitch into a device buffer allocated here (filled with NaN bytes
    first, so an unwritten sample cannot pass), then copy it to the host."""
    H = _hip()
    n = ng * nx * nz
    ptr = C.c_void_p()
    assert H.hipMalloc(C.byref(ptr), n * 4) == 0
    try:
        assert H.hipMemset(ptr, 0xFF, n * 4) == 0
        b.stitch_into(ptr.value, **opts)
        out = np.empty((ng, nx, nz), np.float32)
        assert H.hipMemcpy(out.ctypes.data, ptr, n * 4, 2) == 0  # device to host
    finally:
        H.hipFree(ptr)
    return out


def test_noalloc_stitch_matches_the_allocating_stitch(layers):
    from kirch_datuming_amd import kmig
    ms, b, ws = layers
    start, end = [w.z_range[0] for w in ws], [w.z_range[1] for w in ws]
    nx = ms[0].nx
    # call 1: five layers, every angle, a lateral ramp of 40 samples
    o1 = dict(start=start, end=end, stitch_blend=5, scale=[1.0, 0.8, 1.2, 1.0, 0.9],
              scalex_aper=40, scalex_scale=1.7)
    print("no-alloc stitch: call 1 (5 layers, 61 angles)", flush=True)
    ref1 = b.stitch(**o1)
    got1 = _device_stitch(b, end[-1] + 1, nx, ref1.shape[0], **o1)
    assert np.array_equal(got1.view(np.uint32), ref1.view(np.uint32))
    # call 2: a smaller input on another batch of the first three layers
    b3 = kmig.Batch(ms[:3])
    try:
        o2 = dict(start=start[:3], end=end[:3], stitch_blend=-1, scalex_aper=3,
                  scalex_scale=0.5, angles=(10, 20))
        print("no-alloc stitch: call 2 (3 layers, 20 angles)", flush=True)
        ref2 = b3.stitch(**o2)
        got2 = _device_stitch(b3, end[2] + 1, nx, 20, **o2)
        assert np.array_equal(got2.view(np.uint32), ref2.view(np.uint32))
        assert not np.array_equal(ref2, ref1[10:30, :, :end[2] + 1])
    finally:
        b3.close()


def test_noalloc_stitch_refuses_host_memory(layers):
    from kirch_datuming_amd import kmig
    ms, b, ws = layers
    start, end = [w.z_range[0] for w in ws], [w.z_range[1] for w in ws]
    host = np.empty((61, ms[0].nx, end[-1] + 1), np.float32)
    with pytest.raises(kmig.KmigError, match="not a device pointer"):
        b.stitch_into(host.ctypes.data, start=start, end=end)
