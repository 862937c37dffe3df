"""jsx_visibilities_device_noalloc through the C ABI: the allocation-free
entry point equals jsx_visibilities_device bit for bit, on two calls in a
row (the second smaller, so stale partial sums of the first would show),
in FP64 and mixed precision; it allocates no device memory; a launch
whose partial sums exceed its fixed array is refused before it runs."""
import ctypes as C

import numpy as np
import pytest

from jetshow_amd import abi
from jetshow_amd.host import JsxError, lib

pytestmark = pytest.mark.gpu


def _scatter(n_src, n_vis, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1e-8, 1e-8, n_src)
    y = rng.uniform(-1e-8, 1e-8, n_src)
    c = rng.uniform(0.0, 1.0, n_src)
    u = rng.uniform(-3e8, 3e8, n_vis)
    v = rng.uniform(-3e8, 3e8, n_vis)
    return x, y, c, u, v


def _device(arrays):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def _run(ctx, fn, d, n_src, n_vis, precision):
    import torch
    out = torch.full((2 * n_vis,), -1.0, dtype=torch.float64, device="cuda")
    vp = lambda t: C.c_void_p(t.data_ptr())
    rc = getattr(lib(), fn)(ctx.handle, vp(d[0]), vp(d[1]), vp(d[2]), None, C.c_int64(n_src), vp(d[3]), vp(d[4]),
                            C.c_int64(n_vis), C.c_int32(-1), C.c_int32(precision), vp(out), None)
    assert rc == 0, lib().jsx_last_error(ctx.handle)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_noalloc_equals_device_entry(ctx, precision):
    import torch
    for k, (n_src, n_vis) in enumerate(((20000, 3000), (7000, 1000))):
        print("noalloc call %d: %d pixels, %d visibilities, precision %d" % (k, n_src, n_vis, precision), flush=True)
        d = _device(_scatter(n_src, n_vis, seed=40 + k))
        ref = _run(ctx, "jsx_visibilities_device", d, n_src, n_vis, precision)
        free0 = torch.cuda.mem_get_info()[0]
        got = _run(ctx, "jsx_visibilities_device_noalloc", d, n_src, n_vis, precision)
        free1 = torch.cuda.mem_get_info()[0]
        assert np.array_equal(got, ref), k
        assert np.abs(got).max() > 0
        if k:   # the first call may load the kernels' code objects
            assert free1 == free0


def test_noalloc_refuses_oversized_launch(ctx):
    # 1e6 visibilities of 1e6 pixels need 5 x 1e6 partial sums > JSX_VIS_SCRATCH_MAX;
    # the buffers are real and of full size
    import torch
    n = 1000000
    d = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(5)]
    out = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    with pytest.raises(JsxError) as e:
        ctx.visibilities_device_noalloc(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), None, n,
                                        d[3].data_ptr(), d[4].data_ptr(), n, out.data_ptr())
    assert "partial sums" in str(e.value)
