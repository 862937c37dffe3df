"""Visibilities of model images (fourier.py:11-94, jetmodel.py:114-132).

CPU: the oracle (oracle/fourier.py) pinned against numpy's FFT, the analytic
transform of a Gaussian and Hermitian symmetry; the mixed-precision kernel's
sine/cosine and partial-sum arithmetic (compiled for the CPU) against the
oracle.  GPU: jsx_visibilities through the C ABI against the oracle (FP64 to
1e-12 of sum |c|, mixed to 1e-6 of sum |c|, finufft's eps for
FINUFFT_NUNU), edge sizes, device entry point, JetModelZoom.ft and the
uniform-grid class."""
import math

import numpy as np
import pytest

import lanehost
from jetshow_amd import abi, fourier, zoom
from oracle import fourier as ofour

RNG = np.random.default_rng(20261203)


def _scatter(n_src, n_vis, extent=1e-7, umax=1e9, cplx=False, seed=1):
    r = np.random.default_rng(seed)
    x = r.uniform(-extent, extent, n_src)
    y = r.uniform(-extent, extent, n_src)
    c = r.uniform(0.0, 1.0, n_src)
    if cplx:
        c = c + 1j * r.uniform(-1.0, 1.0, n_src)
    u = r.uniform(-umax, umax, n_vis)
    v = r.uniform(-umax, umax, n_vis)
    return x, y, c, u, v


# ---------------------------------------------------------------- oracle pins

def test_oracle_direct_sum_is_fft2_on_a_grid():
    n1, n2 = 12, 10
    img = RNG.normal(size=(n1, n2)) + 1j * RNG.normal(size=(n1, n2))
    i, j = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
    k1, k2 = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
    got = ofour.direct_sum(i.ravel(), j.ravel(), img.ravel(), k1.ravel() / n1, k2.ravel() / n2, isign=-1)
    ref = np.fft.fft2(img).ravel()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(img).sum()
    # isign +1 is the unnormalised inverse transform
    got = ofour.direct_sum(i.ravel(), j.ravel(), img.ravel(), k1.ravel() / n1, k2.ravel() / n2, isign=1)
    assert np.max(np.abs(got - np.fft.ifft2(img).ravel() * n1 * n2)) <= 1e-12 * np.abs(img).sum()


def test_oracle_gaussian_matches_analytic_transform():
    # exp(-r^2 / 2 s^2) <-> 2 pi s^2 exp(-2 pi^2 s^2 q^2); the Riemann sum on a
    # grid of spacing h is exact up to aliasing exp(-2 pi^2 s^2 / h^2) and the
    # truncation at 8 s
    s, h = 1.0, 0.25
    g = np.arange(-32, 33) * h
    x, y = np.meshgrid(g, g, indexing="ij")
    c = np.exp(-(x ** 2 + y ** 2) / (2 * s * s)) * h * h
    u = RNG.uniform(-1.0, 1.0, 200)
    v = RNG.uniform(-1.0, 1.0, 200)
    got = ofour.direct_sum(x.ravel(), y.ravel(), c.ravel(), u, v, isign=-1)
    ref = 2 * np.pi * s * s * np.exp(-2 * np.pi ** 2 * s * s * (u ** 2 + v ** 2))
    assert np.max(np.abs(got - ref)) < 1e-10


def test_oracle_hermitian_and_shift():
    x, y, c, u, v = _scatter(300, 100)
    a = ofour.direct_sum(x, y, c, u, v)
    b = ofour.direct_sum(x, y, c, -u, -v)
    assert np.max(np.abs(a - np.conj(b))) <= 1e-12 * c.sum()
    # moving every pixel by (x0, y0) multiplies by exp(-2 pi i (u x0 + v y0))
    x0, y0 = 3e-8, -2e-8
    s = ofour.direct_sum(x + x0, y + y0, c, u, v)
    assert np.max(np.abs(s - a * np.exp(-2j * np.pi * (u * x0 + v * y0)))) <= 1e-9 * c.sum()


def test_oracle_reference_forms():
    x, y, c, u, v = _scatter(400, 150, cplx=True)
    uv = np.stack([u, v], axis=1)
    assert np.max(np.abs(ofour.nunu_forward(uv, x, y, c) - ofour.direct_sum(x, y, c, u, v))) <= 1e-9 * np.abs(c).sum()
    # the uniform grid at the grid frequencies: an FFT with the centred modes
    for ms, mt in ((8, 6), (7, 5)):
        img = RNG.normal(size=(ms, mt))
        pix = 2.5e-9
        m1, m2 = np.meshgrid(np.arange(ms), np.arange(mt), indexing="ij")
        uvg = np.stack([m1.ravel() / (ms * pix), m2.ravel() / (mt * pix)], axis=1)
        got = ofour.uniform_forward(uvg, pix, img)
        ph = np.exp(2j * np.pi * ((ms // 2) * m1 / ms + (mt // 2) * m2 / mt))
        ref = (ph * np.fft.fft2(img)).ravel()
        assert np.max(np.abs(got - ref)) <= 1e-9 * np.abs(img).sum()


def test_oracle_zoom_ft_rotation_and_shift():
    z = zoom.JetModelZoom(15.4, 0.0165, 20, 6, -3, -1, central_vfield=False)
    img = RNG.uniform(size=(6, 20))
    pc = 5.0
    uv = RNG.uniform(-2e8, 2e8, size=(50, 2))
    a = ofour.zoom_ft(uv, 0.0, 0.0, 0.0, z.d, z.r_ob, pc, img)
    # rotating by pi mirrors the real image: the conjugate visibilities
    b = ofour.zoom_ft(uv, 0.0, 0.0, math.pi, z.d, z.r_ob, pc, img)
    assert np.max(np.abs(a - np.conj(b))) <= 1e-9 * img.sum()
    # a shift changes phases only
    s = ofour.zoom_ft(uv, 0.3, -0.2, 0.0, z.d, z.r_ob, pc, img)
    assert np.allclose(np.abs(s), np.abs(a), rtol=1e-10, atol=1e-12 * img.sum())


# --------------------------------------- mixed-precision arithmetic on the CPU

def test_mixed_sincos_accuracy():
    ph = np.concatenate([RNG.uniform(-1e4, 1e4, 200000), np.arange(-16, 17) / 8.0,
                         np.arange(-16, 17) / 8.0 + 1e-9, np.arange(-16, 17) / 8.0 - 1e-9])
    s, c = lanehost.vis_sincos(ph)
    f = ph - np.rint(ph)   # exact
    es = np.abs(s - np.sin(2 * np.pi * f))
    ec = np.abs(c - np.cos(2 * np.pi * f))
    # the rounding of the remainder to FP32 and a few FP32 operations
    assert es.max() < 1.2e-7 and ec.max() < 1.2e-7, (es.max(), ec.max())
    assert (np.abs(s) <= 1.0).all() and (np.abs(c) <= 1.0).all()


def test_mixed_sum_on_cpu_matches_oracle():
    x, y, c, u, v = _scatter(2000, 64, umax=5e9)   # phases up to ~1000 revolutions
    got = lanehost.vis_mixed(x, y, c, u, v)
    ref = ofour.direct_sum(x, y, c, u, v)
    err = np.max(np.abs(got - ref)) / c.sum()
    assert err < 1e-6, err


# ------------------------------------------------------------------------ GPU

def _gpu_err(ctx, x, y, c, u, v, isign=-1, precision=abi.PREC_FP64):
    got = ctx.visibilities(x, y, c, u, v, isign=isign, precision=precision)
    ref = ofour.direct_sum(x, y, c, u, v, isign=isign)
    return np.max(np.abs(got - ref)) / np.abs(c).sum(), got


@pytest.mark.gpu
@pytest.mark.parametrize("n_src,n_vis,cplx,isign", [(100, 300, False, -1), (3000, 2500, False, -1),
                                                    (3000, 2500, True, 1), (20000, 700, True, -1)])
def test_visibilities_fp64(ctx, n_src, n_vis, cplx, isign):
    print("vis fp64", n_src, n_vis, cplx, isign, flush=True)
    x, y, c, u, v = _scatter(n_src, n_vis, cplx=cplx, seed=n_src + n_vis)
    err, _ = _gpu_err(ctx, x, y, c, u, v, isign=isign)
    assert err < 1e-12, err


@pytest.mark.gpu
@pytest.mark.parametrize("trig", [0, 1])
def test_visibilities_mixed(ctx, trig):
    saved = ctx.get_option("vis_trig")
    try:
        ctx.set_option("vis_trig", trig)
        errs = []
        for n_src, n_vis, cplx in ((3000, 2500, False), (20000, 700, True)):
            print("vis mixed trig", trig, n_src, n_vis, cplx, flush=True)
            x, y, c, u, v = _scatter(n_src, n_vis, cplx=cplx, umax=5e9, seed=7 + n_src)
            err, _ = _gpu_err(ctx, x, y, c, u, v, precision=abi.PREC_MIXED)
            errs.append(err)
        print("mixed visibility error / sum|c| (vis_trig %d):" % trig, errs, flush=True)
        assert max(errs) < 1e-6, errs
    finally:
        ctx.set_option("vis_trig", saved)


@pytest.mark.gpu
@pytest.mark.parametrize("trig", [0, 1])
def test_mixed_sine_cosine_per_value(ctx, trig):
    """A one-pixel image at x = 1: every visibility is exp(-2 pi i u), one sine
    and one cosine, so the per-value error of either sine/cosine form shows
    (phases up to 1e4 revolutions and the quarter-turn boundaries)."""
    ph = np.concatenate([RNG.uniform(-1e4, 1e4, 100000), np.arange(-16, 17) / 8.0,
                         np.arange(-16, 17) / 8.0 + 1e-9, np.arange(-16, 17) / 8.0 - 1e-9])
    saved = ctx.get_option("vis_trig")
    try:
        ctx.set_option("vis_trig", trig)
        print("sine/cosine per value, vis_trig", trig, flush=True)
        got = ctx.visibilities([1.0], [0.0], [1.0], ph, np.zeros_like(ph), precision=abi.PREC_MIXED)
    finally:
        ctx.set_option("vis_trig", saved)
    f = ph - np.rint(ph)
    es = np.max(np.abs(got.imag + np.sin(2 * np.pi * f)))
    ec = np.max(np.abs(got.real - np.cos(2 * np.pi * f)))
    print("max |sin err| %.3g, |cos err| %.3g (vis_trig %d)" % (es, ec, trig), flush=True)
    assert es < 2e-7 and ec < 2e-7, (es, ec)


@pytest.mark.gpu
def test_visibilities_edges(ctx):
    x, y, c, u, v = _scatter(50, 2049)
    print("vis edges", flush=True)
    # no pixels: zeros; no visibilities: empty
    assert np.array_equal(ctx.visibilities(x[:0], y[:0], c[:0], u, v), np.zeros(2049, complex))
    assert ctx.visibilities(x, y, c, u[:0], v[:0]).size == 0
    # one visibility, block-boundary counts
    for n in (1, 1023, 1024, 1025, 2047, 2048, 2049):
        err, _ = _gpu_err(ctx, x, y, c, u[:n], v[:n])
        assert err < 1e-12, (n, err)
    from jetshow_amd import host
    with pytest.raises(host.JsxError):
        ctx.visibilities(x, y, c, u, v, precision=7)


@pytest.mark.gpu
def test_visibilities_device_matches_host(ctx):
    import torch
    x, y, c, u, v = _scatter(5000, 3000, seed=11)
    print("vis device", flush=True)
    host_vis = ctx.visibilities(x, y, c, u, v, precision=abi.PREC_MIXED)
    d = {k: torch.from_numpy(np.ascontiguousarray(a)).cuda() for k, a in
         (("x", x), ("y", y), ("c", c), ("u", u), ("v", v))}
    out = torch.zeros(2 * u.size, dtype=torch.float64, device="cuda")
    ctx.visibilities_device(d["x"].data_ptr(), d["y"].data_ptr(), d["c"].data_ptr(), None, x.size,
                            d["u"].data_ptr(), d["v"].data_ptr(), u.size, out.data_ptr(), precision=abi.PREC_MIXED)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert np.array_equal(o[0::2] + 1j * o[1::2], host_vis)


@pytest.mark.gpu
def test_zoom_ft_and_uniform_class(ctx):
    z = zoom.JetModelZoom(15.4, 0.0165, 50, 10, -3, -1, central_vfield=False, ctx=ctx)
    z.set_params_vec(np.array([0.2, -0.1, 0.4, 0.663225, 0.068, math.log(0.1), math.log(1e4), math.log(2.0),
                               math.log(1.0), math.log(2.0)]))
    print("zoom ft", flush=True)
    uv = RNG.uniform(-3e8, 3e8, size=(4000, 2))
    img = z.image()
    assert (img > 0).any()
    ref = ofour.zoom_ft(uv, z.dx, z.dy, z.rot, z.d, z.r_ob, z.pc_in_mas, img)
    flux = np.abs(img).sum()
    assert np.max(np.abs(z.ft(uv, precision=abi.PREC_FP64) - ref)) < 1e-12 * flux
    assert np.max(np.abs(z.ft(uv) - ref)) < 1e-6 * flux
    # FINUFFT (uniform grid, fourier.py:11-34)
    im = RNG.uniform(size=(32, 24))
    uvg = RNG.uniform(-1e8, 1e8, size=(3000, 2))
    pix = 1e-9
    got = fourier.FINUFFT(uvg, pix, ctx=ctx).forward(im)
    assert np.max(np.abs(got - ofour.uniform_forward(uvg, pix, im))) < 1e-12 * im.sum()


@pytest.mark.gpu
def test_visibilities_device_two_streams(ctx):
    """Two chunked launches on different streams share the context's
    partial-sum buffer: the second waits for the first (event), so both results
    equal their host-entry values."""
    import torch
    sets = [_scatter(6000, 2100, seed=21), _scatter(9000, 1500, seed=22)]
    print("vis two streams", flush=True)
    ref = [ctx.visibilities(x, y, c, u, v, precision=abi.PREC_MIXED) for x, y, c, u, v in sets]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs, keep = [], []
    for (x, y, c, u, v), st in zip(sets, streams):
        d = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, c, u, v)]
        torch.cuda.synchronize()
        out = torch.zeros(2 * u.size, dtype=torch.float64, device="cuda")
        ctx.visibilities_device(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), None, x.size, d[3].data_ptr(),
                                d[4].data_ptr(), u.size, out.data_ptr(), precision=abi.PREC_MIXED,
                                stream=st.cuda_stream)
        outs.append(out)
        keep.append(d)
    torch.cuda.synchronize()
    for o, r in zip(outs, ref):
        o = o.cpu().numpy()
        assert np.array_equal(o[0::2] + 1j * o[1::2], r)
