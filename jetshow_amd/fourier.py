"""Visibilities of model images on the GPU: the reference's Fourier classes
with their interfaces (fourier.py:11-94), evaluated as the exact direct sum
(jsx_visibilities, one MI355X launch) instead of finufft's approximation.

FINUFFT_NUNU(uv, x, y).forward(img)
    fourier.py:39-94: f[k] = sum_j img[j] exp(-i (2 pi u_k x_j + 2 pi v_k y_j))
    for scattered pixel centres (x, y) in radians and (u, v) in wavelengths --
    finufft nufft2d3 at eps 1e-6, here to ~1e-7 of sum |img| (mixed precision)
    or to FP64 rounding.
FINUFFT(uv, pixsize).forward(img)
    fourier.py:11-34: c[j] = sum_{k1,k2} img[k1, k2] exp(-i (k1 x_j + k2 y_j)),
    x_j = 2 pi u_j pixsize, k1 in [-ms/2, (ms-1)/2] -- finufft nufft2d2 at
    eps 1e-8, here FP64.

Both take an optional context (`ctx`, a host.Context; one is created on
device 0 when omitted) and `precision`.  There is no CPU fallback.
"""
import numpy as np

from . import abi, host

_ctx = None


def _default_ctx():
    global _ctx
    if _ctx is None:
        _ctx = host.Context(0)
    return _ctx


class FINUFFT_NUNU:
    """Scattered pixels -> scattered (u, v) (fourier.py:39-94)."""

    def __init__(self, uv, x, y, ctx=None, precision=abi.PREC_MIXED):
        uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
        self.u = np.ascontiguousarray(uv[:, 0])
        self.v = np.ascontiguousarray(uv[:, 1])
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        self.y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
        if self.x.size != self.y.size:
            raise ValueError("x and y must have the same length")
        self.size_nu = len(uv)
        self.ctx = ctx
        self.precision = precision

    def forward(self, img):
        """Visibilities [len(uv)] (complex128) of pixel values `img` [len(x)]."""
        img = np.asarray(img).reshape(-1)
        if img.size != self.x.size:
            raise ValueError("img has %d values for %d pixels" % (img.size, self.x.size))
        ctx = self.ctx or _default_ctx()
        return ctx.visibilities(self.x, self.y, img, self.u, self.v, isign=-1, precision=self.precision)


class FINUFFT:
    """Uniform pixel grid -> scattered (u, v) (fourier.py:11-34)."""

    def __init__(self, uv, pixsize, ctx=None, precision=abi.PREC_FP64):
        uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
        self.pixsize = pixsize
        # frequencies in cycles per pixel: finufft's x_j = 2 pi u_j pixsize
        self.s = np.ascontiguousarray(uv[:, 0] * pixsize)
        self.t = np.ascontiguousarray(uv[:, 1] * pixsize)
        self.size_nu = len(uv)
        self.ctx = ctx
        self.precision = precision

    def forward(self, img):
        """img [ms, mt]; mode (k1, k2) = (i - ms//2, j - mt//2) as finufft's
        default ordering (-ms/2 <= k1 <= (ms-1)/2)."""
        img = np.asarray(img)
        ms, mt = img.shape
        k1, k2 = np.meshgrid(np.arange(ms, dtype=np.float64) - ms // 2,
                             np.arange(mt, dtype=np.float64) - mt // 2, indexing="ij")
        ctx = self.ctx or _default_ctx()
        return ctx.visibilities(k1.reshape(-1), k2.reshape(-1), img.reshape(-1), self.s, self.t, isign=-1,
                                precision=self.precision)
