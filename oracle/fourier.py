"""CPU oracle of the visibility transform (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the CPU-baseline leg of
bench.py as the checker.  The product never imports it.

The reference evaluates the transform with finufft (a third-party library,
absent here; the reference loads `finufftpy` from a local checkout,
fourier.py:4-8, so no version is pinned).  finufft's published definitions
(type 3: f[k] = sum_j c[j] exp(+-i (s[k] x[j] + t[k] y[j])); type 2: c[j] =
sum_{k1,k2} f[k1,k2] exp(+-i (k1 x[j] + k2 y[j])), modes -ms/2 <= k1 <=
(ms-1)/2) are approximated to eps; this oracle evaluates the same sums
exactly, in the reference's own variables:

  nunu_forward    FINUFFT_NUNU (fourier.py:39-94): u = 2 pi uv / UV_max,
                  x * UV_max (fourier.py:57-63), isign -1 (fourier.py:88)
  uniform_forward FINUFFT (fourier.py:11-34): x = 2 pi u pixsize
                  (fourier.py:14-15), isign -1 (fourier.py:32)
  zoom_ft         JetModelZoom.ft (jetmodel.py:114-132)
  direct_sum      the sum in revolutions, sum_j c_j exp(sgn 2 pi i (u x + v y))

Pins (tests/test_fourier.py): direct_sum on a uniform grid at integer
frequencies against numpy.fft.fft2; a sampled Gaussian against its analytic
Fourier transform; Hermitian symmetry of real images; nunu_forward and
uniform_forward against direct_sum and the FFT.
"""
import math

import numpy as np

MAS_TO_RAD = math.pi / (180.0 * 3600.0 * 1000.0)


def direct_sum(x, y, c, u, v, isign=-1, block=2048):
    """sum_j c_j exp(sgn 2 pi i (u_k x_j + v_k y_j)), FP64, in blocks of
    visibilities."""
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    c = np.asarray(c).reshape(-1).astype(np.complex128)
    u = np.asarray(u, dtype=np.float64).reshape(-1)
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    sgn = 1.0 if isign >= 0 else -1.0
    out = np.zeros(u.size, dtype=np.complex128)
    for k0 in range(0, u.size, block):
        ph = np.outer(u[k0:k0 + block], x) + np.outer(v[k0:k0 + block], y)
        out[k0:k0 + block] = np.exp(sgn * 2j * np.pi * ph) @ c
    return out


def _type3(x, y, c, s, t, isign, block=2048):
    """finufft type-3 definition, radians: sum_j c_j exp(+-i (s_k x_j + t_k y_j))."""
    sgn = 1.0 if isign >= 0 else -1.0
    c = np.asarray(c).reshape(-1).astype(np.complex128)
    out = np.zeros(s.size, dtype=np.complex128)
    for k0 in range(0, s.size, block):
        ph = np.outer(s[k0:k0 + block], x) + np.outer(t[k0:k0 + block], y)
        out[k0:k0 + block] = np.exp(sgn * 1j * ph) @ c
    return out


def nunu_forward(uv, x, y, img):
    """FINUFFT_NUNU(uv, x, y).forward(img) (fourier.py:39-94), exactly."""
    uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
    u = uv[:, 0] * 2 * np.pi
    v = uv[:, 1] * 2 * np.pi
    uv_max = max(np.max(np.abs(u)), np.max(np.abs(v)))
    return _type3(np.asarray(x, dtype=np.float64) * uv_max, np.asarray(y, dtype=np.float64) * uv_max, img,
                  u / uv_max, v / uv_max, -1)


def uniform_forward(uv, pixsize, img):
    """FINUFFT(uv, pixsize).forward(img) (fourier.py:11-34), exactly: modes
    k1 = i - ms//2, k2 = j - mt//2 of img[i, j]."""
    uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
    xs = uv[:, 0] * pixsize * 2 * np.pi
    ys = uv[:, 1] * pixsize * 2 * np.pi
    img = np.asarray(img)
    ms, mt = img.shape
    k1, k2 = np.meshgrid(np.arange(ms) - ms // 2, np.arange(mt) - mt // 2, indexing="ij")
    # type 2 is the type-3 sum with the modes as sources
    return _type3(k1.reshape(-1).astype(np.float64), k2.reshape(-1).astype(np.float64), img.reshape(-1), xs, ys, -1)


def zoom_ft(uv, dx, dy, rot, d_pc, r_ob_pc, pc_in_mas, img):
    """JetModelZoom.ft (jetmodel.py:114-132) with the angular scale given as
    parsecs per mas: img [n_across, n_along] as the model returns it, d_pc and
    r_ob_pc [n_along, n_across] pixel centres."""
    uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
    r = np.array([[np.cos(rot), -np.sin(rot)], [np.sin(rot), np.cos(rot)]])
    shift = [dx * MAS_TO_RAD, dy * MAS_TO_RAD]
    result = np.exp(-2.0 * np.pi * 1j * (uv @ shift))
    uvr = uv @ r
    x = np.asarray(d_pc) / pc_in_mas * MAS_TO_RAD
    y = np.asarray(r_ob_pc) / pc_in_mas * MAS_TO_RAD
    return result * nunu_forward(uvr, x.ravel(), y.ravel(), np.asarray(img).T.ravel())
