"""JetModelZoom on the GPU solver, with a batched run for samplers.

The reference's `JetModelZoom` (jetmodel.py:11-112) wraps one call of
`pyjetshow.get_i_image` per parameter vector: `set_params_vec` (10 parameters,
log-scaled physical ones), `run`, `image` / `image_tau`, the log-spaced pixel
grid `pixsize_array` and the pixel-centre coordinates `r_ob`, `d` in parsec.
This module keeps that interface for the radiative-transfer part and adds
`run_many`, which sends the images of many models -- an ABC population or the
walkers of an ensemble sampler -- through one persistent launch instead of a
Python loop of blocking calls.

Differences: the angular-size distance is the solver's own cosmology
(`mas_to_pc`, src/utils.cpp:186-208, as pyjetshow uses it) instead of
astropy's WMAP9 (astropy is not a dependency here); `ft` evaluates the
visibilities as the exact direct sum on the GPU (jetshow_amd.fourier) instead
of finufft; the plotting helpers are not part of this path.
"""
import math

import numpy as np

from . import abi, fourier, host, models

MAS_TO_RAD = abi.MAS_TO_RAD


class JetModelZoom:
    nparams = 10

    def __init__(self, nu_ghz, redshift, n_along, n_across, lg_pixel_size_mas_min, lg_pixel_size_mas_max, stokes="I",
                 tau_max=1e5, central_vfield=True, tangled_bfield=True, ctx=None):
        if not tangled_bfield:
            raise ValueError("only the tangled (random) B field model, as in jetmodel.py:38-41")
        if stokes != "I":
            raise ValueError("the tangled-field model has Stokes I only")
        self.nu = float(nu_ghz)
        self.z = float(redshift)
        self.n_along, self.n_across = int(n_along), int(n_across)
        self.lg_pixel_size_mas_min = lg_pixel_size_mas_min
        self.lg_pixel_size_mas_max = lg_pixel_size_mas_max
        self.tau_max = tau_max
        self.central_vfield = central_vfield
        self.ctx = ctx
        resolutions = np.logspace(lg_pixel_size_mas_min, lg_pixel_size_mas_max, self.n_along)
        # [n_along, n_across] pixel sizes in mas (jetmodel.py:23-24)
        self.pixsize_array = np.tile(resolutions, self.n_across).reshape(self.n_across, self.n_along).T
        self.pc_in_mas = host.mas_to_pc(self.z)
        self.pix_to_pc = self.pixsize_array * self.pc_in_mas
        self._calculate_grid()
        self._image_i = None
        self._image_tau = None
        self._updated_params = False

    def _calculate_grid(self):
        """Pixel-centre coordinates (jetmodel.py:50-72, non-uniform branch)."""
        p = self.pix_to_pc
        pc_x = np.cumsum(p, axis=0) - p / 2
        up = p[:, :p.shape[1] // 2][::-1]
        low = p[:, p.shape[1] // 2:]
        up = (np.cumsum(up, axis=1) - up / 2)[::-1]
        low = np.cumsum(low, axis=1) - low / 2
        self.r_ob = pc_x
        self.d = -np.hstack((up[:, ::-1], -low))

    def set_params_vec(self, vec):
        """dx, dy, rot, theta, phi, log B1, log K1, log Gamma, log m, log n
        (jetmodel.py:84-92)."""
        self.dx, self.dy, self.rot, self.theta, self.phi, B1, K1, Gamma, m, n = vec
        self.B1, self.K1, self.gamma = math.exp(B1), math.exp(K1), math.exp(Gamma)
        self.m, self.n = math.exp(m), math.exp(n)
        self._updated_params = True

    def descriptors(self):
        """(JetDesc, ImageDesc, SolverParams) of the current parameters: the
        model pyjetshow.get_i_image builds (src/pyjetshow.cpp:69-156)."""
        return models.get_i_image_model(self.theta, self.z, self.n_along, self.n_across,
                                        10 ** self.lg_pixel_size_mas_min, 10 ** self.lg_pixel_size_mas_max,
                                        self.phi, self.B1, self.m, self.K1, self.n, self.gamma, self.nu,
                                        self.tau_max, self.central_vfield)

    def _jy(self, sizes_cm, I):
        """Jy per pixel as pyjetshow returns it (src/pyjetshow.cpp:111-134,163-169)."""
        omega = (sizes_cm / (self.pc_in_mas * abi.PC)) ** 2 * MAS_TO_RAD ** 2
        return I / (1e-23 * (1.0 + self.z) ** 3 / omega)

    def _store(self, planes, image):
        sizes = host.pixel_sizes(image).reshape(image.n_across, image.n_along)
        self._image_tau = planes["tau"]
        self._image_i = self._jy(sizes, planes["I"])
        self._flags = planes["flags"]
        self._updated_params = False

    def _context(self):
        if self.ctx is None:
            self.ctx = host.Context(0)
        return self.ctx

    def run(self):
        j, im, p = self.descriptors()
        self._store(self._context().observe(j, im, p), im)

    def image(self):
        """Stokes I in Jy per pixel, [n_across, n_along] (re-run when the
        parameters changed)."""
        if self._updated_params:
            self.run()
        return self._image_i

    def image_tau(self):
        if self._updated_params:
            self.run()
        return self._image_tau

    @property
    def flags(self):
        return self._flags

    def ft(self, uv, precision=abi.PREC_MIXED):
        """Visibilities of the current image at uv [n, 2] (wavelengths)
        (jetmodel.py:114-132): the image rotated by `rot` (uv @ rot) and shifted
        by (dx, dy) mas (a phase on the unrotated uv); pixel centres d, r_ob
        converted to radians.  mixed precision: ~1e-7 of the total flux."""
        uv = np.asarray(uv, dtype=np.float64).reshape(-1, 2)
        rot = np.array([[math.cos(self.rot), -math.sin(self.rot)],
                        [math.sin(self.rot), math.cos(self.rot)]])
        shift = np.array([self.dx * MAS_TO_RAD, self.dy * MAS_TO_RAD])
        result = np.exp(-2.0 * np.pi * 1j * (uv @ shift))
        rad_per_pc = MAS_TO_RAD / self.pc_in_mas
        x = self.d * rad_per_pc
        y = self.r_ob * rad_per_pc
        img = self.image()
        ft = fourier.FINUFFT_NUNU(uv @ rot, x.ravel(), y.ravel(), ctx=self._context(), precision=precision)
        return result * ft.forward(img.T.ravel())


def run_many(zooms, ctx=None, precision=abi.PREC_FP64, persistent=True):
    """Run every model whose parameters changed in one batched launch per
    precision group (jsx_observe_batch); afterwards `image()` of each returns
    without re-running.  precision=abi.PREC_MIXED uses the FP32 solver
    (relative tolerance floored at 1e-7 for Stokes I)."""
    todo = [z for z in zooms if z._updated_params]
    if not todo:
        return 0
    ctx = ctx or todo[0]._context()
    items = []
    for z in todo:
        j, im, p = z.descriptors()
        if precision == abi.PREC_MIXED:
            p = models.with_precision(p, abi.PREC_MIXED)
        if persistent:
            p.flags |= 0x100 | 0x200
        items.append((j, im, p))
    outs = ctx.observe_batch(items)
    for z, (j, im, p), planes in zip(todo, items, outs):
        z._store(planes, im)
    return len(todo)
