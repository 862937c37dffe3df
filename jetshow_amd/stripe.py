"""Image sizing from a tau stripe (the step before the full run).

`analyze_tau_stripe` restates the decision rule of example_simple.py:105-137
on an optical-depth stripe (Observation::run_stripe, here on the GPU through
jsx_tau_stripe), and `find_image_params` the iteration of
example_simple.py:172-237 that grows or shrinks the pixel size and the number
of pixels until the tau = 1 surface and the optically thin end sit where the
rule wants them -- without writing config files and re-running an executable
for every try.
"""
import math

import numpy as np

from . import abi


def analyze_tau_stripe(tau, log10_tau_small=-4.0, border_tau1=(0.001, 0.3), border_tau_min=(0.5, 1),
                       min_pixels_to_tau1=20):
    """Decision flags for a tau stripe (example_simple.py:105-137)."""
    tau = np.asarray(tau, dtype=np.float64)
    length = len(tau)
    with np.errstate(divide="ignore"):
        lt = np.log10(tau)
    idx_tau1 = int(np.argmin(np.abs(lt - 0.0)))
    idx_tau_ = int(np.argmin(np.abs(lt - log10_tau_small)))
    increase_pixel_size = decrease_pixels_size = False
    increase_number_of_pixels = decrease_number_of_pixels = False
    if float(idx_tau1) / length > border_tau1[0]:
        if float(idx_tau1) / length > border_tau1[1]:
            increase_pixel_size = True
    else:
        decrease_pixels_size = True
    if idx_tau_ < length * border_tau_min[1] - 1:
        if idx_tau_ < border_tau_min[0] * length:
            decrease_number_of_pixels = True
    else:
        increase_number_of_pixels = True
    if idx_tau1 < min_pixels_to_tau1:
        decrease_pixels_size = True
    return {"increase_pixel_size": increase_pixel_size,
            "decrease_pixels_size": decrease_pixels_size,
            "increase_number_of_pixels": increase_number_of_pixels,
            "decrease_number_of_pixels": decrease_number_of_pixels}


def find_image_params(ctx, jet, params, n_pixels, pixel_size_cm, los_angle, max_iter=40, stripe=None):
    """Iterate the pixel count (square image, even) and the uniform pixel size
    until analyze_tau_stripe asks for nothing (example_simple.py:172-237: the
    size moves by factors from linspace(2, 1.5, 20), the count by
    linspace(1.25, 1, 20), one factor per step in order).  `stripe` computes
    the stripe (default ctx.tau_stripe).  Returns (n_pixels, pixel_size_cm,
    history)."""
    stripe = stripe or ctx.tau_stripe
    steps_size = np.linspace(1.5, 2, 20)[::-1]
    steps_n = np.linspace(1, 1.25, 20)[::-1]
    i_size = i_n = 0
    history = []

    def run(n, size):
        im = abi.ImageDesc()
        im.n_across = im.n_along = n
        im.lg_pixel_size_start_cm = im.lg_pixel_size_stop_cm = math.log10(size)
        im.los_angle = los_angle
        return stripe(jet, im, params)

    tau = run(n_pixels, pixel_size_cm)
    decision = analyze_tau_stripe(tau)
    history.append((n_pixels, pixel_size_cm, decision))
    for _ in range(max_iter):
        if not any(decision.values()) or i_size >= 20 or i_n >= 20:
            break
        if decision["increase_pixel_size"]:
            pixel_size_cm *= steps_size[i_size]
            i_size += 1
        elif decision["decrease_pixels_size"]:
            pixel_size_cm /= steps_size[i_size]
            i_size += 1
        elif decision["increase_number_of_pixels"]:
            n = int(steps_n[i_n] * n_pixels)
            i_n += 1
            n_pixels = n + (n % 2)
        elif decision["decrease_number_of_pixels"]:
            n = int(n_pixels / steps_n[i_n])
            i_n += 1
            n_pixels = n + (n % 2)
        tau = run(n_pixels, pixel_size_cm)
        decision = analyze_tau_stripe(tau)
        history.append((n_pixels, pixel_size_cm, decision))
    return n_pixels, pixel_size_cm, history
