"""Image sizing from a tau stripe (the step before the full run).

`analyze_tau_stripe` applies the decision rule of example_simple.py:105-135
to an optical-depth stripe (Observation::run_stripe, here on the GPU through
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
    """Pixel size and count decisions from a tau stripe (the rule of
    example_simple.py:105-135; same arguments and returned keys).

    The stripe runs from the image centre outwards along the jet.  Two
    markers locate the jet on it: m1, the pixel whose log10 tau is nearest 0
    (the tau = 1 surface), and ms, the pixel nearest log10_tau_small (where
    the jet has turned thin).  With n pixels and f1 = m1 / n:

    * the pixels are too coarse when f1 <= border_tau1[0] or when fewer than
      min_pixels_to_tau1 pixels lead up to the tau = 1 surface, and too fine
      when f1 lies beyond both borders of border_tau1;
    * the image is too short when ms reaches n border_tau_min[1] - 1 (the
      thin end is not on it), too long when ms falls below
      n border_tau_min[0].
    """
    n = len(tau)
    with np.errstate(divide="ignore"):
        lg = np.log10(np.asarray(tau, dtype=np.float64))
    m1 = int(np.argmin(np.abs(lg)))
    ms = int(np.argmin(np.abs(lg - log10_tau_small)))
    f1 = float(m1) / n
    too_coarse = not f1 > border_tau1[0] or m1 < min_pixels_to_tau1
    too_fine = f1 > border_tau1[0] and f1 > border_tau1[1]
    too_short = not ms < n * border_tau_min[1] - 1
    too_long = not too_short and ms < border_tau_min[0] * n
    return {"increase_pixel_size": too_fine,
            "decrease_pixels_size": too_coarse,
            "increase_number_of_pixels": too_short,
            "decrease_number_of_pixels": too_long}


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
