"""Golden vectors for the emission / absorption / Faraday coefficients of
jetshow's src/utils.cpp:8-144, evaluated with mpmath at 50 significant digits.

The formulas are restated here from their mathematical form (Gamma functions
and powers as written in src/utils.cpp, with exact rational arguments such as
11/6), not from the oracle's C code, so a slip in the oracle's transcription
(a wrong Gamma argument, a dropped factor, a wrong exponent or sign) shows up
as a disagreement far above the 1e-13 rounding level.

Inputs are seeded draws of (b, n_los, nu, n, s, gamma_min, thermal fraction)
in the plasma frame, including s in {2.05, 2.5, 3.42} and s near 2, where
k_C_r's 2/(s - 2) factor meets a small difference.  Writes
tests/golden/coefficients.json (inputs as exact doubles, outputs rounded to
doubles).  Run: python tests/golden/gen_coefficients.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "coefficients.json")

# include/utils.h:17-36 (decimal values as the header writes them)
Q_E = mp.mpf("4.80320467299766e-10")
M_E = mp.mpf("9.10938356e-28")
C = mp.mpf("29979245800.0")
PI = mp.pi
R = mp.mpf  # exact conversion of a double


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def norm(a):
    return mp.sqrt(sum(x * x for x in a))


def dot(a, b):
    return sum(x * y for x, y in zip(a, b))


def nu_p(n):                      # utils.cpp:8
    return mp.sqrt(n * Q_E ** 2 / (PI * M_E))


def nu_b_perp(b, nl):             # utils.cpp:10-12
    return Q_E * norm(cross(nl, b)) / (2 * PI * M_E * C)


def nu_b_scalar(b):               # utils.cpp:14-16
    return Q_E * b / (2 * PI * M_E * C)


def nu_b_value(b):                # utils.cpp:22-24
    return Q_E * norm(b) / (2 * PI * M_E * C)


def gamma_factor_ki(s):           # utils.cpp:40
    return mp.power(3, (s + 1) / 2) / 4 * mp.gamma(s / 4 + mp.mpf(11) / 6) * mp.gamma(s / 4 + mp.mpf(1) / 6)


def gamma_factor_ei(s):           # utils.cpp:120
    return mp.power(3, s / 2) / (2 * (s + 1)) * mp.gamma(s / 4 + mp.mpf(19) / 12) * mp.gamma(s / 4 - mp.mpf(1) / 12)


def vector_point(b, nl, nu, n, s, gmin, fth):
    """All vector-B functions at one plasma-frame point."""
    nup2 = nu_p(n) ** 2
    nub = nu_b_perp(b, nl)
    nubv = nu_b_value(b)
    cos_t = dot(b, nl) / norm(b)
    x = nub / nu
    k0 = PI * nup2 * nub / (C * nu ** 2)                      # :26-28
    k0v = PI * nup2 * nubv / (C * nu ** 2)                    # :35-37
    e0 = PI * nup2 * nub * M_E / C                            # :107-109
    e0v = PI * nup2 * nubv * M_E / C                          # :116-118
    ki = k0 * mp.power(x, s / 2) * gamma_factor_ki(s)         # :39-43
    ei = e0 * mp.power(x, (s - 1) / 2) * gamma_factor_ei(s)   # :119-122
    kq = (s + 2) / (s + mp.mpf(10) / 3) * ki                  # :63-65
    eq = (s + 1) / (s + mp.mpf(7) / 3) * ei                   # :132-134
    fkv = mp.power(3, s / 2) * (s + 3) * (s + 2) / (4 * (s + 1)) * mp.gamma(s / 4 + mp.mpf(11) / 12) * \
        mp.gamma(s / 4 + mp.mpf(7) / 12)
    kv = -k0v * cos_t * mp.power(x, (s + 1) / 2) * fkv        # :71-75
    fev = mp.power(3, (s - 1) / 2) * (s + 2) / (2 * s) * mp.gamma(s / 4 + mp.mpf(2) / 3) * \
        mp.gamma(s / 4 + mp.mpf(1) / 3)
    ev = -e0v * cos_t * mp.power(x, s / 2) * fev              # :140-144

    def kFc_at(nn):                                           # :77-80
        return 2 * PI * nu_p(nn) ** 2 * nubv * cos_t / (C * nu ** 2)

    def kCc_at(nn):                                           # :82-84
        return -PI * nu_p(nn) ** 2 * nub ** 2 / (C * nu ** 3)
    kFc, kCc = kFc_at(n), kCc_at(n)
    kFr = (s + 2) * mp.log(gmin) / ((s + 1) * mp.power(gmin, s + 1)) * kFc           # :86-91
    kCr = 2 / (s - 2) * (mp.power(gmin, 2 - s) - mp.power(x, (s - 2) / 2)) * kCc    # :93-98
    # the size of k_C_r's two terms: the scale its rounding is measured on
    kCr_scale = abs(2 / (s - 2) * kCc) * max(mp.power(gmin, 2 - s), mp.power(x, (s - 2) / 2))
    kF = kFr + (kFc_at(fth * n) if fth else 0)
    kC = kCr + (kCc_at(fth * n) if fth else 0)
    return [ki, ei, kq, eq, kv, ev, kFc, kCc, kFr, kCr, kF, kC], kCr_scale


def scalar_point(b, nl, nu, n, s):
    """Random-B overloads (utils.cpp:31-33, 46-51, 111-113, 125-130)."""
    nup2 = nu_p(n) ** 2
    nub = nu_b_scalar(b)
    x = nub / nu
    k0 = PI * nup2 * nub / (C * nu ** 2)
    e0 = PI * nup2 * nub * M_E / C
    rk = mp.sqrt(PI / 4) * mp.gamma((6 + s) / 4) / mp.gamma((8 + s) / 4)
    re = mp.sqrt(PI / 4) * mp.gamma((5 + s) / 4) / mp.gamma((7 + s) / 4)
    ki = k0 * mp.power(x, s / 2) * gamma_factor_ki(s) * rk
    ei = e0 * mp.power(x, (s - 1) / 2) * gamma_factor_ei(s) * re
    return [ki, ei] + [mp.mpf(0)] * 10


def main():
    rng = np.random.default_rng(20261204)
    s_fixed = [2.05, 2.5, 3.42, 2.001, 2.2, 2.8]
    cases = []
    for k in range(72):
        s = s_fixed[k] if k < len(s_fixed) else float(rng.uniform(2.01, 3.6))
        gmin = float([10.0, 100.0, 300.0][k % 3])
        fth = float([0.0, 0.1, 1.0][(k // 3) % 3])
        bmag = 10 ** rng.uniform(-3, 0.5)
        bd = rng.normal(size=3)
        b = (bmag * bd / np.linalg.norm(bd)).tolist()
        nd = rng.normal(size=3)
        nl = (nd / np.linalg.norm(nd)).tolist()
        nu = float(10 ** rng.uniform(9, 11.5))
        n = float(10 ** rng.uniform(0, 5))
        scalar = (k % 4 == 3)
        bm = [R(x) for x in b]
        nm = [R(x) for x in nl]
        if scalar:
            vals = scalar_point(R(bmag), nm, R(nu), R(n), R(s))
            scale = mp.mpf(0)
            b_in = [bmag]
        else:
            vals, scale = vector_point(bm, nm, R(nu), R(n), R(s), R(gmin), R(fth))
            b_in = b
        cases.append({"b": b_in, "n_los": nl, "nu": nu, "n": n, "s": s, "gamma_min": gmin, "thermal": fth,
                      "out": [float(v) for v in vals], "kCr_scale": float(scale)})
    names = ["k_i", "eta_i", "k_q", "eta_q", "k_v", "eta_v", "k_F_c", "k_C_c", "k_F_r", "k_C_r", "k_F", "k_C"]
    head = {"source": "src/utils.cpp:8-144 restated, mpmath %s at %d digits" % (mp.__version__, mp.mp.dps),
            "outputs": names}
    with open(OUT, "w") as f:   # one case per line
        f.write(json.dumps(head)[:-1] + ', "cases": [\n')
        f.write(",\n".join(json.dumps(c) for c in cases))
        f.write("\n]}\n")
    print("wrote %d cases to %s" % (len(cases), OUT))


if __name__ == "__main__":
    main()
