"""CPU checks of the C ABI: the HIP library loads without a GPU, exports every
entry point include/jetshow_amd.h declares, its structures match the Python
mirror, host-only helpers agree with the oracle, and compute entry points
refuse to run (no CPU fallback) when no GPU is visible."""
import os
import re

import numpy as np
import pytest

from jetshow_amd import abi, host, models
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "jetshow_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(jsx_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = host.lib()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(L, s), s
    assert set(host.ABI_SYMBOLS) <= set(syms)


def test_struct_layout_matches_header():
    L = host.lib()
    import ctypes
    for i, t in enumerate([abi.JetDesc, abi.ImageDesc, abi.SolverParams, abi.Outputs, abi.GridDesc, abi.Request,
                            abi.GridPayload, abi.RequestTiming]):
        assert L.jsx_sizeof(i) == ctypes.sizeof(t), t
    assert L.jsx_abi_version() == abi.ABI_VERSION == 2


def test_pixel_sizes_and_cosmology_match_oracle():
    j, im, p = models.get_i_image_model(0.663225, 0.0165, 50, 10, 0.001, 0.1, 0.0683, 0.1, 1.0, 1e4, 2.0, 5.0, 15.35,
                                        1e5, False, mas_to_pc=oracle.mas_to_pc)
    sizes, _, _ = oracle.image(im)
    assert np.array_equal(host.pixel_sizes(im).reshape(-1), sizes)
    for z in (0.0165, 0.1, 0.5):
        assert host.mas_to_pc(z) == oracle.mas_to_pc(z)


def test_no_gpu_no_compute():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(host.JsxError):
        host.Context(0)
