"""The reference's text formats (src/utils.cpp:211-236,381-388,
src/SimulationInterpolater.cpp:105-117, src/main.cpp:1447-1462).

The writers are pinned byte for byte against C printf semantics as Python's
%-formatting states them and against the class API's iostream writers (the
reference's own write_2dvector / write_vector form); the reader against the
stream-extraction rules (stop at the first row that does not parse); the grid
builder against known grids in shuffled order and against incomplete and
non-uniform point sets.  On the GPU a field loaded from text gives the image
of the same grid uploaded directly."""
import os
import sys

import numpy as np
import pytest

from jetshow_amd import host, textio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PC = 3.0856775814671913e+18
VALUES = np.array([[0.0, 1.0, -2.5, 1e-300, 123456789.0, 1.5e-7],
                   [3.14159265358979, -0.000123456789, 1e21, 7.0, 65536.0, 2.0 / 3.0]])


def _fmt_map(a, scale=1.0):
    return "".join("".join("%g " % (x / scale) for x in row) + "\n" for row in a)


def test_write_map_format_and_append(tmp_path):
    p = str(tmp_path / "map_i.txt")
    textio.write_map(p, VALUES)
    assert open(p).read() == _fmt_map(VALUES)
    textio.write_map(p, VALUES, scale=2.0)   # appended, as the reference opens the file
    assert open(p).read() == _fmt_map(VALUES) + _fmt_map(VALUES, 2.0)
    textio.write_map(p, VALUES[:1], append=False)
    assert open(p).read() == _fmt_map(VALUES[:1])
    back = textio.read_map(p)
    assert back.shape == (1, 6)
    assert np.allclose(back, VALUES[:1], rtol=5e-6, atol=0)


def test_write_stripe_format(tmp_path):
    p = str(tmp_path / "stripe_tau.txt")
    v = VALUES.ravel()
    textio.write_stripe(p, v, scale=0.5)
    assert open(p).read() == "".join("%g \n" % (x / 0.5) for x in v)


def test_writers_match_class_api_streams(tmp_path):
    sys.path.insert(0, os.path.join(ROOT, "jetshow_amd"))
    import pyjetshow
    a, b = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    pyjetshow.write_2dvector(a, VALUES.tolist(), 3.0)
    pyjetshow.write_2dvector(a, VALUES.tolist())
    textio.write_map(b, VALUES, scale=3.0)
    textio.write_map(b, VALUES)
    assert open(a, "rb").read() == open(b, "rb").read()
    pyjetshow.write_vector(a, VALUES.ravel().tolist(), 1.0, False)
    textio.write_stripe(b, VALUES.ravel(), append=False)
    assert open(a, "rb").read() == open(b, "rb").read()


def test_save_maps_names(tmp_path):
    planes = {"tau": VALUES, "I": VALUES * 2, "l": VALUES * 3, "flags": np.zeros((2, 6), np.uint32)}
    paths = textio.save_maps(str(tmp_path), planes, suffix="_15.4")
    assert [os.path.basename(p) for p in paths] == ["map_tau_15.4.txt", "map_i_15.4.txt", "map_l_15.4.txt"]
    assert np.allclose(textio.read_map(paths[1]), VALUES * 2, rtol=5e-6)


def test_read_field_text_stops_at_first_bad_row(tmp_path):
    p = tmp_path / "nfield.txt"
    p.write_text("1 2 3\n4.5e18\t5e17 6\n  7 8 9 10 11\nx 1 2\n3 4 5\n")
    pts = textio.read_field_text(str(p))
    # numbers are read three at a time regardless of line breaks; "x" ends it
    assert pts.tolist() == [[1, 2, 3], [4.5e18, 5e17, 6], [7, 8, 9]]
    sys.path.insert(0, os.path.join(ROOT, "jetshow_amd"))
    import pyjetshow
    assert pyjetshow.read_from_txt(str(p)) == pts.tolist()
    assert pyjetshow.read_from_txt(str(tmp_path / "missing.txt")) == []
    with pytest.raises(host.JsxError):
        textio.read_field_text(str(tmp_path / "missing.txt"))
    # a number of any length is read whole, as stream extraction reads it
    long_tok = "1" + "0" * 199
    q = tmp_path / "long.txt"
    q.write_text("%s 2 3\n4 5 %s.5e-190\n" % (long_tok, long_tok))
    assert textio.read_field_text(str(q)).tolist() == [[1e199, 2, 3], [4, 5, float(long_tok + ".5e-190")]]


def _grid_points(nz=7, nr=5, seed=3):
    z = np.linspace(0.0, 20.0, nz) * PC
    r = np.linspace(0.0, 0.5, nr) * PC
    Z, R = np.meshgrid(z, r, indexing="ij")
    V = np.random.default_rng(seed).uniform(size=(nz, nr))
    pts = np.stack([Z.ravel(), R.ravel(), V.ravel()], axis=1)
    return pts, V


def test_grid_from_points_regular_shuffled():
    pts, V = _grid_points()
    perm = np.random.default_rng(1).permutation(len(pts))
    vals, (z0, dz, r0, dr) = textio.grid_from_points(pts[perm], nan_value=1.0)
    assert np.array_equal(vals, V)
    assert z0 == 0.0 and r0 == 0.0
    assert abs(dz - 20.0 / 6) < 1e-14 and abs(dr - 0.5 / 4) < 1e-15


def test_grid_from_points_rejects_non_grids():
    pts, _ = _grid_points()
    with pytest.raises(host.JsxError) as e:
        textio.grid_from_points(pts[:-1])          # a node missing
    assert e.value.code == -4
    bad = pts.copy()
    bad[bad[:, 0] == bad[:, 0].max(), 0] *= 1.1    # last z row moved: spacing not uniform
    with pytest.raises(host.JsxError):
        textio.grid_from_points(bad)
    dup = np.concatenate([pts[:-1], pts[:1]])      # right count, one node twice
    with pytest.raises(host.JsxError):
        textio.grid_from_points(dup)


@pytest.mark.gpu
def test_field_grids_from_text_give_same_image(ctx, tmp_path):
    from jetshow_amd import grf, models
    (B, N, G), geo = grf.jet_grids(64, 16)
    print("grid from text", flush=True)
    direct = [ctx.upload_grid(v, *geo, nan_value=nv) for v, nv in ((B, 0.0), (N, 0.0), (G, 1.0))]
    z = geo[0] + geo[1] * np.arange(B.shape[0])
    r = geo[2] + geo[3] * np.arange(B.shape[1])
    Z, R = np.meshgrid(z * PC, r * PC, indexing="ij")
    loaded = []
    for name, v, nv in (("b", B, 0.0), ("n", N, 0.0), ("v", G, 1.0)):
        p = str(tmp_path / ("%sfield.txt" % name))
        pts = np.stack([Z.ravel(), R.ravel(), v.ravel()], axis=1)
        np.savetxt(p, pts[np.random.default_rng(2).permutation(len(pts))], fmt="%.17g")
        loaded.append(textio.load_field_grid(ctx, p, nan_value=nv))
    j1, im, p = models.grid_jet_model(direct, n_side=32, pixel_pc=0.3)
    j2, _, _ = models.grid_jet_model(loaded, n_side=32, pixel_pc=0.3)
    a, b = ctx.observe(j1, im, p), ctx.observe(j2, im, p)
    assert np.array_equal(a["flags"], b["flags"])
    assert (a["I"] > 0).any()
    assert np.allclose(a["I"], b["I"], rtol=1e-9, atol=0)
    assert np.allclose(a["tau"], b["tau"], rtol=1e-9, atol=0)
