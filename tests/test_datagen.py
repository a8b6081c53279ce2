"""Seeded workload generators and the bin_data file layout (CPU only).

The closed-form ellipse projector is pinned by mathematics (area, the circle's
chord, rotation covariance); the pixel projector that makes the other
workloads' data is pinned against it, with an error that must shrink as the
grid is refined."""
import os
import subprocess

import numpy as np
import pytest

from timbir_amd import bindata, datagen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# --------------------------------------------------------------------------
# closed-form projections
# --------------------------------------------------------------------------

@pytest.mark.parametrize("profile", ["box", "triangle"])
def test_ellipse_projection_preserves_area(profile):
    # every view integrates to v * pi * a * b (a unit-area bin profile)
    ell = np.array([[3.0, -2.0, 12.0, 5.0, 0.4, 0.7], [-4.0, 6.0, 3.0, 8.0, 2.0, 0.2]])
    ang = np.linspace(0, np.pi, 11, endpoint=False)
    p = datagen.project_ellipses(ell, ang, 64, 32.0, profile=profile)
    mass = 0.7 * np.pi * 12 * 5 + 0.2 * np.pi * 3 * 8
    np.testing.assert_allclose(p.sum(axis=1), mass, rtol=1e-7)


def test_circle_projection_is_the_chord_length():
    # a circle of radius R: chord 2 sqrt(R^2 - d^2), averaged over each bin
    # by quadrature, against the box profile's closed form
    R, x0, y0 = 9.3, 2.5, -1.25
    ell = np.array([[x0, y0, R, R, 0.0, 1.0]])
    ang = np.array([0.0, 0.3, 1.2, 2.9])
    p = datagen.project_ellipses(ell, ang, 40, 20.0, profile="box")
    for v, th in enumerate(ang):
        r0 = x0 * np.cos(th) - y0 * np.sin(th)
        for j in range(40):
            r = np.linspace(j - 20.0, j - 19.0, 20001)
            chord = 2 * np.sqrt(np.maximum(R * R - (r - r0) ** 2, 0))
            assert abs(np.trapezoid(chord, r) - p[v, j]) < 2e-6


def test_ellipse_projection_rotates_with_the_object():
    # rotating the object by delta turns its view at theta into the
    # original's view at theta + delta (r = Re((x + iy) e^{i theta}))
    x0, y0, a, b, t = 5.0, -3.0, 10.0, 4.0, 0.3
    delta = 0.7
    xr = x0 * np.cos(delta) - y0 * np.sin(delta)
    yr = x0 * np.sin(delta) + y0 * np.cos(delta)
    ang = np.array([0.1, 0.9, 2.0])
    p0 = datagen.project_ellipses(np.array([[x0, y0, a, b, t, 1.0]]), ang + delta, 48, 24.0)
    p1 = datagen.project_ellipses(np.array([[xr, yr, a, b, t + delta, 1.0]]), ang, 48, 24.0)
    # (rounding of the rotated parameters moves the chord ends by ~1e-14,
    # which the square root at a chord end turns into ~1e-6)
    np.testing.assert_allclose(p1, p0, atol=2e-5)


def test_pixel_projector_converges_to_closed_form():
    # project() of the supersampled raster against the closed form with the
    # same triangular bin profile: relative L1 error below 0.7 % at 64
    # pixels, below 0.3 % at 128, and shrinking with the grid
    ang = np.linspace(0, np.pi, 7, endpoint=False)
    errs = []
    for N in (64, 128):
        k = N / 64
        ell = np.array([[3.0 * k, -2.0 * k, 20.0 * k, 12.0 * k, 0.4, 1.0]])
        img = datagen.rasterize_ellipses(ell, N, N / 2, sub=16)
        ps = datagen.project(img[None], ang.astype(np.float32), N, N / 2)[:, 0]
        pt = datagen.project_ellipses(ell, ang, N, N / 2)
        errs.append(float((np.abs(ps - pt).sum(axis=1) / pt.sum(axis=1)).max()))
    assert errs[0] < 7e-3 and errs[1] < 3e-3 and errs[1] < 0.6 * errs[0], errs


def test_raster_keeps_the_area():
    ell = np.array([[1.5, -2.0, 11.0, 6.0, 1.1, 1.0]])
    img = datagen.rasterize_ellipses(ell, 48, 24.0, sub=16)
    assert abs(img.sum() - np.pi * 11 * 6) < 0.5


# --------------------------------------------------------------------------
# the workloads
# --------------------------------------------------------------------------

def test_batched_problems_are_seeded_and_uneven():
    b1 = datagen.batched(n_problems=6, sizes=(32, 48), views=(24, 36))
    b2 = datagen.batched(n_problems=6, sizes=(32, 48), views=(24, 36))
    assert [datagen.sha256(s["projections"]) for s in b1] == \
        [datagen.sha256(s["projections"]) for s in b2]
    shapes = set()
    for s in b1:
        N, V = s["proj_cols"], s["proj_num"]
        assert N in (32, 48) and V in (24, 36)
        assert s["projections"].shape == (V, 4, N) and s["weights"].shape == (V, 4, N)
        assert s["recon_num"] == 1 and s["proj_rows"] == 4 and s["quad_convex"] == 1
        assert 5e3 <= s["I0"] <= 5e4
        # every view lies inside the one frame (half-open intervals)
        assert s["proj_times"].max() < s["recon_times"][-1]
        assert abs(float(s["truth"].max()) - 0.02) < 1e-7 and s["truth"].min() >= 0
        assert np.all(s["weights"] >= 1)
        shapes.add((N, V))
    assert len(shapes) > 1
    b3 = datagen.batched(n_problems=6, sizes=(32, 48), views=(24, 36), seed=1)
    assert datagen.sha256(b3[0]["projections"]) != datagen.sha256(b1[0]["projections"])


def test_precision_scan_matches_its_phantom():
    # every view of every slice integrates to the slice's mass (up to noise)
    s = datagen.precision3d(N=64, nz=6, n_views=60, I0=1e6)
    assert s["projections"].shape == (60, 6, 64) and s["truth"].shape == (1, 6, 64, 64)
    assert s["remove_rings"] == 1 and s["rings"].shape == (6, 64)
    truth = s["truth"][0]
    assert 0 <= truth.min() and truth.max() <= 0.02 + 1e-7
    # slices differ (ellipsoid sections), the microstructure does not
    assert len({float(truth[z].sum()) for z in range(6)}) > 1
    clean = s["projections"] - s["rings"][None]
    for z in range(6):
        np.testing.assert_allclose(clean[:, z].sum(axis=1).mean(), truth[z].sum(), rtol=5e-3)


def test_deploy_scan_is_the_interlaced_model():
    s = datagen.deploy4d(N=32, nz=4, n_frames=2, n_views=64, ch_steps=4, frame_steps=2)
    assert s["projections"].shape == (64, 4, 32) and s["recon_num"] == 2
    assert s["quad_convex"] == 0 and s["remove_rings"] == 1
    i = datagen.interlaced4d(N=32, nz=4, n_frames=2, n_views=64, ch_steps=4, frame_steps=2,
                             I0=1e4, seed=20260005)
    np.testing.assert_array_equal(s["projections"], i["projections"])
    assert set(datagen.WORKLOADS) == {"plain3d", "interlaced4d", "batched", "precision3d",
                                      "deploy4d"}


# --------------------------------------------------------------------------
# bin_data layout
# --------------------------------------------------------------------------

def small():
    return datagen.tiny(N=16, nz=4, n_frames=2, n_views=24, remove_rings=2, quad_convex=0)


def test_bin_files_round_trip(tmp_path):
    s = small()
    bindata.write(s, tmp_path)
    assert os.path.getsize(tmp_path / "projections.bin") == 4 * 24 * 4 * 16
    assert os.path.getsize(tmp_path / "recon_times.bin") == 4 * 3
    got = bindata.read(tmp_path, 4, 16, 24, 2)
    for k in bindata.INPUTS:
        np.testing.assert_array_equal(got[k], np.asarray(s[k], np.float32))


def test_bin_read_rejects_short_files(tmp_path):
    s = small()
    bindata.write(s, tmp_path)
    with pytest.raises(ValueError, match="recon_times"):
        bindata.read(tmp_path, 4, 16, 24, 3)
    bad = dict(s, proj_num=25)
    with pytest.raises(ValueError, match="projections"):
        bindata.write(bad, tmp_path)


def test_command_line_carries_float32_values():
    s = small()
    for k in ("sig_s", "sig_t", "vox_wid", "rot_center"):
        s[k] = float(np.float32(s[k] * 1.1))
    args = bindata.command_line(s, restart=True)
    assert args[0] == bindata.DRIVER and args[-1] == "--restart"
    opts = dict(zip(args[1:-2:2], args[2:-2:2])) if "--quad_convex" in args else \
        dict(zip(args[1:-1:2], args[2:-1:2]))
    for k in ("sig_s", "sig_t", "vox_wid", "rot_center"):
        assert np.float32(float(opts[f"--{k}"])) == np.float32(s[k])
    assert opts["--remove_rings"] == "2" and "--quad_convex" not in args


def test_driver_reports_a_short_weights_file(tmp_path):
    # files from write(), weights cut short: the driver stops while reading,
    # before any device work, with the reference's message and status 1
    if not os.path.exists(bindata.DRIVER):
        pytest.fail("build() makes timbir_amd/tb_bin_data")
    s = small()
    bindata.write(s, tmp_path)
    w = np.fromfile(tmp_path / "weights.bin", np.float32)
    w[:-1].tofile(tmp_path / "weights.bin")
    r = subprocess.run(bindata.command_line(s), cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    assert "read_BinFile" in r.stdout and "weights.bin" in r.stdout
