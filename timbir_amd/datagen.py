"""Seeded synthetic scans in the layout ``reconstruct()`` consumes.

Every array is produced by numpy's PCG64 from an explicit seed, so a scan is
reproducible bit for bit on any machine with the same numpy.  A scan is a dict
holding the 23 ``reconstruct()`` arguments (projections/weights in the
reference input layout ``[proj_num][proj_rows][proj_cols]``, reference
``src/include/mbir4d.h:34``) plus the ground-truth phantom under ``truth``.

The measurement model: parallel beam, detector column ``j`` centred at
``(j + 0.5 - rot_center) * vox_wid``, a pixel at column ``k``, row ``i`` sits at
``x = (k + 0.5 - rot_center) * vox_wid``, ``y = (i + 0.5 - N/2) * vox_wid`` and
projects to ``r = x cos(theta) - y sin(theta)`` (reference
``src/MBIR_4D/XT_AMatrix.c:57-64``).  Projections are computed by 3x3
sub-pixel splatting with linear interpolation; counts are Poisson with mean
``I0 exp(-Ax)``; ``weights = counts`` and ``projections = -log(counts / I0)``.
"""
from __future__ import annotations

import hashlib
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

# ---------------------------------------------------------------------------
# view schedules
# ---------------------------------------------------------------------------


def bitrev(v: int, bits: int) -> int:
    out = 0
    for j in range(bits):
        if v & (1 << j):
            out |= 1 << (bits - 1 - j)
    return out


def interlaced_views(K: int, N_theta: int, N_p: int, proj_start: int = 0):
    """Time-interlaced view schedule of the reference's acquisition driver
    (src/reconstruct/aps_data/XT_Views.c:41-80, with the minimum-angle clip
    disabled): view i has angle index ``i*K + bitrev(floor(i/L) mod K)`` with
    ``L = N_theta/K``, angle ``index * pi / N_theta`` and time ``index / K``."""
    L = N_theta // K
    bits = int(np.floor(np.log(K) / np.log(2.0) + 0.5))
    idx = np.arange(proj_start + N_p, dtype=np.int64)
    sub = (idx // L) % K
    br = np.array([bitrev(int(s), bits) for s in range(K)], dtype=np.int64)[sub]
    a = (idx * K + br).astype(np.float32)
    times = (a / np.float32(K)).astype(np.float32)
    dtheta = np.float32(np.pi / N_theta)
    angles = (a * dtheta).astype(np.float32)
    return angles[proj_start:], times[proj_start:]


# ---------------------------------------------------------------------------
# phantoms
# ---------------------------------------------------------------------------

# modified 3D Shepp-Logan ellipsoids: (value, a, b, c, x0, y0, z0, phi_deg)
_SL3D = [
    (1.0, 0.6900, 0.920, 0.810, 0.0, 0.0, 0.0, 0.0),
    (-0.8, 0.6624, 0.874, 0.780, 0.0, -0.0184, 0.0, 0.0),
    (-0.2, 0.1100, 0.310, 0.220, 0.22, 0.0, 0.0, -18.0),
    (-0.2, 0.1600, 0.410, 0.280, -0.22, 0.0, 0.0, 18.0),
    (0.1, 0.2100, 0.250, 0.410, 0.0, 0.35, -0.15, 0.0),
    (0.1, 0.0460, 0.046, 0.050, 0.0, 0.1, 0.25, 0.0),
    (0.1, 0.0460, 0.046, 0.050, 0.0, -0.1, 0.25, 0.0),
    (0.1, 0.0460, 0.023, 0.050, -0.08, -0.605, 0.0, 0.0),
    (0.1, 0.0230, 0.023, 0.020, 0.0, -0.606, 0.0, 0.0),
    (0.1, 0.0230, 0.046, 0.020, 0.06, -0.605, 0.0, 0.0),
]


def shepp_logan_3d(N: int, nz: int, vmax: float = 0.02) -> np.ndarray:
    """[nz][N][N] modified Shepp-Logan sampled at the central nz slices of
    z in [-0.25, 0.25], scaled to [0, vmax]."""
    c = (np.arange(N) + 0.5) / N * 2 - 1
    zs = np.linspace(-0.25, 0.25, nz) if nz > 1 else np.zeros(1)
    X, Y = np.meshgrid(c, -c)
    out = np.zeros((nz, N, N))
    for s, z in enumerate(zs):
        img = np.zeros((N, N))
        for v, a, b, cc, x0, y0, z0, phi in _SL3D:
            t = np.deg2rad(phi)
            xr = (X - x0) * np.cos(t) + (Y - y0) * np.sin(t)
            yr = -(X - x0) * np.sin(t) + (Y - y0) * np.cos(t)
            inside = (xr / a) ** 2 + (yr / b) ** 2 + ((z - z0) / cc) ** 2 <= 1
            img[inside] += v
        out[s] = img
    out = np.clip(out, 0, None)
    out *= vmax / max(out.max(), 1e-12)
    return out.astype(np.float32)


def random_ellipses(rng: np.random.Generator, N: int, n_min=3, n_max=12, vmax=0.02) -> np.ndarray:
    c = (np.arange(N) + 0.5) / N * 2 - 1
    X, Y = np.meshgrid(c, -c)
    img = np.zeros((N, N))
    for _ in range(int(rng.integers(n_min, n_max + 1))):
        a, b = rng.uniform(0.05, 0.5, 2)
        x0, y0 = rng.uniform(-0.5, 0.5, 2)
        t = rng.uniform(0, np.pi)
        v = rng.uniform(0.1, 1.0)
        xr = (X - x0) * np.cos(t) + (Y - y0) * np.sin(t)
        yr = -(X - x0) * np.sin(t) + (Y - y0) * np.cos(t)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1] += v
    img *= vmax / max(img.max(), 1e-12)
    return img.astype(np.float32)


def _lap(u):
    return (np.roll(u, 1, 0) + np.roll(u, -1, 0) + np.roll(u, 1, 1) + np.roll(u, -1, 1) - 4 * u)


def cahn_hilliard(rng: np.random.Generator, N: int, steps: int, gamma=0.5, dt=0.02,
                  c0: np.ndarray | None = None) -> np.ndarray:
    """Explicit Cahn-Hilliard c_t = lap(c^3 - c - gamma lap c) on a periodic
    N x N grid from small seeded noise (or from c0)."""
    c = rng.normal(0, 0.05, (N, N)) if c0 is None else c0.copy()
    for _ in range(steps):
        mu = c ** 3 - c - gamma * _lap(c)
        c = c + dt * _lap(mu)
    return c


def two_phase(c: np.ndarray, lo=0.005, hi=0.015) -> np.ndarray:
    return np.where(c > 0, hi, lo).astype(np.float32)


def disk(N: int, frac=0.98) -> np.ndarray:
    c = np.arange(N) + 0.5 - N / 2
    X, Y = np.meshgrid(c, c)
    return (X ** 2 + Y ** 2 < (frac * N / 2) ** 2).astype(np.float32)


# ---------------------------------------------------------------------------
# ellipses in detector units, with closed-form projections
# ---------------------------------------------------------------------------
#
# An ellipse is a row (x0, y0, a, b, t, v) in the projector's coordinates
# (pixel units times vox_wid, origin on the rotation axis): centre (x0, y0),
# semi-axis a along (cos t, sin t), b along (-sin t, cos t), value v added
# inside.


def draw_ellipses(rng: np.random.Generator, R: float, n_min=3, n_max=12) -> np.ndarray:
    """Seeded ellipses inside the disk of radius R: n_min..n_max of them,
    semi-axes U(0.05, 0.5) R, centres U(-0.35, 0.35) R per coordinate (so the
    whole ellipse stays inside R), angle U(0, pi), value U(0.1, 1)."""
    n = int(rng.integers(n_min, n_max + 1))
    out = np.empty((n, 6))
    for q in range(n):
        a, b = rng.uniform(0.05, 0.5, 2) * R
        x0, y0 = rng.uniform(-0.35, 0.35, 2) * R
        out[q] = (x0, y0, a, b, rng.uniform(0, np.pi), rng.uniform(0.1, 1.0))
    return out


def rasterize_ellipses(ell: np.ndarray, N: int, rot_center: float, vox_wid: float = 1.0,
                       sub: int = 4) -> np.ndarray:
    """[N][N] image: each pixel the mean over a sub x sub grid of points of
    the summed ellipse values (pixel (i, k) centred at
    x = (k + 0.5 - rot_center) vox_wid, y = (i + 0.5 - N/2) vox_wid)."""
    off = (np.arange(sub) + 0.5) / sub - 0.5
    xs = ((np.arange(N)[:, None] + 0.5 + off[None, :] - rot_center) * vox_wid).reshape(-1)
    ys = ((np.arange(N)[:, None] + 0.5 + off[None, :] - N / 2) * vox_wid).reshape(-1)
    X, Y = np.meshgrid(xs, ys)
    img = np.zeros_like(X)
    for x0, y0, a, b, t, v in ell:
        u = (X - x0) * np.cos(t) + (Y - y0) * np.sin(t)
        w = -(X - x0) * np.sin(t) + (Y - y0) * np.cos(t)
        img += v * ((u / a) ** 2 + (w / b) ** 2 <= 1)
    return img.reshape(N, sub, N, sub).mean(axis=(1, 3))


def _chord_integrals(d, s2):
    """First and second antiderivatives of the unit chord profile
    sqrt(s^2 - d^2) (zero outside |d| < s), both zero at d = -s:
    G(d) = (d sqrt(s^2 - d^2) + s^2 asin(d / s)) / 2 + pi s^2 / 4 and
    H(d) = integral of G = ((2 s^2 + d^2) sqrt(s^2 - d^2)) / 6
           + s^2 (d asin(d / s) + pi d / 2) / 2 + pi s^4 / 12 ... continued
    linearly past d = s."""
    sw = np.sqrt(s2)
    dc = np.clip(d, -sw, sw)
    rt = np.sqrt(np.maximum(s2 - dc * dc, 0))
    asn = np.arcsin(dc / sw)
    G = 0.5 * (dc * rt + s2 * asn) + 0.25 * np.pi * s2
    # integral of G from -s to dc, then the constant G(s) = pi s^2 / 2 beyond s
    H = (0.5 * (-(rt ** 3) / 3 + s2 * (dc * asn + rt)) + 0.25 * np.pi * s2 * dc
         + 0.25 * np.pi * s2 * sw)
    H = H + 0.5 * np.pi * s2 * np.maximum(d - sw, 0)
    return G, H


def project_ellipses(ell: np.ndarray, angles: np.ndarray, n_bins: int, rot_center: float,
                     vox_wid: float = 1.0, profile: str = "triangle") -> np.ndarray:
    """[n_views][n_bins] exact line integrals of the ellipses, weighted per
    detector bin j (centre r_j = (j + 0.5 - rot_center) vox_wid, ray offset
    r = x cos(theta) - y sin(theta)) by a profile of unit area:
    "triangle" (half-width one bin: the linear interpolation that project()
    applies to each sample, so both agree up to rasterization) or "box" (the
    mean over the bin).

    Along direction n the ellipse has half-width s with s^2 = a^2 (n.e1)^2 +
    b^2 (n.e2)^2 and chord length (2ab / s^2) sqrt(s^2 - d^2) at offset d
    from its centre's ray; the profile-weighted integrals of that are
    differences of its antiderivatives G and H."""
    th = np.asarray(angles, np.float64)[:, None]
    c, s_ = np.cos(th), np.sin(th)
    h = float(vox_wid)
    centres = (np.arange(n_bins) + 0.5 - rot_center) * h
    out = np.zeros((len(th), n_bins))
    for x0, y0, a, b, t, v in ell:
        # n = (cos th, -sin th); e1 = (cos t, sin t); e2 = (-sin t, cos t)
        n1 = c * np.cos(t) - s_ * np.sin(t)
        n2 = -c * np.sin(t) - s_ * np.cos(t)
        s2 = (a * n1) ** 2 + (b * n2) ** 2
        d = centres[None, :] - (x0 * c - y0 * s_)
        if profile == "box":
            G1, _ = _chord_integrals(d + 0.5 * h, s2)
            G0, _ = _chord_integrals(d - 0.5 * h, s2)
            w = (G1 - G0) / h
        elif profile == "triangle":
            _, Hp = _chord_integrals(d + h, s2)
            _, Hc = _chord_integrals(d, s2)
            _, Hm = _chord_integrals(d - h, s2)
            w = (Hp - 2 * Hc + Hm) / (h * h)
        else:
            raise ValueError(f"unknown profile {profile!r}")
        out += v * (2 * a * b / s2) * w
    return out


# ---------------------------------------------------------------------------
# projection and noise
# ---------------------------------------------------------------------------


def project(obj_slices: np.ndarray, angles: np.ndarray, n_bins: int, rot_center: float,
            vox_wid: float = 1.0, sub: int = 3) -> np.ndarray:
    """Line integrals of obj_slices [nz][N][N] for every angle ->
    [n_views][nz][n_bins] (reference input layout)."""
    nz, N, _ = obj_slices.shape
    if nz > 1 and all(np.array_equal(obj_slices[0], obj_slices[s]) for s in range(1, nz)):
        one = project(obj_slices[:1], angles, n_bins, rot_center, vox_wid, sub)
        return np.repeat(one, nz, axis=1)
    off = (np.arange(sub) + 0.5) / sub
    k = (np.arange(N)[None, :, None, None] + off[None, None, None, :])
    i = (np.arange(N)[:, None, None, None] + off[None, None, :, None])
    X = ((k - rot_center) * vox_wid) * np.ones_like(i)
    Y = ((i - N / 2) * vox_wid) * np.ones_like(k)
    X = X.reshape(-1)
    Y = Y.reshape(-1)
    vals = np.repeat(obj_slices.reshape(nz, N * N), sub * sub, axis=1) * (vox_wid ** 2 / sub ** 2)
    out = np.zeros((len(angles), nz, n_bins), np.float64)
    ang = angles.astype(np.float64)

    def one_view(v):
        th = ang[v]
        r = X * np.cos(th) - Y * np.sin(th)
        u = r / vox_wid + rot_center - 0.5          # fractional bin index
        j0 = np.floor(u).astype(np.int64)
        fr = u - j0
        for s in range(nz):
            w = vals[s]
            acc = np.zeros(n_bins + 2)
            acc += np.bincount(np.clip(j0 + 1, 0, n_bins + 1), weights=w * (1 - fr), minlength=n_bins + 2)
            acc += np.bincount(np.clip(j0 + 2, 0, n_bins + 1), weights=w * fr, minlength=n_bins + 2)
            out[v, s] = acc[1:n_bins + 1] / vox_wid

    # views are independent and numpy releases the GIL in these loops: a
    # thread per share of the views gives the same numbers as one thread
    workers = min(16, os.cpu_count() or 1, len(ang))
    if workers > 1 and len(ang) * N * N >= (1 << 22):
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(one_view, range(len(ang))))
    else:
        for v in range(len(ang)):
            one_view(v)
    return out


def noisy_scan(rng: np.random.Generator, line_int: np.ndarray, I0) -> tuple:
    lam = np.asarray(I0, np.float64) * np.exp(-line_int)
    counts = np.maximum(rng.poisson(lam), 1).astype(np.float64)
    proj = -np.log(counts / I0)
    return proj.astype(np.float32), counts.astype(np.float32)


def sha256(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _scan(proj, wts, angles, times, recon_times, rows, cols, recon_num, truth, **kw):
    scan = dict(projections=proj, weights=wts, proj_angles=angles.astype(np.float32),
                proj_times=times.astype(np.float32), recon_times=recon_times.astype(np.float32),
                proj_rows=int(rows), proj_cols=int(cols), proj_num=int(len(angles)),
                recon_num=int(recon_num), vox_wid=1.0, rot_center=float(cols) / 2,
                sig_s=0.01, sig_t=1.0, c_s=1e-2, c_t=1e-2, convg_thresh=1.0,
                remove_rings=0, quad_convex=1, huber_delta=0.5, huber_T=4.0, truth=truth)
    scan.update(kw)
    return scan


# ---------------------------------------------------------------------------
# the workloads
# ---------------------------------------------------------------------------


def plain3d(N=128, nz=4, n_views=256, I0=2e4, seed=20260001, **kw) -> dict:
    """Plain 3D MBIR: 3D Shepp-Logan N x N x nz, n_views uniform on [0, pi),
    N bins, Poisson I0, quadratic data term (zinger model off), no rings,
    qGGMRF p=1.2 q=2 c=1e-2.  recon_times covers every view."""
    rng = np.random.default_rng(seed)
    truth = shepp_logan_3d(N, nz) * disk(N)[None]
    angles = (np.arange(n_views) * (np.pi / n_views)).astype(np.float32)
    times = np.arange(n_views, dtype=np.float32)
    li = project(truth, angles, N, N / 2)
    proj, wts = noisy_scan(rng, li, I0)
    rt = np.array([0, n_views], np.float32)
    return _scan(proj, wts, angles, times, rt, nz, N, 1, truth[None], **kw)


def interlaced4d(N=256, nz=8, n_frames=16, n_views=1024, K=8, I0=1e4, zinger_frac=1e-3,
                 zinger_gain=5.0, ring_sigma=0.01, ch_steps=64, frame_steps=8,
                 seed=20260002, **kw) -> dict:
    """4D interlaced-view MBIR: a two-phase Cahn-Hilliard microstructure
    (phases 0.005 / 0.015) stacked to nz slices and evolved over n_frames,
    n_views interlaced views in K sub-frames over one half turn, N bins,
    Poisson I0, zingers (zinger_frac of the bins scaled by zinger_gain) and
    per-detector ring offsets N(0, ring_sigma).  Generalized-Huber data term
    and ring correction on."""
    rng = np.random.default_rng(seed)
    c = cahn_hilliard(rng, N, ch_steps)
    m = disk(N)
    frames = []
    for f in range(n_frames):
        frames.append(np.repeat((two_phase(c) * m)[None], nz, axis=0))
        c = cahn_hilliard(rng, N, frame_steps, c0=c)
    truth = np.stack(frames)  # [T][nz][N][N]
    angles, times = interlaced_views(K, n_views, n_views)
    rt = np.linspace(0, float(n_views), n_frames + 1).astype(np.float32)
    frame_of = np.clip(np.searchsorted(rt, times, side="right") - 1, 0, n_frames - 1)
    li = np.zeros((n_views, nz, N))
    for f in range(n_frames):
        sel = np.nonzero(frame_of == f)[0]
        if len(sel):
            li[sel] = project(truth[f], angles[sel], N, N / 2)
    proj, wts = noisy_scan(rng, li, I0)
    rings = rng.normal(0, ring_sigma, (nz, N)).astype(np.float32)
    proj = proj + rings[None]
    nzing = int(round(zinger_frac * proj.size))
    zi = rng.choice(proj.size, nzing, replace=False)
    flat = proj.reshape(-1)
    flat[zi] = flat[zi] * zinger_gain
    return _scan(proj, wts, angles, times, rt, nz, N, n_frames, truth,
                 sig_s=0.02, sig_t=5e-4, c_s=1e-2, c_t=1e-2, remove_rings=1, quad_convex=0,
                 huber_delta=0.5, huber_T=4.0, rings=rings, zinger_index=zi, **kw)


def tiny(N=16, nz=4, n_frames=2, n_views=24, I0=2e4, seed=7, rings=0.0, zingers=0.0,
         interlace_K=0, **kw) -> dict:
    """Small seeded scan for unit tests: random ellipses per frame."""
    rng = np.random.default_rng(seed)
    base = random_ellipses(rng, N, 3, 6) * disk(N)
    truth = np.stack([np.stack([np.roll(base, f, axis=1) * disk(N) for _ in range(nz)])
                      for f in range(n_frames)])
    if interlace_K:
        angles, times = interlaced_views(interlace_K, n_views, n_views)
    else:
        angles = (np.arange(n_views) * (np.pi / n_views)).astype(np.float32)
        times = np.arange(n_views, dtype=np.float32)
    rt = np.linspace(0, float(n_views), n_frames + 1).astype(np.float32)
    frame_of = np.clip(np.searchsorted(rt, times, side="right") - 1, 0, n_frames - 1)
    li = np.zeros((n_views, nz, N))
    for f in range(n_frames):
        sel = np.nonzero(frame_of == f)[0]
        if len(sel):
            li[sel] = project(truth[f], angles[sel], N, N / 2)
    proj, wts = noisy_scan(rng, li, I0)
    if rings:
        proj = proj + rng.normal(0, rings, (nz, N)).astype(np.float32)[None]
    if zingers:
        zi = rng.choice(proj.size, max(1, int(zingers * proj.size)), replace=False)
        proj.reshape(-1)[zi] *= 5.0
    return _scan(proj.astype(np.float32), wts, angles, times, rt, nz, N, n_frames, truth, **kw)


def _phantom_scale(raster_max: float, vmax: float) -> float:
    return vmax / max(raster_max, 1e-12)


def batched(n_problems=64, sizes=(128, 192, 256, 384, 512), views=(180, 360, 720), rows=4,
            I0_range=(5e3, 5e4), vmax=0.02, seed=20260003, **kw) -> list:
    """Batch of independent 3D problems of uneven size: per problem a square
    image width N drawn from sizes (N bins), a view count from views (uniform
    on [0, pi)), 3-12 seeded ellipses scaled to [0, vmax], and Poisson counts
    at I0 ~ U(I0_range).  rows identical slices (the reference's slab
    partition needs at least 4 rows), recon_num 1, quadratic data term.
    Line integrals are the ellipses' closed-form projections with project()'s
    triangular bin profile; ``truth`` is their 4x4-supersampled raster."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_problems):
        N = int(rng.choice(sizes))
        V = int(rng.choice(views))
        I0 = float(rng.uniform(*I0_range))
        ell = draw_ellipses(rng, 0.95 * N / 2)
        img = rasterize_ellipses(ell, N, N / 2)
        k = _phantom_scale(img.max(), vmax)
        ell[:, 5] *= k
        truth = np.repeat((img * k).astype(np.float32)[None], rows, axis=0)
        angles = (np.arange(V) * (np.pi / V)).astype(np.float32)
        li = project_ellipses(ell, angles, N, N / 2)
        proj, wts = noisy_scan(rng, np.repeat(li[:, None, :], rows, axis=1), I0)
        times = np.arange(V, dtype=np.float32)
        rt = np.array([0, V], np.float32)
        out.append(_scan(proj, wts, angles, times, rt, rows, N, 1, truth[None], I0=I0, **kw))
    return out


def precision3d(N=512, nz=32, n_views=1440, I0=2e4, ring_sigma=0.005, ch_steps=64,
                seed=20260004, **kw) -> dict:
    """Precision-study scan: one frame, N x N x nz, n_views uniform on
    [0, pi), N bins, Poisson I0, per-detector ring offsets N(0, ring_sigma).
    The phantom mixes a two-phase Cahn-Hilliard microstructure (0.002 /
    0.008 inside the reconstruction disk, the same in every slice) with 3-12
    seeded ellipsoids (axes in the slice plane from draw_ellipses, z centre
    and half-height uniform over the slab) scaled so the ellipsoids reach
    0.012: sections differ per slice.  The microstructure is projected by
    project(), the ellipsoid sections in closed form with the same bin
    profile.  Ring correction on, quadratic data term."""
    rng = np.random.default_rng(seed)
    ch = two_phase(cahn_hilliard(rng, N, ch_steps), 0.002, 0.008) * disk(N)
    ell2 = draw_ellipses(rng, 0.95 * N / 2)
    zc = rng.uniform(-nz / 2, nz / 2, len(ell2))
    zh = rng.uniform(0.25, 1.0, len(ell2)) * nz
    zs = np.arange(nz) + 0.5 - nz / 2
    sections = []
    for z in zs:
        k2 = 1 - ((z - zc) / zh) ** 2
        e = ell2[k2 > 0].copy()
        e[:, 2:4] *= np.sqrt(k2[k2 > 0])[:, None]
        sections.append(e)
    rasters = np.stack([rasterize_ellipses(e, N, N / 2) for e in sections])
    k = _phantom_scale(rasters.max(), 0.012)
    truth = (ch[None] + rasters * k).astype(np.float32)
    angles = (np.arange(n_views) * (np.pi / n_views)).astype(np.float32)
    li = np.repeat(project(ch[None], angles, N, N / 2), nz, axis=1)
    for s, e in enumerate(sections):
        if len(e):
            e = e.copy()
            e[:, 5] *= k
            li[:, s, :] += project_ellipses(e, angles, N, N / 2)
    proj, wts = noisy_scan(rng, li, I0)
    rings = rng.normal(0, ring_sigma, (nz, N)).astype(np.float32)
    proj = proj + rings[None]
    times = np.arange(n_views, dtype=np.float32)
    rt = np.array([0, n_views], np.float32)
    return _scan(proj, wts, angles, times, rt, nz, N, 1, truth[None], sig_s=0.01,
                 remove_rings=1, quad_convex=1, rings=rings, **kw)


def deploy4d(N=512, nz=32, n_frames=16, n_views=2048, K=8, I0=1e4, seed=20260005, **kw) -> dict:
    """Deployment-size 4D interlaced scan: the interlaced4d model (evolving
    Cahn-Hilliard microstructure, zingers, ring offsets, generalized-Huber
    data term, ring correction) at N x N x nz, n_frames frames, n_views
    interlaced views in K sub-frames, Poisson I0."""
    return interlaced4d(N=N, nz=nz, n_frames=n_frames, n_views=n_views, K=K, I0=I0, seed=seed,
                        **kw)


# the workloads by name
WORKLOADS = dict(plain3d=plain3d, interlaced4d=interlaced4d, batched=batched,
                 precision3d=precision3d, deploy4d=deploy4d)
