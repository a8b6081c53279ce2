"""Writes the known-answer fixtures under tests/golden/ from the reference.

Run in a container that has the reference tree; it compiles nothing of the
reference itself -- ``make -C oracle ref`` builds ``oracle/_ref/ref_harness``
from the reference's own MPI-free sources -- and records:

  rng.npz        random2() stream after srandom2(761521) (randlib.c:119-174)
  hr_m{1,2,4}.npz DetectorResponseProfile tables (XT_Profile.c:129-176)
  acol_m{1,2,4}.npz calcAMatrixColumnforAngle columns (XT_AMatrix.c:50-110)
  vsc.npz        VSC_based_Voxel_Line_Select lists + next random2()
                 (XT_NHICD.c:164-216)
  demo_views.npz the demo view schedules shipped with the reference
                 (demo/recon_{3d,4d}/*/proj_angles.bin, proj_times.bin,
                 recon_times.bin), copied as data

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference"

from oracle import pyorc  # noqa: E402

# geometry of the fixtures: 16 x 16 detector columns, vox_wid 1, centred
N_R = 16
N_P = 24


def angles():
    # irregular angles over more than a half turn, including 0, pi/4, pi/2
    a = np.linspace(0.0, 1.3 * np.pi, N_P, dtype=np.float64)
    a[3] = np.pi / 4
    a[7] = np.pi / 2
    return a.astype(np.float32)


def geom(m):
    delta_r = 1.0
    delta_xy = m * delta_r
    OffsetR = (delta_xy / np.sqrt(2.0) + delta_r / 2.0) / 64
    R0 = -N_R / 2 * delta_r
    return dict(delta_xy=float(delta_xy), delta_r=float(delta_r), OffsetR=float(OffsetR),
                R0=float(R0), RMax=float((N_R - N_R / 2) * delta_r), x0=float(R0),
                y0=float(-N_R / 2 * delta_r), N=N_R // m)


def run(*args):
    subprocess.run([pyorc.REF_HARNESS, *map(str, args)], check=True)


def main():
    pyorc.build()
    if not pyorc.ref_available():
        raise SystemExit("reference harness unavailable (needs /root/reference)")
    tmp = tempfile.mkdtemp()
    # RNG
    run("rng", 761521, 4096, f"{tmp}/r.bin")
    np.savez(os.path.join(HERE, "rng.npz"), seed=761521,
             values=np.fromfile(f"{tmp}/r.bin", np.float64))
    ang = angles()
    ang.tofile(f"{tmp}/a.bin")
    for m in (1, 2, 4):
        g = geom(m)
        run("hr", f"{tmp}/a.bin", N_P, repr(g["delta_xy"]), repr(g["delta_r"]),
            repr(g["OffsetR"]), repr(g["delta_r"]), f"{tmp}/h.bin")
        hr = np.fromfile(f"{tmp}/h.bin", np.float32).reshape(N_P, 65)
        np.savez(os.path.join(HERE, f"hr_m{m}.npz"), angles=ang, H_r=hr,
                 **{k: np.float64(v) for k, v in g.items()})
        hr.tofile(f"{tmp}/h.bin")
        run("acol", f"{tmp}/h.bin", f"{tmp}/a.bin", N_P, g["N"], g["N"], repr(g["x0"]),
            repr(g["y0"]), repr(g["delta_xy"]), repr(g["R0"]), repr(g["RMax"]),
            repr(g["delta_r"]), N_R, repr(g["OffsetR"]), f"{tmp}/c.bin")
        rec = np.fromfile(f"{tmp}/c.bin", dtype=np.dtype([("n", "<i4"), ("v", "<f8", 8),
                                                         ("i", "<i4", 8)]))
        np.savez_compressed(os.path.join(HERE, f"acol_m{m}.npz"), count=rec["n"], values=rec["v"],
                 index=rec["i"])
    # VSC on a seeded magnitude map with ties
    NT, NB, NY, NX, USEL, SEED = 2, 4, 16, 16, 200, 761521
    rng = np.random.default_rng(3)
    mag = rng.random((NT, NB, NY, NX)).astype(np.float32)
    mag[0, 1, 3:6, :] = 0
    mag[1, 2, :, 4:6] = 0.5
    mag.tofile(f"{tmp}/m.bin")
    run("vsc", f"{tmp}/m.bin", NT, NB, NY, NX, USEL, SEED, f"{tmp}/v.bin")
    raw = np.fromfile(f"{tmp}/v.bin", np.uint8)
    off = 0
    counts, xs, ys = [], [], []
    for _ in range(NT * NB):
        n = int(np.frombuffer(raw[off:off + 4], np.int32)[0])
        off += 4
        xs.append(np.frombuffer(raw[off:off + 4 * n], np.int32).copy())
        off += 4 * n
        ys.append(np.frombuffer(raw[off:off + 4 * n], np.int32).copy())
        off += 4 * n
        counts.append(n)
    nxt = np.frombuffer(raw[off:off + 8], np.float64)[0]
    np.savez(os.path.join(HERE, "vsc.npz"), mag=mag, usel=USEL, seed=SEED,
             counts=np.array(counts, np.int32), xs=np.concatenate(xs), ys=np.concatenate(ys),
             next_random=nxt)
    # demo view schedules (data files of the reference)
    demo = {}
    for tag, d in (("d3", "demo/recon_3d/shepp-logan-3D"), ("d4", "demo/recon_4d/cahn-hilliard-4D")):
        for f in ("proj_angles", "proj_times", "recon_times"):
            demo[f"{tag}_{f}"] = np.fromfile(os.path.join(REF, d, f + ".bin"), np.float32)
    np.savez(os.path.join(HERE, "demo_views.npz"), **demo)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
