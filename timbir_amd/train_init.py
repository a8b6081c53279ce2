"""Training of the learned initialiser (timbir_amd.learned.UNet) on MBIR
reconstructions this pipeline makes from seeded phantoms.

A training scan is a 4D interlaced acquisition at 256 x 256 x 8 slices x 16
frames with 2048 views in 8 sub-frames (128 views per frame, as the
deployment scan), Poisson I0 = 1e4, zingers and ring offsets, the
generalized-Huber data term and ring correction, the deployment's prior.
Its phantom is a Cahn-Hilliard two-phase microstructure evolving over the
frames (phase values drawn per scan) or a set of random ellipses drifting
over the frames.  The line integrals come from the library's forward
projector (tb_forward_project: the reference's system matrix,
XT_AMatrix.c:50-110), the noise, zingers and rings from numpy.

For each scan the pipeline runs reconstruct()'s stages (XT_Engine.c:111-182)
with the finest stage continued for a fixed number of ICD iterations past
the reference's stop rule; that object is the target.  The network's input
is the filtered back-projection of the same scan (tb_init_fbp) and its mean
over frames.  One scan gives 16 samples (its frames).

    python -m timbir_amd.train_init --scans 64 --steps 3000 \\
        --out timbir_amd/weights/unet_init.npz

With --workload deploy4d the training scans are the deployment scan's model
at its own size (datagen.deploy4d: 512 x 512 x 32 slices x 16 frames, 2048
views, I0 = 1e4) with other seeds; the scan the bench and the tests use
(seed 20260005) is never one of them.

    python -m timbir_amd.train_init --workload deploy4d --scans 6 --val 1 \\
        --steps 3000 --out timbir_amd/weights/unet_init_deploy4d.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import Stage, datagen, set_options
from .batch import multires_schedule
from .devmem import DevArray
from .learned import UNet

N, NZ, FRAMES, VIEWS, K, I0 = 256, 8, 16, 2048, 8, 1e4
PRIOR = dict(sig_s=0.02, sig_t=5e-4, c_s=1e-2, c_t=1e-2, remove_rings=1, quad_convex=0,
             huber_delta=0.5, huber_T=4.0)


def log(msg):
    print(msg, flush=True)


def phantom(rng: np.random.Generator) -> np.ndarray:
    """[FRAMES][NZ][N][N] attenuation per pixel, 0 outside the disk."""
    m = datagen.disk(N)
    frames = []
    if rng.random() < 0.75:
        lo, hi = rng.uniform(0.003, 0.007), rng.uniform(0.012, 0.018)
        c = datagen.cahn_hilliard(rng, N, int(rng.integers(32, 97)))
        steps = int(rng.integers(4, 13))
        for _ in range(FRAMES):
            frames.append(np.repeat((datagen.two_phase(c, lo, hi) * m)[None], NZ, axis=0))
            c = datagen.cahn_hilliard(rng, N, steps, c0=c)
    else:
        ell = datagen.draw_ellipses(rng, 0.45 * N)
        ell[:, 5] *= 0.02 / max(ell[:, 5].sum(), 1e-12) * rng.uniform(1.0, 2.5)
        vel = rng.normal(0, 0.6, (len(ell), 2))
        for f in range(FRAMES):
            e = ell.copy()
            e[:, 0:2] += f * vel
            img = datagen.rasterize_ellipses(e, N, N / 2, sub=2)
            frames.append(np.repeat((np.minimum(img, 0.03) * m)[None], NZ, axis=0))
    return np.stack(frames).astype(np.float32)


def synth_scan(seed: int) -> dict:
    """A seeded training scan (see the module docstring)."""
    rng = np.random.default_rng(seed)
    truth = phantom(rng)
    angles, times = datagen.interlaced_views(K, VIEWS, VIEWS)
    rt = np.linspace(0, float(VIEWS), FRAMES + 1).astype(np.float32)
    shell = datagen._scan(np.zeros((VIEWS, NZ, N), np.float32),
                          np.ones((VIEWS, NZ, N), np.float32), angles, times, rt, NZ, N, FRAMES,
                          truth, **PRIOR)
    st = Stage(shell)
    st.set_object(truth)
    li = np.transpose(st.forward_project(), (0, 2, 1)).astype(np.float64)   # [views][rows][bins]
    st.close()
    proj, wts = datagen.noisy_scan(rng, li, I0)
    rings = rng.normal(0, 0.01, (NZ, N)).astype(np.float32)
    proj = proj + rings[None]
    zi = rng.choice(proj.size, int(round(1e-3 * proj.size)), replace=False)
    flat = proj.reshape(-1)
    flat[zi] *= 5.0
    return datagen._scan(proj, wts, angles, times, rt, NZ, N, FRAMES, truth, **PRIOR)


def fbp_and_target(scan: dict, finest_iters: int, coarse_rel: float = 1e-5,
                   coarse_cap: int = 200) -> tuple:
    """The filtered back-projection of the scan at the finest resolution and
    its MBIR object (both [FRAMES][NZ][N][N]): reconstruct()'s stages, the
    coarse ones run on until an iteration changes their MAP cost by less
    than coarse_rel (at most coarse_cap iterations; they are cheap and carry
    the slow, temporally smooth part of the solution), then finest_iters
    ICD iterations of the finest stage."""
    sched = multires_schedule(scan["proj_rows"], scan["proj_cols"])
    f = Stage(scan, mult_xy=1, mult_z=1, stage_idx=len(sched) - 1,
              prev_mult_z=sched[-2][1] if len(sched) > 1 else 0)
    f.init_fbp()
    fbp = f.object()
    f.close()
    prev, var = None, 1.0
    for s, (mxy, mz) in enumerate(sched):
        st = Stage(scan, mult_xy=mxy, mult_z=mz, stage_idx=s,
                   prev_mult_z=sched[s - 1][1] if s else 0, var_est=var)
        if prev is not None:
            st.init_from_stage(prev)
            st.set_offset(prev.offset())
            prev.close()
        if s < len(sched) - 1:
            st.init_error_sino()
            last = st.cost()[0]
            for it in range(1, coarse_cap + 1):
                st.icd_iteration(it)
                c = st.cost()[0]
                if it > 2 and abs(last - c) < coarse_rel * abs(c):
                    break
                last = c
            prev, var = st, st.var_est
        else:
            st.init_error_sino()
            for it in range(1, finest_iters + 1):
                st.icd_iteration(it)
            x = st.object()
            st.close()
    return fbp, x


def crops(data, rng, batch: int, crop: int):
    """A batch of random crops (every slice, crop x crop pixels) of random
    frames: (fbp [B][NZ][c][c], mean, target).  data holds (fbp, its mean
    over frames, target) per scan."""
    fb, mn, tg = [], [], []
    for _ in range(batch):
        fbp, mean, target = data[int(rng.integers(len(data)))]
        frames, n = fbp.shape[0], fbp.shape[-1]
        f = int(rng.integers(frames))
        y0, x0 = rng.integers(0, n - crop + 1, 2)
        sl = (slice(None), slice(y0, y0 + crop), slice(x0, x0 + crop))
        fb.append(fbp[f][sl])
        mn.append(mean[sl])
        tg.append(target[f][sl])
    return np.stack(fb), np.stack(mn), np.stack(tg)


def train(net: UNet, data, steps: int, batch: int = 4, crop: int = 128, lr: float = 1e-3,
          seed: int = 0, log_every: int = 100, budget_s: float = 1e30, checkpoint=None) -> list:
    """Adam on the squared error between the network's output and the MBIR
    target over batches of random crops, cosine learning-rate decay; stops
    early after budget_s seconds; checkpoint(step) every 500 steps."""
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.time()
    for step in range(1, steps + 1):
        if time.time() - t0 > budget_s:
            log(f"  training stopped by its time budget after {step - 1} steps")
            break
        if checkpoint and step % 500 == 0:
            checkpoint(step)
        fb, mn, tg = crops(data, rng, batch, crop)
        cur = lr * 0.5 * (1 + np.cos(np.pi * (step - 1) / steps))
        loss = net.train_step(DevArray.from_numpy(fb), DevArray.from_numpy(mn),
                              DevArray.from_numpy(tg), cur)
        hist.append(loss)
        if step % log_every == 0 or step == 1:
            log(f"  step {step}: loss {np.mean(hist[-log_every:]):.4e} lr {cur:.2e} "
                f"({time.time() - t0:.0f} s)")
    return hist


def evaluate(net: UNet, val) -> dict:
    """Mean squared error to the target of the back-projection and of the
    network's output over the validation scans' full frames."""
    e_fbp, e_net = [], []
    for fbp, mean, target in val:
        f = DevArray.from_numpy(np.ascontiguousarray(fbp))
        m = DevArray.from_numpy(np.ascontiguousarray(mean))
        out = net.forward(f, m).numpy()
        e_fbp.append(float(np.mean((fbp - target) ** 2)))
        e_net.append(float(np.mean((out - target) ** 2)))
    return dict(mse_fbp=float(np.mean(e_fbp)), mse_net=float(np.mean(e_net)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="interlaced256", choices=["interlaced256", "deploy4d"])
    ap.add_argument("--scans", type=int, default=64)
    ap.add_argument("--val", type=int, default=2)
    ap.add_argument("--finest-iters", type=int, default=20)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20261000)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "weights",
                                                  "unet_init.npz"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--data-budget", type=float, default=1e30,
                    help="seconds of scan generation and reconstruction at most")
    ap.add_argument("--train-budget", type=float, default=1e30,
                    help="seconds of training at most")
    a = ap.parse_args(argv)
    set_options(verbose=0)
    t0 = time.time()
    data = []
    for k in range(a.scans + a.val):
        if time.time() - t0 > a.data_budget and len(data) > a.val + 1:
            log(f"data budget reached after {len(data)} scans")
            break
        ts = time.time()
        if a.workload == "deploy4d":
            scan = datagen.deploy4d(seed=a.seed + k)
        else:
            scan = synth_scan(a.seed + k)
        tg = time.time() - ts
        fbp, target = fbp_and_target(scan, a.finest_iters)
        del scan
        data.append((fbp, fbp.mean(axis=0).astype(np.float32), target))
        log(f"scan {k}: generated in {tg:.1f} s, reconstructed in {time.time() - ts - tg:.1f} s "
            f"(mse fbp {np.mean((fbp - target) ** 2):.3e})")
    val, train_set = data[:a.val], data[a.val:]
    t1 = time.time()
    net = UNet(seed=a.seed)
    before = evaluate(net, val)
    os.makedirs(os.path.dirname(a.out) or ".", exist_ok=True)
    hist = train(net, train_set, a.steps, a.batch, a.crop, a.lr, seed=a.seed,
                 budget_s=a.train_budget,
                 checkpoint=lambda st: net.save(a.out, scans=len(train_set), steps=st))
    after = evaluate(net, val)
    t2 = time.time()
    log(f"validation mse: back-projection {before['mse_fbp']:.4e}, untrained {before['mse_net']:.4e}, "
        f"trained {after['mse_net']:.4e}")
    net.save(a.out, scans=len(train_set), finest_iters=a.finest_iters, steps=a.steps, seed=a.seed,
             val_mse_fbp=after["mse_fbp"], val_mse_net=after["mse_net"])
    rep = dict(scans=len(train_set), val=a.val, steps_run=len(hist), finest_iters=a.finest_iters, steps=a.steps,
               batch=a.batch, crop=a.crop, lr=a.lr, seed=a.seed, data_s=t1 - t0,
               train_s=t2 - t1, before=before, after=after, loss=hist[::max(1, len(hist) // 100)])
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(rep, fh, indent=1)
    log(json.dumps({k: v for k, v in rep.items() if k != "loss"}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
