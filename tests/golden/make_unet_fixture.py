"""Known answers of the learned initialiser's network from the CPU oracle
(oracle/unet_ref.py, bf16 rounding points as on the GPU), for the GPU
gradient check: the network's output and the gradients of the squared-error
loss with respect to every parameter, for seeded parameters and inputs.

The gradients have as many entries as the parameters (1.4 M), so the fixture
keeps, per tensor, its norm and its inner products with 8 seeded Gaussian
probe vectors (a wrong entry anywhere moves them), plus the full gradients of
the head and the first layer's biases.

    python tests/golden/make_unet_fixture.py   # writes tests/golden/unet_grads.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import unet_ref as U  # noqa: E402

SHAPE = (2, 8, 16, 20)
SCALE = 50.0
N_PROBES = 8


def case():
    """Seeded parameters (He weights, small random biases and head) and
    inputs: two frames of a smooth random volume, their mean, a target."""
    rng = np.random.default_rng(20261201)
    p = U.init_params(11)
    params = {k: (w, (rng.standard_normal(b.shape) * 0.05).astype(np.float32))
              for k, (w, b) in p.items()}
    params["head"] = ((rng.standard_normal(32) * 0.3).astype(np.float32),
                      np.array([0.02], np.float32))
    base = rng.random((SHAPE[1], SHAPE[2], SHAPE[3]))
    fbp = np.stack([base + 0.3 * rng.random(SHAPE[1:]) for _ in range(SHAPE[0])])
    fbp = (fbp * 0.015).astype(np.float32)
    mean = fbp.mean(axis=0).astype(np.float32)
    target = (fbp + 0.002 * rng.standard_normal(SHAPE)).astype(np.float32)
    return params, fbp, mean, target


def probes(shape, name):
    seed = sum(ord(c) * 31 ** i for i, c in enumerate(name)) % (2 ** 31)
    return np.random.default_rng(seed).standard_normal((N_PROBES,) + tuple(shape))


def oracle_grads(params, fbp, mean, target):
    out, acts = U.forward(params, fbp, mean, SCALE)
    out = out.astype(np.float32)
    n = out.size
    gout = (2.0 * (out - target) / n).astype(np.float32)
    loss = float(np.mean((out.astype(np.float64) - target) ** 2))
    return out, loss, U.backward(params, acts, gout, SCALE)


def summarise(grads):
    s = {}
    for name, (gw, gb) in grads.items():
        for kind, g in (("w", gw), ("b", gb)):
            key = f"{name}.{kind}"
            g = np.asarray(g, np.float64)
            s[key + ".norm"] = np.linalg.norm(g)
            s[key + ".proj"] = probes(g.shape, key).reshape(N_PROBES, -1) @ g.reshape(-1)
    return s


def main():
    params, fbp, mean, target = case()
    out, loss, grads = oracle_grads(params, fbp, mean, target)
    arrays = dict(out_sum=np.float64(out.astype(np.float64).sum()),
                  out_l2=np.float64(np.linalg.norm(out.astype(np.float64))), loss=np.float64(loss),
                  head_w=grads["head"][0].astype(np.float64), head_b=grads["head"][1],
                  e1a_b=grads["e1a"][1].astype(np.float64))
    arrays.update(summarise(grads))
    np.savez_compressed(os.path.join(HERE, "unet_grads.npz"), **arrays)
    print(f"loss {loss:.6e}, {len(arrays)} arrays")


if __name__ == "__main__":
    main()
