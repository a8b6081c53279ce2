"""Independent numpy statements of the MBIR MAP problem, used to pin the oracle.

These follow the mathematics, not the oracle's loops: the system matrix is
assembled densely from A columns, the data term and the qGGMRF prior are
evaluated over all unordered neighbour pairs with numpy slicing, and the MAP
point is found with scipy's L-BFGS on the explicit cost.  A voxel update that
had a wrong sign, dropped a neighbour or mis-weighted a term would make ICD
converge somewhere other than this minimiser.
"""
from __future__ import annotations

import numpy as np

P, Q = 1.2, 2.0


def rho(d, sq, sqp, c):
    a = np.abs(d)
    return (a ** Q / sq) / (c + a ** (Q - P) / sqp)


def drho(d, sq, sqp, c):
    a = np.abs(d)
    t1 = a ** (Q - P) / sqp
    t3 = c + t1
    return np.sign(d) * (a ** (Q - 1) / (t3 * sq)) * (Q - (Q - P) * t1 / t3)


def dense_A(prob) -> np.ndarray:
    """Rows: sinogram entries [N_p][N_r][N_t]; columns: voxels
    [N_time][N_z][N_y][N_x].  Voxel (f,z,y,x) touches the views of frame f,
    the A-column bins r and the detector rows z*m_z .. z*m_z+m_z-1."""
    L = prob.L
    i = prob.info
    Np, Nr, Nt = i["N_p"], i["N_r"], i["N_t"]
    Nf, Nz, Ny, Nx, mz = i["N_time"], i["N_z"], i["N_y"], i["N_x"], i["mult_z"]
    off, idx = prob.views()
    hr = prob.hr()
    A = np.zeros((Np * Nr * Nt, Nf * Nz * Ny * Nx))
    vals = np.zeros(64)
    ind = np.zeros(64, np.int32)
    # recompute cos/sin the same way the reference does (XT_Init.c:193-204)
    ang = prob._keep[2].astype(np.float64)
    cosv, sinv = np.cos(ang), np.sin(ang)
    for f in range(Nf):
        for v in idx[off[f]:off[f + 1]]:
            for y in range(Ny):
                for x in range(Nx):
                    n = L.orc_amatrix_column(np.ascontiguousarray(hr[v]), cosv[v], sinv[v], y, x,
                                             i["x0"], i["y0"], i["delta_xy"], i["R0"], i["RMax"],
                                             i["delta_r"], Nr, i["OffsetR"], vals, ind)
                    for z in range(Nz):
                        col = ((f * Nz + z) * Ny + y) * Nx + x
                        for m in range(n):
                            for o in range(mz):
                                row = (v * Nr + ind[m]) * Nt + z * mz + o
                                A[row, col] += vals[m]
    return A


def neighbour_pairs(shape):
    """All unordered 26-neighbour pairs within a frame plus temporal pairs:
    yields (slice_a, slice_b, kind, (dz,dy,dx))."""
    Nf, Nz, Ny, Nx = shape
    out = []
    for dz in (0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dz == 0 and (dy < 0 or (dy == 0 and dx <= 0)):
                    continue
                za = slice(0, Nz - dz)
                zb = slice(dz, Nz)
                ya = slice(max(0, -dy), Ny - max(0, dy))
                yb = slice(max(0, dy), Ny - max(0, -dy))
                xa = slice(max(0, -dx), Nx - max(0, dx))
                xb = slice(max(0, dx), Nx - max(0, -dx))
                out.append(((slice(None), za, ya, xa), (slice(None), zb, yb, xb), "s", (dz, dy, dx)))
    out.append(((slice(0, Nf - 1),), (slice(1, Nf),), "t", None))
    return out


class MapProblem:
    """MAP cost of one stage with the generalized-Huber mask b frozen
    (b is all true when the zinger model is off)."""

    def __init__(self, prob, A=None):
        i = prob.info
        self.i = i
        self.shape = (i["N_time"], i["N_z"], i["N_y"], i["N_x"])
        self.A = dense_A(prob) if A is None else A
        self.y = prob.proj().astype(np.float64).reshape(-1)
        self.w = prob.wgt().astype(np.float64).reshape(-1)
        self.d = np.repeat(prob.off().astype(np.float64)[None], i["N_p"], axis=0).reshape(-1)
        self.sfilt = prob.sfilt
        self.mask = np.broadcast_to(prob.mask().astype(bool)[None, None], self.shape)
        self.pairs = neighbour_pairs(self.shape)
        self.var = prob.var_est

    def cost_grad(self, xfull):
        i = self.i
        x = xfull.reshape(self.shape)
        e = self.y - self.A @ xfull - self.d
        u = e * np.sqrt(self.w)
        data = 0.5 * np.sum(u * u)
        g = -(self.A.T @ (self.w * e))
        prior = 0.0
        gx = np.zeros(self.shape)
        for a, b, kind, o in self.pairs:
            diff = x[a] - x[b]
            if kind == "s":
                dz, dy, dx = o
                gw = self.sfilt[dz + 1, dy + 1, dx + 1]
                sq, sqp, c = i["sigS_q"], i["sigS_qp"], i["c_s"]
            else:
                gw = i["tfilt"]
                sq, sqp, c = i["sigT_q"], i["sigT_qp"], i["c_t"]
            prior += gw * np.sum(rho(diff, sq, sqp, c))
            dd = gw * drho(diff, sq, sqp, c)
            gx[a] += dd
            gx[b] -= dd
        total = data + prior + i["N_p"] * i["N_r"] * i["N_t"] * np.log(self.var) / 2
        return total, g + gx.reshape(-1)

    def minimise(self, x0=None, iters=20000):
        from scipy.optimize import minimize
        m = self.mask.reshape(-1)
        full = np.zeros(m.size) if x0 is None else x0.astype(np.float64).reshape(-1).copy()

        def f(z):
            full[m] = z
            c, g = self.cost_grad(full)
            return c, g[m]

        res = minimize(f, full[m], jac=True, method="L-BFGS-B",
                       options=dict(maxiter=iters, maxcor=50, ftol=1e-16, gtol=1e-12))
        full[m] = res.x
        return res.fun, full.reshape(self.shape)


# --------------------------------------------------------------------------
# filtered back-projection (the initial object of tb_init_fbp)
# --------------------------------------------------------------------------

def ramp_matrix(n_bins: int, h: float) -> np.ndarray:
    """[n][n] matrix of the Ram-Lak filter sampled at pitch h (Kak and Slaney
    eq. 3.61) times h: q = R @ p is the linear (zero-padded) convolution."""
    n = np.arange(n_bins)[:, None] - np.arange(n_bins)[None, :]
    k = np.zeros(n.shape)
    k[n == 0] = 0.25
    odd = (n % 2) != 0
    k[odd] = -1.0 / (np.pi ** 2 * n[odd].astype(np.float64) ** 2)
    return k / h


def fbp(sino, angles, frame_views, n_x, n_y, x0, y0, dxy, r0, dr, mask=None) -> np.ndarray:
    """sino [views][slices][bins] -> object [frames][slices][n_y][n_x]: each
    frame back-projects its own views (frame_views[f], an index array) with
    weight pi / len(views), the ray offset r = x cos - y sin of pixel centres
    linearly interpolated between bin centres r0 + (j + 0.5) dr (zero
    outside the detector); pixels outside mask are zero."""
    sino = np.asarray(sino, np.float64)
    q = np.einsum("rk,vzk->vzr", ramp_matrix(sino.shape[2], dr), sino)
    x = x0 + (np.arange(n_x) + 0.5) * dxy
    y = y0 + (np.arange(n_y) + 0.5) * dxy
    X, Y = np.meshgrid(x, y)
    nb = sino.shape[2]
    qp = np.concatenate([np.zeros(q.shape[:2] + (1,)), q, np.zeros(q.shape[:2] + (1,))], axis=2)
    out = np.zeros((len(frame_views), sino.shape[1], n_y, n_x))
    for f, views in enumerate(frame_views):
        for v in views:
            th = float(angles[v])
            u = (X * np.cos(th) - Y * np.sin(th) - r0) / dr - 0.5
            j = np.floor(u)
            t = u - j
            j0 = np.clip(j.astype(np.int64) + 1, 0, nb + 1)
            j1 = np.clip(j.astype(np.int64) + 2, 0, nb + 1)
            out[f] += (1 - t)[None] * qp[v][:, j0] + t[None] * qp[v][:, j1]
        if len(views):
            out[f] *= np.pi / len(views)
    if mask is not None:
        out *= mask[None, None]
    return out
