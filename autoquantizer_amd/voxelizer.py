"""Mesh emission of the 3D elastic scenes: the reference's voxelizer
(mpm_exps/engine/voxelizer.py:17-157) and particle seeding
(MPMSolver.add_mesh / seed_from_voxels / cnt_emit_num,
mpm_exps/engine/mpm_solver.py:1356-1441), on the host with numpy.

Voxelization counts, for every column (p, q) of the super-sampled voxel
grid, the mesh surfaces above it: each triangle whose xy projection
covers the column centre adds +1 (normal z > 0) or -1 (normal z < 0) to
the voxels from the padding up to the triangle's height there.  A voxel
with a positive count is inside.  Seeding puts sample_density + 1
particles uniformly at random in each inside voxel, at (index + U[0,1)^3)
times the solver's dx (not the voxel size: the reference seeds in solver
cells, so a mesh comes out super_sample times larger), plus the
translation.
"""
import numpy as np

_JITTER = np.array([-0.057616723909439505, -0.25608986292614977, 0.06716309129743714])


def _cross2d(ax, ay, bx, by):
    return ax * by - ay * bx


def _inside_ccw(px, py, a, b, c):
    """voxelizer.py:10-13, vectorised over the points (px, py)."""
    return ((_cross2d(a[0] - px, a[1] - py, b[0] - px, b[1] - py) >= 0)
            & (_cross2d(b[0] - px, b[1] - py, c[0] - px, c[1] - py) >= 0)
            & (_cross2d(c[0] - px, c[1] - py, a[0] - px, a[1] - py) >= 0))


class Voxelizer:
    """Voxelizer(res, dx, super_sample=2, padding=3) (voxelizer.py:17-38):
    each resolution rounded up to a power of two, times super_sample; voxel
    size dx / super_sample; f64 arithmetic (the reference's default)."""

    def __init__(self, res, dx, super_sample=2, padding=3):
        r = []
        for v in res:
            k = 1
            while k < v:
                k *= 2
            r.append(k * super_sample)
        self.res = tuple(r)
        self.dx = dx / super_sample
        self.inv_dx = 1.0 / self.dx
        self.padding = padding
        self.voxels = None

    def voxelize(self, triangles):
        """voxelize_triangles (voxelizer.py:46-118): the counts, int8
        [res0, res1, res2] (also kept in self.voxels)."""
        tri = np.asarray(triangles, dtype=np.float64).reshape(-1, 9)
        r0, r1, r2 = self.res
        pad = self.padding
        diff = np.zeros((r0, r1, r2 + 1), dtype=np.int16)
        jit = _JITTER * 1e-8
        for t in tri:
            a, b, c = t[0:3] + jit, t[3:6] + jit, t[6:9] + jit
            bmin = np.minimum(np.minimum(a, b), c)
            bmax = np.maximum(np.maximum(a, b), c)
            p0 = max(pad, int(np.floor(bmin[0] * self.inv_dx)))
            p1 = min(r0 - pad, int(np.floor(bmax[0] * self.inv_dx)) + 1)
            q0 = max(pad, int(np.floor(bmin[1] * self.inv_dx)))
            q1 = min(r1 - pad, int(np.floor(bmax[1] * self.inv_dx)) + 1)
            nrm = np.cross(b - a, c - a)
            with np.errstate(invalid="ignore", divide="ignore"):
                nrm = nrm / np.sqrt(np.dot(nrm, nrm))  # NaN for a degenerate triangle, as normalized()
            if abs(nrm[2]) < 1e-10 or p1 <= p0 or q1 <= q0:
                continue
            P, Q = np.meshgrid(np.arange(p0, p1), np.arange(q0, q1), indexing="ij")
            px, py = (P + 0.5) * self.dx, (Q + 0.5) * self.dx
            hit = _inside_ccw(px, py, a, b, c) | _inside_ccw(px, py, a, c, b)
            if not hit.any():
                continue
            px, py, P, Q = px[hit], py[hit], P[hit], Q[hit]
            # height = int(-n . ((x, y, 0) - a) / n_z * inv_dx + 0.5), truncated toward zero
            dot = nrm[0] * (px - a[0]) + nrm[1] * (py - a[1]) + nrm[2] * (0.0 - a[2])
            with np.errstate(invalid="ignore"):
                hf = np.trunc(-dot / nrm[2] * self.inv_dx + 0.5)
            h = np.where(np.isfinite(hf), hf, -1).astype(np.int64)  # a NaN height fills nothing
            h = np.minimum(h, r1 - pad)  # clamped by res[1], as the reference does
            h = np.minimum(h, r2)
            inc = 1 if nrm[2] > 0 else -1
            ok = h > pad
            np.add.at(diff, (P[ok], Q[ok], np.full(ok.sum(), pad)), inc)
            np.add.at(diff, (P[ok], Q[ok], h[ok]), -inc)
        self.voxels = np.cumsum(diff[:, :, :r2], axis=2, dtype=np.int16).astype(np.int8)
        return self.voxels


def count_emit(voxels, sample_density):
    """cnt_emit_num (mpm_solver.py:1403-1412): particles one seeding adds."""
    return int((voxels > 0).sum()) * (sample_density + 1)


def seed_from_voxels(voxels, sample_density, dx, translation, velocity, rng):
    """seed_from_voxels (mpm_solver.py:1414-1441): positions [N, 3] and
    velocities [N, 3] of sample_density + 1 particles per inside voxel,
    x = (U[0,1)^3 + (i, j, k)) dx + translation (seeded numpy draws in
    voxel order; the reference's order is its parallel loop's)."""
    idx = np.argwhere(voxels > 0).astype(np.float64)
    k = sample_density + 1
    base = np.repeat(idx, k, axis=0)
    x = (rng.random(base.shape) + base) * dx + np.asarray(translation, dtype=np.float64)
    v = np.broadcast_to(np.asarray(velocity, dtype=np.float64), x.shape).copy()
    return x, v
