"""The mesh voxelizer and particle seeding (autoquantizer_amd.voxelizer,
restating mpm_exps/engine/voxelizer.py:17-157 and mpm_solver.py:1356-1441),
pinned by closed forms: a closed axis-aligned box marks exactly the voxels
whose column centres lie over the box and whose index lies between the
rounded heights of its bottom and top faces; an open surface above a
column marks the voxels below it; seeding puts sample_density + 1
particles in every inside voxel, in solver cells."""
import numpy as np

from autoquantizer_amd.voxelizer import Voxelizer, count_emit, seed_from_voxels


def box_triangles(lo, hi):
    """12 outward-oriented triangles of the box [lo, hi]."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]])
    faces = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
             (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return np.array([np.concatenate([v[a], v[b], v[c]]) for a, b, c in faces])


def test_box_voxels_closed_form():
    vox = Voxelizer((32, 32, 32), 1.0 / 32)
    lo, hi = (0.21, 0.33, 0.27), (0.58, 0.49, 0.71)
    got = vox.voxelize(box_triangles(lo, hi)) > 0
    dx = vox.dx
    assert vox.res == (64, 64, 64)
    c = (np.arange(64) + 0.5) * dx
    inx = (c > lo[0]) & (c < hi[0])
    iny = (c > lo[1]) & (c < hi[1])
    k = np.arange(64)
    inz = (k >= int(lo[2] / dx + 0.5)) & (k < int(hi[2] / dx + 0.5))
    want = inx[:, None, None] & iny[None, :, None] & inz[None, None, :]
    np.testing.assert_array_equal(got, want)


def test_open_surface_counts_below():
    """One upward triangle: every covered column gets +1 from the padding
    up to the triangle's height (a lone surface: the voxels below it)."""
    vox = Voxelizer((16, 16, 16), 1.0 / 16)
    tri = np.array([[0.2, 0.2, 0.5, 0.8, 0.2, 0.5, 0.2, 0.8, 0.5]])
    v = vox.voxelize(tri)
    h = int(0.5 / vox.dx + 0.5)
    cols = v.sum(axis=2)
    assert set(np.unique(v)) <= {0, 1}
    assert np.all(cols[cols > 0] == h - vox.padding)
    p = (np.arange(32) + 0.5) * vox.dx
    P, Q = np.meshgrid(p, p, indexing="ij")
    inside = (P > 0.2) & (Q > 0.2) & (P + Q < 1.0)
    inside[:3] = inside[:, :3] = False
    np.testing.assert_array_equal(cols > 0, inside)


def test_seeding_in_solver_cells():
    vox = np.zeros((8, 8, 8), np.int8)
    vox[2, 3, 4] = 1
    vox[5, 5, 5] = 2
    vox[1, 1, 1] = -1
    assert count_emit(vox, 3) == 8
    x, v = seed_from_voxels(vox, 3, 0.1, (1.0, 0.0, 0.5), (0.0, -4.0, 0.0), np.random.default_rng(0))
    assert x.shape == (8, 3) and np.all(v == (0.0, -4.0, 0.0))
    cell = np.floor((x - (1.0, 0.0, 0.5)) / 0.1 + 1e-9).astype(int)
    assert sorted(map(tuple, cell.tolist())) == [(2, 3, 4)] * 4 + [(5, 5, 5)] * 4


def test_letters_match_the_reference_emission():
    """The reference's own letters emission (letters_emit_one_frame.npy:
    63,856 particles, two in each of 31,928 solver cells) is exactly what
    voxelizing its letters mesh (auto_quant.ply, scale (0.1, 0.1, 2),
    offset (0.02, 0.02, 0.5)) on the n_grid = 128 voxelizer gives: the same
    inside voxels, sample density 1."""
    import os
    from autoquantizer_amd.scenes import letters_triangles
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "letters_cells.npz"), allow_pickle=False)
    vox = Voxelizer((128,) * 3, 1.0 / 128).voxelize(letters_triangles())
    ours = np.argwhere(vox > 0)
    assert count_emit(vox, 1) == int(g["particles"])
    assert np.all(g["counts"] == 2)
    ref = g["cells"].astype(np.int64)
    key = lambda a: (a[:, 0] * 512 + a[:, 1]) * 512 + a[:, 2]
    np.testing.assert_array_equal(np.sort(key(ours)), np.sort(key(ref)))
