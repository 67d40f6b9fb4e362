"""Writes the scene inputs this package takes from the reference's data
files (read as data only: numpy loads without pickles, a binary PLY):

- propeller_higher_inside.npz: the voxels of the SDF propeller
  (mpm_exps/inputs/propeller_higher.npz and propeller_higher_normal.npz,
  the volume init_sdf_value loads for the range and AD runs,
  mpm_solver.py:300-319) that the collider treats as inside
  (value < -1e-5, sdf() at mpm_solver.py:1447-1452), with their values and
  normals; every other voxel acts as outside, so the sparse copy gives the
  same collisions as the dense volume;
- ranges_elastic_3d.npy: mpm_exps/ranges/elastic_3d/elastic_3d.npy, the
  ranges of the 3D elastic performance configuration
  (run_mpm_elastic_3d.py:366-369);
- letters_triangles.npz: the triangles of mpm_exps/inputs/auto_quant.ply
  (unscaled), the letters the 3D elastic scenes emit (:47-63, 317).

    python autoquantizer_amd/data/make_inputs.py [reference root]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main(root="/root/reference"):
    from autoquantizer_amd.io import load_mesh
    inp = os.path.join(root, "mpm_exps", "inputs")
    sdf = np.load(os.path.join(inp, "propeller_higher.npz"), allow_pickle=False)["sdf"]
    nrm = np.load(os.path.join(inp, "propeller_higher_normal.npz"), allow_pickle=False)["sdf"]
    inside = np.argwhere(sdf < -1e-5)
    np.savez_compressed(os.path.join(HERE, "propeller_higher_inside.npz"),
                        shape=np.array(sdf.shape, dtype=np.int32), idx=inside.astype(np.int16),
                        sdf=sdf[tuple(inside.T)].astype(np.float32),
                        normal=nrm[tuple(inside.T)].astype(np.float32))
    print(f"propeller: {len(inside)} inside voxels of {sdf.shape}")
    r = np.load(os.path.join(root, "mpm_exps", "ranges", "elastic_3d", "elastic_3d.npy"), allow_pickle=False)
    np.save(os.path.join(HERE, "ranges_elastic_3d.npy"), r.astype(np.float64))
    tri = load_mesh(os.path.join(inp, "auto_quant.ply"))
    np.savez_compressed(os.path.join(HERE, "letters_triangles.npz"), triangles=tri)
    print(f"letters: {tri.shape[0]} triangles")


if __name__ == "__main__":
    main(*sys.argv[1:])
