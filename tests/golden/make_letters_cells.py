"""Writes letters_cells.npz: the solver cells (1/128) that hold the
particles of the reference's precomputed letters emission
(mpm_exps/inputs/letters_emit_one_frame.npy, the particles
emit_letters_one_frame places for the AD run of run_mpm_elastic_3d.py,
translation (1.5, 0.8, 0.1), sample density 1 on the n_grid = 128
voxelizer) and the particle count per cell.

    python tests/golden/make_letters_cells.py [reference root]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(root="/root/reference"):
    x = np.load(os.path.join(root, "mpm_exps", "inputs", "letters_emit_one_frame.npy"), allow_pickle=False)
    cells = np.floor((x.astype(np.float64) - np.array([1.5, 0.8, 0.1])) * 128).astype(np.int16)
    u, c = np.unique(cells, axis=0, return_counts=True)
    np.savez_compressed(os.path.join(HERE, "letters_cells.npz"), cells=u, counts=c.astype(np.int16),
                        particles=np.int64(len(x)))
    print(f"{len(x)} particles in {len(u)} cells")


if __name__ == "__main__":
    main(*sys.argv[1:])
