"""On-disk formats of the reference's runs: packed PLY point clouds and npy
state checkpoints, plus exact record checkpoints.

- write_point_cloud_packed / load_particles_only follow
  mpm_exps/engine/mesh_io.py:68-81 and :94-102: binary little-endian PLY,
  header "ply / format binary_little_endian 1.0 / comment Created by taichi
  / element vertex N / property float x, y, z / end_header", then N float32
  triples.  (The reference reads them with plyfile; here a reader for
  exactly this vertex-only float layout, which also accepts any further
  float properties per vertex, e.g. the vel_x.. of write_point_cloud.)
- MPMSolver.dump_states / checkpoint_states / load_states
  (mpm_exps/engine/mpm_solver.py:1768-1788, 1868-1895) write
  data_{frame:06d}.ply (positions) and v.npy [N, d], C.npy [N, d, d],
  J.npy [N] or F.npy [N, d, d] (float32), and read them back, re-encoding
  into the solver's records as the reference's push_* kernels store them.
- load_mesh: triangle meshes for the voxelizer (mesh_io.py:14-33).
- save_records / load_records: the solver's raw records (word planes),
  particle ids, step counter and a layout fingerprint in one npz, restored
  bit for bit (no counterpart in the reference).
"""
import os

import numpy as np

_PLY_TYPES = {"float": "<f4", "float32": "<f4", "double": "<f8", "float64": "<f8", "int": "<i4", "int32": "<i4",
              "uint": "<u4", "uint32": "<u4", "uchar": "u1", "uint8": "u1", "char": "i1", "int8": "i1",
              "short": "<i2", "int16": "<i2", "ushort": "<u2", "uint16": "<u2"}


def write_point_cloud_packed(fn, pos, num_particles=None):
    """Binary PLY of float32 positions [N, 3] (mesh_io.py:68-81)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("positions must be [N, 3]")
    n = pos.shape[0] if num_particles is None else int(num_particles)
    if n != pos.shape[0]:
        raise ValueError("num_particles does not match the positions")
    header = ("ply\nformat binary_little_endian 1.0\ncomment Created by taichi\n"
              f"element vertex {n}\nproperty float x\nproperty float y\nproperty float z\nend_header\n")
    with open(fn, "wb") as f:
        f.write(header.encode())
        f.write(pos.tobytes())


def read_ply_vertices(fn):
    """The vertex element of a binary little-endian PLY file as a structured
    numpy array (scalar properties only)."""
    with open(fn, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError(f"{fn}: not a PLY file")
        props, n, fmt, in_vertex = [], None, None, False
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{fn}: no end_header")
            tok = line.decode("ascii").split()
            if not tok or tok[0] in ("comment", "obj_info"):
                continue
            if tok[0] == "format":
                fmt = tok[1]
            elif tok[0] == "element":
                in_vertex = tok[1] == "vertex"
                if in_vertex:
                    n = int(tok[2])
                elif n is None:
                    raise ValueError(f"{fn}: elements before the vertices are not supported")
            elif tok[0] == "property" and in_vertex:
                if tok[1] == "list":
                    raise ValueError(f"{fn}: list properties are not supported")
                props.append((tok[2], _PLY_TYPES[tok[1]]))
            elif tok[0] == "end_header":
                break
        if fmt != "binary_little_endian" or n is None:
            raise ValueError(f"{fn}: only binary_little_endian vertex data is read")
        data = np.frombuffer(f.read(np.dtype(props).itemsize * n), dtype=np.dtype(props), count=n)
    return data


def load_mesh(fn, scale=1, offset=(0, 0, 0)):
    """Triangles [T, 9] (float32; vertex a, b, c) of a binary little-endian
    PLY mesh, each vertex coordinate scaled and offset per axis
    (mesh_io.py:14-33).  Reads a vertex element of scalar properties and a
    face element whose only property is a list of vertex indices (uchar
    count, then 3 integer indices per face: triangles only)."""
    if isinstance(scale, (int, float)):
        scale = (scale, scale, scale)
    with open(fn, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError(f"{fn}: not a PLY file")
        elems, fmt = [], None
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{fn}: no end_header")
            tok = line.decode("ascii").split()
            if not tok or tok[0] in ("comment", "obj_info"):
                continue
            if tok[0] == "format":
                fmt = tok[1]
            elif tok[0] == "element":
                elems.append([tok[1], int(tok[2]), []])
            elif tok[0] == "property":
                elems[-1][2].append(tok[1:])
            elif tok[0] == "end_header":
                break
        if fmt != "binary_little_endian":
            raise ValueError(f"{fn}: only binary_little_endian meshes are read")
        verts = faces = None
        for name, count, props in elems:
            if name == "vertex":
                dt = np.dtype([(pr[1], _PLY_TYPES[pr[0]]) for pr in props])
                verts = np.frombuffer(f.read(dt.itemsize * count), dtype=dt, count=count)
            elif name == "face":
                if len(props) != 1 or props[0][0] != "list":
                    raise ValueError(f"{fn}: faces must hold one vertex-index list")
                cnt_t, idx_t = _PLY_TYPES[props[0][1]], _PLY_TYPES[props[0][2]]
                dt = np.dtype([("n", cnt_t), ("i", idx_t, (3,))])
                faces = np.frombuffer(f.read(dt.itemsize * count), dtype=dt, count=count)
                if np.any(faces["n"] != 3):
                    raise ValueError(f"{fn}: only triangles are read")
            else:
                raise ValueError(f"{fn}: element {name} is not read")
    if verts is None or faces is None:
        raise ValueError(f"{fn}: vertex and face elements needed")
    xyz = np.stack([verts["x"], verts["y"], verts["z"]], axis=1).astype(np.float32)
    xyz = xyz * np.asarray(scale, dtype=np.float32) + np.asarray(offset, dtype=np.float32)
    return xyz[faces["i"].astype(np.int64)].reshape(-1, 9)


def load_particles_only(fn):
    """Positions [N, 3] of a PLY point cloud (mesh_io.py:94-102)."""
    d = read_ply_vertices(fn)
    return np.vstack([d["x"], d["y"], d["z"]]).T


def dump_states(solver, folder, frame):
    """data_{frame:06d}.ply with the solver's decoded positions (3D)."""
    if solver.dim != 3:
        raise ValueError("the packed point cloud holds 3D positions")
    os.makedirs(folder, exist_ok=True)
    st = solver.get_state()
    fn = os.path.join(folder, f"data_{frame:06d}.ply")
    write_point_cloud_packed(fn, st[:3].T, st.shape[1])
    return fn


def checkpoint_states(solver, folder):
    """v.npy, C.npy and J.npy (fluid) or F.npy (elastic): decoded float32."""
    os.makedirs(folder, exist_ok=True)
    d = solver.dim
    st = solver.get_state().astype(np.float32)
    n = st.shape[1]
    np.save(os.path.join(folder, "v.npy"), np.ascontiguousarray(st[d:2 * d].T))
    np.save(os.path.join(folder, "C.npy"), np.ascontiguousarray(st[2 * d:2 * d + d * d].T.reshape(n, d, d)))
    if solver.fluid:
        np.save(os.path.join(folder, "J.npy"), np.ascontiguousarray(st[-1]))
    else:
        np.save(os.path.join(folder, "F.npy"), np.ascontiguousarray(st[-d * d:].T.reshape(n, d, d)))


def load_states(solver, folder, frame, dither_step=0xFFFFFFFF):
    """Positions from data_{frame:06d}.ply and v, C, J or F from the npy
    files, stored into the solver (encoded into its records when quantized)."""
    x = load_particles_only(os.path.join(folder, f"data_{frame:06d}.ply"))
    d = solver.dim
    n = x.shape[0]
    v = np.load(os.path.join(folder, "v.npy"), allow_pickle=False)
    C = np.load(os.path.join(folder, "C.npy"), allow_pickle=False)
    st = np.zeros((solver.ncomp, n))
    st[:d] = x[:, :d].T
    st[d:2 * d] = v.T
    st[2 * d:2 * d + d * d] = C.reshape(n, d * d).T
    if solver.fluid:
        st[-1] = np.load(os.path.join(folder, "J.npy"), allow_pickle=False)
    else:
        st[-d * d:] = np.load(os.path.join(folder, "F.npy"), allow_pickle=False).reshape(n, d * d).T
    solver.set_state(st, dither_step=dither_step)


def _fingerprint(solver):
    comps = [[c[k] for k in ("kind", "num_bits", "is_signed", "bit_offset", "exp_bits", "exp_offset", "shared",
                             "use_dithering", "scale", "offset")] for c in solver.comps]
    return np.array([solver.dim, int(solver.fluid), solver.n_grid, solver.storage, solver.nwords, solver.ncomp]
                    + [float(x) for row in comps for x in row], dtype=np.float64)


def save_records(solver, fn):
    """The solver's stored state exactly: raw records (or value planes),
    particle ids, step counter and a layout fingerprint (npz)."""
    n = solver.current_idx
    np.savez(fn, state=solver.state[:, :n].cpu().numpy(), ids=solver.ids[:n].cpu().numpy(),
             step=np.int64(solver.step_count), layout=_fingerprint(solver))


def load_records(solver, fn):
    """Restores save_records' state bit for bit into a solver of the same
    scheme (same layout fingerprint)."""
    import torch
    with np.load(fn, allow_pickle=False) as z:
        if not np.array_equal(z["layout"], _fingerprint(solver)):
            raise ValueError("the checkpoint's record layout differs from the solver's")
        state, ids = z["state"], z["ids"]
        n = state.shape[1]
        if n > solver.capacity:
            raise ValueError("more particles than capacity")
        solver.state[:, :n] = torch.as_tensor(state, device=solver.device)
        solver.ids[:n] = torch.as_tensor(ids, device=solver.device)
        solver.current_idx = n
        solver.step_count = int(z["step"])
    torch.cuda.current_stream().synchronize()
