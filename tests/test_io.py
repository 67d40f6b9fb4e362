"""On-disk formats (autoquantizer_amd.io): the reference's packed PLY point
cloud (mpm_exps/engine/mesh_io.py:68-81) and its reader
(mesh_io.py:94-102), and the save_records npz layout check.  CPU only."""
import numpy as np
import pytest


def test_packed_ply_header_and_round_trip(tmp_path):
    from autoquantizer_amd.io import load_particles_only, write_point_cloud_packed
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, 1, (1000, 3)).astype(np.float32)
    fn = tmp_path / "data_000007.ply"
    write_point_cloud_packed(str(fn), pos, 1000)
    raw = fn.read_bytes()
    header = (b"ply\nformat binary_little_endian 1.0\ncomment Created by taichi\nelement vertex 1000\n"
              b"property float x\nproperty float y\nproperty float z\nend_header\n")
    assert raw.startswith(header)
    assert len(raw) == len(header) + 1000 * 12
    assert np.frombuffer(raw[len(header):], dtype="<f4").reshape(1000, 3).tobytes() == pos.tobytes()
    back = load_particles_only(str(fn))
    assert back.dtype == np.float32 and np.array_equal(back, pos)


def test_ply_reader_extra_properties_and_faces(tmp_path):
    """A point cloud with velocities (the reference's write_point_cloud,
    mesh_io.py:40-66) and a mesh with a face list after its vertices (as
    Blender writes the reference's inputs/auto_quant.ply): positions only."""
    from autoquantizer_amd.io import load_particles_only
    rng = np.random.default_rng(1)
    data = rng.standard_normal((50, 6)).astype(np.float32)
    header = ("ply\nformat binary_little_endian 1.0\ncomment Created by taichi\nelement vertex 50\n"
              "property float x\nproperty float y\nproperty float z\nproperty float vel_x \n"
              "property float vel_y \nproperty float vel_z\nend_header\n")
    fn = tmp_path / "pv.ply"
    fn.write_bytes(header.encode() + data.tobytes())
    assert np.array_equal(load_particles_only(str(fn)), data[:, :3])
    header2 = ("ply\nformat binary_little_endian 1.0\ncomment Created by Blender\nelement vertex 50\n"
               "property float x\nproperty float y\nproperty float z\nproperty float nx\nproperty float ny\n"
               "property float nz\nelement face 2\nproperty list uchar uint vertex_indices\nend_header\n")
    faces = np.array([3, 0, 0, 0, 0, 1, 0, 0, 0, 2, 0, 0, 0], dtype=np.uint8).tobytes() * 2
    fn2 = tmp_path / "mesh.ply"
    fn2.write_bytes(header2.encode() + data.tobytes() + faces)
    assert np.array_equal(load_particles_only(str(fn2)), data[:, :3])


def test_ply_rejects_other_formats(tmp_path):
    from autoquantizer_amd.io import load_particles_only, write_point_cloud_packed
    fn = tmp_path / "a.ply"
    fn.write_bytes(b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nend_header\n0\n")
    with pytest.raises(ValueError):
        load_particles_only(str(fn))
    with pytest.raises(ValueError):
        write_point_cloud_packed(str(tmp_path / "b.ply"), np.zeros((4, 2), np.float32))
