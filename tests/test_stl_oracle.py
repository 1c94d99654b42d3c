"""Pins of the oracle's voxel-surface STL (oracle/top3d_oracle.c o_stl_image,
oracle.py stl_volume) on the CPU.

* Against the reference's own code where oracle/_ref is built: extract_faces +
  write_stl_binary (voxel_surface.hpp:12-93, export_stl.hpp:14-33) on random
  volumes, thresholds and NaNs, and export_surface_stl (io.cpp:80-87) on the
  cuboid model with random densities -- byte for byte.
* From the geometry alone (always): one voxel gives the 12 triangles of a unit
  cube; a solid box a x b x c gives 2 (ab + bc + ca) faces; every edge of the
  surface of a random solid is shared by an even number of triangles (a closed
  surface); every triangle lies in a face of a unit cell; the header and the
  triangle count follow the binary STL layout.
"""
import os
import struct

import numpy as np
import pytest

import oracle as O


def _tris(img):
    assert img[:21] == b"Generated by TOP3D_XL" and img[21:80] == bytes(59)
    (n,) = struct.unpack("<I", img[80:84])
    assert len(img) == 84 + 50 * n
    rec = np.frombuffer(img[84:], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    assert np.all(rec["n"] == 0) and np.all(rec["a"] == 0)
    return rec["v"]


def test_single_voxel_cube():
    vol = np.ones(1, np.float32)
    v = _tris(O.stl_image(1, 1, 1, vol, 0.5))
    assert v.shape == (12, 3, 3)
    assert set(np.unique(v)) == {0.0, 1.0}
    # each pair of triangles covers one face of the unit cube: all 4 corners on one coordinate plane
    for f in range(6):
        q = v[2 * f: 2 * f + 2].reshape(-1, 3)
        assert any(np.all(q[:, k] == q[0, k]) for k in range(3))


@pytest.mark.parametrize("a,b,c", [(3, 4, 5), (1, 7, 2), (6, 1, 1)])
def test_solid_box_face_count(a, b, c):
    ny, nx, nz = a + 2, b + 3, c + 1
    vol = np.zeros((nz, nx, ny), np.float32)
    vol[1:1 + c, 2:2 + b, 1:1 + a] = 1.0
    v = _tris(O.stl_image(ny, nx, nz, vol.reshape(-1), 0.3))
    assert len(v) == 2 * 2 * (a * b + b * c + c * a)


def test_closed_surface_and_unit_faces():
    rng = np.random.default_rng(3)
    ny, nx, nz = 9, 7, 6
    vol = rng.uniform(0, 1, ny * nx * nz).astype(np.float32)
    v = _tris(O.stl_image(ny, nx, nz, vol, 0.45))
    edges = {}
    for t in v:
        for i in range(3):
            e = tuple(sorted((tuple(t[i]), tuple(t[(i + 1) % 3]))))
            edges[e] = edges.get(e, 0) + 1
    assert all(c % 2 == 0 for c in edges.values())
    span = v.max(axis=1) - v.min(axis=1)
    assert np.all((span == 0).sum(axis=1) == 1) and np.all(span.max(axis=1) == 1)


def test_empty_and_full():
    vol = np.full(4 * 3 * 2, 0.2, np.float32)
    assert len(_tris(O.stl_image(4, 3, 2, vol, 0.5))) == 0
    assert len(_tris(O.stl_image(4, 3, 2, vol, 0.2))) == 2 * 2 * (4 * 3 + 3 * 2 + 2 * 4)


needs_ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref (the reference build) not built here")


@needs_ref
@pytest.mark.parametrize("dims,iso,seed", [((5, 4, 3), 0.5, 1), ((16, 8, 12), 0.3, 2), ((7, 9, 5), 0.0, 3),
                                           ((6, 6, 6), -0.1, 4), ((11, 3, 8), 0.7, 5)])
def test_volume_matches_reference(tmp_path, dims, iso, seed):
    ny, nx, nz = dims
    rng = np.random.default_rng(seed)
    vol = rng.uniform(0, 1, ny * nx * nz).astype(np.float32)
    vol[rng.random(vol.size) < 0.1] = 0.0
    vol[rng.random(vol.size) < 0.02] = np.nan
    path = str(tmp_path / "ref.stl")
    assert O.ref().ref_stl_volume(ny, nx, nz, vol, iso, path.encode()) == 0
    with open(path, "rb") as f:
        want = f.read()
    assert O.stl_image(ny, nx, nz, vol, iso) == want


@needs_ref
@pytest.mark.parametrize("dims", [(8, 16, 8), (12, 20, 8), (6, 10, 5)])
def test_export_matches_reference(tmp_path, dims):
    nely, nelx, nelz = dims
    m = O.Model(*dims)
    x = np.random.default_rng(7).uniform(0, 1, m.ne).astype(np.float32)
    path = str(tmp_path / "ref.stl")
    assert O.ref().ref_export_stl(nely, nelx, nelz, x, path.encode(), 0.3) == 0
    with open(path, "rb") as f:
        want = f.read()
    got = O.stl_image(m.resY, m.resX, m.resZ, O.stl_volume(m, x), 0.3)
    assert got == want
    assert os.path.getsize(path) > 84
