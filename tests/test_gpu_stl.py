"""Voxel-surface STL built on the device (t3d_stl_size / t3d_stl_image /
t3d_export_surface_stl, the file of io.cpp:80-87 export_surface_stl) against
the oracle's image (oracle/top3d_oracle.c o_stl_image, pinned byte for byte
against the reference's own extract_faces / write_stl_binary in
test_stl_oracle.py): bit-exact bytes on cuboids, padded grids, voxel domains
with voids, thresholds at and below zero (cells without an element then count
as inside), NaN densities, empty and multi-chunk block scans; the optimisation
loop's own STL and summary output."""
import numpy as np
import pytest

import oracle as O
from test_voxels import spheres_mask

pytestmark = pytest.mark.gpu


def _case(t3d, dims, solid=None, seed=0, nan=False):
    p = t3d.Problem(*dims, solid=solid)
    m = O.Model(*dims, solid=solid)
    x = np.random.default_rng(seed).uniform(0, 1, m.ne).astype(np.float32)
    if nan:
        x[::17] = np.nan
    return p, m, x


def _want(m, x, iso):
    return O.stl_image(m.resY, m.resX, m.resZ, O.stl_volume(m, x), iso)


@pytest.mark.parametrize("dims,iso,seed,nan", [((8, 16, 8), 0.5, 1, False), ((12, 20, 8), 0.3, 2, False),
                                               ((6, 10, 5), 0.0, 3, False), ((8, 16, 8), -0.2, 4, False),
                                               ((12, 20, 8), 0.6, 5, True), ((8, 16, 8), 2.0, 6, False),
                                               ((64, 128, 64), 0.7, 7, False)])
def test_image_matches_oracle(t3d, dims, iso, seed, nan):
    p, m, x = _case(t3d, dims, seed=seed, nan=nan)
    print("stl case", dims, iso, seed, nan, flush=True)
    want = _want(m, x, iso)
    dx = t3d.DeviceArray.from_host(x)
    nbytes, ntri = t3d.stl_size(p, dx, iso)
    assert nbytes == len(want) and ntri == (len(want) - 84) // 50
    got = t3d.stl_image(p, dx, iso)
    assert got == want
    if iso > 1.5:
        assert ntri == 0 and nbytes == 84


def test_voxel_domain_with_voids(t3d):
    dims = (16, 32, 16)
    solid = spheres_mask(*dims, [(10.0, 8.0, 8.0, 4.5), (22.0, 4.0, 10.0, 3.0)])
    for iso in (0.4, 0.0):
        p, m, x = _case(t3d, dims, solid=solid, seed=8)
        dx = t3d.DeviceArray.from_host(x)
        assert t3d.stl_image(p, dx, iso) == _want(m, x, iso), iso


def test_export_file_and_loop_output(t3d, tmp_path):
    p, m, x = _case(t3d, (12, 20, 8), seed=9)
    path = tmp_path / "d.stl"
    t3d.export_surface_stl(p, t3d.DeviceArray.from_host(x), str(path), 0.3)
    assert path.read_bytes() == _want(m, x, 0.3)
    # the loop's own export of its final design (driver.cpp:272-275, iso 0.3)
    out = tmp_path / "loop.stl"
    md = tmp_path / "loop.md"
    recs, xl, Cs, _ = t3d.top3d_run(8, 16, 8, V0=0.12, nLoop=2, rMin=1.5, stlPath=str(out).encode(),
                                    summaryPath=str(md).encode())
    m2 = O.Model(8, 16, 8)
    assert len(recs) == 2 and out.read_bytes() == _want(m2, xl, 0.3)
    # and its summary (driver.cpp:225-270): float sums of the per-loop times in loop order
    sums = [np.float32(0)] * 4
    for r in recs:
        sums = [a + np.float32(r[k]) for a, k in zip(sums, ("tSolve", "tOpti", "tFilter", "tTotal"))]
    want = t3d.summary_format(origX=16, origY=8, origZ=8, resX=m2.resX, resY=m2.resY, resZ=m2.resZ, V0=0.12,
                              iterations=2, cgMaxIt=800, Csolid=Cs, objFirst=recs[0]["Cdisp"],
                              objLast=recs[-1]["Cdisp"], sumCG=float(sums[0]), sumOpt=float(sums[1]),
                              sumFilter=float(sums[2]), sumIter=float(sums[3]), threads=1)
    assert md.read_text() == want
