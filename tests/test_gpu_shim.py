"""The reference's own CLI (apps/top3d_cli.cpp) linked with the drop-in shim
(integration/top3d_gpu.cpp) on the GPU: its per-loop lines are the device
loop's records, its STL is the device loop's final design's surface byte for
byte, and its objective values agree with the unmodified reference CLI (strict
CPU build, one thread): loop 1 to the printed digits, later loops within the
reference's own run-to-run spread (1e-5 relative)."""
import glob
import math
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPU_CLI = os.path.join(REPO, "oracle", "_ref", "top3d_xl_cli_gpu")
REF_CLI = os.path.join(REPO, "oracle", "_ref", "top3d_xl_cli")
IT = re.compile(r" It\.:\s*(\d+) Obj\.:\s*(\S+) Vol\.:\s*(\S+) Sharp\.:\s*(\S+) Cons\.:\s*(\S+) Ch\.:\s*(\S+)")


def _run(cli, cwd, args, threads=None):
    os.makedirs(os.path.join(cwd, "out"), exist_ok=True)
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    r = subprocess.run([cli, "GLOBAL"] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return [m.groups() for m in IT.finditer(r.stdout)]


@pytest.mark.skipif(not os.path.exists(GPU_CLI), reason="shim CLI not built (make -C oracle shim)")
def test_reference_cli_on_the_shim(t3d, tmp_path):
    nely, nelx, nelz, V0, nLoop = 8, 16, 8, 0.12, 3
    lines = _run(GPU_CLI, str(tmp_path / "gpu"), [nely, nelx, nelz, V0, nLoop, 1])
    rec, x, Cs, si = t3d.top3d_run(nely, nelx, nelz, nLoop=nLoop, V0=V0)
    assert len(lines) == len(rec) == nLoop
    for g, r in zip(lines, rec):
        assert int(g[0]) == r["loop"]
        assert g[1] == "%.8e" % r["Cdisp"]
        assert g[2] == "%.4e" % r["volFrac"] and g[3] == "%.4e" % r["sharpness"]
    # the STL (out/stl/<tag>.stl, iso 0.3) against the device image of the same final design
    stl = glob.glob(str(tmp_path / "gpu" / "out" / "stl" / "*.stl"))
    md = glob.glob(str(tmp_path / "gpu" / "out" / "comments" / "*.md"))
    assert len(stl) == 1 and len(md) == 1
    p = t3d.Problem(nely, nelx, nelz)
    mine = str(tmp_path / "mine.stl")
    t3d.export_surface_stl(p, t3d.DeviceArray.from_host(x), mine, 0.3)
    assert open(stl[0], "rb").read() == open(mine, "rb").read()
    assert "objective last: %.6e" % np.float32(rec[-1]["Cdisp"]) in open(md[0]).read()
    if os.path.exists(REF_CLI):
        ref = _run(REF_CLI, str(tmp_path / "ref"), [nely, nelx, nelz, V0, nLoop, 1], threads=1)
        assert len(ref) == len(lines)
        for k, (a, b) in enumerate(zip(ref, lines)):
            # loop 1 to the printed digits; later loops inside the 1e-5 relative
            # spread the reference shows against itself across thread counts
            # (its OpenMP reductions and atomics make the loop chaotic in the last bits)
            assert math.isclose(float(a[1]), float(b[1]), rel_tol=1e-8 if k == 0 else 1e-5), (k, a, b)
            assert abs(float(a[2]) - float(b[2])) <= 1e-4 and abs(float(a[3]) - float(b[3])) <= 1e-4


GPU_CHECK = os.path.join(REPO, "oracle", "_ref", "shim_check_gpu")
REF_CHECK = os.path.join(REPO, "oracle", "_ref", "shim_check_ref")
I32 = ("eleMapBack", "eNodMat", "eDofMat", "elemColor", "freeDofIndex", "eDofFreeMat", "iters")
F32 = ("filter1", "filter2")
# bit for bit: integer model data, Ke, the FP64 operator, the layout conversions
# and the PDE filter (each bit-exact against the reference in its own tests)
EXACT = I32 + F32 + ("Fz", "Ke", "Ku", "KKu", "scatter_uy", "restrict", "filter_kernel", "filter_diag", "x_zero")
# round-off: ce by a sequential row dot on both sides but C by the canonical
# tree, the V-cycle's coarsest Cholesky and PCG's dot products in another order
CLOSE = {"ce": 1e-13, "C": 1e-13, "vcycle": 1e-12, "x_mg": 1e-9, "x_jacobi": 1e-9, "x_cg": 1e-10}


def _load(d, name):
    dt = np.int32 if name in I32 else (np.float32 if name in F32 else np.float64)
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=dt)


@pytest.mark.skipif(not (os.path.exists(GPU_CHECK) and os.path.exists(REF_CHECK)),
                    reason="shim checks not built (make -C oracle shim)")
@pytest.mark.parametrize("dims", [(8, 16, 8), (12, 20, 8)])
def test_every_entry_point_through_the_shim(tmp_path, dims):
    """integration/shim_check.cpp through the reference's headers, once on the
    shim (GPU) and once on the reference's own CPU sources: the model arrays,
    Ke, K*u (both overloads), restrict_to_free / scatter_from_free and the PDE
    filter (including its warm start) bit for bit; ComputeCompliance, the
    V-cycle closure and PCG_free with the multigrid, Jacobi and no
    preconditioner to round-off with identical iteration counts; a zero
    right-hand side gives one iteration and x = 0 (solver.cpp:147-149)."""
    outs = {}
    for name, exe in (("gpu", GPU_CHECK), ("ref", REF_CHECK)):
        d = str(tmp_path / name)
        os.makedirs(d)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([exe] + [str(v) for v in dims], cwd=d, capture_output=True, text=True, timeout=300,
                           env=env)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        outs[name] = d
    for name in EXACT:
        a, b = _load(outs["gpu"], name), _load(outs["ref"], name)
        assert a.shape == b.shape and np.array_equal(a, b), name
    for name, bound in CLOSE.items():
        a, b = _load(outs["gpu"], name), _load(outs["ref"], name)
        assert a.shape == b.shape
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        print(dims, name, "rel diff %.2e" % err)
        assert err <= bound, (name, err)
    its = _load(outs["gpu"], "iters")
    assert its[3] == 1 and not np.any(_load(outs["gpu"], "x_zero"))
