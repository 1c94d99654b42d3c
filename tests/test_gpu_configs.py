"""The optimisation loop at the benchmark workloads' own sizes, on the device,
against per-loop traces the CPU oracle wrote (tests/golden/make_config_traces.py).

TOP3D_XL_GLOBAL (cpp/src/driver.cpp:23-218) with the FP64 plain path: every
loop's PCG and filter iteration counts, bisection steps, compliance C and
Cdisp, volume fraction, sharpness, change, fval and the checksum of the design
after the loop must equal the oracle's bit for bit.

  30x60x30    54 000 voxels, 172 980 free DOFs: 30 OC loops, PCG to 1e-8 with
              at most 800 iterations, Emin 1e-9, rMin 1.5, V0 0.12
  96x192x96   1.77M voxels, 5.42M free DOFs, the same parameters, 5 loops
"""
import json
import os
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _f32bits(v):
    return "%08x" % struct.unpack("<I", struct.pack("<f", v))[0]


@pytest.mark.parametrize("trace", ["trace_30x60x30.json", "trace_96x192x96.json"])
def test_loop_trace_bitexact(t3d, trace):
    path = os.path.join(GOLD, trace)
    if not os.path.exists(path):
        pytest.skip("trace not written: " + trace)
    with open(path) as fp:
        T = json.load(fp)
    recs, x, Csolid, solidIters = t3d.top3d_run(*T["dims"], **T["params"])
    assert float(Csolid).hex() == T["Csolid"]
    assert solidIters == T["solidIters"]
    assert len(recs) == T["loops"]
    for g, o in zip(recs, T["records"]):
        got = dict(loop=g["loop"], pcgIters=g["pcgIters"], filterIters=g["filterIters"],
                   bisectSteps=g["bisectSteps"], C=float(g["C"]).hex(), Cdisp=float(g["Cdisp"]).hex(),
                   xCheck=str(g["xCheck"]))
        for k in ("volFrac", "sharpness", "change", "fval"):
            got[k] = _f32bits(g[k])
        assert got == o, (trace, o["loop"])
    # the final design itself against the last record's checksum
    import oracle as O
    assert str(O.lib().o_x_check(np.ascontiguousarray(x, np.float32), x.size)) == T["records"][-1]["xCheck"]
