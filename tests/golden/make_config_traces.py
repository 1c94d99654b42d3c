"""Writes the per-loop parity traces of the optimisation loop at the benchmark
workloads' own sizes, from the repo's CPU oracle (oracle/liboracle.so; run
`make -C oracle` first).  Test data only: tests/test_gpu_configs.py replays the
same loops on the device and requires every record bit for bit.

  trace_30x60x30.json    the CPU-runnable demo: solid 30x60x30 cuboid
                         (54 000 voxels, 172 980 free DOFs), Emin 1e-9,
                         rMin 1.5, V0 0.12, 30 OC loops, PCG to 1e-8 with at
                         most 800 iterations, the reference's multigrid rules
  trace_96x192x96.json   the FP64 drop-in workload: 96x192x96 (1.77M voxels,
                         5.42M free DOFs), same material, load and parameters,
                         nLoop 50, the workload's own count; the loop stops
                         after 20 on its sharpness test (every solve at its
                         800-iteration cap, compliance NaN from loop 16).  51
                         minutes on 7 host threads with the oracle built with
                         `make -C oracle OPT="-O3 -march=native"`, which gives
                         the same bits as the default build (its first 5
                         records equal those of the default build's 5-loop run)

Each record holds the loop's PCG and filter iteration counts, bisection steps,
compliance C and Cdisp (FP64, as float.hex), volume fraction, sharpness,
change and fval (FP32 bit patterns) and o_x_check of the design after the loop.

usage: [O_LOOP_PROGRESS=1] python tests/golden/make_config_traces.py [cfg0|cfg1 ...]
"""
import json
import os
import struct
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle as O  # noqa: E402

WORKLOADS = {
    "cfg0": dict(dims=(30, 60, 30), file="trace_30x60x30.json",
                 params=dict(nLoop=30, cgTol=1e-8, cgMaxIt=800, Emin=1e-9, rMin=1.5, V0=0.12)),
    "cfg1": dict(dims=(96, 192, 96), file="trace_96x192x96.json",
                 params=dict(nLoop=50, cgTol=1e-8, cgMaxIt=800, Emin=1e-9, rMin=1.5, V0=0.12)),
}

F32 = ("volFrac", "sharpness", "change", "fval")


def f32bits(v):
    return "%08x" % struct.unpack("<I", struct.pack("<f", v))[0]


def make(name):
    w = WORKLOADS[name]
    t0 = time.time()
    recs, x, Csolid, solidIters = O.run(*w["dims"], **w["params"])
    out = dict(dims=list(w["dims"]), params=w["params"], Csolid=float(Csolid).hex(), solidIters=solidIters,
               loops=len(recs), seconds=round(time.time() - t0, 1), records=[])
    for r in recs:
        d = dict(loop=r["loop"], pcgIters=r["pcgIters"], filterIters=r["filterIters"],
                 bisectSteps=r["bisectSteps"], C=float(r["C"]).hex(), Cdisp=float(r["Cdisp"]).hex(),
                 xCheck=str(r["xCheck"]))
        for k in F32:
            d[k] = f32bits(r[k])
        out["records"].append(d)
    with open(os.path.join(HERE, w["file"]), "w") as fp:
        json.dump(out, fp, indent=1)
    print(name, "written in %.1f s" % (time.time() - t0), flush=True)


if __name__ == "__main__":
    for n in (sys.argv[1:] or ["cfg0", "cfg1"]):
        make(n)
