"""The run summary (t3d_summary_format, driver.cpp:225-258) against the text the
reference's own CLI prints after a run (oracle/_ref/top3d_xl_cli, built from the
reference's sources where they lie).  The CLI's values are read back from its
text and formatted again: every line must come out identical except the four
derived time figures, which depend on sums the CLI does not print and are
checked for their format.  Host code only: no GPU is touched."""
import os
import re
import subprocess

import pytest

import oracle as O

CLI = os.path.join(O.HERE, "_ref", "top3d_xl_cli")


def test_format_layout(t3d):
    txt = t3d.summary_format(origX=16, origY=8, origZ=8, resX=16, resY=8, resZ=8, V0=0.12, iterations=4,
                             cgMaxIt=800, Csolid=1.2345678901, objFirst=3.25, objLast=2.5, sumCG=0.5, sumOpt=0.1,
                             sumFilter=0.05, sumIter=0.7, threads=8)
    lines = txt.split("\n")
    assert [l for l in lines if l.isupper()] == ["PARAMETERS", "SOLVING", "RESULTS", "TIME", "THREADS", "NOTES"]
    assert "compliance solid: 1.234568e+00" in lines and "v0: 0.12" in lines
    assert "time per iter: 0.1750" in lines and "percentage time spent on cg: 71.43%" in lines


@pytest.mark.skipif(not os.path.exists(CLI), reason="reference CLI not built")
def test_matches_reference_cli(t3d, tmp_path):
    env = dict(os.environ, OMP_NUM_THREADS="2")
    out = subprocess.run([CLI, "GLOBAL", "8", "16", "8", "0.12", "3", "1"], cwd=tmp_path, env=env,
                         capture_output=True, text=True, timeout=600).stdout
    ref = out[out.index("PARAMETERS"):]
    ref = ref[:ref.index("NOTES") + len("NOTES\n")]
    num = r"([-+0-9.eE]+)"
    g = lambda pat: re.search(pat, ref).groups()
    ox, oy, oz = map(int, g(r"original dims: x=(\d+), y=(\d+), z=(\d+)"))
    rx, ry, rz = map(int, g(r"padded dims: x=(\d+), y=(\d+), z=(\d+)"))
    vals = dict(origX=ox, origY=oy, origZ=oz, resX=rx, resY=ry, resZ=rz, V0=float(g(r"v0: " + num)[0]),
                iterations=int(g(r"iterations: (\d+)")[0]), cgMaxIt=int(g(r"cg iterations: (\d+)")[0]),
                Csolid=float(g(r"compliance solid: " + num)[0]), objFirst=float(g(r"objective first: " + num)[0]),
                objLast=float(g(r"objective last: " + num)[0]), sumCG=float(g(r"total solver time: " + num)[0]),
                sumOpt=0.0, sumFilter=0.0, sumIter=1.0, threads=int(g(r"THREADS\n(\d+)")[0]))
    mine = t3d.summary_format(**vals).split("\n")
    theirs = ref.split("\n")
    assert len(mine) == len(theirs)
    timed = ("time per iter:", "percentage time spent on")
    for a, b in zip(mine, theirs):
        if b.startswith(timed):
            assert re.fullmatch(r"[a-z ]+: \d+\.\d{2}%" if "percentage" in b else r"time per iter: \d+\.\d{4}", b), b
        else:
            assert a == b, (a, b)
    assert (ox, oy, oz) == (16, 8, 8) and vals["iterations"] == 3 and vals["threads"] == 2
