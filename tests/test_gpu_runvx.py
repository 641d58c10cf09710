"""runvx's object I/O as the reference's runvx defines it (utilities/runvx/vxImage.cpp:560-640,
1185-1265): COMPARE with rect{}, err{}, checksum and checksum-save-instead-of-test, per-frame file
patterns, and READ / COMPARE of files holding several frames -- on a Gaussian3x3 graph whose expected
output comes from the oracle."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from mivisionx_amd import synth

pytestmark = pytest.mark.gpu

W, H = 200, 120


@pytest.fixture(scope="module")
def runvx(gpu):
    from mivisionx_amd import openvx as vx
    return vx.RUNVX


def _gdf(tmp_path, out_io, in_file="in.u8"):
    text = (f"data input = image:{W},{H},U008:READ,{tmp_path}/{in_file}\n"
            f"data output = image:{W},{H},U008:{out_io}\n"
            "node org.khronos.openvx.gaussian_3x3 input output\n")
    path = tmp_path / "g.gdf"
    path.write_text(text)
    return str(path)


def _run(runvx, *args):
    r = subprocess.run([runvx, *args], capture_output=True, timeout=120)
    return r.returncode, r.stdout.decode() + r.stderr.decode()


@pytest.fixture(scope="module")
def frames(orc):
    ins = [synth.scene_u8(W, H, 60 + k) for k in range(2)]
    return ins, [orc.filter3x3(0, a) for a in ins]


def test_exact_err_and_rect(runvx, frames, tmp_path):
    ins, outs = frames
    (tmp_path / "in.u8").write_bytes(ins[0].tobytes())
    (tmp_path / "ref.u8").write_bytes(outs[0].tobytes())
    rc, log = _run(runvx, _gdf(tmp_path, f"COMPARE,{tmp_path}/ref.u8"))
    assert rc == 0, log
    # perturb 3 pixels by +-1 inside the rectangle x 100..150, y 50..80
    bad = outs[0].copy()
    for (y, x, d) in ((60, 110, 1), (70, 140, -1), (55, 101, 1)):
        bad[y, x] = np.uint8(int(bad[y, x]) + d)
    (tmp_path / "bad.u8").write_bytes(bad.tobytes())
    rc, log = _run(runvx, _gdf(tmp_path, f"COMPARE,{tmp_path}/bad.u8"))
    assert rc == 2 and "3 pixel(s) differ" in log, log
    rc, log = _run(runvx, _gdf(tmp_path, f"COMPARE,{tmp_path}/bad.u8,err{{-1;1}}"))
    assert rc == 0, log
    rc, log = _run(runvx, _gdf(tmp_path, f"COMPARE,{tmp_path}/bad.u8,rect{{0;0;100;120}}"))
    assert rc == 0, log
    rc, log = _run(runvx, _gdf(tmp_path, f"COMPARE,{tmp_path}/bad.u8,rect{{100;50;150;80}}"))
    assert rc == 2 and "3 pixel(s) differ" in log, log


def test_checksums(runvx, frames, tmp_path):
    ins, outs = frames
    (tmp_path / "in.u8").write_bytes(ins[0].tobytes())
    sums = tmp_path / "sums.txt"
    rc, log = _run(runvx, "-frames:2", _gdf(tmp_path, f"COMPARE,{sums},rect{{10;20;110;70}},checksum-save-instead-of-test"))
    assert rc == 0, log
    want = hashlib.md5(np.ascontiguousarray(outs[0][20:70, 10:110]).tobytes()).hexdigest()
    assert sums.read_text().split() == [want, want]
    rc, log = _run(runvx, "-frames:2", _gdf(tmp_path, f"COMPARE,{sums},rect{{10;20;110;70}},checksum"))
    assert rc == 0, log
    sums.write_text(want + "\n" + "0" * 32 + "\n")
    rc, log = _run(runvx, "-frames:2", _gdf(tmp_path, f"COMPARE,{sums},rect{{10;20;110;70}},checksum"))
    assert rc == 2 and "checksum" in log, log


def test_multi_frame_files_and_patterns(runvx, frames, tmp_path):
    ins, outs = frames
    (tmp_path / "in2.u8").write_bytes(ins[0].tobytes() + ins[1].tobytes())
    (tmp_path / "ref2.u8").write_bytes(outs[0].tobytes() + outs[1].tobytes())
    rc, log = _run(runvx, "-frames:2", _gdf(tmp_path, f"COMPARE,{tmp_path}/ref2.u8", in_file="in2.u8"))
    assert rc == 0 and "2 frame(s)" in log, log
    # the input ends after two frames: processing stops there
    rc, log = _run(runvx, "-frames:5", _gdf(tmp_path, f"COMPARE,{tmp_path}/ref2.u8", in_file="in2.u8"))
    assert rc == 0 and "2 frame(s)" in log, log
    # per-frame reference files named by a printf pattern of (frame, width, height)
    for k in range(2):
        (tmp_path / f"ref_{k}_{W}x{H}.u8").write_bytes(outs[k].tobytes())
    rc, log = _run(runvx, "-frames:2", _gdf(tmp_path, f"COMPARE,{tmp_path}/ref_%d_%dx%d.u8", in_file="in2.u8"))
    assert rc == 0, log
