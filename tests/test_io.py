"""Trajectory reader (read_frame, src/trajectory.h:107-150) against frames the
reference's own save_frame wrote (tests/golden/frame.npz); runs on the CPU
(the reader is host code of liborbc_hip.so)."""
import os

import numpy as np
import pytest

import openrbc_amd as orbc

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden(tmp_path_factory):
    g = np.load(os.path.join(HERE, "golden", "frame.npz"))
    path = tmp_path_factory.mktemp("frames") / "ref.orbc"
    path.write_bytes(g["frame"].tobytes())
    return g, str(path)


def test_read_reference_frames(golden):
    g, path = golden
    nl, np_ = len(g["lx"]), len(g["px"])
    f0 = orbc.read_frame(path, 0)
    assert f0["step"] == 0 and f0["natom"] == nl + np_
    assert f0["fields"] == 31
    assert np.array_equal(f0["tag"], np.concatenate([np.arange(int(g["lbase"]), int(g["lbase"]) + nl), g["ptag"]]))
    assert np.array_equal(f0["type"], np.concatenate([np.zeros(nl, np.int32), g["ptype"]]))
    assert np.array_equal(f0["x"], np.concatenate([g["lx"], g["px"]]))
    assert np.array_equal(f0["v"], np.concatenate([g["lv"], g["pv"]]))
    assert np.array_equal(f0["n"], np.concatenate([g["ln"], g["pn"]]))
    assert np.array_equal(f0["aff"], np.concatenate([g["laff"], g["paff"]]))
    assert np.array_equal(f0["f"], np.zeros((nl + np_, 3), np.float32))
    f1 = orbc.read_frame(path, 1)                  # appended frame: step 5, position + rotation + voronoi
    assert f1["step"] == 5 and f1["fields"] == 7 and "v" not in f1
    assert np.array_equal(f1["x"], f0["x"])


def test_read_frame_errors(golden, tmp_path):
    _, path = golden
    with pytest.raises(orbc.OrbcError):
        orbc.read_frame(path, 2)                   # only two frames
    bad = tmp_path / "bad.orbc"
    bad.write_bytes(open(path, "rb").read()[:100])  # truncated
    with pytest.raises(orbc.OrbcError):
        orbc.read_frame(str(bad), 0)
