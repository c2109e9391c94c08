"""The library's native readers and writer (ibem3d_read_inputs / ibem3d_write_stress) against the
Python ones and the reference's files; the native driver's argument handling. No GPU calls."""
import os
import subprocess

import numpy as np
import pytest

from ibemproject_amd import engine
from ibemproject_amd.formats import read_problem, read_stress, write_problem
from ibemproject_amd.generate import make_rve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def same(a, b):
    assert a.title == b.title
    for k in ("E0", "nu0", "E1", "nu1"):
        assert getattr(a, k) == getattr(b, k), k
    for k in ("nodes", "conn", "bc", "tbc", "centres", "radii", "points"):
        x, y = np.asarray(getattr(a, k)), np.asarray(getattr(b, k))
        assert x.shape == y.shape and np.array_equal(x, y), k


@pytest.mark.parametrize("case", ["example", "tiny_rve", "rve5"])
def test_native_reader_matches_python(case):
    d = os.path.join(GOLD, case)
    same(engine.read_inputs_native(d), read_problem(d))


def test_native_reader_crlf_and_generated(tmp_path):
    p = make_rve(5, 3, 0.08, 0.12, 3, nump=7)
    write_problem(p, str(tmp_path))
    for f in ("BEM.txt", "position.txt", "postprocess.txt"):
        path = tmp_path / f
        path.write_bytes(path.read_bytes().replace(b"\n", b"\r\n"))
    q = engine.read_inputs_native(str(tmp_path))
    same(q, read_problem(str(tmp_path)))
    assert np.array_equal(q.nodes, p.nodes) and np.array_equal(q.centres, p.centres)


def test_native_reader_errors(tmp_path):
    with pytest.raises(engine.EngineError, match="position.txt"):
        engine.read_inputs_native(str(tmp_path))
    (tmp_path / "position.txt").write_text("1\n10\n0.25\n0.5 0.5\n")
    with pytest.raises(engine.EngineError, match="end of file"):
        engine.read_inputs_native(str(tmp_path))


def test_stress_writer_reference_format(tmp_path):
    truth = read_stress(os.path.join(GOLD, "example", "stress_ground_truth.txt"))
    out = tmp_path / "stress.txt"
    engine.write_stress_native(str(out), truth)
    want = open(os.path.join(GOLD, "example", "stress_ground_truth.txt")).read().split("\n")
    got = out.read_text().split("\n")
    # the shipped file's rows are what std::scientific wrote; re-writing its values reproduces them
    assert [w.strip() for w in want if w.strip()] == [g for g in got if g]
    engine.write_stress_native(str(out), truth, full=True)
    assert np.array_equal(read_stress(str(out)), truth)


def test_driver_usage_and_missing_input(tmp_path):
    exe = os.path.join(ROOT, "ibemproject_amd", "ibem3d_run")
    r = subprocess.run([exe, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    (tmp_path / "calculationType.txt").write_text("2\n")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 1 and "calculationType 2" in r.stderr
    (tmp_path / "calculationType.txt").write_text("1\n")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 1 and "position.txt" in r.stderr
