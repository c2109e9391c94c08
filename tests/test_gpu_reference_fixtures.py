"""End-to-end parity with the compiled reference on its own solutions (fixtures under tests/golden/,
written by tests/golden/make_golden.py from oracle/_ref): the 64-sphere, 8x8-per-face workload with
its 1000 stress points, and 8-sphere RVEs at particle/matrix contrasts 2 and 100. The FP64 path must
give U, T (78 layout) and the stresses to 1e-10 relative; every refined mixed-precision mode to 1e-8.
"""
import os

import numpy as np
import pytest

from ibemproject_amd import engine as E
from ibemproject_amd.formats import read_problem

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["contrast2", "contrast100", "rve64_m8"]
MIXED = ["SOLVE_MIXED_FP16_1", "SOLVE_MIXED_FP16", "SOLVE_MIXED_BF16", "SOLVE_MIXED_BF16X3", "SOLVE_MIXED_FP32",
         "SOLVE_BJ_GMRES", "SOLVE_BJ_GMRES16"]


def case(msg):
    print(f"[case] {msg}", flush=True)


def compare(eng, ref, tol, label):
    U, T, _ = eng.getUTVector()
    s = eng.getFlux()
    eu = np.linalg.norm(U - ref["U"]) / np.linalg.norm(ref["U"])
    et = np.linalg.norm(T - ref["T"]) / np.linalg.norm(ref["T"])
    es = np.abs(s - ref["stress"]).max() / np.abs(ref["stress"]).max()
    print(f"  {label}: U {eu:.2e} T {et:.2e} stress {es:.2e}")
    assert eu < tol and et < tol and es < tol


@pytest.fixture(scope="module", params=CASES)
def fixture(request):
    d = os.path.join(GOLD, request.param)
    if not os.path.exists(os.path.join(d, "reference.npz")):
        pytest.skip(f"{request.param}: fixture not written")
    prob = read_problem(d)
    ref = np.load(os.path.join(d, "reference.npz"))
    assert ref["stress"].shape == (prob.nump, 6)
    return request.param, prob, ref


def test_fp64_against_reference(fixture):
    name, prob, ref = fixture
    case(f"{name}: fp64 LU, n60={prob.n60}, points={prob.nump}")
    eng = E.Engine(prob)
    eng.assemble()
    rel, _ = eng.solve(E.SOLVE_FP64_LU)
    assert rel < 1e-11
    compare(eng, ref, 1e-10, "fp64")
    eng.close()


@pytest.mark.parametrize("mode_name", MIXED)
def test_mixed_against_reference(fixture, mode_name):
    name, prob, ref = fixture
    case(f"{name}: {mode_name}, n60={prob.n60}")
    eng = E.Engine(prob)
    eng.assemble()
    rel, it = eng.solve(getattr(E, mode_name), 1e-10)
    print(f"  residual {rel:.2e} iterations {it}")
    assert rel <= 1e-10
    compare(eng, ref, 1e-8, mode_name)
    if name == "rve64_m8" and not mode_name.startswith("SOLVE_BJ"):  # the compact layout on the same system
        eng.reset()
        eng.assemble()
        rel, it = eng.solve(getattr(E, mode_name) | E.SOLVE_COMPACT, 1e-10)
        assert rel <= 1e-10
        compare(eng, ref, 1e-8, mode_name + " compact")
    eng.close()
