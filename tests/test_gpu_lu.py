"""Dense LU kernels on random systems (MI355X): sizes that are not multiples of the panel (256),
the panel leaf (32) or the GEMM tile (128), compared with numpy/LAPACK."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("num,m", [(4, 2), (9, 3), (23, 2), (64, 8)])
def test_lu_sizes_through_rve(num, m, oracle_lib):
    from ibemproject_amd.engine import SOLVE_FP64_LU, Engine
    from ibemproject_amd.generate import make_rve
    prob = make_rve(100 + num, num, 0.05, 0.07, m, nump=0)
    case(f"rve num={num} m={m} n60={prob.n60}")
    eng = Engine(prob)
    eng.assemble()
    A, b = eng.system()
    rel, _ = eng.solve(SOLVE_FP64_LU)
    _, _, X = eng.getUTVector()
    x_ref = np.linalg.solve(A, b)
    e = np.linalg.norm(X - x_ref) / np.linalg.norm(x_ref)
    r = np.linalg.norm(b - A @ X) / np.linalg.norm(b)
    berr = np.linalg.norm(b - A @ X) / (np.linalg.norm(A, 1) * np.linalg.norm(X, 1))
    print("n", prob.n60, "err vs LAPACK", e, "residual", r, "backward error", berr,
          "engine residual", rel)
    assert berr < 1e-15
    assert r < 1e-11
    assert rel < 1e-11
    assert e < 1e-10
    eng.close()


@pytest.mark.parametrize("mode_name", ["SOLVE_MIXED_FP16", "SOLVE_MIXED_FP32"])
def test_mixed_lu_mid_size(mode_name):
    """A system large enough for the bulk trailing updates to take the 256 x 256 split-product
    kernel (with partial edge tiles) and for the panels' deep updates to take the split path,
    refined to 1e-10 and compared with LAPACK's solution."""
    from ibemproject_amd import engine as E
    from ibemproject_amd.generate import make_rve
    prob = make_rve(321, 150, 0.035, 0.05, 8, nump=0)
    case(f"mixed mid size: {mode_name} num=150 n60={prob.n60}")
    eng = E.Engine(prob)
    eng.assemble()
    A, b = eng.system()
    rel, it = eng.solve(getattr(E, mode_name), 1e-10)
    _, _, X = eng.getUTVector()
    x_ref = np.linalg.solve(A, b)
    e = np.linalg.norm(X - x_ref) / np.linalg.norm(x_ref)
    r = np.linalg.norm(b - A @ X) / np.linalg.norm(b)
    print("n", prob.n60, "iterations", it, "engine residual", rel, "residual", r, "err vs LAPACK", e)
    assert prob.n60 > 9000
    assert rel <= 1e-10 and r <= 1e-10
    assert it <= 8
    assert e < 1e-8
    eng.close()


@pytest.mark.parametrize("seed,num,m", [(11, 3, 7), (12, 2, 8), (13, 5, 10)])
@pytest.mark.parametrize("mode_name", ["SOLVE_MIXED_FP16", "SOLVE_MIXED_FP32"])
def test_mixed_lu_past_first_panel(seed, num, m, mode_name):
    """Systems just past one 1024-column LU step (n = 1068, 1278, 2106): the equilibrated FP32 copy
    of the columns after the first panel runs on the second stream beside that panel and is narrower
    than a panel; the LU has two or three steps, the last a partial one. Refined to 1e-10 and
    compared with LAPACK's solution."""
    from ibemproject_amd import engine as E
    from ibemproject_amd.generate import make_rve
    prob = make_rve(seed, num, 0.05, 0.07, m, nump=0)
    case(f"mixed past first panel: {mode_name} num={num} m={m} n60={prob.n60}")
    eng = E.Engine(prob)
    eng.assemble()
    A, b = eng.system()
    rel, it = eng.solve(getattr(E, mode_name), 1e-10)
    _, _, X = eng.getUTVector()
    x_ref = np.linalg.solve(A, b)
    e = np.linalg.norm(X - x_ref) / np.linalg.norm(x_ref)
    r = np.linalg.norm(b - A @ X) / np.linalg.norm(b)
    print("n", prob.n60, "iterations", it, "engine residual", rel, "residual", r, "err vs LAPACK", e)
    assert 1024 < prob.n60 < 3 * 1024
    assert rel <= 1e-10 and r <= 1e-10
    assert it <= 8
    assert e < 1e-8
    eng.close()
