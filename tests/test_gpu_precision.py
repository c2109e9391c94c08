"""Factor precisions and storage layouts of the solve (MI355X), against LAPACK on the same assembled
system: single FP16 and BF16 trailing-update products and split (hi/lo) BF16 products (per-row /
per-column power-of-two scaled operands), and compact storage, where the factor takes the place of A (FP64 in place; mixed modes:
FP32 in place, then a 16-bit per-tile-scaled copy and A assembled again). Every refined result must
reach a 1e-10 relative residual and agree with LAPACK to 1e-8; the FP64 path to 1e-10. Each test
prints its error against the FP64 LU solution."""
import numpy as np
import pytest

from ibemproject_amd import engine as E
from ibemproject_amd.generate import make_rve

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def mid():
    """A system past several LU steps whose bulk updates take the 256 x 256 kernel, and whose
    in-place FP32 conversion needs blocks beyond the staged first columns."""
    prob = make_rve(321, 150, 0.035, 0.05, 8, nump=40)
    case(f"mid size num=150 n60={prob.n60}")
    eng = E.Engine(prob)
    eng.assemble()
    A, b = eng.system()
    eng.solve(E.SOLVE_FP64_LU)
    _, _, x64 = eng.getUTVector()
    s64 = eng.getFlux()
    eng.close()
    x_ref = np.linalg.solve(A, b)
    return dict(prob=prob, A=A, b=b, x_ref=x_ref, x64=x64, s64=s64)


def check_refined(eng, d, rel, it, label):
    A, b, x_ref = d["A"], d["b"], d["x_ref"]
    _, _, X = eng.getUTVector()
    e = np.linalg.norm(X - x_ref) / np.linalg.norm(x_ref)
    e64 = np.linalg.norm(X - d["x64"]) / np.linalg.norm(d["x64"])
    r = np.linalg.norm(b - A @ X) / np.linalg.norm(b)
    s = eng.getFlux()
    es = np.abs(s - d["s64"]).max() / np.abs(d["s64"]).max()
    print(f"{label}: n {len(b)} iterations {it} engine residual {rel:.3e} residual {r:.3e} "
          f"err vs LAPACK {e:.3e} vs FP64 LU {e64:.3e} stress vs FP64 LU {es:.3e}")
    assert rel <= 1e-10 and r <= 1.5e-10
    assert e < 1e-8 and e64 < 1e-8 and es < 1e-8
    return X


@pytest.mark.parametrize("mode_name", ["SOLVE_MIXED_FP16_1", "SOLVE_MIXED_BF16", "SOLVE_MIXED_BF16X3"])
def test_single_product_modes(mid, mode_name):
    case(f"single-product LU: {mode_name} n60={mid['prob'].n60}")
    eng = E.Engine(mid["prob"])
    eng.assemble()
    rel, it = eng.solve(getattr(E, mode_name), 1e-10)
    check_refined(eng, mid, rel, it, mode_name)
    eng.close()


@pytest.mark.parametrize("mode_name", ["SOLVE_MIXED_FP32", "SOLVE_MIXED_FP16", "SOLVE_MIXED_FP16_1",
                                       "SOLVE_MIXED_BF16", "SOLVE_MIXED_BF16X3"])
def test_compact_mixed_modes(mid, mode_name):
    """Compact storage: the FP32 factor overwrites A in place, its 16-bit copy preconditions FP64
    GMRES against the re-assembled A, which must equal the first assembly bit for bit."""
    case(f"compact mixed: {mode_name} n60={mid['prob'].n60}")
    eng = E.Engine(mid["prob"])
    eng.assemble()
    rel, it = eng.solve(getattr(E, mode_name) | E.SOLVE_COMPACT, 1e-10)
    check_refined(eng, mid, rel, it, mode_name + " compact")
    A2, b2 = eng.system()
    assert np.array_equal(A2, mid["A"]) and np.array_equal(b2, mid["b"])
    eng.close()


def test_compact_fp64(mid):
    """FP64 LU in place of A: no residual from the solve; ibem3d_residual assembles A again."""
    case(f"compact fp64 n60={mid['prob'].n60}")
    eng = E.Engine(mid["prob"])
    eng.assemble()
    rel, _ = eng.solve(E.SOLVE_FP64_LU | E.SOLVE_COMPACT)
    assert rel == -1.0
    _, _, X = eng.getUTVector()
    s = eng.getFlux()
    with pytest.raises(E.EngineError):
        eng.system()  # A holds the factor
    r_engine = eng.residual()
    A2, b2 = eng.system()
    assert np.array_equal(A2, mid["A"]) and np.array_equal(b2, mid["b"])
    e = np.linalg.norm(X - mid["x_ref"]) / np.linalg.norm(mid["x_ref"])
    e64 = np.linalg.norm(X - mid["x64"]) / np.linalg.norm(mid["x64"])
    es = np.abs(s - mid["s64"]).max() / np.abs(mid["s64"]).max()
    print(f"compact fp64: residual {r_engine:.3e} err vs LAPACK {e:.3e} vs FP64 LU (copy) {e64:.3e} "
          f"stress {es:.3e}")
    assert r_engine < 1e-11 and e < 1e-10 and es < 1e-10
    # the factor was overwritten by the assembly: reuse is refused
    with pytest.raises(E.EngineError):
        eng.solve(E.SOLVE_FP64_LU | E.SOLVE_REUSE)
    eng.close()


@pytest.mark.parametrize("seed,num,m", [(11, 3, 7), (13, 5, 10)])
def test_compact_past_first_panel(seed, num, m):
    """Compact mixed solves of systems just past one LU step (the whole FP32 conversion staged)."""
    prob = make_rve(seed, num, 0.05, 0.07, m, nump=0)
    case(f"compact past first panel: num={num} m={m} n60={prob.n60}")
    eng = E.Engine(prob)
    eng.assemble()
    A, b = eng.system()
    rel, it = eng.solve(E.SOLVE_MIXED_FP16_1 | E.SOLVE_COMPACT, 1e-10)
    _, _, X = eng.getUTVector()
    x_ref = np.linalg.solve(A, b)
    e = np.linalg.norm(X - x_ref) / np.linalg.norm(x_ref)
    print("n", prob.n60, "iterations", it, "residual", rel, "err vs LAPACK", e)
    assert rel <= 1e-10 and e < 1e-8
    eng.close()


def test_compact_reuse_new_load(mid):
    """The 16-bit factor copy stays valid when only the known values change (A is assembled again by
    the caller): a reused solve matches a fresh one."""
    from ibemproject_amd.engine import unit_strain
    from ibemproject_amd.formats import Problem
    from ibemproject_amd.generate import macro_strain_bc
    p = mid["prob"]
    nodes = np.asarray(p.nodes).reshape(-1, 3)
    conn = np.asarray(p.conn).reshape(-1, 4)
    bc0, tbc0 = macro_strain_bc(nodes, conn, unit_strain(0))
    bc1, tbc1 = macro_strain_bc(nodes, conn, unit_strain(4))
    p0 = Problem(p.title, p.E0, p.nu0, p.E1, p.nu1, nodes, conn, bc0, tbc0, p.centres, p.radii,
                 np.zeros((0, 3)))
    p1 = Problem(p.title, p.E0, p.nu0, p.E1, p.nu1, nodes, conn, bc1, tbc1, p.centres, p.radii,
                 np.zeros((0, 3)))
    case(f"compact reuse n60={p.n60}")
    mode = E.SOLVE_MIXED_FP16_1 | E.SOLVE_COMPACT
    eng = E.Engine(p0)
    eng.assemble()
    eng.solve(mode, 1e-10)
    eng.set_known_values(bc1[:, 1], tbc1[:, 12:])
    eng.assemble()
    rel, it = eng.solve(E.SOLVE_MIXED_FP16_1 | E.SOLVE_REUSE, 1e-10)
    assert rel <= 1e-10
    _, _, x_reuse = eng.getUTVector()
    eng.close()
    eng = E.Engine(p1)
    eng.assemble()
    eng.solve(E.SOLVE_FP64_LU)
    _, _, x_fresh = eng.getUTVector()
    eng.close()
    e = np.linalg.norm(x_reuse - x_fresh) / np.linalg.norm(x_fresh)
    print("reuse iterations", it, "residual", rel, "vs fresh FP64", e)
    assert e < 1e-8


def test_block_jacobi_gmres(mid):
    """FP64 GMRES preconditioned by block Jacobi (no n x n factor): 1e-10 residual, the LU
    solution to 1e-8; started from the exact eigenstrains perturbed by 1 %, it needs fewer
    iterations than from zero, and the same solution."""
    case(f"block-Jacobi GMRES n60={mid['prob'].n60}")
    eng = E.Engine(mid["prob"])
    eng.assemble()
    rel, it = eng.solve(E.SOLVE_BJ_GMRES, 1e-10)
    assert not eng.fell_back()
    X = check_refined(eng, mid, rel, it, "bj-gmres cold")
    nb = 3 * mid["prob"].NN
    rng = np.random.default_rng(3)
    guess = mid["x64"][nb:] * (1 + 0.01 * rng.standard_normal(X.size - nb))
    eng.set_initial_guess(guess)
    rel_w, it_w = eng.solve(E.SOLVE_BJ_GMRES | E.SOLVE_REUSE, 1e-10)
    check_refined(eng, mid, rel_w, it_w, "bj-gmres warm 1%")
    print("iterations cold", it, "warm", it_w)
    assert it_w < it
    eng.set_initial_guess(None)
    eng.close()


def test_block_jacobi_gmres16(mid):
    """Flexible FP64 GMRES whose directions come from inner block-Jacobi GMRES on the FP16 copy of A:
    1e-10 residual, the LU solution to 1e-8, and far fewer outer iterations (FP64 products) than
    block-Jacobi GMRES needs iterations (a wrong 16-bit product would leave the inner directions poor
    and the outer count high)."""
    case(f"block-Jacobi GMRES, 16-bit inner, n60={mid['prob'].n60}")
    eng = E.Engine(mid["prob"])
    eng.assemble()
    rel_b, it_b = eng.solve(E.SOLVE_BJ_GMRES, 1e-10)
    rel, it = eng.solve(E.SOLVE_BJ_GMRES16, 1e-10)
    inner = eng.inner_iterations()
    assert not eng.fell_back() and inner > 0
    check_refined(eng, mid, rel, it, "bj-gmres16")
    print("block-Jacobi GMRES iterations", it_b, "| 16-bit inner: outer", it, "inner", inner)
    assert 2 * it <= it_b
    # reuse with the copy in place, then a warm start from the solution perturbed by 0.1 %
    rel_r, it_r = eng.solve(E.SOLVE_BJ_GMRES16 | E.SOLVE_REUSE, 1e-10)
    assert rel_r <= 1e-10 and it_r == it
    nb = 3 * mid["prob"].NN
    rng = np.random.default_rng(4)
    eng.set_initial_guess(mid["x64"][nb:] * (1 + 1e-3 * rng.standard_normal(mid["x64"].size - nb)))
    rel_w, it_w = eng.solve(E.SOLVE_BJ_GMRES16 | E.SOLVE_REUSE, 1e-10)
    check_refined(eng, mid, rel_w, it_w, "bj-gmres16 warm")
    eng.set_initial_guess(None)
    # a LU mode afterwards gets the copy's memory back and still solves
    rel_l, _ = eng.solve(E.SOLVE_MIXED_FP16_1, 1e-10)
    assert rel_l <= 1e-10
    eng.close()


def test_block_jacobi_setup_beside_assembly(mid):
    """After a block-Jacobi solve the next assembly builds the preconditioner on a second stream: the
    next solves reproduce the first bit for bit, also after a reset that follows such an assembly
    directly, and an LU mode after it still solves."""
    case(f"block-Jacobi setup beside the assembly, n60={mid['prob'].n60}")
    eng = E.Engine(mid["prob"])
    xs = []
    for step in range(3):
        eng.reset()
        eng.assemble()
        if step == 1:  # a reset straight after the speculative assembly, then assemble again
            eng.reset()
            eng.assemble()
        rel, it = eng.solve(E.SOLVE_BJ_GMRES16, 1e-10)
        _, _, x = eng.getUTVector()
        xs.append((rel, it, eng.inner_iterations(), x))
        assert rel <= 1e-10 and not eng.fell_back()
    tm = eng.timings()
    print("steps", [(r, i, k) for r, i, k, _ in xs], "aux phases",
          {k: round(v, 2) for k, v in tm.items() if k.endswith("_aux")})
    assert "assemble.bj_boundary_aux" in tm and "assemble.bj_blocks_aux" in tm
    for r, i, k, x in xs[1:]:
        assert (r, i, k) == xs[0][:3] and np.array_equal(x, xs[0][3])
    eng.reset()
    eng.assemble()
    rel, _ = eng.solve(E.SOLVE_MIXED_FP16_1, 1e-10)
    assert rel <= 1e-10
    _, _, x = eng.getUTVector()
    assert np.linalg.norm(x - xs[0][3]) / np.linalg.norm(xs[0][3]) < 1e-8
    eng.close()


def test_bj_boundary_without_interchanges(mid):
    """The block-Jacobi boundary block factored without row interchanges in its natural column order
    in FP32 (the default) or FP64, with partial pivoting, and through the pivoting fallback (a
    threshold no pivot meets), and its triangular solves beside the FP16 product (the default) or
    before it: the preconditioners are the same matrix inverse to FP32 accuracy or better, so each
    solve reaches 1e-10 with the same iteration counts (+-1) and the solutions agree to 1e-8 (the bar
    of the refined modes); the second assembly's setup on the second stream takes the same path,
    and repeats the first bit for bit (outer and inner iteration counts, solution)."""
    case(f"block-Jacobi boundary block without interchanges, n60={mid['prob'].n60}")
    out = {}
    # (IBEM3D_TUNE_BJ_NOPIV, IBEM3D_TUNE_BJ_OVERLAP, IBEM3D_TUNE_BJ_FP32) and the factor expected
    configs = [((1, 1, 1), 2), ((1, 1, 0), 1), ((0, 1, 1), 0), ((2, 1, 1), 0), ((1, 0, 1), 2)]
    try:
        for cfg, want in configs:
            E.set_tuning(E.TUNE_BJ_NOPIV, cfg[0])
            E.set_tuning(E.TUNE_BJ_OVERLAP, cfg[1])
            E.set_tuning(E.TUNE_BJ_FP32, cfg[2])
            eng = E.Engine(mid["prob"])
            res = []
            for step in range(2):  # the second assembly builds the preconditioner beside it
                eng.reset()
                eng.assemble()
                rel, it = eng.solve(E.SOLVE_BJ_GMRES16, 1e-10)
                assert rel <= 1e-10 and not eng.fell_back()
                _, _, X = eng.getUTVector()
                res.append((it, eng.inner_iterations(), X, eng.bj_nopiv()))
            rel_b, it_b = eng.solve(E.SOLVE_BJ_GMRES, 1e-10)
            assert rel_b <= 1e-10
            res.append((it_b, 0, eng.getUTVector()[2], eng.bj_nopiv()))
            out[cfg] = res
            print(f"  setting {cfg}: " + ", ".join(f"its {r[0]}/{r[1]} factor {r[3]}" for r in res))
            assert all(r[3] == want for r in res)
            eng.close()
    finally:
        E.set_tuning(E.TUNE_BJ_NOPIV, 1)
        E.set_tuning(E.TUNE_BJ_OVERLAP, 1)
        E.set_tuning(E.TUNE_BJ_FP32, 1)
    for cfg, _ in configs:  # the two assemblies and solves of each setting agree bit for bit
        assert out[cfg][0][:2] == out[cfg][1][:2] and np.array_equal(out[cfg][0][2], out[cfg][1][2])
    base = out[(0, 1, 1)]
    x0 = base[0][2]
    for cfg, _ in configs:
        for a, b in zip(out[cfg], base):
            assert abs(a[0] - b[0]) <= 1 and abs(a[1] - b[1]) <= 2
            assert np.linalg.norm(a[2] - b[2]) / np.linalg.norm(x0) < 1e-8


def test_stress_beside_final_residual(mid):
    """Stresses computed on the second stream beside a solve's final residual check (the default)
    equal the ones computed after the solve bit for bit, for the 16-bit-inner block-Jacobi mode and
    a GMRES-IR LU mode, and after the known values change the solve's new stresses are returned,
    not the previous ones."""
    p = mid["prob"]
    bc = np.asarray(p.bc).reshape(-1, 2)
    tbc = np.asarray(p.tbc).reshape(-1, 24)
    case(f"stresses beside the final residual, n60={p.n60}, points={len(np.asarray(p.points).reshape(-1, 3))}")
    out = {}
    try:
        for flag in (1, 0):
            E.set_tuning(E.TUNE_STRESS_BESIDE, flag)
            eng = E.Engine(p)
            res = []
            for load, mode in ((1.0, E.SOLVE_BJ_GMRES16), (2.0, E.SOLVE_BJ_GMRES16),
                               (1.0, E.SOLVE_MIXED_FP16_1 | E.SOLVE_GMRES)):
                eng.set_known_values(load * bc[:, 1], load * tbc[:, 12:])
                eng.assemble()
                rel, _ = eng.solve(mode, 1e-10)
                assert rel <= 1e-10
                res.append(np.array(eng.getFlux(), copy=True))
                res.append(np.array(eng.getFlux(), copy=True))  # a second request: the same values
            out[flag] = res
            eng.close()
    finally:
        E.set_tuning(E.TUNE_STRESS_BESIDE, 1)
    for a, b in zip(out[1], out[0]):
        assert np.array_equal(a, b)
    s1, s2 = out[1][0], out[1][2]
    assert not np.array_equal(s1, s2)
    assert np.linalg.norm(s2 - 2.0 * s1) <= 1e-8 * np.linalg.norm(s2)


def test_residual_from_cycle_products(mid):
    """The GMRES cycle that meets the tolerance checks its x with r - (A Z) y from its own FP64
    products (the default) instead of another product b - A x: the same x bit for bit, the same
    iteration counts, and a reported residual that matches b - A x computed by a product to rounding
    (for the 16-bit-inner block-Jacobi mode and a GMRES-IR LU mode)."""
    p = mid["prob"]
    case(f"residual from the cycle's products, n60={p.n60}")
    out = {}
    try:
        for flag in (1, 0):
            E.set_tuning(E.TUNE_RESIDUAL_UPDATE, flag)
            eng = E.Engine(p)
            eng.assemble()
            res = []
            for mode in (E.SOLVE_BJ_GMRES16, E.SOLVE_MIXED_FP16_1 | E.SOLVE_GMRES):
                rel, it = eng.solve(mode, 1e-10)
                _, _, x = eng.getUTVector()
                res.append((rel, it, np.array(x, copy=True), eng.residual()))
            out[flag] = res
            eng.close()
    finally:
        E.set_tuning(E.TUNE_RESIDUAL_UPDATE, 1)
    for (r1, i1, x1, t1), (r0, i0, x0, t0) in zip(out[1], out[0]):
        print(f"reported {r1:.3e} (update) {r0:.3e} (product); by a product afterwards {t1:.3e}")
        assert i1 == i0
        assert np.array_equal(x1, x0)
        assert r1 <= 1e-10 and t1 <= 1e-10
        assert r0 == t0
        assert abs(r1 - t1) <= 1e-2 * t1 + 1e-15
