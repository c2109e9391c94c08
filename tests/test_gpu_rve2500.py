"""The 2,500-sphere RVE (24 x 24 quads per face, n = 160,374) on one MI355X, at full size.

The FP64 system takes 205.8 GB, so the solves use compact storage (the factor in A's place, an FP16
copy of the mixed factor beside A). Size-independent properties: the refined solve reaches a 1e-10
relative FP64 residual at every particle/matrix contrast E1/E0 = 2, 10, 100; the loading fixes the
average stress (<s_xx> = 1, <s_yy> = <s_zz> = 0) and puts the apparent modulus between the Reuss and
Voigt bounds; at contrast 10 the single-product and split-product LU solutions and the block-Jacobi
GMRES solutions (FP64 products only, and with the 16-bit inner matrix) all agree with the plain FP64
LU to 1e-8 in the solution and the stresses."""
import numpy as np
import pytest

from ibemproject_amd import engine as E
from ibemproject_amd.generate import make_workload

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def averages_ok(eng, prob):
    sig, eps, V = eng.average_fields()
    phi = float(np.sum(4.0 / 3.0 * np.pi * prob.radii ** 3)) / V
    Ea = sig[0] / eps[0]
    reuss = 1.0 / (phi / prob.E1 + (1 - phi) / prob.E0)
    voigt = phi * prob.E1 + (1 - phi) * prob.E0
    print(f"  phi {phi:.4f} E_apparent {Ea:.5f} Reuss {reuss:.5f} Voigt {voigt:.5f} "
          f"<s> {sig[0]:.12f} {sig[3]:.2e} {sig[5]:.2e}")
    assert abs(sig[0] - 1.0) < 1e-9 and abs(sig[3]) < 1e-9 and abs(sig[5]) < 1e-9
    assert reuss <= Ea <= voigt


@pytest.mark.parametrize("contrast", [2.0, 100.0])
def test_rve2500_contrast(contrast):
    prob = make_workload("rve2500_m24", seed=1, nump=100, E1=contrast)
    case(f"rve2500_m24 contrast {contrast} n60={prob.n60}")
    eng = E.Engine(prob)
    eng.assemble()
    rel, it = eng.solve(E.SOLVE_MIXED_FP16_1, 1e-10)
    print(f"  fp16 single product, compact: residual {rel:.3e} GMRES iterations {it}")
    assert rel <= 1e-10
    averages_ok(eng, prob)
    _, _, x1 = eng.getUTVector()
    rel, it = eng.solve(E.SOLVE_BJ_GMRES, 1e-10)
    print(f"  block-Jacobi GMRES: residual {rel:.3e} iterations {it} fell back {eng.fell_back()}")
    assert rel <= 1e-10
    _, _, xb = eng.getUTVector()
    assert np.linalg.norm(xb - x1) / np.linalg.norm(x1) < 1e-8
    rel, it = eng.solve(E.SOLVE_BJ_GMRES16, 1e-10)
    print(f"  block-Jacobi GMRES, 16-bit inner: residual {rel:.3e} outer {it} inner {eng.inner_iterations()} "
          f"fell back {eng.fell_back()}")
    assert rel <= 1e-10 and eng.inner_iterations() > 0
    _, _, x16 = eng.getUTVector()
    assert np.linalg.norm(x16 - x1) / np.linalg.norm(x1) < 1e-8
    eng.close()


def test_rve2500_against_fp64_lu():
    prob = make_workload("rve2500_m24", seed=1, nump=200)
    case(f"rve2500_m24 contrast 10 n60={prob.n60}: fp16 single, fp16 split, fp64 LU")
    eng = E.Engine(prob)
    eng.assemble()
    rel1, it1 = eng.solve(E.SOLVE_MIXED_FP16_1, 1e-10)
    _, _, x1 = eng.getUTVector()
    s1 = eng.getFlux()
    print(f"  fp16 single: residual {rel1:.3e} iterations {it1}")
    assert rel1 <= 1e-10
    averages_ok(eng, prob)
    relb, itb = eng.solve(E.SOLVE_BJ_GMRES, 1e-10)
    _, _, xb = eng.getUTVector()
    sb = eng.getFlux()
    print(f"  block-Jacobi GMRES: residual {relb:.3e} iterations {itb} fell back {eng.fell_back()}")
    assert relb <= 1e-10
    rel16, it16 = eng.solve(E.SOLVE_BJ_GMRES16, 1e-10)
    _, _, x16 = eng.getUTVector()
    s16 = eng.getFlux()
    print(f"  block-Jacobi GMRES, 16-bit inner: residual {rel16:.3e} outer {it16} inner {eng.inner_iterations()}")
    assert rel16 <= 1e-10 and eng.inner_iterations() > 0
    eng.reset()
    eng.assemble()
    rel3, it3 = eng.solve(E.SOLVE_MIXED_FP16, 1e-10)
    _, _, x3 = eng.getUTVector()
    s3 = eng.getFlux()
    print(f"  fp16 split: residual {rel3:.3e} iterations {it3}")
    assert rel3 <= 1e-10
    eng.reset()
    eng.assemble()
    rel64, _ = eng.solve(E.SOLVE_FP64_LU)
    assert rel64 == -1.0  # compact: A holds the factor
    _, _, x64 = eng.getUTVector()
    s64 = eng.getFlux()
    rel64 = eng.residual()
    print(f"  fp64 LU (in place): residual {rel64:.3e}")
    assert rel64 <= 1e-10
    for name, x, s in (("fp16 single", x1, s1), ("fp16 split", x3, s3), ("block-Jacobi GMRES", xb, sb),
                       ("block-Jacobi GMRES 16-bit inner", x16, s16)):
        ex = np.linalg.norm(x - x64) / np.linalg.norm(x64)
        es = np.abs(s - s64).max() / np.abs(s64).max()
        print(f"  {name} vs fp64 LU: x {ex:.3e} stress {es:.3e}")
        assert ex < 1e-8 and es < 1e-8
    eng.close()
