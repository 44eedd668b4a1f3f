"""The precision form at its configured size, 256^3 (16.8M elements): FP64
iterative refinement with FP16-stored coarse stencils. Closed forms that
hold at any size, the residual target, and the MMA loop's first steps."""
import sys

import numpy as np
import pytest

import opentm_amd as otm
from opentm_amd import precision as pf
from opentm_amd import workloads as wl

pytestmark = pytest.mark.gpu
N = 256


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def hom256():
    cfg = otm.HomoConfig(N, (1.0, 1e-3))
    cfg.femRelThres = 1e-8
    cfg.fp64_refine = 1
    cfg.fp16_stencils = 1
    h = otm.Homogenization_H(cfg)
    yield h
    h.close()


def test_256_uniform_is_kappa_identity(hom256):
    say("256 uniform")
    hom256.reset_displacement()
    hom256.update(np.full((N, N, N), 0.37, np.float32))
    C = hom256.heatMatrix()
    assert rel(C, (0.37 + 1e-9) * np.eye(3)) < 1e-7


def test_256_laminate_harmonic_and_arithmetic(hom256):
    """x-laminate: the harmonic mean across the layers lands in the
    reference's C[1][1] (load case 1 is along grid x), the arithmetic mean
    in C[0][0] and C[2][2]; within 1e-7."""
    say("256 laminate")
    vals = np.where(np.arange(N) % 32 < 11, 1.0, 0.02).astype(np.float32)
    lam = np.broadcast_to(vals[None, None, :], (N, N, N)).copy()
    hom256.reset_displacement()
    hom256.update(lam)
    C = hom256.heatMatrix()
    its, rr = hom256.solve_stats()
    say("  its", its, "relres", rr)
    ve = vals.astype(np.float64) + 1e-9
    assert C[1, 1] == pytest.approx(1.0 / np.mean(1.0 / ve), rel=1e-7)
    assert C[0, 0] == pytest.approx(np.mean(ve), rel=1e-7)
    assert C[2, 2] == pytest.approx(np.mean(ve), rel=1e-7)


def test_256_cos_mode_residual_and_symmetry(hom256):
    """The config's density (cos modes at volume 0.25): true relative
    residual <= 1e-8 for every load case, kappa^H exactly symmetric,
    positive definite and inside the Wiener bounds."""
    say("256 cos modes")
    rho = wl.plain_density(N, 0, mean=0.25)
    k = wl.simp(rho)
    hom256.reset_displacement()
    hom256.update(k)
    C = hom256.heatMatrix()
    its, rr = hom256.solve_stats()
    say("  its", its, "relres", rr, "C", np.round(C, 6).tolist())
    assert max(rr) <= 1e-8
    np.testing.assert_array_equal(C, C.T)
    w = np.linalg.eigvalsh(C)
    ke = k.astype(np.float64) + 1e-9
    assert w.min() >= 1.0 / np.mean(1.0 / ke) * (1 - 1e-7) and w.max() <= np.mean(ke) * (1 + 1e-7)


def test_256_mma_m3_two_iterations():
    """Two MMA updates (m = 3: C11, C22 <= 0.05 and volume <= 0.25) of the
    precision-form loop: every solve reaches 1e-8, the densities stay in
    [1e-3, 1] and move, and the iterates stay feasible (C11, C22 well below
    their bound; the volume constraint, active from the start, kept to the
    subproblem's tolerance)."""
    say("256 mma")
    rho0 = wl.plain_density(N, 0, mean=0.25)
    loop = pf.PrecisionLoop(N, rho0, rtol=1e-8, refine=True, fp16=True)
    h0 = loop.step()
    h1 = loop.step()
    h2 = loop.step()
    for h in (h0, h1, h2):
        say("  itn", h["itn"], "C00", h["C"][0, 0], "g", np.round(h["g"], 5).tolist(), "relres", max(h["relres"]))
        assert max(h["relres"]) <= 1e-8
    x = loop.design()
    assert x.min() >= 1e-3 - 1e-7 and x.max() <= 1.0
    assert float(np.abs(x - rho0.reshape(-1)).max()) > 1e-3
    for h in (h1, h2):
        assert h["g"][0] < 0 and h["g"][1] < 0 and h["g"][2] <= 1e-3
    loop.mma.close()
    loop.hom.close()
