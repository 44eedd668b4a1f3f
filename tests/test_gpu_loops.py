"""Design loops on the GPU against oracle loops: the 128^3 OC loop of the
plain optimisation workload, and runCustom's OC branch with its adaptive
volume bound (VolumeGovernor) and convergence test."""
import sys

import numpy as np
import pytest

import opentm_amd as otm
from opentm_amd import workloads as wl

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.timeout(900)
def test_oc_loop_128_matches_oracle(orc):
    """Three iterations of the 128^3 OC workload (max trace kappa^H at volume
    0.3, move 0.2, damping 0.5, bisection to 1e-6, Spline4 r=1.5 density
    filter, periodic linear r=2 sensitivity filter) run on the GPU and, from
    the same start, in the FP64 oracle (MG-preconditioned CG to 1e-10):
    densities within 1e-4 L-inf at every iteration, and kappa^H of the GPU's
    own design within 1e-5 of the oracle's on that design."""
    n = 128
    k0, k1 = 1.0, 1e-3
    rho0 = wl.oc_start_density(n, 0)
    cfg = otm.HomoConfig(n, (k0, k1))
    hom = otm.Homogenization_H(cfg)
    hom.set_design(rho0)
    oc = otm.OCOptimizer(n ** 3, 0.001, 0.2, 0.5, bisect_steps=60, bisect_tol=1e-6)
    dC = -np.eye(3)
    r = rho0.astype(np.float64)
    for it in range(3):
        say("oc128 it", it)
        rg_in = hom.get_design().reshape(r.shape).astype(np.float64)
        C = hom.eval()
        hom.backward(dC, want=False)
        oc.step(hom, 2.0, 0.3)
        rg = hom.get_design().reshape(r.shape)
        if it != 1:  # kappa^H of the GPU's own design (first and last iterations)
            kg = k1 + (k0 - k1) * orc.filter_fwd(rg_in, 0, 1.5, False) ** 3
            Cg, _, _, _, rrg = orc.mg_homogenize(kg, rtol=1e-10)
            say("  C err (same design)", rel(C, Cg), "oracle relres", max(rrg))
            assert rel(C, Cg) < 1e-5
        rf = orc.filter_fwd(r, 0, 1.5, False)
        kap = k1 + (k0 - k1) * rf ** 3
        _, _, So, _, rro = orc.mg_homogenize(kap, rtol=1e-10, want_S=True)
        sk = orc.sensitivity(So, dC).reshape(r.shape)
        g = orc.filter_adj(sk * 3 * (k0 - k1) * rf ** 2, 0, 1.5, False)
        gs = orc.oc_filter_sens(r, g, 2.0)
        r, steps, _ = orc.oc_update(gs, r, 0.3, 0.001, 0.2, 0.5, 60, 1e-6)
        r = r.reshape(rho0.shape)
        drho = float(np.abs(rg - r).max())
        say("  rho err", drho, "gpu stats", hom.solve_stats())
        assert drho < 1e-4


def test_run_custom_oc_matches_oracle_loop(orc):
    """runCustom's OC branch (Framework.cu:203-231): objective
    sum (C_ij - t_ij)^2 - 1e-4, VolumeGovernor bound (from the mean SIMP
    conductivity and the mean density), ConvergeChecker, OC with move 0.02
    and the reference's 20-step / 1e-4 bisection: ten iterations against
    the oracle loop with the restated governor, from the same start."""
    from oracle.framework import ConvergeChecker, VolumeGovernor, tensor_objective
    n = 16
    k0, k1 = 1.0, 1e-4
    t = np.array([0.5, 0.4, 0.3, 0.0, 0.0, 0.0])
    rho0 = wl.oc_start_density(n, 4, mean=0.5, amplitude=0.2)
    iters = 10
    say("runCustom oc loop", n, iters)
    rg, rep = otm.run_custom(n, "oc", iters, rho0=rho0, target=t, heat_ratio=(k0, k1), fem_rel_thres=1e-7)
    r = rho0.astype(np.float64)
    gov = VolumeGovernor()
    chk = ConvergeChecker(5e-4)
    bounds = []
    done = iters
    for itn in range(iters):
        rf = orc.filter_fwd(r, 0, 1.5, False)
        kap = k1 + (k0 - k1) * rf ** 3
        C, _, S, _ = orc.homogenize(kap, rtol=1e-12, want_S=True)
        val, dC = tensor_objective(C, t)
        val = np.float32(val)
        gov.volume_check(val, np.float32(kap.mean()), np.float32(r.mean()), itn)
        sk = orc.sensitivity(S, dC).reshape(r.shape)
        g = orc.filter_adj(sk * 3 * (k0 - k1) * rf ** 2, 0, 1.5, False)
        if chk.is_converge(itn, val) and float(gov.get_current_decrease()) < 1e-4:
            done = itn
            break
        gs = orc.oc_filter_sens(r, g, 2.0)
        vb = float(gov.get_volume_bound())
        bounds.append(vb)
        r, steps, _ = orc.oc_update(gs, r, vb, 0.001, 0.02, 0.5, 20, 1e-4)
        r = r.reshape(rho0.shape)
    err = float(np.abs(rg.reshape(r.shape) - r).max())
    say("  iterations gpu", rep["iterations"], "oracle", done, "bounds", [round(b, 4) for b in bounds], "rho err", err)
    assert rep["iterations"] == done
    assert rep["unconverged_solves"] == 0
    assert err < 1e-4
