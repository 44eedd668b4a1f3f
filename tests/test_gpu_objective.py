"""Generic objectives on the GPU: full field reductions (sum, p-norm,
log-sum-exp) against numpy in FP64 with their gradients, an objective
expression's dC fed to the solver's backward, and runCustom's iteration
log (logIter)."""
import ctypes
import sys

import numpy as np
import pytest

import opentm_amd as otm
from opentm_amd import _capi
from opentm_amd import workloads as wl
from opentm_amd.objective import field_reduce, heat_tensor, tensor_grad

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def reduce(x, kind, p=2.0, upstream=1.0, device=False):
    kinds = {"sum": 0, "pnorm": 1, "logsumexp": 2}
    val = ctypes.c_double()
    if device:
        dx = otm.DeviceBuffer.from_numpy(x)
        dg = otm.DeviceBuffer(x.size)
        _capi.check(_capi.load().otm_field_reduce(dx.ptr, x.size, kinds[kind], p, 1, ctypes.byref(val), dg.ptr,
                                                  upstream), "otm_field_reduce")
        g = dg.numpy()
        dx.free()
        dg.free()
        return val.value, g
    g = np.empty(x.size, np.float32)
    _capi.check(_capi.load().otm_field_reduce(x.ctypes.data, x.size, kinds[kind], p, 0, ctypes.byref(val),
                                              g.ctypes.data, upstream), "otm_field_reduce")
    return val.value, g


@pytest.mark.parametrize("kind,p", [("sum", 0.0), ("pnorm", 3.0), ("pnorm", 8.0), ("logsumexp", 0.0)])
@pytest.mark.parametrize("device", [False, True])
def test_field_reductions_match_numpy(kind, p, device):
    rng = np.random.default_rng(41)
    x = rng.uniform(1e-3, 1.0, 300007).astype(np.float32)  # not a multiple of the block
    say("field reduce", kind, p, "device" if device else "host")
    v, g = reduce(x, kind, p, upstream=0.7, device=device)
    xd = x.astype(np.float64)
    if kind == "sum":
        ref, gref = xd.sum(), np.ones_like(xd)
    elif kind == "pnorm":
        ref = (xd ** p).sum() ** (1.0 / p)
        gref = (xd / ref) ** (p - 1.0)
    else:
        m = xd.max()
        ref = m + np.log(np.exp(xd - m).sum())
        gref = np.exp(xd - ref)
    assert v == pytest.approx(ref, rel=1e-12)
    assert np.allclose(g, 0.7 * gref, rtol=2e-6, atol=1e-30)


def test_field_reduction_is_deterministic():
    x = np.random.default_rng(42).standard_normal(1 << 20).astype(np.float32)
    assert reduce(x, "logsumexp")[0] == reduce(x, "logsumexp")[0]


def test_objective_dC_drives_backward(orc):
    """d objective / d C of an expression (log of a squared error plus a
    p-norm of the design) through otm_backward equals the same chain rule
    written by hand: the expression layer only forms dC."""
    n = 16
    hom = otm.Homogenization_H(otm.HomoConfig(n, (1.0, 1e-3)))
    rho = wl.plain_density(n, 3).astype(np.float32).reshape(-1)
    hom.set_design(rho)
    C = hom.eval()
    c = heat_tensor(C)
    g_field = np.zeros(rho.size, np.float32)
    obj = ((c[0][0] - 0.3) ** 2 + (c[1][1] - 0.2) ** 2).log() + 0.05 * field_reduce(rho, "pnorm", 4.0, g_field)
    obj.backward()
    dC = tensor_grad(c)
    e0, e1 = C[0, 0] - 0.3, C[1, 1] - 0.2
    s = e0 ** 2 + e1 ** 2
    dC_hand = np.zeros((3, 3))
    dC_hand[0, 0], dC_hand[1, 1] = 2 * e0 / s, 2 * e1 / s
    assert np.allclose(dC, dC_hand, rtol=1e-13)
    pn = (rho.astype(np.float64) ** 4).sum() ** 0.25
    assert np.allclose(g_field, 0.05 * (rho / pn) ** 3, rtol=2e-6)
    total = np.asarray(hom.backward(dC)).reshape(-1) + g_field
    assert np.array_equal(np.asarray(hom.backward(dC_hand)).reshape(-1) + g_field, total)
    hom.close()


def test_run_custom_iteration_log(tmp_path):
    """logIter's files from otm_run_custom: the design and tensor every 5th
    iteration, rho_%04d / sens_%04d / Clog / objlog at their intervals."""
    n, iters = 16, 7
    pre = str(tmp_path / "run_")
    rho, rep = otm.run_custom(n, model="oc", max_iter=iters, initway="IWP", log_prefix=pre, logrho=3, logc=1,
                              logsens=2, logobj=1)
    say("run_custom log", rep["iterations"])
    assert rep["iterations"] == iters
    clog = (tmp_path / "run_Clog").read_text().splitlines()
    objlog = (tmp_path / "run_objlog").read_text().splitlines()
    assert len(clog) == iters and len(objlog) == iters
    assert clog[0].startswith("iter 0 ") and len(clog[-1].split()) == 2 + 9
    last_C = np.array([float(v) for v in clog[-1].split()[2:]]).reshape(3, 3)
    assert np.allclose(last_C, rep["C"], rtol=1e-8)
    assert float(objlog[-1].split("=")[1]) == pytest.approx(rep["objective"], rel=1e-8)
    for it in (0, 3, 6):
        a = otm.read_density(str(tmp_path / f"run_rho_{it:04d}"))
        assert a.shape == (n, n, n)
        assert a.min() >= 1e-3 - 1e-7 and a.max() <= 1.0 + 1e-7
    # the log holds the design each iteration evaluated (before its update)
    assert not np.array_equal(otm.read_density(str(tmp_path / "run_rho_0003")),
                              otm.read_density(str(tmp_path / "run_rho_0006")))
    for it in (0, 2, 4, 6):
        assert (tmp_path / f"run_sens_{it:04d}").exists()
    assert not (tmp_path / "run_sens_0001").exists()
    assert (tmp_path / "run_rho").exists() and len((tmp_path / "run_C").read_text().split()) == 9
