"""opentm_amd -- MI355X (gfx950) thermal inverse-homogenization engine.

Python mirror of the reference's operator interface, over the C ABI of
include/opentm.h (libopentm.so, hand-written HIP kernels):

* ``HomoConfig``        -- cfg::HomoConfig (cmdline.h:12-33)
* ``Homogenization_H``  -- homo::Homogenization_H (homogenization.h:15-32)
* ``OCOptimizer``       -- homo::OCOptimizer (OCOptimizer.h:7-20)
* ``conv_filter``       -- the radial convolution of conv_tsexp_t
                           (TensorExpression.h:1095-1185)

Host arrays are numpy float32 (x fastest, i.e. shape (nz, ny, nx) in C
order); ``DeviceBuffer`` holds device-resident data for loops that never
leave HBM.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import OTM_DEVICE, OTM_HOST, OtmError

__all__ = [
    "HomoConfig", "Homogenization_H", "OCOptimizer", "MMAOptimizer", "DeviceBuffer", "DensityParams",
    "conv_filter", "OtmError", "load_library", "run_custom", "Batch", "Server", "symmetrize",
    "write_density", "read_density",
]


def load_library():
    return _capi.load()


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceBuffer:
    """A device allocation made through the engine (otm_device_alloc)."""

    def __init__(self, nfloats, device=0):
        lib = _capi.load()
        self.n = int(nfloats)
        self.device = device
        p = ctypes.c_void_p()
        _capi.check(lib.otm_device_alloc(device, 4 * self.n, ctypes.byref(p)), "otm_device_alloc")
        self.ptr = p
        _capi.track(self)

    @classmethod
    def from_numpy(cls, a, device=0):
        a = _f32(a)
        b = cls(a.size, device)
        b.upload(a)
        return b

    def upload(self, a):
        a = _f32(a, self.n)
        _capi.check(_capi.load().otm_memcpy(self.ptr, _ptr(a), 4 * self.n, OTM_DEVICE, OTM_HOST), "otm_memcpy")

    def numpy(self):
        out = np.empty(self.n, np.float32)
        _capi.check(_capi.load().otm_memcpy(_ptr(out), self.ptr, 4 * self.n, OTM_HOST, OTM_DEVICE), "otm_memcpy")
        return out

    def free(self):
        if self.ptr:
            _capi.load().otm_device_free(self.ptr)
            self.ptr = None

    close = free

    def __del__(self):
        if _capi.shutting_down():
            return
        try:
            self.free()
        except Exception:
            pass


def _arg(a, n):
    """(pointer, where, keepalive) for a numpy array or a DeviceBuffer."""
    if isinstance(a, DeviceBuffer):
        if a.n < n:
            raise ValueError("device buffer too small")
        return a.ptr, OTM_DEVICE, a
    arr = _f32(a)
    if arr.size < n:
        raise ValueError(f"expected at least {n} values, got {arr.size}")
    return _ptr(arr), OTM_HOST, arr


class HomoConfig:
    """cfg::HomoConfig (cmdline.h:12-33) with the defaults of
    HomoConfig::init (cmdline.cpp:4-35), plus the multigrid knobs."""

    def __init__(self, reso=128, heatRatio=(1.0, 0.0)):
        r = (reso, reso, reso) if np.isscalar(reso) else tuple(reso)
        self.reso = [int(v) for v in r]
        self.heatRatio = [float(heatRatio[0]), float(heatRatio[1])]
        self.volRatio = 0.3
        self.finthres = 5e-4
        self.filterRadius = 2.0
        self.designStep = 0.05
        self.dampRatio = 0.5
        self.femRelThres = 1e-6
        self.target_tensor = [0.3, 0.2, 0.1, 0.0, 0.0, 0.0]
        self.max_iter = 500
        self.max_cg_iter = 200
        self.pre_sweeps = 2
        self.post_sweeps = 2
        self.min_coarse = 4
        self.coarse_sweeps = 0
        self.fp64_refine = 0
        self.fp16_stencils = 0
        self.exact_smoother = 0
        self.graph_vcycle = 1
        self.fine_tile = -1
        self.device = 0
        self.verbose = 0

    def _c(self):
        c = _capi.OtmConfig()
        _capi.load().otm_config_init(ctypes.byref(c))
        for a in range(3):
            c.reso[a] = self.reso[a]
        c.heat_ratio[0], c.heat_ratio[1] = self.heatRatio
        c.fem_rel_thres = self.femRelThres
        c.max_cg_iter = self.max_cg_iter
        c.pre_sweeps = self.pre_sweeps
        c.post_sweeps = self.post_sweeps
        c.min_coarse = self.min_coarse
        c.coarse_sweeps = self.coarse_sweeps
        c.fp64_refine = self.fp64_refine
        c.fp16_stencils = self.fp16_stencils
        c.exact_smoother = self.exact_smoother
        c.graph_vcycle = self.graph_vcycle
        c.fine_tile = self.fine_tile
        c.device = self.device
        c.verbose = self.verbose
        return c


class DensityParams:
    """Density expression of runCustom (Framework.cu:163): Spline4 r=1.5
    non-periodic filter, then kappa = k1 + (k0-k1) rho^3."""

    def __init__(self, filter_kind=0, filter_radius=1.5, filter_periodic=False):
        self.filter_kind = filter_kind
        self.filter_radius = filter_radius
        self.filter_periodic = filter_periodic

    def _c(self):
        p = _capi.OtmDensityParams()
        _capi.load().otm_density_params_init(ctypes.byref(p))
        p.filter_kind = self.filter_kind
        p.filter_radius = self.filter_radius
        p.filter_periodic = 1 if self.filter_periodic else 0
        return p


class Homogenization_H:
    """homo::Homogenization_H: build / update / heatMatrix / Sensitivity."""

    def __init__(self, config):
        self._ctx = None
        self.build(config)

    def build(self, config):
        lib = _capi.load()
        self.close()
        self.config = config
        ctx = ctypes.c_void_p()
        cfg = config._c()
        _capi.check(lib.otm_create(ctypes.byref(cfg), ctypes.byref(ctx)), "otm_create")
        self._ctx = ctx
        _capi.track(self)
        self.reso = tuple(config.reso)
        self.ne = self.reso[0] * self.reso[1] * self.reso[2]

    def close(self):
        if self._ctx:
            _capi.load().otm_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        if _capi.shutting_down():
            return
        try:
            self.close()
        except Exception:
            pass

    # -- hierarchy -----------------------------------------------------
    def levels(self):
        lib = _capi.load()
        n = ctypes.c_int()
        _capi.check(lib.otm_levels(self._ctx, ctypes.byref(n)), "otm_levels")
        out = []
        for l in range(n.value):
            r = (ctypes.c_int * 3)()
            _capi.check(lib.otm_level_reso(self._ctx, l, r), "otm_level_reso")
            out.append(tuple(r))
        return out

    # -- Homogenization_H interface --------------------------------------
    def update(self, rho=None, pitchT=-1):
        """Homogenization_H::update: rho is the SIMP-mapped conductivity."""
        lib = _capi.load()
        if rho is None:
            _capi.check(lib.otm_update(self._ctx, None, self.reso[0], OTM_HOST), "otm_update")
            return
        pitch = self.reso[0] if pitchT is None or pitchT < 0 else int(pitchT)
        p, where, keep = _arg(rho, pitch * self.reso[1] * self.reso[2])
        _capi.check(lib.otm_update(self._ctx, p, pitch, where), "otm_update")

    def heatMatrix(self):
        C = (ctypes.c_double * 9)()
        _capi.check(_capi.load().otm_heat_matrix(self._ctx, C), "otm_heat_matrix")
        return np.array(C[:], dtype=np.float64).reshape(3, 3)

    def solve_stats(self):
        it = (ctypes.c_int * 3)()
        rr = (ctypes.c_double * 3)()
        _capi.check(_capi.load().otm_solve_stats(self._ctx, it, rr), "otm_solve_stats")
        return list(it), list(rr)

    def fine_tile(self):
        """The fine smoother's tile variant and shape (bx, by, bz, threads),
        or (-1, None) when the fine level uses exact colour sweeps."""
        v = ctypes.c_int()
        _capi.check(_capi.load().otm_fine_tile(self._ctx, ctypes.byref(v)), "otm_fine_tile")
        return v.value, (fine_tile_shape(v.value) if v.value >= 0 else None)

    def autotune(self, warm=False):
        """Times one solve of the current design per candidate (fine tile
        shape x coarse sweeps 1, 2), from a zero start or (warm) from the
        current iterate, and keeps the fastest; returns the candidates as
        dicts and the index of the chosen one."""
        rep = _capi.OtmTuneReport()
        _capi.check(_capi.load().otm_autotune(self._ctx, 1 if warm else 0, ctypes.byref(rep)), "otm_autotune")
        rows = [{"fine_tile": rep.fine_tile[i], "shape": fine_tile_shape(rep.fine_tile[i]),
                 "coarse_sweeps": rep.coarse_sweeps[i], "ms": rep.ms[i], "iters": rep.iters[i], "ok": bool(rep.ok[i])}
                for i in range(rep.n)]
        return rows, rep.chosen

    def Sensitivity(self, dC, sens=None, pitchT=-1, lexiOrder=True):
        """Homogenization_H::Sensitivity with lexiOrder=true. Returns the
        sensitivity array (host) or writes into a DeviceBuffer."""
        if not lexiOrder:
            raise ValueError("only lexicographic output is provided (the reference callers use lexiOrder=true)")
        dCa = (ctypes.c_double * 9)(*np.asarray(dC, dtype=np.float64).reshape(9).tolist())
        pitch = self.reso[0] if pitchT is None or pitchT < 0 else int(pitchT)
        n = pitch * self.reso[1] * self.reso[2]
        if sens is None:
            sens = np.zeros(n, np.float32)
        p, where, keep = _arg(sens, n)
        _capi.check(_capi.load().otm_sensitivity(self._ctx, dCa, p, pitch, where), "otm_sensitivity")
        return keep if where == OTM_HOST else sens

    def element_energy(self):
        out = np.empty(6 * self.ne, np.float32)
        _capi.check(_capi.load().otm_element_energy(self._ctx, _ptr(out), OTM_HOST), "otm_element_energy")
        return out.reshape(6, self.ne)

    def get_chi(self):
        out = np.empty(3 * self.ne, np.float32)
        _capi.check(_capi.load().otm_get_chi(self._ctx, _ptr(out), OTM_HOST), "otm_get_chi")
        return out.reshape(3, self.ne)

    def set_chi(self, chi):
        a = _f32(chi, 3 * self.ne)
        _capi.check(_capi.load().otm_set_chi(self._ctx, _ptr(a), OTM_HOST), "otm_set_chi")

    def reset_displacement(self):
        _capi.check(_capi.load().otm_reset_displacement(self._ctx), "otm_reset_displacement")

    def apply_level(self, level, u):
        n = 1
        for v in self.levels()[level]:
            n *= v
        a = _f32(u, 3 * n)
        q = np.empty(3 * n, np.float32)
        _capi.check(_capi.load().otm_apply_level(self._ctx, level, _ptr(a), _ptr(q), OTM_HOST), "otm_apply_level")
        return q

    def apply_fine_pcg(self, u):
        """The PCG's fine operator apply: (q = K u, [u.q per load case])."""
        a = _f32(u, 3 * self.ne)
        q = np.empty(3 * self.ne, np.float32)
        pq = (ctypes.c_double * 3)()
        _capi.check(_capi.load().otm_apply_fine_pcg(self._ctx, _ptr(a), _ptr(q), pq, OTM_HOST), "otm_apply_fine_pcg")
        return q, list(pq)

    def vcycle(self, r):
        a = _f32(r, 3 * self.ne)
        z = np.empty(3 * self.ne, np.float32)
        _capi.check(_capi.load().otm_vcycle(self._ctx, _ptr(a), _ptr(z), OTM_HOST), "otm_vcycle")
        return z

    def stencil(self, level):
        n = 1
        for v in self.levels()[level]:
            n *= v
        st = np.empty(27 * n, np.float32)
        _capi.check(_capi.load().otm_get_stencil(self._ctx, level, _ptr(st), OTM_HOST), "otm_get_stencil")
        return st.reshape(27, n)

    def _level_nodes(self, level):
        n = 1
        for v in self.levels()[level]:
            n *= v
        return n

    def solve(self, rtol=0.0, warm=True):
        """MG_H::solveEquation for the three load cases (no kappa^H); returns
        the three final relative residuals."""
        rr = (ctypes.c_double * 3)()
        _capi.check(_capi.load().otm_solve(self._ctx, float(rtol), 1 if warm else 0, rr), "otm_solve")
        return list(rr)

    def mg_cycles(self, cycles=1):
        """MG_H::v_cycle iterated as a stationary method; returns the relative
        residual measured at the start of the last cycle."""
        rr = (ctypes.c_double * 3)()
        _capi.check(_capi.load().otm_mg_cycles(self._ctx, int(cycles), rr), "otm_mg_cycles")
        return list(rr)

    def restrict(self, level, r_fine):
        """P^T r onto `level` from three level-(level-1) vectors."""
        nf, nc = self._level_nodes(level - 1), self._level_nodes(level)
        a = _f32(r_fine, 3 * nf)
        out = np.empty(3 * nc, np.float32)
        _capi.check(_capi.load().otm_transfer(self._ctx, level, 0, _ptr(a), _ptr(out), OTM_HOST), "otm_transfer")
        return out

    def prolong_add(self, level, u_coarse, u_fine):
        """u_fine + P u_coarse (returns a new array)."""
        nf, nc = self._level_nodes(level - 1), self._level_nodes(level)
        a = _f32(u_coarse, 3 * nc)
        out = np.array(_f32(u_fine, 3 * nf), copy=True)
        _capi.check(_capi.load().otm_transfer(self._ctx, level, 1, _ptr(a), _ptr(out), OTM_HOST), "otm_transfer")
        return out

    def smooth_level(self, level, f, u, sweeps=1, mode=0):
        """Gauss-Seidel sweeps on a level (mode 0/1 exact in forward/reference
        colour order, 2/3 the brick smoother); returns the new u."""
        n = self._level_nodes(level)
        fa = _f32(f, 3 * n)
        out = np.array(_f32(u, 3 * n), copy=True)
        _capi.check(_capi.load().otm_smooth_level(self._ctx, level, _ptr(fa), _ptr(out), int(sweeps), int(mode),
                                                  OTM_HOST), "otm_smooth_level")
        return out

    def coarsest_solve(self, f):
        n = self._level_nodes(len(self.levels()) - 1)
        fa = _f32(f, 3 * n)
        out = np.empty(3 * n, np.float32)
        _capi.check(_capi.load().otm_coarsest_solve(self._ctx, _ptr(fa), _ptr(out), OTM_HOST), "otm_coarsest_solve")
        return out

    # -- density pipeline (AutoDiff parts of runCustom) ------------------
    def set_design(self, rho):
        p, where, keep = _arg(rho, self.ne)
        _capi.check(_capi.load().otm_set_design(self._ctx, p, where), "otm_set_design")

    def init_design(self, initway="IWP", seed=0, volratio=0.3, n_period=10, sym="NONE", coeffs=None):
        """initDensity (Framework.cu:61-145) on the design; returns the randTri
        coefficients used (random, randcenter, rep_randcenter) or None."""
        lib = _capi.load()
        p = _capi.OtmInitParams()
        lib.otm_init_params_init(ctypes.byref(p))
        iw = _INITWAYS.index(initway) if isinstance(initway, str) else int(initway)
        p.initway, p.seed, p.volratio, p.n_period = iw, int(seed), float(volratio), int(n_period)
        p.sym = _SYMMETRIES.index(sym) if isinstance(sym, str) else int(sym)
        keep = None
        if coeffs is not None:
            keep = _f32(coeffs)
            p.coeffs, p.ncoeffs = keep.ctypes.data, keep.size
        nb = 6 * n_period + 36 * n_period ** 2 + (4 if iw in (1, 5) else 0)
        out = np.empty(nb, np.float32)
        _capi.check(lib.otm_init_design(self._ctx, ctypes.byref(p), _ptr(out)), "otm_init_design")
        return out if iw in (0, 1, 5) else None

    def get_design(self, out=None):
        """Design density (host array, or written into a DeviceBuffer)."""
        if isinstance(out, DeviceBuffer):
            _capi.check(_capi.load().otm_get_design(self._ctx, out.ptr, OTM_DEVICE), "otm_get_design")
            return out
        out = np.empty(self.ne, np.float32)
        _capi.check(_capi.load().otm_get_design(self._ctx, _ptr(out), OTM_HOST), "otm_get_design")
        return out

    def eval(self, params=None):
        params = params or DensityParams()
        p = params._c()
        C = (ctypes.c_double * 9)()
        _capi.check(_capi.load().otm_eval(self._ctx, ctypes.byref(p), C), "otm_eval")
        return np.array(C[:], dtype=np.float64).reshape(3, 3)

    def apply_design(self, params=None):
        """The density pipeline of eval (filter, SIMP, operator rebuild)
        without the solve."""
        params = params or DensityParams()
        p = params._c()
        _capi.check(_capi.load().otm_apply_design(self._ctx, ctypes.byref(p)), "otm_apply_design")

    def backward(self, dC, want=True, out=None):
        """d(sum dC_ij C_ij)/d rho through the SIMP map and the filter:
        returned as a host array, written into a DeviceBuffer `out`, or only
        kept on the device (want=False, for the OC step)."""
        dCa = (ctypes.c_double * 9)(*np.asarray(dC, dtype=np.float64).reshape(9).tolist())
        if not want:
            _capi.check(_capi.load().otm_backward(self._ctx, dCa, None, OTM_HOST), "otm_backward")
            return None
        if isinstance(out, DeviceBuffer):
            _capi.check(_capi.load().otm_backward(self._ctx, dCa, out.ptr, OTM_DEVICE), "otm_backward")
            return out
        out = np.empty(self.ne, np.float32)
        _capi.check(_capi.load().otm_backward(self._ctx, dCa, _ptr(out), OTM_HOST), "otm_backward")
        return out

    def design_stats(self):
        """(mean design density, mean SIMP conductivity) on the device."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _capi.check(_capi.load().otm_design_stats(self._ctx, ctypes.byref(a), ctypes.byref(b)), "otm_design_stats")
        return a.value, b.value

    def synchronize(self):
        _capi.check(_capi.load().otm_synchronize(self._ctx), "otm_synchronize")

    # -- live kernel timing --------------------------------------------
    def profile_enable(self, on=True, classes=None):
        """Time kernel classes with HIP events (all classes by default)."""
        mask = 0
        if on:
            names = _capi.PROF_CLASSES if classes is None else classes
            for nm in names:
                mask |= 1 << _capi.PROF_CLASSES.index(nm)
        _capi.check(_capi.load().otm_profile_enable(self._ctx, mask), "otm_profile_enable")

    def profile_read(self):
        """{class: (launches, total_ms, algorithmic_bytes)} since enabling."""
        lib = _capi.load()
        out = {}
        for k, name in enumerate(_capi.PROF_CLASSES):
            n = ctypes.c_longlong()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            _capi.check(lib.otm_profile_read(self._ctx, k, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)),
                        "otm_profile_read")
            out[name] = (n.value, ms.value, by.value)
        return out


class OCOptimizer:
    """homo::OCOptimizer(int ne, float min_density, float stepLimit,
    float dampExponent) with filterSens and update."""

    def __init__(self, ne, min_density=0.001, stepLimit=0.02, dampExponent=0.5, device=0,
                 bisect_steps=20, bisect_tol=1e-4):
        lib = _capi.load()
        self.ne = int(ne)
        h = ctypes.c_void_p()
        _capi.check(lib.otm_oc_create(self.ne, min_density, stepLimit, dampExponent, device, ctypes.byref(h)),
                    "otm_oc_create")
        self._h = h
        _capi.track(self)
        _capi.check(lib.otm_oc_set_bisection(h, int(bisect_steps), float(bisect_tol)), "otm_oc_set_bisection")
        self.last_steps = 0
        self.last_volume = 0.0

    def close(self):
        if getattr(self, "_h", None):
            _capi.load().otm_oc_destroy(self._h)
            self._h = None

    def __del__(self):
        if _capi.shutting_down():
            return
        try:
            self.close()
        except Exception:
            pass

    def filterSens(self, sens, rho, pitchT, reso, radius=2.0):
        """In place on ``sens`` (numpy float32 array or DeviceBuffer)."""
        reso3 = (ctypes.c_int * 3)(*[int(v) for v in reso])
        n = int(pitchT) * reso3[1] * reso3[2]
        if isinstance(sens, np.ndarray):
            if sens.dtype != np.float32 or not sens.flags.c_contiguous:
                raise ValueError("sens must be a contiguous float32 array (updated in place)")
        ps, ws, ks = _arg(sens, n)
        pr, wr, kr = _arg(rho, n)
        if ws != wr:
            raise ValueError("sens and rho must both be host or both device")
        _capi.check(_capi.load().otm_oc_filter_sens(self._h, ps, pr, int(pitchT), reso3, float(radius), ws),
                    "otm_oc_filter_sens")
        return sens

    def update(self, sens, rho, volratio):
        """In place on ``rho``; returns it."""
        if isinstance(rho, np.ndarray):
            if rho.dtype != np.float32 or not rho.flags.c_contiguous:
                raise ValueError("rho must be a contiguous float32 array (updated in place)")
        ps, ws, ks = _arg(sens, self.ne)
        pr, wr, kr = _arg(rho, self.ne)
        if ws != wr:
            raise ValueError("sens and rho must both be host or both device")
        st = ctypes.c_int()
        vol = ctypes.c_double()
        _capi.check(_capi.load().otm_oc_update(self._h, ps, pr, float(volratio), ws, ctypes.byref(st),
                                               ctypes.byref(vol)), "otm_oc_update")
        self.last_steps, self.last_volume = st.value, vol.value
        return rho

    def step(self, hom, filter_radius, volratio):
        """filterSens + update on the design held by a Homogenization_H."""
        st = ctypes.c_int()
        vol = ctypes.c_double()
        _capi.check(_capi.load().otm_oc_step(hom._ctx, self._h, float(filter_radius), float(volratio),
                                             ctypes.byref(st), ctypes.byref(vol)), "otm_oc_step")
        self.last_steps, self.last_volume = st.value, vol.value


class MMAOptimizer:
    """homo::MMAOptimizer(int m, int n, double a0, double a, double c,
    double d) with setBound and update (mmaOptimizer.h:12-35)."""

    def __init__(self, m, n, a0=1.0, a=0.0, c=1e3, d=1.0, device=0):
        lib = _capi.load()
        self.m, self.n = int(m), int(n)
        h = ctypes.c_void_p()
        _capi.check(lib.otm_mma_create(self.m, self.n, a0, a, c, d, device, ctypes.byref(h)), "otm_mma_create")
        self._h = h
        _capi.track(self)
        self.last_newton = 0

    def close(self):
        if getattr(self, "_h", None):
            _capi.load().otm_mma_destroy(self._h)
            self._h = None

    def __del__(self):
        if _capi.shutting_down():
            return
        try:
            self.close()
        except Exception:
            pass

    def setBound(self, xmin, xmax):
        if np.isscalar(xmin) and np.isscalar(xmax):
            _capi.check(_capi.load().otm_mma_set_bounds(self._h, float(xmin), float(xmax)), "otm_mma_set_bounds")
            return
        a = _f32(xmin, self.n)
        b = _f32(xmax, self.n)
        _capi.check(_capi.load().otm_mma_set_bounds_array(self._h, _ptr(a), _ptr(b), OTM_HOST),
                    "otm_mma_set_bounds_array")

    def set_move(self, move):
        _capi.check(_capi.load().otm_mma_set_move(self._h, float(move)), "otm_mma_set_move")

    MMA_PASSES = ("trial", "bound", "point")

    def profile_enable(self, enable=True):
        """HIP-event timing of the subsolver passes (later trial points of a
        line search, each Newton step's first trial with the step bound,
        start-of-level point)."""
        _capi.check(_capi.load().otm_mma_profile_enable(self._h, 1 if enable else 0), "otm_mma_profile_enable")

    def profile_read(self):
        """{pass: (launches, ms, algorithmic bytes)} since profile_enable."""
        out = {}
        for k, name in enumerate(self.MMA_PASSES):
            nl, ms, by = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
            _capi.check(_capi.load().otm_mma_profile_read(self._h, k, ctypes.byref(nl), ctypes.byref(ms),
                                                          ctypes.byref(by)), "otm_mma_profile_read")
            out[name] = (nl.value, ms.value, by.value)
        return out

    def update(self, itn, x, dfdx, gval, dgdx):
        """In place on x (float32 array or DeviceBuffer); dgdx is a sequence
        of m arrays (all host or all device, like x)."""
        px, wx, kx = _arg(x, self.n)
        pf, wf, kf = _arg(dfdx, self.n)
        if isinstance(x, np.ndarray) and (x.dtype != np.float32 or not x.flags.c_contiguous):
            raise ValueError("x must be a contiguous float32 array (updated in place)")
        keep = []
        ptrs = (ctypes.c_void_p * self.m)()
        for i in range(self.m):
            p, w, k = _arg(dgdx[i], self.n)
            if w != wx:
                raise ValueError("all arrays must be host or all device")
            ptrs[i] = p
            keep.append(k)
        if wf != wx:
            raise ValueError("all arrays must be host or all device")
        g = (ctypes.c_double * self.m)(*[float(v) for v in np.asarray(gval).reshape(-1)[:self.m]])
        newton = ctypes.c_int()
        _capi.check(_capi.load().otm_mma_update(self._h, int(itn), px, pf, g, ptrs, wx, ctypes.byref(newton)),
                    "otm_mma_update")
        self.last_newton = newton.value
        return x


def fine_tile_shape(variant):
    """(bx, by, bz, threads) of a fine smoother tile variant."""
    sh = (ctypes.c_int * 4)()
    _capi.check(_capi.load().otm_fine_tile_shape(int(variant), sh), "otm_fine_tile_shape")
    return tuple(sh)


def conv_filter(x, kind=0, radius=1.5, periodic=False, adjoint=False):
    """Radial renormalised convolution of a (nz, ny, nx) float32 field."""
    a = _f32(x)
    if a.ndim != 3:
        raise ValueError("x must be 3-D (nz, ny, nx)")
    nz, ny, nx = a.shape
    reso = (ctypes.c_int * 3)(nx, ny, nz)
    out = np.empty_like(a)
    _capi.check(_capi.load().otm_conv_filter(reso, int(kind), float(radius), 1 if periodic else 0,
                                             1 if adjoint else 0, _ptr(a), _ptr(out), OTM_HOST),
                "otm_conv_filter")
    return out


_INITWAYS = ["random", "randcenter", "noise", "manual", "interp", "rep_randcenter", "P", "G", "D", "IWP", "example"]


def write_density(path, rho):
    """Grid_H::writeDensity binary format (grid.cpp:330-345): int32 reso[3]
    then float32 values, x fastest; rho has shape (nz, ny, nx)."""
    a = _f32(rho)
    nz, ny, nx = a.shape
    reso = (ctypes.c_int * 3)(nx, ny, nz)
    _capi.check(_capi.load().otm_write_density(str(path).encode(), reso, _ptr(a), OTM_HOST), "otm_write_density")


def read_density(path):
    """Reads the binary density format; returns an array of shape (nz, ny, nx)."""
    lib = _capi.load()
    reso = (ctypes.c_int * 3)()
    _capi.check(lib.otm_read_density(str(path).encode(), reso, None, OTM_HOST), "otm_read_density")
    nx, ny, nz = reso[0], reso[1], reso[2]
    out = np.empty((nz, ny, nx), np.float32)
    _capi.check(lib.otm_read_density(str(path).encode(), reso, _ptr(out), OTM_HOST), "otm_read_density")
    return out


_SYMMETRIES = ["reflect3", "reflect6", "rotate3", "rotate2", "NONE"]


def symmetrize(field, sym):
    """symmetrizeField (Framework.cuh:221-235) on the GPU: the orbit mean
    under the named symmetry group ("reflect3", "reflect6", "rotate3",
    "rotate2"); field has shape (nz, ny, nx). Returns a new array."""
    a = np.array(field, dtype=np.float32, order="C", copy=True)
    nz, ny, nx = a.shape
    reso = (ctypes.c_int * 3)(nx, ny, nz)
    _capi.check(_capi.load().otm_symmetrize(reso, _SYMMETRIES.index(sym), _ptr(a), OTM_HOST), "otm_symmetrize")
    return a


def run_custom(reso, model="oc", max_iter=500, rho0=None, initway="IWP", seed=0,
               target=(0.3, 0.2, 0.1, 0.0, 0.0, 0.0), heat_ratio=(1.0, 0.0), fem_rel_thres=1e-6,
               finthres=5e-4, volratio=0.3, coarse_sweeps=0, sym="NONE", log_prefix=None, logrho=0, logc=0,
               logsens=0, logobj=0):
    """runCustom (Framework.cu:147-244) on the engine: returns the final
    design density (x fastest) and a report dict. log_prefix (a path
    prefix) turns on the iteration log of logIter (Framework.cu:7-59) with
    the given intervals."""
    lib = _capi.load()
    p = _capi.OtmCustomParams()
    lib.otm_custom_params_init(ctypes.byref(p))
    r = (reso, reso, reso) if np.isscalar(reso) else tuple(reso)
    for a in range(3):
        p.base.reso[a] = int(r[a])
    p.base.heat_ratio[0], p.base.heat_ratio[1] = float(heat_ratio[0]), float(heat_ratio[1])
    p.base.fem_rel_thres = fem_rel_thres
    p.base.coarse_sweeps = coarse_sweeps
    p.model = 0 if model == "mma" else 1
    p.initway = _INITWAYS.index(initway)
    p.seed = int(seed)
    p.max_iter = int(max_iter)
    p.finthres = finthres
    p.volratio = volratio
    p.sym = _SYMMETRIES.index(sym)
    for i in range(6):
        p.target[i] = float(target[i])
    prefix = None
    if log_prefix is not None:
        prefix = str(log_prefix).encode()
        p.log.prefix = prefix
        p.log.logrho, p.log.logc, p.log.logsens, p.log.logobj = int(logrho), int(logc), int(logsens), int(logobj)
    n = r[0] * r[1] * r[2]
    out = np.empty(n, np.float32)
    rep = _capi.OtmCustomReport()
    r0 = None
    if rho0 is not None:
        r0a = _f32(rho0, n)
        r0 = _ptr(r0a)
    _capi.check(lib.otm_run_custom(ctypes.byref(p), r0, _ptr(out), ctypes.byref(rep)), "otm_run_custom")
    return out, {"iterations": rep.iterations, "converged": bool(rep.converged), "objective": rep.objective,
                 "volume": rep.volume, "C": np.array(rep.C[:]).reshape(3, 3),
                 "unconverged_solves": rep.unconverged_solves}


class Batch:
    """Batched solves of many cells (otm_batch_*): every launch covers one
    operation for all member cells, so small cells fill the GPU together.
    Members are Homogenization_H objects on the batch's device."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _capi.check(_capi.load().otm_batch_create(int(device), ctypes.byref(h)), "otm_batch_create")
        self._h = h
        _capi.track(self)

    def close(self):
        if getattr(self, "_h", None):
            _capi.load().otm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if _capi.shutting_down():
            return
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ctxs(homs):
        arr = (ctypes.c_void_p * len(homs))()
        for i, h in enumerate(homs):
            arr[i] = h._ctx
        return arr

    def update(self, homs, kappas, pitchT=-1):
        """Homogenization_H::update of every member together (host arrays or
        DeviceBuffers, all of one kind)."""
        n = len(homs)
        ptrs = (ctypes.c_void_p * max(n, 1))()
        keep, wheres = [], set()
        for i, (h, k) in enumerate(zip(homs, kappas)):
            pitch = h.reso[0] if pitchT is None or pitchT < 0 else int(pitchT)
            p, w, kk = _arg(k, pitch * h.reso[1] * h.reso[2])
            ptrs[i] = p
            keep.append(kk)
            wheres.add(w)
        if len(wheres) > 1:
            raise ValueError("conductivities must be all host arrays or all device buffers")
        where = wheres.pop() if wheres else OTM_HOST
        _capi.check(_capi.load().otm_batch_update(self._h, self._ctxs(homs), n, ptrs, int(pitchT), where),
                    "otm_batch_update")

    def heat_matrix(self, homs, status=False):
        """kappa^H of every member, shape (n, 3, 3); with status=True also the
        per-cell status codes (failed cells hold NaN)."""
        n = len(homs)
        C = (ctypes.c_double * (9 * max(n, 1)))()
        st = (ctypes.c_int * max(n, 1))()
        _capi.check(_capi.load().otm_batch_heat_matrix(self._h, self._ctxs(homs), n, C, st if status else None),
                    "otm_batch_heat_matrix")
        out = np.array(C[:9 * n], dtype=np.float64).reshape(n, 3, 3)
        return (out, list(st[:n])) if status else out

    def sensitivity(self, homs, dCs, outs=None):
        """Contracted sensitivities of every member (host arrays, or written
        into the DeviceBuffers of `outs`)."""
        n = len(homs)
        dC = (ctypes.c_double * (9 * max(n, 1)))(*np.asarray(dCs, dtype=np.float64).reshape(-1).tolist())
        ptrs = (ctypes.c_void_p * max(n, 1))()
        keep = []
        where = OTM_DEVICE if outs is not None else OTM_HOST
        for i, h in enumerate(homs):
            if outs is not None:
                ptrs[i] = outs[i].ptr
            else:
                a = np.empty(h.ne, np.float32)
                keep.append(a)
                ptrs[i] = a.ctypes.data
        _capi.check(_capi.load().otm_batch_sensitivity(self._h, self._ctxs(homs), n, dC, ptrs, -1, where),
                    "otm_batch_sensitivity")
        return outs if outs is not None else keep

    def stats(self):
        L, S = ctypes.c_longlong(), ctypes.c_int()
        _capi.check(_capi.load().otm_batch_stats(self._h, ctypes.byref(L), ctypes.byref(S)), "otm_batch_stats")
        return L.value, S.value


class Server:
    """Serving form (otm_server_*): a worker thread admits submitted
    requests into a running batch between PCG steps, within an HBM budget,
    and keeps hierarchies resident for later requests of the same shape."""

    def __init__(self, hbm_budget_gb=64.0, max_active=64, device=0):
        h = ctypes.c_void_p()
        _capi.check(_capi.load().otm_server_create(int(device), float(hbm_budget_gb) * 2 ** 30, int(max_active),
                                                   ctypes.byref(h)), "otm_server_create")
        self._h = h
        _capi.track(self)
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            _capi.load().otm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        if _capi.shutting_down():
            return
        try:
            self.close()
        except Exception:
            pass

    def submit(self, kappa, cfg=None, dC=None, want_sens=True):
        """Queues one homogenization (+ sensitivity) of the SIMP-mapped
        conductivity `kappa` (shape (nz, ny, nx)); returns the request id."""
        k = _f32(kappa)
        req = _capi.OtmRequest()
        _capi.load().otm_request_init(ctypes.byref(req))
        if cfg is None:
            nz, ny, nx = k.shape
            cfg = HomoConfig((nx, ny, nz), (1.0, 1e-3))
        req.cfg = cfg._c()
        req.cfg.device = self.device
        req.kappa = k.ctypes.data
        d = np.zeros(9) if dC is None else np.asarray(dC, dtype=np.float64).reshape(9)
        for i in range(9):
            req.dC[i] = float(d[i])
        req.want_sens = 1 if want_sens else 0
        rid = ctypes.c_longlong()
        _capi.check(_capi.load().otm_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)), "otm_server_submit")
        return rid.value

    def wait(self, rid, ne=None):
        """(kappa^H (3, 3), sensitivity or None, info dict) of request `rid`."""
        res = _capi.OtmResult()
        sens = np.empty(ne, np.float32) if ne else None
        _capi.check(_capi.load().otm_server_wait(self._h, ctypes.c_longlong(rid), ctypes.byref(res),
                                                 _ptr(sens) if sens is not None else None), "otm_server_wait")
        info = {"iters": list(res.iters), "relres": list(res.relres), "queue_ms": res.queue_ms,
                "latency_ms": res.latency_ms}
        return np.array(res.C[:]).reshape(3, 3), sens, info

    def stats(self):
        st = _capi.OtmServerCounters()
        _capi.check(_capi.load().otm_server_stats(self._h, ctypes.byref(st)), "otm_server_stats")
        return {f[0]: getattr(st, f[0]) for f in st._fields_}
