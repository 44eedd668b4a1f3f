"""Precision-form design loop: MMA with three constraints on a cell solved
with FP64 iterative refinement and FP16 coarse stencils.

Problem (reference tensor indices, see the load-case convention of
homogenization.cu:180-252): maximise C[0][0] subject to C[1][1] <= c_max,
C[2][2] <= c_max and mean density <= vol, over the runCustom density
expression (Spline4 r=1.5 filter, cubic SIMP, Framework.cu:163). The MMA
subproblem is the reference's (mmaOpt.cu:834-1019) with m = 3; objective
and constraints are scaled to O(1):

    f0 = -C00,  g1 = C11/c_max - 1,  g2 = C22/c_max - 1,  g3 = mean(rho)/vol - 1.

Everything stays on the device between iterations: the design, the four
gradient fields and the MMA state; per iteration the host sees the 3x3
tensor and the volume.
"""
import time

import numpy as np

from . import DeviceBuffer, HomoConfig, Homogenization_H, MMAOptimizer

E00 = np.diag([1.0, 0.0, 0.0])
E11 = np.diag([0.0, 1.0, 0.0])
E22 = np.diag([0.0, 0.0, 1.0])

MMA_C = 1e3       # MMAOptimizer(m, n, a0=1, a=0, c, d=1)
MMA_MOVE = 0.2
RHO_MIN = 1e-3


def objective(C, vol_mean, c_max=0.05, vol=0.25):
    """(f0, [g1, g2, g3]) of the scaled problem."""
    return -C[0, 0], np.array([C[1, 1] / c_max - 1.0, C[2, 2] / c_max - 1.0, vol_mean / vol - 1.0])


class PrecisionLoop:
    """One cell and its MMA optimiser; `step()` runs one design iteration."""

    def __init__(self, reso, rho0, k1=1e-3, rtol=1e-8, refine=True, fp16=True, c_max=0.05, vol=0.25, device=0,
                 graph=True):
        cfg = HomoConfig(reso, (1.0, k1))
        cfg.femRelThres = rtol
        cfg.graph_vcycle = 1 if graph else 0
        cfg.fp64_refine = 1 if refine else 0
        cfg.fp16_stencils = 1 if fp16 else 0
        cfg.device = device
        self.hom = Homogenization_H(cfg)
        self.ne = self.hom.ne
        self.c_max, self.vol = c_max, vol
        self.hom.set_design(np.asarray(rho0, np.float32).reshape(-1))
        self.mma = MMAOptimizer(3, self.ne, 1.0, 0.0, MMA_C, 1.0, device=device)
        self.mma.setBound(RHO_MIN, 1.0)
        self.mma.set_move(MMA_MOVE)
        self.x = DeviceBuffer(self.ne, device)
        self.df = DeviceBuffer(self.ne, device)
        self.dg = [DeviceBuffer(self.ne, device) for _ in range(3)]
        self.dg[2].upload(np.full(self.ne, 1.0 / (vol * self.ne), np.float32))
        self.itn = 0
        self.history = []

    def step(self):
        t0 = time.perf_counter()
        C = self.hom.eval()
        its, relres = self.hom.solve_stats()
        self.hom.get_design(self.x)
        vol_mean = self.hom.design_stats()[0]
        f0, g = objective(C, vol_mean, self.c_max, self.vol)
        self.hom.backward(-E00, out=self.df)
        self.hom.backward(E11 / self.c_max, out=self.dg[0])
        self.hom.backward(E22 / self.c_max, out=self.dg[1])
        self.mma.update(self.itn, self.x, self.df, g, self.dg)
        self.hom.set_design(self.x)
        self.hom.synchronize()
        dt = time.perf_counter() - t0
        self.history.append({"itn": self.itn, "C": C, "f0": f0, "g": g, "pcg_iters": its, "relres": relres,
                             "newton": self.mma.last_newton, "seconds": dt})
        self.itn += 1
        return self.history[-1]

    def design(self):
        return self.hom.get_design()
