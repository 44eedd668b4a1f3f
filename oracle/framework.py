"""runCustom's control logic (Framework.cu:147-244, Framework.cuh:312-411)
restated for the oracle loops: the convergence test, the adaptive volume
bound of the OC branch and the tensor objective.

TEST INFRASTRUCTURE ONLY (the checker for the engine's otm_run_custom).
The FP32 state of the reference's VolumeGovernor is kept with numpy
float32 scalars so that its thresholds see the same values.
"""
import numpy as np

f32 = np.float32

# objective pairs in the order of Framework.cu:168-170: (00, 11, 22, 01, 21, 02)
PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (2, 1), (0, 2)]


def tensor_objective(C, target):
    """sum_q (C_q - t_q)^2 - 1e-4 and its gradient with respect to C (3x3)."""
    f = -1e-4
    dC = np.zeros((3, 3))
    for q, (i, j) in enumerate(PAIRS):
        e = C[i, j] - target[q]
        f += e * e
        dC[i, j] += 2.0 * e
    return f, dC


class ConvergeChecker:
    """Framework.cuh:312-343: after iteration 11, converged when the last
    three relative changes of the objective are all <= thres."""

    def __init__(self, thres=1e-3):
        self.thres = thres
        self.hist = {}

    def is_converge(self, it, obj):
        self.hist[it] = float(obj)
        if it <= 11:
            return False
        for i in range(3):
            a, b = self.hist[it - i], self.hist[it - i - 1]
            if abs(a - b) / abs(b) > self.thres:
                return False
        return True


class VolumeGovernor:
    """Framework.cuh:345-411 (float members, double literals)."""

    def __init__(self):
        self.lowBound = f32(0)
        self.volume_bound = f32(1.0)
        self.decrease = f32(0)
        self.decrease_factor = f32(1.0)
        self.val_last = f32(1)
        self.vibrate = f32(0)
        self.vibrate_count = 0
        self.count = 0

    def _shrink(self):
        self.decrease = f32(self.volume_bound - self.lowBound)
        self.volume_bound = f32(self.volume_bound - f32(self.decrease_factor * self.decrease))
        if float(self.decrease_factor) > 0.05:
            self.decrease_factor = f32(float(self.decrease_factor) * 0.8)

    def _expand(self):
        self.volume_bound = f32(float(self.volume_bound) + 0.3 * float(self.decrease) * float(self.decrease_factor))

    def get_volume_bound(self):
        return self.volume_bound

    def get_current_decrease(self):
        return f32(self.decrease_factor * self.decrease)

    def volume_check(self, value, lowBound, volfrac, itn):
        value, volfrac = f32(value), f32(volfrac)
        self.lowBound = f32(lowBound)
        if float(value) + 0.0001 < 1e-4 or itn == 100:
            self._shrink()
            if float(self.decrease_factor) <= 0.05:
                return 1
            self.count = 0
        reach = abs(float(f32(self.volume_bound - volfrac))) < 1e-2
        diff = f32(self.val_last - value)
        if float(f32(self.vibrate * diff)) < 0:
            self.vibrate_count += 1
        else:
            self.vibrate_count = 0
        if self.vibrate_count >= 8:
            if reach:
                self._expand()
                self.vibrate_count = 0
            else:
                self._shrink()
        self.vibrate = diff
        if float(diff) < 1e-4 and float(value) > 1e-4 and reach:
            self.count += 1
        else:
            self.count = 0
        if self.count >= 5:
            self.volume_bound = f32(float(self.volume_bound) + 0.3 * float(self.decrease) *
                                    float(self.decrease_factor))
            self.count = 0
        self.val_last = value
        return 0
