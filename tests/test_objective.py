"""Objective expressions (opentm_amd.objective, include/opentm/objective.hpp)
against central differences of the same composite expression, which uses
every operation of the reference's scalar AutoDiff (AutoDiff.h:717-849):
+ - * /, pow, log, exp, sgm, abs."""
import math
import os
import subprocess

import numpy as np
import pytest

from opentm_amd.objective import Var, heat_tensor, tensor_grad

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def py_objective(c):
    f = (c[0][0] - 0.3) ** 2 + (c[1][1] - 0.2) ** 2 + (c[2][2] - 0.1) ** 2
    f = f.log() + 0.1 * c[0][1].abs() + c[1][2].sgm(4.0, 0.05) * c[2][0].exp() / (c[0][0] + 2.0)
    return f - 1e-4


def plain_objective(C):
    f = (C[0, 0] - 0.3) ** 2 + (C[1, 1] - 0.2) ** 2 + (C[2, 2] - 0.1) ** 2
    sg = 1.0 / (1.0 + math.exp(-4.0 * (C[1, 2] - 0.05)))
    return math.log(f) + 0.1 * abs(C[0, 1]) + sg * math.exp(C[2, 0]) / (C[0, 0] + 2.0) - 1e-4


def central(C, h=1e-6):
    g = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            Cp, Cm = C.copy(), C.copy()
            Cp[i, j] += h
            Cm[i, j] -= h
            g[i, j] = (plain_objective(Cp) - plain_objective(Cm)) / (2 * h)
    return g


POINTS = [np.array([[0.9, 0.05, 0.02], [-0.03, 0.5, 0.11], [0.2, -0.01, 0.35]]),
          np.array([[0.25, -0.2, 0.0], [0.0, 0.15, -0.3], [-0.4, 0.7, 0.05]])]


@pytest.mark.parametrize("C", POINTS)
def test_python_objective_gradient(C):
    c = heat_tensor(C)
    obj = py_objective(c)
    v = obj.backward()
    assert v == pytest.approx(plain_objective(C), rel=1e-14)
    assert np.allclose(tensor_grad(c), central(C), rtol=1e-6, atol=1e-8)


def test_shared_subexpression_accumulates():
    x = Var(1.5)
    y = x * x + x.exp() * x   # x appears four times
    y.backward()
    assert x.grad == pytest.approx(2 * 1.5 + math.exp(1.5) * (1 + 1.5), rel=1e-14)


def test_abs_at_zero_has_zero_derivative():
    x = Var(0.0)
    x.abs().backward()
    assert x.grad == 0.0


@pytest.mark.parametrize("C", POINTS)
def test_cxx_objective_header(tmp_path, C):
    exe = tmp_path / "objective_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cxx", "objective_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)] + [repr(float(v)) for v in C.reshape(-1)], capture_output=True, text=True,
                         check=True).stdout.split()
    vals = np.array([float(v) for v in out])
    assert vals[0] == pytest.approx(plain_objective(C), rel=1e-14)
    assert np.allclose(vals[1:].reshape(3, 3), central(C), rtol=1e-6, atol=1e-8)
