"""Scalar objective expressions with reverse-mode gradients.

The reference composes its design objectives from an expression-template
AutoDiff library (AutoDiff.h:400-850): scalar nodes with value and diff,
operators + - * /, pow, log, exp, sgm (logistic with slope and centre), abs,
and full reductions of tensor expressions (TensorExpression.h:1187-1402:
sum, p-norm, log-sum-exp). Here an objective over the homogenized tensor
entries (and field reductions of the design) is built from `Var` nodes;
`backward()` gives d objective / d C as the 3x3 matrix that
`Homogenization_H.backward` / `otm_backward` takes, and the design gradient
of any field reductions used.

    C = hom.eval()
    c = heat_tensor(C)
    obj = ((c[0][0] - 0.3) ** 2 + (c[1][1] - 0.2) ** 2).log() + 0.1 * c[0][1].abs()
    obj.backward()
    dC = tensor_grad(c)
"""
import math

import numpy as np

__all__ = ["Var", "heat_tensor", "tensor_grad", "field_reduce"]


class Var:
    """A node of a scalar expression: value, accumulated gradient, and the
    local derivatives with respect to its operands."""

    __slots__ = ("value", "grad", "_parents", "_on_grad")

    def __init__(self, value, parents=(), on_grad=None):
        self.value = float(value)
        self.grad = 0.0
        self._parents = parents  # ((node, d self / d node), ...)
        self._on_grad = on_grad  # leaf hook: called with the final gradient

    # -- construction ------------------------------------------------------
    @staticmethod
    def _lift(x):
        return x if isinstance(x, Var) else Var(x)

    def __add__(self, o):
        o = Var._lift(o)
        return Var(self.value + o.value, ((self, 1.0), (o, 1.0)))

    __radd__ = __add__

    def __sub__(self, o):
        o = Var._lift(o)
        return Var(self.value - o.value, ((self, 1.0), (o, -1.0)))

    def __rsub__(self, o):
        return Var._lift(o) - self

    def __neg__(self):
        return Var(-self.value, ((self, -1.0),))

    def __mul__(self, o):
        o = Var._lift(o)
        return Var(self.value * o.value, ((self, o.value), (o, self.value)))

    __rmul__ = __mul__

    def __truediv__(self, o):
        # div_exp_t (AutoDiff.h:717-727)
        o = Var._lift(o)
        return Var(self.value / o.value, ((self, 1.0 / o.value), (o, -self.value / (o.value * o.value))))

    def __rtruediv__(self, o):
        return Var._lift(o) / self

    def pow(self, e):
        # pow_exp_t (AutoDiff.h:738-748): scalar exponent
        e = float(e)
        return Var(self.value ** e, ((self, e * self.value ** (e - 1.0)),))

    __pow__ = pow

    def log(self):
        # ln_exp_t (AutoDiff.h:782-793)
        return Var(math.log(self.value), ((self, 1.0 / self.value),))

    def exp(self):
        # exp_exp_t (AutoDiff.h:771-780)
        v = math.exp(self.value)
        return Var(v, ((self, v),))

    def sgm(self, s=1.0, c=0.0):
        # sgm_exp_t (AutoDiff.h:833-849): 1 / (1 + exp(-s (x - c)))
        v = 1.0 / (1.0 + math.exp(-s * (self.value - c)))
        return Var(v, ((self, s * v * (1.0 - v)),))

    def abs(self):
        # abs_exp_t (AutoDiff.h:751-768): derivative 0 at 0
        d = 1.0 if self.value > 0 else (-1.0 if self.value < 0 else 0.0)
        return Var(abs(self.value), ((self, d),))

    # -- gradients -----------------------------------------------------------
    def backward(self, seed=1.0):
        """Reverse sweep: every node reachable from this one gets
        d self / d node in .grad (leaves call their hook once)."""
        order, seen = [], set()
        stack = [(self, False)]
        while stack:
            node, done = stack.pop()
            if done:
                order.append(node)
                continue
            if id(node) in seen:
                continue
            seen.add(id(node))
            stack.append((node, True))
            for p, _ in node._parents:
                if id(p) not in seen:
                    stack.append((p, False))
        for node in order:
            node.grad = 0.0
        self.grad = float(seed)
        for node in reversed(order):  # operands after every consumer
            for p, d in node._parents:
                p.grad += node.grad * d
        for node in order:
            if node._on_grad is not None:
                node._on_grad(node.grad)
        return self.value


def heat_tensor(C):
    """The 3x3 homogenized tensor as leaf variables (heat_tensor_t's entries,
    homoExpression.h:68-150); C[i][j] and C[j][i] are distinct leaves, as the
    reference differentiates each entry."""
    C = np.asarray(C, dtype=np.float64).reshape(3, 3)
    return [[Var(C[i, j]) for j in range(3)] for i in range(3)]


def tensor_grad(c):
    """d objective / d C after backward(): the dC argument of otm_backward."""
    return np.array([[c[i][j].grad for j in range(3)] for i in range(3)], dtype=np.float64)


def field_reduce(x, kind="sum", p=2.0, grad_out=None):
    """A full reduction of a field (sum, pnorm, logsumexp) computed on the
    GPU (otm_field_reduce) as a leaf: after backward(), grad_out (a float32
    array of x's size, host) holds d objective / d x."""
    import ctypes
    from . import _capi, _f32, _ptr
    kinds = {"sum": 0, "pnorm": 1, "logsumexp": 2}
    a = _f32(x).reshape(-1)
    val = ctypes.c_double()
    _capi.check(_capi.load().otm_field_reduce(_ptr(a), a.size, kinds[kind], float(p), 0, ctypes.byref(val), None,
                                              0.0), "otm_field_reduce")

    def hook(g):
        if grad_out is None:
            return
        tmp = np.empty_like(a)
        v = ctypes.c_double()
        _capi.check(_capi.load().otm_field_reduce(_ptr(a), a.size, kinds[kind], float(p), 0, ctypes.byref(v),
                                                  _ptr(tmp), float(g)), "otm_field_reduce")
        grad_out.reshape(-1)[:] += tmp

    return Var(val.value, on_grad=hook)
