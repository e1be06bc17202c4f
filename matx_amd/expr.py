"""Fused elementwise expressions: MatX's ``(lhs = expr).run(exec)``.

An expression is built from device tensors, scalars, operators
(``+ - * /``, comparisons) and the MatX functions the radar path uses
(``conj``, ``abs2``, ``abs``, ``sqrt``, ``pow``, ``hamming``, ``ones``,
``select``), then ``run(out, expr)`` lowers it to a postfix program and
evaluates it in one kernel through ``mxr_ew_eval`` (include/mxr.h).  Inputs
broadcast like MatX ``Clone``: a tensor of lower rank is aligned to the
output's trailing dimensions and gets zero strides elsewhere.

References (MatX): operator tables include/matx/operators/binary_operators.h:
585-599 and unary_operators.h:551-584; hamming include/matx/generators/
hamming.h:96-99; set/run include/matx/operators/set.h and base_operator.h:
180-285.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib as L

OP = {"load": 1, "const": 2, "hamming": 3, "index": 4, "+": 10, "-": 11, "*": 12, "/": 13,
      "pow": 14, ">": 15, "<": 16, ">=": 17, "<=": 18, "==": 19, "max": 20, "min": 21,
      "conj": 30, "abs2": 31, "abs": 32, "sqrt": 33, "neg": 34, "real": 35, "imag": 36,
      "exp": 37, "log": 38, "cos": 39, "sin": 40, "select": 50}


class mxr_ew_instr(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("a", ctypes.c_int32), ("b", ctypes.c_int32)]


class Expr:
    def _bin(self, o, op, rev=False):
        o = wrap(o)
        return Node(op, [o, self] if rev else [self, o])

    def __add__(self, o): return self._bin(o, "+")
    def __radd__(self, o): return self._bin(o, "+", True)
    def __sub__(self, o): return self._bin(o, "-")
    def __rsub__(self, o): return self._bin(o, "-", True)
    def __mul__(self, o): return self._bin(o, "*")
    def __rmul__(self, o): return self._bin(o, "*", True)
    def __truediv__(self, o): return self._bin(o, "/")
    def __rtruediv__(self, o): return self._bin(o, "/", True)
    def __gt__(self, o): return self._bin(o, ">")
    def __lt__(self, o): return self._bin(o, "<")
    def __ge__(self, o): return self._bin(o, ">=")
    def __le__(self, o): return self._bin(o, "<=")
    def __neg__(self): return Node("neg", [self])


class Ref(Expr):
    def __init__(self, t):
        self.t = t


class Const(Expr):
    def __init__(self, v):
        self.v = complex(v)


class Gen(Expr):
    def __init__(self, op, dim, n=0):
        self.op, self.dim, self.n = op, dim, n


class Node(Expr):
    def __init__(self, op, args):
        self.op, self.args = op, args


def wrap(x) -> Expr:
    if isinstance(x, Expr):
        return x
    if isinstance(x, torch.Tensor):
        return Ref(x)
    return Const(x)


def conj(x): return Node("conj", [wrap(x)])
def abs2(x): return Node("abs2", [wrap(x)])
def abs(x): return Node("abs", [wrap(x)])  # noqa: A001
def sqrt(x): return Node("sqrt", [wrap(x)])
def pow(x, y): return Node("pow", [wrap(x), wrap(y)])  # noqa: A001
def exp(x): return Node("exp", [wrap(x)])
def real(x): return Node("real", [wrap(x)])
def imag(x): return Node("imag", [wrap(x)])
def select(c, x, y): return Node("select", [wrap(x), wrap(y), wrap(c)])


def hamming(dim: int, n: int) -> Expr:
    """``hamming<dim>`` generator of length n along output dimension dim."""
    return Gen("hamming", dim, n)


def index(dim: int) -> Expr:
    return Gen("index", dim)


def ones() -> Expr:
    return Const(1.0)


def _broadcast_view(t: torch.Tensor, shape):
    """Align t to the trailing dims of shape; zero strides where t has
    extent 1 or no dimension (MatX Clone)."""
    r = len(shape)
    if t.dim() > r:
        raise L.MatxError(3, f"input rank {t.dim()} exceeds output rank {r}")
    st = [0] * (r - t.dim()) + list(t.stride())
    sh = [1] * (r - t.dim()) + list(t.shape)
    for d in range(r):
        if sh[d] != shape[d]:
            if sh[d] != 1:
                raise L.MatxError(4, f"cannot broadcast extent {sh[d]} to {shape[d]}")
            st[d] = 0
    return t.as_strided(shape, st)


def lower(e: Expr, shape):
    prog, ins, consts = [], [], []

    def emit(x):
        if isinstance(x, Ref):
            for k, t in enumerate(ins):
                if t is x.t:
                    prog.append((OP["load"], k, 0))
                    return
            ins.append(x.t)
            prog.append((OP["load"], len(ins) - 1, 0))
        elif isinstance(x, Const):
            consts.append(x.v)
            prog.append((OP["const"], len(consts) - 1, 0))
        elif isinstance(x, Gen):
            d = x.dim % len(shape)
            prog.append((OP[x.op], d, x.n))
        else:
            for a in x.args:
                emit(a)
            prog.append((OP[x.op], 0, 0))

    emit(e)
    views = [_broadcast_view(t, shape) for t in ins]
    return prog, views, consts


def run(out: torch.Tensor, e, stream=None) -> torch.Tensor:
    """``(out = e).run(exec)``: evaluate e into the view out, one kernel."""
    e = wrap(e)
    prog, views, consts = lower(e, tuple(out.shape))
    descs = [L.desc(v) for v in views]
    arr = (ctypes.POINTER(L.mxr_tensor) * max(1, len(descs)))(*[ctypes.pointer(d) for d in descs])
    instrs = (mxr_ew_instr * len(prog))(*[mxr_ew_instr(*p) for p in prog])
    cvals = (ctypes.c_float * max(2, 2 * len(consts)))(
        *[v for c in consts for v in (c.real, c.imag)])
    L.check(L.lib().mxr_ew_eval(ctypes.byref(L.desc(out)), arr, len(descs), instrs, len(prog),
                                cvals, len(consts), L.stream_handle(stream)))
    return out


def sum(x: torch.Tensor, out=None, stream=None) -> torch.Tensor:  # noqa: A001
    """``sum(x)`` to a rank-0 tensor (include/matx/operators/sum.h)."""
    if out is None:
        out = torch.empty((), dtype=x.dtype, device=x.device)
    L.check(L.lib().mxr_sum(ctypes.byref(L.desc(out)), ctypes.byref(L.desc(x)),
                            L.stream_handle(stream)))
    return out
