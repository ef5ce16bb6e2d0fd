"""Loader for the golden fixtures written by tests/golden/make_golden.py."""

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "torchcvdv_golden.npz")


def load():
    z = np.load(GOLDEN, allow_pickle=False)
    cases = json.loads(bytes(z["cases_json"]).decode())
    return z, cases


def c(v):
    if isinstance(v, list) and len(v) == 2 and all(isinstance(u, (int, float)) for u in v):
        return complex(v[0], v[1])
    return v


def build_sep(SepCls, regs, init, **kw):
    sep = SepCls(regs, **kw)
    for item in init:
        meth, reg, *args = item
        if meth == "setCoherent":
            sep.setCoherent(reg, c(args[0]))
        elif meth == "setFocks":
            sep.setFocks(reg, [c(a) for a in args[0]])
        elif meth == "setCoeffs":
            sep.setCoeffs(reg, [c(a) for a in args[0]])
        elif meth == "setCat":
            sep.setCat(reg, [(c(a), c(b)) for a, b in args[0]])
        else:
            getattr(sep, meth)(reg, *args)
    return sep


def apply_ops(sim, ops):
    for op in ops:
        meth, *args = op
        getattr(sim, meth)(*[c(a) for a in args])
