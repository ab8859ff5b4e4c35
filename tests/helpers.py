"""Shared helpers: build matching oracle / HIP states from the same inputs."""
import numpy as np

FIELDS_EXACT = ["mask", "valid_x", "valid_y", "ub_x", "ub_y"]


def oracle_state(O, g_args, raw, p_kw=None, kfs=0, C=0.5, gap=0):
    g = O.Grid(*g_args)
    p = O.ModelParameters(**(p_kw or {}))
    m = O.Model(kfs=kfs, gap_scheme=gap, sliding_C=C)
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    return g, p, m, s


def hip_state(H, g_args, raw, p_kw=None, C=0.5):
    g = H.Grid(*g_args)
    p = H.ModelParameters(**(p_kw or {}))
    s = H.State(g, p)
    H.set_initial_conditions(s, g, p, H.RegularizedCoulombSlidingLaw(C), **raw)
    return g, p, s


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)


def field_errors(hs, os_, names):
    out = {}
    for n in names:
        out[n] = rel_err(hs.get(n), getattr(os_, n))
    return out
