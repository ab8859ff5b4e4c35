"""Shared helpers: build matching oracle / HIP states from the same inputs."""
import numpy as np

FIELDS_EXACT = ["mask", "valid_x", "valid_y", "ub_x", "ub_y"]


def oracle_state(O, g_args, raw, p_kw=None, kfs=0, C=0.5, gap=0, model=None):
    """``model(g)``, when given, builds the oracle Model (the field sliding laws need the grid)."""
    g = O.Grid(*g_args)
    p = O.ModelParameters(**(p_kw or {}))
    m = model(g) if model is not None else O.Model(kfs=kfs, gap_scheme=gap, sliding_C=C)
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    return g, p, m, s


def hip_state(H, g_args, raw, p_kw=None, C=0.5, law=None):
    """``law(g)``, when given, builds the sliding law (the field laws need the grid)."""
    g = H.Grid(*g_args)
    p = H.ModelParameters(**(p_kw or {}))
    s = H.State(g, p)
    sl = law(g) if law is not None else H.RegularizedCoulombSlidingLaw(C)
    H.set_initial_conditions(s, g, p, sl, **raw)
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


def true_residuals(O, hs, h_prev):
    """The FP64 residual of the head a solve returned, recomputed on the host
    with the oracle's five-array operator on the library's own assembly:
    (||b - A h|| / ||b||, ||b - A h|| / ||b - A h_prev||)."""
    sysd = hs.download_system()
    co = [sysd[k] for k in ("aP", "aE", "aW", "aN", "aS")]
    b = sysd["rhs"]
    r = b - O.matvec(co, hs.get("h"))
    r0 = b - O.matvec(co, h_prev)
    nr = float(np.linalg.norm(r))
    return nr / float(np.linalg.norm(b)), nr / float(np.linalg.norm(r0))
