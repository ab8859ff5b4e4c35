"""Writes tests/golden/config1_64x32_48steps.npz: the oracle's trajectory of
the 64x32 configuration (Cholesky solves, Picard 500 / 1e-6, 48 steps of
3600 s, implicit gap, arithmetic K_face, ConstantMeltInput 1e-8 m/s,
regularised-Coulomb C = 0.5; synthetic glacier seed 1, bed amplitude 10 m).

Stored: the inputs' checksum, the Picard (converged, last_iter) of every
step, h and b after steps 1, 12 and 48, and the face masks.  The reference
itself cannot run here (no Julia), so this is the oracle's output, kept so
that a change of the oracle or of the synthetic inputs is caught.

    python tests/golden/make_config1_fixture.py
"""
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "shakti.jl_amd"))

OUT = pathlib.Path(__file__).resolve().parent / "config1_64x32_48steps.npz"
KEEP = (1, 12, 48)


def trajectory():
    import oracle as O
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(64, 32, bed_amp=10.0, seed=1)
    g = O.Grid(gg.nx, gg.ny, gg.lx, gg.ly)
    p = O.ModelParameters()
    m = O.Model(kfs=0, gap_scheme=0, sliding_C=0.5)
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    sv = O.Solver(g, "cholesky")
    out = {"zb": raw["zb"], "valid_x": s.valid_x.copy(), "valid_y": s.valid_y.copy()}
    picard = []
    for t in range(1, 49):
        conv, last, _ = O.step(sv, s, g, p, m, 3600.0)
        picard.append((int(conv), int(last)))
        if t in KEEP:
            out[f"h_{t}"] = s.h.copy()
            out[f"b_{t}"] = s.b.copy()
    out["picard"] = np.array(picard, dtype=np.int64)
    return out


if __name__ == "__main__":
    import oracle as O
    O.build()
    np.savez_compressed(OUT, **trajectory())
    print(f"wrote {OUT}")
