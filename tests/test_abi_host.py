"""C-ABI library and host-logic checks that need no GPU."""
import ctypes as C
import pathlib
import re

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _declared_symbols(header="shakti_hip.h"):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(hip):
    from shakti_hip import _lib as L
    lib = C.CDLL(str(L.LIB_PATH))
    declared = _declared_symbols()
    assert len(declared) >= 25
    learned = _declared_symbols("shakti_learned.h")
    assert len(learned) >= 18
    headers = sorted(p.name for p in (ROOT / "include").glob("*.h"))
    assert headers == ["shakti_hip.h", "shakti_learned.h"]
    missing = [s for s in declared + learned if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(L.EXPORTS) == declared
    assert sorted(L.EXPORTS_LEARNED) == learned


def test_unet_parameter_layout(hip):
    """Parameter count and per-layer offsets of the 16/32/64/128 U-Net (no device needed)."""
    from shakti_hip import _lib as L
    lib = L.load()
    chans = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (192, 64), (64, 64),
             (96, 32), (32, 32), (48, 16), (16, 16)]
    off = 0
    for l, (ci, co) in enumerate(chans):
        o, a, b = C.c_int64(), C.c_int32(), C.c_int32()
        assert lib.shk_unet_layer_info(l, C.byref(o), C.byref(a), C.byref(b)) == 0
        assert (o.value, a.value, b.value) == (off, ci, co)
        off += co * 9 * ci + co
    o, a, b = C.c_int64(), C.c_int32(), C.c_int32()
    assert lib.shk_unet_layer_info(14, C.byref(o), C.byref(a), C.byref(b)) == 0
    assert (o.value, a.value, b.value) == (off, 16, 1)
    assert lib.shk_unet_num_params() == off + 17
    assert lib.shk_unet_layer_info(15, None, None, None) == -1


def test_abi_version_and_struct_layouts(hip, tmp_path):
    """The ctypes structures match the header's: every field's offset and each
    struct's size as gcc lays them out from include/shakti_hip.h."""
    import subprocess
    from shakti_hip import _lib as L
    assert L.load().shk_abi_version() == L.ABI_VERSION
    structs = {"shk_grid": L.ShkGrid, "shk_params": L.ShkParams, "shk_model": L.ShkModel,
               "shk_solver_cfg": L.ShkSolverCfg, "shk_solve_stats": L.ShkSolveStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "shakti_hip.h"', "int main(void) {",
             '  printf("ABI %d\\n", SHK_ABI_VERSION);']
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    assert out[0] == f"ABI {L.ABI_VERSION}"
    got = {(l.split()[0], l.split()[1]): int(l.split()[2]) for l in out[1:] if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_null_and_bad_arguments_are_errors(hip):
    from shakti_hip import _lib as L
    lib = L.load()
    assert lib.shk_destroy(None) == -1
    out = C.c_void_p()
    assert lib.shk_create(None, None, None, None, C.byref(out)) == -1
    g = L.ShkGrid(1, 5, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)
    p = L.ShkParams()
    m = L.ShkModel()
    cfg = L.ShkSolverCfg(0, 4, 15, 0, 0, 8, 0.0, 1e-8, 1e-3)
    assert lib.shk_create(C.byref(g), C.byref(p), C.byref(m), C.byref(cfg), C.byref(out)) == -1


def test_reference_constructor_errors(hip):
    """ErrorException contract of the reference constructors
    (src/linear_solver.jl:804-834, src/preconditioner.jl:181-182,
    src/simulation.jl:113-145), raised before any device call."""
    H = hip
    g = H.Grid(6, 6, 1e3, 1e3)
    with pytest.raises(H.ShaktiHIPError):
        H.CGIterativeSolver(g, H.SparseAssembledLinearSystem)
    with pytest.raises(H.ShaktiHIPError):
        H.CGIterativeSolver(g, H.HIPMatrixFreeLinearSystem, amg=True)
    with pytest.raises(H.ShaktiHIPError):
        H.CGIterativeSolver(g, H.HIPMatrixFreeLinearSystem, chebyshev_degree=0)
    with pytest.raises(H.ShaktiHIPError):
        H.CGIterativeSolver(g, H.HIPMatrixFreeLinearSystem, chebyshev_nsteps_estimate=1)
    ls = H.CGIterativeSolver(g, H.HIPMatrixFreeLinearSystem, chebyshev_degree=4)
    assert ls.precond.lambda_min == 1.0 and ls.precond.lambda_max == 1.0
    ps = H.PicardSolver(500, 1e-6, ls, g)
    assert ps.converged is False and ps.last_iter == 0


def test_seasonal_melt_input_host_value(hip, oracle):
    H = hip
    mi = H.SeasonalMeltInput()
    for t in [0.0, 0.45 * 365 * 86400, 0.69 * 365 * 86400, 1.5 * 365 * 86400]:
        assert mi.value(t) == pytest.approx(oracle.seasonal_ieb(t), rel=1e-14)


def test_synthetic_inputs_deterministic(hip):
    from shakti_hip import synthetic as S
    g1, r1 = S.glacier(128, 32, bed_amp=10.0, seed=11)
    g2, r2 = S.glacier(128, 32, bed_amp=10.0, seed=11)
    for k in r1:
        assert np.array_equal(r1[k], r2[k])
    assert abs(r1["zb"].std() - 10.0) < 1e-9 and abs(r1["zb"].mean()) < 1e-9
    assert np.all(r1["mask"][0] == H_OCEAN) and np.all(r1["mask"][1:] == 0.0)
    m = S.moulin_field(g1, r1["mask"], n=20, seed=3)
    assert np.count_nonzero(m) == 20
    q = m[m > 0] * g1.dx * g1.dy
    assert np.all((q >= 1.0) & (q <= 10.0))


H_OCEAN = 1.0


# ---- the Julia binding (shakti.jl_amd/julia/ShaktiHIP.jl) against the C ABI ----
# Julia is not in the image, so the binding is checked as text: every ccall
# names a symbol the library exports and the headers declare, with as many
# argument types as the declaration has parameters, and every mirrored struct
# has the C struct's fields, in order and of the same widths.

JL = ROOT / "shakti.jl_amd" / "julia" / "ShaktiHIP.jl"


def _header_text():
    text = "".join((ROOT / "include" / h).read_text() for h in ("shakti_hip.h", "shakti_learned.h"))
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _top_level_items(s):
    """Comma-separated items of s at bracket depth 0 (empty items dropped)."""
    items, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            items.append(cur.strip())
            cur = ""
        else:
            cur += ch
    items.append(cur.strip())
    return [i for i in items if i]


def _c_arity():
    out = {}
    for m in re.finditer(r"\b(shk_\w+)\s*\(([^()]*)\)\s*;", _header_text()):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else len(_top_level_items(args))
    return out


def _jl_ccalls():
    src = JL.read_text()
    calls = []
    for m in re.finditer(r"ccall\(\(:(\w+),\s*libshk\),\s*[\w{}]+,\s*\(", src):
        i = j = m.end() - 1
        depth = 0
        while True:
            depth += {"(": 1, ")": -1}.get(src[j], 0)
            if depth == 0:
                break
            j += 1
        calls.append((m.group(1), len(_top_level_items(src[i + 1:j]))))
    return calls


def test_julia_ccalls_match_the_exported_declarations(hip):
    from shakti_hip import _lib as L
    lib = C.CDLL(str(L.LIB_PATH))
    arity = _c_arity()
    calls = _jl_ccalls()
    names = {n for n, _ in calls}
    # the plug-in's solve path, the device-resident simulation and the learned form
    for need in ("shk_create", "shk_solve_elliptic", "shk_solve_parabolic", "shk_step", "shk_checkpoint_save",
                 "shk_checkpoint_load", "shk_unet_create", "shk_attach_unet", "shk_set_moulin_field"):
        assert need in names, need
    bad = [(n, k, arity.get(n)) for n, k in calls if arity.get(n) != k or not hasattr(lib, n)]
    assert not bad, bad


_JL_WIDTH = {"Int32": ("int32_t", 4), "Int64": ("int64_t", 8), "Float64": ("double", 8), "Float32": ("float", 4),
             "UInt64": ("uint64_t", 8)}


def _c_struct(name):
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r"\s*;", _header_text())
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, rest = decl.split(None, 1)
        fields += [(f.strip(), ctype) for f in rest.split(",")]
    return fields


def _jl_struct(name):
    m = re.search(r"(?:mutable )?struct " + name + r"\n(.*?)\nend", JL.read_text(), flags=re.S)
    assert m, name
    fields = []
    for line in m.group(1).splitlines():
        if "=" in line:  # an inner constructor
            continue
        for item in line.split(";"):
            item = item.strip()
            if "::" in item:
                f, t = item.split("::")
                fields.append((f.strip(), _JL_WIDTH[t.strip()][0]))
    return fields


@pytest.mark.parametrize("jl,c", [("ShkGrid", "shk_grid"), ("ShkParams", "shk_params"), ("ShkModel", "shk_model"),
                                  ("ShkSolverCfg", "shk_solver_cfg"), ("ShkSolveStats", "shk_solve_stats"),
                                  ("ShkUnetCfg", "shk_unet_cfg")])
def test_julia_structs_mirror_the_c_structs(jl, c):
    assert _jl_struct(jl) == _c_struct(c)


def test_julia_state_fields_follow_the_field_ids():
    """STATE_FIELDS (the device-resident simulation's upload/download order)
    is enum shk_field's order, src/state.jl:13-57, and names each field as
    the reference's State does."""
    text = _header_text()
    enum = re.search(r"enum shk_field \{(.*?)\};", text, flags=re.S).group(1)
    ids = [e.strip().split("=")[0].strip() for e in enum.split(",") if e.strip()]
    assert ids[-1] == "SHK_NUM_FIELDS"
    m = re.search(r"const STATE_FIELDS = \((.*?)\)\n", JL.read_text(), flags=re.S)
    jl = [f.strip().lstrip(":") for f in m.group(1).replace("\n", " ").split(",") if f.strip()]
    assert len(jl) == len(ids) - 1 == 37
    # the ice thickness is State.H, SHK_HICE in the header
    assert ["SHK_HICE" if f == "H" else "SHK_" + f.upper() for f in jl] == ids[:-1]
