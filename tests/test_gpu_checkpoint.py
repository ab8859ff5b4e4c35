"""Checkpoint / restart and observers on the device State (runs on an MI355X).

The reference's own restart test (test/checkpoint_test.jl:1-150) on its 6x6
mixed-mask case: a full 8-step run; a run stopped at step 5 that
checkpointed every 3 steps; a fresh Simulation resumed from that checkpoint
to step 8.  The resumed State must equal the full run's -- here bit for bit,
since every solve recomputes its Chebyshev bounds from its own system -- and
so must the observer files the resumed run appended to.  Then the file
itself: the library's file read by the host reader (every checksum), a
host-written file loaded by the library, damaged files refused with the State
untouched, snapshots isolated from later steps, and the device min / max /
mean against numpy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make_sim(H, O, tsteps, writer=None, path=None, tracked=("h", "b")):
    g, _, _, raw = O.test_fixture_6x6(sliding_C=0.25)
    hg = H.Grid(g.nx, g.ny, g.lx, g.ly)
    hp = H.ModelParameters(e_v=0.0)
    s = H.State(hg, hp)
    sl = H.RegularizedCoulombSlidingLaw(0.25)
    H.set_initial_conditions(s, hg, hp, sl, **raw)
    ls = H.CGIterativeSolver(hg, H.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-13)
    ps = H.PicardSolver(500, 1e-6, ls, hg)
    kw = {}
    if writer is not None:
        kw = dict(which_observer="IO", which_file_writer=writer, tracked_times=range(0, 9), path=str(path))
    return H.Simulation(hg, s, tsteps, 3600.0, hp, "implicit", list(tracked) if writer else [],
                        H.ConstantMeltInput(), sl, ps=ps, **kw)


@pytest.mark.parametrize("writer", [None, "CSV", "NetCDF"])
def test_resumed_run_reproduces_the_full_run(oracle, hip, tmp_path, writer):
    from shakti_hip import observer as Obs
    ext = {None: "", "CSV": ".csv", "NetCDF": ".nc"}[writer]
    full = _make_sim(hip, oracle, 8, writer, tmp_path / f"full{ext}")
    hip.run(full)
    ckpt = tmp_path / "ckpt.shk"
    crash = _make_sim(hip, oracle, 5, writer, tmp_path / f"crash{ext}")
    hip.run(crash, checkpoint_every=3, checkpoint_path=ckpt)
    resumed = _make_sim(hip, oracle, 8, writer, tmp_path / f"crash{ext}")
    hip.run(resumed, restart_path=ckpt)
    for name in ["h", "b", "N", "mdot", "taub_x", "q_y"]:
        assert np.array_equal(resumed.state.get(name), full.state.get(name)), name
    assert resumed.total_time == full.total_time == 8 * 3600.0
    assert not (tmp_path / "ckpt.shk.tmp").exists()
    if writer == "CSV":
        rows_full = Obs.read_csv(tmp_path / "full.csv")
        rows_res = Obs.read_csv(tmp_path / "crash.csv")
        assert len(rows_res) == len(rows_full) == 9  # no row of t = 4, 5 twice
        assert rows_res == rows_full
        h = full.state.get("h")
        last = rows_full[-1]
        assert (last["h_min"], last["h_max"]) == (h.min(), h.max())
        assert last["h_mean"] == pytest.approx(h.mean(), rel=1e-14)
    if writer == "NetCDF":
        for name in ["h", "b"]:
            a_full = Obs.read_netcdf(tmp_path / "full.nc", name)
            assert a_full.shape == (6, 6, 9)
            assert np.array_equal(Obs.read_netcdf(tmp_path / "crash.nc", name), a_full)
        assert np.array_equal(a_full[..., -1], full.state.get("b"))


def test_checkpoint_file_holds_the_state_and_host_files_load(oracle, hip, tmp_path):
    from shakti_hip import checkpoint as K
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(700, 300, bed_amp=10.0, seed=2)  # 37 fields of 210k values: multi-block checksums
    hg = hip.Grid(gg.nx, gg.ny, gg.lx, gg.ly)
    hp = hip.ModelParameters()
    s = hip.State(hg, hp)
    hip.set_initial_conditions(s, hg, hp, hip.RegularizedCoulombSlidingLaw(0.5), **raw)
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-10)
    sim = hip.Simulation(hg, s, 1, 3600.0, hp, "implicit", [], hip.ConstantMeltInput(),
                         hip.RegularizedCoulombSlidingLaw(0.5), ps=hip.PicardSolver(50, 1e-6, ls, hg))
    hip.step(sim)
    sim.total_time = 3600.0
    want = s.to_dict()
    path = tmp_path / "big.shk"
    hip.save_checkpoint(path, sim, 1)
    ck = K.read(path)  # verifies every checksum against the device's
    assert (ck.nx, ck.ny, ck.t, ck.total_time, ck.lx, ck.ly) == (700, 300, 1, 3600.0, gg.lx, gg.ly)
    for n in K.FIELD_NAMES:
        assert np.array_equal(ck.fields[n], want[n]), n
    # a host-written file (every field changed) loads into a fresh State
    rng = np.random.default_rng(5)
    ck2 = K.Checkpoint(nx=700, ny=300, t=42, total_time=1.5e5, lx=gg.lx, ly=gg.ly,
                       fields={n: np.asfortranarray(a + rng.standard_normal(a.shape)) for n, a in want.items()})
    K.write(tmp_path / "host.shk", ck2)
    s2 = hip.State(hg, hp)
    sim2 = hip.Simulation(hg, s2, 1, 3600.0, hp, "implicit", [], hip.ConstantMeltInput(),
                          hip.RegularizedCoulombSlidingLaw(0.5), ps=hip.PicardSolver(50, 1e-6, ls, hg))
    assert hip.load_checkpoint(sim2, tmp_path / "host.shk") == 42
    assert sim2.total_time == 1.5e5
    got = s2.to_dict()
    for n in K.FIELD_NAMES:
        assert np.array_equal(got[n], ck2.fields[n]), n


@pytest.mark.parametrize("damage", ["flip_data", "truncate", "other_grid", "missing"])
def test_damaged_checkpoints_leave_the_state_unchanged(oracle, hip, tmp_path, damage):
    from shakti_hip import _lib as L
    from shakti_hip import checkpoint as K
    sim = _make_sim(hip, oracle, 2)
    hip.run(sim)
    path = tmp_path / "c.shk"
    hip.save_checkpoint(path, sim, 2)
    raw = bytearray(path.read_bytes())
    target = _make_sim(hip, oracle, 2)
    if damage == "flip_data":
        rec = K._ENTRY.unpack_from(raw, 64 + 3 * 64)  # b's record
        raw[rec[5] + 17] ^= 0x04
        path.write_bytes(bytes(raw))
    elif damage == "truncate":
        path.write_bytes(bytes(raw[:len(raw) // 2]))
    elif damage == "other_grid":
        g = hip.Grid(8, 6, 1e3, 1e3)
        st = hip.State(g, hip.ModelParameters())
        target = hip.Simulation(g, st, 1, 3600.0, hip.ModelParameters(), "implicit", [], hip.ConstantMeltInput(),
                                hip.RegularizedCoulombSlidingLaw(0.25), ps=sim.hs.ps)
    else:
        path = tmp_path / "absent.shk"
    before = target.state.to_dict()
    with pytest.raises(L.ShaktiHIPError) as e:
        hip.load_checkpoint(target, path)
    if damage == "flip_data":
        assert "checksum" in str(e.value)
    after = target.state.to_dict()
    for n in before:
        assert np.array_equal(before[n], after[n]), n


def test_failed_writes_are_reported(oracle, hip, tmp_path):
    from shakti_hip import _lib as L
    sim = _make_sim(hip, oracle, 1)
    bad = tmp_path / "no_such_dir" / "c.shk"
    with pytest.raises(L.ShaktiHIPError):
        hip.save_checkpoint(bad, sim, 0)
    hip.save_checkpoint(bad, sim, 0, asynchronous=True)  # fails on the writer thread
    with pytest.raises(L.ShaktiHIPError, match="cannot open"):
        hip.wait_checkpoint(sim)
    hip.wait_checkpoint(sim)  # reported once
    hip.save_checkpoint(tmp_path / "ok.shk", sim, 0, asynchronous=True)
    hip.wait_checkpoint(sim)
    assert (tmp_path / "ok.shk").exists()


def test_snapshot_is_isolated_from_later_steps(oracle, hip):
    sim = _make_sim(hip, oracle, 3)
    hip.step(sim)
    s = sim.state
    before = {n: s.get(n) for n in ["h", "b", "taub_x"]}
    s.snapshot(["h", "b", "taub_x"])
    hip.step(sim)  # queued behind the snapshot's device copy
    assert not np.array_equal(s.get("h"), before["h"])
    for n, a in before.items():
        assert np.array_equal(s.fetch(n), a), n
    from shakti_hip import _lib as L
    with pytest.raises(L.ShaktiHIPError):
        s.fetch("N")  # not in the snapshot


def test_field_stats_match_numpy(hip):
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(1031, 517, bed_amp=10.0, seed=3)
    hg = hip.Grid(gg.nx, gg.ny, gg.lx, gg.ly)
    hp = hip.ModelParameters()
    s = hip.State(hg, hp)
    hip.set_initial_conditions(s, hg, hp, hip.RegularizedCoulombSlidingLaw(0.5), **raw)
    for name in ["zb", "po", "N", "mask", "q_x", "taub_y", "valid_y"]:
        a = s.get(name)
        mn, mx, mean = s.stats(name)
        assert (mn, mx) == (a.min(), a.max()), name
        # a zero-mean field (zb) has a mean that is all rounding: bound by the mean magnitude
        assert abs(mean - a.mean()) <= 1e-13 * np.abs(a).mean(), name
    g = s.get("G")
    g[5, 7] = np.nan
    s.set("G", g)
    assert all(np.isnan(v) for v in s.stats("G"))
