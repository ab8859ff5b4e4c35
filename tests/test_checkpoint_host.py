"""Host logic of checkpoints and observers (CPU).

The checkpoint format's reader and writer (shakti_hip.checkpoint) against
SplitMix64's published outputs, round trips, and every corruption the
library's loader must refuse; the observers' CSV / NetCDF / live bookkeeping
(src/observer.jl) on a host stand-in for the device State, including the
resume rules: CSV rows past the resumed step are dropped, NetCDF slots are
overwritten in place (src/observer.jl:225-247)."""
import numpy as np
import pytest

from shakti_hip import checkpoint as K


def _fields(nx, ny, seed=0):
    rng = np.random.default_rng(seed)
    return {n: np.asfortranarray(rng.standard_normal(K.field_shape(n, nx, ny))) for n in K.FIELD_NAMES}


def test_splitmix64_published_outputs():
    """SplitMix64 from state 0: its first three outputs (Vigna's reference
    implementation) are splitmix64(k * golden) for k = 0, 1, 2."""
    z = np.array([0, 0x9E3779B97F4A7C15, (2 * 0x9E3779B97F4A7C15) % 2 ** 64], dtype=np.uint64)
    assert [int(v) for v in K.splitmix64(z)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_checksum_sees_values_and_positions():
    a = np.asfortranarray(np.arange(12.0).reshape(3, 4, order="F"))
    c0 = K.checksum(a)
    # the same sum in Python integers, one value at a time
    M = 2 ** 64

    def sm(z):
        z = (z + 0x9E3779B97F4A7C15) % M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) % M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) % M
        return z ^ (z >> 31)

    words = [int(w) for w in a.ravel(order="F").view(np.uint64)]
    assert c0 == sum(sm((w + i * 0x9E3779B97F4A7C15) % M) for i, w in enumerate(words)) % M
    b = a.copy(order="F")
    b[0, 0], b[1, 0] = a[1, 0], a[0, 0]  # swapped values
    assert K.checksum(b) != c0
    b = a.copy(order="F")
    b.ravel(order="F").view(np.uint64)[5] ^= np.uint64(1)  # one flipped bit
    assert K.checksum(b) != c0
    big = np.asfortranarray(np.random.default_rng(1).standard_normal((2048, 2100)))  # crosses the chunk size
    w = big.ravel(order="F").view(np.uint64)
    i = np.arange(w.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        assert K.checksum(big) == int(np.sum(K.splitmix64(w + i * K.GOLDEN), dtype=np.uint64))


def test_round_trip(tmp_path):
    nx, ny = 7, 5
    ck = K.Checkpoint(nx=nx, ny=ny, t=3, total_time=10800.0, lx=1e3, ly=2e3, fields=_fields(nx, ny))
    p = tmp_path / "c.shk"
    K.write(p, ck)
    assert not (tmp_path / "c.shk.tmp").exists()
    r = K.read(p)
    assert (r.nx, r.ny, r.t, r.total_time, r.lx, r.ly) == (nx, ny, 3, 10800.0, 1e3, 2e3)
    for n in K.FIELD_NAMES:
        assert np.array_equal(r.fields[n], ck.fields[n]) and r.fields[n].flags.f_contiguous
    # layout: 512-byte aligned blocks after a 64-byte header and 64-byte records
    raw = p.read_bytes()
    assert raw[:8] == b"SHKCKPT1"
    assert len(raw) % 512 == 0
    npz = tmp_path / "c.npz"
    K.to_npz(p, npz)
    with np.load(npz) as z:
        assert int(z["t"]) == 3 and np.array_equal(z["taub_y"], ck.fields["taub_y"])


@pytest.mark.parametrize("damage", ["flip_data", "truncate", "magic", "version", "shape"])
def test_damaged_files_are_refused(tmp_path, damage):
    nx, ny = 6, 6
    ck = K.Checkpoint(nx=nx, ny=ny, t=1, total_time=3600.0, fields=_fields(nx, ny, seed=2))
    p = tmp_path / "c.shk"
    K.write(p, ck)
    raw = bytearray(p.read_bytes())
    if damage == "flip_data":
        last = K._ENTRY.unpack_from(raw, 64 + 36 * 64)  # valid_y's record
        raw[last[5] + last[6] - 3] ^= 0x10  # a byte of its last value
    elif damage == "truncate":
        raw = raw[:len(raw) - 1000]
    elif damage == "magic":
        raw[0:8] = b"NOTACKPT"
    elif damage == "version":
        raw[8] = 9
    elif damage == "shape":
        raw[64 + 20] = 99  # the first record's rows
    p.write_bytes(bytes(raw))
    with pytest.raises(K.CheckpointError):
        K.read(p)


def test_writer_refuses_incomplete_states(tmp_path):
    f = _fields(4, 3)
    del f["valid_y"]
    with pytest.raises(K.CheckpointError):
        K.write(tmp_path / "x", K.Checkpoint(nx=4, ny=3, t=0, total_time=0.0, fields=f))
    f = _fields(4, 3)
    f["h"] = np.zeros((3, 4))
    with pytest.raises(K.CheckpointError):
        K.write(tmp_path / "x", K.Checkpoint(nx=4, ny=3, t=0, total_time=0.0, fields=f))


# ------------------------------------------------------------------ observers

class _HostState:
    """The parts of State the observers use, on host arrays."""

    def __init__(self, nx, ny, seed=0):
        self.nx, self.ny = nx, ny
        self.fields = _fields(nx, ny, seed)
        self._snap = {}

    def shape_of(self, name):
        return K.field_shape(name, self.nx, self.ny)

    def snapshot(self, names):
        self._snap = {n: self.fields[n].copy(order="F") for n in names}

    def fetch(self, name):
        return self._snap[name]

    def stats(self, name):
        a = self.fields[name]
        return float(a.min()), float(a.max()), float(a.mean())

    def advance(self, seed):
        self.fields = _fields(self.nx, self.ny, seed)


def _run(obs, state, t_from, t_to, O):
    for t in range(t_from, t_to + 1):
        state.advance(100 + t)
        O.observe(obs, state, t, 3600.0 * t)


def test_csv_observer_rows_and_resume(tmp_path):
    from shakti_hip import observer as O
    times = list(range(0, 9))
    full = O.IOObserver(["h", "b"], times, O.CSVFileWriter(), tmp_path / "full.csv")
    s = _HostState(6, 5)
    O.prepare(full, s)
    s.advance(100)
    O.observe(full, s, 0, 0.0)
    _run(full, s, 1, 8, O)
    O.finalize(full, s)
    rows = O.read_csv(tmp_path / "full.csv")
    assert list(rows[0]) == ["t", "total_time", "h_min", "h_max", "h_mean", "b_min", "b_max", "b_mean"]
    assert [r["t"] for r in rows] == times
    # a run that stopped after t = 5 with its last checkpoint at t = 3, resumed
    crash = O.IOObserver(["h", "b"], times, O.CSVFileWriter(), tmp_path / "crash.csv")
    O.prepare(crash, s)
    s.advance(100)
    O.observe(crash, s, 0, 0.0)
    _run(crash, s, 1, 5, O)
    O.finalize(crash, s)
    O.resume(crash, s, 3)
    assert [r["t"] for r in O.read_csv(tmp_path / "crash.csv")] == [0, 1, 2, 3]
    _run(crash, s, 4, 8, O)
    O.finalize(crash, s)
    assert O.read_csv(tmp_path / "crash.csv") == rows


def test_netcdf_observer_layout_and_resume(tmp_path):
    from shakti_hip import observer as O
    times = [0, 2, 4, 6, 8]
    s = _HostState(7, 4)
    full = O.IOObserver(["h", "q_x", "taub_y"], times, O.NetCDFFileWriter(), tmp_path / "full.nc")
    O.prepare(full, s)
    hist = {}
    for t in range(0, 9):
        s.advance(100 + t)
        O.observe(full, s, t, 3600.0 * t)
        if t in times:
            hist[t] = {n: s.fields[n].copy() for n in ["h", "q_x", "taub_y"]}
    O.finalize(full, s)
    for n in ["h", "q_x", "taub_y"]:
        a = O.read_netcdf(tmp_path / "full.nc", n)
        assert a.shape == s.shape_of(n) + (len(times),)
        for k, t in enumerate(times):
            assert np.array_equal(a[..., k], hist[t][n]), (n, t)
    assert np.array_equal(O.read_netcdf(tmp_path / "full.nc", "time"), 3600.0 * np.array(times))
    # resumed: the slots of t = 4 (already written by the stopped run) are overwritten
    crash = O.IOObserver(["h", "q_x", "taub_y"], times, O.NetCDFFileWriter(), tmp_path / "crash.nc")
    O.prepare(crash, s)
    for t in range(0, 6):
        s.advance(100 + t if t != 4 else 999)  # the stopped run saw another t = 4
        O.observe(crash, s, t, 3600.0 * t)
    O.finalize(crash, s)
    O.resume(crash, s, 3)
    for t in range(4, 9):
        s.advance(100 + t)
        O.observe(crash, s, t, 3600.0 * t)
    O.finalize(crash, s)
    for n in ["h", "q_x", "taub_y"]:
        assert np.array_equal(O.read_netcdf(tmp_path / "crash.nc", n), O.read_netcdf(tmp_path / "full.nc", n))


def test_live_observer_history():
    from shakti_hip import observer as O
    s = _HostState(5, 5)
    obs = O.LiveObserver(["N"], [1, 3])
    O.prepare(obs, s)
    seen = {}
    for t in range(0, 5):
        s.advance(t)
        O.observe(obs, s, t, 0.0)
        seen[t] = s.fields["N"].copy()
    assert obs.history["N"].shape == (5, 5, 2)
    assert np.array_equal(obs.history["N"][..., 0], seen[1]) and np.array_equal(obs.history["N"][..., 1], seen[3])


def test_observer_choices_raise_as_the_reference():
    """Simulation's observer set-up errors (src/simulation.jl:148-181)."""
    from shakti_hip import _lib as L
    from shakti_hip.api import _make_observer
    from shakti_hip import observer as O
    assert isinstance(_make_observer([], None, None, None, None), O.NoObserver)
    with pytest.raises(L.ShaktiHIPError, match="which_observer should be specified"):
        _make_observer(["h"], None, None, None, None)
    with pytest.raises(L.ShaktiHIPError, match="which_file_writer should be specified"):
        _make_observer(["h"], "IO", None, [0], "x")
    with pytest.raises(L.ShaktiHIPError, match="Unknown which_file_writer"):
        _make_observer(["h"], "IO", "XLS", [0], "x")
    with pytest.raises(L.ShaktiHIPError, match="tracked_times should be specified"):
        _make_observer(["h"], "IO", "CSV", None, "x")
    with pytest.raises(L.ShaktiHIPError, match="path should be specified"):
        _make_observer(["h"], "IO", "CSV", [0], None)
    with pytest.raises(L.ShaktiHIPError, match="Unknown which_observer"):
        _make_observer(["h"], "Disk", None, [0], None)
    with pytest.raises(L.ShaktiHIPError, match="HDF5 library"):
        _make_observer(["h"], "IO", "HDF5", [0], "x")
    assert isinstance(_make_observer(["h"], "Live", None, [0, 1], None), O.LiveObserver)
