"""Simulation output: the observers of src/observer.jl.

NoObserver records nothing; LiveObserver keeps the tracked fields at the
tracked times in host memory; IOObserver writes them to a file.  The field
values come from a device snapshot of all tracked fields at once
(State.snapshot / State.fetch); the CSV writer's min / max / mean are reduced
on the device (State.stats), so a tracked time moves three numbers per field
over PCIe instead of the field.

File writers:
  CSVFileWriter     one row per tracked time: t, total_time and
                    <name>_min, <name>_max, <name>_mean per field
                    (src/observer.jl:142, 185-190, 324-339);
  NetCDFFileWriter  a dataset per field of shape (field..., ntimes) and a
                    "time" variable (src/observer.jl:151-162, 286-294),
                    written as NetCDF classic (64-bit offset) with scipy; the
                    dimensions are stored in C order (time, <name>_y,
                    <name>_x), which NetCDF readers of column-major languages
                    present as the reference's (<name>_x, <name>_y, time).
  HDF5FileWriter, JLD2FileWriter
                    need an HDF5 library, which this installation does not
                    have: constructing one raises.
"""
from __future__ import annotations

import csv

import numpy as np

from . import _lib as L


class AbstractFileWriter:
    pass


class CSVFileWriter(AbstractFileWriter):
    """src/observer.jl:40."""


class NetCDFFileWriter(AbstractFileWriter):
    """src/observer.jl:17."""


class HDF5FileWriter(AbstractFileWriter):
    """src/observer.jl:24 -- not available here (no HDF5 library)."""

    def __init__(self):
        raise L.ShaktiHIPError("HDF5FileWriter needs an HDF5 library, which is not installed here; "
                               "use NetCDF or CSV, or a checkpoint (save_checkpoint) for the full State")


class JLD2FileWriter(AbstractFileWriter):
    """src/observer.jl:32 -- not available here (JLD2 files are HDF5)."""

    def __init__(self):
        raise L.ShaktiHIPError("JLD2FileWriter needs an HDF5 library, which is not installed here; "
                               "use NetCDF or CSV, or a checkpoint (save_checkpoint) for the full State")


class AbstractObserver:
    pass


class NoObserver(AbstractObserver):
    """src/observer.jl:57."""


class LiveObserver(AbstractObserver):
    """src/observer.jl:87-94: history[name] has shape (field..., ntimes)."""

    def __init__(self, tracked_obs, tracked_times):
        self.tracked_obs = list(tracked_obs)
        self.tracked_times = [int(t) for t in tracked_times]
        self.history: dict[str, np.ndarray] = {}


class IOObserver(AbstractObserver):
    """src/observer.jl:66-79."""

    def __init__(self, tracked_obs, tracked_times, fr: AbstractFileWriter, path):
        self.tracked_obs = list(tracked_obs)
        self.tracked_times = [int(t) for t in tracked_times]
        self.fr = fr
        self.path = str(path)
        self.handle = None


def csv_stat_colnames(tracked_obs) -> list[str]:
    """src/observer.jl:142."""
    return ["t", "total_time"] + [name + suffix for name in tracked_obs for suffix in ("_min", "_max", "_mean")]


def _index(observer, t: int):
    try:
        return observer.tracked_times.index(int(t))
    except ValueError:
        return None


def _fields(observer, state) -> dict:
    state.snapshot(observer.tracked_obs)
    return {name: state.fetch(name) for name in observer.tracked_obs}


# ------------------------------------------------------------ prepare / resume

def prepare(observer: AbstractObserver, state) -> None:
    """prepare! (src/observer.jl:121-134)."""
    if isinstance(observer, LiveObserver):
        nt = len(observer.tracked_times)
        observer.history = {name: np.empty(state.shape_of(name) + (nt,), order="F") for name in observer.tracked_obs}
    elif isinstance(observer, IOObserver):
        observer.handle = _openfile(observer.fr, observer, state)


def resume(observer: AbstractObserver, state, resume_t: int) -> None:
    """resume! (src/observer.jl:214-219): the IO observer reopens its file for
    further writes; the live observer starts a fresh buffer."""
    if isinstance(observer, LiveObserver):
        prepare(observer, state)
    elif isinstance(observer, IOObserver):
        observer.handle = _reopenfile(observer.fr, observer, state, int(resume_t))


def _openfile(fr, observer: IOObserver, state):
    if isinstance(fr, CSVFileWriter):
        with open(observer.path, "w", newline="") as f:
            csv.writer(f).writerow(csv_stat_colnames(observer.tracked_obs))
        return observer.path
    if isinstance(fr, NetCDFFileWriter):
        from scipy.io import netcdf_file
        nc = netcdf_file(observer.path, "w", version=2, mmap=False)
        nc.createDimension("time", len(observer.tracked_times))
        nc.createVariable("time", "d", ("time",))
        for name in observer.tracked_obs:
            rows, cols = state.shape_of(name)
            nc.createDimension(f"{name}_x", rows)
            nc.createDimension(f"{name}_y", cols)
            nc.createVariable(name, "d", ("time", f"{name}_y", f"{name}_x"))
        nc.flush()
        return nc
    raise L.ShaktiHIPError(f"unsupported file writer {type(fr).__name__}")


def _reopenfile(fr, observer: IOObserver, state, resume_t: int):
    if isinstance(fr, CSVFileWriter):
        # rows past resume_t came from the run that stopped; the resumed run
        # writes them again (src/observer.jl:242-247)
        with open(observer.path, newline="") as f:
            rows = list(csv.reader(f))
        keep = [rows[0]] + [r for r in rows[1:] if r and int(r[0]) <= resume_t]
        with open(observer.path, "w", newline="") as f:
            csv.writer(f).writerows(keep)
        return observer.path
    if isinstance(fr, NetCDFFileWriter):
        from scipy.io import netcdf_file
        # preallocated slots: a rewritten time index is overwritten in place
        return netcdf_file(observer.path, "a", mmap=False)
    raise L.ShaktiHIPError(f"unsupported file writer {type(fr).__name__}")


# ------------------------------------------------------------ observe / finalize

def observe(observer: AbstractObserver, state, t: int, total_time: float) -> None:
    """observe! (src/observer.jl:261-280)."""
    if isinstance(observer, NoObserver):
        return
    idx = _index(observer, t)
    if idx is None:
        return
    if isinstance(observer, LiveObserver):
        for name, a in _fields(observer, state).items():
            observer.history[name][..., idx] = a
        return
    fr = observer.fr
    if isinstance(fr, CSVFileWriter):
        row = [observer.tracked_times[idx], repr(float(total_time))]
        for name in observer.tracked_obs:
            row += [repr(v) for v in state.stats(name)]
        with open(observer.path, "a", newline="") as f:
            csv.writer(f).writerow(row)
    elif isinstance(fr, NetCDFFileWriter):
        nc = observer.handle
        nc.variables["time"][idx] = total_time
        for name, a in _fields(observer, state).items():
            nc.variables[name][idx, :, :] = a.T
        nc.flush()  # the slot is on disk before the next step


def finalize(observer: AbstractObserver, state) -> None:
    """finalize! (src/observer.jl:350-353)."""
    if isinstance(observer, IOObserver) and isinstance(observer.fr, NetCDFFileWriter) and observer.handle is not None:
        observer.handle.close()
        observer.handle = None


def read_csv(path) -> list[dict]:
    """Rows of a CSV observer file, numbers parsed."""
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    return [{k: (int(v) if k == "t" else float(v)) for k, v in r.items()} for r in rows]


def read_netcdf(path, name: str) -> np.ndarray:
    """A NetCDF observer dataset as (field..., ntimes), the reference's order."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        a = np.array(nc.variables[name][:], dtype=np.float64)
    if a.ndim == 1:
        return a
    return np.asfortranarray(np.transpose(a, (2, 1, 0)))


__all__ = ["AbstractFileWriter", "CSVFileWriter", "NetCDFFileWriter", "HDF5FileWriter", "JLD2FileWriter",
           "AbstractObserver", "NoObserver", "LiveObserver", "IOObserver", "prepare", "resume", "observe", "finalize",
           "read_csv", "read_netcdf", "csv_stat_colnames"]
