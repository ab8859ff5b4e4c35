"""Host-side reader and writer of the library's checkpoint files.

The library writes and reads these files itself (shk_checkpoint_save /
shk_checkpoint_load); this module lets a host program inspect one, convert it
(``to_npz``) or build one from fields it already holds, e.g. the State of a
JLD2 checkpoint written by the reference's save_checkpoint
(src/checkpoint.jl:26-38), which stores the same field names.

Layout (little-endian): a 64-byte header -- magic b"SHKCKPT1", version (u32),
field count (u32), nx, ny (i32), t (i64), total_time, lx, ly (f64), data
offset (u64) -- then one 64-byte record per field -- name (16 bytes, NUL
padded), field id, rows, cols, reserved (i32), byte offset, byte count,
checksum, reserved (u64) -- then the fields as FP64, column-major, each block
starting on a 512-byte boundary.  The checksum of n values with bit patterns
w_i is sum_i splitmix64(w_i + i * 0x9E3779B97F4A7C15) mod 2^64 (Vigna's
SplitMix64 finaliser).
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np

MAGIC = b"SHKCKPT1"
VERSION = 1
_HEADER = struct.Struct("<8sIIiiqdddQ")
_ENTRY = struct.Struct("<16siiiiQQQQ")
assert _HEADER.size == 64 and _ENTRY.size == 64

GOLDEN = np.uint64(0x9E3779B97F4A7C15)

# the reference State's fields (src/state.jl:13-57) in the library's id order
FIELD_NAMES = ["h", "pw", "po", "b", "beta", "abs_ub", "mdot", "shear", "potential", "sensible", "Re", "K", "G", "zb",
               "zs", "H", "ieb", "lambda", "A_visc", "N", "mask",
               "dhdx", "q_x", "Re_x", "b_x", "ub_x", "taub_x", "dpwdx", "valid_x",
               "dhdy", "q_y", "Re_y", "b_y", "ub_y", "taub_y", "dpwdy", "valid_y"]
N_CENTER, N_XFACE = 21, 8


class CheckpointError(ValueError):
    pass


def field_shape(name: str, nx: int, ny: int) -> tuple[int, int]:
    i = FIELD_NAMES.index(name)
    if i < N_CENTER:
        return nx, ny
    if i < N_CENTER + N_XFACE:
        return nx + 1, ny
    return nx, ny + 1


def splitmix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64's output function of state z (uint64, wrapping)."""
    with np.errstate(over="ignore"):
        z = np.asarray(z, dtype=np.uint64) + GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def checksum(a: np.ndarray) -> int:
    """The field checksum of an FP64 array in column-major order."""
    w = np.asfortranarray(a, dtype=np.float64).ravel(order="F").view(np.uint64)
    out = np.uint64(0)
    step = 1 << 22
    with np.errstate(over="ignore"):
        for s in range(0, w.size, step):
            i = np.arange(s, min(s + step, w.size), dtype=np.uint64)
            out = out + np.sum(splitmix64(w[s:s + step] + i * GOLDEN), dtype=np.uint64)
    return int(out)


@dataclass
class Checkpoint:
    nx: int
    ny: int
    t: int
    total_time: float
    lx: float = 0.0
    ly: float = 0.0
    fields: dict = field(default_factory=dict)


def read(path, verify: bool = True) -> Checkpoint:
    """Read a checkpoint; with verify, every field's checksum is checked."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        head = f.read(_HEADER.size)
        if len(head) < _HEADER.size:
            raise CheckpointError(f"{path} is not a checkpoint (short header)")
        magic, version, nf, nx, ny, t, total_time, lx, ly, data_off = _HEADER.unpack(head)
        if magic != MAGIC:
            raise CheckpointError(f"{path} is not a checkpoint of this library")
        if version != VERSION:
            raise CheckpointError(f"{path}: checkpoint version {version}, this reader knows {VERSION}")
        table = f.read(nf * _ENTRY.size)
        if len(table) < nf * _ENTRY.size:
            raise CheckpointError(f"{path} is truncated (field table)")
        ck = Checkpoint(nx=nx, ny=ny, t=t, total_time=total_time, lx=lx, ly=ly)
        for k in range(nf):
            name, fid, rows, cols, _, off, nbytes, cks, _ = _ENTRY.unpack_from(table, k * _ENTRY.size)
            name = name.rstrip(b"\0").decode()
            if nbytes != rows * cols * 8 or off + nbytes > size:
                raise CheckpointError(f"{path}: field {name} is truncated or malformed")
            f.seek(off)
            a = np.frombuffer(f.read(nbytes), dtype="<f8").reshape((rows, cols), order="F").copy(order="F")
            if verify and checksum(a) != cks:
                raise CheckpointError(f"{path}: field {name} fails its checksum")
            ck.fields[name] = a
    return ck


def write(path, ck: Checkpoint) -> None:
    """Write ``ck`` (every State field present, shaped for nx x ny) as the
    library does: into path + ".tmp", flushed, then renamed over path."""
    missing = [n for n in FIELD_NAMES if n not in ck.fields]
    if missing:
        raise CheckpointError(f"a checkpoint holds every State field; missing {missing}")
    arrays = []
    for n in FIELD_NAMES:
        a = np.asfortranarray(ck.fields[n], dtype=np.float64)
        if a.shape != field_shape(n, ck.nx, ck.ny):
            raise CheckpointError(f"field {n} has shape {a.shape}, expected {field_shape(n, ck.nx, ck.ny)}")
        arrays.append(a)
    nf = len(FIELD_NAMES)
    table_end = _HEADER.size + nf * _ENTRY.size
    data_off = (table_end + 511) // 512 * 512
    entries, pos = [], data_off
    for i, (n, a) in enumerate(zip(FIELD_NAMES, arrays)):
        nbytes = a.size * 8
        entries.append(_ENTRY.pack(n.encode(), i, a.shape[0], a.shape[1], 0, pos, nbytes, checksum(a), 0))
        pos += (nbytes + 511) // 512 * 512
    tmp = str(path) + ".tmp"
    with open(tmp, "wb") as f:
        f.write(_HEADER.pack(MAGIC, VERSION, nf, ck.nx, ck.ny, int(ck.t), float(ck.total_time), float(ck.lx),
                             float(ck.ly), data_off))
        f.write(b"".join(entries))
        f.write(b"\0" * (data_off - table_end))
        for a in arrays:
            b = a.ravel(order="F").astype("<f8").tobytes()
            f.write(b)
            f.write(b"\0" * ((512 - len(b) % 512) % 512))
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)


def to_npz(src, dst) -> None:
    """Convert a checkpoint to a numpy .npz archive: every field by name, plus
    t, total_time, nx, ny, lx and ly."""
    ck = read(src)
    np.savez(dst, t=np.int64(ck.t), total_time=np.float64(ck.total_time), nx=np.int64(ck.nx), ny=np.int64(ck.ny),
             lx=np.float64(ck.lx), ly=np.float64(ck.ly), **ck.fields)
