"""NairnMPM archive files, format version 6 (ArchiveData, System/ArchiveData.cpp): writer through
the C ABI, and a reader that lays records out by the rules of ArchiveData::CalcArchiveSize
(:266-424) and ArchiveResults (:721-1000)."""
import ctypes
import os

import numpy as np

from . import Material, MPMError, library, _dptr, _iptr

HEADER_LENGTH = 64
NUM_MPM_ITEMS = 25
(A_BYTE_ORDER, A_DEFAULTS, A_VELOCITY, A_STRESS, A_STRAIN, A_PLASTIC_STRAIN, A_OLD_ORIG_POSITION,
 A_WORK_ENERGY, A_DELTA_TEMP, A_PLASTIC_ENERGY, A_VER2_EMPTY, A_SHEAR_COMPONENTS, A_STRAIN_ENERGY,
 A_HISTORY, A_CONCENTRATION, A_HEAT_ENERGY, A_ELEMENT_CROSSINGS, A_ROT_STRAIN, A_DAMAGE_NORMAL,
 A_SPIN_MOMENTUM, A_SPIN_VELOCITY, A_HISTORY59, A_HISTORY1014, A_HISTORY1519, A_SIZE) = range(25)


def _err():
    m = library().mpmgpu_archive_error()
    return MPMError(m.decode() if m else "archive error")


def normalize_order(order: str) -> str:
    """The order as an archive stores it (padding, fixed items, rotation angles with the strain)"""
    out = ctypes.create_string_buffer(NUM_MPM_ITEMS + 1)
    if library().mpmgpu_archive_order(order.encode("latin-1"), out) != 0:
        raise _err()
    return out.raw[:NUM_MPM_ITEMS].decode("latin-1")


def record_size(order: str, dim: int) -> int:
    n = library().mpmgpu_archive_record_size(order.encode("latin-1"), int(dim))
    if n < 0:
        raise _err()
    return n


def write(path, dim, order, time_ms, dfields, ifields, mats, thickness=None, default_thickness=1.):
    """Archive of particles given in the C ABI layout (host arrays, mpm[] order)"""
    d = np.ascontiguousarray(dfields, dtype=np.float64)
    i = np.ascontiguousarray(ifields, dtype=np.int32)
    arr = (Material * len(mats))(*mats)
    t = None if thickness is None else np.ascontiguousarray(thickness, dtype=np.float64)
    if library().mpmgpu_archive_write(os.fsencode(path), int(dim), order.encode("latin-1"), float(time_ms),
                                      d.shape[1], _dptr(d), _iptr(i), arr, len(mats),
                                      None if t is None else _dptr(t), float(default_thickness)) != 0:
        raise _err()


def _history_numbers(c: str):
    if c == "N":
        return []
    if c == "Y":
        return [1]
    b = ord(c)
    return [k + 1 for k in range(4) if b & (1 << k)]


def record_dtype(order: str, dim: int):
    """numpy dtype of one particle record (little endian), padded to the record size"""
    three = dim == 3
    f = [("elem", "<i4"), ("mass", "<f8"), ("mat", "<i2"), ("pad", "<i2")]
    if three:
        f += [("angle_z", "<f8"), ("angle_y", "<f8"), ("angle_x", "<f8")]
    else:
        f += [("angle_z", "<f8"), ("thickness", "<f8")]
    axes = "xyz" if three else "xy"
    comps = ["xx", "yy", "zz", "xy"] + (["xz", "yz"] if three else [])
    f += [("pos_" + a, "<f8") for a in axes] + [("orig_" + a, "<f8") for a in axes]
    on = lambda k: order[k] != "N"
    if on(A_VELOCITY):
        f += [("vel_" + a, "<f8") for a in axes]
    if on(A_STRESS):
        f += [("stress_" + c, "<f8") for c in comps]
    if on(A_STRAIN):
        f += [("strain_" + c, "<f8") for c in comps]
    if on(A_PLASTIC_STRAIN):
        f += [("plastic_" + c, "<f8") for c in comps]
    if on(A_SHEAR_COMPONENTS):
        f += [("dudy", "<f8"), ("dvdx", "<f8")]
    f += [("history%d" % h, "<f8") for h in _history_numbers(order[A_HISTORY])]
    if on(A_ROT_STRAIN):
        f += [("angle0_z", "<f8")] + ([("angle0_y", "<f8"), ("angle0_x", "<f8")] if three else [])
    dt = np.dtype(f)
    rec = record_size(order, dim)
    if dt.itemsize > rec:
        raise MPMError("record layout is larger than the record size")
    f += [("fill", "V%d" % (rec - dt.itemsize))] if rec > dt.itemsize else []
    return np.dtype(f)


def read(path):
    """(header dict, records) of an archive file written in format version 6, little endian"""
    with open(path, "rb") as fh:
        raw = fh.read()
    h = raw[:HEADER_LENGTH]
    if h[:4] != b"ver6":
        raise MPMError("not a version-6 archive")
    k = 4
    nm = h[k]; k += 1
    order = h[k:k + nm].decode("latin-1"); k += nm
    nc = h[k]; k += 1
    crack = h[k:k + nc].decode("latin-1"); k += nc
    dim = int(chr(h[k])); k += 1
    structured = chr(h[k]) == "1"; k += 1
    time_ms = float(np.frombuffer(h[k:k + 4], dtype="<f4")[0])
    if order[A_BYTE_ORDER] != "i":
        raise MPMError("only little-endian ('i') archives are read here")
    dt = record_dtype(order, dim)
    body = raw[HEADER_LENGTH:]
    if len(body) % dt.itemsize:
        raise MPMError("archive length is not a whole number of records")
    recs = np.frombuffer(body, dtype=dt)
    return {"version": 6, "mpm_order": order, "crack_order": crack, "dim": dim, "structured": structured,
            "time_ms": time_ms, "record_size": dt.itemsize}, recs
