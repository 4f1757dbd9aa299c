"""State files through the C ABI (include/xdem_io.h): the reference's
part%04d.bin and <name>_final.bfr formats (dem_simulation::saveResult /
saveFinalResult, source/simulation/dem_simulation.cpp:491-583; read back by
startingModel::setDEMData, source/startingModel.cpp:32-67) and the
checkpoint container check.  The simulation and batch checkpoints are methods
of metakonga_amd.xdem.Simulation and metakonga_amd.batch.Batch.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi


def _p(path) -> bytes:
    return os.fsencode(path)


def write_part(path, ct: float, pos4: np.ndarray, precision: int = 2, lib=None):
    """part%04d.bin: double ct, uint32 np, pos[np*4] (FP64, or FP32 for precision 1)."""
    lib = lib or abi.load()
    dt = np.float64 if precision == 2 else np.float32
    pos = np.ascontiguousarray(pos4, dtype=dt).reshape(-1, 4)
    f = lib.xdem_write_part_f64 if precision == 2 else lib.xdem_write_part_f32
    abi.check(lib, f(_p(path), float(ct), pos.shape[0], C.c_void_p(pos.ctypes.data)))


def read_part(path, lib=None):
    """(ct, pos4 (np,4) float64, precision)."""
    lib = lib or abi.load()
    ct, n, prec = C.c_double(), C.c_uint32(), C.c_int32()
    abi.check(lib, lib.xdem_read_part(_p(path), C.byref(ct), C.byref(n), C.byref(prec), None, 0))
    pos = np.empty((n.value, 4))
    abi.check(lib, lib.xdem_read_part(_p(path), None, None, None, abi.ptr(pos), n.value))
    return ct.value, pos, prec.value


def write_bfr(path, ct: float, pos4, vel3, avel3, precision: int = 2, lib=None):
    """<name>_final.bfr: double ct, int32 precision, int32 flag = 1, pos, vel, avel."""
    lib = lib or abi.load()
    dt = np.float64 if precision == 2 else np.float32
    pos = np.ascontiguousarray(pos4, dtype=dt).reshape(-1, 4)
    n = pos.shape[0]
    vel = np.ascontiguousarray(vel3, dtype=dt).reshape(n, 3)
    av = np.ascontiguousarray(avel3, dtype=dt).reshape(n, 3)
    f = lib.xdem_write_bfr_f64 if precision == 2 else lib.xdem_write_bfr_f32
    abi.check(lib, f(_p(path), float(ct), n, *(C.c_void_p(a.ctypes.data) for a in (pos, vel, av))))


def read_bfr(path, np_: int, lib=None):
    """(ct, pos4, vel3, avel3, precision) of a .bfr holding np_ particles."""
    lib = lib or abi.load()
    ct, prec = C.c_double(), C.c_int32()
    pos, vel, av = np.empty((np_, 4)), np.empty((np_, 3)), np.empty((np_, 3))
    abi.check(lib, lib.xdem_read_bfr(_p(path), np_, C.byref(ct), C.byref(prec), abi.ptr(pos),
                                     abi.ptr(vel), abi.ptr(av)))
    return ct.value, pos, vel, av, prec.value


def verify_checkpoint(path, lib=None) -> int:
    """Checks a checkpoint container; returns its kind (1 simulation, 2 batch)."""
    lib = lib or abi.load()
    kind = C.c_uint32()
    abi.check(lib, lib.xdem_checkpoint_verify(_p(path), C.byref(kind)))
    return kind.value
