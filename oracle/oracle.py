"""TEST INFRASTRUCTURE — ctypes handle on the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only to check or time against the product; the product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
sys.path.insert(0, os.path.join(_REPO, "endobeams.jl_amd"))

from endobeams_amd import abi  # noqa: E402
from endobeams_amd.engine import Engine  # noqa: E402

LIB = os.path.join(_HERE, "liboracle.so")
# the same restatement with its element and contact loops over OpenMP threads (timing only)
LIB_OMP = os.path.join(_HERE, "liboracle_omp.so")
_libs = {}


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def load(timing=False):
    path = LIB_OMP if timing else LIB
    if path not in _libs:
        if not os.path.exists(path):
            build()
        _lib = C.CDLL(path)
        _libs[path] = _lib
        dp, lp = C.POINTER(C.c_double), C.POINTER(C.c_int64)
        _lib.ebo_compute_beam.argtypes = [dp, dp, C.c_int, dp, dp, dp]
        _lib.ebo_compute_contact_beam.argtypes = [C.c_void_p, dp, C.c_double, dp]
        _lib.ebo_sdf_eval.argtypes = [C.c_void_p, dp, dp]
        _lib.ebo_sdf_coefs.argtypes = [C.c_void_p, dp, C.c_int64]
        _lib.ebo_contact_gap.argtypes = [C.c_void_p, dp, C.c_double, dp]
        _lib.ebo_rot_util.argtypes = [C.c_int, dp, dp]
        _lib.ebo_regularization.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        _lib.ebo_band_info.argtypes = [C.c_void_p, lp, lp]
    return _libs[path]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleEngine(Engine):
    """Same interface as endobeams_amd.HipEngine, computed on the CPU by the restatement."""
    prefix = "ebo_"

    def __init__(self, conf, params, inter=None, precision=abi.EB_PRECISION_FP64, timing=False):
        super().__init__(load(timing), conf, params, inter, precision)

    def band_info(self):
        kl, ku = C.c_int64(0), C.c_int64(0)
        self.lib.ebo_band_info(self.h, C.byref(kl), C.byref(ku))
        return kl.value, ku.value

    def contact_beam(self, elem82, radius):
        out = np.zeros(302)
        e = np.ascontiguousarray(elem82, dtype=np.float64)
        self._check(self.lib.ebo_compute_contact_beam(self.h, _dp(e), radius, _dp(out)), "compute_contact_beam")
        return {"energy": out[0], "active": int(out[1]), "Tc": out[2:14],
                "K": out[14:158].reshape(12, 12, order="F"), "C": out[158:302].reshape(12, 12, order="F")}

    def sdf_eval(self, p):
        out = np.zeros(13)
        pp = np.ascontiguousarray(p, dtype=np.float64)
        self._check(self.lib.ebo_sdf_eval(self.h, _dp(pp), _dp(out)), "sdf_eval")
        return out[0], out[1:4], out[4:13].reshape(3, 3, order="F")

    def sdf_coefs(self, n):
        out = np.zeros((n[0] + 2) * (n[1] + 2) * (n[2] + 2))
        self._check(self.lib.ebo_sdf_coefs(self.h, _dp(out), out.size), "sdf_coefs")
        return out.reshape(n[2] + 2, n[1] + 2, n[0] + 2)

    def contact_gap(self, p, radius):
        out = np.zeros(14)
        pp = np.ascontiguousarray(p, dtype=np.float64)
        self._check(self.lib.ebo_contact_gap(self.h, _dp(pp), radius, _dp(out)), "contact_gap")
        return bool(out[0]), out[1], out[2:5], out[5:14].reshape(3, 3, order="F")


def pack_elem(X1, X2, u1, u2, ud1, ud2, udd1, udd2, wd1, wd2, wdd1, wdd2, R1, R2, dR1, dR2, l0, Re0):
    """Packs one element's inputs in the 82-double layout of ebo_compute_beam (matrices given as
    3×3 numpy arrays are stored column-major)."""
    vecs = [X1, X2, u1, u2, ud1, ud2, udd1, udd2, wd1, wd2, wdd1, wdd2]
    mats = [R1, R2, dR1, dR2]
    parts = [np.asarray(v, dtype=np.float64).reshape(3) for v in vecs]
    parts += [np.asarray(m, dtype=np.float64).reshape(3, 3).reshape(-1, order="F") for m in mats]
    parts += [np.array([l0], dtype=np.float64), np.asarray(Re0, dtype=np.float64).reshape(3, 3).reshape(-1, order="F")]
    return np.concatenate(parts)


def compute_beam(elem82, props14, gw=(5 / 9, 8 / 9, 5 / 9), gz=(-np.sqrt(3 / 5), 0.0, np.sqrt(3 / 5))):
    lib = load()
    e = np.ascontiguousarray(elem82, dtype=np.float64)
    p = np.ascontiguousarray(props14, dtype=np.float64)
    w = np.ascontiguousarray(gw, dtype=np.float64)
    z = np.ascontiguousarray(gz, dtype=np.float64)
    out = np.zeros(458)
    lib.ebo_compute_beam(_dp(e), _dp(p), len(w), _dp(w), _dp(z), _dp(out))
    return {"strain": out[0], "kinetic": out[1], "Tint": out[2:14], "Tk": out[14:26],
            "K": out[26:170].reshape(12, 12, order="F"), "M": out[170:314].reshape(12, 12, order="F"),
            "C": out[314:458].reshape(12, 12, order="F")}


def rot_util(op, x):
    lib = load()
    xi = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, order="F"))
    out = np.zeros(9)
    lib.ebo_rot_util(op, _dp(xi), _dp(out))
    return out


def rotation_matrix(v):
    return rot_util(0, v).reshape(3, 3, order="F")


def toangle(R):
    return rot_util(1, np.asarray(R).reshape(-1, order="F"))[:3]


def Ts_inv(v):
    return rot_util(2, v).reshape(3, 3, order="F")


def Ts(v):
    return rot_util(3, v).reshape(3, 3, order="F")


def eta_mu(v):
    return tuple(rot_util(4, v)[:2])


def get_Re0(X1, X2):
    return rot_util(5, np.concatenate([np.asarray(X1, float), np.asarray(X2, float)])).reshape(3, 3, order="F")


def regularization(g, rb, eta):
    lib = load()
    out = np.zeros(5)
    lib.ebo_regularization(g, rb, eta, _dp(out))
    return out
