"""ctypes binding of oracle/sph_cpu.c, the OpenMP C restatement of one
si_euler step (standard WCSPH, density summation, Tait EOS, QSK or WC2K,
optional transport velocity, periodic box) used as the benchmark's
multi-core CPU baseline.

TEST INFRASTRUCTURE ONLY: loaded by ``bench.py``'s ``cpu_baseline`` leg and
``tests/``, never by the product package.  ``build()`` compiles it with gcc
(``__graft_entry__.build()`` calls it).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "sph_cpu.c")
LIB = os.path.join(HERE, "libsph_cpu.so")
# x86-64-v2 (SSE4.2) runs on any host this is copied to; no FMA contraction, so
# the neighbour predicate sums squares exactly as the reference does
FLAGS = ["-O3", "-fopenmp", "-march=x86-64-v2", "-ffp-contract=off", "-fPIC", "-shared"]


def build(verbose=False) -> str:
    cmd = ["gcc", *FLAGS, "-o", LIB + ".tmp", SRC, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


class Params(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("n", ctypes.c_int32), ("box", ctypes.c_float * 3),
                ("cutoff", ctypes.c_float), ("kernel", ctypes.c_int32),
                ("one_over_h", ctypes.c_float), ("sigma", ctypes.c_float),
                ("p_ref", ctypes.c_float), ("rho_ref", ctypes.c_float),
                ("p_bg", ctypes.c_float), ("gamma", ctypes.c_float), ("dt", ctypes.c_float),
                ("ctvf", ctypes.c_float), ("g", ctypes.c_float * 3),
                ("threads", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} not built (oracle.cpu_port.build())")
        L = ctypes.CDLL(LIB)
        fp = ctypes.POINTER(ctypes.c_float)
        L.sph_cpu_step.argtypes = [ctypes.POINTER(Params)] + [fp] * 9
        L.sph_cpu_step.restype = ctypes.c_int
        L.sph_cpu_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


class CPUStep:
    """si_euler(tvf, WCSPH(...).forward) of the oracle as one C call per step,
    for the configurations above; ``kernel`` is an oracle ``Kernel``, ``eos``
    an oracle ``TaitEoS``; the state dict is updated in place."""

    def __init__(self, box, kernel, eos, dt, tvf, g=(0.0, 0.0, 0.0), threads=0):
        if kernel.name not in ("QSK", "WC2K"):
            raise NotImplementedError(kernel.name)
        f = np.float32
        p = Params()
        p.dim = kernel.dim
        for d in range(3):
            p.box[d] = float(box[d]) if d < kernel.dim else 1.0
            p.g[d] = float(g[d]) if d < len(g) else 0.0
        p.cutoff = float(f(kernel.cutoff))
        p.kernel = 1 if kernel.name == "QSK" else 2
        p.one_over_h = float(f(kernel.one_over_h))
        p.sigma = float(f(kernel.sigma))
        p.p_ref, p.rho_ref = float(f(eos.p_ref)), float(f(eos.rho_ref))
        p.p_bg, p.gamma = float(f(eos.p_bg)), float(eos.gamma)
        dtf = f(dt)
        p.dt = float(dtf)
        p.ctvf = float(f(f(tvf * 0.5) * dtf))
        p.threads = int(threads)
        self.p = p

    def __call__(self, st):
        L = lib()
        self.p.n = len(st["r"])
        keys = ("r", "u", "v", "dudt", "dvdt", "rho", "p", "mass", "eta")
        for k in keys:
            if st[k].dtype != np.float32 or not st[k].flags.c_contiguous:
                st[k] = np.ascontiguousarray(st[k], dtype=np.float32)
        ptr = [st[k].ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for k in keys]
        rc = L.sph_cpu_step(ctypes.byref(self.p), *ptr)
        if rc != 0:
            raise ValueError("sph_cpu_step: unsupported configuration")
        return st
