"""ctypes binding of libsph.so (the gfx950 HIP library) and small device helpers.

The library is mandatory: importing a module that needs it fails loudly when
``libsph.so`` is missing or no HIP device is visible.  There is no CPU path.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_double, c_float, c_int, c_int32, c_int64
from ctypes import c_void_p

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SPH_LIB", os.path.join(_HERE, "libsph.so"))

# PartitionErrorCode bits (jax_md/partition.py:434-457)
NEIGHBOR_LIST_OVERFLOW = 1
CELL_LIST_OVERFLOW = 2
CELL_SIZE_TOO_SMALL = 4
MALFORMED_BOX = 8

KERNEL_IDS = {"CSK": 0, "QSK": 1, "WC2K": 2, "WC4K": 3, "WC6K": 4, "GK": 5, "SGK": 6}


class NLConfig(Structure):
    _fields_ = [
        ("dim", c_int32), ("n", c_int32), ("box", c_float * 3), ("cutoff", c_float),
        ("cutoff_sq", c_float), ("use_cells", c_int32), ("cps", c_int32 * 3),
        ("cell_size", c_float * 3), ("cell_capacity", c_int32), ("mask_self", c_int32),
        ("ordered", c_int32), ("ell_width", c_int32),
    ]


class SolverParams(Structure):
    _fields_ = [
        ("dim", c_int32), ("kernel", c_int32), ("solver", c_int32),
        ("density_evolution", c_int32), ("bc_trick", c_int32), ("free_slip", c_int32),
        ("heat_conduction", c_int32), ("eos_kind", c_int32),
        ("one_over_h", c_float), ("sigma", c_float),
        ("dt", c_float), ("c_ref", c_float), ("eta_limiter", c_float),
        ("artificial_alpha", c_float), ("av_h", c_float),
        ("p_ref", c_float), ("rho_ref", c_float), ("p_bg", c_float), ("gamma", c_float),
        ("inv_gamma", c_float), ("riemann_k", c_float),
        ("box", c_float * 3), ("g_const", c_float * 3),
        ("tvf_stress", c_int32), ("rho_renorm", c_int32),
        ("diff_delta", c_float), ("diff_alpha", c_float), ("support", c_float),
        ("pad_", c_int32),
    ]


_IN_FIELDS = ["r", "u", "v", "dudt", "dvdt", "rho", "p", "mass", "eta", "T", "dTdt", "kappa",
              "Cp", "nw", "drhodt", "g_ext", "tag"]
_OUT_FIELDS = ["u", "v", "dudt", "dvdt", "rho", "p", "T", "dTdt", "drhodt"]


class StateIn(Structure):
    _fields_ = [(k, c_void_p) for k in _IN_FIELDS]


class StateOut(Structure):
    _fields_ = [(k, c_void_p) for k in _OUT_FIELDS]


class BCSpec(Structure):
    _fields_ = [("kind", c_int32), ("u_lid", c_float * 3)]


class CotangentIn(Structure):
    _fields_ = [("r", c_void_p), ("u", c_void_p), ("a", c_void_p), ("v", c_void_p),
                ("b", c_void_p), ("rho", c_void_p)]


class CotangentOut(Structure):
    _fields_ = [("r", c_void_p), ("u", c_void_p), ("a", c_void_p), ("b", c_void_p),
                ("rho", c_void_p)]


class EnsembleMember(Structure):
    """One member of sph_ensemble_create: the arguments of its sph_step on bank 2."""
    _fields_ = [("ws", c_void_p), ("prm", POINTER(SolverParams)), ("dt", c_float),
                ("tvf_coef", c_float), ("bc", POINTER(BCSpec)), ("in_", POINTER(StateIn)),
                ("r_out", c_void_p), ("out", POINTER(StateOut)), ("prev_info", c_void_p),
                ("max_occupancy", c_int64), ("info", c_void_p)]


class SphError(RuntimeError):
    """A libsph call returned an error code."""


_lib = None


def lib():
    """Load libsph.so once; raise if it is missing (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libsph.so not found at {LIB_PATH}; build it with "
            "`python jax-sph_amd/build.py` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64p = c_void_p, POINTER(c_int64)
    sig = {
        "sph_last_error": (c_char_p, []),
        "sph_version": (c_int, []),
        "sph_ws_create": (c_int, [POINTER(NLConfig), POINTER(c_void_p)]),
        "sph_ws_destroy": (c_int, [vp]),
        "sph_ws_bytes": (c_int64, [vp]),
        "sph_ws_tune": (c_int, [vp, c_char_p, c_int64]),
        "sph_nl_fill_max": (c_int, [c_int32, c_int32, vp, POINTER(c_float), POINTER(c_int32), vp,
                                    vp, vp]),
        "sph_nl_fill_max_static": (c_int, [c_int32, c_int32, vp, POINTER(c_float),
                                           POINTER(c_int32), vp, vp]),
        "sph_cell_list_work_bytes": (ctypes.c_size_t, [c_int64, c_int64]),
        "sph_cell_list": (c_int, [c_int32, c_int32, vp, POINTER(c_float), POINTER(c_int32),
                                  c_int64, vp, vp, c_int32, vp, vp, vp, vp, vp,
                                  ctypes.c_size_t, vp, vp]),
        "sph_nl_needs_rebuild": (c_int, [vp, vp, vp, c_float, vp, vp]),
        "sph_nl_build": (c_int, [vp, c_int, vp, vp, c_int64, vp, vp]),
        "sph_nl_finalize": (c_int, [vp, c_int64, vp, vp]),
        "sph_nl_scan_budget": (c_int, [vp, c_int, c_int64, c_int64, vp, vp]),
        "sph_nl_write_idx": (c_int, [vp, c_int, vp, c_int64, vp, vp]),
        "sph_skin_enable": (c_int, [vp, c_float, c_int32]),
        "sph_nl_write_idx_skin": (c_int, [vp, vp, vp, c_int64, vp]),
        "sph_kernel_eval": (c_int, [c_int32, c_int32, c_float, c_float, vp, c_int64, vp, vp, vp]),
        "sph_forward": (c_int, [vp, c_int, POINTER(SolverParams), POINTER(StateIn),
                                POINTER(StateOut), vp]),
        "sph_step": (c_int, [vp, c_int, POINTER(SolverParams), c_float, c_float,
                             POINTER(BCSpec), POINTER(StateIn), vp, POINTER(StateOut), vp,
                             c_int64, vp, vp]),
        "sph_integrate": (c_int, [c_int32, c_int64, vp, vp, vp, vp, vp, vp, vp, c_float, c_float,
                                  POINTER(c_float), vp]),
        "sph_step_vjp": (c_int, [vp, c_int, POINTER(SolverParams), c_float, c_float,
                                 POINTER(BCSpec), POINTER(StateIn), POINTER(CotangentIn),
                                 POINTER(CotangentOut), c_int32, vp]),
        "sph_ensemble_create": (c_int, [POINTER(EnsembleMember), c_int32, POINTER(c_void_p)]),
        "sph_ensemble_step": (c_int, [vp, vp]),
        "sph_ensemble_tune": (c_int, [vp, c_char_p, c_int64]),
        "sph_ensemble_info": (c_int, [vp, i64p, i64p, i64p]),
        "sph_ensemble_destroy": (c_int, [vp]),
        "sph_prof_enable": (c_int, [vp, c_int]),
        "sph_prof_reset": (c_int, [vp]),
        "sph_prof_read": (c_int, [vp, c_char_p, c_int, POINTER(c_double), i64p, c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int, what: str = "libsph"):
    if rc != 0:
        msg = lib().sph_last_error()
        raise SphError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("jax_sph_amd needs a HIP device (MI355X); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> c_void_p:
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor | None) -> c_void_p | None:
    if t is None:
        return None
    return c_void_p(t.data_ptr())


def as_f32(x, dev=None) -> torch.Tensor:
    """Contiguous float32 tensor on the HIP device (copies host data once)."""
    dev = dev or device()
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.as_tensor(x)
    if t.dtype != torch.float32 or t.device != dev or not t.is_contiguous():
        t = t.to(device=dev, dtype=torch.float32).contiguous()
    return t


def as_i32(x, dev=None) -> torch.Tensor:
    dev = dev or device()
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(x)
    if t.dtype != torch.int32 or t.device != dev or not t.is_contiguous():
        t = t.to(device=dev, dtype=torch.int32).contiguous()
    return t


SKIN_BANK = 2


class Workspace:
    """Owns one ``sph_ws`` (cell arrays, two banks of exact pairs, optional
    skin lists as bank 2, sorted records)."""

    def __init__(self, cfg: NLConfig):
        self.cfg = cfg
        h = c_void_p()
        check(lib().sph_ws_create(ctypes.byref(cfg), ctypes.byref(h)), "sph_ws_create")
        self.handle = h
        self.generation = [0, 0, 0]
        self.skin = 0.0

    def enable_skin(self, skin: float, width: int) -> bool:
        """Skin lists for the fused step (sph_skin_enable); False if this
        geometry cannot use them."""
        rc = lib().sph_skin_enable(self.handle, float(skin), int(width))
        if rc < 0:
            check(rc, "sph_skin_enable")
        if rc == 1:
            self.skin = float(skin)
        return rc == 1

    @property
    def skin_on(self) -> bool:
        return self.skin > 0.0

    def tune(self, key: str, value: int):
        """Launch-shape knob of the library (sph_ws_tune), e.g. ``xcd_remap``."""
        check(lib().sph_ws_tune(self.handle, key.encode(), int(value)), "sph_ws_tune")

    @property
    def bytes(self) -> int:
        return int(lib().sph_ws_bytes(self.handle))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            try:
                torch.cuda.synchronize()
            except Exception:
                pass
            _lib.sph_ws_destroy(h)
            self.handle = None

    # profiling hooks
    def prof_enable(self, on=True):
        check(lib().sph_prof_enable(self.handle, 1 if on else 0))

    def prof_reset(self):
        check(lib().sph_prof_reset(self.handle))

    def prof_read(self):
        cap = 64
        names = ctypes.create_string_buffer(8192)
        ms = (c_double * cap)()
        cnt = (c_int64 * cap)()
        n = lib().sph_prof_read(self.handle, names, 8192, ms, cnt, cap)
        raw = names.raw.split(b"\0")
        return {raw[k].decode(): (ms[k], int(cnt[k])) for k in range(min(n, cap))}
