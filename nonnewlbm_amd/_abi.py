"""ctypes mirror of include/lbm355.h and the loader of the native library.

The product path is the HIP library ``nonnewlbm_amd/liblbm355.so``, built
in-tree by ``__graft_entry__.build()``.  There is no CPU fallback: if the
library is missing, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "liblbm355.so"

# collision operators
CUM, BGK, MRT, KBC = 0, 1, 2, 3
COLLISIONS = {"cumulant": CUM, "cum": CUM, "bgk": BGK, "mrt": MRT, "mrt_les": MRT, "kbc": KBC, "kbc_n4": KBC}
COLLISION_IDS = {CUM: "CUM", BGK: "BGK", MRT: "MRT_LES", KBC: "KBC_N4"}
# storage
F64, F32, F16_OFFSET, BF16_OFFSET = 0, 1, 2, 3
STORAGES = {"fp64": F64, "f64": F64, "fp32": F32, "f32": F32, "fp16": F16_OFFSET, "f16": F16_OFFSET,
            "fp16-offset": F16_OFFSET, "bf16": BF16_OFFSET, "bf16-offset": BF16_OFFSET}
STORAGE_BYTES = {F64: 8, F32: 4, F16_OFFSET: 2, BF16_OFFSET: 2}
STORAGE_NAMES = {F64: "f64", F32: "f32", F16_OFFSET: "f16-offset", BF16_OFFSET: "bf16-offset"}
# Carreau-Yasuda coupling
CY_OFF, CY_NEQ, CY_FD_FORCE = 0, 1, 2
CY_MODELS = {"off": CY_OFF, "neq": CY_NEQ, "fd-force": CY_FD_FORCE}
# inflow profiles
PROFILE_PLUG, PROFILE_PARABOLIC = 0, 1
# face boundary kinds (x-, x+, y-, y+, z-, z+)
FACE_TCPC, FACE_WALL, FACE_INFLOW, FACE_OUTFLOW = 0, 1, 2, 3
# the 64x32x32 channel: inflow at x-min, pressure outlet at x-max, walls elsewhere
CHANNEL_FACES = (FACE_INFLOW, FACE_OUTFLOW, FACE_WALL, FACE_WALL, FACE_WALL, FACE_WALL)

# fields
FIELD = {
    "rho": 0, "vx": 1, "vy": 2, "vz": 3,
    "sum_rux": 4, "sum_ruy": 5, "sum_ruz": 6, "sum_rruxx": 7, "sum_rruyy": 8, "sum_rruzz": 9,
    "s11": 10, "s12": 11, "s13": 12, "s22": 13, "s32": 14, "s33": 15,
    "label": 16, "wallmask": 17, "fx": 18, "fy": 19, "fz": 20,
}

OK = 0
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ETERMINATE", -5: "ENOTFOUND", -6: "EBUSY"}
EINVAL, ENOMEM, EDEVICE, ETERMINATE, ENOTFOUND, EBUSY = -1, -2, -3, -4, -5, -6


class CaseDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("geometry", C.POINTER(C.c_int32)),
        ("dl", C.c_double), ("dt", C.c_double),
        ("collision", C.c_int32), ("storage", C.c_int32),
        ("nu_ref", C.c_double),
        ("cy_enabled", C.c_int32), ("cy_model", C.c_int32),
        ("cy_nu0", C.c_double), ("cy_lambda", C.c_double), ("cy_a", C.c_double), ("cy_n", C.c_double),
        ("ivc_flux", C.c_double), ("svc_flux", C.c_double),
        ("profile", C.c_int32),
        ("ivc_umax", C.c_double), ("svc_umax", C.c_double),
        ("ivc_center", C.c_double * 2), ("svc_center", C.c_double * 2),
        ("ivc_radius", C.c_double), ("svc_radius", C.c_double),
        ("pulse_amp", C.c_double), ("pulse_period", C.c_double),
        ("stats_every", C.c_int32),
        ("face_bc", C.c_int32 * 6),
        ("reserved", C.c_int32 * 1),
    ]


class CaseInfo(C.Structure):
    _fields_ = [
        ("fluid", C.c_int64), ("wall", C.c_int64),
        ("area_ivc", C.c_int64), ("area_svc", C.c_int64), ("area_out5", C.c_int64), ("area_out6", C.c_int64),
        ("input_label3", C.c_int64), ("input_label4", C.c_int64), ("input_label5", C.c_int64),
        ("input_label6", C.c_int64),
        ("active_cells", C.c_int64), ("bricks_active", C.c_int64), ("bricks_stored", C.c_int64),
        ("device_bytes", C.c_int64), ("steps", C.c_int64),
        ("terminate", C.c_int32), ("reserved", C.c_int32 * 5),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class StateDesc(C.Structure):
    _fields_ = [
        ("phys_final_time", C.c_double),
        ("counter_count", C.c_int32 * 11), ("counter_period", C.c_double * 11),
        ("stat_prev_iter", C.c_int64), ("iterations", C.c_int64), ("reserved", C.c_int32 * 4),
    ]


class OutputDesc(C.Structure):
    _fields_ = [
        ("phys_density", C.c_double), ("phys_nu_inf", C.c_double), ("time", C.c_double),
        ("cycle", C.c_int32), ("stat_steps", C.c_int32), ("material_cy", C.c_int32), ("reserved", C.c_int32 * 5),
    ]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3
REQ_STATES = {REQ_QUEUED: "queued", REQ_RUNNING: "running", REQ_DONE: "done", REQ_FAILED: "failed"}


class RequestStatus(C.Structure):
    _fields_ = [
        ("state", C.c_int32), ("error", C.c_int32),
        ("steps_done", C.c_int64), ("steps", C.c_int64),
        ("active_cells", C.c_int64), ("device_bytes", C.c_int64),
        ("t_submit", C.c_double), ("t_start", C.c_double), ("t_first_report", C.c_double), ("t_done", C.c_double),
        ("flux", C.c_double * 4), ("mean_flux", C.c_double * 4),
        ("message", C.c_char * 128),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["flux"] = list(self.flux)
        d["mean_flux"] = list(self.mean_flux)
        d["message"] = self.message.decode(errors="replace")
        d["state_name"] = REQ_STATES.get(self.state, "?")
        return d


class PoolInfo(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("queued", C.c_int64), ("running", C.c_int64), ("done", C.c_int64),
        ("failed", C.c_int64), ("budget_bytes", C.c_int64), ("used_bytes", C.c_int64),
        ("cell_updates", C.c_int64), ("ticks", C.c_int64), ("busy_seconds", C.c_double), ("clock", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def make_desc(geometry_flat: np.ndarray, dims, **kw) -> tuple[CaseDesc, np.ndarray]:
    """Build a CaseDesc; returns (desc, keepalive geometry array)."""
    g = np.ascontiguousarray(geometry_flat, dtype=np.int32)
    d = CaseDesc()
    d.nx, d.ny, d.nz = (int(v) for v in dims)
    d.geometry = g.ctypes.data_as(C.POINTER(C.c_int32))
    d.dl = kw.pop("dl", 1.0e-3)
    d.dt = kw.pop("dt", 1.0)
    d.collision = kw.pop("collision", CUM)
    d.storage = kw.pop("storage", F32)
    d.nu_ref = kw.pop("nu_ref", 1.0e-4)
    d.cy_enabled = int(kw.pop("cy_enabled", 0))
    d.cy_model = kw.pop("cy_model", CY_OFF)
    d.cy_nu0 = kw.pop("cy_nu0", 0.0)
    d.cy_lambda = kw.pop("cy_lambda", 1.0)
    d.cy_a = kw.pop("cy_a", 2.0)
    d.cy_n = kw.pop("cy_n", 1.0)
    d.ivc_flux = kw.pop("ivc_flux", 0.0)
    d.svc_flux = kw.pop("svc_flux", 0.0)
    d.profile = kw.pop("profile", PROFILE_PLUG)
    d.ivc_umax = kw.pop("ivc_umax", 0.0)
    d.svc_umax = kw.pop("svc_umax", 0.0)
    ic = kw.pop("ivc_center", (-1.0, -1.0))
    sc = kw.pop("svc_center", (-1.0, -1.0))
    d.ivc_center[0], d.ivc_center[1] = ic
    d.svc_center[0], d.svc_center[1] = sc
    d.ivc_radius = kw.pop("ivc_radius", 0.0)
    d.svc_radius = kw.pop("svc_radius", 0.0)
    d.pulse_amp = kw.pop("pulse_amp", 0.0)
    d.pulse_period = kw.pop("pulse_period", 0.0)
    d.stats_every = kw.pop("stats_every", 1)
    faces = kw.pop("face_bc", (0, 0, 0, 0, 0, 0))
    for i in range(6):
        d.face_bc[i] = int(faces[i])
    if kw:
        raise TypeError(f"unknown descriptor fields {sorted(kw)}")
    return d, g


_lib = None


def load() -> C.CDLL:
    """Load liblbm355.so (the HIP library).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("LBM355_LIB", LIB_PATH))
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: the HIP library is not built (run __graft_entry__.build()); "
            "there is no CPU fallback")
    lib = C.CDLL(str(path))
    P = C.c_void_p
    sig = {
        "lbm355_version": (C.c_char_p, []),
        "lbm355_last_error": (C.c_char_p, []),
        "lbm355_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "lbm355_case_create": (C.c_int, [C.POINTER(CaseDesc), C.POINTER(P)]),
        "lbm355_case_destroy": (C.c_int, [P]),
        "lbm355_case_info_get": (C.c_int, [P, C.POINTER(CaseInfo)]),
        "lbm355_case_set_inflow": (C.c_int, [P, C.c_double, C.c_double]),
        "lbm355_case_step": (C.c_int, [P, C.c_int64]),
        "lbm355_case_sync": (C.c_int, [P]),
        "lbm355_case_set_shape": (C.c_int, [P, C.c_int]),
        "lbm355_case_set_nontemporal": (C.c_int, [P, C.c_int]),
        "lbm355_case_set_occupancy": (C.c_int, [P, C.c_int]),
        "lbm355_case_autotune": (C.c_int, [P, C.POINTER(C.c_int)]),
        "lbm355_case_get_field": (C.c_int, [P, C.c_int, C.POINTER(C.c_double), C.c_size_t]),
        "lbm355_case_get_populations": (C.c_int, [P, C.POINTER(C.c_double), C.c_size_t]),
        "lbm355_case_compute_strain": (C.c_int, [P]),
        "lbm355_case_flux": (C.c_int, [P, C.c_int, C.POINTER(C.c_double)]),
        "lbm355_case_reset_stats": (C.c_int, [P]),
        "lbm355_case_write_vtk": (C.c_int, [P, C.c_char_p, C.POINTER(OutputDesc)]),
        "lbm355_batch_create": (C.c_int, [C.POINTER(P), C.c_int, C.c_int, C.POINTER(P)]),
        "lbm355_batch_step": (C.c_int, [P, C.c_int64]),
        "lbm355_batch_destroy": (C.c_int, [P]),
        "lbm355_batch_set_shape": (C.c_int, [P, C.c_int]),
        "lbm355_batch_set_occupancy": (C.c_int, [P, C.c_int]),
        "lbm355_batch_set_group_bytes": (C.c_int, [P, C.c_int64]),
        "lbm355_batch_autotune": (C.c_int, [P, C.POINTER(C.c_int)]),
        "lbm355_case_compare": (C.c_int, [P, P, C.POINTER(C.c_double)]),
        "lbm355_case_compare_noalloc": (C.c_int, [P, P, C.POINTER(C.c_double)]),
        "lbm355_batch_add": (C.c_int, [P, P]),
        "lbm355_batch_remove": (C.c_int, [P, P]),
        "lbm355_batch_size": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
        "lbm355_batch_set_per_wave": (C.c_int, [P, C.c_int]),
        "lbm355_batch_config": (C.c_int, [P, C.POINTER(C.c_int)]),
        "lbm355_case_footprint": (C.c_int, [C.POINTER(CaseDesc), C.POINTER(C.c_int64)]),
        "lbm355_case_labels": (C.c_int, [C.POINTER(CaseDesc), C.POINTER(C.c_int32), C.POINTER(CaseInfo)]),
        "lbm355_case_time_kernel": (C.c_int, [P, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "lbm355_copy_bandwidth": (C.c_int, [C.c_int64, C.c_int, C.POINTER(C.c_double)]),
        "lbm355_case_save_state": (C.c_int, [P, C.c_char_p, C.POINTER(StateDesc)]),
        "lbm355_case_load_state": (C.c_int, [P, C.c_char_p, C.POINTER(StateDesc)]),
        "lbm355_pool_create": (C.c_int, [C.c_int64, C.c_int, C.c_int, C.POINTER(P)]),
        "lbm355_pool_submit": (C.c_int, [P, C.POINTER(CaseDesc), C.c_int64, C.c_int64, C.c_int,
                                         C.POINTER(C.c_int64)]),
        "lbm355_pool_poll": (C.c_int, [P, C.c_int64, C.POINTER(RequestStatus)]),
        "lbm355_pool_wait": (C.c_int, [P, C.c_int64, C.c_double, C.POINTER(RequestStatus)]),
        "lbm355_pool_drain": (C.c_int, [P, C.c_double]),
        "lbm355_pool_take": (C.c_int, [P, C.c_int64, C.POINTER(P)]),
        "lbm355_pool_stats": (C.c_int, [P, C.POINTER(PoolInfo)]),
        "lbm355_pool_destroy": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


# every symbol include/lbm355.h declares
EXPORTED = [
    "lbm355_version", "lbm355_last_error", "lbm355_device_count",
    "lbm355_case_create", "lbm355_case_destroy", "lbm355_case_info_get", "lbm355_case_set_inflow",
    "lbm355_case_step", "lbm355_case_sync", "lbm355_case_set_shape", "lbm355_case_set_nontemporal", "lbm355_case_set_occupancy", "lbm355_case_autotune", "lbm355_case_get_field", "lbm355_case_get_populations",
    "lbm355_case_compute_strain", "lbm355_case_flux", "lbm355_case_reset_stats", "lbm355_case_write_vtk",
    "lbm355_batch_create", "lbm355_batch_step", "lbm355_batch_destroy", "lbm355_batch_set_shape",
    "lbm355_batch_autotune", "lbm355_batch_set_occupancy", "lbm355_batch_set_group_bytes", "lbm355_case_footprint", "lbm355_case_labels", "lbm355_case_time_kernel",
    "lbm355_batch_add", "lbm355_batch_remove", "lbm355_batch_size", "lbm355_batch_set_per_wave",
    "lbm355_batch_config", "lbm355_case_compare", "lbm355_case_compare_noalloc",
    "lbm355_copy_bandwidth",
    "lbm355_case_save_state", "lbm355_case_load_state",
    "lbm355_pool_create", "lbm355_pool_submit", "lbm355_pool_poll", "lbm355_pool_wait", "lbm355_pool_drain",
    "lbm355_pool_take", "lbm355_pool_stats", "lbm355_pool_destroy",
]


class LbmError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        super().__init__(f"{what}: {ERRORS.get(code, code)}")


def check(code: int, what: str = "lbm355") -> int:
    if code != OK:
        msg = load().lbm355_last_error()
        raise LbmError(code, f"{what} failed ({msg.decode() if msg else ''})")
    return code
