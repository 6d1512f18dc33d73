"""Python face of liblbm355 (ctypes over include/lbm355.h).

``Case`` holds one simulation on the GPU: the sim_tcpc label setup, the sparse
brick index and the resident populations/statistics.  It mirrors what the
reference's StateTcpc/LBM pair exposes to the sim_tcpc driver
(solver/sim_tcpc/sim_tcpc.cu:438-838): per-step inflow hook, stepping,
macroscopic fields, statistics reset and the 3D VTK output.

There is no CPU fallback: every call goes to the HIP library and raises if it
is not built or the device call fails.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Sequence

import numpy as np

from . import _abi
from ._abi import FIELD, CaseDesc, CaseInfo, OutputDesc, check, load, make_desc


class Case:
    """One resident case.  ``geometry_flat`` is the x-fastest wall-mask label
    array of the geometry VTK; keyword arguments are the CaseDesc fields."""

    def __init__(self, geometry_flat: np.ndarray, dims: Sequence[int], **kw):
        self.lib = load()
        self.dims = tuple(int(v) for v in dims)
        self.desc, self._geom = make_desc(geometry_flat, dims, **kw)
        h = C.c_void_p()
        rc = self.lib.lbm355_case_create(C.byref(self.desc), C.byref(h))
        self.terminate = rc == _abi.ETERMINATE
        if rc not in (_abi.OK, _abi.ETERMINATE):
            check(rc, "lbm355_case_create")
        self.h = h
        self._geom = None  # the library copied what it needs

    @classmethod
    def from_handle(cls, handle: C.c_void_p, dims: Sequence[int]) -> "Case":
        """Wrap a case created elsewhere (a finished pool request); the Case
        owns it from now on."""
        self = cls.__new__(cls)
        self.lib = load()
        self.dims = tuple(int(v) for v in dims)
        self.desc = None
        self._geom = None
        self.terminate = False
        self.h = handle
        return self

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.lbm355_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stepping
    def set_inflow(self, ivc: float, svc: float) -> None:
        check(self.lib.lbm355_case_set_inflow(self.h, float(ivc), float(svc)), "set_inflow")

    def step(self, n: int = 1) -> None:
        check(self.lib.lbm355_case_step(self.h, int(n)), "step")

    def sync(self) -> None:
        check(self.lib.lbm355_case_sync(self.h), "sync")

    def set_shape(self, bricks_per_group: int) -> None:
        """64-cell bricks (wavefronts) per workgroup of the fused step (1, 2, 4 or 8)."""
        check(self.lib.lbm355_case_set_shape(self.h, int(bricks_per_group)), "case_set_shape")

    def set_nontemporal(self, mode: int) -> None:
        """-1 auto (by working set), 0 default cache policy, 1 non-temporal even step."""
        check(self.lib.lbm355_case_set_nontemporal(self.h, int(mode)), "case_set_nontemporal")

    def set_occupancy(self, waves_per_cu: int) -> None:
        """Cap the fused step at waves_per_cu wavefronts per CU (4..32); 0 = no cap."""
        check(self.lib.lbm355_case_set_occupancy(self.h, int(waves_per_cu)), "case_set_occupancy")

    def autotune(self) -> tuple[int, int, int]:
        """Time launch shapes x cache policy x wavefront cap on real steps
        (700 with statistics every step); keeps and returns the fastest
        (bricks_per_group, nontemporal, waves_per_cu)."""
        best = C.c_int()
        check(self.lib.lbm355_case_autotune(self.h, C.byref(best)), "case_autotune")
        return best.value & 0xFF, (best.value >> 8) & 0xFF, best.value >> 16

    def time_kernel(self, n: int) -> tuple[float, float]:
        """Run n steps; returns (ms per whole step, ms per fused-kernel launch)."""
        step_ms, kern_ms = C.c_double(), C.c_double()
        check(self.lib.lbm355_case_time_kernel(self.h, int(n), C.byref(step_ms), C.byref(kern_ms)), "time_kernel")
        return step_ms.value, kern_ms.value

    # -- outputs
    def info(self) -> dict:
        i = CaseInfo()
        check(self.lib.lbm355_case_info_get(self.h, C.byref(i)), "info")
        return i.as_dict()

    def field(self, name: str) -> np.ndarray:
        X, Y, Z = self.dims
        out = np.empty(X * Y * Z, dtype=np.float64)
        check(self.lib.lbm355_case_get_field(self.h, FIELD[name], out.ctypes.data_as(C.POINTER(C.c_double)),
                                             out.nbytes), f"get_field({name})")
        return out.reshape(Z, Y, X).transpose(2, 1, 0)

    def populations(self) -> np.ndarray:
        """[27, X, Y, Z] post-collision populations of the last step (A/B sense)."""
        X, Y, Z = self.dims
        out = np.empty(27 * X * Y * Z, dtype=np.float64)
        check(self.lib.lbm355_case_get_populations(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), out.nbytes),
              "get_populations")
        return out.reshape(27, Z, Y, X).transpose(0, 3, 2, 1)

    def compute_strain(self) -> None:
        check(self.lib.lbm355_case_compute_strain(self.h), "compute_strain")

    def flux(self, mean: bool = False) -> np.ndarray:
        out = (C.c_double * 4)()
        check(self.lib.lbm355_case_flux(self.h, 1 if mean else 0, out), "flux")
        return np.array(out[:], dtype=np.float64)

    def compare(self, ref: "Case") -> dict:
        """Relative L2 errors against ``ref`` (same geometry), reduced on the
        device from both cases' last macro output: velocity, rho - 1, and the
        four raw sums."""
        out = (C.c_double * 4)()
        check(self.lib.lbm355_case_compare(self.h, ref.h, out), "case_compare")
        s = list(out)
        return {"u_rel_l2": (s[0] / s[1]) ** 0.5 if s[1] > 0 else s[0] ** 0.5,
                "rho_rel_l2": (s[2] / s[3]) ** 0.5 if s[3] > 0 else s[2] ** 0.5, "sums": s}

    def reset_stats(self) -> None:
        check(self.lib.lbm355_case_reset_stats(self.h), "reset_stats")

    def save_state(self, directory: str | Path, final_time: float = 0.0, counters=None,
                   stat_prev_iter: int = 0) -> None:
        """Checkpoint in the reference's current_state layout; counters is a
        list of 11 (count, period) pairs (STAT_RESET ... VTK3DCUT)."""
        s = _abi.StateDesc()
        s.phys_final_time = float(final_time)
        for k in range(11):
            cnt, per = (counters[k] if counters else (0, -1.0))
            s.counter_count[k] = int(cnt)
            s.counter_period[k] = float(per)
        s.stat_prev_iter = int(stat_prev_iter)
        check(self.lib.lbm355_case_save_state(self.h, str(directory).encode(), C.byref(s)), "save_state")

    def load_state(self, directory: str | Path) -> dict:
        s = _abi.StateDesc()
        check(self.lib.lbm355_case_load_state(self.h, str(directory).encode(), C.byref(s)), "load_state")
        return {"iterations": s.iterations, "final_time": s.phys_final_time, "stat_prev_iter": s.stat_prev_iter,
                "counters": [(s.counter_count[k], s.counter_period[k]) for k in range(11)]}

    def write_vtk(self, path: str | Path, *, phys_density: float, phys_nu_inf: float, time: float, cycle: int,
                  stat_steps: int, material_cy: bool = False) -> None:
        o = OutputDesc()
        o.phys_density = phys_density
        o.phys_nu_inf = phys_nu_inf
        o.time = time
        o.cycle = int(cycle)
        o.stat_steps = int(stat_steps)
        o.material_cy = 1 if material_cy else 0
        check(self.lib.lbm355_case_write_vtk(self.h, str(path).encode(), C.byref(o)), "write_vtk")


class Batch:
    """Cases of one collision and storage type stepped by one launch per step."""

    def __init__(self, cases: Sequence[Case], graph: bool = True):
        self.lib = load()
        self.cases = list(cases)
        arr = (C.c_void_p * len(self.cases))(*[c.h for c in self.cases])
        h = C.c_void_p()
        check(self.lib.lbm355_batch_create(arr, len(self.cases), 1 if graph else 0, C.byref(h)), "batch_create")
        self.h = h

    def step(self, n: int) -> None:
        check(self.lib.lbm355_batch_step(self.h, int(n)), "batch_step")

    def set_shape(self, bricks_per_group: int) -> None:
        check(self.lib.lbm355_batch_set_shape(self.h, int(bricks_per_group)), "batch_set_shape")

    def set_occupancy(self, waves_per_cu: int) -> None:
        check(self.lib.lbm355_batch_set_occupancy(self.h, int(waves_per_cu)), "batch_set_occupancy")

    def set_group_bytes(self, nbytes: int) -> None:
        """Step members in groups of at most nbytes of device memory (0: one group)."""
        check(self.lib.lbm355_batch_set_group_bytes(self.h, int(nbytes)), "batch_set_group_bytes")

    def add(self, case: Case) -> None:
        """Make ``case`` a member (between batch calls); it keeps its own step count."""
        check(self.lib.lbm355_batch_add(self.h, case.h), "batch_add")
        self.cases.append(case)

    def remove(self, case: Case) -> None:
        check(self.lib.lbm355_batch_remove(self.h, case.h), "batch_remove")
        self.cases = [c for c in self.cases if c is not case]

    def size(self) -> tuple[int, int]:
        """(members, bricks in the concatenated list)."""
        m, b = C.c_int(), C.c_int64()
        check(self.lib.lbm355_batch_size(self.h, C.byref(m), C.byref(b)), "batch_size")
        return m.value, b.value

    def set_per_wave(self, per_wave: int) -> None:
        """Bricks per wavefront: 1 one each, k > 1 the persistent loop."""
        check(self.lib.lbm355_batch_set_per_wave(self.h, int(per_wave)), "batch_set_per_wave")

    def config(self) -> dict:
        out = (C.c_int * 8)()
        check(self.lib.lbm355_batch_config(self.h, out), "batch_config")
        keys = ("bricks_per_group", "nontemporal", "waves_per_cu", "group_mib", "per_wave", "groups", "uploads",
                "captures")
        return dict(zip(keys, list(out)))

    def autotune(self) -> tuple[int, int, int, int]:
        """Returns the fastest (bricks_per_group, nontemporal, waves_per_cu,
        group MiB with 0 = one group), now in use (``config()`` also gives the
        bricks per wavefront); real steps of every member."""
        best = C.c_int()
        check(self.lib.lbm355_batch_autotune(self.h, C.byref(best)), "batch_autotune")
        v = best.value
        return v & 0xFF, (v >> 8) & 0xFF, (v >> 16) & 0xFF, ((v >> 24) & 0xFF) * 16

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.lbm355_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pool:
    """Serving pool (lbm355_pool_*): requests are admitted against a device
    memory budget and stepped together, one fused launch per step and type."""

    def __init__(self, budget_bytes: int = 0, steps_per_tick: int = 64, graph: bool = True):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.lbm355_pool_create(int(budget_bytes), int(steps_per_tick), 1 if graph else 0, C.byref(h)),
              "pool_create")
        self.h = h
        self._dims: dict[int, tuple] = {}

    def submit(self, geometry_flat: np.ndarray, dims: Sequence[int], steps: int, report_every: int = 0,
               keep: bool = False, **kw) -> int:
        d, keep_alive = make_desc(geometry_flat, dims, **kw)
        rid = C.c_int64()
        check(self.lib.lbm355_pool_submit(self.h, C.byref(d), int(steps), int(report_every), 1 if keep else 0,
                                          C.byref(rid)), "pool_submit")
        del keep_alive  # the pool copied the geometry
        self._dims[rid.value] = tuple(int(v) for v in dims)
        return rid.value

    def poll(self, rid: int) -> dict:
        st = _abi.RequestStatus()
        rc = self.lib.lbm355_pool_poll(self.h, int(rid), C.byref(st))
        if rc not in (_abi.OK, _abi.EBUSY):
            check(rc, "pool_poll")
        return st.as_dict()

    def wait(self, rid: int, timeout: float = -1.0) -> dict:
        st = _abi.RequestStatus()
        rc = self.lib.lbm355_pool_wait(self.h, int(rid), float(timeout), C.byref(st))
        if rc not in (_abi.OK, _abi.EBUSY):
            check(rc, "pool_wait")
        return st.as_dict()

    def drain(self, timeout: float = -1.0) -> bool:
        rc = self.lib.lbm355_pool_drain(self.h, float(timeout))
        if rc not in (_abi.OK, _abi.EBUSY):
            check(rc, "pool_drain")
        return rc == _abi.OK

    def take(self, rid: int) -> Case:
        h = C.c_void_p()
        check(self.lib.lbm355_pool_take(self.h, int(rid), C.byref(h)), "pool_take")
        return Case.from_handle(h, self._dims[int(rid)])

    def stats(self) -> dict:
        st = _abi.PoolInfo()
        check(self.lib.lbm355_pool_stats(self.h, C.byref(st)), "pool_stats")
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.lbm355_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count() -> int:
    n = C.c_int()
    check(load().lbm355_device_count(C.byref(n)), "device_count")
    return n.value


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 10) -> float:
    """GB/s of a plain device copy kernel (read + write bytes per second)."""
    g = C.c_double()
    check(load().lbm355_copy_bandwidth(int(nbytes), int(iters), C.byref(g)), "copy_bandwidth")
    return g.value


def footprint(geometry_flat: np.ndarray, dims: Sequence[int], **kw) -> int:
    d, keep = make_desc(geometry_flat, dims, **kw)
    b = C.c_int64()
    check(load().lbm355_case_footprint(C.byref(d), C.byref(b)), "footprint")
    del keep
    return b.value


def labels(geometry_flat: np.ndarray, dims: Sequence[int], **kw) -> tuple[np.ndarray, dict]:
    """The label setup of a case without a device: (labels[x, y, z], info).

    A zero inlet/outlet area is reported as info["terminate"] = 1 (no raise)."""
    d, keep = make_desc(geometry_flat, dims, **kw)
    X, Y, Z = (int(v) for v in dims)
    out = np.empty(X * Y * Z, dtype=np.int32)
    info = CaseInfo()
    rc = load().lbm355_case_labels(C.byref(d), out.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(info))
    if rc not in (_abi.OK, _abi.ETERMINATE):
        check(rc, "labels")
    del keep
    return out.reshape(Z, Y, X).transpose(2, 1, 0), info.as_dict()
