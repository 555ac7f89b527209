"""ctypes binding of libnexd_gpu.so (include/nexd_gpu.h).

The library is loaded from the package directory (built in-tree by
``make -C nexd-3d_amd``).  Nothing here falls back to CPU code: if the
library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libnexd_gpu.so")

NEXD_EULER, NEXD_RK2, NEXD_RK3 = 1, 2, 3
COMPAT_RK3_FLOAT_WEIGHTS = 1
COMPAT_PML_LAST_ELEMENT = 2
COMPAT_ALL = 3

EXPORTS = ["nexd_create", "nexd_upload_operators", "nexd_upload_mesh", "nexd_upload_material",
           "nexd_upload_anelastic", "nexd_upload_pml", "nexd_upload_sources", "nexd_upload_receivers",
           "nexd_reset", "nexd_run", "nexd_set_pml_active", "nexd_get_seismograms", "nexd_get_state",
           "nexd_set_state",
           "nexd_last_run_ms", "nexd_time_stage_kernel", "nexd_destroy", "nexd_last_error"]


class NexdConfig(ctypes.Structure):
    _fields_ = [("nelem", ctypes.c_int), ("order", ctypes.c_int), ("nsrc", ctypes.c_int),
                ("nrec", ctypes.c_int), ("nt", ctypes.c_int), ("timeint", ctypes.c_int),
                ("attenuation", ctypes.c_int), ("pml", ctypes.c_int), ("compat", ctypes.c_int),
                ("track_displacement", ctypes.c_int), ("energy", ctypes.c_int), ("device", ctypes.c_int),
                ("dt", ctypes.c_double)]


class NexdError(RuntimeError):
    pass


_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def load(path: str = LIB_PATH):
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise NexdError(f"{path} is not built (run make -C nexd-3d_amd)")
        lib = ctypes.CDLL(path)
        lib.nexd_create.argtypes = [ctypes.POINTER(NexdConfig), ctypes.POINTER(ctypes.c_void_p)]
        lib.nexd_last_error.restype = ctypes.c_char_p
        lib.nexd_last_error.argtypes = [ctypes.c_void_p]
        lib.nexd_destroy.restype = None
        lib.nexd_destroy.argtypes = [ctypes.c_void_p]
        for name in EXPORTS:
            f = getattr(lib, name)
            if name not in ("nexd_last_error", "nexd_destroy"):
                f.restype = ctypes.c_int
        _lib = lib
    return _lib


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def _i(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I)


class Context:
    """One library context holding one partition on one GPU."""

    def __init__(self, arrs: dict, compat: int = COMPAT_ALL, track_displacement: bool = False,
                 energy: bool = True, device: int = 0):
        self.lib = load()
        self.arrs = arrs
        cfg = NexdConfig(nelem=arrs["nelem"], order=arrs["order"], nsrc=arrs["nsrc"], nrec=arrs["nrec"],
                         nt=arrs["nt"], timeint=arrs["timeint"], attenuation=arrs["attenuation"],
                         pml=arrs["pml"], compat=compat, track_displacement=int(track_displacement),
                         energy=int(energy), device=device, dt=arrs["dt"])
        h = ctypes.c_void_p()
        rc = self.lib.nexd_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise NexdError(f"nexd_create: {rc}: {self.lib.nexd_last_error(None).decode()}")
        self.h = h
        self.cfg = cfg
        self._upload()

    def _chk(self, rc, what):
        if rc != 0:
            raise NexdError(f"{what}: {rc}: {self.lib.nexd_last_error(self.h).decode()}")

    def _upload(self):
        a, L = self.arrs, self.lib
        self._chk(L.nexd_upload_operators(self.h, _d(a["dr"]), _d(a["ds"]), _d(a["dt_"]), _d(a["lift"]),
                                          _d(a["mass"]), _i(a["fmask"])), "upload_operators")
        self._chk(L.nexd_upload_mesh(self.h, _i(a["ibool"]), _d(a["rx"]), _d(a["ry"]), _d(a["rz"]), _d(a["sx"]),
                                     _d(a["sy"]), _d(a["sz"]), _d(a["tx"]), _d(a["ty"]), _d(a["tz"]),
                                     _d(a["jac"]), _d(a["fscale"]), _i(a["neighbor"]), _i(a["face"]),
                                     _i(a["ibt"]), _i(a["ibn"]), _d(a["nnormal"]), _d(a["stangential"]),
                                     _d(a["ttangential"])), "upload_mesh")
        self._chk(L.nexd_upload_material(self.h, _d(a["vpu"]), _d(a["vsu"]), _d(a["rho"]), _d(a["lambdau"]),
                                         _d(a["muu"]), _d(a["mu"]), _d(a["lambda"])), "upload_material")
        if a["attenuation"]:
            self._chk(L.nexd_upload_anelastic(self.h, _d(a["wl"]), _d(a["anelasticvar"])), "upload_anelastic")
        if a["pml"]:
            self._chk(L.nexd_upload_pml(self.h, _i(a["pmlcheck"]), *[_d(a[k]) for k in (
                "ddx", "ddy", "ddz", "alphax", "alphay", "alphaz", "kx", "ky", "kz")]), "upload_pml")
        if a["nsrc"]:
            self._chk(L.nexd_upload_sources(self.h, _i(a["srcelem"]), _i(a["srctype"]), _d(a["src_array"]),
                                            _d(a["src_array_m"])), "upload_sources")
        if a["nrec"]:
            self._chk(L.nexd_upload_receivers(self.h, _i(a["recelem"]), _d(a["rec_int"])), "upload_receivers")

    def reset(self):
        self._chk(self.lib.nexd_reset(self.h), "reset")

    def run(self, it_begin: int, it_end: int, want_energy: bool = False):
        en = np.zeros(it_end - it_begin + 1) if want_energy else None
        self._chk(self.lib.nexd_run(self.h, it_begin, it_end, _d(en)), "run")
        return en

    def set_pml_active(self, active: bool):
        self._chk(self.lib.nexd_set_pml_active(self.h, int(active)), "set_pml_active")

    def seismograms(self) -> np.ndarray:
        """(nrec, 9, nt): ux uy uz vx vy vz ax ay az."""
        nrec, nt = self.cfg.nrec, self.cfg.nt
        outs = [np.zeros(nrec * nt) for _ in range(9)]
        self._chk(self.lib.nexd_get_seismograms(self.h, *[_d(o) for o in outs]), "get_seismograms")
        # each output is Fortran (nrec,nt): element (r,it) at it*nrec + r
        return np.stack([o.reshape(nt, nrec).T for o in outs], axis=1)

    def state(self, displacement: bool = False):
        """q as (9, nglob); with displacement=True also u as (3, nglob)
        (the context must track displacement)."""
        ng = self.arrs["nelem"] * self.arrs["np"]
        q = np.zeros(9 * ng)
        if not displacement:
            self._chk(self.lib.nexd_get_state(self.h, _d(q), None, None, None), "get_state")
            return q.reshape(9, ng)
        u = np.zeros((3, ng))
        self._chk(self.lib.nexd_get_state(self.h, _d(q), _d(u[0]), _d(u[1]), _d(u[2])), "get_state")
        return q.reshape(9, ng), u

    def set_state(self, it: int, q=None, theta=None, fprime=None, gprime=None, hprime=None, u=None):
        """Continue from a state after `it` steps (layouts as nexd_set_state)."""
        def arr(x):
            return None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        q, theta, fprime, gprime, hprime = map(arr, (q, theta, fprime, gprime, hprime))
        ux = uy = uz = None
        if u is not None:
            ux, uy, uz = (np.ascontiguousarray(u[k], dtype=np.float64) for k in range(3))
        self._keep = (q, theta, fprime, gprime, hprime, ux, uy, uz)
        self._chk(self.lib.nexd_set_state(self.h, it, *[_d(x) for x in self._keep]), "set_state")

    def last_run_ms(self) -> float:
        ms = ctypes.c_double()
        self._chk(self.lib.nexd_last_run_ms(self.h, ctypes.byref(ms)), "last_run_ms")
        return ms.value

    def time_stage_kernel(self, lst: int, nrep: int = 20):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        self._chk(self.lib.nexd_time_stage_kernel(self.h, lst, nrep, ctypes.byref(ms), ctypes.byref(n)),
                  "time_stage_kernel")
        return ms.value, n.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.nexd_destroy(self.h)
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


# ---------------------------------------------------------------------------
# batched shots (nexd_batch_*)
# ---------------------------------------------------------------------------
NEXD_FP64, NEXD_FP32 = 0, 1

BATCH_EXPORTS = ["nexd_batch_create", "nexd_batch_upload_operators", "nexd_batch_upload_mesh",
                 "nexd_batch_upload_material", "nexd_batch_upload_anelastic", "nexd_batch_upload_pml",
                 "nexd_batch_upload_sources", "nexd_batch_upload_receivers", "nexd_batch_reset", "nexd_batch_run",
                 "nexd_batch_get_seismograms", "nexd_batch_get_state", "nexd_batch_get_energy", "nexd_batch_set_pml_active",
                 "nexd_batch_autotune", "nexd_batch_set_tile", "nexd_batch_upload_element_dt",
                 "nexd_batch_last_run_ms", "nexd_batch_time_stage_kernel",
                 "nexd_batch_destroy", "nexd_batch_last_error"]
GROUP_EXPORTS = ["nexd_group_create", "nexd_group_run", "nexd_group_last_run_ms", "nexd_group_destroy",
                 "nexd_group_last_error"]


class NexdBatchConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nelem", "order", "nshot", "nsrc", "nrec", "nt", "timeint",
                                            "attenuation", "pml", "compat", "precision", "shared_material",
                                            "compensated", "energy", "device")] + [("dt", ctypes.c_double)]


def _load_batch():
    lib = load()
    if not getattr(lib, "_batch_ready", False):
        lib.nexd_batch_create.argtypes = [ctypes.POINTER(NexdBatchConfig), ctypes.POINTER(ctypes.c_void_p)]
        lib.nexd_batch_last_error.restype = ctypes.c_char_p
        lib.nexd_batch_last_error.argtypes = [ctypes.c_void_p]
        lib.nexd_batch_destroy.restype = None
        lib.nexd_batch_destroy.argtypes = [ctypes.c_void_p]
        for n in BATCH_EXPORTS:
            if n not in ("nexd_batch_last_error", "nexd_batch_destroy"):
                getattr(lib, n).restype = ctypes.c_int
        lib._batch_ready = True
    return lib


class BatchContext:
    """Many shots on one mesh (nexd_amd.batch.ShotBatch), FP32 or FP64."""

    def __init__(self, sb, precision: int = NEXD_FP32, compat: int = COMPAT_ALL, device: int = 0,
                 compensated: bool = False, energy: bool = False):
        from . import layout
        self.lib = _load_batch()
        base = sb.base
        a = layout.reference_arrays(base)
        self.nshot, self.nt, self.nrec = sb.nshot, base.nt, a["nrec"]
        self.nglob = a["nelem"] * a["np"]
        cfg = NexdBatchConfig(nelem=a["nelem"], order=a["order"], nshot=sb.nshot, nsrc=len(sb.shots[0].sources),
                              nrec=a["nrec"], nt=a["nt"], timeint=a["timeint"], attenuation=a["attenuation"],
                              pml=a["pml"], compat=compat, precision=precision,
                              shared_material=int(sb.shared_material), compensated=int(compensated),
                              energy=int(energy), device=device, dt=base.dt)
        h = ctypes.c_void_p()
        rc = self.lib.nexd_batch_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise NexdError(f"nexd_batch_create: {rc}: {self.lib.nexd_batch_last_error(None).decode()}")
        self.h = h
        L = self.lib
        self._chk(L.nexd_batch_upload_operators(h, _d(a["dr"]), _d(a["ds"]), _d(a["dt_"]), _d(a["lift"]), _d(a["mass"]),
                                                _i(a["fmask"])), "upload_operators")
        self._chk(L.nexd_batch_upload_mesh(h, _i(a["ibool"]), *[_d(a[k]) for k in (
            "rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "jac", "fscale")], _i(a["neighbor"]), _i(a["face"]),
            _i(a["ibt"]), _i(a["ibn"]), _d(a["nnormal"]), _d(a["stangential"]), _d(a["ttangential"])), "upload_mesh")
        nsets = 1 if sb.shared_material else sb.nshot
        for k in range(nsets):
            ak = layout.reference_arrays(sb.shots[k]) if k else layout.reference_arrays(sb.shots[0])
            self._chk(L.nexd_batch_upload_material(h, k, *[_d(ak[x]) for x in ("vpu", "vsu", "rho", "lambdau", "muu",
                                                                               "mu", "lambda")]),
                      "upload_material")
            if ak["attenuation"]:
                self._chk(L.nexd_batch_upload_anelastic(h, k, _d(ak["wl"]), _d(ak["anelasticvar"])), "upload_anelastic")
            if ak["pml"]:
                self._chk(L.nexd_batch_upload_pml(h, k, _i(ak["pmlcheck"]), *[_d(ak[x]) for x in (
                    "ddx", "ddy", "ddz", "alphax", "alphay", "alphaz", "kx", "ky", "kz")]), "upload_pml")
        for k, s in enumerate(sb.shots):
            so = s.src
            srcelem = np.ascontiguousarray(so.srcelem + 1, dtype=np.int32)
            srctype = np.ascontiguousarray(so.srctype, dtype=np.int32)
            spatial = np.ascontiguousarray(so.spatial, dtype=np.float64)
            stf = np.ascontiguousarray(so.stf, dtype=np.float64)
            self._chk(L.nexd_batch_upload_sources(h, k, _i(srcelem), _i(srctype), _d(spatial), _d(stf)),
                      "upload_sources")
        if a["nrec"]:
            self._chk(L.nexd_batch_upload_receivers(h, _i(a["recelem"]), _d(a["rec_int"])), "upload_receivers")

    def _chk(self, rc, what):
        if rc != 0:
            raise NexdError(f"{what}: {rc}: {self.lib.nexd_batch_last_error(self.h).decode()}")

    def reset(self):
        self._chk(self.lib.nexd_batch_reset(self.h), "reset")

    def run(self, it_begin: int, it_end: int):
        self._chk(self.lib.nexd_batch_run(self.h, it_begin, it_end), "run")

    def seismograms(self, shot: int) -> np.ndarray:
        outs = [np.zeros(self.nrec * self.nt) for _ in range(9)]
        self._chk(self.lib.nexd_batch_get_seismograms(self.h, shot, *[_d(o) for o in outs]), "get_seismograms")
        return np.stack([o.reshape(self.nt, self.nrec).T for o in outs], axis=1)

    def autotune(self, nrep: int = 3):
        """Times the block shapes of the stage launches on this batch and keeps
        the fastest (resets the state).  Returns (shape, [ms per shape])."""
        t = ctypes.c_int()
        ms = (ctypes.c_double * 2)()
        self._chk(self.lib.nexd_batch_autotune(self.h, nrep, ctypes.byref(t), ms), "autotune")
        return t.value, [ms[0], ms[1]]

    def set_tile(self, tile: int):
        self._chk(self.lib.nexd_batch_set_tile(self.h, tile), "set_tile")

    def state(self, shot: int) -> np.ndarray:
        """q of one shot as (9, nglob)."""
        out = np.zeros(9 * self.nglob)
        self._chk(self.lib.nexd_batch_get_state(self.h, shot, _d(out)), "get_state")
        return out.reshape(9, self.nglob)

    def energy(self, shot: int) -> np.ndarray:
        """all_energy of every step of one shot (nt,)."""
        out = np.zeros(self.nt)
        self._chk(self.lib.nexd_batch_get_energy(self.h, shot, _d(out)), "get_energy")
        return out

    def set_pml_active(self, shot: int, active: bool):
        self._chk(self.lib.nexd_batch_set_pml_active(self.h, shot, int(active)), "set_pml_active")

    def last_run_ms(self) -> float:
        ms = ctypes.c_double()
        self._chk(self.lib.nexd_batch_last_run_ms(self.h, ctypes.byref(ms)), "last_run_ms")
        return ms.value

    def time_stage_kernel(self, lst: int, nrep: int = 20):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        self._chk(self.lib.nexd_batch_time_stage_kernel(self.h, lst, nrep, ctypes.byref(ms), ctypes.byref(n)),
                  "time_stage_kernel")
        return ms.value, n.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.nexd_batch_destroy(self.h)
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


class Group:
    """Independent batches advanced together, one parent HIP graph per step
    (nexd_group_*).  The members keep their own state and getters."""

    def __init__(self, members):
        self.lib = _load_batch()
        L = self.lib
        if not getattr(L, "_group_ready", False):
            L.nexd_group_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_void_p)]
            L.nexd_group_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            L.nexd_group_last_run_ms.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
            L.nexd_group_destroy.restype = None
            L.nexd_group_destroy.argtypes = [ctypes.c_void_p]
            L.nexd_group_last_error.restype = ctypes.c_char_p
            L.nexd_group_last_error.argtypes = [ctypes.c_void_p]
            for n in ("nexd_group_create", "nexd_group_run", "nexd_group_last_run_ms"):
                getattr(L, n).restype = ctypes.c_int
            L._group_ready = True
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h for m in self.members])
        h = ctypes.c_void_p()
        rc = L.nexd_group_create(arr, len(self.members), ctypes.byref(h))
        if rc != 0:
            raise NexdError(f"nexd_group_create: {rc}: {L.nexd_group_last_error(None).decode()}")
        self.h = h

    def run(self, it_begin: int, it_end: int):
        rc = self.lib.nexd_group_run(self.h, it_begin, it_end)
        if rc != 0:
            raise NexdError(f"nexd_group_run: {rc}: {self.lib.nexd_group_last_error(self.h).decode()}")

    def last_run_ms(self) -> float:
        ms = ctypes.c_double()
        self.lib.nexd_group_last_run_ms(self.h, ctypes.byref(ms))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.nexd_group_destroy(self.h)
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
