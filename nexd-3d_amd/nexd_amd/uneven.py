"""Uneven batches packed by order: independent instances of different mesh,
size, order and time step (no reference counterpart).

Instances of one polynomial order (and physics) are concatenated into one
mesh, a disjoint union: element, node and neighbour indices are offset,
every per-element array is stacked, and each element advances with its own
instance's time step (nexd_batch_upload_element_dt).  One batch of that
union advances all of them with one stage launch per element list, instead
of one per instance; a group (nexd_group_*) of the per-order unions
advances the whole uneven set with one graph launch per step.  The seismograms
of instance k are the rows of its receivers.

``pack`` builds the unions, ``PackedBatch`` runs one (FP32 or FP64, with
the block-shape autotuner if asked); several advance together as a
``capi.Group``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import capi, layout

# per-element arrays stacked along their first (element) axis
_STACK = ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "jac", "fscale", "face", "nnormal",
          "stangential", "ttangential", "vpu", "vsu", "rho", "lambdau", "muu", "mu", "lambda", "pmlcheck")


@dataclass
class Packed:
    """One union of instances of the same order."""
    arrs: dict                 # reference-layout arrays of the union
    members: list              # instance indices
    elem_off: list             # first element of each member
    rec_off: list              # first receiver of each member
    nrec: list                 # receivers of each member
    dt_elem: np.ndarray        # (nelem,) step of each element
    src_spatial: np.ndarray    # (nsrc, 6, Np)
    src_stf: np.ndarray        # (nsrc, nt)


def _stack(arr_list, key):
    return np.ascontiguousarray(np.concatenate([a[key] for a in arr_list], axis=0))


def pack(instances) -> list:
    """Unions of ``instances`` (nexd_amd.batch.ShotBatch, one shot each, no
    PML) grouped by (order, attenuation, time scheme); nt must agree."""
    groups = {}
    for k, sb in enumerate(instances):
        p = sb.base
        if p.set_pml:
            # the reference's PML update reads the fluxes of the mesh's last
            # element (src/timeloopMod.f90:399-405, reproduced in compat
            # mode): in a union every instance would read the union's last
            # element instead of its own
            raise ValueError("instances with PML cannot be packed: their PML update reads the mesh's last element")
        key = (p.order, int(p.attenuation), int(p.set_pml), p.timeint, p.nt)
        groups.setdefault(key, []).append(k)
    nts = {key[4] for key in groups}
    if len(nts) > 1:
        raise ValueError("instances of a packed set share nt")
    out = []
    for key, ks in sorted(groups.items()):
        arr_list = [layout.reference_arrays(instances[k].shots[0]) for k in ks]
        a0 = arr_list[0]
        np_ = a0["np"]
        d = {k: a0[k] for k in ("np", "npf", "order", "nmb", "nt", "timeint", "attenuation", "pml", "dr", "ds",
                                 "dt_", "mass", "lift", "fmask", "vdm", "wl")}
        for a in arr_list[1:]:
            if not np.array_equal(a["wl"], a0["wl"]):
                raise ValueError("relaxation frequencies must agree within a packed order")
        nel = [a["nelem"] for a in arr_list]
        eoff = np.concatenate([[0], np.cumsum(nel)[:-1]]).astype(np.int64)
        d["nelem"] = int(sum(nel))
        for k2 in _STACK:
            d[k2] = _stack(arr_list, k2)
        d["anelasticvar"] = np.ascontiguousarray(np.concatenate([a["anelasticvar"] for a in arr_list]))
        d["ibool"] = np.ascontiguousarray(np.concatenate(
            [a["ibool"] + o * np_ for a, o in zip(arr_list, eoff)])).astype(np.int32)
        d["neighbor"] = np.ascontiguousarray(np.concatenate(
            [np.where(a["neighbor"] > 0, a["neighbor"] + o, 0) for a, o in zip(arr_list, eoff)])).astype(np.int32)
        d["ibt"] = np.ascontiguousarray(np.concatenate(
            [a["ibt"] + o * np_ for a, o in zip(arr_list, eoff)])).astype(np.int32)
        d["ibn"] = np.ascontiguousarray(np.concatenate(
            [np.where(a["ibn"] > 0, a["ibn"] + o * np_, 0) for a, o in zip(arr_list, eoff)])).astype(np.int32)
        if d["pml"]:
            for k2 in ("ddx", "ddy", "ddz", "alphax", "alphay", "alphaz", "kx", "ky", "kz"):
                d[k2] = np.ascontiguousarray(np.concatenate([a[k2] for a in arr_list]))
        d["nsrc"] = int(sum(a["nsrc"] for a in arr_list))
        d["srcelem"] = np.ascontiguousarray(np.concatenate(
            [a["srcelem"] + o for a, o in zip(arr_list, eoff)])).astype(np.int32)
        d["srctype"] = np.ascontiguousarray(np.concatenate([a["srctype"] for a in arr_list])).astype(np.int32)
        d["nrec"] = int(sum(a["nrec"] for a in arr_list))
        d["recelem"] = np.ascontiguousarray(np.concatenate(
            [a["recelem"] + o for a, o in zip(arr_list, eoff)])).astype(np.int32)
        d["rec_int"] = np.ascontiguousarray(np.concatenate([a["rec_int"] for a in arr_list]))
        d["dt"] = float(min(a["dt"] for a in arr_list))
        dt_elem = np.concatenate([np.full(n, a["dt"]) for n, a in zip(nel, arr_list)])
        spatial = np.ascontiguousarray(np.concatenate([a["src_spatial"] for a in arr_list]))
        stf = np.ascontiguousarray(np.concatenate([a["src_stf"] for a in arr_list]))
        nrec = [a["nrec"] for a in arr_list]
        roff = np.concatenate([[0], np.cumsum(nrec)[:-1]]).astype(int).tolist()
        out.append(Packed(d, ks, eoff.tolist(), roff, nrec, dt_elem, spatial, stf))
    return out


class PackedBatch:
    """One union as a batch of one shot: FP32 or FP64, own step per element."""

    def __init__(self, pk: Packed, precision: int = capi.NEXD_FP32, device: int = 0, autotune: bool = False):
        self.lib = capi._load_batch()
        self.pk = pk
        a = pk.arrs
        L = self.lib
        cfg = capi.NexdBatchConfig(nelem=a["nelem"], order=a["order"], nshot=1, nsrc=a["nsrc"], nrec=a["nrec"],
                                   nt=a["nt"], timeint=a["timeint"], attenuation=a["attenuation"], pml=a["pml"],
                                   compat=capi.COMPAT_ALL, precision=precision, shared_material=1, compensated=0,
                                   energy=0, device=device, dt=a["dt"])
        h = ctypes.c_void_p()
        rc = L.nexd_batch_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise capi.NexdError(f"nexd_batch_create: {rc}: {L.nexd_batch_last_error(None).decode()}")
        self.h = h
        self.nt, self.nrec = a["nt"], a["nrec"]
        d, i = capi._d, capi._i
        self._chk(L.nexd_batch_upload_operators(h, d(a["dr"]), d(a["ds"]), d(a["dt_"]), d(a["lift"]), d(a["mass"]),
                                                i(a["fmask"])), "upload_operators")
        self._chk(L.nexd_batch_upload_mesh(h, i(a["ibool"]), *[d(a[k]) for k in (
            "rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "jac", "fscale")], i(a["neighbor"]), i(a["face"]),
            i(a["ibt"]), i(a["ibn"]), d(a["nnormal"]), d(a["stangential"]), d(a["ttangential"])), "upload_mesh")
        self._chk(L.nexd_batch_upload_material(h, 0, *[d(a[k]) for k in ("vpu", "vsu", "rho", "lambdau", "muu", "mu",
                                                                          "lambda")]), "upload_material")
        if a["attenuation"]:
            self._chk(L.nexd_batch_upload_anelastic(h, 0, d(a["wl"]), d(a["anelasticvar"])), "upload_anelastic")
        if a["pml"]:
            self._chk(L.nexd_batch_upload_pml(h, 0, i(a["pmlcheck"]), *[d(a[k]) for k in (
                "ddx", "ddy", "ddz", "alphax", "alphay", "alphaz", "kx", "ky", "kz")]), "upload_pml")
        self._chk(L.nexd_batch_upload_sources(h, 0, i(a["srcelem"]), i(a["srctype"]), d(pk.src_spatial),
                                              d(pk.src_stf)), "upload_sources")
        self._chk(L.nexd_batch_upload_receivers(h, i(a["recelem"]), d(a["rec_int"])), "upload_receivers")
        dte = np.ascontiguousarray(pk.dt_elem, dtype=np.float64)
        self._chk(L.nexd_batch_upload_element_dt(h, d(dte)), "upload_element_dt")
        self.tile, self.tile_ms = 0, None
        if autotune:
            self.autotune()

    def _chk(self, rc, what):
        if rc != 0:
            raise capi.NexdError(f"{what}: {rc}: {self.lib.nexd_batch_last_error(self.h).decode()}")

    def autotune(self, nrep: int = 3):
        t = ctypes.c_int()
        ms = (ctypes.c_double * 2)()
        self._chk(self.lib.nexd_batch_autotune(self.h, nrep, ctypes.byref(t), ms), "autotune")
        self.tile, self.tile_ms = t.value, [ms[0], ms[1]]
        return self.tile

    def reset(self):
        self._chk(self.lib.nexd_batch_reset(self.h), "reset")

    def run(self, it_begin, it_end):
        self._chk(self.lib.nexd_batch_run(self.h, it_begin, it_end), "run")

    def last_run_ms(self):
        ms = ctypes.c_double()
        self._chk(self.lib.nexd_batch_last_run_ms(self.h, ctypes.byref(ms)), "last_run_ms")
        return ms.value

    def seismograms(self):
        """Per member: (nrec_k, 9, nt)."""
        outs = [np.zeros(self.nrec * self.nt) for _ in range(9)]
        self._chk(self.lib.nexd_batch_get_seismograms(self.h, 0, *[capi._d(o) for o in outs]), "get_seismograms")
        s = np.stack([o.reshape(self.nt, self.nrec).T for o in outs], axis=1)
        return [s[r:r + n] for r, n in zip(self.pk.rec_off, self.pk.nrec)]

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
