"""Packs the reference's `simulations/brick` example into one fixture.

The example holds the mesher inputs (cubit/{coord,mesh,mat,matprop,free,
absorb}), the run inputs (data/{parfile,source,stations}) and the
seismograms the reference wrote (ref/seismo.{x,y,z}.%07d.sd{u,v,a}, FP32
build, 12 ranks, RK2, N=4, 1000 steps).  This script reads those data files
(as text) and stores their numbers in tests/golden/brick.npz, and the nine
input files unchanged in tests/golden/brick_inputs.tar.xz (for the tests of
the readers); no reference source is involved.

Needs /root/reference (this container only).
"""
import os
import re
import tarfile

import numpy as np

REF = os.environ.get("NEXD_REFERENCE", "/root/reference")
EX = os.path.join(REF, "simulations", "brick")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "brick.npz")
TEXT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "brick_inputs.tar.xz")


def rows(path, skip_count=True):
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    return lines[1:] if skip_count else lines


def parfile(path):
    vals = {}
    for ln in open(path):
        ln = ln.split("#")[0]
        if "=" in ln:
            k, v = ln.split("=", 1)
            vals[k.strip()] = v.strip()
    return vals


def main():
    cub = os.path.join(EX, "cubit")
    coord = np.array([[float(v) for v in r[1:4]] for r in rows(os.path.join(cub, "coord"))])
    mesh = rows(os.path.join(cub, "mesh"))
    elem = np.array([[int(v) for v in r[1:5]] for r in mesh], dtype=np.int32)
    props = np.array([[float(v) for v in r[5:10]] for r in mesh])
    mat = np.array([[int(v) for v in r[:3]] for r in rows(os.path.join(cub, "mat"), False)], dtype=np.int32)
    matprop = np.array([[float(v) for v in r[:7]] for r in rows(os.path.join(cub, "matprop"))])
    free = np.array([[int(v) for v in r[:4]] for r in rows(os.path.join(cub, "free"))], dtype=np.int32)
    absorb = np.array([[int(v) for v in r[:4]] for r in rows(os.path.join(cub, "absorb"))], dtype=np.int32)
    par = parfile(os.path.join(EX, "data", "parfile"))
    src = parfile(os.path.join(EX, "data", "source"))
    st = rows(os.path.join(EX, "data", "stations"))
    stations = np.array([[float(v) for v in r[1:4]] for r in st])
    nrec = stations.shape[0]
    nt = int(par["nt"])
    seis = np.zeros((nrec, 9, nt), dtype=np.float32)
    time = None
    for r in range(nrec):
        for k, (c, q) in enumerate([(c, q) for q in ("sdu", "sdv", "sda") for c in "xyz"]):
            d = np.loadtxt(os.path.join(EX, "ref", f"seismo.{c}.{r + 1:07d}.{q}"))
            seis[r, k] = d[:nt, 1]
            time = d[:nt, 0]
    keys = ["timeint", "nt", "cfl", "nproc"]
    srck = ["xsource", "ysource", "zsource", "sourcetype", "stf", "f0", "factor", "xangle_force",
            "yangle_force", "zangle_force"]
    np.savez_compressed(OUT, coord=coord, elem=elem, props=props, mat=mat, matprop=matprop, free=free,
                        absorb=absorb, stations=stations, seismo=seis, time=time.astype(np.float32),
                        par=np.array([float(re.sub(r"[^0-9.eE+-]", "", par[k])) for k in keys]),
                        source=np.array([float(src[k]) for k in srck]))
    print(f"wrote {OUT}: {elem.shape[0]} tets, {nrec} receivers, {nt} samples")
    # the input files themselves, as they are, for the readers' tests
    with tarfile.open(TEXT, "w:xz") as tf:
        for sub, names in (("cubit", ("coord", "mesh", "mat", "matprop", "free", "absorb")),
                           ("data", ("parfile", "source", "stations"))):
            for n in names:
                tf.add(os.path.join(EX, sub, n), arcname=f"{sub}/{n}")
    print(f"wrote {TEXT}")


if __name__ == "__main__":
    main()
