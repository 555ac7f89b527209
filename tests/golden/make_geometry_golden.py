"""Writes the geometry, location and attenuation goldens from the reference's
own routines (oracle/ref_geom_probe.f90, linked against the reference
modules that oracle/build_ref.sh compiles):

  ref_geometry_mesh.npz   neighbours and face connections of the brick mesh
                          (tet_mesh_neighbor_tets and the mesher's face loop);
  ref_geometry_N<n>.npz   for N = 2..5 (FP64 builds): metric terms, J,
                          normals, sj, fscale, face frames, ibt/ibn of 48
                          elements, dtfactor, and findSource/findReceiver on
                          the brick's source, its 11 stations and 14 more
                          points (random, a mesh vertex, a face centroid,
                          an edge midpoint);
  ref_sls_f32.npz         anelasticPreCalc and calcAnelasticCoefficients of
                          the FP32 build (the FP64 build solves the least
                          squares with single-precision LAPACK on REAL(8)
                          arrays) for six materials and two frequency bands.

The mesh is the reference's simulations/brick example as packed in
tests/golden/brick.npz.  Needs the builds under oracle/_ref (this container
only): oracle/build_ref.sh <n> f64 for n = 2..5 and oracle/build_ref.sh 3 f32.
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")

MATERIALS = np.array([  # qp, qs, vp, vs, rho
    [50.0, 25.0, 3000.0, 1730.0, 2200.0],
    [200.0, 100.0, 5000.0, 2890.0, 2700.0],
    [9000.0, 9000.0, 3200.0, 1847.5, 2200.0],
    [30.0, 15.0, 2000.0, 1000.0, 2000.0],
    [120.0, 60.0, 6500.0, 3750.0, 3000.0],
    [80.0, 40.0, 2500.0, 1443.0, 2100.0],
])
BANDS = [(2.0, 100.0), (1.0, 10.0)]   # (f0_att, fr_att)


def brick():
    d = np.load(os.path.join(HERE, "brick.npz"))
    return d["coord"].astype(np.float64), d["elem"].astype(np.int32), d["props"][:, 0].astype(np.float64), d


def probe_points(coord, elem, d):
    """The brick's source and stations, 10 random points inside, one mesh
    vertex inside the cube, one face centroid, one edge midpoint and one
    point near a corner."""
    rng = np.random.default_rng(7)
    pts = [d["source"][:3]]
    pts += list(d["stations"])
    pts += list(rng.uniform(0.5, 9.5, (10, 3)))
    interior = np.nonzero(np.all((coord > 1.0) & (coord < 9.0), axis=1))[0]
    pts.append(coord[interior[0]])
    e = elem[123] - 1
    pts.append(coord[e[[0, 1, 2]]].mean(axis=0))
    pts.append(coord[e[[1, 3]]].mean(axis=0))
    pts.append(np.array([0.2, 0.3, 9.7]))
    return np.array(pts, dtype=np.float64)


def selection(nelem):
    rng = np.random.default_rng(3)
    return np.unique(np.concatenate([np.arange(16), np.arange(nelem - 16, nelem),
                                     rng.choice(nelem, 16, replace=False)])).astype(np.int32)


def run_probe(exe, coord, elem, src, rec, sel, vpu, band, mats):
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            np.array([elem.shape[0], coord.shape[0], len(src), len(rec), len(sel), len(mats)], "<i4").tofile(f)
            coord.astype("<f8").tofile(f)                      # (ncoord, 3) C order == (3, ncoord) Fortran
            elem.astype("<i4").tofile(f)
            np.asarray(src, "<f8").tofile(f)
            np.asarray(rec, "<f8").tofile(f)
            (sel + 1).astype("<i4").tofile(f)
            vpu.astype("<f8").tofile(f)
            np.array(band, "<f8").tofile(f)
            np.asarray(mats, "<f8").tofile(f)
        subprocess.run([exe, fin, fout], check=True)
        return parse(open(fout, "rb").read(), elem.shape[0], len(src), len(rec), len(sel), len(mats))


def parse(buf, nelem, nsrc, nrec, nsel, nmat):
    pos = 0

    def take(dtype, n):
        nonlocal pos
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=pos)
        pos += a.nbytes
        return a.copy()

    np_, npf, kind, nmb = take("<i4", 4)
    out = {"np": np_, "npf": npf, "real_kind": kind}
    out["neighbor"] = take("<i4", 4 * nelem).reshape(nelem, 4)
    out["face"] = take("<i4", 4 * nelem).reshape(nelem, 4)
    out["dtfactor"], out["min_gll_jac"] = take("<f8", 2)
    out["srcelem"] = take("<i4", nsrc)
    out["srcrst"] = take("<f8", 3 * nsrc).reshape(nsrc, 3)
    out["recelem"] = take("<i4", nrec)
    out["recrst"] = take("<f8", 3 * nrec).reshape(nrec, 3)
    names = ["vx", "vy", "vz", "rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "jac"]
    fnames = ["nx", "ny", "nz", "sj", "fscale"]
    per = {k: [] for k in names + fnames + ["nnormal", "stangential", "ttangential", "ibt", "ibn"]}
    for _ in range(nsel):
        for k in names:
            per[k].append(take("<f8", np_))
        for k in fnames:
            per[k].append(take("<f8", 4 * npf))
        for k in ("nnormal", "stangential", "ttangential"):
            per[k].append(take("<f8", 12).reshape(4, 3))
        per["ibt"].append(take("<i4", 4 * npf).reshape(4, npf))
        per["ibn"].append(take("<i4", 4 * npf).reshape(4, npf))
    out.update({k: np.array(v) for k, v in per.items()})
    out["wl"] = take("<f8", nmb)
    out["wk"] = take("<f8", 2 * nmb - 1)
    sls = take("<f8", nmat * (2 * nmb + 2)).reshape(nmat, 2 * nmb + 2)
    out["ylambda"], out["ymu"] = sls[:, :nmb], sls[:, nmb:2 * nmb]
    out["lambda_u"], out["mu_u"] = sls[:, 2 * nmb], sls[:, 2 * nmb + 1]
    assert pos == len(buf), (pos, len(buf))
    return out


def main():
    coord, elem, vp, d = brick()
    pts = probe_points(coord, elem, d)
    sel = selection(elem.shape[0])
    mesh_saved = False
    for n in (2, 3, 4, 5):
        exe = os.path.join(REF, f"o{n}_f64", "ref_geom_probe")
        o = run_probe(exe, coord, elem, pts, pts, sel, vp, BANDS[0], MATERIALS[:1])
        if not mesh_saved:
            np.savez_compressed(os.path.join(HERE, "ref_geometry_mesh.npz"), neighbor=o["neighbor"].astype(np.int16),
                                face=o["face"].astype(np.int8))
            mesh_saved = True
        keep = {k: v for k, v in o.items() if k not in ("neighbor", "face", "wl", "wk", "ylambda", "ymu",
                                                          "lambda_u", "mu_u")}
        np.savez_compressed(os.path.join(HERE, f"ref_geometry_N{n}.npz"), points=pts, sel=sel, **keep)
        print(f"N={n}: dtfactor {o['dtfactor']!r}, sources located in {o['srcelem'][:4]}...")
    exe = os.path.join(REF, "o3_f32", "ref_geom_probe")
    sls = {}
    for b, band in enumerate(BANDS):
        o = run_probe(exe, coord, elem, pts[:1], pts[:1], sel[:1], vp, band, MATERIALS)
        for k in ("wl", "wk", "ylambda", "ymu", "lambda_u", "mu_u"):
            sls[f"{k}_{b}"] = o[k]
    np.savez_compressed(os.path.join(HERE, "ref_sls_f32.npz"), materials=MATERIALS, bands=np.array(BANDS), **sls)
    print("wrote", sorted(sls))


if __name__ == "__main__":
    main()
