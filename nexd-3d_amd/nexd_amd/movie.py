"""Movie frames and their VTK files.

The solver writes, every ``frame`` steps, the displacement magnitude
sqrt(ux^2 + uy^2 + uz^2) at every node as one Fortran unformatted record,
out/moviedata%06d_it%07d.bin (src/timeloopMod.f90:576-582).  The movie
program turns them into one legacy-ASCII VTK file per frame,
out/movie_elem_it%07d.vtk, with the mean over each element's nodes as cell
data on the mesh's tetrahedra (src/collectMovieMod.f90:30-204,
src/vtkMod.f90:70-111).  ``collect`` does the latter for a one-partition
run; the frames come from the GPU state (nexd_get_state) or from the
reference solver alike.
"""
from __future__ import annotations

import os
import struct

import numpy as np


def frame_path(outdir: str, it: int, rank: int = 1) -> str:
    return os.path.join(outdir, f"moviedata{rank:06d}_it{it:07d}.bin")


def vtk_path(outdir: str, frame_no: int) -> str:
    """Frames are numbered 1, 2, ... in the VTK names, not by step."""
    return os.path.join(outdir, f"movie_elem_it{frame_no:07d}.vtk")


def magnitude(ux, uy, uz) -> np.ndarray:
    return np.sqrt(ux ** 2 + uy ** 2 + uz ** 2)


def write_frame(path: str, uplot, real: str = "<f8") -> None:
    """One unformatted sequential record: length marker, values, marker."""
    data = np.ascontiguousarray(uplot, dtype=real).tobytes()
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(data)))
        f.write(data)
        f.write(struct.pack("<i", len(data)))


def read_frame(path: str, real: str = "<f8") -> np.ndarray:
    with open(path, "rb") as f:
        buf = f.read()
    n = struct.unpack("<i", buf[:4])[0]
    if len(buf) != n + 8 or struct.unpack("<i", buf[4 + n:8 + n])[0] != n:
        raise ValueError(f"{path}: not one unformatted record")
    return np.frombuffer(buf[4:4 + n], dtype=real).astype(np.float64)


def element_means(uplot: np.ndarray, nelem: int, np_: int) -> np.ndarray:
    """Sum over each element's nodes in node order, then divided by Np
    (src/collectMovieMod.f90:133-147); nodes are numbered element by
    element."""
    u = np.asarray(uplot, np.float64).reshape(nelem, np_)
    acc = np.zeros(nelem)
    for j in range(np_):
        acc = acc + u[:, j]
    return acc / np_


def fortran_e(x: float, w: int = 18, d: int = 6) -> str:
    """Fortran Ew.d: 0.ddd...E+ee, right-justified in w columns."""
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
    else:
        s = f"{abs(x):.{d - 1}e}"
        mant, ex = s.split("e")
        digits = mant.replace(".", "")
        e10 = int(ex) + 1
        if abs(e10) > 99:
            raise ValueError("exponent beyond two digits")
        body = f"0.{digits}E{'+' if e10 >= 0 else '-'}{abs(e10):02d}"
        if x < 0:
            body = "-" + body
    return body.rjust(w)


def write_vtk_cells(path: str, elem0: np.ndarray, coord: np.ndarray, data: np.ndarray) -> None:
    """Legacy ASCII unstructured grid of tetrahedra (cell type 10) with one
    scalar per cell; elem0 holds 0-based vertex ids (src/vtkMod.f90:70-111)."""
    nelem, ncoord = elem0.shape[0], coord.shape[0]
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.1\nMesh with Data\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {ncoord:12d} float\n")
        for x, y, z in coord:
            f.write(f"{fortran_e(x)}{fortran_e(y)}{fortran_e(z)}\n")
        f.write(" \n")
        f.write(f"CELLS {nelem:12d}{nelem * 5:12d}\n")
        for v in elem0:
            f.write(f"{4:12d}{v[0]:12d}{v[1]:12d}{v[2]:12d}{v[3]:12d}\n")
        f.write(" \n")
        f.write(f"CELL_TYPES {nelem:12d}\n")
        for k in range(0, nelem, 20):
            f.write(" " + " ".join(["10"] * min(20, nelem - k)) + "\n")
        f.write(" \n")
        f.write(f"CELL_DATA {nelem:12d}\nSCALARS data float\nLOOKUP_TABLE default\n")
        for v in data:
            f.write(f" {float(v)!r}\n")
        f.write(" \n")


def kept_coordinates(coord: np.ndarray) -> np.ndarray:
    """The movie program keeps the mesh vertices whose x is above -1 (it
    marks unused slots with -1, src/collectMovieMod.f90:160-178)."""
    return coord[coord[:, 0] > -1.0]


def collect(outdir: str, elem0: np.ndarray, coord: np.ndarray, nt: int, frame: int, np_: int,
            real: str = "<f8") -> list:
    """VTK files of every frame of a one-partition run: frames at steps
    frame, 2 frame, ... <= nt.  Returns the paths written."""
    if frame <= 0:
        raise ValueError("frame must be positive")
    nelem = elem0.shape[0]
    pts = kept_coordinates(np.asarray(coord, np.float64))
    paths = []
    for c, it in enumerate(range(frame, nt + 1, frame), start=1):
        u = read_frame(frame_path(outdir, it), real)
        if u.size != nelem * np_:
            raise ValueError(f"frame {it}: {u.size} values for {nelem * np_} nodes")
        p = vtk_path(outdir, c)
        write_vtk_cells(p, elem0, pts, element_means(u, nelem, np_))
        paths.append(p)
    return paths


def read_vtk_cells(path: str):
    """(points (n,3), cells (m,4) 0-based, cell types (m,), data (m,)) of a
    file written as above, by either writer."""
    tok = open(path).read().split()
    i = tok.index("POINTS")
    npts = int(tok[i + 1])
    pts = np.array([float(t.replace("D", "E")) for t in tok[i + 3:i + 3 + 3 * npts]]).reshape(npts, 3)
    i = tok.index("CELLS")
    nc = int(tok[i + 1])
    cells = np.array([int(t) for t in tok[i + 3:i + 3 + 5 * nc]]).reshape(nc, 5)
    if not np.all(cells[:, 0] == 4):
        raise ValueError("cells are not tetrahedra")
    i = tok.index("CELL_TYPES")
    types = np.array([int(t) for t in tok[i + 2:i + 2 + nc]])
    i = tok.index("LOOKUP_TABLE")
    data = np.array([float(t.replace("D", "E")) for t in tok[i + 2:i + 2 + nc]])
    return pts, cells[:, 1:], types, data


__all__ = ["frame_path", "vtk_path", "magnitude", "write_frame", "read_frame", "element_means", "fortran_e",
           "write_vtk_cells", "kept_coordinates", "collect", "read_vtk_cells"]
