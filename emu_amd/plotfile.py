"""Readers of the plotfiles WritePlotFile writes (Source/IO.cpp:11-36; the
format is in emu_amd/cpp/PlotFile.cpp): the grid of a single-level AMReX
plotfile and the "neutrinos" particle checkpoint.  The particle reader follows
the reference's own (Scripts/visualization/amrex_plot_tools.py:89-210,
Scripts/tests/EmuReader.py): ints (id, cpu) of a grid's particles, then their
reals, at each grid's offset in its DATA file.  Pure numpy; no GPU.
"""
from __future__ import annotations

import os

import numpy as np


def read_header(plotdir: str) -> dict:
    """Plotfile Header: component names, time, step, domain and cell size."""
    with open(os.path.join(plotdir, "Header")) as f:
        lines = [l.rstrip("\n") for l in f]
    ncomp = int(lines[1])
    names = lines[2:2 + ncomp]
    i = 2 + ncomp
    dim = int(lines[i])
    time = float(lines[i + 1])
    finest = int(lines[i + 2])
    plo = [float(x) for x in lines[i + 3].split()]
    phi = [float(x) for x in lines[i + 4].split()]
    box = lines[i + 6].strip()
    hi = [int(x) for x in box.split(") (")[1].split(",")]
    step = int(lines[i + 7].split()[0])
    dx = [float(x) for x in lines[i + 8].split()]
    return {"names": names, "dim": dim, "time": time, "finest_level": finest, "plo": plo, "phi": phi,
            "ncell": [h + 1 for h in hi], "step": step, "dx": dx}


def read_grid(plotdir: str):
    """(names, array[ncomp, nz, ny, nx]) of the valid cells."""
    h = read_header(plotdir)
    n = h["ncell"]
    with open(os.path.join(plotdir, "Level_0", "Cell_H")) as f:
        cell_h = f.read().split("\n")
    fab = [l for l in cell_h if l.startswith("FabOnDisk:")][0].split()
    fn, offset = fab[1], int(fab[2])
    with open(os.path.join(plotdir, "Level_0", fn), "rb") as f:
        f.seek(offset)
        header = f.readline().decode()
        if "(8, (1 2 3 4 5 6 7 8))" not in header:
            raise ValueError(f"{fn}: not little-endian IEEE double data")
        ncomp = int(header.split()[-1])
        data = np.fromfile(f, dtype="<f8", count=ncomp * n[0] * n[1] * n[2])
    return h["names"], data.reshape(ncomp, n[2], n[1], n[0])


def read_particles(plotdir: str, ptype: str = "neutrinos"):
    """(ids[np], rdata[np, 3 + nattribs], real names): positions first, then
    the PIdx attributes, as the reference's read_particle_data returns them."""
    base = os.path.join(plotdir, ptype)
    with open(os.path.join(base, "Header")) as f:
        version = f.readline().strip()
        if not version.endswith("_double"):
            raise ValueError(f"{base}: not double-precision particles")
        dim = int(f.readline())
        nreal = int(f.readline())
        names = [f.readline().strip() for _ in range(nreal)]
        nint = int(f.readline())
        for _ in range(nint):
            f.readline()
        checkpoint = bool(int(f.readline()))
        np_ = int(f.readline())
        f.readline()  # next id
        finest = int(f.readline())
        grids = []
        for _ in range(finest + 1):
            grids.append(int(f.readline()))
        entries = [[int(v) for v in f.readline().split()] for _ in range(sum(grids))]
    ni = 2 + nint if checkpoint else 0
    nr = dim + nreal
    ids = np.zeros(np_, dtype=np.int64)
    rdata = np.zeros((np_, nr))
    ip = 0
    for which, count, where in entries:
        if count == 0:
            continue
        with open(os.path.join(base, "Level_0", "DATA_%05d" % which), "rb") as f:
            f.seek(where)
            ints = np.fromfile(f, dtype="<i4", count=count * ni).reshape(count, ni) if ni else None
            reals = np.fromfile(f, dtype="<f8", count=count * nr).reshape(count, nr)
        ids[ip:ip + count] = ints[:, 0] if ni else np.arange(ip + 1, ip + count + 1)
        rdata[ip:ip + count] = reals
        ip += count
    if ip != np_:
        raise ValueError(f"{base}: {ip} particles read, header says {np_}")
    return ids, rdata, ["pos_x", "pos_y", "pos_z"] + names
