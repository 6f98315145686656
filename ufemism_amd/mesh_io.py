"""Restart files in the reference's NetCDF layout: the primary mesh and the
ice state that a UFEMISM run restarts from.

The reference writes a classic-format NetCDF file
(``nf90_create(IOR(nf90_clobber, nf90_share))``,
src/netcdf_module.f90:490-620) and reads it back in three steps:
``inquire_restart_file_mesh`` (:3012-3049) checks the dimensions vi, ti, ci,
two, three and the variables V, nC, C, niTri, iTri, edge_index, Tri, Tricc,
TriC, Tri_edge_index; ``inquire_restart_file_init`` (:3050-3103) checks zeta,
time and month and Hi, Hb, Hs, Ti, U_SSA, V_SSA, MeltPreviousYear and
FirnDepth; ``read_restart_file_mesh`` / ``read_restart_file_init``
(:3104-3185) read them, the latter at the time frame nearest to
``time_to_restart_from``.  After reading, the reference takes the domain
bounds from the vertex coordinates and tol_dist = ((xmax-xmin)+(ymax-ymin))
* 1e-9 / 2 (src/restart_module.f90:75-85) before it computes the secondary
mesh data.

NetCDF lists dimensions slowest first, the reverse of Fortran: the
reference's V(vi, two) is the file variable V(two, vi).  Arrays here keep
the reference (Fortran) shapes; the transposition happens at the file
boundary.
"""
from __future__ import annotations

import numpy as np
from scipy.io import netcdf_file

from .meshgen import Mesh

# (variable, Fortran dimensions, type, long_name, units) of the mesh part,
# src/netcdf_module.f90:549-558, 564-565
MESH_VARS = [
    ("V", ("vi", "two"), "d", "Vertex coordinates", "m"),
    ("Tri", ("ti", "three"), "i", "Vertex indices", None),
    ("nC", ("vi",), "i", "Number of connected vertices", None),
    ("C", ("vi", "ci"), "i", "Indices of connected vertices", None),
    ("niTri", ("vi",), "i", "Number of inverse triangles", None),
    ("iTri", ("vi", "ci"), "i", "Indices of inverse triangles", None),
    ("edge_index", ("vi",), "i", "Edge index", None),
    ("Tricc", ("ti", "two"), "d", "Triangle circumcenter", "m"),
    ("TriC", ("ti", "three"), "i", "Triangle neighbours", None),
    ("Tri_edge_index", ("ti",), "i", "Triangle edge index", None),
]
MESH_OPTIONAL = [
    ("A", ("vi",), "d", "Vertex Voronoi cell area", "m^2"),
    ("R", ("vi",), "d", "Vertex resolution", "m"),
]
# model data, src/netcdf_module.f90:598-607 (time is the unlimited dimension)
INIT_VARS = [
    ("Hi", ("vi", "time"), "Ice Thickness", "m"),
    ("Hb", ("vi", "time"), "Bedrock Height", "m"),
    ("Hs", ("vi", "time"), "Surface Height", "m"),
    ("U_SSA", ("vi", "time"), "SSA ice x-velocity", "m/yr"),
    ("V_SSA", ("vi", "time"), "SSA ice y-velocity", "m/yr"),
    ("Ti", ("vi", "zeta", "time"), "Ice temperature", "K"),
    ("FirnDepth", ("vi", "month", "time"), "Firn depth", "m"),
    ("MeltPreviousYear", ("vi", "time"), "Melt during previous year", "mie"),
]


def _to_file(a):
    """Fortran-shaped array -> NetCDF (slowest-first) order."""
    return np.ascontiguousarray(np.asarray(a).T)


def _from_file(a, dtype):
    return np.asfortranarray(np.array(a, dtype=dtype).T)


def write_restart(path, mesh, zeta=None, frames=()):
    """Write a restart file in the reference's layout.

    mesh    a primary mesh (meshgen.Mesh or anything with the same attributes;
            A and R are written when present, as the reference does);
    zeta    the scaled vertical coordinate (required with frames);
    frames  sequence of (time, {name: array}) with the Aa fields of INIT_VARS
            in their Fortran shapes; absent fields are written as zeros.
    """
    nV, nTri, nCm = int(mesh.V.shape[0]), int(mesh.Tri.shape[0]), int(mesh.nC_mem)
    big = nV * nCm * 8 * 2 > 2 ** 30
    with netcdf_file(path, "w", version=2 if big else 1) as f:
        if frames:
            # the record dimension first (the reference finds dimensions by name)
            f.createDimension("time", None)
        for name, n in (("vi", nV), ("ti", nTri), ("ci", nCm), ("two", 2), ("three", 3)):
            f.createDimension(name, n)
        for name, dims, kind, long_name, units in MESH_VARS + MESH_OPTIONAL:
            a = getattr(mesh, name, None)
            if a is None:
                continue
            v = f.createVariable(name, "f8" if kind == "d" else "i4", tuple(reversed(dims)))
            v.long_name = long_name
            if units:
                v.units = units
            v[...] = _to_file(np.asarray(a, dtype=np.float64 if kind == "d" else np.int32))
        if frames:
            if zeta is None:
                raise ValueError("write_restart: frames need zeta")
            zeta = np.asarray(zeta, np.float64)
            f.createDimension("zeta", zeta.size)
            f.createDimension("month", 12)
            vt = f.createVariable("time", "f8", ("time",))
            vt.long_name, vt.units = "Time", "years"
            vz = f.createVariable("zeta", "f8", ("zeta",))
            vz.long_name = "Vertical scaled coordinate"
            vz.units = "unitless (0 = ice surface, 1 = bedrock)"
            vz[:] = zeta
            vm = f.createVariable("month", "f8", ("month",))
            vm.long_name, vm.units = "Month", "1-12"
            vm[:] = np.arange(1, 13, dtype=np.float64)
            shapes = {"vi": nV, "zeta": zeta.size, "month": 12}
            vs = {}
            for name, dims, long_name, units in INIT_VARS:
                v = f.createVariable(name, "f8", tuple(reversed(dims)))
                v.long_name, v.units = long_name, units
                vs[name] = (v, tuple(shapes[d] for d in dims[:-1]))
            for k, (t, fields) in enumerate(frames):
                vt[k] = float(t)
                for name, (v, shape) in vs.items():
                    a = np.zeros(shape) if name not in fields else np.asarray(fields[name], np.float64)
                    if a.shape != shape:
                        raise ValueError(f"write_restart: {name} has shape {a.shape}, expected {shape}")
                    v[k] = _to_file(a)


def _require(f, dims, variables):
    for d in dims:
        if d not in f.dimensions:
            raise ValueError(f"restart file lacks dimension '{d}'")
    for name, fdims in variables:
        if name not in f.variables:
            raise ValueError(f"restart file lacks variable '{name}'")
        if tuple(f.variables[name].dimensions) != tuple(reversed(fdims)):
            raise ValueError(f"restart variable '{name}' has dimensions {f.variables[name].dimensions}")


def read_restart_mesh(path) -> Mesh:
    """The primary mesh of a restart file (inquire_restart_file_mesh +
    read_restart_file_mesh), with the domain bounds taken from the vertices.
    A is read when the file holds it; Cw is left zero (both are secondary data
    that ufe_mesh_geometry recomputes)."""
    with netcdf_file(path, "r", mmap=False) as f:
        _require(f, ("vi", "ti", "ci", "two", "three"), [(n, d) for n, d, *_ in MESH_VARS])
        arr = {}
        for name, dims, kind, *_ in MESH_VARS:
            arr[name] = _from_file(f.variables[name][...], np.float64 if kind == "d" else np.int32)
        nV, nCm = arr["V"].shape[0], int(f.dimensions["ci"])
        A = (_from_file(f.variables["A"][...], np.float64) if "A" in f.variables
             else np.zeros(nV))
    V = arr["V"]
    return Mesh(xmin=float(V[:, 0].min()), xmax=float(V[:, 0].max()), ymin=float(V[:, 1].min()),
                ymax=float(V[:, 1].max()), nC_mem=nCm, V=V, nC=arr["nC"], C=arr["C"],
                niTri=arr["niTri"], iTri=arr["iTri"], edge_index=arr["edge_index"], Tri=arr["Tri"],
                TriC=arr["TriC"], Tricc=arr["Tricc"], Tri_edge_index=arr["Tri_edge_index"], A=A,
                Cw=np.zeros((nV, nCm), order="F"))


def read_restart_init(path, time_to_restart_from, nZ=None):
    """Hi, Hb, Hs, Ti, U_SSA and V_SSA (plus FirnDepth and MeltPreviousYear)
    at the frame nearest to time_to_restart_from: the first frame with the
    smallest |time - t| (read_restart_file_init).  Raises where the reference
    aborts: nZ differing from the file's, or a time outside the file's range."""
    with netcdf_file(path, "r", mmap=False) as f:
        _require(f, ("zeta", "time", "month"), [("zeta", ("zeta",)), ("time", ("time",)),
                                                ("month", ("month",))]
                 + [(n, d) for n, d, *_ in INIT_VARS])
        if nZ is not None and int(f.dimensions["zeta"]) != nZ:
            raise ValueError("nZ in restart file doesnt match nZ in config")
        time = np.array(f.variables["time"][:], np.float64)
        t = float(time_to_restart_from)
        if t < time.min() or t > time.max():
            raise ValueError(f"time_to_restart_from {t} outside range of restart file "
                             f"[{time.min()}, {time.max()}]")
        ti, dt_min = 0, 1e8
        for k, tk in enumerate(time):
            if abs(tk - t) < dt_min:
                ti, dt_min = k, abs(tk - t)
        out = {"time": float(time[ti]), "zeta": np.array(f.variables["zeta"][:], np.float64)}
        for name, *_ in INIT_VARS:
            out[name] = _from_file(f.variables[name][ti], np.float64)
    return out
