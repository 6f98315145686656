"""Restart files in the reference's NetCDF layout: the structure that
inquire_restart_file_mesh / inquire_restart_file_init check
(src/netcdf_module.f90:3012-3103), round trips of the mesh and of the ice
state, the time-frame choice of read_restart_file_init (:3132-3185) and its
failure paths."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from ufemism_amd import geometry, mesh_io
from ufemism_amd.fields import ZETA_DEFAULT

po = pytest.importorskip("oracle.pyoracle")


@pytest.fixture(scope="module")
def mesh():
    return geometry.marine(3000, seed=4).mesh


def test_file_structure_matches_the_reference_reader(tmp_path, mesh):
    p = tmp_path / "restart.nc"
    H = np.linspace(0, 1000, mesh.nV)
    mesh_io.write_restart(p, mesh, zeta=ZETA_DEFAULT, frames=[(0.0, {"Hi": H})])
    with netcdf_file(p, "r", mmap=False) as f:
        assert f.version_byte == 1                                  # classic format
        dims = {k: (v if v is not None else None) for k, v in f.dimensions.items()}
        assert dims["vi"] == mesh.nV and dims["ti"] == mesh.nTri and dims["ci"] == mesh.nC_mem
        assert dims["two"] == 2 and dims["three"] == 3 and dims["zeta"] == len(ZETA_DEFAULT)
        assert dims["month"] == 12 and dims["time"] is None            # unlimited
        # Fortran (vi, two) is (two, vi) in the file; int variables are NF90_INT
        expect = {"V": (("two", "vi"), "d"), "Tri": (("three", "ti"), "i"), "nC": (("vi",), "i"),
                  "C": (("ci", "vi"), "i"), "niTri": (("vi",), "i"), "iTri": (("ci", "vi"), "i"),
                  "edge_index": (("vi",), "i"), "Tricc": (("two", "ti"), "d"),
                  "TriC": (("three", "ti"), "i"), "Tri_edge_index": (("ti",), "i"),
                  "Hi": (("time", "vi"), "d"), "Ti": (("time", "zeta", "vi"), "d"),
                  "FirnDepth": (("time", "month", "vi"), "d"), "time": (("time",), "d")}
        for name, (d, kind) in expect.items():
            v = f.variables[name]
            assert tuple(v.dimensions) == d, name
            assert v.typecode() == kind, name
        assert f.variables["V"].long_name == b"Vertex coordinates"
        # the file's V(two, vi) holds x in its first row
        assert np.array_equal(f.variables["V"][0], mesh.V[:, 0])


def test_mesh_round_trip_and_geometry(tmp_path, mesh):
    g = po.mesh_geometry(mesh)
    mesh.Tricc = g["Tricc"]                 # the circumcentres the reference stores
    mesh.R = g["R"]
    p = tmp_path / "mesh.nc"
    mesh_io.write_restart(p, mesh)
    m2 = mesh_io.read_restart_mesh(p)
    for k in ("V", "nC", "C", "niTri", "iTri", "edge_index", "Tri", "Tricc", "TriC", "Tri_edge_index"):
        a, b = getattr(m2, k), np.asarray(getattr(mesh, k))
        assert a.dtype == b.dtype and np.array_equal(a, b), k
        assert a.flags.f_contiguous, k
    assert (m2.xmin, m2.xmax, m2.ymin, m2.ymax) == (mesh.xmin, mesh.xmax, mesh.ymin, mesh.ymax)
    assert m2.nC_mem == mesh.nC_mem and m2.tol_dist == mesh.tol_dist
    # secondary data from the file's mesh with its stored circumcentres
    g2 = po.mesh_geometry(m2, compute_Tricc=False)
    for k in ("A", "Cw", "TriA", "R"):
        assert np.array_equal(g2[k], g[k]), k


def test_init_frames_and_time_choice(tmp_path, mesh):
    nV, nZ = mesh.nV, len(ZETA_DEFAULT)
    rng = np.random.default_rng(0)
    frames = []
    for t in (0.0, 100.0, 200.0):
        frames.append((t, {"Hi": rng.random(nV) + t, "Hb": rng.random(nV), "Hs": rng.random(nV),
                           "U_SSA": rng.random(nV), "V_SSA": rng.random(nV),
                           "Ti": np.asfortranarray(rng.random((nV, nZ)) + 250)}))
    p = tmp_path / "init.nc"
    mesh_io.write_restart(p, mesh, zeta=ZETA_DEFAULT, frames=frames)
    # 150 is equally far from 100 and 200: the first nearest frame wins
    r = mesh_io.read_restart_init(p, 150.0, nZ=nZ)
    assert r["time"] == 100.0
    for k in ("Hi", "Hb", "Hs", "U_SSA", "V_SSA", "Ti"):
        assert np.array_equal(r[k], frames[1][1][k]), k
    assert r["Ti"].shape == (nV, nZ) and r["Ti"].flags.f_contiguous
    assert np.all(r["FirnDepth"] == 0.0) and r["FirnDepth"].shape == (nV, 12)
    assert mesh_io.read_restart_init(p, 190.0)["time"] == 200.0
    with pytest.raises(ValueError, match="outside range"):
        mesh_io.read_restart_init(p, 250.0)
    with pytest.raises(ValueError, match="nZ"):
        mesh_io.read_restart_init(p, 0.0, nZ=nZ + 1)


def test_missing_variable_is_refused(tmp_path, mesh):
    p = tmp_path / "bad.nc"
    with netcdf_file(p, "w") as f:
        for name, n in (("vi", mesh.nV), ("ti", mesh.nTri), ("ci", mesh.nC_mem), ("two", 2), ("three", 3)):
            f.createDimension(name, n)
        f.createVariable("V", "f8", ("two", "vi"))[...] = mesh.V.T
    with pytest.raises(ValueError, match="lacks variable"):
        mesh_io.read_restart_mesh(p)
