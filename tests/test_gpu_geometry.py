"""Primary-mesh geometry on the GPU (ufe_mesh_geometry) against the oracle:
circumcentres, Voronoi cell areas with the boundary cropping, connection
widths, triangle areas and resolution, on the reference's dummy mesh and on
generator meshes; the degenerate-triangle error path; and mesh_build taking
its A and Cw from it."""
import numpy as np
import pytest

from ufemism_amd import abi, geometry, meshgen

pytestmark = pytest.mark.gpu

TOL = 1e-14   # elementwise relative; only sqrt and division may round differently


def _oracle():
    from oracle import pyoracle
    return pyoracle


MESHES = {
    "dummy": lambda: meshgen.dummy_mesh(),
    "dome8k": lambda: geometry.dome(8000, seed=1).mesh,
    "marine20k": lambda: geometry.marine(20000, seed=5).mesh,
}


def _cmp(name, a, b, scale=None):
    a, b = np.asarray(a), np.asarray(b)
    s = np.maximum(np.abs(b), 1e-300) if scale is None else scale
    e = float(np.max(np.abs(a - b) / s)) if a.size else 0.0
    print(f"  {name}: bit-exact={np.array_equal(a, b)} max rel {e:.1e}", flush=True)
    return e


@pytest.mark.parametrize("name", list(MESHES))
def test_geometry_matches_oracle(ctx, name):
    print(f"[geometry] {name}", flush=True)
    m = MESHES[name]()
    ref = _oracle().mesh_geometry(m)
    got = ctx.mesh_geometry(m)
    st = got["stats"]
    print(f"  area error {st.area_error:.1e} resolution {st.resolution_min:.4g}..{st.resolution_max:.4g}",
          flush=True)
    assert st.invalid_circumcentres == 0 and st.voronoi_outside == 0 and st.connection_failures == 0
    assert st.rotated_triangles == ref["rotated"]
    assert np.array_equal(got["Tri"], ref["Tri"])
    L = max(m.xmax - m.xmin, m.ymax - m.ymin)
    assert _cmp("Tricc", got["Tricc"], ref["Tricc"], scale=L) <= TOL
    for k in ("A", "TriA", "R"):
        assert _cmp(k, got[k], ref[k]) <= TOL, k
    assert _cmp("Cw", got["Cw"], ref["Cw"], scale=np.maximum(np.abs(ref["Cw"]), L * 1e-12)) <= TOL
    # same operations in the same order, no contraction: identical bits
    for k in ("Tricc", "A", "TriA", "R", "Cw"):
        assert np.array_equal(got[k], ref[k]), k
    assert st.area_error < 1e-12
    assert st.resolution_min == ref["resolution_min"] or abs(st.resolution_min / ref["resolution_min"] - 1) < TOL


def test_given_circumcentres_and_mesh_build(ctx):
    m = geometry.dome(2000, seed=3).mesh
    ref = _oracle().mesh_geometry(m, compute_Tricc=False)
    got = ctx.mesh_geometry(m, compute_Tricc=False)
    assert np.array_equal(got["Tricc"], np.asarray(m.Tricc))
    assert _cmp("A (given Tricc)", got["A"], ref["A"]) <= TOL
    ma = ctx.mesh_build(m)
    full = ctx.mesh_geometry(m)
    assert np.array_equal(ma["A"], full["A"]) and np.array_equal(ma["Cw"], full["Cw"])


def test_degenerate_triangle_is_an_error(ctx):
    m = meshgen.dummy_mesh()
    m.V = np.asfortranarray(m.V.copy())
    m.V[4] = [0.0, -1.0]
    with pytest.raises(abi.UFEError) as e:
        ctx.mesh_geometry(m)
    assert "circumcentre" in str(e.value)


def test_mesh_read_from_a_restart_file(ctx, tmp_path):
    """The restart path of the reference: primary mesh from the file, stored
    circumcentres, then the secondary data on the GPU."""
    from ufemism_amd import mesh_io
    m = geometry.marine(20000, seed=5).mesh
    ref = _oracle().mesh_geometry(m)
    m.Tricc = ref["Tricc"]
    p = tmp_path / "restart.nc"
    mesh_io.write_restart(p, m)
    m2 = mesh_io.read_restart_mesh(p)
    got = ctx.mesh_geometry(m2, compute_Tricc=False)
    for k in ("A", "Cw", "TriA", "R"):
        assert np.array_equal(got[k], ref[k]), k
    ma = ctx.mesh_build(m2)
    om = _oracle().OracleMesh(m2).build()
    for k in ("iAci", "Aci", "colour", "colour_vi"):
        assert np.array_equal(ma[k], om[k]), k
