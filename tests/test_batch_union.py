"""The batch union mesh: joining separately built members must give exactly
what the oracle builds from the union's primary mesh, and the joined colour
lists must colour the union validly."""
import numpy as np
import pytest

from ufemism_amd import batch, geometry, meshgen

po = pytest.importorskip("oracle.pyoracle")


@pytest.fixture(scope="module")
def union():
    meshes = [geometry.dome(600, seed=31).mesh, geometry.marine(900, seed=32).mesh,
              meshgen.dummy_mesh(), geometry.dome(400, seed=33).mesh]
    built = [po.OracleMesh(m).build() for m in meshes]
    for b in built:
        assert b.colour_status == 0
    members = batch.members_of(built)
    return built, members, batch.join_meshes(built, members)


def test_union_secondary_data_equals_a_fresh_build(union):
    built, members, ua = union
    fresh = po.OracleMesh(ua.mesh, geometry=False).build(colour=False)
    assert fresh.nAc == ua.nAc
    for k in ("iAci", "Aci", "edge_index_Ac", "nCAaAc", "CAaAc"):
        assert np.array_equal(fresh[k], ua[k]), k
    for k in ("VAc", "Nx_Ac", "Ny_Ac", "No_Ac", "Np_Ac", "VAaAc", "Nx", "Ny", "Nxx", "Nxy", "Nyy",
              "Nx_AaAc", "Ny_AaAc", "Nxx_AaAc", "Nxy_AaAc", "Nyy_AaAc"):
        assert np.array_equal(fresh[k], ua[k], equal_nan=True), k


def test_union_colouring_is_valid_and_block_local(union):
    built, members, ua = union
    nA = ua.nVAaAc
    lists = [ua["colour_vi"][: ua.colour_nV[c], c] for c in range(5)]
    allp = np.sort(np.concatenate(lists))
    assert np.array_equal(allp, np.arange(1, nA + 1))
    col = np.zeros(nA + 1, int)
    for c in range(5):
        col[lists[c]] = c + 1
    assert np.array_equal(col[1:], ua["colour"])
    owner = np.zeros(nA, int)
    for b, m in enumerate(members):
        owner[batch.aaac_map(m, ua.nV)] = b
    C, nC = ua["CAaAc"], ua["nCAaAc"]
    for a in range(nA):
        nb = C[a, : nC[a]] - 1
        assert np.all(owner[nb] == owner[a])          # no edge between members
        assert np.all(col[nb + 1] != col[a + 1])      # proper colouring


def test_union_keeps_every_members_corners(union):
    built, members, ua = union
    e = ua.mesh.edge_index
    corners = np.flatnonzero(np.isin(e, (2, 4, 6, 8)))
    expect = np.concatenate([m.vbase + np.arange(4) for m in members])
    assert np.array_equal(corners, expect)
