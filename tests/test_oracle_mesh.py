"""Pins the oracle's mesh, CSR, neighbour and quadrature restatements (CPU only).

Known answers come from the reference's own fixtures (the SALOME DAT meshes it ships, read under the
reader's rules) and from closed forms: nnz = V + 2E, polynomial exactness of the Gauss rules,
J * J^-1 = I and the shoelace area.
"""
import math

import numpy as np
import pytest

from conftest import golden
from helpers import face_neighbors_by_hashing

# (file, dim, V, T, CSR nnz, neighbour entries, max neighbours) -- counts derived from the fixtures
FIXTURES = [
    ("TestProblem2.dat", 2, 917, 1712, 6173, 19682, 15),
    ("ChannelMesh.dat", 2, 4599, 8756, 31307, 101602, 14),
    ("CylinderMesh.dat", 2, 2838, 5351, 19216, 61794, 15),
    ("CylinderMesh3.dat", 2, 17473, 34185, 120789, 404512, 15),
    ("BoxMesh.dat", 3, 1361, 5393, 16541, None, None),
]


def unique_edges(cells):
    nper = cells.shape[1]
    e = set()
    for c in cells:
        for a in range(nper):
            for b in range(a + 1, nper):
                u, v = int(c[a]), int(c[b])
                e.add((min(u, v), max(u, v)))
    return len(e)


@pytest.mark.parametrize("fname,dim,V,T,nnz,nbr,maxnbr", FIXTURES)
def test_fixture_counts(oracle, fname, dim, V, T, nnz, nbr, maxnbr):
    v, c, ndev = oracle.load_dat(golden(fname), dim)
    assert v.shape[0] == V and c.shape[0] == T and ndev == T
    rp, col = oracle.csr_pattern(V, c)
    assert col.shape[0] == nnz
    # closed form: diagonal + both directions of every unique edge
    assert nnz == V + 2 * unique_edges(c)
    # sorted, diagonal present, symmetric
    for i in range(V):
        row = col[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0) and i in row
    pairs = set(zip(np.repeat(np.arange(V), np.diff(rp)).tolist(), col.tolist()))
    assert all((j, i) in pairs for (i, j) in pairs)
    if nbr is not None and T < 20000:
        off, idx = oracle.neighbors_bruteforce(c)
        assert idx.shape[0] == nbr and np.diff(off).max() == maxnbr


def test_dat_trailing_newline(oracle, tmp_path):
    """Reference reader (src/mesh_2d.cu:56-72): trailing whitespace costs the device mesh its last cell."""
    src = open(golden("TestProblem2.dat")).read()
    p = tmp_path / "nl.dat"
    p.write_text(src + "\n")
    v, c, ndev = oracle.load_dat(str(p), 2)
    assert c.shape[0] == 1712 and ndev == 1711


@pytest.mark.parametrize("index,degree", [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 5), (6, 6), (7, 7)])
def test_face_rules_exact(oracle, index, degree):
    """sum w xi^a eta^b = 2 a! b! / (a+b+2)! on the reference triangle (weights normalised to area 1)."""
    q = oracle.quadrature(2, index)
    assert abs(q[:, 3].sum() - 1.0) < 1e-14
    assert np.allclose(q[:, 0] + q[:, 1] + q[:, 2], 1.0, atol=1e-15)
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            exact = 2.0 * math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            got = np.sum(q[:, 3] * q[:, 0] ** a * q[:, 1] ** b)
            assert abs(got - exact) < 2e-12, (index, a, b, got, exact)


@pytest.mark.parametrize("index", range(5))
def test_edge_rules_exact(oracle, index):
    q = oracle.quadrature(1, index)
    n = q.shape[0]
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(q[:, 1] * q[:, 0] ** k) - exact) < 1e-12


@pytest.mark.parametrize("index,degree", [(0, 2), (1, 3)])
def test_tet_rules_exact(oracle, index, degree):
    """sum w x^a y^b z^c = 6 a! b! c! / (a+b+c+3)! on the reference tetrahedron (volume-normalised)."""
    q = oracle.quadrature(3, index)
    assert abs(q[:, 4].sum() - 1.0) < 1e-12
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            for c in range(degree + 1 - a - b):
                exact = 6.0 * math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
                got = np.sum(q[:, 4] * q[:, 0] ** a * q[:, 1] ** b * q[:, 2] ** c)
                assert abs(got - exact) < 1e-10


def test_geometry2d_closed_forms(oracle):
    v, c, _ = oracle.load_dat(golden("CylinderMesh.dat"), 2)
    area, invJ = oracle.geometry2d(v, c)
    p = v[c]
    shoelace = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                            (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    assert np.allclose(area, shoelace, rtol=1e-13)
    J = np.stack([p[:, 0] - p[:, 2], p[:, 1] - p[:, 2]], axis=1)   # rows v31, v32
    prod = np.einsum("tij,tjk->tik", J, invJ.reshape(-1, 2, 2))
    assert np.allclose(prod, np.eye(2)[None], atol=1e-9)


def test_geometry3d_closed_forms(oracle):
    from gpupfem2_amd import meshgen
    xyz, tet = meshgen.box3d(4)
    vol, invJ = oracle.geometry3d(xyz, tet)
    assert np.allclose(vol, 1.0 / (6 * 64))
    assert abs(vol.sum() - 1.0) < 1e-13
    p = xyz[tet]
    J = np.stack([p[:, 0] - p[:, 3], p[:, 1] - p[:, 3], p[:, 2] - p[:, 3]], axis=1)
    prod = np.einsum("tij,tjk->tik", J, invJ.reshape(-1, 3, 3))
    assert np.allclose(prod, np.eye(3)[None], atol=1e-12)


def test_generated_meshes_closed_forms(oracle):
    from gpupfem2_amd import meshgen
    xy, tri = meshgen.unit_square(64)
    assert xy.shape[0] == 4225 and tri.shape[0] == 8192
    rp, col = oracle.csr_pattern(4225, tri)
    assert col.shape[0] == 29057          # V + 2E with E = 12,416
    xy, tri = meshgen.rectangle(0, 4, 0, 1, 400, 100, jitter=0.2, seed=2)
    rp, col = oracle.csr_pattern(xy.shape[0], tri)
    assert xy.shape[0] == 40501 and tri.shape[0] == 80000 and col.shape[0] == 281501
    xyz, tet = meshgen.box3d(8)
    rp, col = oracle.csr_pattern(xyz.shape[0], tet)
    E = 3 * 8 * 9 * 9 + 3 * 8 * 8 * 9 + 8 ** 3
    assert col.shape[0] == 9 ** 3 + 2 * E


def test_face_neighbors3d_pinned(oracle):
    """The oracle's all-pairs face test (src/mesh_3d.cu:37-82) on BoxMesh.dat equals the face-hashing
    construction, and the count of boundary faces (4T minus the neighbour entries) equals Euler's 2V_b - 4
    for the closed triangulated surface of the box, V_b its vertices on the bounding planes."""
    v, c, ndev = oracle.load_dat(golden("BoxMesh.dat"), 3)
    tet = c[:ndev]
    off, idx = oracle.face_neighbors3d(tet)
    hoff, hidx, interior = face_neighbors_by_hashing(tet)
    assert np.array_equal(off, hoff) and np.array_equal(idx, hidx)
    lo, hi = v.min(axis=0), v.max(axis=0)
    vb = int(np.sum(np.any((np.abs(v - lo) < 1e-12) | (np.abs(v - hi) < 1e-12), axis=1)))
    assert 4 * tet.shape[0] - idx.shape[0] == 2 * vb - 4
    assert idx.shape[0] == 2 * interior and np.diff(off).max() == 4


def test_face_neighbors3d_degenerate(oracle):
    """Cells with a repeated vertex follow the pair count of the reference's kernel: (a, a, b, c) against
    a cell with face (a, b, c) has 4 equal pairs and is not a neighbour; against a cell holding only a and b
    it has 2 + 1 = 3 and is one."""
    tet = np.array([[0, 1, 2, 3], [0, 1, 2, 4], [0, 0, 1, 5], [0, 0, 1, 2]], np.int32)
    off, idx = oracle.face_neighbors3d(tet)
    got = [sorted(idx[off[i]:off[i + 1]].tolist()) for i in range(4)]
    assert got == [[1, 2], [0, 2], [0, 1], []]
