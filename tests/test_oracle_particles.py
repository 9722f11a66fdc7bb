"""Pins the oracle's particle handler restatement (CPU only).

Philox4x32-10 against the published known-answer vectors of Random123; determineSubcell against brute
force point location in the sub-triangles; advection against the closed form of a uniform field;
projection against its invariants (sum of weights = number of particles, constant fields reproduced);
refill against completeness (no empty subcell remains).
"""
import numpy as np
import pytest

from conftest import golden


def test_philox_known_answers(oracle):
    # Random123 kat_vectors, philox4x32_10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, out in kat:
        assert tuple(int(v) for v in oracle.philox(ctr, key)) == out


@pytest.mark.parametrize("level", [1, 2, 3, 4])
def test_subcells(oracle, level):
    c = oracle.subcell_centers(level)
    assert c.shape == (level * level, 3)
    assert np.allclose(c.sum(axis=1), 1.0, atol=1e-15)
    for k in range(level * level):
        assert oracle.determine_subcell(level, *c[k]) == k
    # brute force: random points, located by testing every sub-triangle
    rng = np.random.default_rng(level)
    h = 1.0 / level
    tris = []
    for i in range(level):
        for j in range(2 * i + 1):
            xmin = (j // 2) * h
            xmax, ymin = xmin + h, (level - 1 - i) * h
            ymax = ymin + h
            v0 = (xmin, ymax)
            v1 = (xmin, ymin) if j % 2 == 0 else (xmax, ymax)
            v2 = (xmax, ymin)
            tris.append((v0, v1, v2))
    for _ in range(2000):
        r1, r2 = rng.random(2)
        s = np.sqrt(r1)
        l1, l2 = 1 - s, s * (1 - r2)
        l3 = 1 - l1 - l2
        k = oracle.determine_subcell(level, l1, l2, l3)
        (ax, ay), (bx, by), (cx, cy) = tris[k]
        d = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy)
        a = ((by - cy) * (l1 - cx) + (cx - bx) * (l2 - cy)) / d
        b = ((cy - ay) * (l1 - cx) + (ax - cx) * (l2 - cy)) / d
        assert a > -1e-12 and b > -1e-12 and 1 - a - b > -1e-12


def small_mesh(oracle):
    v, c, _ = oracle.load_dat(golden("CylinderMesh.dat"), 2)
    area, invJ = oracle.geometry2d(v, c)
    off, idx = oracle.neighbors_bruteforce(c)
    return v, c, area, invJ, off, idx


@pytest.mark.parametrize("mode,ppc", [(0, 4), (1, 20)])
def test_seed(oracle, mode, ppc):
    v, c, area, invJ, off, idx = small_mesh(oracle)
    ph = oracle.Particles(v, c, invJ, off, idx, 2)
    ph.seed(mode, ppc, 103)
    st = ph.get()
    T = c.shape[0]
    assert st["cell"].shape[0] == T * ppc
    assert np.array_equal(st["cell"], np.repeat(np.arange(T), ppc))
    assert np.array_equal(st["id"], np.arange(T * ppc))
    lam = st["lam"]
    assert np.all(lam > -1e-15) and np.allclose(lam.sum(1), 1.0, atol=1e-14)
    pos = np.einsum("pk,pkd->pd", lam, v[c[st["cell"]]])
    assert np.allclose(pos, st["pos"], atol=1e-14)


def test_uniform_advection_is_translation(oracle):
    v, c, area, invJ, off, idx = small_mesh(oracle)
    ph = oracle.Particles(v, c, invJ, off, idx, 2)
    ph.seed(1, 8, 7)
    before = ph.get()
    nv = v.shape[0]
    u0, u1 = np.full(nv, 0.3), np.full(nv, -0.1)
    dt = 1e-3
    ph.advect(u0, u1, dt, 3)
    after = ph.get()
    # every original particle moved by dt * u (up to rounding); match by ID
    pos_by_id = dict(zip(after["id"].tolist(), after["pos"]))
    moved = np.array([pos_by_id[i] for i in before["id"].tolist() if i in pos_by_id])
    orig = np.array([p for i, p in zip(before["id"].tolist(), before["pos"]) if i in pos_by_id])
    assert np.allclose(moved - orig, [0.3 * dt, -0.1 * dt], atol=1e-14)
    # each particle lies in the cell it is assigned to and lambda sums to one
    lam = after["lam"]
    assert np.all(lam > -2.1e-6) and np.allclose(lam.sum(1), 1.0, atol=1e-14)
    pos = np.einsum("pk,pkd->pd", lam, v[c[after["cell"]]])
    assert np.allclose(pos, after["pos"], atol=1e-12)
    # stable order by cell
    assert np.all(np.diff(after["cell"]) >= 0)


def test_refill_and_projection_invariants(oracle):
    v, c, area, invJ, off, idx = small_mesh(oracle)
    nv, T = v.shape[0], c.shape[0]
    ph = oracle.Particles(v, c, invJ, off, idx, 4)
    ph.seed(1, 3, 11)      # 3 random particles per cell: most of the 16 subcells start empty
    u0, u1 = np.full(nv, 2.0), np.full(nv, -1.0)
    ph.init_velocity(u0, u1)
    st = ph.get()
    assert np.allclose(st["vel"], [2.0, -1.0])
    ph.advect(np.zeros(nv), np.zeros(nv), 1e-3, 3)
    deleted, added = ph.stats()
    st = ph.get()
    assert deleted == 0 and added > 0
    # completeness: every subcell of every cell holds a particle
    occ = np.zeros((T, 16), bool)
    for cell, lam in zip(st["cell"], st["lam"]):
        occ[cell, oracle.determine_subcell(4, *lam)] = True
    assert occ.all()
    # refilled particles got the grid velocity (here the zero field), the old ones kept theirs
    new = st["id"] >= 3 * T
    assert np.all(st["vel"][new] == 0.0) and np.allclose(st["vel"][~new], [2.0, -1.0])
    # projection: constant particle velocity is reproduced exactly where all weights agree
    ph2 = oracle.Particles(v, c, invJ, off, idx, 2)
    ph2.seed(1, 6, 5)
    ph2.init_velocity(np.full(nv, 0.75), np.full(nv, -0.25))
    p0, p1 = ph2.project(nv)
    assert np.allclose(p0, 0.75, rtol=1e-14) and np.allclose(p1, -0.25, rtol=1e-14)
