"""Host-only mesh renumbering for memory locality (pfem_mesh_locality_order), without a GPU: the orders are
permutations, the renumbered mesh is the same mesh (coordinates, orientation, boundary), its vertex
adjacency has the small bandwidth of a reverse Cuthill-McKee order even when the input is numbered at
random, and the cells are listed by their smallest new vertex index."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee


def adjacency(nv, cells):
    k = cells.shape[1]
    r = np.repeat(cells, k, axis=1).ravel()
    c = np.tile(cells, (1, k)).ravel()
    return sp.csr_matrix((np.ones(r.size), (r, c)), shape=(nv, nv))


def bandwidth(nv, cells):
    a = adjacency(nv, cells).tocoo()
    return int(np.abs(a.row - a.col).max())


def shuffled(xy, cells, seed):
    rng = np.random.default_rng(seed)
    new_id = rng.permutation(xy.shape[0])
    xy2 = np.empty_like(xy)
    xy2[new_id] = xy
    return xy2, new_id[cells][rng.permutation(cells.shape[0])].astype(np.int32)


def signed_volumes(x, cells):
    p = x[cells]
    if cells.shape[1] == 3:
        return 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                      - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    return np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])) / 6.0


@pytest.mark.parametrize("dim", [2, 3])
def test_locality_order(dim):
    from gpupfem2_amd import capi, meshgen
    if dim == 2:
        xy, cells = meshgen.unit_square(40)
    else:
        xy, cells = meshgen.box3d(8)
    xs, cs = shuffled(xy, cells, seed=dim)
    nv = xs.shape[0]
    nodes, order = capi.locality_order(nv, cs)
    assert np.array_equal(np.sort(nodes), np.arange(nv))
    assert np.array_equal(np.sort(order), np.arange(cs.shape[0]))
    xr, cr, _, _ = capi.renumber_mesh(xs, cs)
    # the same mesh: the same cells as vertex coordinates, the same orientations
    key = lambda x, c: np.sort(np.round(x[c].reshape(c.shape[0], -1), 12).view(np.dtype((np.void, 8 * c.shape[1] * x.shape[1]))).ravel())
    assert np.array_equal(key(xs, cs), key(xr, cr))
    assert np.allclose(np.sort(signed_volumes(xs, cs)), np.sort(signed_volumes(xr, cr)))
    assert np.all(signed_volumes(xr, cr) > 0) == np.all(signed_volumes(xs, cs) > 0)
    # bandwidth: from the random numbering's O(n) to an RCM-sized one (scipy's RCM as the yardstick)
    b_in, b_out = bandwidth(nv, cs), bandwidth(nv, cr)
    p = reverse_cuthill_mckee(adjacency(nv, cs).tocsr(), symmetric_mode=True)
    inv = np.empty(nv, np.int64)
    inv[p] = np.arange(nv)
    b_ref = bandwidth(nv, inv[cs])
    assert b_in > 5 * b_out, (b_in, b_out)
    assert b_out <= 1.5 * b_ref + 2, (b_out, b_ref)
    # cells listed by their smallest new vertex index
    first = np.sort(cr, axis=1)[:, 0]
    assert np.all(np.diff(first) >= 0)


def test_locality_order_rejects():
    from gpupfem2_amd import capi
    with pytest.raises(capi.PfemError):
        capi.locality_order(3, np.array([[0, 1, 5]], np.int32))   # vertex out of range
    with pytest.raises(capi.PfemError):
        capi.locality_order(4, np.zeros((2, 5), np.int32))        # 5 vertices per cell
