"""pfem_body_forces_noalloc (the allocation-free variant of pfem_body_forces) against pfem_body_forces
through the C-ABI: the same bits on two calls in a row, the second on a smaller mesh (the static scratch
array keeps the first call's contents), and zeros for a body ID no edge carries."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case_fields(ctx, h, seed):
    from gpupfem2_amd import capi, cases
    case = cases.cylinder_case(h=h)
    m = capi.Mesh2D(ctx, case.xy, case.tri, fill_neighbors=False)
    m.set_edge_ids(case.edge_ids)
    rng = np.random.default_rng(seed)
    nv = case.xy.shape[0]
    u0, u1, p = (ctx.to_device(rng.standard_normal(nv)) for _ in range(3))
    prm = capi.make_params(case.rho, case.mu, case.dt, case.scheme, case.traction_mask, case.normals)
    return case, m, u0, u1, p, prm


def test_body_forces_noalloc_matches(ctx):
    from gpupfem2_amd import capi
    for h, seed in ((1.0 / 112, 1), (1.0 / 16, 2)):
        case, m, u0, u1, p, prm = _case_fields(ctx, h, seed)
        print("body forces noalloc", case.name, flush=True)
        ref = capi.body_forces(m, case.body_id, case.body_center, u0, u1, p, prm)
        got = capi.body_forces_noalloc(m, case.body_id, case.body_center, u0, u1, p, prm)
        assert np.any(ref != 0.0)
        assert np.array_equal(got, ref), (got, ref)
    none = capi.body_forces_noalloc(m, 7, case.body_center, u0, u1, p, prm)
    assert np.array_equal(none, np.zeros(4))
