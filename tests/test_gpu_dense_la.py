"""Dense BLAS-1/2 of LinearAlgebra (src/linear_algebra/linear_algebra.h:27-54): dot, the Euclidean norm
(normSquared, which is cublasDnrm2), the upper-triangular solve and the general matrix-vector product, with
their scalars in device memory, against exact host sums (math.fsum), scipy's triangular solve and numpy.
Covers the BLAS conventions the reference inherits (n = 0, beta = 0 not reading y, alpha = 0 not reading A,
entries outside the triangle never read), the norm's overflow/underflow rescaling, run-to-run bit
reproducibility and argument errors."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def la(ctx):
    from gpupfem2_amd import capi
    return capi.LinearAlgebra(ctx)


@pytest.mark.parametrize("n", [0, 1, 255, 256, 1025, 4097, 262145, 4_000_037])
def test_dot_and_norm(ctx, n):
    print("dot/nrm2 n =", n, flush=True)
    rng = np.random.default_rng(n + 1)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    L = la(ctx)
    dx, dy = ctx.to_device(x), ctx.to_device(y)
    res = ctx.zeros(4)
    L.dot(dx, dy, res)                  # lengths from the arrays
    L.norm_squared(dx, res.ptr + 8)     # raw device addresses, the pointer overloads
    L.dot(dx.ptr, dy.ptr, res.ptr + 16, n)
    L.norm_squared(dx, res.ptr + 24)
    got = res.download()
    ref_dot = math.fsum(x * y)
    ref_nrm = math.sqrt(math.fsum(x * x))
    assert abs(got[0] - ref_dot) <= 1e-14 * max(np.abs(x * y).sum(), 1e-300), (got[0], ref_dot)
    assert abs(got[1] - ref_nrm) <= 1e-14 * max(ref_nrm, 1e-300), (got[1], ref_nrm)
    # the same call twice gives the same bits (fixed grid and summation order)
    assert got[0] == got[2] and got[1] == got[3]
    if n == 0:
        assert got[0] == 0.0 and got[1] == 0.0


def test_dot_length_mismatch(ctx):
    from gpupfem2_amd import capi
    with pytest.raises(capi.PfemError):
        la(ctx).dot(ctx.zeros(3), ctx.zeros(4), ctx.zeros(1))


@pytest.mark.parametrize("scale", [1e200, 1e-200, 1e-310, 1.0])
def test_norm_extreme_range(ctx, scale):
    """Squares that overflow (1e200) or underflow (1e-200, subnormal 1e-310) take the rescaled pass; the
    norm is the scale times the norm of the unscaled vector."""
    print("nrm2 scale", scale, flush=True)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(100_003)
    ref = scale * math.sqrt(math.fsum(x * x))
    xs = x * scale
    res = ctx.zeros(1)
    la(ctx).norm_squared(ctx.to_device(xs), res)
    got = res.download()[0]
    # subnormal inputs carry fewer significant bits themselves
    rel = 1e-13 if scale >= 1e-300 else 1e-5
    assert abs(got - ref) <= rel * ref, (got, ref)


def test_norm_special_values(ctx):
    L = la(ctx)
    res = ctx.zeros(3)
    L.norm_squared(ctx.to_device(np.zeros(1000)), res.ptr)
    x = np.ones(1000)
    x[500] = np.inf
    L.norm_squared(ctx.to_device(x), res.ptr + 8)
    x[500] = np.nan
    L.norm_squared(ctx.to_device(x), res.ptr + 16)
    got = res.download()
    assert got[0] == 0.0 and np.isposinf(got[1]) and np.isnan(got[2]), got


@pytest.mark.parametrize("n", [1, 5, 37, 256, 257, 700])
def test_upper_triangular_solve(ctx, n):
    import scipy.linalg as sl
    print("trsv n =", n, flush=True)
    rng = np.random.default_rng(n)
    lda = n + 3
    a = np.full((lda, n), np.nan, order="F")    # column-major storage; NaN wherever it must not be read
    u = np.triu(rng.uniform(-1.0, 1.0, (n, n)))
    u[np.diag_indices(n)] = rng.uniform(1.0, 2.0, n) * np.sign(rng.uniform(-1, 1, n))
    a[:n, :] = np.where(np.triu(np.ones((n, n), bool)), u, np.nan)
    b = rng.standard_normal(n)
    ref = sl.solve_triangular(u, b, lower=False)
    da = ctx.to_device(np.asfortranarray(a).ravel(order="F"))
    dx = ctx.to_device(b)
    la(ctx).solve_upper_triangular(da, dx, n, lda)
    got = dx.download()
    assert np.all(np.isfinite(got))
    err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1.0)
    assert err <= 1e-11, err


@pytest.mark.parametrize("m,n", [(1, 1), (1000, 5), (100_003, 37), (500_000, 30)])
def test_general_mv(ctx, m, n):
    print("gemv", m, n, flush=True)
    rng = np.random.default_rng(m + n)
    lda = m + 1
    a = np.zeros((lda, n), order="F")
    a[:m] = rng.standard_normal((m, n))
    a[m:] = np.nan                               # the row beyond m in each column is never read
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    da = ctx.to_device(a.ravel(order="F"))
    dxv = ctx.to_device(x)
    L = la(ctx)
    bound = np.abs(a[:m]) @ np.abs(x) + np.abs(y)
    # as the GMRES update of the reference calls it: alpha = beta = 1
    for alpha, beta in ((1.0, 1.0), (2.0, -0.5)):
        ab = ctx.to_device(np.array([alpha, beta]))
        dy = ctx.to_device(y)
        L.general_mv(da, dxv, dy, ab.ptr, ab.ptr + 8, m, n, lda)
        ref = alpha * (a[:m] @ x) + beta * y
        assert np.all(np.abs(dy.download() - ref) <= 1e-13 * (abs(alpha) + abs(beta)) * bound), (alpha, beta)


def test_general_mv_blas_conventions(ctx):
    m, n = 3000, 4
    rng = np.random.default_rng(3)
    a = rng.standard_normal((m, n))
    x = rng.standard_normal(n)
    ynan = np.full(m, np.nan)
    y = rng.standard_normal(m)
    L = la(ctx)
    da = ctx.to_device(np.asfortranarray(a).ravel(order="F"))
    danan = ctx.to_device(np.full(m * n, np.nan))
    dxv = ctx.to_device(x)

    def run(dA, yin, alpha, beta, nn=n):
        ab = ctx.to_device(np.array([alpha, beta]))
        dy = ctx.to_device(yin)
        L.general_mv(dA, dxv, dy, ab.ptr, ab.ptr + 8, m, nn, m)
        return dy.download()

    # beta = 0: y is written, not read (a NaN there does not propagate)
    assert np.allclose(run(da, ynan, 1.5, 0.0), 1.5 * (a @ x), rtol=1e-13, atol=1e-13)
    # alpha = 0: A and x are not read
    assert np.array_equal(run(danan, y, 0.0, 0.5), 0.5 * y)
    # alpha = 0 and beta = 1: y untouched, bit for bit
    assert np.array_equal(run(danan, y, 0.0, 1.0), y)
    # n = 0: y untouched whatever beta is
    assert np.array_equal(run(da, y, 1.0, 0.25, nn=0), y)


def test_dense_argument_errors(ctx):
    from gpupfem2_amd import capi
    L = la(ctx)
    a, v, r = ctx.zeros(16), ctx.zeros(4), ctx.zeros(2)
    for call in (lambda: L.solve_upper_triangular(a, v, 4, 3),             # lda < n
                 lambda: L.general_mv(a, v, v, r.ptr, r.ptr + 8, 4, 4, 3),   # lda < m
                 lambda: L.dot(v.ptr, v.ptr, r.ptr, -1),
                 lambda: L.norm_squared(v, 0, 4)):                          # null result
        with pytest.raises(capi.PfemError) as e:
            call()
        assert e.value.code == capi.PFEM_EINVAL
