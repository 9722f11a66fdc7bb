"""The C++ facade (include/gpupfem2.hpp) on the GPU: the reference's Poisson2D driver sequence written
against it reproduces the oracle's run of the same sequence, and the facade's own checks pass (a
Preconditioner subclass plugged into SolverCG, IC0 orderings, restarted GMRES, particles, parameters,
quadrature)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import golden, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "gpupfem2_amd", "examples")


def run(args, cwd):
    print("[facade] running", " ".join(os.path.basename(a) for a in args), flush=True)
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout + p.stderr


def test_facade_checks(tmp_path):
    rc, out = run([os.path.join(EX, "facade_check"), golden("ChannelMesh.dat")], tmp_path)
    print(out)
    assert rc == 0 and "FAIL" not in out and out.strip().endswith("ok 0 failures")


def test_poisson2d_driver_matches_oracle(tmp_path, oracle):
    rc, out = run([os.path.join(EX, "poisson2d"), golden("TestProblem2.dat"), str(tmp_path)], tmp_path)
    print(out)
    assert rc == 0, out
    O = oracle
    xy, cells, ndev = O.load_dat(golden("TestProblem2.dat"))
    tri = cells[:ndev]
    area, invJ = O.geometry2d(xy, tri)
    rp, col = O.csr_pattern(xy.shape[0], tri)
    val, rhs = O.assemble_poisson2d(xy, tri, area, invJ, rp, col, coeff=0.5, rhs_kind=0, quad=1)
    eps = 2e-6
    node, gval = [], []
    for i, (x, y) in enumerate(xy):
        for cond, g in ((abs(x + 1.0) < eps, -1.0), (abs(x - 1.0) < eps, 1.0), (abs(y) < eps, 0.0),
                        (abs(y - 1.0) < eps, 2.0)):
            if cond:
                node.append(i)
                gval.append(g)
                break
    O.dirichlet_apply(rp, col, val, rhs, np.asarray(node, np.int32), np.asarray(gval), 0)
    # exported matrix and vectors carry 6 significant digits (default stream formatting, as the reference)
    m = np.loadtxt(tmp_path / "matrix.dat")
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    assert np.array_equal(m[:, 0].astype(int) - 1, rows) and np.array_equal(m[:, 1].astype(int) - 1, col)
    assert np.allclose(m[:, 2], val, rtol=1e-5, atol=1e-12)
    assert rel_l2(np.loadtxt(tmp_path / "rhs.dat"), rhs) < 1e-5
    x = np.zeros(rhs.size)
    st = O.pcg(rp, col, val, x, rhs, 1e-8, 1000, 0, "jacobi")
    assert st.converged
    assert rel_l2(np.loadtxt(tmp_path / "solution_cg.dat"), x) < 1e-5
    O.gmres(rp, col, val, x, rhs, 1e-8, 1000, 1000, 0, "jacobi")
    assert rel_l2(np.loadtxt(tmp_path / "solution.dat"), x) < 1e-5
    assert (tmp_path / "solution.vtu").stat().st_size > 1000
