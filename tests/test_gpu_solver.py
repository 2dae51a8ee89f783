"""Explicit CSR transpose, the device BiCGStab LinearSolver (solve and
solveT) and the reduced-functional gradient loop on the GPU: transpose
bitwise against femx's MatrixHandler::transpose (fixture), solves against
SciPy's sparse direct solver on the host, the adjoint gradient against
central differences of the reduced functional."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from conftest import golden
from femx_amd import assembly as A
from femx_amd.inverse import PoissonLogKappaResidual
from femx_amd.linalg import CsrMatrix, CsrPattern, MatrixHandler
from femx_amd.solver import BiCGStabSolver, NewtonStateSolver, ReducedFunctional, StateMisfit

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", ["rect", "p1_n3"])
def test_transpose_equals_reference(name):
    g = golden("transpose.npz")
    nr, nc = (int(v) for v in g[f"{name}_shape"])
    pat = CsrPattern(nr, nc, g[f"{name}_row_ptr"], g[f"{name}_col_ind"])
    M = CsrMatrix(pat, DEV, dev(g[f"{name}_vals"]))
    mh = MatrixHandler()
    T = mh.transpose(M)
    assert np.array_equal(host(T.vals), g[f"{name}_t_vals"])
    assert np.array_equal(T.pattern.row_ptr, g[f"{name}_t_row_ptr"])
    T2 = mh.transpose(M, T)
    assert T2 is T
    with pytest.raises(RuntimeError, match="in-place"):
        mh.transpose(M, M)


def poisson(n, seed=1):
    geom = A.structured_p1(n, DEV)
    x, y = geom.h_x, geom.h_y
    bc = A.unit_square_edge_nodes(x, y)
    rng = np.random.default_rng(seed)
    f = dev(rng.standard_normal(geom.num_nodes))
    R = PoissonLogKappaResidual(geom, bc, f=f)
    m = dev(0.5 * rng.standard_normal(geom.num_elems))
    return geom, R, m


def scipy_csr(M):
    return sp.csr_matrix((host(M.vals), M.pattern.col_ind, M.pattern.row_ptr),
                         shape=(M.rows, M.cols))


def test_bicgstab_solve_and_solve_transpose():
    """The row-replaced Poisson Jacobian (non-symmetric: constrained columns
    kept): solve and solveT against SciPy's sparse LU, 1e-9 relative; repeated
    solves bitwise identical."""
    geom, R, m = poisson(24)
    N = geom.num_nodes
    J = CsrMatrix(R.pattern(), DEV)
    zero = torch.zeros(N, dtype=torch.float64, device=DEV)
    R.assemble_state_jac(zero, m, J)
    b = dev(np.random.default_rng(5).standard_normal(N))
    S = BiCGStabSolver(tol=1e-13)
    x = torch.zeros(N, dtype=torch.float64, device=DEV)
    S.solve(J, b, x)
    Jh = scipy_csr(J)
    ref = spla.spsolve(Jh.tocsc(), host(b))
    assert np.linalg.norm(host(x) - ref) <= 1e-9 * np.linalg.norm(ref)
    x2 = torch.zeros(N, dtype=torch.float64, device=DEV)
    S.solve(J, b, x2)
    assert np.array_equal(host(x), host(x2))
    y = torch.zeros(N, dtype=torch.float64, device=DEV)
    S.solveT(J, b, y)
    reft = spla.spsolve(Jh.T.tocsc(), host(b))
    assert np.linalg.norm(host(y) - reft) <= 1e-9 * np.linalg.norm(reft)
    assert not np.allclose(ref, reft)  # the system is not symmetric


def test_reduced_gradient_loop_on_gpu_matches_finite_differences():
    """femx's ReducedFunctional::gradAt on the P1 log-kappa inverse problem:
    state by Newton with the device solver, adjoint by solveT, parameter VJP
    by fx_p1_vjp; g . dm against central differences of J(m)."""
    geom, R, m = poisson(16, seed=3)
    N, E = geom.num_nodes, geom.num_elems
    rng = np.random.default_rng(9)
    u_obs = dev(0.1 * rng.standard_normal(N))
    S = BiCGStabSolver(tol=1e-14)
    F = ReducedFunctional(NewtonStateSolver(R, S), R, CsrMatrix(R.pattern(), DEV), S,
                          StateMisfit(u_obs, beta=1e-3))
    g = torch.empty(E, dtype=torch.float64, device=DEV)
    val = F.value_grad(m, g)
    dm = rng.standard_normal(E)
    h = 1e-5
    jp = F.value(m + h * dev(dm))
    jm = F.value(m - h * dev(dm))
    fd = (jp - jm) / (2 * h)
    gd = host(g) @ dm
    assert abs(gd - fd) <= 1e-6 * max(1.0, abs(fd)), (gd, fd)
    assert val == F.value(m)
