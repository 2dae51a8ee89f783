"""P1 log-kappa adjoint VJP on the GPU: against the oracle (1e-12 mixed
norm: exp differs by an ulp between libm and the device), central finite
differences of the GPU residual (h = 1e-6, as femx's inverse tests use),
torch.autograd.gradcheck through ResidualVJP, and a reduced gradient in the
form of ReducedFunctional::gradAt (femx/inverse/ReducedFunctional.hpp:94-117)
against finite differences of the reduced functional."""
import numpy as np
import pytest
import torch

from conftest import close
from femx_amd import assembly as A
from femx_amd.inverse import PoissonLogKappaResidual, ResidualVJP
from femx_amd.linalg import CsrMatrix
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
RNG = np.random.default_rng(5)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def problem(n, with_bc=True, f=True):
    geom = A.structured_p1(n, DEV)
    x, y = geom.h_x, geom.h_y
    dofs = A.unit_square_edge_nodes(x, y)
    ff = dev(RNG.standard_normal(geom.num_nodes)) if f else None
    g = dev(RNG.standard_normal(dofs.size)) if with_bc else None
    R = PoissonLogKappaResidual(geom, dofs if with_bc else None, g, ff)
    return geom, R


@pytest.mark.parametrize("n,with_bc", [(4, True), (8, False), (64, True)])
def test_vjp_matches_oracle(n, with_bc):
    geom, R = problem(n, with_bc)
    N, E = geom.num_nodes, geom.num_elems
    m = 0.5 * RNG.standard_normal(E)
    u = RNG.standard_normal(N)
    lam = RNG.standard_normal(N)
    jt, g = R.vjp(dev(u), dev(m), dev(lam))
    mask = R.bmap.bc_mask if with_bc else None
    ojt, og = O.p1_vjp(geom.h_x, geom.h_y, geom.h_conn.reshape(-1), m, u, lam, mask)
    assert close(host(jt), ojt) and close(host(g), og)
    jt2, g2 = R.vjp(dev(u), dev(m), dev(lam))
    assert np.array_equal(host(jt2), host(jt)) and np.array_equal(host(g2), host(g))


@pytest.mark.parametrize("seed", [1, 2])
def test_vjp_scrambled_mesh_matches_oracle(seed):
    """The VJP on a relabelled, shuffled, rotated mesh (each node at every
    local index of its elements), with boundary conditions."""
    from test_gpu_p1 import scrambled_p1
    geom = scrambled_p1(12, seed)
    x, y = geom.h_x, geom.h_y
    dofs = A.unit_square_edge_nodes(x, y)
    R = PoissonLogKappaResidual(geom, dofs, dev(RNG.standard_normal(dofs.size)))
    N, E = geom.num_nodes, geom.num_elems
    m = 0.5 * RNG.standard_normal(E)
    u = RNG.standard_normal(N)
    lam = RNG.standard_normal(N)
    jt, g = R.vjp(dev(u), dev(m), dev(lam))
    ojt, og = O.p1_vjp(x, y, geom.h_conn.reshape(-1), m, u, lam, R.bmap.bc_mask)
    assert close(host(jt), ojt) and close(host(g), og)


def test_vjp_full_size_matches_oracle():
    """Config 4 size: 1024^2 cells, 2.1M elements."""
    n = 1024
    geom, R = problem(n, True, False)
    N, E = geom.num_nodes, geom.num_elems
    rng = np.random.default_rng(6)
    m = 0.5 * rng.standard_normal(E)
    lam = rng.standard_normal(N)
    u = rng.standard_normal(N)
    jt, g = R.vjp(dev(u), dev(m), dev(lam))
    ojt, og = O.p1_vjp(geom.h_x, geom.h_y, geom.h_conn.reshape(-1), m, u, lam, R.bmap.bc_mask)
    assert close(host(jt), ojt) and close(host(g), og)


def test_vjp_finite_differences():
    geom, R = problem(6)
    N, E = geom.num_nodes, geom.num_elems
    m = dev(0.5 * RNG.standard_normal(E))
    u = dev(RNG.standard_normal(N))
    lam = dev(RNG.standard_normal(N))
    jt, g = (host(t) for t in R.vjp(u, m, lam))
    out = torch.empty(N, dtype=torch.float64, device=DEV)

    def lr(uu, mm):
        R.res(uu, mm, out)
        return float(host(torch.dot(lam, out)))

    h = 1e-6
    for k in RNG.choice(E, 8, replace=False):
        e = torch.zeros(E, dtype=torch.float64, device=DEV)
        e[k] = h
        fd = (lr(u, m + e) - lr(u, m - e)) / (2 * h)
        assert abs(g[k] - fd) <= 1e-6 * max(1.0, abs(fd))
    for k in list(RNG.choice(N, 8, replace=False)) + [0]:
        e = torch.zeros(N, dtype=torch.float64, device=DEV)
        e[k] = h
        fd = (lr(u + e, m) - lr(u - e, m)) / (2 * h)
        assert abs(jt[k] - fd) <= 1e-6 * max(1.0, abs(fd))


def test_autograd_gradcheck():
    geom, R = problem(3)
    u = dev(RNG.standard_normal(geom.num_nodes)).requires_grad_()
    m = dev(0.3 * RNG.standard_normal(geom.num_elems)).requires_grad_()
    assert torch.autograd.gradcheck(lambda uu, mm: ResidualVJP.apply(uu, mm, R), (u, m),
                                    eps=1e-6, atol=1e-7, rtol=1e-6)


def test_reduced_gradient_matches_finite_differences():
    """J(m) = 1/2 |u(m) - d|^2 with R(u(m), m) = 0: grad = dJ/dm - (dR/dm)^T lam,
    lam solving J^T lam = dJ/du (ReducedFunctional::gradAt)."""
    geom, R = problem(6, True, True)
    N, E = geom.num_nodes, geom.num_elems
    d = RNG.standard_normal(N)
    mat = CsrMatrix(R.map.pattern, DEV)
    zero = torch.zeros(N, dtype=torch.float64, device=DEV)
    r0 = torch.empty(N, dtype=torch.float64, device=DEV)

    def solve_state(m):
        R.assemble_state_jac(zero, m, mat)
        R.res(zero, m, r0)  # R(0) = -F on free rows, -g on constrained rows
        return np.linalg.solve(mat.to_dense(), -host(r0)), mat.to_dense()

    def J(m):
        u, _ = solve_state(m)
        return 0.5 * np.sum((u - d) ** 2)

    m = dev(0.4 * RNG.standard_normal(E))
    u, Jm = solve_state(m)
    lam = np.linalg.solve(Jm.T, u - d)
    g = host(R.apply_param_jac_t(dev(u), m, dev(lam)))
    grad = 0.0 - g
    h = 1e-6
    for k in RNG.choice(E, 6, replace=False):
        e = torch.zeros(E, dtype=torch.float64, device=DEV)
        e[k] = h
        fd = (J(m + e) - J(m - e)) / (2 * h)
        assert abs(grad[k] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, grad[k], fd)
