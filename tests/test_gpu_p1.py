"""HIP path vs the oracle and the compiled reference's fixtures: P1 Poisson
assembly, Dirichlet operations, SpMV and vector ops.  Integer outputs are
bit-exact; the FP64 outputs of these kernels are bitwise equal to the
oracle's (same operation order, no contraction), which is stricter than the
1e-12 mixed-norm bar of conftest.close."""
import numpy as np
import pytest
import torch

from conftest import close, golden
from femx_amd import assembly as A
from femx_amd.linalg import CsrMatrix, CsrPattern, MatrixHandler, VectorHandler
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
RNG = np.random.default_rng(3)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def p1_setup(n):
    geom = A.structured_p1(n, DEV)
    amap = A.make_assembly_map(geom.num_nodes, geom.num_nodes, geom.h_conn)
    return geom, amap


@pytest.mark.parametrize("n", [1, 2, 8, 33])
def test_p1_assemble_bitwise_vs_oracle(n):
    geom, amap = p1_setup(n)
    N, E = geom.num_nodes, geom.num_elems
    kappa = np.exp(0.5 * RNG.standard_normal(E))
    f = RNG.standard_normal(N)
    u = RNG.standard_normal(N)
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(N, dtype=torch.float64, device=DEV)
    A.assemble(A.PoissonP1(dev(kappa), dev(f)), geom, amap, dev(u), res, jac)
    _, jm = amap.jac()
    ov, orr = O.p1_assemble(geom.h_x, geom.h_y, geom.h_conn.reshape(-1), kappa, f, u, jm,
                            amap.pattern.nnz)
    assert np.array_equal(host(jac.vals), ov)
    assert np.array_equal(host(res), orr)
    # residual-only pass gives the same residual
    res2 = torch.full((N,), np.nan, dtype=torch.float64, device=DEV)
    A.assemble(A.PoissonP1(dev(kappa), dev(f)), geom, amap, dev(u), res2, None)
    assert np.array_equal(host(res2), orr)
    # log-kappa input
    res3 = torch.empty(N, dtype=torch.float64, device=DEV)
    jac3 = CsrMatrix(amap.pattern, DEV)
    A.assemble(A.PoissonP1(f=dev(f), log_kappa=dev(np.log(kappa))), geom, amap, dev(u), res3, jac3)
    assert close(host(jac3.vals), ov) and close(host(res3), orr)


def scrambled_p1(n, seed):
    """The structured P1 mesh with its nodes relabelled, its elements shuffled
    and each element's vertex list rotated (orientation kept): every local
    index meets every kind of neighbourhood, in no particular order."""
    rng = np.random.default_rng(seed)
    x, y, conn = O.mesh_p1(n)
    conn = conn.reshape(-1, 3)
    perm = rng.permutation(x.size)  # new label of each old node
    inv = np.empty_like(perm)
    inv[perm] = np.arange(x.size)
    conn = perm[conn][rng.permutation(conn.shape[0])]
    shift = rng.integers(0, 3, conn.shape[0])
    conn = np.stack([np.roll(c, s) for c, s in zip(conn, shift)]).astype(np.int32)
    return A.Geometry(x[inv], y[inv], conn, DEV)


@pytest.mark.parametrize("n,seed", [(5, 1), (16, 2)])
def test_p1_scrambled_mesh_bitwise_vs_oracle(n, seed):
    """Assembly on a relabelled, shuffled, rotated mesh: the lane's own node
    is at every local index of its elements, in every position of its row."""
    geom = scrambled_p1(n, seed)
    amap = A.make_assembly_map(geom.num_nodes, geom.num_nodes, geom.h_conn)
    N, E = geom.num_nodes, geom.num_elems
    kappa = np.exp(0.5 * RNG.standard_normal(E))
    f = RNG.standard_normal(N)
    u = RNG.standard_normal(N)
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(N, dtype=torch.float64, device=DEV)
    A.assemble(A.PoissonP1(dev(kappa), dev(f)), geom, amap, dev(u), res, jac)
    _, jm = amap.jac()
    ov, orr = O.p1_assemble(geom.h_x, geom.h_y, geom.h_conn.reshape(-1), kappa, f, u, jm,
                            amap.pattern.nnz)
    assert np.array_equal(host(jac.vals), ov)
    assert np.array_equal(host(res), orr)


@pytest.mark.parametrize("n", [2, 8])
def test_p1_laplace_bitwise_vs_reference_fixture(n):
    """kappa = 1, f = 0: femx's own PoissonComponents assembly."""
    g = golden(f"poisson_p1_n{n}.npz")
    geom = A.Geometry(g["x"], g["y"], g["conn"].reshape(-1, 3), DEV)
    amap = A.make_assembly_map(geom.num_nodes, geom.num_nodes, geom.h_conn)
    assert np.array_equal(amap.pattern.row_ptr, g["row_ptr"])
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(geom.num_nodes, dtype=torch.float64, device=DEV)
    A.assemble(A.PoissonP1(), geom, amap, dev(g["state"]), res, jac)
    assert np.array_equal(host(jac.vals), g["vals"])
    assert np.array_equal(host(res), g["res"])


def test_p1_full_size_step_bitwise_and_deterministic():
    """The whole 2048^2 step (8.4M elements) with the benchmark's inputs --
    kappa_e = exp(0.5 z_e) (seed 2), f ~ N(0,1) (seed 3), u ~ N(0,1) (seed 4),
    x_k ~ U(-1,1) (seeds 10..19): assembly, rhs = -res, prepareForwardSolve
    with g = 0 on all 8 192 edge DOFs, and the 10-vector product.  Every
    output is bitwise equal to the oracle (femx's host order), and a second
    assembly launch is bitwise equal to the first."""
    n = 2048
    geom, amap = p1_setup(n)
    N, E = geom.num_nodes, geom.num_elems
    assert amap.pattern.nnz == 7 * n * n + 6 * n + 1
    kappa = np.exp(0.5 * np.random.default_rng(2).standard_normal(E))
    f = np.random.default_rng(3).standard_normal(N)
    u = np.random.default_rng(4).standard_normal(N)
    xs = np.stack([np.random.default_rng(10 + k).uniform(-1.0, 1.0, N) for k in range(10)])
    op = A.PoissonP1(dev(kappa), dev(f))
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(N, dtype=torch.float64, device=DEV)
    A.assemble(op, geom, amap, dev(u), res, jac)
    v1, r1 = host(jac.vals), host(res)
    A.assemble(op, geom, amap, dev(u), res, jac)
    assert np.array_equal(host(jac.vals), v1) and np.array_equal(host(res), r1)
    _, jm = amap.jac()
    # the serial oracle: femx's host assembly order (the OpenMP one scatters
    # with atomics, in no fixed order)
    ov, orr = O.p1_assemble(geom.h_x, geom.h_y, geom.h_conn.reshape(-1), kappa, f, u, jm,
                            amap.pattern.nnz)
    assert np.array_equal(v1, ov) and np.array_equal(r1, orr)

    x, y = geom.h_x, geom.h_y
    dofs = A.unit_square_edge_nodes(x, y)
    assert dofs.size == 4 * n
    bmap = A.make_boundary_map(dofs, amap.pattern, DEV)
    rhs = torch.empty(N, dtype=torch.float64, device=DEV)
    VectorHandler().axpby(-1.0, res, 0.0, rhs)
    A.prepare_forward_solve(bmap, jac, rhs, torch.zeros(dofs.size, dtype=torch.float64,
                                                        device=DEV))
    Y = torch.empty((10, N), dtype=torch.float64, device=DEV)
    MatrixHandler().matvec_multi(jac, dev(xs), Y)
    bc = dict(bc_rows=bmap.bc_rows, diag=bmap.diag, col_offsets=bmap.col_offsets,
              col_entries=bmap.col_entries, col_rows=bmap.col_rows, bc_mask=bmap.bc_mask)
    orhs = -orr
    O.prepare_forward(amap.pattern.row_ptr, bc, ov, orhs, np.zeros(dofs.size))
    assert np.array_equal(host(jac.vals), ov), "eliminated matrix"
    assert np.array_equal(host(rhs), orhs), "eliminated right-hand side"
    Yh = host(Y)
    oy = np.empty(N)
    for k in range(10):
        O.matvec(amap.pattern.row_ptr, amap.pattern.col_ind, ov, xs[k], oy)
        assert np.array_equal(Yh[k], oy), f"product {k}"


def test_bc_ops_bitwise_vs_reference_fixture():
    g = golden("bcops_p1_n4.npz")
    N = g["row_ptr"].size - 1
    pat = CsrPattern(N, N, g["row_ptr"], g["col_ind"])
    b = A.make_boundary_map(g["dofs"], pat, DEV)
    for name, diag in (("rows1", 1.0), ("rows0", 0.0)):
        m = CsrMatrix(pat, DEV, dev(g["vals0"]))
        A.replace_rows(b, m, diag)
        assert np.array_equal(host(m.vals), g[name + "_vals"])
    m = CsrMatrix(pat, DEV, dev(g["vals0"]))
    rhs = dev(g["rhs0"])
    A.prepare_forward_solve(b, m, rhs, dev(g["bc_vals"]))
    assert np.array_equal(host(m.vals), g["forward_vals"])
    assert np.array_equal(host(rhs), g["forward_rhs"])
    r = dev(g["rhs0"])
    A.replace_res(b, dev(g["state"]), dev(g["bc_vals"]), r)
    assert np.array_equal(host(r), g["res_rhs"])
    r = dev(g["rhs0"])
    A.zero_boundary(b, r)
    assert np.array_equal(host(r), g["zero_rhs"])


def test_bc_kat_and_layout_check():
    # tests/unit/backend/runBackendStateTests.cpp:80-114
    pat = CsrPattern(3, 3, [0, 3, 6, 9], [0, 1, 2] * 3)
    b = A.make_boundary_map([1], pat, DEV)
    m = CsrMatrix(pat, DEV, dev(np.array([4.0, 1.0, 2.0, 3.0, 5.0, 6.0, 7.0, 8.0, 9.0])))
    rhs = dev(np.array([10.0, 20.0, 30.0]))
    A.prepare_forward_solve(b, m, rhs, dev(np.array([2.0])))
    assert host(m.vals).tolist() == [4.0, 0.0, 2.0, 0.0, 1.0, 0.0, 7.0, 0.0, 9.0]
    assert host(rhs).tolist() == [8.0, 2.0, 14.0]
    other = CsrMatrix(CsrPattern(3, 3, [0, 3, 6, 9], [0, 1, 2] * 3), DEV)
    with pytest.raises(RuntimeError, match="layout"):
        A.replace_rows(b, other, 1.0)


def test_spmv_bitwise_vs_reference_fixture():
    g = golden("spmv_p1_n4.npz")
    N = g["row_ptr"].size - 1
    pat = CsrPattern(N, N, g["row_ptr"], g["col_ind"])
    m = CsrMatrix(pat, DEV, dev(g["vals"]))
    mh = MatrixHandler()
    for tr in (0, 1):
        for alpha, beta in ((1.0, 0.0), (1.5, -0.5)):
            y = dev(g["y0"])
            (mh.matvecT if tr else mh.matvec)(m, dev(g["x"]), y, alpha, beta)
            assert np.array_equal(host(y), g[f"y_t{tr}_a{alpha}_b{beta}"])


def test_linalg_kats():
    # tests/unit/backend/runCudaLinalgTests.cu:70-181
    pat = CsrPattern(3, 4, [0, 2, 4, 7], [0, 2, 1, 3, 0, 2, 3])
    m = CsrMatrix(pat, DEV, dev(np.array([2.0, -1.0, 3.0, 4.0, -2.0, 5.0, 1.0])))
    mh, vh = MatrixHandler(), VectorHandler()
    x = dev(np.array([1.0, 2.0, 3.0, 4.0]))
    y = torch.empty(3, dtype=torch.float64, device=DEV)
    mh.matvec(m, x, y)
    assert host(y).tolist() == [-1.0, 22.0, 17.0]
    assert host(vh.dot(x, x)).tolist() == [30.0]
    vh.axpby(-2.0, dev(np.array([1.0, 2.0, 3.0])), 0.5, y)
    assert host(y).tolist() == [-2.5, 7.0, 2.5]
    t = torch.empty(4, dtype=torch.float64, device=DEV)
    mh.matvecT(m, dev(np.array([2.0, -1.0, 0.5])), t)
    ref = np.zeros(4)
    O.matvecT(pat.row_ptr, pat.col_ind, np.array([2.0, -1.0, 3.0, 4.0, -2.0, 5.0, 1.0]), 4,
              np.array([2.0, -1.0, 0.5]), ref)
    assert np.array_equal(host(t), ref)
    buf = torch.zeros(9, dtype=torch.float64, device=DEV)
    with pytest.raises(RuntimeError, match="in-place"):
        from femx_amd import _lib
        from femx_amd.linalg import ptr, stream_handle
        rp, ci = pat.device_arrays(DEV)
        _lib.call("fx_csr_matvec", 3, 4, ptr(rp), ptr(ci), ptr(m.vals), buf.data_ptr() + 24,
                  buf.data_ptr() + 32, 1.0, 0.0, stream_handle())


@pytest.mark.parametrize("bandwidth", [50, 80000])
def test_spmv_16bit_offsets_match_int32_path(bandwidth):
    """MatrixHandler.matvec takes the 16-bit column-offset path when the
    pattern allows it; both paths give bitwise equal results (random
    rectangular pattern with empty rows, negative and positive offsets)."""
    from femx_amd import _lib
    from femx_amd.linalg import ptr, stream_handle
    rng = np.random.default_rng(11)
    rows, cols = 3000, 3000 + bandwidth
    counts = rng.integers(0, 9, rows)
    counts[::17] = 0
    rp = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(np.arange(max(0, r - bandwidth // 2),
                                                      min(cols, r + bandwidth // 2)),
                                            c, replace=False)) for r, c in enumerate(counts)])
    pat = CsrPattern(rows, cols, rp, ci.astype(np.int32))
    assert (pat.col_off16(DEV) is None) == (bandwidth > 32767 * 2)
    m = CsrMatrix(pat, DEV, dev(rng.standard_normal(pat.nnz)))
    x, y0 = dev(rng.standard_normal(cols)), dev(rng.standard_normal(rows))
    for alpha, beta in ((1.0, 0.0), (-0.7, 1.3)):
        y1, y2 = y0.clone(), y0.clone()
        MatrixHandler().matvec(m, x, y1, alpha, beta)
        d_rp, d_ci = pat.device_arrays(DEV)
        _lib.call("fx_csr_matvec", rows, cols, ptr(d_rp), ptr(d_ci), ptr(m.vals), ptr(x), ptr(y2),
                  alpha, beta, stream_handle())
        assert torch.equal(y1, y2)
        ref = np.zeros(rows)
        O.matvec(rp, pat.col_ind, host(m.vals), host(x), ref)
        assert np.array_equal(host(y1), alpha * ref + beta * host(y0) if beta else alpha * ref)


@pytest.mark.parametrize("bandwidth,nv", [(50, 10), (80000, 3), (50, 17)])
def test_multi_vector_spmv_equals_single_products(bandwidth, nv):
    """matvec_multi: nv products with one matrix, each bitwise equal to its
    single-vector matvec (16-bit and int32 column paths, two 16-vector chunks
    at nv = 17, alpha/beta), rows with more entries than the LDS stage."""
    rng = np.random.default_rng(12 + nv)
    rows, cols = 2500, 2500 + bandwidth
    counts = rng.integers(0, 9, rows)
    counts[::13] = 0
    counts[100:110] = 400  # one block over the 2048-entry stage: the direct-read path
    rp = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(np.arange(max(0, r - max(bandwidth, 800) // 2),
                                                      min(cols, r + max(bandwidth, 800) // 2)),
                                            c, replace=False)) for r, c in enumerate(counts)])
    pat = CsrPattern(rows, cols, rp, ci.astype(np.int32))
    m = CsrMatrix(pat, DEV, dev(rng.standard_normal(pat.nnz)))
    X = dev(rng.standard_normal((nv, cols)))
    Y0 = dev(rng.standard_normal((nv, rows)))
    mh = MatrixHandler()
    for alpha, beta in ((1.0, 0.0), (-0.7, 1.3)):
        Y = Y0.clone()
        mh.matvec_multi(m, X, Y, alpha, beta)
        Ys = Y0.clone()  # the stencil kernel: no stencil block here, every row gathers
        mh.matvec_multi(m, X, Ys, alpha, beta, stencil=True)
        for j in range(nv):
            yj = Y0[j].clone()
            mh.matvec(m, X[j], yj, alpha, beta)
            assert torch.equal(Y[j], yj), (j, alpha)
            assert torch.equal(Ys[j], yj), ("stencil kernel", j, alpha)


@pytest.mark.parametrize("n,nv", [(300, 10), (300, 3), (129, 17), (512, 1)])
def test_multi_vector_spmv_stencil_blocks_equal_single_products(n, nv):
    """The stencil-block kernel on structured P1 patterns (interior 7-point and
    edge-row 5-point stencil blocks beside gathered blocks, a partial last
    block; 16 + 1 vectors at nv = 17): every product bitwise equal to its
    single-vector matvec, with alpha/beta, and to the gathering kernel."""
    geom, amap = p1_setup(n)
    pat = amap.pattern
    code, _, frac, hcode = pat.stencil_plan(DEV)
    assert frac > 0.3 and np.any(hcode >= 0) and np.any(hcode < 0)
    rng = np.random.default_rng(40 + n + nv)
    m = CsrMatrix(pat, DEV, dev(rng.standard_normal(pat.nnz)))
    X = dev(rng.uniform(-1.0, 1.0, (nv, pat.cols)))
    Y0 = dev(rng.standard_normal((nv, pat.rows)))
    mh = MatrixHandler()
    for alpha, beta in ((1.0, 0.0), (0.3, -2.0)):
        Ys, Yg = Y0.clone(), Y0.clone()
        mh.matvec_multi(m, X, Ys, alpha, beta, stencil=True)
        mh.matvec_multi(m, X, Yg, alpha, beta, stencil=False)
        assert torch.equal(Ys, Yg), alpha
        for j in range(nv):
            yj = Y0[j].clone()
            mh.matvec(m, X[j], yj, alpha, beta)
            assert torch.equal(Ys[j], yj), (j, alpha)


def test_dot_is_deterministic_and_accurate():
    vh = VectorHandler()
    x = dev(RNG.standard_normal(3_000_001))
    y = dev(RNG.standard_normal(3_000_001))
    a = host(vh.dot(x, y))[0]
    b = host(vh.dot(x, y))[0]
    assert a == b
    ref = float(np.dot(host(x), host(y)))
    assert abs(a - ref) <= 1e-12 * np.sqrt(3e6) * 2


def test_config1_pipeline_bitwise():
    """Config 1: 32x32 P1, kappa = 1, f = 2 pi^2 sin sin, homogeneous
    Dirichlet on all edges; assemble, apply BCs, one SpMV."""
    n = 32
    geom, amap = p1_setup(n)
    N = geom.num_nodes
    x, y = geom.h_x, geom.h_y
    f = 2 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y)
    dofs = A.unit_square_edge_nodes(x, y)
    assert dofs.size == 128
    b = A.make_boundary_map(dofs, amap.pattern, DEV)
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(N, dtype=torch.float64, device=DEV)
    A.assemble(A.PoissonP1(None, dev(f)), geom, amap, None, res, jac)
    rhs = -res
    A.prepare_forward_solve(b, jac, rhs, torch.zeros(dofs.size, dtype=torch.float64, device=DEV))
    xs = dev(np.sin(np.pi * x) * np.sin(np.pi * y))
    out = torch.empty(N, dtype=torch.float64, device=DEV)
    MatrixHandler().matvec(jac, xs, out)
    _, jm = amap.jac()
    ov, orr = O.p1_assemble(x, y, geom.h_conn.reshape(-1), np.ones(geom.num_elems), f,
                            np.zeros(N), jm, amap.pattern.nnz)
    orhs = -orr
    bc = O.bc_build(amap.pattern.row_ptr, amap.pattern.col_ind, dofs)
    O.prepare_forward(amap.pattern.row_ptr, bc, ov, orhs, np.zeros(dofs.size))
    oy = np.zeros(N)
    O.matvec(amap.pattern.row_ptr, amap.pattern.col_ind, ov, np.sin(np.pi * x) * np.sin(np.pi * y),
             oy)
    assert np.array_equal(host(jac.vals), ov)
    assert np.array_equal(host(rhs), orhs)
    assert np.array_equal(host(out), oy)
    # A u_exact approximates the right-hand side (consistency, O(h^2))
    assert np.abs(host(out) - orhs).max() < 0.05 * np.abs(orhs).max()


def test_empty_inputs():
    pat = CsrPattern(0, 0, [0], [])
    m = CsrMatrix(pat, DEV)
    y = torch.empty(0, dtype=torch.float64, device=DEV)
    MatrixHandler().matvec(m, torch.empty(0, dtype=torch.float64, device=DEV), y)
    b = A.make_boundary_map([], CsrPattern(2, 2, [0, 1, 2], [0, 1]), DEV)
    A.zero_boundary(b, torch.ones(2, dtype=torch.float64, device=DEV))
    torch.cuda.synchronize()
