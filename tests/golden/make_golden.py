"""Write the golden fixtures of tests/golden/ from the compiled femx reference.

Runs in the development container only: it needs oracle/_ref/libfemx_ref.so,
which `make -C oracle ref` compiles from femx's own sources.  Each fixture
holds the inputs (structured triangle meshes built by this repository's
conventions, seeded values) and the reference's outputs:

  map_p1_n{1,2,4,8}.npz   makeAssemblyMap on P1 DOF tables
  map_th_n{1,2}.npz       makeAssemblyMap on Taylor-Hood DOF tables
  bc_p1_n4.npz            makeBoundaryMap on the n=4 boundary DOFs
  bcops_p1_n4.npz         replaceRows(1 and 0), replaceRes, zeroBoundary,
                          prepareForwardSolve on seeded values
  poisson_p1_n{2,8}.npz   assemble(PoissonComponents) K values and K u
  spmv_p1_n4.npz          matvec / matvecT with alpha, beta
  gls_tri1.npz, gls_p1_n4_perturbed.npz, gls_p1_n6.npz
                          NavierOperator (GLS P1/P1) time assemble: CSR
                          values and residual per (step, num_hist, state)
  transpose.npz           MatrixHandler::transpose on a rectangular unsorted
                          matrix and the P1 n=3 pattern
  gls_constrained_n4.npz  NavierStokesModel residual under ConstrainedTimeResidual:
                          res, assembleNext, applyJacT(param), and
                          TimeIntegrator::solveStep's rhs/prepareLinearSolve
  gmsh_*.npz              femx's own Gmsh meshes (data/meshes) through its
                          GmshReader: makeAssemblyMap, PoissonComponents,
                          makeBoundaryMap on physical groups +
                          prepareForwardSolve, NavierOperator time assembly

    python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

I32 = np.int32


def p(a):
    return O._p(a)


def ref_map(R, num_res, num_states, off, dofs):
    E = off.size - 1
    h = C.c_void_p()
    assert R.ref_map_build(num_res, num_states, E, p(off), p(dofs), p(off), p(dofs), C.byref(h)) == 0
    nnz, jl, rows = C.c_int32(), C.c_int32(), C.c_int32()
    R.ref_map_sizes(h, C.byref(nnz), C.byref(jl), C.byref(rows))
    rp = np.empty(num_res + 1, I32)
    ci = np.empty(nnz.value, I32)
    jo = np.empty(E + 1, I32)
    jm = np.empty(jl.value, I32)
    R.ref_map_export(h, p(rp), p(ci), p(jo), p(jm))
    R.ref_map_free(h)
    return rp, ci, jo, jm


def boundary_dofs_p1(n):
    """Marker-based DirichletBC order: ascending node id (femx/fem/
    DirichletBC.cpp:138-145) over nodes on the unit-square boundary."""
    x, y, _ = O.mesh_p1(n)
    tol = 1e-12  # x_min + i*dx can miss 1.0 by an ulp (femx/fem/Mesh.cpp:19-29)
    on = (abs(x) <= tol) | (abs(x - 1.0) <= tol) | (abs(y) <= tol) | (abs(y - 1.0) <= tol)
    return np.nonzero(on)[0].astype(I32)


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libfemx_ref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(20240611)

    for n in (1, 2, 4, 8):
        x, y, conn = O.mesh_p1(n)
        N, E = x.size, conn.size // 3
        off = O.uniform_offsets(E, 3)
        rp, ci, jo, jm = ref_map(R, N, N, off, conn)
        np.savez(os.path.join(HERE, f"map_p1_n{n}.npz"), n=n, conn=conn, row_ptr=rp, col_ind=ci,
                 jac_offsets=jo, jac_map=jm)

    for n in (1, 2):
        xv, yv, c6, v3 = O.mesh_p2(n)
        d = O.th_dofs(n, c6, v3)
        E = 2 * n * n
        ndof = 2 * (2 * n + 1) ** 2 + (n + 1) ** 2
        off = O.uniform_offsets(E, 15)
        rp, ci, jo, jm = ref_map(R, ndof, ndof, off, d)
        np.savez(os.path.join(HERE, f"map_th_n{n}.npz"), n=n, dofs15=d, row_ptr=rp, col_ind=ci,
                 jac_offsets=jo, jac_map=jm)

    # Boundary map and boundary operations on the P1 n=4 pattern.
    n = 4
    x, y, conn = O.mesh_p1(n)
    N, E = x.size, conn.size // 3
    off = O.uniform_offsets(E, 3)
    rp, ci, _, _ = ref_map(R, N, N, off, conn)
    dofs = boundary_dofs_p1(n)
    h = C.c_void_p()
    assert R.ref_bc_build(N, p(rp), p(ci), dofs.size, p(dofs), C.byref(h)) == 0
    nce = R.ref_bc_num_col_entries(h)
    arrs = dict(bc_rows=np.empty(dofs.size, I32), diag=np.empty(dofs.size, I32),
                col_offsets=np.empty(dofs.size + 1, I32), col_entries=np.empty(nce, I32),
                col_rows=np.empty(nce, I32), bc_mask=np.empty(N, I32))
    R.ref_bc_export(h, *(p(arrs[k]) for k in ("bc_rows", "diag", "col_offsets", "col_entries",
                                              "col_rows", "bc_mask")))
    R.ref_bc_free(h)
    np.savez(os.path.join(HERE, "bc_p1_n4.npz"), row_ptr=rp, col_ind=ci, dofs=dofs, **arrs)

    nnz = ci.size
    vals0 = rng.standard_normal(nnz)
    rhs0 = rng.standard_normal(N)
    g = rng.standard_normal(dofs.size)
    state = rng.standard_normal(N)
    out = {}
    for name, op, diag in (("rows1", 0, 1.0), ("rows0", 0, 0.0), ("forward", 1, 1.0),
                           ("res", 2, 0.0), ("zero", 3, 0.0)):
        v = vals0.copy()
        r = rhs0.copy()
        assert R.ref_bc_apply(op, N, p(rp), p(ci), dofs.size, p(dofs), p(v), p(r), p(g), p(state),
                              diag) == 0, R.ref_last_error()
        out[name + "_vals"], out[name + "_rhs"] = v, r
    np.savez(os.path.join(HERE, "bcops_p1_n4.npz"), row_ptr=rp, col_ind=ci, dofs=dofs,
             vals0=vals0, rhs0=rhs0, bc_vals=g, state=state, **out)

    # Poisson (Laplace) assembly through the reference's own pipeline.
    for n in (2, 8):
        x, y, conn = O.mesh_p1(n)
        N, E = x.size, conn.size // 3
        u = rng.standard_normal(N)
        cap = 7 * n * n + 6 * n + 1
        rp = np.empty(N + 1, I32)
        ci = np.empty(cap, I32)
        vals = np.empty(cap)
        res = np.empty(N)
        assert R.ref_poisson_p1(N, p(x), p(y), E, p(conn), p(u), cap, p(rp), p(ci), p(vals),
                                p(res)) == 0, R.ref_last_error()
        np.savez(os.path.join(HERE, f"poisson_p1_n{n}.npz"), n=n, x=x, y=y, conn=conn, state=u,
                 row_ptr=rp, col_ind=ci, vals=vals, res=res)

    # SpMV / transpose apply on the P1 n=4 pattern with seeded values.
    n = 4
    x, y, conn = O.mesh_p1(n)
    N, E = x.size, conn.size // 3
    off = O.uniform_offsets(E, 3)
    rp, ci, _, _ = ref_map(R, N, N, off, conn)
    vals = rng.standard_normal(ci.size)
    xin = rng.uniform(-1, 1, N)
    y0 = rng.standard_normal(N)
    res = {}
    for tr in (0, 1):
        for alpha, beta in ((1.0, 0.0), (1.5, -0.5)):
            yy = y0.copy()
            assert R.ref_spmv(tr, N, N, p(rp), p(ci), p(vals), p(xin), p(yy), alpha, beta) == 0
            res[f"y_t{tr}_a{alpha}_b{beta}"] = yy
    np.savez(os.path.join(HERE, "spmv_p1_n4.npz"), row_ptr=rp, col_ind=ci, vals=vals, x=xin,
             y0=y0, **res)
    # GLS P1/P1 Navier-Stokes time assembly (femx's NavierOperator through
    # the host time assemble).  The single triangle and the state pattern are
    # those of tests/unit/model/ns/runNavierOperatorTests.cpp:142-178.
    def gls_case(x, y, conn, rho, mu, dt, cases, name):
        out = dict(x=x, y=y, conn=conn, rho=rho, mu=mu, dt=dt, num_cases=len(cases))
        for k, (step, nh, hist, nxt) in enumerate(cases):
            d9, rp, ci, vals, res = O.ref_ns_gls(x, y, conn, rho, mu, dt, step, nh, hist, nxt)
            out.update({f"step{k}": step, f"num_hist{k}": nh, f"hist{k}": hist, f"nxt{k}": nxt,
                        f"vals{k}": vals, f"res{k}": res})
        out.update(dofs9=d9, row_ptr=rp, col_ind=ci)
        np.savez_compressed(os.path.join(HERE, name), **out)

    x = np.array([0.0, 1.0, 0.0])
    y = np.array([0.0, 0.0, 1.0])
    conn = np.array([0, 1, 2], I32)
    ns_ = 9
    i = np.arange(ns_)
    hist = np.concatenate([0.01 * (i + 1), -0.005 * (i + 1)])
    nxt = 0.02 - 0.003 * i
    gls_case(x, y, conn, 1.2, 0.03, 0.05,
             [(1, 2, hist, nxt), (0, 1, hist[:ns_], nxt), (1, 2, hist, np.zeros(ns_))],
             "gls_tri1.npz")

    n = 4
    x, y, conn = O.mesh_p1(n)
    inner = (x > 0) & (x < 1) & (y > 0) & (y < 1)
    x[inner] += 0.2 / n * rng.uniform(-1, 1, inner.sum())
    y[inner] += 0.2 / n * rng.uniform(-1, 1, inner.sum())
    nd = 3 * x.size
    still = rng.standard_normal(2 * nd)
    still[:2 * x.size] = 0.0  # zero velocity in both lags: elemLength's fallback direction
    gls_case(x, y, conn, 1.2, 0.03, 0.05,
             [(0, 1, rng.standard_normal(nd), rng.standard_normal(nd)),
              (1, 2, rng.standard_normal(2 * nd), rng.standard_normal(nd)),
              (2, 3, rng.standard_normal(3 * nd), rng.standard_normal(nd)),
              (1, 2, still, rng.standard_normal(nd))],
             "gls_p1_n4_perturbed.npz")

    n = 6
    x, y, conn = O.mesh_p1(n)
    nd = 3 * x.size
    gls_case(x, y, conn, 1.0, 1.0e-3, 0.01,
             [(1, 2, rng.standard_normal(2 * nd), rng.standard_normal(nd))], "gls_p1_n6.npz")

    # femx's NavierStokesModel residual under ConstrainedTimeResidual: an
    # inflow control with two parameters per level (profile weights y(1-y) and
    # y, each row's entries given column 1 first), walls and a pressure gauge
    # fixed with per-step values, four steps interpolating three control
    # levels, and one leading parameter outside the control (ctr_off = 1).
    n = 4
    x, y, conn = O.mesh_p1(n)
    inner = (x > 0) & (x < 1) & (y > 0) & (y < 1)
    x[inner] += 0.15 / n * rng.uniform(-1, 1, inner.sum())
    y[inner] += 0.15 / n * rng.uniform(-1, 1, inner.sum())
    N = x.size
    left = [i for i in range(N) if x[i] == 0.0 and 0.0 < y[i] < 1.0]
    ctr_dofs = np.array([2 * i for i in left], I32)
    ents = []
    for r, i in enumerate(left):
        ents += [(r, 1, y[i]), (r, 0, y[i] * (1.0 - y[i]))]
    ent = np.array(ents, dtype=np.float64)
    wall = [i for i in range(N) if (x[i] in (0.0, 1.0) or y[i] in (0.0, 1.0)) and i not in left]
    fixed = np.array(sorted([2 * i + c for i in wall for c in (0, 1)] + [2 * i + 1 for i in left]
                            + [2 * N]), I32)
    nstep = 4
    fvals = rng.standard_normal(nstep * fixed.size)
    time = np.array([(s // 2, s // 2 + 1, 0.5 * (s % 2)) for s in range(nstep)], np.float64)
    ns_ = 3 * N
    out = dict(x=x, y=y, conn=conn, nstep=nstep, dt=0.05, rho=1.2, mu=0.03, ctr_dofs=ctr_dofs,
               num_ctr_params=2, entries=ent, fixed_dofs=fixed, fixed_vals=fvals, time=time,
               ctr_off=1, num_prm=8, num_cases=2)
    for k, step in enumerate((1, 2)):
        hist, nxt = rng.standard_normal(2 * ns_), rng.standard_normal(ns_)
        prm, adj = rng.standard_normal(8), rng.standard_normal(ns_)
        r = O.ref_ns_constrained(x, y, conn, nstep, 0.05, 1.2, 0.03, ctr_dofs, 2,
                                 [(int(a), int(b), w) for a, b, w in ents], fixed, fvals,
                                 [tuple(t) for t in time], 1, 8, step, hist, nxt, prm, adj)
        out.update({f"step{k}": step, f"hist{k}": hist, f"nxt{k}": nxt, f"prm{k}": prm,
                    f"adj{k}": adj})
        for key in ("res", "asm_res", "asm_vals", "grad", "prep_vals", "prep_rhs"):
            out[f"{key}{k}"] = r[key]
    np.savez_compressed(os.path.join(HERE, "gls_constrained_n4.npz"), **out)

    # Explicit CSR transpose (MatrixHandler::transpose): a rectangular matrix
    # with unsorted columns and an empty row, and the P1 n=3 pattern.
    rp = np.array([0, 3, 3, 5, 9], I32)
    ci = np.array([4, 0, 2, 1, 4, 3, 0, 4, 1], I32)
    cases = {"rect": (4, 5, rp, ci)}
    x, y, conn = O.mesh_p1(3)
    N = x.size
    off = O.uniform_offsets(conn.size // 3, 3)
    prp, pci, _, _ = ref_map(R, N, N, off, conn)
    cases["p1_n3"] = (N, N, prp, pci)
    tout = {}
    for name, (nr, nc, rp, ci) in cases.items():
        v = rng.standard_normal(ci.size)
        trp = np.empty(nc + 1, I32)
        tci = np.empty(ci.size, I32)
        tv = np.empty(ci.size)
        assert R.ref_transpose(nr, nc, p(rp), p(ci), p(v), p(trp), p(tci), p(tv)) == 0
        tout.update({f"{name}_shape": np.array([nr, nc]), f"{name}_row_ptr": rp,
                     f"{name}_col_ind": ci, f"{name}_vals": v, f"{name}_t_row_ptr": trp,
                     f"{name}_t_col_ind": tci, f"{name}_t_vals": tv})
    np.savez_compressed(os.path.join(HERE, "transpose.npz"), **tout)

    gmsh_fixtures(R)
    print("fixtures written to", HERE)


MESHES = "/root/reference/data/meshes"


def gmsh_read(R, path):
    """femx's GmshReader::read: node coordinates, triangles, and boundary
    facets (physical tag, 2 node ids)."""
    dims = np.zeros(3, I32)
    assert R.ref_gmsh_read(path.encode(), 1, p(dims), None, None, None, None, None) == 0, \
        R.ref_last_error()
    N, E, F = (int(v) for v in dims)
    x, y = np.empty(N), np.empty(N)
    conn, fp, fn = np.empty(3 * E, I32), np.empty(F, I32), np.empty(2 * F, I32)
    assert R.ref_gmsh_read(path.encode(), 0, p(dims), p(x), p(y), p(conn), p(fp), p(fn)) == 0
    return x, y, conn, fp, fn


def gmsh_bc(R, path, ptag, ncomp, comp):
    """DirichletBC::addBoundary(space, ptag, ...) on the P1 space with ncomp
    components: the DOFs in femx's order."""
    n = C.c_int32()
    assert R.ref_gmsh_bc_dofs(path.encode(), ptag, ncomp, comp, C.byref(n), None) == 0, \
        R.ref_last_error()
    d = np.empty(n.value, I32)
    assert R.ref_gmsh_bc_dofs(path.encode(), ptag, ncomp, comp, C.byref(n), p(d)) == 0
    return d


def ref_bcmap(R, rp, ci, dofs):
    N = rp.size - 1
    h = C.c_void_p()
    assert R.ref_bc_build(N, p(rp), p(ci), dofs.size, p(dofs), C.byref(h)) == 0, R.ref_last_error()
    nce = R.ref_bc_num_col_entries(h)
    arrs = dict(bc_rows=np.empty(dofs.size, I32), diag=np.empty(dofs.size, I32),
                col_offsets=np.empty(dofs.size + 1, I32), col_entries=np.empty(nce, I32),
                col_rows=np.empty(nce, I32), bc_mask=np.empty(N, I32))
    R.ref_bc_export(h, *(p(arrs[k]) for k in ("bc_rows", "diag", "col_offsets", "col_entries",
                                              "col_rows", "bc_mask")))
    R.ref_bc_free(h)
    return arrs


def ref_forward(R, rp, ci, dofs, vals, rhs, g):
    """prepareForwardSolve (BoundaryMap.cpp:218-247) on copies."""
    v, r = vals.copy(), rhs.copy()
    assert R.ref_bc_apply(1, rp.size - 1, p(rp), p(ci), dofs.size, p(dofs), p(v), p(r), p(g),
                          p(np.zeros(rp.size - 1)), 1.0) == 0, R.ref_last_error()
    return v, r


def gmsh_fixtures(R):
    """femx's own unstructured meshes (data/meshes, Gmsh 4.1) read by femx's
    GmshReader: makeAssemblyMap on the P1 DOF table, PoissonComponents
    assembly (kappa = 1), makeBoundaryMap on physical groups with
    prepareForwardSolve, and NavierOperator time assembly with the Navier-
    Stokes model's boundary DOFs (velocity on walls and inlet, one pressure
    DOF; ascending as femx's orderedBC leaves them, model/ns/Boundary.cpp:
    19-33, 242-299) and prepareForwardSolve on that system.
        gmsh_2d_square_res_low.npz  (walls: bottom 12, then top 9)
        gmsh_2d_rectangle.npz       (wall 6; NS: wall 6 and inlet 4)
        gmsh_2d_square_res_very_low.npz  (NS: all four walls)
    and the mesh alone (coordinates, triangles, facets) of the next finer
    square, for refinement studies:
        gmsh_2d_square_res_mid.npz"""
    rng = np.random.default_rng(20261204)
    cases = {"2d_square_res_low": ([12, 9], None),
             "2d_rectangle": ([6], [6, 4]),
             "2d_square_res_very_low": ([12, 9], [9, 10, 11, 12])}
    for name, (p1_tags, ns_tags) in cases.items():
        path = os.path.join(MESHES, name + ".msh")
        x, y, conn, fp, fn = gmsh_read(R, path)
        N, E = x.size, conn.size // 3
        out = dict(x=x, y=y, conn=conn, facet_ptag=fp, facet_nids=fn)
        off = O.uniform_offsets(E, 3)
        rp, ci, jo, jm = ref_map(R, N, N, off, conn)
        out.update(row_ptr=rp, col_ind=ci, jac_offsets=jo, jac_map=jm)
        u = rng.standard_normal(N)
        vals = np.empty(ci.size)
        res = np.empty(N)
        rp2 = np.empty(N + 1, I32)
        ci2 = np.empty(ci.size, I32)
        assert R.ref_poisson_p1(N, p(x), p(y), E, p(conn), p(u), ci.size, p(rp2), p(ci2),
                                p(vals), p(res)) == 0, R.ref_last_error()
        assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci)
        out.update(state=u, vals=vals, res=res)
        dofs = np.concatenate([gmsh_bc(R, path, t, 1, 0) for t in p1_tags]).astype(I32)
        bc = ref_bcmap(R, rp, ci, dofs)
        out.update(bc_dofs=dofs, **{"bc_" + k: v for k, v in bc.items()})
        rhs0, g = rng.standard_normal(N), rng.standard_normal(dofs.size)
        fv, fr = ref_forward(R, rp, ci, dofs, vals, rhs0, g)
        out.update(fwd_rhs0=rhs0, fwd_g=g, fwd_vals=fv, fwd_rhs=fr)
        if ns_tags is not None:
            nd = 3 * N
            hist, nxt = rng.standard_normal(2 * nd), rng.standard_normal(nd)
            d9, grp, gci, gvals, gres = O.ref_ns_gls(x, y, conn, 1.0, 1e-3, 0.01, 1, 2, hist, nxt)
            vel = [gmsh_bc(R, path, t, 2, c) for t in ns_tags for c in (0, 1)]
            nsd = np.unique(np.concatenate(vel + [np.array([2 * N], I32)])).astype(I32)
            gbc = ref_bcmap(R, grp, gci, nsd)
            grhs0, gg = rng.standard_normal(nd), rng.standard_normal(nsd.size)
            gfv, gfr = ref_forward(R, grp, gci, nsd, gvals, grhs0, gg)
            out.update(ns_hist=hist, ns_nxt=nxt, ns_dofs9=d9, ns_row_ptr=grp, ns_col_ind=gci,
                       ns_vals=gvals, ns_res=gres, ns_bc_dofs=nsd, ns_bc_mask=gbc["bc_mask"],
                       ns_bc_diag=gbc["diag"], ns_fwd_rhs0=grhs0, ns_fwd_g=gg, ns_fwd_vals=gfv,
                       ns_fwd_rhs=gfr)
        np.savez_compressed(os.path.join(HERE, f"gmsh_{name}.npz"), **out)
    x, y, conn, fp, fn = gmsh_read(R, os.path.join(MESHES, "2d_square_res_mid.msh"))
    np.savez_compressed(os.path.join(HERE, "gmsh_2d_square_res_mid.npz"), x=x, y=y, conn=conn,
                        facet_ptag=fp, facet_nids=fn)


if __name__ == "__main__":
    main()
