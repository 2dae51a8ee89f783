"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py for its cpu_baseline leg; never by the product.

Also exposes the compiled femx reference (oracle/_ref/libfemx_ref.so) when it
has been built, as `ref()`; callers must treat its absence as "not
available" and skip.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

I32 = np.int32
F64 = np.float64
_pi = C.POINTER(C.c_int32)
_pd = C.POINTER(C.c_double)


def _p(a):
    if a is None:
        return None
    if a.dtype == np.int32:
        return a.ctypes.data_as(_pi)
    if a.dtype == np.float64:
        return a.ctypes.data_as(_pd)
    raise TypeError(a.dtype)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle oracle)")
        _LIB = C.CDLL(path)
        _LIB.or_quad_tri.restype = C.c_int
        _LIB.or_map_build.restype = C.c_int
        _LIB.or_bc_build.restype = C.c_int
        _LIB.or_omp_threads.restype = C.c_int
        for name in ("or_p1_assemble", "or_p1_assemble_omp", "or_p1_vjp", "or_p1_vjp_omp",
                     "or_matvec", "or_matvecT"):
            getattr(_LIB, name).restype = None
        _LIB.or_th_assemble.argtypes = None
        for name in ("or_gls_element", "or_gls_assemble", "or_gls_assemble_omp",
                     "or_gls_element_jvp", "or_gls_element_hvjp", "or_gls_hist_vjp",
                     "or_gls_hist_jvp"):
            getattr(_LIB, name).restype = None
    return _LIB


def ref():
    """The compiled femx reference, or None when it was not built here."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libfemx_ref.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
        _REF.ref_last_error.restype = C.c_char_p
        _REF.ref_bc_num_col_entries.restype = C.c_int32
        _REF.ref_map_build.argtypes = [C.c_int32] * 3 + [_pi] * 4 + [C.POINTER(C.c_void_p)]
        _REF.ref_map_sizes.argtypes = [C.c_void_p, _pi, _pi, _pi]
        _REF.ref_map_export.argtypes = [C.c_void_p, _pi, _pi, _pi, _pi]
        _REF.ref_map_free.argtypes = [C.c_void_p]
        _REF.ref_bc_build.argtypes = [C.c_int32, _pi, _pi, C.c_int32, _pi, C.POINTER(C.c_void_p)]
        _REF.ref_bc_num_col_entries.argtypes = [C.c_void_p]
        _REF.ref_bc_export.argtypes = [C.c_void_p] + [_pi] * 6
        _REF.ref_bc_free.argtypes = [C.c_void_p]
        _REF.ref_bc_apply.argtypes = ([C.c_int32, C.c_int32, _pi, _pi, C.c_int32, _pi]
                                      + [_pd] * 4 + [C.c_double])
        _REF.ref_spmv.argtypes = ([C.c_int32] * 3 + [_pi, _pi] + [_pd] * 3
                                  + [C.c_double, C.c_double])
        _REF.ref_poisson_p1.argtypes = ([C.c_int32, _pd, _pd, C.c_int32, _pi, _pd, C.c_int32,
                                         _pi, _pi, _pd, _pd])
        _REF.ref_gmsh_read.argtypes = [C.c_char_p, C.c_int, _pi, _pd, _pd, _pi, _pi, _pi]
        _REF.ref_gmsh_bc_dofs.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _pi, _pi]
        _REF.ref_ns_gls.argtypes = ([C.c_int32, _pd, _pd, C.c_int32, _pi] + [C.c_double] * 3
                                    + [C.c_int32, C.c_int32, _pd, _pd, C.c_int32]
                                    + [_pi, _pi, _pi, _pd, _pd])
        _REF.ref_transpose.argtypes = [C.c_int32, C.c_int32, _pi, _pi, _pd, _pi, _pi, _pd]
        _REF.ref_ns_constrained.argtypes = (
            [C.c_int, C.c_int32, _pd, _pd, C.c_int32, _pi, C.c_int32] + [C.c_double] * 3
            + [C.c_int32, _pi, C.c_int32, C.c_int32, _pi, _pi, _pd, C.c_int32, _pi, _pd, _pi, _pi,
               _pd, C.c_int32, C.c_int32, C.c_int32] + [_pd] * 4 + [_pi] + [_pd] * 6)
    return _REF


# --------------------------------------------------------------------------
# meshes and DOF tables

def mesh_p1(n):
    N, E = (n + 1) ** 2, 2 * n * n
    x = np.empty(N, F64)
    y = np.empty(N, F64)
    conn = np.empty(3 * E, I32)
    lib().or_mesh_p1(C.c_int(n), _p(x), _p(y), _p(conn))
    return x, y, conn


def mesh_p2(n):
    Nv, E = (2 * n + 1) ** 2, 2 * n * n
    x = np.empty(Nv, F64)
    y = np.empty(Nv, F64)
    conn6 = np.empty(6 * E, I32)
    vconn3 = np.empty(3 * E, I32)
    lib().or_mesh_p2(C.c_int(n), _p(x), _p(y), _p(conn6), _p(vconn3))
    return x, y, conn6, vconn3


def th_dofs(n, conn6, vconn3):
    E = conn6.size // 6
    d = np.empty(15 * E, I32)
    lib().or_th_dofs(C.c_int(n), C.c_int32(E), _p(conn6), _p(vconn3), _p(d))
    return d


def _take(ptr, n, dtype):
    if n == 0:
        out = np.empty(0, dtype)
    else:
        out = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
    lib().or_free(ptr)
    return out


def map_build(num_res, num_states, res_offsets, res_dofs, state_offsets, state_dofs):
    """makeAssemblyMap restatement -> (row_ptr, col_ind, jac_offsets, jac_map)."""
    res_offsets = np.ascontiguousarray(res_offsets, I32)
    res_dofs = np.ascontiguousarray(res_dofs, I32)
    state_offsets = np.ascontiguousarray(state_offsets, I32)
    state_dofs = np.ascontiguousarray(state_dofs, I32)
    E = res_offsets.size - 1
    rp, ci, jo, jm = (_pi(), _pi(), _pi(), _pi())
    nnz = C.c_int32()
    rc = lib().or_map_build(C.c_int32(num_res), C.c_int32(num_states), C.c_int32(E),
                            _p(res_offsets), _p(res_dofs), _p(state_offsets), _p(state_dofs),
                            C.byref(rp), C.byref(ci), C.byref(jo), C.byref(jm), C.byref(nnz))
    if rc != 0:
        raise RuntimeError(f"or_map_build failed ({rc})")
    total = int(np.sum(np.diff(res_offsets).astype(np.int64) * np.diff(state_offsets)))
    return (_take(rp, num_res + 1, I32), _take(ci, nnz.value, I32), _take(jo, E + 1, I32),
            _take(jm, total, I32))


def uniform_offsets(E, k):
    return (np.arange(E + 1, dtype=np.int64) * k).astype(I32)


def bc_build(row_ptr, col_ind, dofs):
    rows = row_ptr.size - 1
    dofs = np.ascontiguousarray(dofs, I32)
    nb = dofs.size
    diag = np.empty(nb, I32)
    col_off = np.empty(nb + 1, I32)
    mask = np.empty(rows, I32)
    ce, cr = _pi(), _pi()
    rc = lib().or_bc_build(C.c_int32(rows), _p(row_ptr), _p(col_ind), C.c_int32(nb), _p(dofs),
                           _p(diag), _p(col_off), C.byref(ce), C.byref(cr), _p(mask))
    if rc != 0:
        raise RuntimeError(f"or_bc_build failed ({rc})")
    nce = int(col_off[-1])
    return dict(bc_rows=dofs.copy(), diag=diag, col_offsets=col_off,
                col_entries=_take(ce, nce, I32), col_rows=_take(cr, nce, I32), bc_mask=mask)


def replace_rows(row_ptr, bc, vals, diag_val):
    lib().or_replace_rows(_p(row_ptr), C.c_int32(bc["bc_rows"].size), _p(bc["bc_rows"]),
                          _p(bc["diag"]), _p(vals), C.c_double(diag_val))


def replace_res(bc, state, bc_vals, res):
    lib().or_replace_res(C.c_int32(bc["bc_rows"].size), _p(bc["bc_rows"]), _p(state),
                         _p(bc_vals), _p(res))


def zero_boundary(bc, v):
    lib().or_zero_boundary(C.c_int32(bc["bc_rows"].size), _p(bc["bc_rows"]), _p(v))


def prepare_forward(row_ptr, bc, vals, rhs, bc_vals):
    lib().or_prepare_forward(_p(row_ptr), C.c_int32(bc["bc_rows"].size), _p(bc["bc_rows"]),
                             _p(bc["diag"]), _p(bc["col_offsets"]), _p(bc["col_entries"]),
                             _p(bc["col_rows"]), _p(bc["bc_mask"]), _p(vals), _p(rhs),
                             _p(np.ascontiguousarray(bc_vals, F64)))


def matvec(row_ptr, col_ind, vals, x, y, alpha=1.0, beta=0.0):
    lib().or_matvec(C.c_int32(row_ptr.size - 1), _p(row_ptr), _p(col_ind), _p(vals), _p(x),
                    _p(y), C.c_double(alpha), C.c_double(beta))


def matvecT(row_ptr, col_ind, vals, ncols, x, y, alpha=1.0, beta=0.0):
    lib().or_matvecT(C.c_int32(row_ptr.size - 1), C.c_int32(ncols), _p(row_ptr), _p(col_ind),
                     _p(vals), _p(x), _p(y), C.c_double(alpha), C.c_double(beta))


def quad_tri(order):
    rs = np.zeros(14, F64)
    w = np.zeros(7, F64)
    nq = lib().or_quad_tri(C.c_int(order), _p(rs), _p(w))
    if nq < 0:
        raise RuntimeError("unsupported order")
    return rs[: 2 * nq].reshape(nq, 2), w[:nq]


def p2_shape(r, s):
    N = np.zeros(6, F64)
    d = np.zeros(12, F64)
    lib().or_p2_shape(C.c_double(r), C.c_double(s), _p(N), _p(d))
    return N, d.reshape(6, 2)


def p1_element(xe, ye, kappa, fe, ue):
    Ke = np.zeros(9, F64)
    Re = np.zeros(3, F64)
    lib().or_p1_element(_p(np.ascontiguousarray(xe, F64)), _p(np.ascontiguousarray(ye, F64)),
                        C.c_double(kappa), _p(np.ascontiguousarray(fe, F64)),
                        _p(np.ascontiguousarray(ue, F64)), _p(Ke), _p(Re))
    return Ke.reshape(3, 3), Re


def p1_assemble(x, y, conn, kappa, f, u, jac_map, nnz, omp=False):
    N = x.size
    E = conn.size // 3
    vals = np.empty(nnz, F64)
    res = np.empty(N, F64)
    fn = lib().or_p1_assemble_omp if omp else lib().or_p1_assemble
    fn(C.c_int32(N), _p(x), _p(y), C.c_int32(E), _p(conn), _p(kappa), _p(f), _p(u),
       _p(jac_map), C.c_int32(nnz), _p(vals), _p(res))
    return vals, res


def p1_vjp(x, y, conn, m, u, lam, bc_mask=None, omp=False):
    N = x.size
    E = conn.size // 3
    jt = np.empty(N, F64)
    g = np.empty(E, F64)
    fn = lib().or_p1_vjp_omp if omp else lib().or_p1_vjp
    fn(C.c_int32(N), _p(x), _p(y), C.c_int32(E), _p(conn), _p(m), _p(u), _p(lam),
       _p(bc_mask), _p(jt), _p(g))
    return jt, g


def th_element(xe, ye, nu, se):
    Ke = np.zeros(225, F64)
    Re = np.zeros(15, F64)
    lib().or_th_element(_p(np.ascontiguousarray(xe, F64)), _p(np.ascontiguousarray(ye, F64)),
                        C.c_double(nu), _p(np.ascontiguousarray(se, F64)), _p(Ke), _p(Re))
    return Ke.reshape(15, 15), Re


def th_assemble(xv, yv, vconn3, dofs15, ndof, nu, state, jac_map, nnz, omp=False, nu_e=None):
    E = vconn3.size // 3
    vals = np.empty(nnz, F64)
    res = np.empty(ndof, F64)
    fn = lib().or_th_assemble_omp if omp else lib().or_th_assemble
    fn(_p(xv), _p(yv), C.c_int32(E), _p(vconn3), _p(dofs15), C.c_int32(ndof), C.c_double(nu),
       _p(nu_e), _p(state), _p(jac_map), C.c_int32(nnz), _p(vals), _p(res))
    return vals, res


def th_vjp(xv, yv, vconn3, dofs15, ndof, nu, state, lam, bc_mask=None, nu_e=None):
    """(J^T lam, g[E]) with g_e = lam~_e^T dR_e/dnu_e."""
    E = vconn3.size // 3
    jt = np.empty(ndof, F64)
    g = np.empty(E, F64)
    lib().or_th_vjp(_p(xv), _p(yv), C.c_int32(E), _p(vconn3), _p(dofs15), C.c_int32(ndof),
                    C.c_double(nu), _p(nu_e), _p(state), _p(lam), _p(bc_mask), _p(jt), _p(g))
    return jt, g


def gls_dofs(num_nodes, conn):
    """Element DOF table (9 per element) of femx's mixed P1/P1 space
    (femx/model/ns/Helper.cpp:40-50, femx/fem/MixedFESpace.cpp:57-84): the
    vector velocity field first (global 2*node + c), then pressure
    (2*num_nodes + node); local order u0 v0 u1 v1 u2 v2 p0 p1 p2."""
    c = conn.reshape(-1, 3).astype(np.int64)
    d = np.empty((c.shape[0], 9), np.int64)
    d[:, 0:6:2] = 2 * c
    d[:, 1:6:2] = 2 * c + 1
    d[:, 6:9] = 2 * num_nodes + c
    return d.reshape(-1).astype(I32)


def gls_element(xe, ye, rho, mu, dt, step, num_hist, hist_e, nxt_e):
    Ke = np.empty(81, F64)
    Re = np.empty(9, F64)
    lib().or_gls_element(_p(np.ascontiguousarray(xe, F64)), _p(np.ascontiguousarray(ye, F64)),
                         C.c_double(rho), C.c_double(mu), C.c_double(dt), C.c_int(step),
                         C.c_int(num_hist), _p(np.ascontiguousarray(hist_e, F64)),
                         _p(np.ascontiguousarray(nxt_e, F64)), _p(Ke), _p(Re))
    return Ke.reshape(9, 9), Re


def gls_assemble(x, y, conn, dofs9, num_states, rho, mu, dt, step, num_hist, hist, nxt,
                 jac_map, nnz, omp=False, jac=True):
    """GLS P1/P1 time assemble; hist is lag-major (num_hist x num_states)."""
    E = conn.size // 3
    vals = np.empty(nnz, F64) if jac else None
    res = np.empty(num_states, F64)
    fn = lib().or_gls_assemble_omp if omp else lib().or_gls_assemble
    fn(C.c_int32(num_states), _p(x), _p(y), C.c_int32(E), _p(conn), _p(dofs9), C.c_double(rho),
       C.c_double(mu), C.c_double(dt), C.c_int(step), C.c_int(num_hist), _p(hist), _p(nxt),
       _p(jac_map), C.c_int32(nnz), _p(vals), _p(res))
    return vals, res


def gls_element_jvp(xe, ye, rho, mu, dt, step, num_hist, hist_e, dhist_e, nxt_e):
    """(R_e, dR_e/dhist . dhist) by dual numbers; hist arrays lag-major x 9."""
    Re = np.empty(9, F64)
    dRe = np.empty(9, F64)
    f = lambda a: _p(np.ascontiguousarray(a, F64))  # noqa: E731
    lib().or_gls_element_jvp(f(xe), f(ye), C.c_double(rho), C.c_double(mu), C.c_double(dt),
                             C.c_int(step), C.c_int(num_hist), f(hist_e), f(dhist_e), f(nxt_e),
                             _p(Re), _p(dRe))
    return Re, dRe


def gls_element_hvjp(xe, ye, rho, mu, dt, step, num_hist, hist_e, nxt_e, adj_e):
    """(dR_e/dhist)^T adj for both lags: array [2, 9]."""
    out = np.empty(18, F64)
    h = np.zeros(18, F64)
    h[:np.asarray(hist_e).size] = hist_e
    f = lambda a: _p(np.ascontiguousarray(a, F64))  # noqa: E731
    lib().or_gls_element_hvjp(f(xe), f(ye), C.c_double(rho), C.c_double(mu), C.c_double(dt),
                              C.c_int(step), C.c_int(num_hist), _p(h), f(nxt_e), f(adj_e),
                              _p(out))
    return out.reshape(2, 9)


def gls_hist_vjp(x, y, conn, dofs9, num_states, rho, mu, dt, step, num_hist, lag, hist, nxt,
                 adj):
    out = np.empty(num_states, F64)
    E = conn.size // 3
    f = lambda a: _p(np.ascontiguousarray(a, F64))  # noqa: E731
    lib().or_gls_hist_vjp(C.c_int32(num_states), f(x), f(y), C.c_int32(E), _p(conn), _p(dofs9),
                          C.c_double(rho), C.c_double(mu), C.c_double(dt), C.c_int(step),
                          C.c_int(num_hist), C.c_int(lag), f(hist), f(nxt), f(adj), _p(out))
    return out


def gls_hist_jvp(x, y, conn, dofs9, num_states, rho, mu, dt, step, num_hist, lag, hist, nxt,
                 dirv):
    """(dR/dhist_lag) dir assembled from the dual-number element JVP."""
    out = np.empty(num_states, F64)
    E = conn.size // 3
    f = lambda a: _p(np.ascontiguousarray(a, F64))  # noqa: E731
    lib().or_gls_hist_jvp(C.c_int32(num_states), f(x), f(y), C.c_int32(E), _p(conn), _p(dofs9),
                          C.c_double(rho), C.c_double(mu), C.c_double(dt), C.c_int(step),
                          C.c_int(num_hist), C.c_int(lag), f(hist), f(nxt), f(dirv), _p(out))
    return out


def ref_ns_constrained(x, y, conn, nstep, dt, rho, mu, ctr_dofs, num_ctr_params, entries,
                       fixed_dofs, fixed_vals, time, ctr_off, num_prm, step, hist, nxt, prm, adj):
    """femx's NavierStokesModel residual under ConstrainedTimeResidual at one
    step (oracle/_ref).  entries: (row, col, weight) of the DirichletControl
    map; time: per-step (lower, upper, upper_weight).  Returns a dict."""
    R = ref()
    i32 = lambda a: np.ascontiguousarray(a, I32)  # noqa: E731
    f64 = lambda a: np.ascontiguousarray(a, F64)  # noqa: E731
    ent = np.asarray(entries, dtype=object).reshape(-1, 3)
    er, ec, ew = i32(ent[:, 0].astype(int)), i32(ent[:, 1].astype(int)), f64(ent[:, 2].astype(float))
    tl, tu, tw = (i32([t[0] for t in time]), i32([t[1] for t in time]),
                  f64([t[2] for t in time]))
    cd, fd, fv = i32(ctr_dofs), i32(fixed_dofs), f64(fixed_vals)
    dims = np.zeros(4, I32)
    z = np.zeros(1)
    args = [x.size, _p(f64(x)), _p(f64(y)), conn.size // 3, _p(i32(conn)), nstep, dt, rho, mu,
            cd.size, _p(cd), num_ctr_params, er.size, _p(er), _p(ec), _p(ew), fd.size, _p(fd),
            _p(fv), _p(tl), _p(tu), _p(tw), ctr_off, num_prm, step]
    rc = R.ref_ns_constrained(1, *args, _p(z), _p(z), _p(z), _p(z), _p(dims), *([_p(z)] * 6))
    if rc != 0:
        raise RuntimeError(R.ref_last_error().decode())
    ns_, nh, npar, nnz = (int(v) for v in dims)
    out = {k: np.empty(n, F64) for k, n in (("res", ns_), ("asm_res", ns_), ("asm_vals", nnz),
                                             ("grad", npar), ("prep_vals", nnz),
                                             ("prep_rhs", ns_))}
    rc = R.ref_ns_constrained(0, *args, _p(f64(hist)), _p(f64(nxt)), _p(f64(prm)), _p(f64(adj)),
                              _p(dims), _p(out["res"]), _p(out["asm_res"]), _p(out["asm_vals"]),
                              _p(out["grad"]), _p(out["prep_vals"]), _p(out["prep_rhs"]))
    if rc != 0:
        raise RuntimeError(R.ref_last_error().decode())
    out.update(num_states=ns_, num_hist=nh, num_param=npar, nnz=nnz)
    return out


def ref_ns_gls(x, y, conn, rho, mu, dt, step, num_hist, hist, nxt):
    """femx's own GLS time assembly (oracle/_ref); returns (dofs9, row_ptr,
    col_ind, vals, res)."""
    R = ref()
    N, E = x.size, conn.size // 3
    nd = 3 * N
    cap = 81 * E
    d9 = np.empty(9 * E, I32)
    rp = np.empty(nd + 1, I32)
    ci = np.empty(cap, I32)
    vals = np.empty(cap, F64)
    res = np.empty(nd, F64)
    rc = R.ref_ns_gls(N, _p(x), _p(y), E, _p(conn), rho, mu, dt, step, num_hist,
                      _p(np.ascontiguousarray(hist, F64)), _p(np.ascontiguousarray(nxt, F64)),
                      cap, _p(d9), _p(rp), _p(ci), _p(vals), _p(res))
    if rc != 0:
        raise RuntimeError(R.ref_last_error().decode())
    return d9, rp, ci[:rp[-1]].copy(), vals[:rp[-1]].copy(), res


# --------------------------------------------------------------------------
# Dirichlet control (femx/fem/DirichletControl.cpp:32-73, ControlMap.cpp:
# 36-342) and the constrained time residual (femx/assembly/
# ConstrainedTimeResidual.cpp:318-405), composed from the restated CSR and
# boundary operations above.

def control_matrix(rows, cols, entries):
    """makeControlMatrix: CSR with each row's entries in input order."""
    rp = np.zeros(rows + 1, np.int64)
    for r, _, _ in entries:
        rp[int(r) + 1] += 1
    rp = np.cumsum(rp)
    nxt = rp[:-1].copy()
    ci = np.empty(len(entries), I32)
    v = np.empty(len(entries), F64)
    for r, c, w in entries:
        k = nxt[int(r)]
        nxt[int(r)] += 1
        ci[k], v[k] = int(c), w
    return rp.astype(I32), ci, v


def control_vals(cm, step, prm):
    """controlVals (ControlMap.cpp:165-199)."""
    rp, ci, v = cm["P"]
    lo, hi, w = cm["time"][step]
    lo, hi = int(lo), int(hi)
    b = cm["num_ctr_params"]
    off = cm["ctr_off"]
    out = np.empty(cm["ctr_dofs"].size + cm["fixed_dofs"].size)
    ctl = np.empty(cm["ctr_dofs"].size)
    matvec(rp, ci, v, np.ascontiguousarray(prm[off + lo * b: off + lo * b + b]), ctl, 1.0 - w, 0.0)
    if hi != lo and w != 0.0:
        matvec(rp, ci, v, np.ascontiguousarray(prm[off + hi * b: off + hi * b + b]), ctl, w, 1.0)
    nf = cm["fixed_dofs"].size
    out[:ctl.size] = ctl
    out[ctl.size:] = cm["fixed_vals"][step * nf:(step + 1) * nf]
    return out


def add_control_jac_t(cm, step, adj, grad):
    """addControlJacT (ControlMap.cpp:309-342): grad += -P^T (adj on the
    controlled DOFs) at the step's levels."""
    rp, ci, v = cm["P"]
    lo, hi, w = cm["time"][step]
    lo, hi = int(lo), int(hi)
    b = cm["num_ctr_params"]
    off = cm["ctr_off"]
    compact = np.ascontiguousarray(adj[cm["ctr_dofs"]])
    for lev, wt in ((lo, 1.0 - w), (hi, w)):
        if lev == hi and (hi == lo or w == 0.0):
            continue
        g = np.ascontiguousarray(grad[off + lev * b: off + lev * b + b])
        matvecT(rp, ci, v, b, compact, g, -wt, 1.0)
        grad[off + lev * b: off + lev * b + b] = g
    return grad


def omp_threads():
    return lib().or_omp_threads()
