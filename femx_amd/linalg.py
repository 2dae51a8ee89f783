"""Device CSR storage and handlers, mirroring femx::linalg.

`CsrPattern` follows femx's immutable, shared CSR pattern with a layout
identity (femx/linalg/CsrPattern.hpp:28-32,43-217); `CsrMatrix` owns device
values over a shared pattern (femx/linalg/CsrMatrix.hpp:17-105);
`MatrixHandler` and `VectorHandler` mirror the CUDA handlers
(femx/linalg/handler/MatrixHandler.hpp:193-378, VectorHandler.hpp:190-477)
and run on the HIP kernels of libfemx_amd.so.  Device memory is torch's; the
arithmetic is the library's.
"""
from __future__ import annotations

import ctypes as C
import itertools

import numpy as np
import torch

from . import _lib

_layout_ids = itertools.count(1)


def ptr(a):
    """Raw address of a torch tensor or numpy array (None stays None)."""
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        return a.data_ptr() if a.numel() > 0 else None
    return a.ctypes.data if a.size > 0 else None


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _require(cond, msg):
    if not cond:
        raise RuntimeError(msg)


def _check_dev_vec(t, n, dtype, what):
    _require(isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous(),
             f"{what} must be a contiguous CUDA tensor of {dtype}")
    _require(t.numel() == n, f"{what} has size {t.numel()}, expected {n}")


class CsrPattern:
    """Immutable CSR structure with a layout identity shared by matrices."""

    def __init__(self, rows, cols, row_ptr, col_ind, layout_id=None):
        row_ptr = np.ascontiguousarray(row_ptr, np.int32)
        col_ind = np.ascontiguousarray(col_ind, np.int32)
        _require(rows >= 0 and cols >= 0, "CsrPattern dimensions must be non-negative")
        _require(row_ptr.size == rows + 1, "CsrPattern row_ptr size must be rows + 1")
        _require(row_ptr[0] == 0 and np.all(np.diff(row_ptr) >= 0),
                 "CsrPattern row_ptr must start at zero and be non-decreasing")
        _require(row_ptr[-1] == col_ind.size, "CsrPattern nnz does not match col_ind")
        _require(col_ind.size == 0 or (col_ind.min() >= 0 and col_ind.max() < cols),
                 "CsrPattern column index is out of range")
        self.rows, self.cols = int(rows), int(cols)
        self.row_ptr, self.col_ind = row_ptr, col_ind
        self.layout_id = next(_layout_ids) if layout_id is None else layout_id
        self._dev = {}
        self._tmap = {}
        self._off16 = {}
        self._stencil = {}

    @property
    def nnz(self):
        return int(self.col_ind.size)

    def device_arrays(self, device="cuda"):
        dev = torch.device(device)
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = (torch.from_numpy(self.row_ptr).to(dev),
                              torch.from_numpy(self.col_ind).to(dev))
        return self._dev[key]

    def col_off16(self, device="cuda"):
        """16-bit column offsets (col_ind - row) on the device, built once per
        device; None when the pattern's bandwidth does not fit 16 bits."""
        dev = torch.device(device)
        key = str(dev)
        if key not in self._off16:
            off = np.empty(max(self.nnz, 1), np.int16)
            rc = _lib.lib().fx_csr_col_off16(self.rows, ptr(self.row_ptr), ptr(self.col_ind), ptr(off))
            self._off16[key] = torch.from_numpy(off[: self.nnz]).to(dev) if rc == 0 else None
        return self._off16[key]

    STENCIL_MAX_LEN = 8  # FX_STENCIL_MAX_LEN

    def stencil_plan(self, device="cuda"):
        """The 64-row stencil blocks of the multi-vector product
        (fx_csr_stencil_plan), built once per device: (blk_code, st_off) device
        tensors and the fraction of rows in stencil blocks."""
        dev = torch.device(device)
        key = str(dev)
        if key not in self._stencil:
            nb, no = C.c_int32(), C.c_int32()
            args = (self.rows, self.cols, ptr(self.row_ptr), ptr(self.col_ind),
                    self.STENCIL_MAX_LEN)
            _lib.call("fx_csr_stencil_plan", *args, C.byref(nb), C.byref(no), None, None)
            code = np.empty(max(nb.value, 1), np.int32)
            off = np.empty(max(no.value, 1), np.int32)
            _lib.call("fx_csr_stencil_plan", *args, C.byref(nb), C.byref(no), ptr(code), ptr(off))
            code = code[: nb.value]
            frac = 64.0 * np.count_nonzero(code >= 0) / max(self.rows, 1)
            self._stencil[key] = (torch.from_numpy(code).to(dev), torch.from_numpy(off).to(dev),
                                  frac, code)
        return self._stencil[key]

    def _host_transpose(self):
        if "host" not in self._tmap:
            t_ptr = np.empty(self.cols + 1, np.int32)
            t_row = np.empty(max(self.nnz, 1), np.int32)
            t_perm = np.empty(max(self.nnz, 1), np.int32)
            _lib.call("fx_csr_transpose_map", self.rows, self.cols, ptr(self.row_ptr),
                      ptr(self.col_ind), ptr(t_ptr), ptr(t_row), ptr(t_perm))
            self._tmap["host"] = (t_ptr, t_row[: self.nnz], t_perm[: self.nnz])
        return self._tmap["host"]

    def transpose_map(self, device="cuda"):
        """(t_ptr, t_row, t_perm) device tensors, built once per device."""
        dev = torch.device(device)
        key = str(dev)
        if key not in self._tmap:
            self._tmap[key] = tuple(torch.from_numpy(a).to(dev) for a in self._host_transpose())
        return self._tmap[key]

    def transposed(self):
        """The transpose's pattern (its own layout), built once: rows are this
        pattern's columns, each listing its source rows in ascending order, as
        MatrixHandler::transpose lays it out (MatrixHandler.cpp:37-70)."""
        if "pattern" not in self._tmap:
            t_ptr, t_row, _ = self._host_transpose()
            self._tmap["pattern"] = CsrPattern(self.cols, self.rows, t_ptr, t_row)
        return self._tmap["pattern"]


class CsrMatrix:
    """Device CSR matrix: shared pattern, owned FP64 values."""

    def __init__(self, pattern: CsrPattern, device="cuda", vals=None):
        self.pattern = pattern
        self.device = torch.device(device)
        if vals is None:
            vals = torch.zeros(pattern.nnz, dtype=torch.float64, device=self.device)
        _check_dev_vec(vals, pattern.nnz, torch.float64, "CsrMatrix values")
        self.vals = vals

    @property
    def rows(self):
        return self.pattern.rows

    @property
    def cols(self):
        return self.pattern.cols

    @property
    def nnz(self):
        return self.pattern.nnz

    def clone(self):
        return CsrMatrix(self.pattern, self.device, self.vals.clone())

    def to_dense(self):
        """Dense host copy (tests and small problems only)."""
        v = self.vals.detach().cpu().numpy()
        out = np.zeros((self.rows, self.cols))
        rp, ci = self.pattern.row_ptr, self.pattern.col_ind
        for r in range(self.rows):
            out[r, ci[rp[r]:rp[r + 1]]] += v[rp[r]:rp[r + 1]]
        return out


class MatrixHandler:
    """Stream-ordered CSR operations (MatrixHandler<CudaCsrBackend>)."""

    def __init__(self, stream=None):
        self.stream = stream

    def zero(self, mat: CsrMatrix):
        mat.vals.zero_()

    def copy(self, src: CsrMatrix, dst: CsrMatrix):
        _require(src.pattern.layout_id == dst.pattern.layout_id,
                 "CSR copy requires matrices with the same layout")
        dst.vals.copy_(src.vals)

    def matvec(self, mat: CsrMatrix, x, y, alpha=1.0, beta=0.0):
        """y = alpha*A*x + beta*y (femx MatrixHandler::matvec)."""
        _check_dev_vec(x, mat.cols, torch.float64, "CSR matvec input")
        _check_dev_vec(y, mat.rows, torch.float64, "CSR matvec output")
        rp, ci = mat.pattern.device_arrays(mat.device)
        off = mat.pattern.col_off16(mat.device)
        if off is not None:  # the pattern's column offsets fit 16 bits: narrower index stream
            _lib.call("fx_csr_matvec16", mat.rows, mat.cols, ptr(rp), ptr(off), ptr(mat.vals), ptr(x),
                      ptr(y), float(alpha), float(beta), stream_handle(self.stream))
        else:
            _lib.call("fx_csr_matvec", mat.rows, mat.cols, ptr(rp), ptr(ci), ptr(mat.vals), ptr(x),
                      ptr(y), float(alpha), float(beta), stream_handle(self.stream))
        return y

    STENCIL_MIN_FRAC = 0.5  # the stencil kernel when at least this share of rows is in stencil blocks

    def matvec_multi(self, mat: CsrMatrix, X, Y, alpha=1.0, beta=0.0, stencil=None):
        """Y[j] = alpha*A*X[j] + beta*Y[j] for each row j of the 2-D tensors
        X [nv, cols] and Y [nv, rows]: nv products that read the matrix once
        (per 16 vectors); each equals matvec's result bit for bit.  When at
        least half of the rows lie in 64-row blocks that share one
        column-offset stencil (structured meshes; the plan is built on first
        use), those blocks read each band of the vectors once, coalesced
        (fx_csr_matvec_multi_st); otherwise every row gathers its entries
        (fx_csr_matvec_multi).  `stencil` True/False forces either kernel."""
        _require(X.dim() == 2 and Y.dim() == 2 and X.shape[0] == Y.shape[0]
                 and X.shape[1] == mat.cols and Y.shape[1] == mat.rows
                 and X.is_contiguous() and Y.is_contiguous() and X.dtype == torch.float64
                 and Y.dtype == torch.float64 and X.is_cuda and Y.is_cuda,
                 "CSR multi-vector product needs contiguous CUDA float64 [nv, n] tensors")
        rp, ci = mat.pattern.device_arrays(mat.device)
        if stencil is not False and mat.rows > 0:
            code, st_off, frac, _ = mat.pattern.stencil_plan(mat.device)
            if stencil or frac >= self.STENCIL_MIN_FRAC:
                _lib.call("fx_csr_matvec_multi_st", mat.rows, mat.cols, ptr(rp), ptr(ci),
                          ptr(mat.vals), ptr(code), ptr(st_off), X.shape[0], ptr(X), mat.cols,
                          ptr(Y), mat.rows, float(alpha), float(beta), stream_handle(self.stream))
                return Y
        off = mat.pattern.col_off16(mat.device)
        _lib.call("fx_csr_matvec_multi", mat.rows, mat.cols, ptr(rp),
                  None if off is not None else ptr(ci), ptr(off) if off is not None else None,
                  ptr(mat.vals), X.shape[0], ptr(X), mat.cols, ptr(Y), mat.rows, float(alpha),
                  float(beta), stream_handle(self.stream))
        return Y

    def transpose(self, src: CsrMatrix, dst: CsrMatrix = None):
        """dst = src^T (MatrixHandler::transpose, MatrixHandler.cu:401-500):
        the transposed pattern is built once per source layout and cached;
        values move through the cached permutation (one gather).  Returns dst,
        allocated on the transposed layout when None or of another layout."""
        _require(dst is None or dst is not src, "CSR transpose does not support in-place output")
        tp = src.pattern.transposed()
        if dst is None or dst.pattern.layout_id != tp.layout_id:
            dst = CsrMatrix(tp, src.device)
        _, _, t_perm = src.pattern.transpose_map(src.device)
        _lib.call("fx_vec_gather", src.nnz, ptr(t_perm), ptr(src.vals), src.nnz, ptr(dst.vals),
                  stream_handle(self.stream))
        return dst

    def matvecT(self, mat: CsrMatrix, x, y, alpha=1.0, beta=0.0):
        """y = alpha*A^T*x + beta*y (femx MatrixHandler::matvecT)."""
        _check_dev_vec(x, mat.rows, torch.float64, "CSR matvecT input")
        _check_dev_vec(y, mat.cols, torch.float64, "CSR matvecT output")
        t_ptr, t_row, t_perm = mat.pattern.transpose_map(mat.device)
        _lib.call("fx_csr_matvec_t", mat.rows, mat.cols, ptr(t_ptr), ptr(t_row), ptr(t_perm),
                  ptr(mat.vals), ptr(x), ptr(y), float(alpha), float(beta),
                  stream_handle(self.stream))
        return y


class VectorHandler:
    """Stream-ordered vector operations (VectorHandler<CudaCsrBackend>)."""

    def __init__(self, stream=None):
        self.stream = stream
        self._work = {}

    def axpby(self, a, x, b, y):
        """y = a*x + b*y."""
        _check_dev_vec(x, y.numel(), torch.float64, "axpby input")
        _lib.call("fx_vec_axpby", y.numel(), float(a), ptr(x), float(b), ptr(y),
                  stream_handle(self.stream))
        return y

    def dot(self, x, y, out=None):
        """Deterministic dot product into a one-element device tensor."""
        _check_dev_vec(y, x.numel(), torch.float64, "dot input")
        key = str(x.device)
        if key not in self._work:
            self._work[key] = torch.empty(_lib.lib().fx_dot_work_len(), dtype=torch.float64,
                                          device=x.device)
        if out is None:
            out = torch.empty(1, dtype=torch.float64, device=x.device)
        _lib.call("fx_vec_dot", x.numel(), ptr(x), ptr(y), ptr(self._work[key]), ptr(out),
                  stream_handle(self.stream))
        return out
