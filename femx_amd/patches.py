"""Patch plans of the single-pass VJPs (fx_patches_create): elements in
Morton-ordered patches whose per-DOF sums are finished in LDS, with the seam
DOFs' partials summed by a second pass.  Used by the Taylor-Hood VJP (15 DOFs
per element, femx_amd.ns) and the P1 log-kappa VJP (3, femx_amd.inverse);
`P1AsmPlan` is the P1 assembly's owner-computes plan (femx_amd.assembly)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .linalg import ptr

_KEYS = ("patch_off", "patch_elem", "elem_vert", "elem_ldof", "ldof_off", "ldof_dof",
         "ldof_part", "ldof_end", "seam_off", "seam_dof")


class Patches:
    """A patch plan on one device (host copies in `h`, device copies in `d`,
    the C-ABI struct in `plan`), with its seam-partials workspace: VJPs that
    share it (one mesh) on different streams must not overlap."""

    def __init__(self, ndof, num_elems, dofs_per_elem, h_dofs, h_vconn3, h_xv, h_yv, device,
                 max_patch=256, part_width=1):
        dofs = np.ascontiguousarray(h_dofs, np.int32)
        vconn = np.ascontiguousarray(h_vconn3, np.int32)
        xv = np.ascontiguousarray(h_xv, np.float64)
        yv = np.ascontiguousarray(h_yv, np.float64)
        h = C.c_void_p()
        _lib.call("fx_patches_create", int(ndof), int(num_elems), int(dofs_per_elem), ptr(dofs),
                  ptr(vconn), xv.size, ptr(xv), ptr(yv), int(max_patch), C.byref(h))
        try:
            info = _lib.PatchInfo()
            _lib.call("fx_patches_info_get", h, C.byref(info))
            P, L, S, E = info.num_patches, info.num_ldofs, info.num_seam, int(num_elems)
            a = {"patch_off": np.empty(P + 1, np.int32), "patch_elem": np.empty(max(E, 1), np.int32),
                 "elem_vert": np.empty(max(3 * E, 1), np.int32),
                 "elem_ldof": np.empty(max(info.ldof_stride * E, 1), np.uint16),
                 "ldof_off": np.empty(P + 1, np.int32), "ldof_dof": np.empty(max(L, 1), np.int32),
                 "ldof_part": np.empty(max(L, 1), np.int32),
                 "ldof_end": np.empty(max(L, 1), np.uint16),
                 "seam_off": np.empty(S + 1, np.int32), "seam_dof": np.empty(max(S, 1), np.int32)}
            _lib.call("fx_patches_export", h, *[ptr(a[k]) for k in _KEYS])
        finally:
            _lib.lib().fx_patches_destroy(h)
        self.h = a
        self.num_patches, self.num_ldofs, self.num_seam = P, L, S
        self.num_parts, self.max_patch_ldofs = info.num_parts, info.max_patch_ldofs
        self.ldof_stride = info.ldof_stride
        dev = torch.device(device)
        # uint16 tables travel as int16 bit patterns (torch has no uint16 copy path)
        self.d = {k: torch.from_numpy(v.view(np.int16) if v.dtype == np.uint16 else v).to(dev)
                  for k, v in a.items()}
        # part_width doubles per seam partial (4 for the GLS history VJP's two lags)
        self.parts = torch.empty(max(part_width * self.num_parts, 1), dtype=torch.float64,
                                 device=dev)
        d = self.d
        self.plan = _lib.PatchPlan(P, S, self.num_parts, self.max_patch_ldofs,
                                   *[ptr(d[k]) for k in _KEYS], ptr(self.parts))


class P1AsmPlan:
    """The P1 assembly's owner-computes patch plan (fx_p1_asm_create) on one
    device: host copies in `h`, device copies in `d`, the C-ABI struct in
    `plan`."""

    MAX_TILE = 4096  # CSR values per patch tile (32 KiB of LDS)
    _KEYS = ("patch_off", "patch_elem", "elem_ldof", "elem_pos", "ldof_off", "ldof_dof",
             "ldof_tile", "patch_tile", "run_off", "runs", "rows_left")

    def __init__(self, h_conn, h_x, h_y, pattern, device, max_patch=256, max_tile=MAX_TILE):
        conn = np.ascontiguousarray(h_conn, np.int32)
        x = np.ascontiguousarray(h_x, np.float64)
        y = np.ascontiguousarray(h_y, np.float64)
        E, N = conn.size // 3, pattern.rows
        hd = C.c_void_p()
        _lib.call("fx_p1_asm_create", N, E, ptr(conn), ptr(x), ptr(y), ptr(pattern.row_ptr),
                  ptr(pattern.col_ind), int(max_patch), int(max_tile), C.byref(hd))
        try:
            info = _lib.P1AsmInfo()
            _lib.call("fx_p1_asm_info_get", hd, C.byref(info))
            P, M, L = info.num_patches, info.num_entries, info.num_ldofs
            a = {"patch_off": np.empty(P + 1, np.int32), "patch_elem": np.empty(max(M, 1), np.int32),
                 "elem_ldof": np.empty(max(4 * M, 1), np.uint16),
                 "elem_pos": np.empty(max(4 * M, 1), np.uint16),
                 "ldof_off": np.empty(P + 1, np.int32), "ldof_dof": np.empty(max(L, 1), np.int32),
                 "ldof_tile": np.empty(max(L, 1), np.uint16),
                 "patch_tile": np.empty(max(P, 1), np.int32),
                 "run_off": np.empty(P + 1, np.int32),
                 "runs": np.empty(max(2 * info.num_runs, 2), np.int32),
                 "rows_left": np.empty(max(info.num_left, 1), np.int32)}
            _lib.call("fx_p1_asm_export", hd, *[ptr(a[k]) for k in self._KEYS])
        finally:
            _lib.lib().fx_p1_asm_destroy(hd)
        self.h = a
        self.num_patches, self.num_entries, self.num_ldofs = P, M, L
        self.num_left, self.max_patch_ldofs = info.num_left, info.max_patch_ldofs
        self.max_tile, self.max_runs = info.max_tile, info.max_runs
        dev = torch.device(device)
        self.d = {k: torch.from_numpy(v.view(np.int16) if v.dtype == np.uint16 else v).to(dev)
                  for k, v in a.items()}
        d = self.d
        self.plan = _lib.P1AsmPlan(P, self.num_left, self.max_patch_ldofs, self.max_tile,
                                   self.max_runs, *[ptr(d[k]) for k in self._KEYS])
