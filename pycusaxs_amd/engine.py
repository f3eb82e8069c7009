"""Python handle on one engine plan: a system's geometry, its shell tables and
the device workspace, driven through the C ABI (include/saxs_engine.h).

This is the per-run object the reference builds in RunSaxs::Run
(RunSaxs.cu:117-133: saxsKernel construction, scaledCell, resetHistogram-
Parameters, createMemory, cuFFT plan).  Device buffers are torch tensors;
launches go to torch's current HIP stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Mapping, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .geometry import (Histogram, border_planes, bspline_moduli, fft_size, form_factor_points,
                       form_factor_table, form_factor_vectors, half_spectrum_counts, hist_params,
                       inverse_box_matrices, inverse_box_matrix, is_triclinic, max_given_pad,
                       point_bins, point_kmax, reciprocal_bases, reciprocal_basis,
                       shell_bins, triclinic_bins, triclinic_kmax)

_NO_BIN_LIMIT = 2 ** 31 - 1     # bins past num_bins still mark q <= qcut


@dataclass
class SystemSpec:
    """Everything that fixes one plan."""
    n: Tuple[int, int, int]
    nn: Tuple[int, int, int]
    order: int
    symbols: Tuple[str, ...]
    sigma: float
    dq_req: float
    qcut_req: float
    dq_policy: str = "reference"


def resolve_grid(n: Sequence[int], sigma: float, scaled: Optional[Sequence[int]] = None):
    n = tuple(int(v) for v in n)
    if scaled:
        nn = tuple(int(v) for v in scaled)
    else:
        nn = tuple(fft_size(v, sigma) for v in n)
    return n, nn


def _ptr(t: Optional[torch.Tensor]):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _stream_handle():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class SaxsEngine:
    """One plan of the MI355X engine for a fixed system and first-frame box.

    forward(xyz, elem) accumulates the S6 histogram of a batch of frames into
    the engine's device accumulators H, C (float64), like saxsKernel's
    d_histogram / d_nhist (saxsKernel.cu:248-254)."""

    def __init__(self, spec: SystemSpec, box, max_atoms: int, max_frames: int = 1,
                 tile: Tuple[int, int] = (0, 0), device: Optional[torch.device] = None,
                 flags: int = 0, pad_given: Optional[Mapping[str, float]] = None):
        if not torch.cuda.is_available():
            raise RuntimeError("the MI355X SAXS engine needs a ROCm GPU (torch.cuda is unavailable)")
        self.lib = _lib.load()
        self.spec = spec
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.max_atoms = int(max_atoms)
        self.max_frames = int(max_frames)
        self.cell = inverse_box_matrix(box)
        self.oc = np.ascontiguousarray(np.diag(self.cell), dtype=np.float32)
        self.hist: Histogram = hist_params(spec.n, spec.nn, self.oc, spec.sigma, spec.dq_req,
                                           spec.qcut_req, spec.dq_policy)
        self.border = border_planes(spec.n[2])
        self.moduli = [bspline_moduli(spec.nn[a], spec.order) for a in range(3)]
        self._tile = (int(tile[0]), int(tile[1]))
        self._flags = int(flags)
        self._timing = False
        bins, ff, kmax = self._q_tables(self.hist, self.cell)
        self._box_key = tuple(float(v) for v in self.cell.ravel())
        self._plan = C.c_void_p()
        self._create_plan(self.cell, bins, ff, kmax)
        nb = self.hist.num_bins
        self.H = torch.zeros(nb, dtype=torch.float64, device=self.device)
        self.C = torch.zeros(nb, dtype=torch.float64, device=self.device)
        self.frames_done = 0
        self._generation = 0
        self._last = None
        self.pad_given: Optional[Dict[str, float]] = None
        if pad_given is not None:
            self.set_padding(pad_given)

    def _create_plan(self, cell, bins, ff, kmax):
        """saxs_plan_create for `cell` with its q tables (the previous plan,
        if any, is destroyed first; H, C and the frame count are kept)."""
        spec = self.spec
        mods = [np.ascontiguousarray(m, dtype=np.float32) for m in self.moduli]
        desc = _lib.SaxsDesc()
        desc.n[:] = list(spec.n)
        desc.nn[:] = list(spec.nn)
        desc.order = spec.order
        desc.num_elements = len(spec.symbols)
        desc.num_bins = self.hist.num_bins
        desc.border = self.border
        desc.lut_len = len(bins)
        desc.bin_lut = bins.ctypes.data_as(C.POINTER(C.c_int32))
        desc.ff_lut = ff.ctypes.data_as(C.POINTER(C.c_float))
        desc.mod_x = mods[0].ctypes.data_as(C.POINTER(C.c_float))
        desc.mod_y = mods[1].ctypes.data_as(C.POINTER(C.c_float))
        desc.mod_z = mods[2].ctypes.data_as(C.POINTER(C.c_float))
        desc.max_atoms = self.max_atoms
        desc.max_frames = self.max_frames
        desc.tile_x, desc.tile_y = self._tile
        desc.flags = self._flags
        desc.kmax[:] = list(kmax) if kmax is not None else [0, 0, 0]
        desc.cell[:] = [float(v) for v in cell.ravel()] if is_triclinic(cell) else [0.0] * 9
        plan = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.saxs_plan_create(C.byref(desc), C.byref(plan)), "saxs_plan_create")
        self.close()
        self._plan = plan
        self.plans_created = getattr(self, "plans_created", 0) + 1

    def _rebuild_plan(self, cell, bins, ff, kmax):
        """A box whose q tables do not fit the workspace (it keeps more
        frequencies than the first box): a plan sized for it replaces the
        current one; accumulated H, C, frame count, padding and timing carry
        over."""
        torch.cuda.current_stream().synchronize()
        self._create_plan(cell, bins, ff, kmax)
        if self.pad_given is not None:
            pads = self.pad_given
            self.pad_given = None
            self.set_padding(pads)
        if self._timing:
            self.set_timing(True)
        self._last = None

    # -- tables ---------------------------------------------------------
    def _q_tables(self, hist: Histogram, cell):
        """Bin and form-factor tables: per shell n = |m|^2 when the cell is
        orthorhombic with equal reciprocal spacings (kmax None), per point
        (|m_x|, |m_y|, |m_z|) of the box |m_a| <= kmax[a] for other
        orthorhombic cells, per signed point for triclinic cells."""
        s = hist.spacing
        if is_triclinic(cell):
            basis = reciprocal_basis(cell, self.spec.n, self.spec.nn)
            kmax = triclinic_kmax(basis, hist.qcut, self.spec.nn)
            bins = triclinic_bins(basis, hist.dq, hist.qcut, hist.num_bins, kmax)
            ff = form_factor_vectors(self.spec.symbols, basis, kmax, bins >= 0)
        elif s[0] == s[1] == s[2]:
            bins = shell_bins(s[0], hist.dq, hist.qcut, hist.num_bins)
            ff = form_factor_table(self.spec.symbols, s[0], len(bins))
            kmax = None
        else:
            kmax = point_kmax(s, hist.qcut)
            kmax = tuple(min(k, nn // 2) for k, nn in zip(kmax, self.spec.nn))
            bins = point_bins(s, hist.dq, hist.qcut, hist.num_bins, kmax)
            ff = form_factor_points(self.spec.symbols, s, kmax)
        return (np.ascontiguousarray(bins, dtype=np.int32),
                np.ascontiguousarray(ff, dtype=np.float32), kmax)

    def _q_tables_binned(self, h: Histogram, cell, h0: Histogram, cell0):
        """Tables for a frame in box `cell` binned by the spacings of box
        `cell0` (NVT: the reference sums every frame's spectrum and bins the
        sum once with the first box, saxsKernel.cu:184-189 and RunSaxs.cu:
        124-130,208-211).  The kept range and the form factors are the
        frame's own: its amplitude is zero where its own q exceeds qcut
        (saxsDeviceKernels.cu:177-216), so those points carry f = 0."""
        nn, syms = self.spec.nn, self.spec.symbols
        s, s0 = h.spacing, h0.spacing
        if is_triclinic(cell) or is_triclinic(cell0):
            if not is_triclinic(cell):
                raise ValueError("NVT: a frame with a rectangular box after a triclinic first "
                                 "box is not supported")
            basis = reciprocal_basis(cell, self.spec.n, nn)
            kmax = triclinic_kmax(basis, h.qcut, nn)
            bins = triclinic_bins(reciprocal_basis(cell0, self.spec.n, nn), h.dq, h.qcut,
                                  h.num_bins, kmax)
            keep = triclinic_bins(basis, h.dq, h.qcut, _NO_BIN_LIMIT, kmax) >= 0
            ff = form_factor_vectors(syms, basis, kmax, keep)
        elif s[0] == s[1] == s[2] and s0[0] == s0[1] == s0[2]:
            n_shells = len(shell_bins(s[0], h.dq, h.qcut, h.num_bins))
            b0 = shell_bins(s0[0], h.dq, h.qcut, h.num_bins)[:n_shells]
            bins = np.full(n_shells, -1, dtype=np.int32)
            bins[:len(b0)] = b0
            ff = form_factor_table(syms, s[0], n_shells)
            kmax = None
        else:
            kmax = point_kmax(s, h.qcut)
            kmax = tuple(min(k, n // 2) for k, n in zip(kmax, nn))
            bins = point_bins(s0, h.dq, h.qcut, h.num_bins, kmax)
            keep = point_bins(s, h.dq, h.qcut, _NO_BIN_LIMIT, kmax) >= 0
            ff = form_factor_points(syms, s, kmax) * keep
        return (np.ascontiguousarray(bins, dtype=np.int32),
                np.ascontiguousarray(ff, dtype=np.float32), kmax)

    def _hist_for(self, cell) -> Histogram:
        """This run's dq, qcut and num_bins (those of the first frame, as in
        the reference) with the reciprocal spacings of `cell`."""
        oc = np.ascontiguousarray(np.diag(cell), dtype=np.float32)
        h = hist_params(self.spec.n, self.spec.nn, oc, self.spec.sigma, self.spec.dq_req,
                        self.spec.qcut_req, self.spec.dq_policy)
        return Histogram(self.hist.dq, self.hist.qcut, self.hist.num_bins, h.spacing,
                         self.hist.dq_reset, self.hist.qcut_reset)

    def set_box(self, box, bin_box=None):
        """Rebuilds the q tables for a new box (NPT).  dq, qcut and num_bins
        stay those of the first frame, as in the reference.  With `bin_box`
        the frame's points are binned by that box's q instead (NVT)."""
        cell = inverse_box_matrix(box)
        oc = np.ascontiguousarray(np.diag(cell), dtype=np.float32)
        cell0 = None if bin_box is None else inverse_box_matrix(bin_box)
        if cell0 is not None and np.array_equal(cell0, cell):
            cell0 = None
        key = tuple(float(v) for v in cell.ravel())
        if cell0 is not None:
            key += tuple(float(v) for v in cell0.ravel())
        if key == self._box_key:
            return oc
        h = self._hist_for(cell)
        if cell0 is None:
            bins, ff, kmax = self._q_tables(h, cell)
        else:
            bins, ff, kmax = self._q_tables_binned(h, cell, self._hist_for(cell0), cell0)
        torch.cuda.current_stream().synchronize()
        try:
            if kmax is None:
                _lib.check(self.lib.saxs_plan_set_tables(self._plan, len(bins),
                                                         bins.ctypes.data_as(C.c_void_p),
                                                         ff.ctypes.data_as(C.c_void_p)),
                           "saxs_plan_set_tables")
            else:
                km = np.array(kmax, dtype=np.int32)
                cm = np.ascontiguousarray(cell, dtype=np.float32)
                cptr = cm.ctypes.data_as(C.c_void_p) if is_triclinic(cell) else None
                _lib.check(self.lib.saxs_plan_set_point_tables(self._plan, cptr,
                                                               km.ctypes.data_as(C.c_void_p),
                                                               len(bins),
                                                               bins.ctypes.data_as(C.c_void_p),
                                                               ff.ctypes.data_as(C.c_void_p)),
                           "saxs_plan_set_point_tables")
        except _lib.CapacityError:
            self._rebuild_plan(cell, bins, ff, kmax)
        self._box_key = key
        self.oc = oc
        self.cell = cell
        return oc

    def set_padding(self, pad_given: Optional[Mapping[str, float]]):
        """Bulk-solvent pad mode.  None: each frame's pad per element is the
        mean density of the border z-slabs (padding::avg, the reference's
        mode).  A mapping {symbol: atoms per cell voxel}: that constant for
        every frame (padding::given; elements it does not name get 0), e.g.
        geometry.explicit_solvent_pads(...) for explicit water and ions."""
        torch.cuda.current_stream().synchronize()
        if pad_given is None:
            _lib.check(self.lib.saxs_plan_set_padding(self._plan, None), "saxs_plan_set_padding")
            self.pad_given = None
            return
        pads = {str(k): float(v) for k, v in dict(pad_given).items()}
        arr = np.array([pads.get(s, 0.0) for s in self.spec.symbols], dtype=np.float64)
        limit = max_given_pad(self.spec.n[2])
        bad = [s for s, v in zip(self.spec.symbols, arr) if not 0.0 <= v <= limit]
        if bad:
            raise ValueError(f"given pad of {bad} outside [0, 2^31 / n_z = {limit:g}] atoms per voxel")
        _lib.check(self.lib.saxs_plan_set_padding(self._plan, arr.ctypes.data_as(C.c_void_p)),
                   "saxs_plan_set_padding")
        self.pad_given = pads
        self._last = None

    # -- compute --------------------------------------------------------
    def _prep(self, xyz: torch.Tensor, elem: torch.Tensor):
        if xyz.dim() == 2:
            xyz = xyz.unsqueeze(0)
        if xyz.dim() != 3 or xyz.shape[-1] != 3:
            raise ValueError(f"positions must be [frames, atoms, 3], got {tuple(xyz.shape)}")
        xyz = xyz.to(device=self.device, dtype=torch.float32).contiguous()
        elem = elem.to(device=self.device, dtype=torch.int32).contiguous()
        if elem.dim() != 1 or elem.shape[0] != xyz.shape[1]:
            raise ValueError("element ids must be [atoms] matching the positions")
        return xyz, elem

    def forward(self, xyz: torch.Tensor, elem: torch.Tensor, box=None, bin_box=None):
        """Accumulates frames into self.H / self.C (asynchronous).  With
        `bin_box` (NVT) the frames' points are binned by that box's q."""
        self._last = None
        xyz, elem = self._prep(xyz, elem)
        oc = self.set_box(box, bin_box) if box is not None else self.oc
        F = xyz.shape[0]
        for f0 in range(0, F, self.max_frames):
            chunk = xyz[f0:f0 + self.max_frames]
            _lib.check(self.lib.saxs_forward(
                self._plan, _ptr(chunk), _ptr(elem), int(chunk.shape[1]), int(chunk.shape[0]),
                oc.ctypes.data_as(C.c_void_p), _ptr(self.H), _ptr(self.C), _stream_handle()),
                "saxs_forward")
        self.frames_done += F

    # -- frames whose boxes differ ----------------------------------------
    def _batch_geometry(self, cells, bin_cell=None):
        """Key type, kept box, table length and per-frame q geometry for a
        stack of inverse cells [F, 3, 3] (saxs_forward_boxes), or None when a
        geometry spans more than 2^19 in magnitude (the device's exact
        comparisons)."""
        n, nn = self.spec.n, self.spec.nn
        h = self.hist
        F = len(cells)
        allc = cells if bin_cell is None else np.concatenate([cells, bin_cell[None]])
        bases = reciprocal_bases(allc, n, nn)
        a = np.abs(bases.reshape(len(bases), 9)).astype(np.float64)
        big = a.max(axis=1)
        small = np.where(a > 0, a, np.inf).min(axis=1)
        if np.any(np.log2(big / small) > 19.0):
            return None
        offd = bases.copy()
        offd[:, [0, 1, 2], [0, 1, 2]] = 0
        diag = np.diagonal(bases, axis1=1, axis2=2)
        bgeom = None
        qn, qd = float(h.qcut).as_integer_ratio()
        if np.any(offd != 0):
            kmax = [0, 0, 0]
            for bm in bases[:F]:
                kmax = [max(x, y) for x, y in zip(kmax, triclinic_kmax(bm, h.qcut, nn))]
            geom = bases[:F].reshape(F, 9)
            if bin_cell is not None:
                bgeom = bases[F].ravel()
            return 2, kmax, int(np.prod([2 * k + 1 for k in kmax])), geom, bgeom
        if np.all((diag[:, 0] == diag[:, 1]) & (diag[:, 1] == diag[:, 2])):
            n_max = 0
            for s in set(float(v) for v in diag[:F, 0]):
                sn, sd = s.as_integer_ratio()
                n_max = max(n_max, (qn * qn * sd * sd) // (sn * sn * qd * qd))
            geom = np.zeros((F, 9), dtype=np.float32)
            geom[:, 0] = diag[:F, 0]
            if bin_cell is not None:
                bgeom = np.zeros(9, dtype=np.float32)
                bgeom[0] = diag[F, 0]
            return 0, [0, 0, 0], int(n_max) + 1, geom, bgeom
        kmax = [0, 0, 0]
        for d in set(tuple(float(v) for v in row) for row in diag[:F]):
            km = point_kmax(d, h.qcut)
            kmax = [max(x, min(y, v // 2)) for x, y, v in zip(kmax, km, nn)]
        geom = np.zeros((F, 9), dtype=np.float32)
        geom[:, :3] = diag[:F]
        if bin_cell is not None:
            bgeom = np.zeros(9, dtype=np.float32)
            bgeom[:3] = diag[F]
        return 1, kmax, int(np.prod([k + 1 for k in kmax])), geom, bgeom

    def _form_coefficients(self) -> np.ndarray:
        if getattr(self, "_coef", None) is None:
            from .formfactors import coefficients
            self._coef = np.ascontiguousarray([coefficients(s) for s in self.spec.symbols],
                                              dtype=np.float32)
        return self._coef

    def forward_boxes(self, xyz: torch.Tensor, elem: torch.Tensor, boxes, bin_box=None):
        """Accumulates frames with one box each ([frames, 3, 3]) into H, C in
        one launch chain per max_frames chunk (saxs_forward_boxes): each frame
        is spread with its own inverse cell and its points take its own q,
        binned with the run's dq, qcut and bins (NVT: binned by bin_box's q).
        A batch that needs more kept frequencies than the plan holds makes a
        plan for its largest box; geometries outside the device's exact range
        go frame by frame through forward(box=...)."""
        self._last = None
        xyz, elem = self._prep(xyz, elem)
        F = xyz.shape[0]
        boxes = np.asarray(boxes, dtype=np.float32).reshape(F, 3, 3)
        cells = inverse_box_matrices(boxes)
        bin_cell = None if bin_box is None else inverse_box_matrix(bin_box)
        geo = self._batch_geometry(cells, bin_cell)
        if geo is None:
            for f in range(F):
                self.forward(xyz[f:f + 1], elem, box=boxes[f], bin_box=bin_box)
            return
        key_type, kmax, lut_len, geom, bgeom = geo
        coef = self._form_coefficients()
        cm = np.ascontiguousarray(cells.reshape(F, 9), dtype=np.float32)
        geom = np.ascontiguousarray(geom, dtype=np.float32)
        bg = None if bgeom is None else np.ascontiguousarray(bgeom, dtype=np.float32)
        for f0 in range(0, F, self.max_frames):
            k = min(self.max_frames, F - f0)
            fb = _lib.SaxsFrameBoxes()
            fb.cells = cm[f0:f0 + k].ctypes.data_as(C.POINTER(C.c_float))
            fb.qgeom = geom[f0:f0 + k].ctypes.data_as(C.POINTER(C.c_float))
            fb.bin_qgeom = bg.ctypes.data_as(C.POINTER(C.c_float)) if bg is not None else None
            fb.key_type = key_type
            fb.kmax[:] = [int(v) for v in kmax]
            fb.lut_len = lut_len
            fb.dq = float(self.hist.dq)
            fb.qcut = float(self.hist.qcut)
            fb.form_coef = coef.ctypes.data_as(C.POINTER(C.c_float))
            chunk = xyz[f0:f0 + k]
            args = (self._plan, _ptr(chunk), _ptr(elem), int(chunk.shape[1]), int(k), C.byref(fb),
                    _ptr(self.H), _ptr(self.C), _stream_handle())
            try:
                _lib.check(self.lib.saxs_forward_boxes(*args), "saxs_forward_boxes")
            except _lib.CapacityError:
                # a plan for the batch's largest box (its most kept frequencies)
                big = f0 + int(np.argmax([np.prod(np.diag(b)) for b in boxes[f0:f0 + k]]))
                self._rebuild_for_box(boxes[big])
                args = (self._plan,) + args[1:]
                try:
                    _lib.check(self.lib.saxs_forward_boxes(*args), "saxs_forward_boxes")
                except _lib.CapacityError:
                    frames_done = self.frames_done
                    for f in range(f0, f0 + k):
                        self.forward(xyz[f:f + 1], elem, box=boxes[f], bin_box=bin_box)
                    self.frames_done = frames_done
        self.frames_done += F

    def _rebuild_for_box(self, box):
        """A plan whose workspace is sized for `box` (its q tables)."""
        cell = inverse_box_matrix(box)
        bins, ff, kmax = self._q_tables(self._hist_for(cell), cell)
        self._rebuild_plan(cell, bins, ff, kmax)
        self._box_key = tuple(float(v) for v in cell.ravel())
        self.cell = cell
        self.oc = np.ascontiguousarray(np.diag(cell), dtype=np.float32)

    def forward_frames(self, xyz: torch.Tensor, elem: torch.Tensor, box=None):
        """Per-frame (H, C) as [frames, num_bins] float64 tensors."""
        xyz, elem = self._prep(xyz, elem)
        oc = self.set_box(box) if box is not None else self.oc
        F, nb = xyz.shape[0], self.hist.num_bins
        self._generation += 1
        # the plan keeps the spectra of this call when it fits in one launch
        self._last = ((xyz.data_ptr(), xyz._version, elem.data_ptr(), elem._version,
                       tuple(oc.tolist()), F), self._generation) if F <= self.max_frames else None
        Hf = torch.empty(F, nb, dtype=torch.float64, device=self.device)
        Cf = torch.empty(F, nb, dtype=torch.float64, device=self.device)
        for f0 in range(0, F, self.max_frames):
            chunk = xyz[f0:f0 + self.max_frames]
            _lib.check(self.lib.saxs_forward_frames(
                self._plan, _ptr(chunk), _ptr(elem), int(chunk.shape[1]), int(chunk.shape[0]),
                oc.ctypes.data_as(C.c_void_p), _ptr(Hf[f0:]), _ptr(Cf[f0:]), _stream_handle()),
                "saxs_forward_frames")
        return Hf, Cf

    def forward_plain(self, xyz: torch.Tensor, elem: torch.Tensor, box=None):
        """Reference-structured per-frame path (one launch chain per frame)."""
        self._last = None
        xyz, elem = self._prep(xyz, elem)
        oc = self.set_box(box) if box is not None else self.oc
        for f in range(xyz.shape[0]):
            _lib.check(self.lib.saxs_forward_plain(
                self._plan, _ptr(xyz[f]), _ptr(elem), int(xyz.shape[1]),
                oc.ctypes.data_as(C.c_void_p), _ptr(self.H), _ptr(self.C), _stream_handle()),
                "saxs_forward_plain")
        self.frames_done += xyz.shape[0]

    def backward(self, xyz: torch.Tensor, elem: torch.Tensor, gH: torch.Tensor, box=None,
                 generation: Optional[int] = None):
        """dL/dx [frames, atoms, 3] float32 for per-frame dL/dH [frames, num_bins].
        With `generation` equal to the value returned by the immediately
        preceding forward_frames on the same, unmodified inputs, the plan's
        stored spectra are reused instead of recomputed."""
        xyz, elem = self._prep(xyz, elem)
        oc = self.set_box(box) if box is not None else self.oc
        gH = gH.to(device=self.device, dtype=torch.float64).contiguous()
        F = xyz.shape[0]
        grad = torch.empty_like(xyz)
        key = (xyz.data_ptr(), xyz._version, elem.data_ptr(), elem._version,
               tuple(oc.tolist()), F)
        reuse = (generation is not None and self._last is not None and self._last[1] == generation
                 and self._last[0] == key)
        self._last = None
        for f0 in range(0, F, self.max_frames):
            chunk = xyz[f0:f0 + self.max_frames]
            _lib.check(self.lib.saxs_backward_ex(
                self._plan, _ptr(chunk), _ptr(elem), int(chunk.shape[1]), int(chunk.shape[0]),
                oc.ctypes.data_as(C.c_void_p), _ptr(gH[f0:]), _ptr(grad[f0:]),
                _lib.BWD_REUSE_FORWARD if reuse else 0, _stream_handle()), "saxs_backward")
        return grad

    @property
    def generation(self) -> int:
        return self._generation

    def reset(self):
        self.H.zero_()
        self.C.zero_()
        self.frames_done = 0

    def kept(self):
        kx, ky, kz = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(self.lib.saxs_plan_kept(self._plan, C.byref(kx), C.byref(ky), C.byref(kz)),
                   "saxs_plan_kept")
        return kx.value, ky.value, kz.value

    def set_timing(self, on: bool):
        _lib.check(self.lib.saxs_plan_set_timing(self._plan, 1 if on else 0), "saxs_plan_set_timing")
        self._timing = bool(on)

    def kernel_times(self):
        """{kernel name: (total ms, launches)} since the last call."""
        ms = np.zeros(len(_lib.KERNEL_NAMES), dtype=np.float64)
        n = np.zeros(len(_lib.KERNEL_NAMES), dtype=np.int64)
        _lib.check(self.lib.saxs_plan_kernel_times(self._plan, ms.ctypes.data_as(C.c_void_p),
                                                   n.ctypes.data_as(C.c_void_p)),
                   "saxs_plan_kernel_times")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(_lib.KERNEL_NAMES) if n[i]}

    def workspace_bytes(self) -> int:
        return int(self.lib.saxs_plan_workspace_bytes(self._plan))

    def rows(self):
        """Output rows (b*dq, H_b/C_b) for b >= 1 with C_b != 0
        (saxsKernel.cu:210-225); q is the float32 product b*dq."""
        H = self.H.cpu().numpy()
        Cn = self.C.cpu().numpy()
        dq = np.float32(self.hist.dq)
        return [(float(np.float32(np.float32(b) * dq)), float(H[b] / Cn[b]))
                for b in range(1, len(H)) if Cn[b] != 0.0]

    def counts(self, box) -> np.ndarray:
        """C_b of one binning of the half spectrum with the q tables of `box`
        (host arithmetic on the same tables the kernels read)."""
        cell = inverse_box_matrix(box)
        bins, _, kmax = self._q_tables(self._hist_for(cell), cell)
        return half_spectrum_counts(self.spec.nn, bins, kmax, self.hist.num_bins,
                                    signed=is_triclinic(cell))

    def nvt_rows(self, box0, frames: int):
        """NVT output rows (saxsKernel.cu:17-39,184-189 and :210-225): the
        spectrum summed over `frames` frames, binned by box0's q (every frame
        was run with bin_box=box0), times 1/frames, over the counts of one
        binning with box0."""
        H = self.H.cpu().numpy() / float(max(frames, 1))
        Cn = self.counts(box0)
        dq = np.float32(self.hist.dq)
        return [(float(np.float32(np.float32(b) * dq)), float(H[b] / Cn[b]))
                for b in range(1, len(H)) if Cn[b] != 0.0]

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            self.lib.saxs_plan_destroy(self._plan)
            self._plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
