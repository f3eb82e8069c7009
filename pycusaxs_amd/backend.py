"""Drop-in implementation of the ``pycusaxs_cuda.run`` entry point on the
MI355X engine.

Mirrors, in order:
  argument validation and error mapping   pybind/cuda_bindings.cpp:11-168
  config, banner and summary string       cudaSAXS.cu:48-232
  frame loop                              Exec/RunSaxs.cu:94-271
      get_atom_index -> element groups (std::map order, RunSaxs.cu:108-115)
      read_frame(0) / get_box -> first box (RunSaxs.cu:124-127)
      iter_frames_stream(begin, end+1, stride) (RunSaxs.cu:137-138)
      per frame: "--> Frame: ..." line, forward (saxsKernel.cu:110-113)
      saxs.dat rows "%10.5f%12.5e" (RunSaxs.cu:217-240), timing banner
Frames are batched into one engine launch chain per batch and staged through
two pinned buffers and a copy stream (FrameStager): the host reads and
uploads batch k+1 while the GPU computes batch k, the reference's double
buffer (RunSaxs.cu:141-170).

Deliberate differences (the reference's behaviour there is a silent failure
or undefined): a malformed frame dict, an exception raised by the frame
iterator and an element with no form factor raise instead of ending the run
early / reading garbage; a triclinic box uses the true inverse of the cell
(Cell.cpp:111-126 drops its off-diagonal terms).
"""
from __future__ import annotations

import math
import os
import sys
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from .formfactors import coefficients

# Defaults of the reference's static Options (Exec/Options.cpp:3-19).
DEFAULT_OUTPUT = "saxs.dat"
DEFAULT_SIMULATION = "npt"
FRAMES_PER_BATCH = int(os.environ.get("PYCUSAXS_AMD_BATCH", "16"))


def _validate(grid, scaled_grid, begin, end, stride, order, scale_factor, bin_size, qcut):
    """cuda_bindings.cpp:32-122: std::invalid_argument -> ValueError."""
    grid = list(grid)
    if len(grid) == 0:
        raise ValueError("Grid must have at least one dimension")
    if len(grid) == 1:
        if grid[0] <= 0:
            raise ValueError("Grid dimension must be positive")
        grid = [int(grid[0])] * 3
    elif len(grid) == 3:
        if grid[0] <= 0 or grid[1] <= 0 or grid[2] <= 0:
            raise ValueError("Grid dimensions must be positive")
        grid = [int(v) for v in grid]
    else:
        raise ValueError("Grid must have either 1 or 3 dimensions (nx, ny, nz)")
    scaled = list(scaled_grid or [])
    if scaled:
        if len(scaled) == 1:
            if scaled[0] <= 0:
                raise ValueError("Scaled grid dimension must be positive")
            scaled = [int(scaled[0])] * 3
        elif len(scaled) == 3:
            if scaled[0] <= 0 or scaled[1] <= 0 or scaled[2] <= 0:
                raise ValueError("Scaled grid dimensions must be positive")
            scaled = [int(v) for v in scaled]
        else:
            raise ValueError("Scaled grid must have either 1 or 3 dimensions")
    if stride <= 0:
        raise ValueError("Stride must be positive")
    if order < 1 or order > 8:
        raise ValueError("Spline order must be between 1 and 8")
    if bin_size < 0.0:
        raise ValueError("Bin size must be non-negative")
    if qcut < 0.0:
        raise ValueError("Q cutoff must be non-negative")
    if scale_factor < 0.0:
        raise ValueError("Scale factor must be non-negative")
    if not scaled and scale_factor == 0.0:
        raise ValueError("Scale factor must be positive when scaled grid is not provided")
    if begin < 0 or end < 0:
        raise ValueError("Frame indices must be non-negative")
    if begin > end:
        raise ValueError("Begin frame must be less than or equal to end frame")
    if order not in (4, 6):
        # orders other than 4 are undefined in the reference's device code
        # (Splines.h:16 fixes 4); this engine implements 4 and 6 exactly
        raise ValueError(f"Spline order {order} is not supported by the MI355X engine (4 or 6)")
    return grid, scaled


def summary_text(topology, trajectory, grid, scaled, begin, end, stride, order, sigma, dq, qcut,
                 water_model, ions: Dict[str, int], padding: str = "avg") -> str:
    """Configuration summary (cudaSAXS.cu:97-217), plain text."""
    bar = "═" * 64
    s = ["", "╔" + bar + "╗", "║                  CuSAXS Configuration                         ║",
         "╚" + bar + "╝", "", "Input Files", f"  Topology      : {topology}",
         f"  Trajectory    : {trajectory}", "", "Frame Selection", f"  Start Frame   : {begin}",
         f"  End Frame     : {end}", f"  Stride        : {stride}",
         f"  Total Frames  : {(end - begin) // stride + 1}", "", "Grid Parameters"]
    if scaled:
        s += [f"  Primary Grid  : ({grid[0]}, {grid[1]}, {grid[2]})",
              f"  Scaled Grid   : ({scaled[0]}, {scaled[1]}, {scaled[2]})"]
    else:
        s += [f"  Grid Size     : ({grid[0]}, {grid[1]}, {grid[2]})"]
    s += [f"  Spline Order  : {order}", "", "SAXS Parameters", f"  Scale Factor  : {sigma:.3f}",
          f"  Bin Size      : {dq:.4f}", f"  Q-cutoff      : {qcut:.3f}", "", "Solvent Model"]
    if water_model:
        s.append(f"  Water Model   : {water_model}")
    for ion, count in sorted(ions.items()):
        if count > 0:
            sym = ion + ("⁺" if ion in ("Na", "K") else "⁻" if ion == "Cl" else "")
            s.append(f"  {sym} atoms     : {count}")
    s.append("  Padding Mode  : " + ("average" if padding == "avg" else "given"))
    return "\n".join(s) + "\n\n" + "-" * 66 + "\n"


def format_rows(rows) -> str:
    """One line per bin: '%10.5f%12.5e' (RunSaxs.cu:224-227)."""
    return "".join("%10.5f%12.5e\n" % (q, v) for q, v in rows)


def element_groups(index_map) -> tuple:
    """(symbols in std::map order, atom indices, element id per gathered atom)."""
    if not isinstance(index_map, dict):
        raise ValueError("get_atom_index() must return a dict of element -> atom indices")
    symbols = sorted(str(k) for k in index_map.keys())
    for s in symbols:
        coefficients(s)                     # unknown symbols raise ValueError
    idx, ids = [], []
    for e, s in enumerate(symbols):
        key = s if s in index_map else next(k for k in index_map if str(k) == s)
        v = np.asarray(list(index_map[key]), dtype=np.int64)
        idx.append(v)
        ids.append(np.full(len(v), e, dtype=np.int32))
    idx = np.concatenate(idx) if idx else np.zeros(0, dtype=np.int64)
    ids = np.concatenate(ids) if ids else np.zeros(0, dtype=np.int32)
    return symbols, idx, ids


def _frame_fields(fr):
    if not isinstance(fr, dict):
        raise ValueError(f"frame record must be a dict, got {type(fr).__name__}")
    try:
        frame = int(fr["frame"])
        t = float(fr["time"])
        pos = np.asarray(fr["positions"], dtype=np.float32)
        box = np.asarray(fr["box"], dtype=np.float32).reshape(3, 3)
    except (KeyError, TypeError, ValueError) as exc:
        raise ValueError(f"malformed frame record: {exc}") from exc
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError(f"frame positions must be [N, 3], got {pos.shape}")
    return frame, t, pos, box


def run(obj_topology, topology: str, trajectory: str, grid: Sequence[int],
        scaled_grid: Sequence[int] = (), begin: int = 0, end: int = 0, stride: int = 1,
        output: str = "", order: int = 4, scale_factor: float = 1.0, bin_size: float = 0.0,
        qcut: float = 0.0, water_model: str = "", sodium: int = 0, chlorine: int = 0,
        simulation: str = "", *, dq_policy: str = "reference", quiet: bool = False,
        padding: str = "avg"):
    """pycusaxs_cuda.run (cuda_bindings.cpp:172-195) on the MI355X engine.

    Keyword-only extensions (defaults give the reference's behaviour):
    dq_policy "exact" keeps the requested bin width; padding "given" pads
    with explicit water of `water_model` plus the sodium/chlorine counts
    (padding::given, Options.h:7-11, AtomCounter.cpp:77-103), computed for
    the first frame's box, instead of the border-slab mean.

    Returns {"summary": str, "grid": [nx, ny, nz], "scaled_grid": [...]} where
    scaled_grid echoes the input ([0, 0, 0] when it was derived from
    scale_factor), as the reference does (cudaSAXS.cu:219-229)."""
    grid3, scaled3 = _validate(grid, scaled_grid, int(begin), int(end), int(stride), int(order),
                               float(scale_factor), float(bin_size), float(qcut))
    if not topology:
        raise ValueError("Topology path must not be empty.")
    if not trajectory:
        raise ValueError("Trajectory path must not be empty.")
    if padding not in ("avg", "given"):
        raise ValueError(f"padding must be 'avg' or 'given', got {padding!r}")
    if padding == "given":
        from .geometry import WATER_MODELS
        if water_model not in WATER_MODELS:
            raise ValueError(f"padding 'given' needs a water model ({', '.join(sorted(WATER_MODELS))}),"
                             f" got {water_model!r}")
    sigma = float(np.float32(scale_factor))
    out_path = output or DEFAULT_OUTPUT
    sim = simulation or DEFAULT_SIMULATION
    if sim not in ("npt", "nvt"):
        # the reference accepts any string and then bins nothing
        # (saxsKernel.cu:184-196), writing an empty saxs.dat
        raise ValueError(f"simulation must be 'npt' or 'nvt', got {simulation!r}")
    ions = {"Na": int(sodium), "Cl": int(chlorine)}
    summary = summary_text(topology, trajectory, grid3, scaled3, int(begin), int(end), int(stride),
                           int(order), sigma, float(bin_size), float(qcut), water_model, ions,
                           padding)
    log = (lambda *a, **k: None) if quiet else print
    log(summary, end="")
    result = {"summary": summary, "grid": list(grid3),
              "scaled_grid": list(scaled3) if scaled3 else [0, 0, 0]}
    _run_frames(obj_topology, grid3, scaled3, int(begin), int(end), int(stride), int(order), sigma,
                float(bin_size), float(qcut), out_path, sim, water_model, dq_policy, log,
                ions if padding == "given" else None)
    return result


def _run_frames(topo, grid3, scaled3, begin, end, stride, order, sigma, dq_req, qcut_req,
                out_path, sim, water_model, dq_policy, log, given_ions=None):
    import torch

    from .engine import SaxsEngine, SystemSpec, resolve_grid

    t_start = time.perf_counter()
    symbols, idx, ids = element_groups(topo.get_atom_index())
    topo.read_frame(0)
    box0 = np.asarray(topo.get_box(), dtype=np.float32).reshape(3, 3)
    n, nn = resolve_grid(grid3, sigma, scaled3 or None)
    spec = SystemSpec(n, nn, order, tuple(symbols), sigma, dq_req, qcut_req, dq_policy)
    n_atoms = len(idx)
    batch = max(1, FRAMES_PER_BATCH)
    pads = None
    if given_ions is not None:
        from .geometry import explicit_solvent_pads
        pads = explicit_solvent_pads(np.diag(box0), n, water_model, given_ions)
    eng = SaxsEngine(spec, box0, max_atoms=max(n_atoms, 1), max_frames=batch, pad_given=pads)
    h = eng.hist
    if h.qcut_reset:
        log("\n--------------------------------------------------")
        log(f"----- Qcut had to be reset to {h.qcut:.2f} from  {qcut_req:.2f} ----")
        log("--------------------------------------------------\n")
    if h.dq_reset and dq_policy == "reference":
        log("\n--------------------------------------------------")
        log(f"----- Dq had to be reset to {h.dq:.3f} from  {dq_req:.3f} ----")
        log("--------------------------------------------------\n")
    log(banner(n, nn, order, sigma, h.dq, h.qcut, water_model))
    dev = eng.device
    ids_t = torch.from_numpy(ids).to(dev)
    stager = FrameStager(n_atoms, batch, idx, dev)
    gpu_ms = 0.0
    n_done = 0

    # NVT bins the summed spectrum once with the first frame's box
    # (RunSaxs.cu:124-130,208-211): every frame is binned by box0's q
    bin_box = box0 if sim == "nvt" else None

    def flush():
        # RunSaxs.cu:195-199 logs a failing frame and goes on, so saxs.dat
        # silently lacks it; here a batch that cannot be processed ends the
        # run with the error (a box that needs a larger workspace does not
        # fail: the engine makes a plan for it)
        nonlocal gpu_ms
        t0 = time.perf_counter()
        try:
            stager.launch(eng, ids_t, bin_box)
        except (ValueError, RuntimeError) as exc:
            print(f"Error processing frames: {exc}", file=sys.stderr)
            eng.close()
            raise
        gpu_ms += (time.perf_counter() - t0) * 1e3

    stage = getattr(topo, "stage_frames", None)
    if callable(stage) and callable(getattr(topo, "frame_indices", None)):
        # a topology that decodes frames itself (trajectory.XdrTopology):
        # each batch is written straight into the free pinned buffer
        chosen = topo.frame_indices(begin, end + 1, stride)
        for b0 in range(0, len(chosen), batch):
            chunk = list(chosen[b0:b0 + batch])
            xyz_host, box_host = stager.reserve(len(chunk), int(topo.n_atoms))
            times = stage(chunk, xyz_host, box_host)
            for f, t in zip(chunk, times):
                log(f"--> Frame: {f:<7}  Time Step: {t:.2f} ps")
            flush()
            n_done += len(chunk)
    else:
        frames_iter = topo.iter_frames_stream(begin, end + 1, stride)
        for fr in frames_iter:
            frame, t, pos, box = _frame_fields(fr)
            log(f"--> Frame: {frame:<7}  Time Step: {t:.2f} ps")
            if stager.count >= batch:
                flush()
            stager.add(pos, box)
            n_done += 1
        if stager.count:
            flush()
    torch.cuda.synchronize(dev)
    rows = eng.nvt_rows(box0, eng.frames_done) if sim == "nvt" else eng.rows()
    with open(out_path, "w") as fh:
        fh.write(format_rows(rows))
    n_frames = len(range(begin, end + 1, stride))
    log(f"Done {n_frames} Steps")
    log(f"Results written to {out_path}")
    total_ms = (time.perf_counter() - t_start) * 1e3 / max(n_frames, 1)
    cuda_ms = gpu_ms / max(n_done, 1)
    log(timing_banner(cuda_ms, total_ms - cuda_ms, total_ms))
    eng.close()
    return rows


class FrameStager:
    """Double-buffered frame upload (the reference's double buffer,
    RunSaxs.cu:141-170, without its per-atom host copies): each frame is
    copied whole into one of two pinned buffers [batch, atoms, 3], its box
    kept beside it (a batch whose boxes differ runs with per-frame cells and
    q tables, saxs_forward_boxes); a batch is
    uploaded to its device buffer on a separate copy stream, and the compute
    stream waits for that copy only, gathers the atoms into element order
    (one index_select) and runs the engine.  So batch k+1 is read and uploaded
    while batch k computes.  Events guard the reuse of each host buffer (its
    copy has finished) and each device buffer (the kernels reading it have)."""

    # host threads of the per-frame copy into the pinned buffer (one frame of
    # 100,000 atoms is 1.2 MB: 37 us on one thread of the MI355X box's host)
    COPY_THREADS = 4

    def __init__(self, n_atoms: int, batch: int, idx: np.ndarray, device):
        import torch
        self.batch = batch
        self.n_atoms = n_atoms
        self.idx = np.ascontiguousarray(idx, dtype=np.int64)
        self.idx_t = torch.from_numpy(self.idx).to(device)
        self.device = device
        self.n_frame_atoms = None
        self.host = self.host_np = self.dev = None
        self.sel = torch.empty((batch, n_atoms, 3), dtype=torch.float32, device=device)
        self.copy_stream = torch.cuda.Stream(device=device)
        self.copied = [None, None]
        self.consumed = [None, None]
        self.slot = 0
        self.count = 0
        self.boxes = np.zeros((batch, 3, 3), dtype=np.float32)
        self.copier = None

    def _allocate(self, n_frame_atoms: int):
        import torch
        if len(self.idx) and int(self.idx.max()) >= n_frame_atoms:
            raise ValueError(f"frame has {n_frame_atoms} atoms, the topology indexes atom "
                             f"{int(self.idx.max())}")
        shape = (self.batch, n_frame_atoms, 3)
        self.host = [torch.empty(shape, dtype=torch.float32, pin_memory=True) for _ in range(2)]
        self.host_np = [h.numpy() for h in self.host]
        self.dev = [torch.empty(shape, dtype=torch.float32, device=self.device) for _ in range(2)]
        self.n_frame_atoms = n_frame_atoms
        if self.COPY_THREADS > 1 and n_frame_atoms * 12 >= 1 << 18:
            try:
                from .trajectory import HostCopier
                self.copier = HostCopier(self.COPY_THREADS)
            except (RuntimeError, OSError):           # reader library not built: numpy copy
                self.copier = None

    def reserve(self, k: int, n_frame_atoms: int):
        """The free pinned buffer's first k frames [k, atoms, 3] and their
        boxes [k, 3, 3], for a reader to fill in place; the batch then holds
        k frames."""
        if not 0 < k <= self.batch:
            raise ValueError(f"batch of {k} frames (at most {self.batch})")
        if self.count:
            raise RuntimeError("reserve() on a batch that already holds frames")
        if self.n_frame_atoms is None:
            self._allocate(n_frame_atoms)
        elif n_frame_atoms != self.n_frame_atoms:
            raise ValueError(f"frames have {n_frame_atoms} atoms, earlier frames "
                             f"{self.n_frame_atoms}")
        if self.copied[self.slot] is not None:
            self.copied[self.slot].synchronize()          # host buffer is free again
        self.count = k
        return self.host_np[self.slot][:k], self.boxes[:k]

    def add(self, pos: np.ndarray, box: np.ndarray):
        if self.n_frame_atoms is None:
            self._allocate(len(pos))
        elif len(pos) != self.n_frame_atoms:
            raise ValueError(f"frame has {len(pos)} atoms, earlier frames {self.n_frame_atoms}")
        if self.count == 0 and self.copied[self.slot] is not None:
            self.copied[self.slot].synchronize()          # host buffer is free again
        dst = self.host_np[self.slot][self.count]
        if (self.copier is not None and isinstance(pos, np.ndarray) and pos.dtype == np.float32
                and pos.flags.c_contiguous and pos.shape == dst.shape):
            self.copier.copy(dst, pos)
        else:
            np.copyto(dst, pos)
        self.boxes[self.count] = box
        self.count += 1

    def launch(self, eng, ids_t, bin_box=None):
        import torch
        k, s = self.count, self.slot
        self.count = 0
        self.slot ^= 1
        compute = torch.cuda.current_stream(self.device)
        with torch.cuda.stream(self.copy_stream):
            if self.consumed[s] is not None:
                self.copy_stream.wait_event(self.consumed[s])
            self.dev[s][:k].copy_(self.host[s][:k], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
        self.copied[s] = ev
        compute.wait_event(ev)
        try:
            xyz = self.sel[:k]
            if self.n_atoms:
                torch.index_select(self.dev[s][:k], 1, self.idx_t, out=xyz)
            boxes = self.boxes[:k]
            same = bool(np.all(boxes == boxes[0]))
            if same and (bin_box is None or np.array_equal(boxes[0], bin_box)):
                eng.forward(xyz, ids_t, box=boxes[0], bin_box=bin_box)
            else:
                # boxes that change within the batch (or, in NVT, differ from
                # the first frame's): one launch chain with per-frame cells
                eng.forward_boxes(xyz, ids_t, boxes.copy(), bin_box=bin_box)
        finally:
            done = torch.cuda.Event()
            done.record(compute)
            self.consumed[s] = done


def banner(n, nn, order, sigma, dq, qcut, water_model) -> str:
    """saxsKernel::writeBanner (saxsKernel.cu:527-544)."""
    return ("*************************************************\n"
            f"* {'Running CuSAXS':^45} *\n"
            f"* {'Cell Grid':<19} {n[0]:>4} * {n[1]:>4} * {n[2]:>4}        *\n"
            f"* {'Supercell Grid':<19} {nn[0]:>4} * {nn[1]:>4} * {nn[2]:>4}        *\n"
            f"* {'Order':<10} {order:>4}      {'Sigma':<10} {sigma:>4.3g}          *\n"
            f"* {'Bin Size':<10} {dq:>4.3f}     {'Q Cutoff ':<10}  {qcut:>3.1f}          *\n"
            f"* {'Water Model':<19} {water_model:<25} *\n"
            "*************************************************\n")


def timing_banner(cuda_ms, read_ms, total_ms) -> str:
    """RunSaxs.cu:253-263."""
    return ("\n=========================================================\n"
            "=                                                       =\n"
            "=                    CuSAXS Timing                      =\n"
            "=                                                       =\n"
            f"=           CUDA Time:     {cuda_ms:<10.2f} ms/per step       =\n"
            f"=           Read Time:     {read_ms:<10.2f} ms/per step       =\n"
            f"=           Total Time:    {total_ms:<10.2f} ms/per step       =\n"
            "=                                                       =\n"
            "=========================================================\n")
