#!/usr/bin/env python3
"""Throughput benchmark of the MI355X SAXS engine.

Workload (the trajectory configuration the headline metric is quoted on):
100,000 atoms uniform in a 100 A cubic box, element mix H 50 % / C 30 % /
O 15 % / N 4 % / S 1 %, 100 frames = seed structure + i.i.d. N(0, 0.5^2 A^2)
displacements, order-4 B-splines, 128^3 cell grid, sigma = 2 (256^3 FFT),
q to 0.6 A^-1 with requested bins of 0.005 A^-1 (reset to the grid spacing
as the reference does), float32.  Inputs are generated with numpy seeds
(base seed 2, frame stream default_rng([2, frame])) and moved to HBM before
the timed region.

A step is the forward pass of all 100 frames (batched launches of
--batch frames).  value = frames/s over --steps steps.  The roofline entry
uses per-kernel hipEvent times recorded by the engine in a second pass over
the same steps; the CPU baseline runs the FP64 oracle on a bounded sample.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pycusaxs_amd import workloads as W  # noqa: E402  (numpy-only module)

N_ATOMS = W.TRAJ["n_atoms"]
BOX = W.TRAJ["box"]
GRID = W.TRAJ["grid"]
SIGMA = W.TRAJ["sigma"]
DQ = W.TRAJ["dq"]
QCUT = W.TRAJ["qcut"]
ORDER = W.TRAJ["order"]
FRAMES = W.TRAJ["frames"]
SEED = W.TRAJ["seed"]
HBM_PEAK_GBS = 8000.0


# HBM bytes per launch of each kernel for the default workload (100 frames
# per launch), from rocprofv3 counters on MI355X: reads = 128 B x
# TCC_EA0_RDREQ_128B + 64 B x TCC_EA0_RDREQ_64B, writes = WRITE_SIZE; the same
# counters give a 512 MiB elementwise copy exactly.  Commands and the
# per-kernel breakdown: profiles/hbm_traffic_trajectory_100k_128.md.
MEASURED_TRAFFIC = {"prep": 229.7e6, "scatter": 440.2e6, "spread_z": 1483.2e6, "yfft": 1710.1e6,
                    "xfft_epi": 401.0e6, "adj_x": 439.9e6, "adj_y": 1829.0e6, "adj_z": 1794.6e6,
                    "grad_sum": 773.1e6}

# Wave64 VALU instructions per launch of each kernel for the default workload,
# from rocprofv3 SQ_INSTS_VALU on MI355X (commands and the per-kernel table:
# profiles/sq_counters_trajectory_100k_128.md).  A CU issues at most one wave64
# VALU instruction per clock, so the chip's issue peak is 256 x 2.4e9 per second.
MEASURED_VALU = {"prep": 35.2e6, "scatter": 13.3e6, "spread_z": 739.9e6, "yfft": 145.8e6,
                 "xfft_epi": 41.1e6, "adj_x": 42.7e6, "adj_y": 162.5e6, "adj_z": 790.1e6,
                 "grad_sum": 57.3e6}
VALU_ISSUE_PEAK = 256 * 2.4e9          # wave64 VALU instructions per second


def make_inputs(n_atoms=N_ATOMS, frames=FRAMES):
    return W.trajectory(n_atoms, frames)


def leg_batched_uneven(torch, steps):
    from pycusaxs_amd.batch import BatchedSystems, SystemInput
    from pycusaxs_amd.engine import SystemSpec, resolve_grid
    from pycusaxs_amd.tuning import Autotuner
    work = W.uneven_systems()
    tuner = Autotuner(frames=2, repeats=1)
    systems = []
    for w in work:
        n3, nn = resolve_grid((w["n"],) * 3, SIGMA)
        systems.append(SystemInput(SystemSpec(n3, nn, ORDER, tuple(w["symbols"]), SIGMA, DQ, 0.5),
                                   np.diag([w["L"]] * 3), w["n_atoms"], w["xyz"].shape[0]))
    t0 = time.perf_counter()
    bs = BatchedSystems(systems, tuner=tuner)
    tune_s = time.perf_counter() - t0
    for i, w in enumerate(work):
        bs.load(i, w["xyz"], w["ids"])
    bs.capture()
    bs.forward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bs.forward()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    frames = bs.frames_total()
    atoms = sum(w["n_atoms"] for w in work)
    grids = sorted(set(w["n"] for w in work))
    bs.close()
    return {"frames_per_s": round(frames / dt, 2), "ms_per_graph": round(dt * 1e3, 3),
            "systems": len(work), "frames_per_graph": frames, "atoms_total": atoms,
            "grids": grids, "setup_incl_autotune_s": round(tune_s, 1),
            "workload": "32 systems, 10k-300k atoms at 0.1/A^3, grids 64-192 (sigma 2), "
                        "8 frames each, one hipGraph replay"}


def leg_large(torch, steps):
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    n_atoms, L, n, F = W.LARGE["n_atoms"], W.LARGE["box"], W.LARGE["grid"], W.LARGE["frames"]
    base, symbols, ids = W.large_base()
    xyz = torch.empty((F, n_atoms, 3), dtype=torch.float32, device="cuda")
    for f in range(F):
        xyz[f].copy_(torch.from_numpy(W.large_frame(base, f)))
    idt = torch.from_numpy(ids).cuda()
    n3, nn = resolve_grid((n,) * 3, SIGMA)
    eng = SaxsEngine(SystemSpec(n3, nn, ORDER, tuple(symbols), SIGMA, 0.0, 0.5),
                     np.diag([L] * 3), max_atoms=n_atoms, max_frames=F, tile=(4, 4))
    eng.forward(xyz, idt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.forward(xyz, idt)
    torch.cuda.synchronize()
    fwd = steps * F / (time.perf_counter() - t0)
    Hf, _ = eng.forward_frames(xyz, idt)
    gH = (torch.randn(Hf.shape, dtype=torch.float64, device="cuda") / Hf.abs().clamp_min(1.0))
    eng.backward(xyz, idt, gH)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.forward_frames(xyz, idt)
        eng.backward(xyz, idt, gH, generation=eng.generation)
    torch.cuda.synchronize()
    fb = steps * F / (time.perf_counter() - t0)
    out = {"fwd_frames_per_s": round(fwd, 2), "fwd_bwd_frames_per_s": round(fb, 2),
           "fwd_bwd_time_over_fwd": round(fwd / fb, 3), "kept": list(eng.kept()),
           "workspace_GiB": round(eng.workspace_bytes() / 2**30, 2),
           "workload": "1,000,000 atoms, 215 A box, 9 elements, 256^3 -> 512^3 FFT, "
                       "q <= 0.5, 16 frames per launch"}
    eng.close()
    return out


def leg_restraint(torch, steps, keep):
    from pycusaxs_amd.autograd import chi2, saxs_profile
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    w = W.restraint()
    n3, nn = resolve_grid((w["n"],) * 3, SIGMA)
    box = np.diag([w["L"]] * 3)
    eng = SaxsEngine(SystemSpec(n3, nn, w["order"], tuple(w["symbols"]), SIGMA, w["dq"],
                                w["qcut"], "exact"), box, len(w["base"]), 1)
    idt = torch.from_numpy(w["ids"]).cuda()
    with torch.no_grad():
        T, C = saxs_profile(torch.from_numpy(w["target"]).cuda()[None], eng, idt)
    sel = (C[0] > 0) & (torch.arange(C.shape[1], device=C.device) >= 1)
    s = 0.01 * T[0].abs() + 1e-30
    x0 = torch.from_numpy(w["base"]).cuda()[None]

    def step():
        x = x0.clone().requires_grad_(True)
        I, _ = saxs_profile(x, eng, idt)
        loss = chi2(I[0], T[0], s, sel)
        loss.backward()
        return loss, x.grad

    step()
    torch.cuda.synchronize()
    times = []
    for _ in range(max(steps, 20)):        # one step is ~1 ms: median of synchronized steps
        t0 = time.perf_counter()
        loss, g = step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ms = float(np.median(times)) * 1e3
    keep.update(T=T[0].cpu().numpy(), s=s.cpu().numpy(), sel=sel.cpu().numpy(),
                grad=g[0].cpu().numpy(), loss=loss.item(), w=w)
    # the same step captured once as a graph and replayed (autograd.RestraintStep)
    from pycusaxs_amd.autograd import RestraintStep
    gstep = RestraintStep(eng, idt, x0, lambda I: chi2(I[0], T[0], s, sel))
    gstep(x0)
    torch.cuda.synchronize()
    gtimes = []
    for _ in range(max(steps, 20)):
        t0 = time.perf_counter()
        gstep(x0)
        torch.cuda.synchronize()
        gtimes.append(time.perf_counter() - t0)
    gms = float(np.median(gtimes)) * 1e3
    del gstep
    eng.close()
    return {"fwd_bwd_ms": round(ms, 3), "fwd_bwd_graph_ms": round(gms, 3), "chi2": loss.item(),
            "bins_used": int(sel.sum().item()),
            "workload": "50,000 atoms, 80 A, 128^3 -> 256^3, order 6, 100 requested bins "
                        "0.005-0.5, chi^2 vs displaced (1 A) structure, 1 % errors"}


def restraint_grad_check(keep, threads):
    """Engine chi^2 gradient vs the FP64 oracle gradient on the same target."""
    from oracle import saxs_oracle as O
    O.set_fft_workers(threads)
    w = keep["w"]
    box = np.diag([w["L"]] * 3)
    plan = O.make_plan((w["n"],) * 3, box, w["order"], SIGMA, None, w["dq"], w["qcut"], "exact")
    H, C = O.frame_histogram(plan, w["base"], w["ids"], w["symbols"], box, literal=False)
    I = np.where(C > 0, H / np.where(C > 0, C, 1), 0.0)
    r = np.where(keep["sel"], (I - keep["T"]) / keep["s"], 0.0)
    gH = np.where(C > 0, np.where(keep["sel"], 2 * r / keep["s"], 0.0) / np.where(C > 0, C, 1), 0.0)
    go = O.grad_linear_H(plan, w["base"], w["ids"], w["symbols"], box, gH)
    return float(np.linalg.norm(keep["grad"] - go) / np.linalg.norm(go))


class _MemoryTopology:
    """The duck-typed topology callbacks pycusaxs_cuda.run uses
    (RunSaxs.cu:109,124-125,137-138) over frames already in host memory, so
    the drop-in path is timed without trajectory decoding.  With `boxes`,
    frame i carries boxes[i] (a box that changes every frame, as in NPT)."""

    def __init__(self, frames, symbols, box, boxes=None):
        self.frames, self.box, self.ts = frames, np.asarray(box, dtype=np.float32), None
        self.boxes = boxes
        self.index = {}
        for i, s in enumerate(symbols):
            self.index.setdefault(s, []).append(i)

    def get_atom_index(self):
        return self.index

    def read_frame(self, i):
        self.ts = i

    def get_box(self):
        return self.box if self.boxes is None else self.boxes[self.ts or 0]

    def iter_frames_stream(self, start, stop, step=1):
        for i in range(start, min(stop, len(self.frames)), step):
            yield {"frame": i, "time": float(i), "positions": self.frames[i],
                   "box": self.box if self.boxes is None else self.boxes[i]}


def changing_boxes(kind, F, L=BOX):
    """A box that changes every frame (NPT-like fluctuations of ~0.5 %):
    cubic, orthorhombic (each axis its own), or triclinic (tilted b and c)."""
    out = np.empty((F, 3, 3), dtype=np.float32)
    for f in range(F):
        a = 1.0 + 0.005 * np.sin(0.37 * f + np.array([0.0, 1.1, 2.3]))
        if kind == "cubic":
            out[f] = np.diag([L * a[0]] * 3)
        elif kind == "orthorhombic":
            out[f] = np.diag(L * a)
        else:
            out[f] = [[L * a[0], 0.0, 0.0], [0.03 * L * a[1], L * a[1], 0.0],
                      [-0.02 * L * a[2], 0.025 * L * a[2], L * a[2]]]
    return out


def leg_run_path(torch, xyz_h, symbols, ids_h, engine_fps, box_kind=None, simulation="",
                 defaults=False):
    """pycusaxs_cuda.run end to end on the trajectory workload: host frames
    (the 100 frames cycled to 1,000) staged through pinned double buffers and
    a copy stream, element gather on the device, batches of 16 frames.
    Steady-state rate from runs of 200 and 1,000 frames, the best of three
    of each (the difference removes plan setup and output).  With box_kind,
    the box changes every frame (changing_boxes)."""
    import tempfile
    import pycusaxs_cuda
    syms = [symbols[i] for i in ids_h]
    F = 10 * len(xyz_h)                 # the host frames cycled, so setup is a small share
    boxes = changing_boxes(box_kind, F) if box_kind else None
    topo = _MemoryTopology([xyz_h[i % len(xyz_h)] for i in range(F)], syms, np.diag([BOX] * 3),
                           boxes)
    out = os.path.join(tempfile.mkdtemp(), "saxs.dat")

    def run(end):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if defaults:      # the reference CLI's defaults (saxs_defaults.py:17-31)
            pycusaxs_cuda.run(topo, "memory", "memory", [GRID], end=end, output=out,
                              scale_factor=2.5, bin_size=0.01, qcut=2.0, quiet=True)
        else:
            pycusaxs_cuda.run(topo, "memory", "memory", [GRID], end=end, output=out,
                              scale_factor=SIGMA, bin_size=DQ, qcut=QCUT, quiet=True,
                              simulation=simulation)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    run(F - 1)                                              # warm-up
    short = F // 5
    # best of three of each: one pair of runs of a few tens of milliseconds
    # swung by 2x with the host's timing noise
    t_short = min(run(short - 1) for _ in range(3))
    t_full = min(run(F - 1) for _ in range(3))
    steady = (F - short) / max(t_full - t_short, 1e-9)
    return {"frames_per_s": round(steady, 1), "frames": F, "wall_s_full": round(t_full, 3),
            "engine_only_frames_per_s": round(engine_fps, 1),
            "fraction_of_engine_rate": round(steady / engine_fps, 3),
            "workload": "pycusaxs_cuda.run on the trajectory workload from host memory "
                        "(pinned double buffer, copy stream, 16 frames per launch)"
                        + (f", {box_kind} box changing every frame ({simulation or 'npt'})"
                           if box_kind else "")}


def _trr_bytes(x_nm, box_nm, step, t):
    """One float32 TRR frame (the layout include/saxs_xdr.h reads)."""
    n = len(x_nm)
    head = np.array([1993, 13, 12], dtype=">i4").tobytes() + b"GMX_trn_file"
    sizes = np.array([0, 0, 36, 0, 0, 0, 0, 12 * n, 0, 0, n, step, 0], dtype=">i4").tobytes()
    return (head + sizes + np.array([t, 0.0], dtype=">f4").tobytes()
            + np.asarray(box_nm, dtype=">f4").tobytes() + np.asarray(x_nm, dtype=">f4").tobytes())


def leg_run_file(torch, xyz_h, symbols, ids_h, engine_fps, fmt):
    """pycusaxs_cuda.run fed by the MDAnalysis-free reader: the trajectory
    workload written to a TRR or XTC file (the 100 host frames cycled to
    500; XTC: four frames packed by the test suite's encoder, cycled, since
    packing in Python is slow; decoding costs the same per frame), read
    through trajectory.XdrTopology, whose batches are decoded on host
    threads straight into the pinned staging buffers.  Steady-state rate
    from runs of 100 and 500 frames, and the reader's decode rate alone."""
    import tempfile
    import pycusaxs_cuda
    from pycusaxs_amd.trajectory import XdrTopology, default_threads
    syms = [symbols[i] for i in ids_h]
    F = 5 * len(xyz_h)
    box_nm = np.diag([BOX / 10.0] * 3)
    path = os.path.join(tempfile.mkdtemp(), "traj." + fmt)
    with open(path, "wb") as fh:
        if fmt == "trr":
            for f in range(F):
                fh.write(_trr_bytes(xyz_h[f % len(xyz_h)] / 10.0, box_nm, f, 0.5 * f))
        else:
            from tests import xdr_files as X
            rng = np.random.default_rng(0)
            blobs = [X.xtc_frame(X.quantize(xyz_h[f] / 10.0, 1000.0), box_nm, f, 0.5 * f, rng=rng)
                     for f in range(4)]
            for f in range(F):
                fh.write(blobs[f % 4])
    topo = XdrTopology(None, path, elements=syms)
    buf = np.empty((len(xyz_h), topo.n_atoms, 3), dtype=np.float32)
    box = np.empty((len(xyz_h), 3, 3), dtype=np.float32)
    topo.stage_frames(range(len(xyz_h)), buf, box)         # page the file in
    t0 = time.perf_counter()
    topo.stage_frames(range(len(xyz_h)), buf, box)
    decode = len(xyz_h) / (time.perf_counter() - t0)
    out = os.path.join(tempfile.mkdtemp(), "saxs.dat")

    def run(end):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pycusaxs_cuda.run(topo, "memory", path, [GRID], end=end, output=out,
                          scale_factor=SIGMA, bin_size=DQ, qcut=QCUT, quiet=True)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    run(F - 1)                                              # warm-up
    short = F // 5
    # best of three of each: one pair of runs of a few tens of milliseconds
    # swung by 2x with the host's timing noise
    t_short = min(run(short - 1) for _ in range(3))
    t_full = min(run(F - 1) for _ in range(3))
    steady = (F - short) / max(t_full - t_short, 1e-9)
    size = os.path.getsize(path)
    topo.close()
    os.remove(path)
    return {"frames_per_s": round(steady, 1), "frames": F, "decode_frames_per_s": round(decode, 1),
            "decode_threads": default_threads(), "file_bytes_per_frame": size // F,
            "engine_only_frames_per_s": round(engine_fps, 1),
            "workload": f"pycusaxs_cuda.run on the trajectory workload from a {fmt.upper()} file "
                        "(trajectory.XdrTopology: batches decoded on host threads into the "
                        "pinned staging buffers, 16 frames per launch)"}


def leg_reference_defaults(torch, xyz_h, symbols, ids_h, steps):
    """The drop-in at the reference CLI's defaults (saxs_defaults.py:17-31:
    grid 128, scale 2.5 -> 320^3 FFT, bin 0.01, qcut 2.0, order 4) on the
    trajectory workload: the engine on HBM-resident frames (50 per launch,
    with per-kernel times) and pycusaxs_cuda.run from host memory."""
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    F = 50
    xyz = torch.from_numpy(xyz_h[:F]).cuda()
    ids = torch.from_numpy(ids_h).cuda()
    n3, nn = resolve_grid((GRID,) * 3, 2.5)
    eng = SaxsEngine(SystemSpec(n3, nn, ORDER, tuple(symbols), 2.5, 0.01, 2.0),
                     np.diag([BOX] * 3), max_atoms=N_ATOMS, max_frames=F)
    eng.forward(xyz, ids)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.forward(xyz, ids)
    torch.cuda.synchronize()
    fps = steps * F / (time.perf_counter() - t0)
    eng.set_timing(True)
    eng.forward(xyz, ids)
    torch.cuda.synchronize()
    kt = {k: round(v[0] / v[1], 4) for k, v in eng.kernel_times().items()}
    eng.set_timing(False)
    # forward + backward (the backward reusing the forward's spectra), as a
    # restraint at these defaults would run it
    Hf, _ = eng.forward_frames(xyz, ids)
    gH = torch.ones_like(Hf) / Hf.abs().clamp_min(1.0)
    eng.backward(xyz, ids, gH, generation=eng.generation)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.forward_frames(xyz, ids)
        eng.backward(xyz, ids, gH, generation=eng.generation)
    torch.cuda.synchronize()
    fb = steps * F / (time.perf_counter() - t0)
    out = {"engine_frames_per_s": round(fps, 1), "fwd_bwd_frames_per_s": round(fb, 1),
           "kept": list(eng.kept()), "fft_grid": nn[0],
           "kernel_ms_per_50_frames": kt,
           "workload": "reference CLI defaults: 100,000 atoms, 100 A, grid 128, scale 2.5 "
                       "(320^3 FFT), bin 0.01 (reset to 0.02513), qcut 2.0, 80 bins"}
    eng.close()
    out["run_path_frames_per_s"] = leg_run_path(torch, xyz_h, symbols, ids_h, fps,
                                                defaults=True)["frames_per_s"]
    return out


def kernel_bytes(eng, frames_per_launch, n_atoms):
    """Algorithmic bytes per launch of each kernel: the pruned byte model
    B_fwd = 16N + 8E n^3 + 16E n^2 (M+1) + 16E n (2M+1)(M+1) per frame and its
    backward B_bwd = 24E n (2M+1)(M+1) + 16E n^2 (M+1) + 8E n^3 + 28N (each
    array that must round-trip memory between the stages counted once read and
    once written; float32 grids, complex64 spectra), split over the stages each
    kernel implements.  The spreading kernels carry the density-grid term
    8E n^3 although the fused implementation keeps that grid on chip: the
    measured HBM bytes are reported beside it.  The atom sort (prep, scan,
    scatter) is not part of the model (None)."""
    kx, ky, kz = eng.kept()
    n = eng.spec.n
    E = len(eng.spec.symbols)
    F = frames_per_launch
    grid = 8.0 * E * n[0] * n[1] * n[2]      # density grids: spread writes, z pass reads (f32)
    zb = 8.0 * E * n[0] * n[1] * kz          # z-pass output (complex64), per direction
    yb = 8.0 * E * n[0] * ky * kz            # y-pass output (complex64), per direction
    return {
        "spread_z": F * (16.0 * n_atoms + grid + zb),     # atoms in, grid, z output out
        "yfft": F * (zb + yb),                            # z output in, y output out
        "xfft_epi": F * yb,                               # y output in (+ the bins)
        # the forward-reusing backward's pruned x adjoint reads the amplitudes
        # saved by the forward at the kept points instead of the y tiles
        "adj_x": F * (8.0 * kx * ky * kz + yb),           # amplitudes in, adjoint x out
        "adj_y": F * (yb + zb),                           # adjoint x in, adjoint y out
        "adj_z": F * (zb + grid + 16.0 * n_atoms),        # adjoint y in, potential grid, atoms in
        "grad_sum": F * 12.0 * n_atoms,                   # gradient out
    }


def fixed_grad_H(torch, eng, xyz, ids, dev):
    """A fixed, seeded dL/dH for the backward legs (scaled by 1/|H|)."""
    gen = torch.Generator(device="cpu").manual_seed(SEED)
    Hf, _ = eng.forward_frames(xyz, ids)
    return (torch.randn(Hf.shape, generator=gen, dtype=torch.float64)
            / Hf.abs().clamp_min(1.0).cpu()).to(dev)


def host_cpu():
    """CPU model of this host, its logical CPU count and the CPUs this process
    may run on."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def cpu_baseline(xyz, symbols, ids, frames, threads):
    from oracle import saxs_oracle as O
    used = O.set_fft_workers(threads)
    box = np.diag([BOX] * 3)
    plan = O.make_plan((GRID,) * 3, box, ORDER, SIGMA, None, DQ, QCUT)
    t0 = time.perf_counter()
    first = None
    for f in range(frames):
        h, c = O.frame_histogram(plan, xyz[f], ids, symbols, box, literal=True)
        if first is None:
            first = (h, c)
    dt = time.perf_counter() - t0
    return frames / dt, used, dt, first[0], first[1]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=100, help="frames per engine launch chain")
    ap.add_argument("--frames", type=int, default=FRAMES)
    ap.add_argument("--cpu-frames", type=int, default=10)
    ap.add_argument("--plain-frames", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = every CPU this process may run on (its affinity set)")
    ap.add_argument("--cpu-frames-1t", type=int, default=2,
                    help="frames of the single-thread CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and 1M-atom legs")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-frame plain leg")
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup + steps only (for rocprofv3), print the line without "
                         "roofline/cpu legs")
    ap.add_argument("--profile-backward", action="store_true",
                    help="with --profile-only: also run forward_frames + backward per step")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="with --profile-only: also run one elementwise copy of this many MiB "
                         "(known HBM bytes, to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--no-graph", action="store_true",
                    help="time the forward as eager launches only (no hipGraph replay)")
    ap.add_argument("--leg", choices=("batched_uneven", "large_1M_512", "restraint_50k_order6",
                                      "run_path_npt_cubic", "run_path_npt_orthorhombic",
                                      "run_path_npt_triclinic", "reference_defaults",
                                      "run_path_trr", "run_path_xtc"),
                    help="run only this secondary workload (for rocprofv3) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this engine's "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.leg:
        legs = {"batched_uneven": leg_batched_uneven, "large_1M_512": leg_large,
                "restraint_50k_order6": lambda t, n: leg_restraint(t, n, {})}
        legs["reference_defaults"] = lambda t, n: leg_reference_defaults(t, *make_inputs()[:3], n)
        for kind in ("cubic", "orthorhombic", "triclinic"):
            legs[f"run_path_npt_{kind}"] = (
                lambda t, n, k=kind: leg_run_path(t, *make_inputs()[:3], float("nan"), k))
        for fmt in ("trr", "xtc"):
            legs[f"run_path_{fmt}"] = (
                lambda t, n, k=fmt: leg_run_file(t, *make_inputs()[:3], float("nan"), k))
        print(json.dumps({"leg": args.leg, **legs[args.leg](torch, max(1, args.steps))}), flush=True)
        return
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid

    xyz_h, symbols, ids_h, _ = make_inputs(frames=args.frames)
    dev = torch.device("cuda", 0)
    xyz = torch.from_numpy(xyz_h).to(dev)
    ids = torch.from_numpy(ids_h).to(dev)
    n3, nn = resolve_grid((GRID,) * 3, SIGMA)
    spec = SystemSpec(n3, nn, ORDER, tuple(symbols), SIGMA, DQ, QCUT)
    batch = min(args.batch, args.frames)
    box = np.diag([BOX] * 3)
    eng = SaxsEngine(spec, box, max_atoms=N_ATOMS, max_frames=batch)
    print(f"# plan: n={n3} nn={nn} kept={eng.kept()} bins={eng.hist.num_bins} "
          f"workspace={eng.workspace_bytes() / 2**20:.1f} MiB", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        eng.forward(xyz, ids)
    torch.cuda.synchronize()
    frames_done = args.steps * args.frames
    # eager launch chain (one host call per step)
    eng.reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.forward(xyz, ids)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    eager_value = frames_done / elapsed
    print(f"# forward, eager launches: {eager_value:.1f} frames/s "
          f"({elapsed * 1e3 / args.steps:.3f} ms/step)", file=sys.stderr, flush=True)
    value = eager_value
    graph_used = False
    if not args.no_graph and not args.profile_only:
        # the same launch chain captured once as a hipGraph and replayed per
        # step: the same kernels on the same inputs without the host launch
        # gaps between them
        cap = torch.cuda.Stream()
        cap.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(cap):
            eng.forward(xyz, ids)
        torch.cuda.current_stream().wait_stream(cap)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=cap):
            eng.forward(xyz, ids)
        for _ in range(args.warmup):
            graph.replay()
        torch.cuda.synchronize()
        eng.reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            graph.replay()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        value = frames_done / elapsed
        graph_used = True
        del graph
        print(f"# forward, hipGraph replay: {value:.1f} frames/s "
              f"({elapsed * 1e3 / args.steps:.3f} ms/step)", file=sys.stderr, flush=True)

    line = {
        "metric": "forward frames/s, 100k atoms, 128^3 grid, sigma=2 (256^3 FFT), batched",
        "value": round(value, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "forward_launch": "hipGraph replay of the launch chain" if graph_used else "eager",
        "eager_frames_per_s": round(eager_value, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform atoms + Gaussian frame displacements)",
        "config": {"workload": "trajectory: 100,000 atoms, 100 A box, 5 elements, 100 frames, "
                               "order 4, 128^3 -> 256^3 FFT, q <= 0.6, dq 0.005 (reset)",
                   "frames_per_step": args.frames, "frames_per_launch": batch,
                   "n_atoms": N_ATOMS, "grid": GRID, "fft_grid": nn[0]},
    }
    if args.profile_only:
        if args.profile_backward:
            gH = fixed_grad_H(torch, eng, xyz, ids, dev)
            for _ in range(args.steps):
                eng.forward_frames(xyz, ids)
                eng.backward(xyz, ids, gH, generation=eng.generation)
        if args.copy_calibration_mib > 0:
            src = torch.ones(args.copy_calibration_mib * 2**18, dtype=torch.float32, device=dev)
            dst = torch.empty_like(src)
            torch.mul(src, 1.0, out=dst)        # reads and writes exactly 4 B per element
        torch.cuda.synchronize()
        print(json.dumps(line))
        return

    # forward + backward (per-frame histograms, then dL/dx for a fixed dL/dH)
    gH = fixed_grad_H(torch, eng, xyz, ids, dev)
    # the backward reuses the spectra of the forward it follows (autograd use)
    for _ in range(max(1, args.warmup)):
        eng.forward_frames(xyz, ids)
        eng.backward(xyz, ids, gH, generation=eng.generation)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.forward_frames(xyz, ids)
        eng.backward(xyz, ids, gH, generation=eng.generation)
    torch.cuda.synchronize()
    fb = frames_done / (time.perf_counter() - t0)
    line["fwd_bwd_eager_frames_per_s"] = round(fb, 2)
    if graph_used:
        # the same forward + backward chain as one replayed hipGraph (the
        # partial-gradient slots were allocated by the warm-up above)
        cap = torch.cuda.Stream()
        cap.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(cap):
            eng.forward_frames(xyz, ids)
            eng.backward(xyz, ids, gH, generation=eng.generation)
        torch.cuda.current_stream().wait_stream(cap)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=cap):
            eng.forward_frames(xyz, ids)
            eng.backward(xyz, ids, gH, generation=eng.generation)
        for _ in range(max(1, args.warmup)):
            graph.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            graph.replay()
        torch.cuda.synchronize()
        fb = frames_done / (time.perf_counter() - t0)
        del graph
    line["fwd_bwd_frames_per_s"] = round(fb, 2)
    line["fwd_bwd_time_over_fwd"] = round(value / fb, 3)
    print(f"# fwd+bwd: {fb:.1f} frames/s ({value / fb:.2f}x forward time)", file=sys.stderr,
          flush=True)

    # reference-structured per-frame path (per element, full-spectrum FFT)
    if not args.no_plain:
        n_plain = min(args.plain_frames, args.frames)
        eng.forward_plain(xyz[:1], ids)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.forward_plain(xyz[:n_plain], ids)
        torch.cuda.synchronize()
        pl_fps = n_plain / (time.perf_counter() - t0)
        line["plain_frames_per_s"] = round(pl_fps, 2)
        line["batched_over_plain"] = round(value / pl_fps, 2)
        print(f"# plain per-frame path: {pl_fps:.1f} frames/s; batched/plain = "
              f"{value / pl_fps:.1f}x", file=sys.stderr, flush=True)

    # batched uneven systems (one hipGraph) and the 1M-atom / 512^3 system
    if not args.no_extra:
        restraint = {}
        legs = (("batched_uneven", leg_batched_uneven), ("large_1M_512", leg_large),
                ("restraint_50k_order6", lambda t, n: leg_restraint(t, n, restraint)),
                ("run_path", lambda t, n: leg_run_path(t, xyz_h, symbols, ids_h, value)),
                ("run_path_npt_cubic", lambda t, n: leg_run_path(t, xyz_h, symbols, ids_h, value,
                                                                 "cubic")),
                ("run_path_npt_orthorhombic", lambda t, n: leg_run_path(
                    t, xyz_h, symbols, ids_h, value, "orthorhombic")),
                ("run_path_npt_triclinic", lambda t, n: leg_run_path(
                    t, xyz_h, symbols, ids_h, value, "triclinic")),
                ("run_path_nvt_triclinic", lambda t, n: leg_run_path(
                    t, xyz_h, symbols, ids_h, value, "triclinic", "nvt")),
                ("run_path_trr", lambda t, n: leg_run_file(t, xyz_h, symbols, ids_h, value, "trr")),
                ("reference_defaults", lambda t, n: leg_reference_defaults(
                    t, xyz_h, symbols, ids_h, n)))
        for name, fn in legs:
            try:
                line[name] = fn(torch, max(1, min(args.steps, 3)))
                print(f"# {name}: {line[name]}", file=sys.stderr, flush=True)
            except Exception as exc:                      # report, keep the main line
                line[name] = {"error": f"{type(exc).__name__}: {exc}"}
                print(f"# {name} failed: {exc}", file=sys.stderr, flush=True)

    # per-kernel hipEvent times over the same number of steps
    eng.set_timing(True)
    for _ in range(args.steps):
        eng.forward(xyz, ids)
    torch.cuda.synchronize()
    eng.forward_frames(xyz, ids)
    eng.backward(xyz, ids, gH, generation=eng.generation)
    torch.cuda.synchronize()
    times = eng.kernel_times()
    eng.set_timing(False)
    kb = kernel_bytes(eng, batch, N_ATOMS)
    per_kernel = {}
    for k, (ms, cnt) in times.items():
        avg = ms / cnt
        gbs = kb[k] / (avg * 1e-3) / 1e9 if k in kb else None
        per_kernel[k] = {"avg_ms": round(avg, 5), "launches": cnt,
                         "alg_bytes": kb.get(k), "GBps": round(gbs, 1) if gbs else None}
    dom = max((k for k in per_kernel if k in kb), key=lambda k: times[k][0])
    d = per_kernel[dom]
    default_workload = args.frames == FRAMES and batch == FRAMES
    if default_workload:
        # counter-based fractions: VALU issue (instructions per launch from
        # rocprofv3 over the live launch time at the peak issue rate) and
        # measured HBM bytes over the live launch time
        for k, v in per_kernel.items():
            avg_s = v["avg_ms"] * 1e-3
            if k in MEASURED_VALU:
                v["valu_issue_frac"] = round(MEASURED_VALU[k] / avg_s / VALU_ISSUE_PEAK, 4)
            if k in MEASURED_TRAFFIC:
                v["hbm_traffic_frac"] = round(MEASURED_TRAFFIC[k] / avg_s / 1e9 / HBM_PEAK_GBS, 4)
    valu_frac = per_kernel[dom].get("valu_issue_frac")
    hbm_frac = per_kernel[dom].get("hbm_traffic_frac")
    bound = "hbm"
    if valu_frac is not None and hbm_frac is not None and valu_frac > hbm_frac:
        bound = "valu"
    # achieved / frac: the byte model (algorithmic bytes over the live launch
    # time); valu_issue_frac and hbm_traffic_frac: what the counters say of
    # the same launches; bound: whichever counter-based fraction is larger
    line["roofline"] = {"bound": bound, "kernel": dom, "achieved": d["GBps"],
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(d["GBps"] / HBM_PEAK_GBS, 4),
                        "traffic": MEASURED_TRAFFIC.get(dom) if default_workload else None,
                        "valu_issue_frac": valu_frac, "hbm_traffic_frac": hbm_frac,
                        "frac_model": "pruned byte model (density grids counted although they "
                                      "stay in LDS)"}
    line["kernels"] = per_kernel
    print("# kernels: " + json.dumps(per_kernel), file=sys.stderr, flush=True)

    # parity of frame 0 against the oracle + CPU baseline
    if not args.no_cpu:
        model, ncpu, usable = host_cpu()
        threads = args.cpu_threads or usable
        Hf, Cf = eng.forward_frames(xyz[:1], ids)
        torch.cuda.synchronize()
        cps, used, dt, Ho, Co = cpu_baseline(xyz_h, symbols, ids_h, args.cpu_frames, threads)
        cps1, _, dt1, _, _ = cpu_baseline(xyz_h, symbols, ids_h, args.cpu_frames_1t, 1)
        line["cpu_baseline"] = {"value": round(cps, 5), "unit": "frames/s", "cores": used,
                                "kind": "port",
                                "single_thread_value": round(cps1, 5),
                                "cpu_model": model, "os_cpu_count": ncpu,
                                "cpus_usable": usable,
                                "sample": f"FP64 oracle (literal reference pipeline, numpy "
                                          f"spreading + scipy FFT) on frames "
                                          f"0..{args.cpu_frames - 1} of the workload with "
                                          f"{used} FFT threads ({dt:.1f} s), and frames "
                                          f"0..{args.cpu_frames_1t - 1} with 1 thread "
                                          f"({dt1:.1f} s); host {model}, os.cpu_count() "
                                          f"{ncpu}, {usable} usable"}
        h0, c0 = Hf[0].cpu().numpy(), Cf[0].cpu().numpy()
        o_h, o_c = Ho, Co
        sel = (np.arange(len(c0)) >= 1) & (o_c > 0)
        rel = np.abs(h0[sel] / c0[sel] - o_h[sel] / o_c[sel]) / np.abs(o_h[sel] / o_c[sel])
        line["iq_max_rel_err"] = float(rel.max())
        line["counts_exact"] = bool(np.array_equal(c0, o_c))
        if not args.no_extra and "grad" in restraint:
            try:
                line["restraint_50k_order6"]["grad_rel_l2_vs_fp64"] = restraint_grad_check(
                    restraint, threads)
            except Exception as exc:
                line["restraint_50k_order6"]["grad_check_error"] = f"{type(exc).__name__}: {exc}"
    print(json.dumps(line))


if __name__ == "__main__":
    main()
