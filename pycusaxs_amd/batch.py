"""Many independent systems (uneven atom counts, boxes and grids) computed
together: one engine plan per system, inputs in static device buffers, and
the launch chains of every system captured once into a single hipGraph that
is replayed per batch.  The systems are independent, so their chains are
captured on parallel branches (a few default-priority streams forked from
and joined to the capture stream): the small kernels of one system overlap
the others' work instead of running one after another.

The reference has no batching: it computes one frame of one system at a
time (RunSaxs.cu:154-203).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .engine import SaxsEngine, SystemSpec


@dataclass
class SystemInput:
    spec: SystemSpec
    box: np.ndarray            # 3x3, orthorhombic
    n_atoms: int
    frames: int


class BatchedSystems:
    """forward() replays one hipGraph covering every system's frames and
    leaves per-system (H, C) in histograms()."""

    def __init__(self, systems: Sequence[SystemInput], tuner=None, branches: int = 4):
        self.systems = list(systems)
        self.engines: List[SaxsEngine] = []
        self.xyz: List[torch.Tensor] = []
        self.elem: List[torch.Tensor] = []
        self.shapes = []
        for s in self.systems:
            # the graphs only run forward: tiles fitted to the forward kernels
            tile, flags = (0, 0), _lib.FLAG_FORWARD_ONLY
            if tuner is not None:
                tile, flags = tuner.tune(s.spec, s.n_atoms, s.box, forward_only=True)
            eng = SaxsEngine(s.spec, s.box, max_atoms=s.n_atoms, max_frames=s.frames,
                             tile=tile, flags=flags)
            self.engines.append(eng)
            self.shapes.append((tile, flags))
            self.xyz.append(torch.zeros((s.frames, s.n_atoms, 3), dtype=torch.float32,
                                        device=eng.device))
            self.elem.append(torch.zeros(s.n_atoms, dtype=torch.int32, device=eng.device))
        # every system's H and C as views of one buffer: one fill kernel per
        # replay clears them all
        sizes = [eng.hist.num_bins for eng in self.engines]
        dev = self.engines[0].device if self.engines else torch.device("cuda", 0)
        self._hist = torch.zeros(2 * sum(sizes), dtype=torch.float64, device=dev)
        o = 0
        for eng, nb in zip(self.engines, sizes):
            eng.H = self._hist[o:o + nb]
            eng.C = self._hist[o + nb:o + 2 * nb]
            o += 2 * nb
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self._stream = torch.cuda.Stream()
        # independent systems on parallel graph branches (default-priority
        # streams joined by events): longest first onto the least-loaded branch
        nbr = max(1, min(branches, len(self.engines)))
        self._branches = [torch.cuda.Stream() for _ in range(nbr)]
        load = [0.0] * nbr
        self._branch_of = [0] * len(self.engines)
        cost = [float(np.prod(s.spec.nn)) * s.frames + 50.0 * s.n_atoms * s.frames
                for s in self.systems]
        for i in sorted(range(len(cost)), key=lambda j: -cost[j]):
            b = min(range(nbr), key=lambda k: load[k])
            self._branch_of[i] = b
            load[b] += cost[i]

    def load(self, i: int, xyz, elem):
        """Copies system i's frames [frames, atoms, 3] and element ids into
        the static input buffers (the graph reads them in place)."""
        x = torch.as_tensor(xyz, dtype=torch.float32)
        e = torch.as_tensor(elem, dtype=torch.int32)
        if tuple(x.shape) != tuple(self.xyz[i].shape) or tuple(e.shape) != tuple(self.elem[i].shape):
            raise ValueError(f"system {i}: expected {tuple(self.xyz[i].shape)} positions")
        self.xyz[i].copy_(x)
        self.elem[i].copy_(e)

    def _enqueue(self):
        main = torch.cuda.current_stream()
        self._hist.zero_()
        for st in self._branches:
            st.wait_stream(main)
        for i, (eng, x, e) in enumerate(zip(self.engines, self.xyz, self.elem)):
            with torch.cuda.stream(self._branches[self._branch_of[i]]):
                eng.forward(x, e)
        for st in self._branches:
            main.wait_stream(st)

    def forward_eager(self):
        self._enqueue()

    def capture(self):
        """Captures every system's launch chain into one graph (inputs and
        accumulators are the static buffers above)."""
        torch.cuda.synchronize()
        self._stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self._stream):
            self._enqueue()                  # warm-up outside the graph
        torch.cuda.current_stream().wait_stream(self._stream)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self._stream):
            self._enqueue()
        torch.cuda.synchronize()
        self.graph = g
        return g

    def forward(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def histograms(self):
        return [(eng.H, eng.C) for eng in self.engines]

    def frames_total(self) -> int:
        return sum(s.frames for s in self.systems)

    def close(self):
        self.graph = None
        for eng in self.engines:
            eng.close()
