"""In-memory stand-in for pycusaxs.topology.Topology: the duck-typed
callbacks RunSaxs.cu:109,124-125,137-138 uses (get_atom_index, read_frame,
get_box, iter_frames_stream)."""
import numpy as np


class FakeTopology:
    def __init__(self, frames, symbols, box, dt=1.0):
        self.frames = [np.asarray(f, dtype=np.float32) for f in frames]
        self.symbols = list(symbols)
        self.box = np.asarray(box, dtype=np.float32).reshape(3, 3)
        self.boxes = None
        self.dt = dt
        self.ts = None
        self.requested = []

    def get_atom_index(self):
        idx = {}
        for i, s in enumerate(self.symbols):
            idx.setdefault(s, []).append(i)
        return idx

    def read_frame(self, i):
        if not 0 <= i < len(self.frames):
            raise IndexError(i)
        self.ts = i
        return i

    def get_box(self):
        if self.ts is None:
            raise RuntimeError("Call read_frame() before get_box()")
        return self.box_of(self.ts)

    def box_of(self, i):
        return self.boxes[i] if self.boxes is not None else self.box

    def iter_frames_stream(self, start, stop, step=1):
        self.requested.append((start, stop, step))
        for i in range(start, min(stop, len(self.frames)), step):
            yield {"frame": i, "time": i * self.dt, "positions": self.frames[i],
                   "box": self.box_of(i)}
