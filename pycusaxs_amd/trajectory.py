"""MDAnalysis-free frame ingestion: the reference's Topology callbacks over
GROMACS XTC/TRR trajectories and GRO/PDB structure files.

The reference opens ``mda.Universe(tpr_file, xtc_file)``
(pycusaxs/topology.py:52-74) and hands the engine one frame at a time through
``read_frame`` / ``get_box`` / ``iter_frames_stream`` (topology.py:536-737,
driven by RunSaxs.cu:48-92,137-170).  ``XdrTopology`` keeps those methods,
names and units (positions and box in Angstrom, time in ps) and reads the
trajectory with the package's C++ frame reader (include/saxs_xdr.h): the
file is memory-mapped and indexed once, and a batch of frames is decoded on
several host threads straight into the engine's pinned staging buffer
(``stage_frames``, the path ``pycusaxs_cuda.run`` takes when the topology
offers it).

Atoms and elements come from a GRO or PDB file of the same system (the
reference takes them from the TPR through MDAnalysis, which this image
lacks; TPR is not parsed here) or from an explicit ``elements`` argument.
Element symbols: a PDB element column when present, else the atom name's
leading letter, or the two-letter element for single-atom ion residues
(resname equal to the atom name, e.g. NA, CL, MG, ZN, and the common ion
residue names SOD, CLA, POT, CAL).

The box is the file's 3x3 matrix (rows = box vectors) times 10.  MDAnalysis
rebuilds ``triclinic_dimensions`` from the box lengths and angles, which
gives the same matrix for rectangular boxes and may differ in the last
float bits for triclinic ones.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Iterator, List, Optional, Sequence

import numpy as np

from ._build import XDR_LIB_PATH
from .formfactors import table

XDR_OK, XDR_EINVAL, XDR_EIO, XDR_EFORMAT = 0, -1, -2, -3
XDR_TRR, XDR_XTC = 1, 2
EXPORTED = ("saxs_xdr_open", "saxs_xdr_info", "saxs_xdr_read", "saxs_xdr_last_error",
            "saxs_xdr_close", "saxs_copier_create", "saxs_copier_copy", "saxs_copier_destroy")

_lib = None


def load():
    """The frame reader library (built in-tree by ``_build.build_xdr_library``);
    raises when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not XDR_LIB_PATH.exists():
        raise RuntimeError(f"trajectory reader not built ({XDR_LIB_PATH} missing): "
                           "run python -m pycusaxs_amd._build")
    lib = C.CDLL(str(XDR_LIB_PATH))
    vp = C.c_void_p
    lib.saxs_xdr_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.saxs_xdr_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int64)]
    lib.saxs_xdr_read.argtypes = [vp, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_float),
                                  C.POINTER(C.c_float), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int64), C.c_int]
    lib.saxs_xdr_last_error.restype = C.c_char_p
    lib.saxs_xdr_close.argtypes = [vp]
    lib.saxs_xdr_close.restype = None
    lib.saxs_copier_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.saxs_copier_create.restype = C.c_int
    lib.saxs_copier_copy.argtypes = [vp, vp, vp, C.c_size_t]
    lib.saxs_copier_copy.restype = C.c_int
    lib.saxs_copier_destroy.argtypes = [vp]
    lib.saxs_copier_destroy.restype = None
    for name in ("saxs_xdr_open", "saxs_xdr_info", "saxs_xdr_read"):
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _check(rc: int):
    if rc == XDR_OK:
        return
    msg = load().saxs_xdr_last_error().decode(errors="replace")
    if rc == XDR_EINVAL:
        raise ValueError(msg)
    if rc == XDR_EIO:
        raise OSError(msg)
    raise ValueError(f"trajectory format error: {msg}")


def _ptr(a: Optional[np.ndarray], ctype):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


class XdrTrajectory:
    """An indexed XTC or TRR file: ``n_frames``, ``n_atoms``, ``kind``
    ("xtc" / "trr") and ``read`` of any frames."""

    def __init__(self, path: str):
        lib = load()
        h = C.c_void_p()
        _check(lib.saxs_xdr_open(os.fsencode(path), C.byref(h)))
        self._h = h
        self.path = path
        kind, na, nf = C.c_int(), C.c_int(), C.c_int64()
        _check(lib.saxs_xdr_info(h, C.byref(kind), C.byref(na), C.byref(nf)))
        self.kind = "xtc" if kind.value == XDR_XTC else "trr"
        self.n_atoms = int(na.value)
        self.n_frames = int(nf.value)

    def read(self, frames: Sequence[int], xyz: Optional[np.ndarray] = None,
             box: Optional[np.ndarray] = None, threads: int = 0, want_xyz: bool = True):
        """Decodes `frames` into xyz [k, n_atoms, 3] and box [k, 3, 3] float32
        (Angstrom; given arrays must be C-contiguous float32 of those shapes,
        e.g. views of a pinned buffer) and returns (xyz, box, time [k] ps,
        step [k])."""
        if self._h is None:
            raise ValueError("trajectory is closed")
        f = np.ascontiguousarray(np.asarray(frames, dtype=np.int64).reshape(-1))
        k = len(f)
        if want_xyz:
            if xyz is None:
                xyz = np.empty((k, self.n_atoms, 3), dtype=np.float32)
            elif (xyz.dtype != np.float32 or not xyz.flags.c_contiguous
                  or xyz.shape != (k, self.n_atoms, 3)):
                raise ValueError(f"xyz must be C-contiguous float32 {(k, self.n_atoms, 3)}")
        else:
            xyz = None
        if box is None:
            box = np.empty((k, 3, 3), dtype=np.float32)
        elif box.dtype != np.float32 or not box.flags.c_contiguous or box.shape != (k, 3, 3):
            raise ValueError(f"box must be C-contiguous float32 {(k, 3, 3)}")
        time = np.empty(k, dtype=np.float64)
        step = np.empty(k, dtype=np.int64)
        nt = threads if threads > 0 else default_threads()
        _check(load().saxs_xdr_read(self._h, _ptr(f, C.c_int64), k, _ptr(xyz, C.c_float),
                                    _ptr(box, C.c_float), _ptr(time, C.c_double),
                                    _ptr(step, C.c_int64), nt))
        return xyz, box, time, step

    def close(self):
        if self._h is not None:
            load().saxs_xdr_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ structure ---

_ION_RESIDUES = {"SOD": "Na", "CLA": "Cl", "POT": "K", "CAL": "Ca", "NA+": "Na", "CL-": "Cl",
                 "K+": "K", "MG2+": "Mg", "ZN2+": "Zn"}


def _neutral_symbols():
    return {s for s in table() if s.isalpha() and s not in ("Du", "Cval", "Hiso", "Sival")}


def guess_element(name: str, resname: str = "") -> str:
    """Element symbol of an atom from its name and residue name (the rule in
    the module docstring)."""
    n = name.strip().lstrip("0123456789")
    letters = ""
    for ch in n:
        if not ch.isalpha():
            break
        letters += ch
    if not letters:
        raise ValueError(f"cannot guess the element of atom {name!r}")
    res = resname.strip().upper()
    if res in _ION_RESIDUES:
        return _ION_RESIDUES[res]
    neutral = _neutral_symbols()
    if res == letters.upper() and len(letters) <= 2 and letters.title() in neutral:
        return letters.title()
    return letters[0].upper()


class Structure:
    """Atoms of a GRO or PDB file: names, residue names, elements and the
    file's coordinates (Angstrom) and box (Angstrom, 3x3 rows)."""

    def __init__(self, names, resnames, elements, xyz, box):
        self.names = list(names)
        self.resnames = list(resnames)
        self.elements = list(elements)
        self.xyz = np.asarray(xyz, dtype=np.float32).reshape(-1, 3)
        self.box = np.asarray(box, dtype=np.float32).reshape(3, 3)

    @property
    def n_atoms(self) -> int:
        return len(self.names)


def read_gro(path: str) -> Structure:
    """GRO: title, atom count, one fixed-column line per atom (residue number
    5, residue name 5, atom name 5, atom number 5, then x y z in nm in fields
    whose width is the distance between the decimal points), box line
    v1x v2y v3z [v1y v1z v2x v2z v3x v3y] in nm."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    if len(lines) < 3:
        raise ValueError(f"{path}: not a GRO file")
    try:
        n = int(lines[1].strip())
    except ValueError as exc:
        raise ValueError(f"{path}: bad atom count line") from exc
    if len(lines) < n + 3:
        raise ValueError(f"{path}: {n} atoms announced, file has {len(lines) - 3} atom lines")
    names, resnames, elements = [], [], []
    xyz = np.zeros((n, 3), dtype=np.float32)
    for i in range(n):
        ln = lines[2 + i]
        res, name = ln[5:10], ln[10:15]
        first = ln.find(".", 20)
        second = ln.find(".", first + 1)
        w = second - first if first >= 0 and second > first else 8
        try:
            vals = [float(ln[20 + k * w:20 + (k + 1) * w]) for k in range(3)]
        except ValueError as exc:
            raise ValueError(f"{path}: bad coordinates on atom line {i + 1}") from exc
        xyz[i] = np.float32(vals) * np.float32(10.0)
        names.append(name.strip())
        resnames.append(res.strip())
        elements.append(guess_element(name, res))
    b = [float(v) for v in lines[2 + n].split()]
    if len(b) not in (3, 9):
        raise ValueError(f"{path}: box line needs 3 or 9 numbers")
    m = np.diag(b[:3])
    if len(b) == 9:
        m[0, 1], m[0, 2], m[1, 0], m[1, 2], m[2, 0], m[2, 1] = b[3:9]
    return Structure(names, resnames, elements, xyz,
                     m.astype(np.float32) * np.float32(10.0))


def _cell_matrix(a, b, c, alpha, beta, gamma):
    """Box rows from lengths and angles (degrees): a along x, b in the xy plane."""
    if alpha == 90.0 and beta == 90.0 and gamma == 90.0:
        return np.diag([a, b, c])
    ca, cb, cg = (np.cos(np.radians(t)) for t in (alpha, beta, gamma))
    sg = np.sin(np.radians(gamma))
    cx = c * cb
    cy = c * (ca - cb * cg) / sg
    cz = np.sqrt(max(c * c - cx * cx - cy * cy, 0.0))
    return np.array([[a, 0.0, 0.0], [b * cg, b * sg, 0.0], [cx, cy, cz]])


def read_pdb(path: str) -> Structure:
    """PDB: ATOM/HETATM records (name 13-16, residue 18-20, x y z 31-54 in
    Angstrom, element 77-78 when present) and CRYST1 for the box."""
    names, resnames, elements, xyz = [], [], [], []
    box = np.zeros((3, 3))
    with open(path) as fh:
        for ln in fh:
            rec = ln[:6]
            if rec == "CRYST1":
                a, b, c = float(ln[6:15]), float(ln[15:24]), float(ln[24:33])
                al, be, ga = float(ln[33:40]), float(ln[40:47]), float(ln[47:54])
                box = _cell_matrix(a, b, c, al, be, ga)
            elif rec in ("ATOM  ", "HETATM"):
                name, res = ln[12:16], ln[17:21]
                el = ln[76:78].strip() if len(ln) >= 78 else ""
                names.append(name.strip())
                resnames.append(res.strip())
                elements.append(el.title() if el else guess_element(name, res))
                xyz.append((float(ln[30:38]), float(ln[38:46]), float(ln[46:54])))
            elif rec.startswith("ENDMDL"):
                break
    if not names:
        raise ValueError(f"{path}: no ATOM/HETATM records")
    return Structure(names, resnames, elements, np.asarray(xyz, dtype=np.float32), box)


def read_structure(path: str) -> Structure:
    ext = os.path.splitext(path)[1].lower()
    if ext == ".gro":
        return read_gro(path)
    if ext in (".pdb", ".ent"):
        return read_pdb(path)
    if ext == ".tpr":
        raise ValueError("TPR topologies are not parsed here; pass the system as a GRO or PDB "
                         "file, or give elements= explicitly")
    raise ValueError(f"unknown structure format {ext!r} (GRO or PDB)")


# ------------------------------------------------------------- topology ---

class XdrTopology:
    """pycusaxs.topology.Topology's frame interface over an XTC/TRR file.

    XdrTopology(topology_file, trajectory_file, elements=None): topology_file
    is a GRO or PDB file of the system (or None when `elements` is given);
    elements is a sequence of per-atom symbols or a dict symbol -> atom
    indices, and overrides the file's."""

    def __init__(self, tpr_file: Optional[str], xtc_file: str, elements=None):
        self.tpr_file = tpr_file
        self.xtc_file = xtc_file
        self.traj = XdrTrajectory(xtc_file)
        self.structure = read_structure(tpr_file) if tpr_file else None
        n = self.traj.n_atoms
        if elements is not None:
            if isinstance(elements, dict):
                per_atom: List[Optional[str]] = [None] * n
                for sym, ids in elements.items():
                    for i in ids:
                        if not 0 <= int(i) < n:
                            raise ValueError(f"atom index {i} out of range [0, {n})")
                        per_atom[int(i)] = str(sym)
                if any(s is None for s in per_atom):
                    raise ValueError("elements dict does not cover every atom")
                self.elements = per_atom
            else:
                self.elements = [str(s) for s in elements]
        elif self.structure is not None:
            self.elements = list(self.structure.elements)
        else:
            raise ValueError("give a GRO/PDB topology file or elements=")
        if len(self.elements) != n:
            raise ValueError(f"topology has {len(self.elements)} atoms, trajectory {n}")
        self.ts = None
        self._pos = self._box = None
        self._time = 0.0
        self._step = 0

    # -- the reference's callbacks (topology.py:273-286, 536-619, 740-747)
    def get_atom_index(self) -> Dict[str, List[int]]:
        idx: Dict[str, List[int]] = {}
        for i, s in enumerate(self.elements):
            idx.setdefault(s, []).append(i)
        return idx

    def read_frame(self, frame_number: int):
        if not 0 <= frame_number < self.n_frames:
            raise IndexError(f"Frame {frame_number} out of range [0, {self.n_frames})")
        xyz, box, time, step = self.traj.read([frame_number])
        self._pos, self._box = xyz[0], box[0]
        self._time, self._step = float(time[0]), int(step[0])
        self.ts = int(frame_number)
        return self.ts

    def _need_frame(self, what):
        if self.ts is None:
            raise RuntimeError(f"Call read_frame() before {what}()")

    def get_box(self) -> np.ndarray:
        self._need_frame("get_box")
        return self._box

    def get_step(self) -> int:
        self._need_frame("get_step")
        return self.ts

    def get_time(self) -> float:
        self._need_frame("get_time")
        return self._time

    def get_coordinates(self) -> np.ndarray:
        self._need_frame("get_coordinates")
        return self._pos

    def frame_indices(self, start: int, stop: int, step: int = 1) -> range:
        """The frames trajectory[start:stop:step] selects (Python slice rules)."""
        return range(self.n_frames)[start:stop:step]

    def iter_frames_stream(self, start: int, stop: int, step: int = 1,
                           show_progress: bool = False, cluster_selection=None,
                           cluster_cutoff: float = 15.0, batch: int = 64) -> Iterator[Dict]:
        """Yields {'frame', 'time', 'positions' [N, 3] A, 'box' [3, 3] A} for
        trajectory[start:stop:step], decoding `batch` frames at a time.
        Cluster reconstruction (topology.py:689-727, MDAnalysis + DBSCAN) is
        not available: a cluster_selection raises."""
        if cluster_selection is not None:
            raise ValueError("cluster reconstruction needs MDAnalysis, which is not used here")
        ids = self.frame_indices(start, stop, step)
        for b0 in range(0, len(ids), max(1, batch)):
            chunk = list(ids[b0:b0 + batch])
            xyz, box, time, _ = self.traj.read(chunk)
            for k, f in enumerate(chunk):
                yield {"frame": f, "time": float(time[k]), "positions": xyz[k], "box": box[k]}

    def stage_frames(self, frames: Sequence[int], xyz_out: np.ndarray, box_out: np.ndarray,
                     threads: int = 0) -> np.ndarray:
        """Decodes `frames` straight into xyz_out [k, N, 3] and box_out
        [k, 3, 3] (float32, C-contiguous: the engine's pinned staging
        buffers) on host threads; returns their times (ps)."""
        _, _, time, _ = self.traj.read(frames, xyz=xyz_out, box=box_out, threads=threads)
        return time

    @property
    def n_frames(self) -> int:
        return self.traj.n_frames

    @property
    def n_atoms(self) -> int:
        return self.traj.n_atoms

    def close(self):
        self.traj.close()

    def __repr__(self) -> str:
        return (f"XdrTopology(tpr='{self.tpr_file}', n_atoms={self.n_atoms}, "
                f"n_frames={self.n_frames})")


class HostCopier:
    """Host threads kept for the copier's life that split one memory copy
    (``saxs_copier_*`` in the reader library): the run path copies each
    frame into its pinned staging buffer with it.  ``copy(dst, src)`` takes
    two C-contiguous numpy arrays of the same byte size and returns when the
    copy is done."""

    def __init__(self, threads: int = 4):
        self.lib = load()
        self._h = C.c_void_p()
        _check(self.lib.saxs_copier_create(int(threads), C.byref(self._h)))
        self.threads = int(threads)

    def copy(self, dst: np.ndarray, src: np.ndarray):
        if dst.nbytes != src.nbytes or not dst.flags.c_contiguous or not src.flags.c_contiguous:
            raise ValueError("copier needs two C-contiguous arrays of the same byte size")
        _check(self.lib.saxs_copier_copy(self._h, dst.ctypes.data, src.ctypes.data, dst.nbytes))

    def close(self):
        if self._h:
            self.lib.saxs_copier_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
