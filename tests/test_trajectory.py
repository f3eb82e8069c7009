"""The MDAnalysis-free trajectory reader (include/saxs_xdr.h,
pycusaxs_amd/trajectory.py) on files written from the published formats
(tests/xdr_files.py).

Parity unpinned: the reference ships no trajectory and MDAnalysis is not
importable, so no file written by GROMACS or read by MDAnalysis checks the
reader.  What is pinned: exact round trips of TRR (float and double) and of
XTC packing choices made by an independent encoder (runs, kept run lengths,
level changes both ways, separate fields for ranges above 2^24, the
uncompressed n <= 9 form), the bit budget the level table must satisfy,
frame indexing and errors on corrupt files, and the Topology callbacks'
contract (pycusaxs/topology.py:273-286, 536-737)."""
import os

import numpy as np
import pytest

from tests import xdr_files as X

pytest.importorskip("pycusaxs_amd.trajectory")
from pycusaxs_amd import trajectory as T  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _lib():
    try:
        T.load()
    except RuntimeError as exc:
        pytest.fail(str(exc))


def _write(tmp_path, name, blobs):
    p = tmp_path / name
    p.write_bytes(b"".join(blobs))
    return str(p)


def test_library_exports_every_declared_symbol():
    lib = T.load()
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "saxs_xdr.h")).read()
    for name in T.EXPORTED:
        assert name + "(" in hdr
        assert hasattr(lib, name)


def test_level_table_bit_budget():
    """A run triplet of level i is packed in i bits, so MAGIC[i]^3 <= 2^i;
    the levels grow by about 2^(1/3)."""
    for i in range(X.FIRST, X.LAST + 1):
        assert X.MAGIC[i] ** 3 <= 2 ** i, i
        assert 0.85 < X.MAGIC[i] / 2 ** (i / 3) <= 1.0, i
    assert all(a < b for a, b in zip(X.MAGIC[X.FIRST:], X.MAGIC[X.FIRST + 1:]))


@pytest.mark.parametrize("double", [False, True])
def test_trr_round_trip(tmp_path, double):
    rng = np.random.default_rng(3)
    n, F = 257, 5
    xs = [rng.uniform(-1, 6, size=(n, 3)) for _ in range(F)]
    boxes = [np.array([[5.0 + 0.01 * f, 0, 0], [0.3, 5.1, 0], [-0.2, 0.4, 4.9]]) for f in range(F)]
    blobs = [X.trr_frame(x, b, step=100 * f, t=0.5 * f, double=double,
                         v=(x * 0.1 if f % 2 else None)) for f, (x, b) in enumerate(zip(xs, boxes))]
    tr = T.XdrTrajectory(_write(tmp_path, "t.trr", blobs))
    assert (tr.kind, tr.n_atoms, tr.n_frames) == ("trr", n, F)
    xyz, box, time, step = tr.read(range(F))
    for f in range(F):
        xf = np.asarray(xs[f], dtype=np.float64 if double else np.float32).astype(np.float32)
        assert np.array_equal(xyz[f], xf * np.float32(10.0))
        bf = np.asarray(boxes[f], dtype=np.float64 if double else np.float32).astype(np.float32)
        assert np.array_equal(box[f], bf * np.float32(10.0))
    assert list(step) == [100 * f for f in range(F)]
    assert np.allclose(time, [0.5 * f for f in range(F)])
    tr.close()


def _xtc(tmp_path, frames_ints, boxes, precision=1000.0, seed=0, start_idx=12, max_small=8,
         stats=None):
    rng = np.random.default_rng(seed)
    blobs = [X.xtc_frame(q, b, step=10 * f, t=0.02 * f, precision=precision, rng=rng,
                         start_idx=start_idx, max_small=max_small, stats=stats)
             for f, (q, b) in enumerate(zip(frames_ints, boxes))]
    return _write(tmp_path, "t.xtc", blobs)


@pytest.mark.parametrize("seed,start_idx,max_small,spread", [(0, 12, 8, 0.06), (1, 9, 3, 0.004),
                                                            (2, 20, 10, 0.02), (3, 15, 1, 0.06)])
def test_xtc_round_trip_runs_and_levels(tmp_path, seed, start_idx, max_small, spread):
    rng = np.random.default_rng(seed)
    F, nmol = 4, 300
    box = np.diag([4.0, 4.2, 3.9])
    ints = [X.quantize(X.water_like(nmol, 4.0, rng, spread), 1000.0) for _ in range(F)]
    stats = {}
    path = _xtc(tmp_path, ints, [box] * F, seed=seed, start_idx=start_idx, max_small=max_small,
                stats=stats)
    # the frames exercise every packing choice the decoder follows
    assert stats["runs"] > 20 and stats["kept"] > 20 and stats["up"] > 20 and stats["down"] > 20
    assert stats["longest"] == max_small
    tr = T.XdrTrajectory(path)
    assert (tr.kind, tr.n_atoms, tr.n_frames) == ("xtc", 3 * nmol, F)
    xyz, bx, time, step = tr.read(range(F))
    for f in range(F):
        assert np.array_equal(xyz[f], X.expected_angstrom(ints[f], 1000.0))
        assert np.array_equal(bx[f], box.astype(np.float32) * np.float32(10.0))
    assert list(step) == [10 * f for f in range(F)]
    assert np.allclose(time, np.float32([0.02 * f for f in range(F)]))


def test_xtc_quantization_error_bound(tmp_path):
    """Decoded positions lie within half a quantum (0.5/precision nm) of
    the coordinates the encoder was given."""
    rng = np.random.default_rng(5)
    x = X.water_like(200, 5.0, rng)
    path = _xtc(tmp_path, [X.quantize(x, 1000.0)], [np.diag([5.0] * 3)])
    xyz, _, _, _ = T.XdrTrajectory(path).read([0])
    assert np.max(np.abs(xyz[0].astype(np.float64) / 10.0 - x)) <= 0.5e-3 + 1e-6


def test_xtc_large_range_separate_fields(tmp_path):
    """A coordinate range above 2^24 quanta takes the three-field path."""
    rng = np.random.default_rng(6)
    x = X.water_like(50, 30.0, rng)
    prec = 1.0e6                                   # 30 nm * 1e6 > 2^24
    q = X.quantize(x, prec)
    assert (q.max(0) - q.min(0) + 1).max() > 0xFFFFFF
    path = _xtc(tmp_path, [q], [np.diag([30.0] * 3)], precision=prec, start_idx=40)
    xyz, _, _, _ = T.XdrTrajectory(path).read([0])
    assert np.array_equal(xyz[0], X.expected_angstrom(q, prec))


def test_xtc_small_systems_uncompressed(tmp_path):
    x = np.array([[0.1, 0.2, 0.3], [1.5, 2.5, 3.25], [0.0, -1.0, 2.0]])
    path = _write(tmp_path, "s.xtc", [X.xtc_frame(x, np.diag([3.0] * 3), 0, 0.0)])
    xyz, _, _, _ = T.XdrTrajectory(path).read([0])
    assert np.array_equal(xyz[0], x.astype(np.float32) * np.float32(10.0))


def test_threaded_batch_equals_serial_and_any_order(tmp_path):
    rng = np.random.default_rng(7)
    F = 12
    ints = [X.quantize(X.water_like(150, 3.0, rng), 1000.0) for _ in range(F)]
    tr = T.XdrTrajectory(_xtc(tmp_path, ints, [np.diag([3.0] * 3)] * F))
    order = [11, 0, 5, 5, 3, 9, 1]
    a, _, ta, _ = tr.read(order, threads=1)
    b, _, tb, _ = tr.read(order, threads=8)
    assert np.array_equal(a, b) and np.array_equal(ta, tb)
    for k, f in enumerate(order):
        assert np.array_equal(a[k], X.expected_angstrom(ints[f], 1000.0))


def test_concurrent_reads_on_one_handle(tmp_path):
    """Reads from several Python threads on one handle (ctypes releases the
    GIL) are serialized and each gets its own frames."""
    import threading
    rng = np.random.default_rng(12)
    F = 8
    ints = [X.quantize(X.water_like(100, 2.5, rng), 1000.0) for _ in range(F)]
    tr = T.XdrTrajectory(_xtc(tmp_path, ints, [np.diag([2.5] * 3)] * F))
    out, errs = {}, []

    def work(first):
        try:
            for rep in range(5):
                fr = [(first + rep + k) % F for k in range(3)]
                xyz, _, _, _ = tr.read(fr, threads=4)
                for k, f in enumerate(fr):
                    if not np.array_equal(xyz[k], X.expected_angstrom(ints[f], 1000.0)):
                        errs.append((first, rep, f))
            out[first] = True
        except Exception as exc:          # reported below
            errs.append(exc)

    ths = [threading.Thread(target=work, args=(i,)) for i in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errs and len(out) == 4


def test_read_into_given_buffers(tmp_path):
    rng = np.random.default_rng(8)
    ints = [X.quantize(X.water_like(40, 2.0, rng), 1000.0) for _ in range(3)]
    tr = T.XdrTrajectory(_xtc(tmp_path, ints, [np.diag([2.0] * 3)] * 3))
    buf = np.zeros((5, tr.n_atoms, 3), dtype=np.float32)
    boxes = np.zeros((5, 3, 3), dtype=np.float32)
    tr.read([2, 0], xyz=buf[:2], box=boxes[:2])
    assert np.array_equal(buf[0], X.expected_angstrom(ints[2], 1000.0))
    assert np.array_equal(buf[1], X.expected_angstrom(ints[0], 1000.0))
    assert not buf[2:].any()
    with pytest.raises(ValueError):
        tr.read([0], xyz=np.zeros((1, tr.n_atoms, 3), dtype=np.float64))


def test_errors(tmp_path):
    rng = np.random.default_rng(9)
    q = X.quantize(X.water_like(30, 2.0, rng), 1000.0)
    good = X.xtc_frame(q, np.diag([2.0] * 3), 0, 0.0)
    with pytest.raises(OSError):
        T.XdrTrajectory(str(tmp_path / "missing.xtc"))
    with pytest.raises(ValueError, match="magic"):
        T.XdrTrajectory(_write(tmp_path, "bad.xtc", [b"\0\0\0\7" + good[4:]]))
    with pytest.raises(ValueError, match="truncated|past"):
        T.XdrTrajectory(_write(tmp_path, "cut.xtc", [good, good[:-8]]))
    other = X.xtc_frame(q[:-3], np.diag([2.0] * 3), 1, 0.1)
    with pytest.raises(ValueError, match="atoms"):
        T.XdrTrajectory(_write(tmp_path, "mixed.xtc", [good, other]))
    tr = T.XdrTrajectory(_write(tmp_path, "ok.xtc", [good]))
    with pytest.raises(ValueError, match="out of range"):
        tr.read([1])
    # a frame whose packed data (4 bytes) ends before its atoms do: header
    # words magic, natoms, step, time, box[9], natoms, then precision,
    # minint[3], maxint[3], smallidx, byte count
    n = len(q)
    hdr_len = 4 * (1 + 1 + 1 + 1 + 9 + 1) + 4 + 24 + 4
    body = good[:hdr_len] + (4).to_bytes(4, "big") + good[hdr_len + 4:hdr_len + 8]
    tr = T.XdrTrajectory(_write(tmp_path, "short.xtc", [body]))
    assert tr.n_atoms == n
    with pytest.raises(ValueError, match="ends early|run past"):
        tr.read([0])


def test_trr_without_positions_is_an_error(tmp_path):
    blob = X.trr_frame(None, np.diag([3.0] * 3), 0, 0.0, v=np.zeros((4, 3)))
    tr = T.XdrTrajectory(_write(tmp_path, "v.trr", [blob]))
    with pytest.raises(ValueError, match="no positions"):
        tr.read([0])
    _, box, _, _ = tr.read([0], want_xyz=False)
    assert np.array_equal(box[0], np.diag([30.0] * 3).astype(np.float32))


def test_guess_element():
    g = T.guess_element
    assert [g("OW", "SOL"), g("HW1", "SOL"), g("CA", "ALA"), g("1HB", "ALA"), g("NA", "NA"),
            g("CL", "CL"), g("CLA", "CLA"), g("SOD", "SOD"), g("CA", "CA"), g("N", "LYS"),
            g("MG", "MG"), g("SG", "CYS")] == \
        ["O", "H", "C", "H", "Na", "Cl", "Cl", "Na", "Ca", "N", "Mg", "S"]


def _system(tmp_path, F=6, nmol=80, seed=11):
    rng = np.random.default_rng(seed)
    L = 3.5
    x0 = X.water_like(nmol, L, rng)
    names, res = [], []
    for m in range(nmol):
        names += ["OW", "HW1", "HW2"]
        res += ["SOL"] * 3
    x0 = np.vstack([x0, [[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]]])
    names += ["NA", "CL"]
    res += ["NA", "CL"]
    box = np.diag([L] * 3)
    gro = tmp_path / "sys.gro"
    gro.write_text(X.gro_text(names, res, x0, box))
    ints = [X.quantize(x0 + rng.normal(0, 0.01, x0.shape), 1000.0) for _ in range(F)]
    xtc = _xtc(tmp_path, ints, [box * (1 + 0.002 * f) for f in range(F)], seed=seed)
    return str(gro), xtc, ints, names


def test_topology_callbacks(tmp_path):
    gro, xtc, ints, names = _system(tmp_path)
    topo = T.XdrTopology(gro, xtc)
    assert (topo.n_atoms, topo.n_frames) == (len(names), 6)
    idx = topo.get_atom_index()
    assert sorted(idx) == ["Cl", "H", "Na", "O"]
    assert idx["Na"] == [len(names) - 2] and idx["Cl"] == [len(names) - 1]
    assert len(idx["H"]) == 2 * len(idx["O"])
    with pytest.raises(RuntimeError, match="read_frame"):
        topo.get_box()
    topo.read_frame(3)
    assert topo.get_step() == 3
    assert np.array_equal(topo.get_coordinates(), X.expected_angstrom(ints[3], 1000.0))
    assert np.array_equal(topo.get_box(),
                          (np.diag([3.5] * 3) * (1 + 0.002 * 3)).astype(np.float32) * np.float32(10))
    assert abs(topo.get_time() - 0.06) < 1e-6
    with pytest.raises(IndexError):
        topo.read_frame(6)
    got = list(topo.iter_frames_stream(1, 100, 2, batch=2))
    assert [g["frame"] for g in got] == [1, 3, 5]
    for g in got:
        assert np.array_equal(g["positions"], X.expected_angstrom(ints[g["frame"]], 1000.0))
        assert g["box"].shape == (3, 3)
    assert list(topo.frame_indices(0, 7, 1)) == list(range(6))
    with pytest.raises(ValueError, match="MDAnalysis"):
        next(topo.iter_frames_stream(0, 2, 1, cluster_selection="resname SOL"))


def test_topology_elements_override_and_mismatch(tmp_path):
    gro, xtc, ints, names = _system(tmp_path)
    n = len(names)
    topo = T.XdrTopology(None, xtc, elements={"C": range(n - 1), "S": [n - 1]})
    assert sorted(topo.get_atom_index()) == ["C", "S"]
    with pytest.raises(ValueError, match="atoms"):
        T.XdrTopology(None, xtc, elements=["C"] * (n - 1))
    with pytest.raises(ValueError, match="TPR"):
        T.XdrTopology(str(tmp_path / "topol.tpr"), xtc)


def test_stage_frames_into_staging_buffers(tmp_path):
    gro, xtc, ints, names = _system(tmp_path)
    topo = T.XdrTopology(gro, xtc)
    xyz = np.zeros((4, topo.n_atoms, 3), dtype=np.float32)
    box = np.zeros((4, 3, 3), dtype=np.float32)
    times = topo.stage_frames([0, 2, 4], xyz[:3], box[:3], threads=3)
    assert np.allclose(times, [0.0, 0.04, 0.08])
    for k, f in enumerate([0, 2, 4]):
        assert np.array_equal(xyz[k], X.expected_angstrom(ints[f], 1000.0))


def test_pdb_structure(tmp_path):
    p = tmp_path / "s.pdb"
    p.write_text(
        "CRYST1   40.000   41.000   42.000  90.00  90.00  90.00 P 1           1\n"
        "ATOM      1  N   ALA A   1      11.104   6.134  -6.504  1.00  0.00           N\n"
        "ATOM      2  CA  ALA A   1      11.639   6.071  -5.147  1.00  0.00           C\n"
        "HETATM    3 NA    NA A   2       1.000   2.000   3.000  1.00  0.00          NA\n"
        "ATOM      4  OW  SOL A   3       5.000   5.000   5.000  1.00  0.00\n")
    s = T.read_pdb(str(p))
    assert s.elements == ["N", "C", "Na", "O"]
    assert np.array_equal(s.box, np.diag([40.0, 41.0, 42.0]).astype(np.float32))
    assert np.allclose(s.xyz[1], [11.639, 6.071, -5.147])


def test_corrupted_files_under_address_sanitizer(tmp_path):
    """1,500 corrupted copies each of an XTC and a TRR file (bit flips,
    random bytes, truncations) through open/index/decode in a driver built
    with AddressSanitizer and UndefinedBehaviorSanitizer (host code): no
    memory error, out-of-bounds read or undefined behaviour."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    root = os.path.join(os.path.dirname(__file__), "..")
    exe = str(tmp_path / "xdr_fuzz")
    build = subprocess.run([cxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                            "-fno-sanitize-recover=all", "-pthread",
                            "-I" + os.path.join(root, "include"), "-o", exe,
                            os.path.join(root, "tests", "xdr_fuzz.cpp"),
                            os.path.join(root, "pycusaxs_amd", "csrc", "xdr_frames.cpp")],
                           capture_output=True, text=True)
    if build.returncode != 0 and "sanitize" in build.stderr:
        pytest.skip("sanitizer runtime unavailable")
    assert build.returncode == 0, build.stderr
    rng = np.random.default_rng(13)
    q = [X.quantize(X.water_like(120, 3.0, rng, 0.05), 1000.0) for _ in range(3)]
    xtc = _write(tmp_path, "seed.xtc", [X.xtc_frame(q[f], np.diag([3.0] * 3), f, 0.1 * f, rng=rng,
                                                   start_idx=15) for f in range(3)])
    x = rng.uniform(0, 3, (40, 3))
    trr = _write(tmp_path, "seed.trr", [X.trr_frame(x, np.diag([3.0] * 3), f, 0.1 * f,
                                                   double=(f == 1)) for f in range(3)])
    for seed in (xtc, trr):
        run = subprocess.run([exe, seed, "1500", str(tmp_path / "case.bin")],
                             capture_output=True, text=True, timeout=300)
        assert run.returncode == 0, run.stderr[-2000:]
        assert "1500 cases" in run.stdout


def test_host_copier_copies_exactly():
    """The run path's parallel host copy: byte-exact for sizes that are not
    multiples of the 64-byte split, below the parallel threshold, and over
    repeated calls (workers spinning between calls and, after a pause,
    woken from their sleep)."""
    import time
    from pycusaxs_amd.trajectory import HostCopier
    rng = np.random.default_rng(5)
    for threads in (1, 2, 4, 7):
        cp = HostCopier(threads)
        for nbytes in (0, 1, 63, 4096, 65537, 1_200_012, 3 * 100_000 * 4):
            src = rng.integers(0, 256, size=nbytes, dtype=np.uint8)
            dst = np.zeros(nbytes, dtype=np.uint8)
            cp.copy(dst, src)
            assert np.array_equal(dst, src), (threads, nbytes)
        src = rng.standard_normal((100_000, 3)).astype(np.float32)
        for k in range(50):
            dst = np.empty_like(src)
            cp.copy(dst, src + k)
            assert np.array_equal(dst, src + k)
            if k == 25:
                time.sleep(0.01)                    # workers go to sleep
        cp.close()
    with pytest.raises(ValueError):
        HostCopier(2).copy(np.zeros(10, np.float32), np.zeros(11, np.float32))
