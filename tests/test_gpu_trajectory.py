"""pycusaxs_cuda.run fed by the MDAnalysis-free reader (XdrTopology over a
GRO + XTC/TRR pair written by tests/xdr_files.py): its staged path (frames
decoded on host threads straight into the pinned staging buffers) writes the
same saxs.dat, byte for byte, as the per-frame iterator path over the same
positions and boxes, for cubic and triclinic NPT boxes that change every
frame, float and double TRR, and batches that cross the run's batch size;
and against the FP64 oracle for a fixed box."""
import numpy as np
import pytest

import pycusaxs_cuda
from oracle import saxs_oracle as O
from pycusaxs_amd import trajectory as T
from tests import xdr_files as X
from tests.fake_topology import FakeTopology

pytestmark = pytest.mark.gpu


TRIC = np.array([[3.6, 0.0, 0.0], [0.4, 3.5, 0.0], [-0.3, 0.2, 3.7]])


def _system(tmp_path, F, fmt, grow, seed=31, nmol=700, tric=False, double=False):
    rng = np.random.default_rng(seed)
    L = 3.6
    x0 = X.water_like(nmol, L, rng, 0.08)
    names = ["OW", "HW1", "HW2"] * nmol + ["NA", "CL", "CA", "N", "C", "O"]
    res = ["SOL"] * (3 * nmol) + ["NA", "CL", "ALA", "ALA", "ALA", "ALA"]
    x0 = np.vstack([x0, rng.uniform(0, L, size=(6, 3))])
    base = TRIC if tric else np.diag([L] * 3)
    gro = tmp_path / "sys.gro"
    gro.write_text(X.gro_text(names, res, x0, base))
    boxes = [base * (1 + grow * f) for f in range(F)]
    # fractional coordinates in each frame's cell (wrapped), then Cartesian
    frac = (x0 / L) % 1.0
    xs = [((frac + rng.normal(0, 0.003, frac.shape)) % 1.0) @ b for b in boxes]
    if fmt == "xtc":
        blobs = [X.xtc_frame(X.quantize(x, 1000.0), b, 5 * f, 0.1 * f, rng=rng)
                 for f, (x, b) in enumerate(zip(xs, boxes))]
    else:
        blobs = [X.trr_frame(x, b, 5 * f, 0.1 * f, double=double)
                 for f, (x, b) in enumerate(zip(xs, boxes))]
    traj = tmp_path / f"traj.{fmt}"
    traj.write_bytes(b"".join(blobs))
    return str(gro), str(traj)


@pytest.mark.parametrize("fmt,grow,stride,tric,double",
                         [("xtc", 0.002, 1, False, False), ("trr", 0.0, 2, False, False),
                          ("xtc", 0.0, 3, False, False), ("xtc", 0.002, 1, True, False),
                          ("trr", 0.003, 1, True, True)])
def test_run_staged_reader_matches_iterator_path(gpu, tmp_path, fmt, grow, stride, tric, double):
    F = 37                                             # crosses the batch size (16)
    gro, traj = _system(tmp_path, F, fmt, grow, tric=tric, double=double)
    topo = T.XdrTopology(gro, traj)
    xyz, box, time, _ = topo.traj.read(range(F))
    elements = topo.elements
    fake = FakeTopology(list(xyz), elements, box[0])
    fake.boxes = list(box)
    fake.dt = float(time[1] - time[0])
    kw = dict(grid=[40], begin=2, end=F - 1, stride=stride, order=4, scale_factor=2.0,
              bin_size=0.02, qcut=0.9)
    out_a, out_b = tmp_path / "a.dat", tmp_path / "b.dat"
    pycusaxs_cuda.run(topo, gro, traj, output=str(out_a), **kw)
    pycusaxs_cuda.run(fake, gro, traj, output=str(out_b), **kw)
    a, b = out_a.read_text(), out_b.read_text()
    print(f"[{fmt} grow {grow} stride {stride} tric {tric}] {len(a.splitlines())} rows", flush=True)
    assert len(a.splitlines()) >= 8
    assert a == b


def test_run_from_xtc_against_oracle(gpu, tmp_path):
    F = 5
    gro, traj = _system(tmp_path, F, "xtc", 0.0, seed=32, nmol=500)
    topo = T.XdrTopology(gro, traj)
    out = tmp_path / "o.dat"
    pycusaxs_cuda.run(topo, gro, traj, grid=[40], begin=0, end=F - 1, stride=1,
                      output=str(out), order=4, scale_factor=2.0, bin_size=0.02, qcut=0.9)
    xyz, box, _, _ = topo.traj.read(range(F))
    symbols, ids = O.element_ids(topo.elements)
    plan = O.make_plan((40,) * 3, box[0], 4, 2.0, None, 0.02, 0.9)
    H, C = O.trajectory_histogram(plan, [(x, box[0]) for x in xyz], ids, symbols)
    want = O.saxs_rows(plan, H, C)
    got = [(float(ln[:10]), float(ln[10:22])) for ln in out.read_text().splitlines()]
    assert len(got) == len(want) and len(got) >= 8
    for (qg, ig), (qw, iw) in zip(got, want):
        assert ("%10.5f" % qg) == ("%10.5f" % qw)
        assert abs(ig - iw) <= 1e-5 * abs(iw) + 1e-5 * 10 ** np.floor(np.log10(abs(iw)))
