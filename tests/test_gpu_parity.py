"""GPU parity of the engine's forward path against the FP64 oracle.

Every case runs the HIP engine through the C ABI (via pycusaxs_amd.engine)
and the oracle on the same seeded inputs.  Bar: bin counts C_b bit-exact;
I(q) = H_b/C_b within a relative error of 1e-5 (FP32 engine vs FP64 oracle).
"""
import sys

import numpy as np
import pytest

from oracle import saxs_oracle as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def engine_hist(frames, syms, box, n, sigma=2.0, dq=0.0, qcut=0.0, order=4, scaled=None,
                policy="reference", max_frames=None, tile=(0, 0), flags=0):
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    symbols, ids = O.element_ids(syms)
    n3, nn = resolve_grid((n,) * 3 if np.isscalar(n) else n, sigma, scaled)
    spec = SystemSpec(n3, nn, order, tuple(symbols), sigma, dq, qcut, policy)
    xyz = torch.from_numpy(np.stack(frames).astype(np.float32))
    eng = SaxsEngine(spec, box, max_atoms=xyz.shape[1], max_frames=max_frames or len(frames),
                     tile=tile, flags=flags)
    eng.forward(xyz, torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    eng.close()
    return H, C


def oracle_hist(frames, syms, box, n, sigma=2.0, dq=0.0, qcut=0.0, order=4, scaled=None,
                policy="reference"):
    symbols, ids = O.element_ids(syms)
    plan = O.make_plan((n,) * 3 if np.isscalar(n) else n, box, order, sigma, scaled, dq, qcut,
                       policy)
    return O.trajectory_histogram(plan, [(f, box) for f in frames], ids, symbols)


def compare(H, C, Ho, Co, tol=REL_TOL, label=""):
    print(f"[{label}] counts {C[:8]} ...", flush=True)
    assert np.array_equal(C, Co), f"bin counts differ: {C} vs {Co}"
    sel = np.arange(len(C)) >= 1
    sel &= Co > 0
    I, Io = H[sel] / C[sel], Ho[sel] / Co[sel]
    rel = np.abs(I - Io) / np.abs(Io)
    print(f"[{label}] max rel err {rel.max():.3e} over {sel.sum()} bins", flush=True)
    assert rel.max() <= tol, f"max rel err {rel.max():.3e} > {tol} ({rel})"


def test_single_atom_on_node(gpu):
    box = np.diag([32.0] * 3)
    pos = np.array([[8.0, 12.0, 20.0]], dtype=np.float32)
    H, C = engine_hist([pos], ["Du"], box, 16, 2.0, 0.0, 100.0)
    Ho, Co = oracle_hist([pos], ["Du"], box, 16, 2.0, 0.0, 100.0)
    compare(H, C, Ho, Co, label="single atom")


def test_5k_atoms_37A_box_64_grid(gpu):
    box = np.diag([37.0] * 3)
    pos, syms = O.synthetic_system(5000, 37.0, seed=1)
    H, C = engine_hist([pos], syms, box, 64, 2.0, 0.01, 0.5)
    assert list(C[1:6].astype(int)) == [34, 82, 178, 258, 402]
    Ho, Co = oracle_hist([pos], syms, box, 64, 2.0, 0.01, 0.5)
    compare(H, C, Ho, Co, label="5k atoms, 37 A, 64 -> 128")


def test_full_spectrum_small_grid(gpu):
    box = np.diag([20.0] * 3)
    pos, syms = O.synthetic_system(300, 20.0, seed=7)
    H, C = engine_hist([pos], syms, box, 16, 2.0, 0.0, 100.0)
    Ho, Co = oracle_hist([pos], syms, box, 16, 2.0, 0.0, 100.0)
    compare(H, C, Ho, Co, label="full 16->32")


def test_order6(gpu):
    box = np.diag([30.0] * 3)
    pos, syms = O.synthetic_system(2000, 30.0, seed=11)
    H, C = engine_hist([pos], syms, box, 32, 2.0, 0.0, 100.0, order=6)
    Ho, Co = oracle_hist([pos], syms, box, 32, 2.0, 0.0, 100.0, order=6)
    compare(H, C, Ho, Co, label="order 6")


@pytest.mark.parametrize("n,sigma", [(30, 2.0), (28, 2.5), (27, 2.0), (21, 2.5)])
def test_mixed_radix_and_partial_tiles(gpu, n, sigma):
    box = np.diag([25.0] * 3)
    pos, syms = O.synthetic_system(1500, 25.0, seed=n)
    H, C = engine_hist([pos], syms, box, n, sigma, 0.0, 100.0)
    Ho, Co = oracle_hist([pos], syms, box, n, sigma, 0.0, 100.0)
    compare(H, C, Ho, Co, label=f"n={n} sigma={sigma}")


def test_batched_frames_accumulate(gpu):
    box = np.diag([40.0] * 3)
    base, syms = O.synthetic_system(3000, 40.0, seed=3)
    frames = [O.frame_positions(base, 3, f, 0.5) for f in range(4)]
    H, C = engine_hist(frames, syms, box, 32, 2.0, 0.0, 1.0, max_frames=3)
    Ho, Co = oracle_hist(frames, syms, box, 32, 2.0, 0.0, 1.0)
    compare(H, C, Ho, Co, label="4 frames, batch 3")


def test_exact_dq_policy(gpu):
    box = np.diag([40.0] * 3)
    pos, syms = O.synthetic_system(2000, 40.0, seed=5)
    H, C = engine_hist([pos], syms, box, 32, 2.0, 0.01, 0.6, policy="exact")
    Ho, Co = oracle_hist([pos], syms, box, 32, 2.0, 0.01, 0.6, policy="exact")
    compare(H, C, Ho, Co, label="exact dq")


def test_atoms_outside_box_wrap(gpu):
    box = np.diag([30.0] * 3)
    pos, syms = O.synthetic_system(1000, 30.0, seed=9)
    shifted = pos + np.float32(45.0) * np.sign(np.random.default_rng(1).normal(size=pos.shape)).astype(np.float32)
    H, C = engine_hist([shifted], syms, box, 32, 2.0, 0.0, 100.0)
    Ho, Co = oracle_hist([shifted], syms, box, 32, 2.0, 0.0, 100.0)
    compare(H, C, Ho, Co, label="wrapped")


def plain_hist(frames, syms, box, n, sigma=2.0, dq=0.0, qcut=0.0, order=4, flags=0):
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    symbols, ids = O.element_ids(syms)
    n3, nn = resolve_grid((n,) * 3, sigma)
    eng = SaxsEngine(SystemSpec(n3, nn, order, tuple(symbols), sigma, dq, qcut), box,
                     max_atoms=len(frames[0]), max_frames=1, flags=flags)
    eng.forward_plain(torch.from_numpy(np.stack(frames)), torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    eng.close()
    return H, C


def test_plain_path_5k_atoms(gpu):
    box = np.diag([37.0] * 3)
    pos, syms = O.synthetic_system(5000, 37.0, seed=1)
    H, C = plain_hist([pos], syms, box, 64, 2.0, 0.01, 0.5)
    Ho, Co = oracle_hist([pos], syms, box, 64, 2.0, 0.01, 0.5)
    compare(H, C, Ho, Co, label="plain, 5k atoms")


@pytest.mark.parametrize("n,sigma,order", [(16, 2.0, 4), (30, 2.0, 4), (32, 2.0, 6)])
def test_plain_path_full_spectrum(gpu, n, sigma, order):
    box = np.diag([24.0] * 3)
    base, syms = O.synthetic_system(800, 24.0, seed=n)
    frames = [O.frame_positions(base, n, f, 0.4) for f in range(2)]
    H, C = plain_hist(frames, syms, box, n, sigma, 0.0, 100.0, order)
    Ho, Co = oracle_hist(frames, syms, box, n, sigma, 0.0, 100.0, order)
    compare(H, C, Ho, Co, label=f"plain n={n} P={order}")


def test_per_frame_histograms(gpu):
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    box = np.diag([30.0] * 3)
    base, syms = O.synthetic_system(1000, 30.0, seed=12)
    frames = [O.frame_positions(base, 12, f, 0.5) for f in range(3)]
    symbols, ids = O.element_ids(syms)
    n3, nn = resolve_grid((32,) * 3, 2.0)
    eng = SaxsEngine(SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.0, 1.0), box, 1000, 2)
    Hf, Cf = eng.forward_frames(torch.from_numpy(np.stack(frames)), torch.from_numpy(ids))
    Hf, Cf = Hf.cpu().numpy(), Cf.cpu().numpy()
    plan = O.make_plan((32,) * 3, box, 4, 2.0, None, 0.0, 1.0)
    for f in range(3):
        Ho, Co = O.frame_histogram(plan, frames[f], ids, symbols, box)
        compare(Hf[f], Cf[f], Ho, Co, label=f"frame {f}")
    eng.close()


def test_invalid_element_ids_are_ignored(gpu):
    """Atoms whose element id is outside [0, E) are skipped by the kernels
    (bounds safety), equal to leaving them out."""
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    box = np.diag([30.0] * 3)
    pos, syms = O.synthetic_system(600, 30.0, seed=13)
    symbols, ids = O.element_ids(syms)
    bad = ids.copy()
    bad[::5] = 99
    n3, nn = resolve_grid((32,) * 3, 2.0)
    eng = SaxsEngine(SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.0, 1.0), box, 600, 1)
    eng.forward(torch.from_numpy(pos[None]), torch.from_numpy(bad))
    torch.cuda.synchronize()
    keep = np.ones(len(pos), bool)
    keep[::5] = False
    plan = O.make_plan((32,) * 3, box, 4, 2.0, None, 0.0, 1.0)
    Ho, Co = O.frame_histogram(plan, pos[keep], ids[keep], symbols, box)
    compare(eng.H.cpu().numpy(), eng.C.cpu().numpy(), Ho, Co, label="invalid ids")
    eng.close()


@pytest.mark.parametrize("flags,tile,qcut", [(1, (4, 8), 1.5), (0, (4, 4), 1.5), (1, (2, 8), 1.5),
                                            (0, (8, 4), 1.5), (2, (4, 8), 1.0), (0, (4, 8), 1.0),
                                            (0, (4, 4), 0.6), (0, (2, 8), 0.6)])
def test_spread_variants_and_tiles(gpu, flags, tile, qcut):
    """Per-element / all-element spreading workgroups, the pruned radix-16 z
    transform (small qcut) and the full one, several tiles."""
    box = np.diag([40.0] * 3)
    base, syms = O.synthetic_system(4000, 40.0, seed=17)
    frames = [O.frame_positions(base, 17, f, 0.5) for f in range(2)]
    H, C = engine_hist(frames, syms, box, 64, 2.0, 0.0, qcut, tile=tile, flags=flags)
    Ho, Co = oracle_hist(frames, syms, box, 64, 2.0, 0.0, qcut)
    compare(H, C, Ho, Co, label=f"flags={flags} tile={tile} qcut={qcut}")


def test_pruned_z_n256(gpu):
    box = np.diag([60.0] * 3)
    pos, syms = O.synthetic_system(20000, 60.0, seed=19)
    H, C = engine_hist([pos], syms, box, 128, 2.0, 0.0, 0.5)
    Ho, Co = oracle_hist([pos], syms, box, 128, 2.0, 0.0, 0.5)
    compare(H, C, Ho, Co, label="pruned 128->256")


def test_pruned_z_n512_matches_full_transform(gpu):
    """N = 512 (sigma 4 on a 128 cell): the pruned radix-16 transform against
    the engine's full Stockham transform of the same frame."""
    box = np.diag([80.0] * 3)
    pos, syms = O.synthetic_system(20000, 80.0, seed=20)
    H1, C1 = engine_hist([pos], syms, box, 128, 4.0, 0.0, 0.4, tile=(4, 4))
    H2, C2 = engine_hist([pos], syms, box, 128, 4.0, 0.0, 0.4, tile=(4, 4), flags=2)
    compare(H1, C1, H2, C2, tol=2e-6, label="pruned vs full, N=512")


def test_pruned_x_pass_chunked_kz(gpu):
    """FFT grid 256 with kz = 29 kept z planes: the pruned x pass splits the
    kz lines over two workgroups (LDS), the pruned z and y passes run at
    their largest kept range (kz <= N / 8)."""
    box = np.diag([60.0] * 3)
    pos, syms = O.synthetic_system(5000, 60.0, seed=23)
    H, C = engine_hist([pos], syms, box, 128, 2.0, 0.0, 1.5)
    Ho, Co = oracle_hist([pos], syms, box, 128, 2.0, 0.0, 1.5)
    compare(H, C, Ho, Co, label="pruned 128->256, kz 29")


def test_counting_sort_global_fallback(gpu):
    """Column tile 2x2 on a 128^3 cell: E * 64 * 64 (element, tile) keys per
    frame exceed the LDS histogram of the segmented sort (12288 keys), so
    the counting sort with one global counter per key runs."""
    box = np.diag([60.0] * 3)
    pos, syms = O.synthetic_system(8000, 60.0, seed=25)
    assert len(set(syms)) * 64 * 64 > 12288
    H, C = engine_hist([pos], syms, box, 128, 2.0, 0.0, 0.5, tile=(2, 2))
    Ho, Co = oracle_hist([pos], syms, box, 128, 2.0, 0.0, 0.5)
    compare(H, C, Ho, Co, label="tile 2x2, global-counter sort")


def test_pruned_forward_independent_of_atomic_order(gpu):
    """The pruned path sums the spread columns and the border slabs in fixed
    point, so the spectra do not depend on the order of the atomics: repeated
    launches and other column tilings (other atom-to-workgroup assignments)
    agree to FP64 rounding of the histogram sums (float atomics differ at
    ~1e-7)."""
    box = np.diag([40.0] * 3)
    base, syms = O.synthetic_system(4000, 40.0, seed=21)
    frames = [O.frame_positions(base, 21, f, 0.5) for f in range(2)]
    runs = [engine_hist(frames, syms, box, 64, 2.0, 0.0, 0.6, tile=t)
            for t in [(8, 4), (8, 4), (4, 4), (2, 8)]]
    H0, C0 = runs[0]
    sel = C0 > 0
    for H, C in runs[1:]:
        assert np.array_equal(C, C0)
        rel = np.abs(H[sel] - H0[sel]) / np.abs(H0[sel])
        print(f"max rel diff {rel.max():.3e}", flush=True)
        assert rel.max() <= 1e-12


def test_dense_tiles_multiple_rounds(gpu):
    """More atoms per tile than one staging round holds (512 records)."""
    box = np.diag([20.0] * 3)
    pos, syms = O.synthetic_system(60000, 20.0, seed=18)
    H, C = engine_hist([pos], syms, box, 16, 2.0, 0.0, 100.0)
    Ho, Co = oracle_hist([pos], syms, box, 16, 2.0, 0.0, 100.0)
    compare(H, C, Ho, Co, label="dense")


def test_pruned_z_n512_exact_half_matches_full_transform(gpu):
    """N = 512 on a 256 cell (sigma 2, n_z = N/2: the bounds-free stage 0):
    pruned transforms against the engine's full Stockham transforms."""
    box = np.diag([160.0] * 3)
    pos, syms = O.synthetic_system(30000, 160.0, seed=22)
    H1, C1 = engine_hist([pos], syms, box, 256, 2.0, 0.0, 0.3, tile=(4, 4))
    H2, C2 = engine_hist([pos], syms, box, 256, 2.0, 0.0, 0.3, tile=(4, 4), flags=2)
    compare(H1, C1, H2, C2, tol=2e-6, label="pruned vs full, N=512, n_z = N/2")


@pytest.mark.parametrize("n,sigma,qcut", [(27, 2.0, 100.0),    # nn 54 = 9 x 6
                                          (30, 2.0, 100.0),    # 60 = 15 x 4
                                          (36, 2.0, 1.2),      # 72 = 12 x 6
                                          (35, 2.0, 100.0),    # 70 = 10 x 7
                                          (40, 2.0, 1.0),      # 80 = 16 x 5
                                          (45, 2.0, 1.0),      # 90 = 15 x 6
                                          (32, 2.0, 100.0)])   # 64 = 16 x 4 (pow2)
def test_wide_radix_factorisation(gpu, n, sigma, qcut):
    """Stockham transforms factored with the widest register radices
    (16, 15, 12, 10, 9, 6 as composite register DFTs) against the oracle,
    forward and the plain per-frame path."""
    box = np.diag([30.0] * 3)
    pos, syms = O.synthetic_system(1500, 30.0, seed=n + 50)
    H, C = engine_hist([pos], syms, box, n, sigma, 0.0, qcut, flags=4 | 2)
    Ho, Co = oracle_hist([pos], syms, box, n, sigma, 0.0, qcut)
    compare(H, C, Ho, Co, label=f"wide radix n={n}")
    Hp, Cp = plain_hist([pos], syms, box, n, sigma, 0.0, qcut, flags=4)
    compare(Hp, Cp, Ho, Co, label=f"plain n={n}")


@pytest.mark.parametrize("n,sigma,qcut,order", [(35, 2.0, 100.0, 4),   # N 70 = 14 x 5, all outputs
                                                (45, 2.0, 1.0, 4),     # 90 = 18 x 5
                                                (50, 2.0, 0.8, 4),     # 100 = 20 x 5
                                                (70, 2.0, 0.5, 4),     # 140 = 20 x 7
                                                (81, 2.0, 0.5, 4),     # 162 = 18 x 9
                                                (30, 2.0, 100.0, 6),   # 60 = 10 x 6, order 6
                                                (48, 2.0, 0.9, 6),     # 96 = 16 x 6, order 6
                                                (28, 2.5, 100.0, 4),   # 70, n_z < N/2
                                                (40, 2.5, 0.7, 4)])    # 100, n_z < N/2
def test_generic_pruned_z_against_oracle(gpu, n, sigma, qcut, order):
    """The output-pruned z transform for N = 2H R2 (H = 5..10): half-input
    first stage of radix 2H, kept outputs as direct sums over R2."""
    L = 0.75 * n
    box = np.diag([L] * 3)
    pos, syms = O.synthetic_system(max(500, int(0.1 * L ** 3)), L, seed=n + 70)
    H, C = engine_hist([pos], syms, box, n, sigma, 0.0, qcut, order=order)
    Ho, Co = oracle_hist([pos], syms, box, n, sigma, 0.0, qcut, order=order)
    compare(H, C, Ho, Co, label=f"generic pruned z n={n} sigma={sigma} order={order}")


@pytest.mark.parametrize("n,tile", [(189, (4, 8)), (105, (4, 4)), (144, (2, 8))])
def test_generic_pruned_z_matches_full_transform(gpu, n, tile):
    """Grids of the uneven-systems workload (N = 378 = 18 x 21, 210 = 14 x 15,
    288 = 18 x 16): the generic pruned z kernel against the full Stockham z
    transform (flag 2) at 0.1 atoms/A^3."""
    L = 0.75 * n
    box = np.diag([L] * 3)
    pos, syms = O.synthetic_system(int(0.1 * L ** 3), L, seed=n + 90)
    H1, C1 = engine_hist([pos], syms, box, n, 2.0, 0.0, 0.5, tile=tile)
    H2, C2 = engine_hist([pos], syms, box, n, 2.0, 0.0, 0.5, tile=tile, flags=2)
    compare(H1, C1, H2, C2, tol=2e-6, label=f"generic pruned vs full, n={n}")


@pytest.mark.parametrize("n,qcut", [((48, 70, 64), 0.6),     # y: N 140 = 20 x 7; z 128 radix 16
                                    ((40, 81, 45), 0.5),     # y: 162 = 18 x 9; z: 90 = 18 x 5
                                    ((36, 105, 50), 0.45),   # y: 210 = 14 x 15
                                    ((48, 70, 64), 2.0),     # K_y = 33 >= N/8: most outputs kept
                                    # most outputs kept on N = 2H x 16: every stage-1 output
                                    # formed by radix-16 DFTs (k_yfft_gf)
                                    ((32, 80, 32), 2.2),     # 160 = 10 x 16, K_y = 21
                                    ((32, 96, 32), 2.5),     # 192 = 12 x 16, K_y = 28
                                    ((32, 160, 32), 2.5)])   # 320 = 20 x 16, K_y = 47
def test_generic_pruned_y_against_oracle(gpu, n, qcut):
    """The output-pruned y transform for N = 2H R2 on non-cubic cells (each
    axis its own transform length), against the oracle."""
    lengths = [0.75 * v for v in n]
    box = np.diag(lengths)
    natoms = int(0.1 * np.prod(lengths))
    rng = np.random.default_rng(sum(n))
    pos = (rng.random((natoms, 3)) * np.array(lengths)).astype(np.float32)
    syms = list(rng.choice(["C", "N", "O", "H", "S"], size=natoms, p=[0.3, 0.04, 0.15, 0.5, 0.01]))
    H, C = engine_hist([pos], syms, box, n, 2.0, 0.0, qcut)
    Ho, Co = oracle_hist([pos], syms, box, n, 2.0, 0.0, qcut)
    compare(H, C, Ho, Co, label=f"generic pruned y n={n}")


@pytest.mark.parametrize("n,qcut", [((70, 64, 48), 0.6),     # x: N 140 = 14 x 10
                                    ((81, 40, 45), 0.5),     # x: 162 = 18 x 9
                                    ((105, 36, 50), 0.45),   # x: 210 = 14 x 15
                                    ((70, 64, 48), 2.0),     # K_x = 33 >= N/8: most outputs kept
                                    # most outputs kept on N = 2H x 16: every stage-1 output
                                    # formed by radix-16 DFTs (k_xfft_gf)
                                    ((80, 32, 32), 2.2),     # 160 = 10 x 16, K_x = 21
                                    ((96, 32, 32), 2.5),     # 192 = 12 x 16, K_x = 28
                                    ((160, 32, 32), 2.5)])   # 320 = 20 x 16, K_x = 47
def test_generic_pruned_x_against_oracle(gpu, n, qcut):
    """The output-pruned x transform fused with the form-factor sum, moduli
    and histogram, for N = 2H R2 on non-cubic cells, against the oracle."""
    lengths = [0.75 * v for v in n]
    box = np.diag(lengths)
    natoms = int(0.1 * np.prod(lengths))
    rng = np.random.default_rng(sum(n) + 1)
    pos = (rng.random((natoms, 3)) * np.array(lengths)).astype(np.float32)
    syms = list(rng.choice(["C", "N", "O", "H", "S"], size=natoms, p=[0.3, 0.04, 0.15, 0.5, 0.01]))
    H, C = engine_hist([pos], syms, box, n, 2.0, 0.0, qcut)
    Ho, Co = oracle_hist([pos], syms, box, n, 2.0, 0.0, qcut)
    compare(H, C, Ho, Co, label=f"generic pruned x n={n}")


def test_forward_only_plan_tile_and_backward_refusal(gpu):
    """A forward-only plan (FLAG_FORWARD_ONLY) may take a column tile that
    no z adjoint fits (4 x 8 on a 189^3 cell, 378^3 FFT, with q <= 3.0 so
    that kz = 68 exceeds the generic pruned adjoint's 64 and only the
    Stockham adjoint, too large for that tile, would remain): its forward
    equals the default plan's (the fixed-point columns do not depend on the
    tile: counts exact, H to FP64 summation order), and a backward call is
    refused before any launch with a ValueError."""
    import torch
    from pycusaxs_amd import _lib
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    n, L = 189, 141.75
    box = np.diag([L] * 3)
    pos, syms = O.synthetic_system(60000, L, seed=189)
    symbols, ids = O.element_ids(syms)
    n3, nn = resolve_grid((n,) * 3, 2.0)
    spec = SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.0, 3.0)
    xyz = torch.from_numpy(pos[None])
    idt = torch.from_numpy(ids)
    out = []
    for flags in (0, _lib.FLAG_FORWARD_ONLY):
        eng = SaxsEngine(spec, box, max_atoms=len(pos), max_frames=1, tile=(4, 8), flags=flags)
        eng.forward(xyz, idt)
        torch.cuda.synchronize()
        out.append((eng.H.cpu().numpy(), eng.C.cpu().numpy()))
        if flags:
            gH = torch.zeros((1, eng.hist.num_bins), dtype=torch.float64, device=eng.device)
            with pytest.raises(ValueError, match="does not fit the z adjoint"):
                eng.backward(xyz, idt, gH)
        eng.close()
    assert np.array_equal(out[0][1], out[1][1])
    # the spread columns agree bit for bit; the histogram's FP64 LDS sums may
    # add in another order
    assert np.allclose(out[0][0], out[1][0], rtol=1e-12, atol=0)


def test_segmented_sort_with_16bit_counters_matches_32bit(gpu):
    """A 200^3 cell with 5 elements in 4 x 4 column tiles has 12,500 keys a
    frame: the segment histogram only fits the LDS with 16-bit counters
    (segments of at most 65,535 atoms).  The same system in 4 x 8 tiles
    (6,250 keys) sorts with 32-bit counters; the fixed-point columns agree
    bit for bit, so the counts are equal and H agrees to FP64 summation order."""
    import torch
    from pycusaxs_amd import _lib
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    n, L, F = 200, 150.0, 2
    box = np.diag([L] * 3)
    pos, syms = O.synthetic_system(150000, L, seed=200)
    symbols, ids = O.element_ids(syms)
    frames = np.stack([O.frame_positions(pos, 200, f, 0.5) for f in range(F)]).astype(np.float32)
    n3, nn = resolve_grid((n,) * 3, 2.0)
    spec = SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.0, 0.3)
    out = []
    for tile in ((4, 4), (4, 8)):
        eng = SaxsEngine(spec, box, max_atoms=len(pos), max_frames=F, tile=tile,
                         flags=_lib.FLAG_FORWARD_ONLY)
        eng.forward(torch.from_numpy(frames), torch.from_numpy(ids))
        torch.cuda.synchronize()
        out.append((eng.H.cpu().numpy(), eng.C.cpu().numpy()))
        eng.close()
    assert np.array_equal(out[0][1], out[1][1])
    # the spread columns agree bit for bit; the histogram's FP64 LDS sums may
    # add in another order
    assert np.allclose(out[0][0], out[1][0], rtol=1e-12, atol=0)


def test_segmented_sort_in_key_passes_matches_single_pass(gpu):
    """32 element types on a 64^3 cell in 2 x 2 column tiles: 32,768 keys a
    frame, counted in two passes over key ranges with 16-bit counters; in
    4 x 8 tiles the 4,096 keys fit one pass of 32-bit counters.  Same
    fixed-point columns, so equal counts and H to FP64 summation order."""
    import json
    import os
    import torch
    from pycusaxs_amd import _lib
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    table = json.load(open(os.path.join(os.path.dirname(_lib.__file__), "data",
                                        "cromer_mann_it1992.json")))
    elems = [s for s in table if s.isalpha() and s != "Du"][:32]
    n, L, F = 64, 48.0, 2
    rng = np.random.default_rng(64)
    pos = rng.uniform(0.0, L, size=(40000, 3)).astype(np.float32)
    syms = list(rng.choice(elems, size=len(pos)))
    symbols, ids = O.element_ids(syms)
    assert len(symbols) == 32
    frames = np.stack([O.frame_positions(pos, 64, f, 0.5) for f in range(F)]).astype(np.float32)
    n3, nn = resolve_grid((n,) * 3, 2.0)
    spec = SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.0, 0.8)
    out = []
    for tile in ((2, 2), (4, 8)):
        eng = SaxsEngine(spec, np.diag([L] * 3), max_atoms=len(pos), max_frames=F, tile=tile,
                         flags=_lib.FLAG_FORWARD_ONLY)
        eng.forward(torch.from_numpy(frames), torch.from_numpy(ids))
        torch.cuda.synchronize()
        out.append((eng.H.cpu().numpy(), eng.C.cpu().numpy()))
        eng.close()
    assert np.array_equal(out[0][1], out[1][1])
    assert np.allclose(out[0][0], out[1][0], rtol=1e-12, atol=0)


def _lambda2(m, nn, p=4):
    """lambda(m)^2 = (G_p(x) / G_2p(x))^2, x = pi m / nn, by the finite series
    G_r(x) = 1 + sum_{k=1}^{50} (x/(x + pi k))^r + (x/(x - pi k))^r."""
    if m == 0:
        return 1.0
    x = np.pi * m / nn
    k = np.arange(1, 51)
    g = lambda r: 1.0 + np.sum((x / (x + np.pi * k)) ** r + (x / (x - np.pi * k)) ** r)
    return (g(p) / g(2 * p)) ** 2


@pytest.mark.parametrize("n,sigma,tile", [(256, 2.0, (4, 4)),    # 512, n_z = N/2, pruned
                                          (128, 4.0, (4, 4)),    # 512, n_z = N/4, pruned
                                          (256, 2.0, (0, 0))])   # 512, default plan
def test_512_grid_on_node_atoms_direct_structure_factor(gpu, n, sigma, tile):
    """FFT grid 512 pinned against mathematics rather than against the
    engine's own transforms: atoms on grid nodes whose stencils reach no
    border slab (pad 0) give, at every half-spectrum point m,
    I(m) = |sum_j f_j(q) exp(-2 pi i m.g_j / nn)|^2 prod_a lambda(m_a)^2
    (theta(0) = (1/6, 4/6, 1/6, 0) per axis is b(m) times a phase, and the
    modulus table divides |b|^2 out), so H_b / C_b is the mean of that
    direct sum over the half-spectrum points of bin b."""
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    h = 0.75                                   # cell size, exact in float32
    L = n * h
    box = np.diag([L] * 3)
    n3, nn = resolve_grid((n,) * 3, sigma)
    assert nn == (512, 512, 512)
    qcut = 0.4
    rng = np.random.default_rng(77)
    d = O.border_planes(n)
    k_at = 16
    nodes = np.stack([rng.integers(4, n, k_at), rng.integers(4, n, k_at),
                      rng.integers(d + 5, n - d, k_at)], axis=1)
    pos = (nodes * h).astype(np.float32)
    syms = [("C", "O", "N", "S")[i % 4] for i in range(k_at)]
    symbols, ids = O.element_ids(syms)
    spec = SystemSpec(n3, nn, 4, tuple(symbols), sigma, 0.0, qcut)
    eng = SaxsEngine(spec, box, max_atoms=k_at, max_frames=1, tile=tile)
    eng.forward(torch.from_numpy(pos[None]), torch.from_numpy(ids))
    torch.cuda.synchronize()
    H, C = eng.H.cpu().numpy(), eng.C.cpu().numpy()
    kept = eng.kept()
    eng.close()
    hp = O.hist_params(n3, nn, [np.float32(1.0) / np.float32(L)] * 3, sigma, 0.0, qcut)
    s = float(hp.spacing[0])
    K = int(qcut / s) + 1
    assert kept[2] <= 512 // 8
    mx = np.arange(-K, K + 1)
    mz = np.arange(0, K + 1)
    M = np.stack(np.meshgrid(mx, mx, mz, indexing="ij"), axis=-1).reshape(-1, 3)
    nsq = (M ** 2).sum(axis=1)
    M, nsq = M[nsq * s * s <= float(hp.qcut) ** 2 * (1 + 1e-9)], nsq[nsq * s * s <= float(hp.qcut) ** 2 * (1 + 1e-9)]
    bins = np.array([O.bin_of_n(int(v), hp.spacing[0], hp.dq, hp.qcut, hp.num_bins) for v in nsq])
    keep = bins >= 0
    M, nsq, bins = M[keep], nsq[keep], bins[keep]
    q = s * np.sqrt(nsq)
    phase = np.exp(-2j * np.pi * (M @ nodes.T) / 512.0)            # [points, atoms]
    ff = np.stack([O.form_factor(sy, q) for sy in syms], axis=1)   # [points, atoms]
    A = (ff * phase).sum(axis=1)
    lam = {m: _lambda2(m, 512) for m in range(-K, K + 1)}
    lam2 = np.array([lam[a] * lam[b] * lam[c] for a, b, c in M])
    I = np.abs(A) ** 2 * lam2
    Hx = np.bincount(bins, weights=I, minlength=hp.num_bins)
    Cx = np.bincount(bins, minlength=hp.num_bins).astype(float)
    sel = (np.arange(hp.num_bins) >= 1) & (Cx > 0)
    assert np.array_equal(C[sel], 2 * Cx[sel])
    rel = np.abs(H[sel] / C[sel] - Hx[sel] / Cx[sel]) / np.abs(Hx[sel] / Cx[sel])
    print(f"[512 on-node n={n} tile={tile}] max rel err {rel.max():.3e} over {sel.sum()} bins",
          flush=True)
    assert rel.max() <= REL_TOL


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_512_point_axis_forward_and_adjoint_against_oracle(gpu, axis):
    """One axis with a 512-point FFT (256 cells, sigma 2) and two short
    ones (16 cells -> 32): the 512-point pruned forward and adjoint kernels of
    that axis against the FP64 oracle and its gradient (the oracle's
    transform is then only 32 x 32 x 512)."""
    import torch
    from pycusaxs_amd.engine import SaxsEngine, SystemSpec, resolve_grid
    n3 = [16, 16, 16]
    n3[axis] = 256
    Ls = [16.0, 16.0, 16.0]
    Ls[axis] = 256.0
    box = np.diag(Ls)
    n3, nn = resolve_grid(tuple(n3), 2.0)
    assert nn[axis] == 512
    rng = np.random.default_rng(30 + axis)
    n_atoms = 600
    pos = (rng.uniform(0.0, 1.0, size=(n_atoms, 3)) * np.array(Ls)).astype(np.float32)
    syms = [("H", "C", "O")[i % 3] for i in range(n_atoms)]
    symbols, ids = O.element_ids(syms)
    qcut = 0.5
    spec = SystemSpec(n3, nn, 4, tuple(symbols), 2.0, 0.0, qcut)
    eng = SaxsEngine(spec, box, max_atoms=n_atoms, max_frames=1, tile=(4, 4))
    xyz = torch.from_numpy(pos[None])
    idt = torch.from_numpy(ids)
    Hf, Cf = eng.forward_frames(xyz, idt)
    Hn = Hf.cpu().numpy()
    gH = np.random.default_rng(5).normal(size=Hn.shape) / np.maximum(np.abs(Hn), 1.0)
    g = eng.backward(xyz, idt, torch.from_numpy(gH)).cpu().numpy()[0]
    kept = eng.kept()
    eng.close()
    assert kept[axis] < (512 if axis < 2 else 257)        # the 512-point axis is pruned
    plan = O.make_plan(n3, box, 4, 2.0, None, 0.0, qcut)
    Ho, Co = O.frame_histogram(plan, pos, ids, symbols, box)
    compare(Hn[0], Cf[0].cpu().numpy(), Ho, Co, label=f"512 on axis {axis}")
    go = O.grad_linear_H(plan, pos, ids, symbols, box, gH[0])
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    mx = np.abs(g - go).max() / np.abs(go).max()
    print(f"[512 on axis {axis}] kept {kept}, adjoint rel-L2 {rel:.2e} max {mx:.2e}", flush=True)
    assert rel <= 2e-4 and mx <= 1e-3
