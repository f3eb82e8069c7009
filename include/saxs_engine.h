/*
 * saxs_engine.h -- C ABI of the MI355X SAXS scattering engine.
 *
 * Plain pointers and sizes only.  All device pointers are HIP device memory
 * owned by the caller; the plan owns its own workspace.  Every call is
 * asynchronous on the given stream (hipStream_t passed as void*; NULL is the
 * default stream).  Functions return 0 on success or a negative code:
 *   SAXS_EINVAL  (-1) invalid argument   -> Python ValueError
 *   SAXS_ENOMEM  (-2) allocation failed  -> Python RuntimeError
 *   SAXS_EHIP    (-3) HIP runtime error  -> Python RuntimeError
 *   SAXS_EUNSUP  (-4) unsupported shape  -> Python ValueError
 *   SAXS_ECAPACITY (-5) new q tables need more workspace than the plan
 *                holds (a box change that keeps more frequencies); the
 *                plan is unchanged and a plan for the new box is needed
 *                -> Python ValueError subclass CapacityError
 * saxs_last_error() returns the calling thread's last message.
 *
 * One plan is used by one stream at a time; calls on one plan must not
 * overlap.  No global mutable state is shared between plans.
 *
 * Reference interfaces this ABI replaces (octupole/pyCuSaxs, cpp-src/):
 *   saxsKernel class and its per-run setup      Saxs/saxsKernel.h:30-114,
 *     scaledCell / createMemory / setcufftPlan    Saxs/saxsKernel.cu:240-366,470-488
 *   per-frame pipeline runPKernel               Saxs/saxsKernel.cu:59-209
 *   histogram read-out getSaxs                  Saxs/saxsKernel.cu:210-225
 *   NVT spectrum accumulation getHistogram       Saxs/saxsKernel.cu:17-39
 * The Python-visible entry point pycusaxs_cuda.run (pybind/cuda_bindings.cpp:
 * 172-195) is kept by the Python host layer on top of this ABI.
 */
#ifndef SAXS_ENGINE_H
#define SAXS_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAXS_OK 0
#define SAXS_EINVAL (-1)
#define SAXS_ENOMEM (-2)
#define SAXS_EHIP (-3)
#define SAXS_EUNSUP (-4)
#define SAXS_ECAPACITY (-5)

typedef struct saxs_plan saxs_plan;

/* Geometry and host tables of one system (one cell grid, one FFT grid).
 * Tables are copied into plan-owned device memory by saxs_plan_create.
 *
 *  n[3], nn[3]  cell grid and FFT grid (x, y, z; z is the fastest axis).
 *               nn comes from the 2,3,5,7-smooth search of
 *               saxsKernel.cu:449-469 or the caller's explicit scaled grid.
 *  order        B-spline order, 4 or 6 (Splines.cu:63-74 defines only 4).
 *  num_elements element types E; atoms carry ids 0..E-1.
 *  num_bins     histogram bins (int(qcut/dq)+1, saxsKernel.cu:247).
 *  border       z-slab half width d of the bulk-solvent border
 *               (saxsKernel.h:107-110).
 *  lut_len      entries of the q tables.  Shell keys (kmax all 0): indexed
 *               by n = |m|^2, shells n >= lut_len are outside q <= qcut.
 *               Point keys (kmax not all 0): indexed by
 *               (|m_x| (kmax[1]+1) + |m_y|) (kmax[2]+1) + |m_z| over the box
 *               |m_a| <= kmax[a]; lut_len = prod (kmax[a]+1).  Point keys are
 *               for orthorhombic cells whose reciprocal spacings differ
 *               (q^2 = sum_a (2 pi oc_aa m_a n_a/nn_a)^2, saxsDeviceKernels.cu:
 *               84-91), where q is not a function of |m|^2.  Triclinic cells
 *               (cell with off-diagonal terms) key the signed point:
 *               ((m_x+kmax[0]) (2kmax[1]+1) + m_y+kmax[1]) (2kmax[2]+1) +
 *               m_z+kmax[2], lut_len = prod (2 kmax[a]+1).
 *  bin_lut      [lut_len] bin of each key, or -1 (outside qcut/bins).
 *  ff_lut       [num_elements * lut_len] form factor f_e(q) per key (opsfact.cu:29-35).
 *  mod_x/y/z    [nn[a]] B-spline modulus tables lambda^2/|b|^2 (BSpmod.cpp:45-129).
 *  max_atoms, max_frames  capacity of the plan's workspace.
 *  tile_x, tile_y spreading column tile (0 = engine default).
 *  cell         row-major inverse cell M (fractional coordinates u = M x);
 *               all zero or diagonal: orthorhombic, the per-call oc_diag is
 *               used.  With off-diagonal terms the plan is triclinic: every
 *               call's oc_diag must equal diag(M), and the q tables use
 *               signed point keys.  The reference computes u = oc x with oc
 *               from Cell.cpp:111-126, whose inverse drops the off-diagonal
 *               terms of MDAnalysis's lower-triangular boxes; M is that
 *               formula applied to the transposed box (the true inverse).
 *  flags        SAXS_FLAG_PER_ELEMENT_SPREAD: one spreading workgroup per
 *               (tile, element) instead of one per tile for all elements;
 *               SAXS_FLAG_NO_PRUNED_Z: full Stockham z and y transforms even
 *               where the output-pruned radix-16 ones apply;
 *               SAXS_FLAG_WIDE_RADIX: Stockham transforms factor nn with the
 *               widest register radices first (12, 10, 9, 8, ...; fewer
 *               stages) instead of 8/4/2 then 3, 5, 7.  The autotuner picks
 *               among these per grid.
 *               SAXS_FLAG_FORWARD_ONLY: the column tile is fitted to the
 *               forward kernels alone; saxs_backward / saxs_backward_ex on
 *               such a plan return SAXS_EUNSUP when the z adjoint does not
 *               fit that tile (no counterpart in the reference, whose plans
 *               have no backward; for forward-only batches).
 *               SAXS_FLAG_NO_KEY_CLASSES: the counting sort keys atoms by
 *               home tile only, and each tile fetches every atom of the
 *               tiles whose stencils can reach it; by default atoms whose
 *               stencil leaves the home tile are keyed apart, so a tile
 *               fetches only atoms that reach it (identical results).
 */
#define SAXS_FLAG_PER_ELEMENT_SPREAD 1
#define SAXS_FLAG_NO_PRUNED_Z 2
#define SAXS_FLAG_WIDE_RADIX 4
#define SAXS_FLAG_FORWARD_ONLY 8
#define SAXS_FLAG_NO_KEY_CLASSES 16
typedef struct saxs_desc {
  int32_t n[3];
  int32_t nn[3];
  int32_t order;
  int32_t num_elements;
  int32_t num_bins;
  int32_t border;
  int32_t lut_len;
  const int32_t* bin_lut;
  const float* ff_lut;
  const float* mod_x;
  const float* mod_y;
  const float* mod_z;
  int32_t max_atoms;
  int32_t max_frames;
  int32_t tile_x;
  int32_t tile_y;
  int32_t flags;
  int32_t kmax[3];
  float cell[9];
} saxs_desc;

/* Creates a plan: copies tables, builds FFT radix plans and twiddles,
 * allocates the workspace for max_frames x max_atoms.  Replaces
 * saxsKernel::scaledCell/createMemory/setcufftPlan (saxsKernel.cu:240-366,
 * 470-488; saxsKernel.h:55-58). */
int saxs_plan_create(const saxs_desc* desc, saxs_plan** out);

/* Replaces the shell tables (bin_lut, ff_lut) after a box change; lut_len
 * may differ from the plan's (kept-frequency ranges are recomputed).  The
 * workspace holds tables up to the point box (K+2)^3 and one more kept
 * frequency per axis than the creating box; beyond that the call returns
 * SAXS_ECAPACITY and leaves the plan as it was.  Synchronous. */
int saxs_plan_set_tables(saxs_plan* plan, int32_t lut_len, const int32_t* bin_lut,
                         const float* ff_lut);

/* saxs_plan_set_tables for point keys over the box |m_a| <= kmax[a] (host
 * int32 [3]) and the cell of the new box (host float [9] as in saxs_desc, or
 * NULL for orthorhombic); lut_len must be prod (kmax[a]+1), or
 * prod (2 kmax[a]+1) for a triclinic cell.  Switches the plan to point keys
 * (saxs_plan_set_tables switches it back to shell keys and an orthorhombic
 * cell).  Replaces the per-frame histogram setup for a new box,
 * saxsKernel.cu:489-526. */
int saxs_plan_set_point_tables(saxs_plan* plan, const float* cell, const int32_t* kmax,
                               int32_t lut_len, const int32_t* bin_lut, const float* ff_lut);

/* Bulk-solvent pad mode.  pad == NULL: the pad of each (frame, element) is
 * the mean density over the border z-slabs (padding::avg, the reference's
 * only live mode: Options.cpp:13, saxsDeviceKernels.cu:446-465,
 * saxsKernel.cu:148-161).  Otherwise pad[num_elements] (host, atoms per
 * cell voxel, each in [0, 2^31 / n_z]) is used as a constant for every frame
 * (padding::given, declared in Options.h:7-11 with its values from
 * AtomCounter::calculateAtomCounts, AtomCounter.cpp:77-103); the adjoint
 * then has no pad-mean term.  Synchronous; applies to later calls. */
int saxs_plan_set_padding(saxs_plan* plan, const double* pad);

/* Forward scattering of n_frames frames of one system, accumulated into the
 * caller's histogram: H[b] += sum over frames of the S6 weighted intensity,
 * C[b] += matching point counts (NPT mode, saxsDeviceKernels.cu:67-125).
 *   xyz     device float32 [n_frames][n_atoms][3], Angstrom
 *   elem    device int32 [n_atoms], element id per atom
 *   oc_diag host float32 [3], diagonal of the inverse box (1/L_a)
 *   H, C    device float64 [num_bins]
 * Replaces saxsKernel::runPKernel (saxsKernel.cu:59-209) for whole batches. */
int saxs_forward(saxs_plan* plan, const float* xyz, const int32_t* elem, int32_t n_atoms,
                 int32_t n_frames, const float* oc_diag, double* H, double* C,
                 void* stream);

/* saxs_forward that allocates nothing (no counterpart in the reference: for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * per-call buffer (sort keys and records, z and y spectra, block
 * histograms) comes from one file-scope device array of 128 MiB that each
 * GPU holds from the library's load until the process ends; the plan
 * supplies only its tables, and its max_atoms / max_frames do not bound the
 * call.  A call whose buffers exceed 128 MiB returns SAXS_EINVAL (about 5
 * frames of 100,000 atoms at a 128^3 cell, q <= 0.6; call with fewer
 * frames).  Results equal saxs_forward's.  All calls on one GPU share the
 * array: calls that overlap on one GPU (other streams, other threads, other
 * plans) race and must be serialised by the caller. */
int saxs_forward_static(saxs_plan* plan, const float* xyz, const int32_t* elem,
                        int32_t n_atoms, int32_t n_frames, const float* oc_diag, double* H,
                        double* C, void* stream);

/* Frames whose boxes differ, in one launch chain (NPT trajectories, where
 * the box changes every frame; NVT with the first frame's binning).
 * Replaces the per-frame runPKernel(..., oc) of RunSaxs.cu:180-193 with
 * each frame's own inverse cell (saxsDeviceKernels.cu:355-361) and q
 * (:84-91, 196-205), dq / qcut / bin count fixed by the plan (those of the
 * first frame, saxsKernel.cu:489-526).  The per-frame q tables are built on
 * the device (exact bin decisions on the float32 values).
 *   cells      host [n_frames][9] row-major inverse cell M of each frame
 *   qgeom      host [n_frames][9] q geometry of each frame: key_type 0 (all
 *              reciprocal spacings equal): [0] = s; 1 (orthorhombic): [0..2]
 *              = s_a; 2 (triclinic): rows b_aj of q = sum_a m_a b_a
 *   bin_qgeom  host [9] geometry the points are binned with (NVT: the first
 *              frame's), or NULL: each frame's own
 *   key_type, kmax, lut_len  the table keys as for the plan's tables
 *              (shell n = |m|^2 < lut_len; or the point box kmax), covering
 *              every frame of the batch
 *   dq, qcut   float32 bin width and cutoff of the run
 *   form_coef  host [num_elements][9] Cromer-Mann a1..a4, b1..b4, c
 *              (Scattering.cpp:9-189)
 * The nonzero entries of a geometry must span at most 2^20 (the exact
 * comparisons stay within 128 bits).  Accumulates into H, C like
 * saxs_forward.  SAXS_ECAPACITY when the batch keeps more frequencies than
 * the plan's workspace holds (make the plan with the largest box).  The
 * first call allocates the per-frame tables (max_frames x the plan's table
 * capacity) and a pinned staging buffer; a call waits for the previous
 * call's upload before reusing the staging buffer. */
typedef struct saxs_frame_boxes {
  const float* cells;
  const float* qgeom;
  const float* bin_qgeom;
  int32_t key_type;
  int32_t kmax[3];
  int32_t lut_len;
  float dq;
  float qcut;
  const float* form_coef;
} saxs_frame_boxes;
int saxs_forward_boxes(saxs_plan* plan, const float* xyz, const int32_t* elem, int32_t n_atoms,
                       int32_t n_frames, const saxs_frame_boxes* boxes, double* H, double* C,
                       void* stream);

/* Per-frame histograms instead of an accumulated one: Hf, Cf are device
 * float64 [n_frames][num_bins] and are overwritten.  Used by the adjoint
 * (per-frame binned intensities).  On a plan with a backward (no
 * SAXS_FLAG_FORWARD_ONLY) whose x pass is one of the output-pruned ones
 * (power-of-two, generic, or generic with most outputs kept), it also keeps
 * that pass's amplitudes at the kept outputs for a following
 * SAXS_BWD_REUSE_FORWARD backward (a buffer of the plan, allocated on first
 * use and grown when a new box keeps more frequencies: at most
 * max_frames x ky x nchx x E bytes with E = 16 nn_x x 8 for the
 * power-of-two pass, 4096 x 8 for the generic pass with most outputs kept
 * and 2048 x 8 for the other generic one). */
int saxs_forward_frames(saxs_plan* plan, const float* xyz, const int32_t* elem,
                        int32_t n_atoms, int32_t n_frames, const float* oc_diag,
                        double* Hf, double* Cf, void* stream);

/* Adjoint of saxs_forward_frames: given gH = dL/dH per frame
 * (device float64 [n_frames][num_bins]), writes dL/dx per atom into grad
 * (device float32 [n_frames][n_atoms][3]; overwritten).  Recomputes the
 * forward spectra internally (no stored state between calls).  Repeated
 * calls on the same inputs return the same bits (no floating-point atomics).
 * No reference counterpart: pyCuSaxs has no gradient. */
int saxs_backward(saxs_plan* plan, const float* xyz, const int32_t* elem, int32_t n_atoms,
                  int32_t n_frames, const float* oc_diag, const double* gH, float* grad,
                  void* stream);

/* saxs_backward with options.  SAXS_BWD_REUSE_FORWARD: skip the forward
 * recomputation and use the sorted atoms and y-pass spectra (and the x-pass
 * amplitudes, where saxs_forward_frames kept them) left in the plan's
 * workspace by the immediately preceding saxs_forward_frames call on this
 * plan, which must have had the same positions, elements, n_atoms, n_frames
 * and box (the caller guarantees it; nothing else may run on the plan in
 * between). */
#define SAXS_BWD_REUSE_FORWARD 1
int saxs_backward_ex(saxs_plan* plan, const float* xyz, const int32_t* elem, int32_t n_atoms,
                     int32_t n_frames, const float* oc_diag, const double* gH, float* grad,
                     int32_t flags, void* stream);

/* Reference-structured per-frame path: per element, spread into the n^3
 * grid with global atomics, border mean, explicit embed into nn^3, full
 * 3-D R2C FFT, form-factor accumulate, modulus, histogram -- the structure
 * of saxsKernel::runPKernel (saxsKernel.cu:59-209) on this engine's
 * kernels.  One frame per call.  Kept as the per-frame baseline. */
int saxs_forward_plain(saxs_plan* plan, const float* xyz, const int32_t* elem,
                       int32_t n_atoms, const float* oc_diag, double* H, double* C,
                       void* stream);

/* Kernel ids of the measurement hook. */
#define SAXS_K_CLEAR 0
#define SAXS_K_PREP 1
#define SAXS_K_SCAN 2
#define SAXS_K_SCATTER 3
#define SAXS_K_SPREAD_Z 4
#define SAXS_K_YFFT 5
#define SAXS_K_XFFT 6
#define SAXS_K_REDUCE 7
#define SAXS_K_ADJ_X 8
#define SAXS_K_ADJ_Y 9
#define SAXS_K_ADJ_Z 10
#define SAXS_K_GRAD_SUM 11
#define SAXS_K_PLAIN 12
#define SAXS_NKERNELS 13

/* Measurement hook (no reference counterpart; the reference times whole
 * frames with cudaEvents, saxsKernel.cu:64-67,197-208).  When enabled,
 * every launch of the plan is bracketed by a pair of hipEvents on its
 * stream.  saxs_plan_kernel_times waits for the recorded events, adds the
 * elapsed milliseconds and launch counts per kernel id into ms[] and
 * launches[] (SAXS_NKERNELS entries each) and forgets the events. */
int saxs_plan_set_timing(saxs_plan* plan, int32_t enable);
int saxs_plan_kernel_times(saxs_plan* plan, double* ms, int64_t* launches);

/* Workspace bytes held by the plan (device memory). */
int64_t saxs_plan_workspace_bytes(const saxs_plan* plan);

/* Kept frequency counts (x, y, z) of the pruned transform, for measurement. */
int saxs_plan_kept(const saxs_plan* plan, int32_t* kx, int32_t* ky, int32_t* kz);

/* Frees the plan and its device memory (saxsKernel::~saxsKernel,
 * saxsKernel.cu:546-553). */
int saxs_plan_destroy(saxs_plan* plan);

/* Thread-local message of the last failing call ("" if none). */
const char* saxs_last_error(void);

/* ABI version, for the loader's compatibility check. */
int32_t saxs_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SAXS_ENGINE_H */
