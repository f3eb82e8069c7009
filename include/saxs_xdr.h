/* Frame reader for GROMACS XTC and TRR trajectories (host code, no GPU).
 *
 * Replaces the reference's trajectory access through MDAnalysis:
 *   pycusaxs/topology.py:63    mda.Universe(tpr_file, xtc_file)
 *   pycusaxs/topology.py:536-619  read_frame / get_box / get_time /
 *                              get_coordinates (ts.positions in A,
 *                              ts.triclinic_dimensions in A)
 *   pycusaxs/topology.py:621-737  iter_frames_stream (frame, time,
 *                              positions, box per frame)
 * which RunSaxs.cu:48-92,137-170 drives frame by frame.  The library maps
 * the file, indexes its frames once (random access, any stride), and decodes
 * frames, several at a time on host threads, straight into caller buffers
 * (the engine's pinned staging buffers).
 *
 * Units as the reference's Topology delivers them: positions and box in
 * Angstrom (the file's nm times 10, in float32), time in ps.  The box is the
 * file's 3x3 matrix (rows = box vectors) times 10.
 *
 * Return codes: 0 ok, negative on error (saxs_xdr_last_error() describes
 * the last error of the calling thread).
 */
#ifndef SAXS_XDR_H
#define SAXS_XDR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAXS_XDR_OK 0
#define SAXS_XDR_EINVAL (-1)  /* bad argument (null pointer, frame out of range) */
#define SAXS_XDR_EIO (-2)     /* the file cannot be opened or mapped */
#define SAXS_XDR_EFORMAT (-3) /* not an XTC/TRR file, truncated or corrupt frame */

#define SAXS_XDR_TRR 1
#define SAXS_XDR_XTC 2

typedef struct saxs_xdr saxs_xdr;

/* Opens and indexes a trajectory (kind detected from the first frame's
 * magic number: 1995 XTC, 1993 TRR).  Every frame must hold the same atom
 * count; a trailing partial frame is an error. */
int saxs_xdr_open(const char* path, saxs_xdr** out);

/* kind (SAXS_XDR_TRR / SAXS_XDR_XTC), atoms per frame, frame count. */
int saxs_xdr_info(const saxs_xdr* h, int* kind, int* n_atoms, int64_t* n_frames);

/* Decodes `count` frames (indices frames[0..count)) into
 *   xyz  [count][n_atoms][3] float, Angstrom (NULL: positions not wanted)
 *   box  [count][9] float, Angstrom (NULL: not wanted)
 *   time [count] double, ps (NULL: not wanted)
 *   step [count] int64 (NULL: not wanted)
 * on up to `threads` host threads (<= 0: one; the threads are kept for the
 * handle's life).  Calls on one handle from several threads run one after
 * another.  A TRR frame without positions is an error when xyz is
 * requested. */
int saxs_xdr_read(saxs_xdr* h, const int64_t* frames, int count, float* xyz, float* box,
                  double* time, int64_t* step, int threads);

const char* saxs_xdr_last_error(void);

/* Host-side copy helper with no counterpart in the reference (its frames
 * reach the GPU through one host copy each): `threads` host threads (the
 * caller included, 1..64) kept for the copier's life split one memory copy
 * into whole 64-byte lines each.  Between copies the workers spin for about
 * 0.2 ms before they block, so a frame loop that copies every few tens of
 * microseconds wakes no thread.  saxs_copier_copy returns when the copy is
 * done; calls on one copier must not overlap. */
typedef struct saxs_copier saxs_copier;
int saxs_copier_create(int threads, saxs_copier** out);
int saxs_copier_copy(saxs_copier* c, void* dst, const void* src, size_t bytes);
void saxs_copier_destroy(saxs_copier* c);

void saxs_xdr_close(saxs_xdr* h);

#ifdef __cplusplus
}
#endif

#endif /* SAXS_XDR_H */
