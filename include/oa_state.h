/*
 * oa_state.h -- the reference's STATES files, read into and written from
 * the stored order of the step library (openatom_amd/liboa_step.so).
 *
 * A state file holds the unique half sphere of one state: a header
 * (n_points, Nx, Ny, Nz) and one (re, im, kx, ky, kz) record per point,
 * G = 0 last.  Four encodings (the ibinary_opt / ibinary_write_opt values of
 * the reference): 0 text ("%d %d %d %d" header, "%g %g %d %d %d " records),
 * 1 binary (four ints, then per point two doubles and three ints, native
 * byte order, no padding), 2 and 3 the same gzip-compressed, with ".gz"
 * appended to the path (util.C:1107-1290 readState, 1955-2090
 * writeStateFile).  Text records carry six significant digits, as the
 * reference writes them; the binary encodings are exact.
 *
 * The stored order is the one of oa_kvecs / oa_system (kx = 0 plane first,
 * its redundant half ky < 0 or ky = 0, kz < 0 first).  Reading matches the
 * file's points to it by (kx, ky, kz), whatever their order in the file,
 * and fills the redundant half with the complex conjugates of its partners
 * (flip_data_set, para_grp_parse.C:178-285, through processState,
 * util.C:1325-1365).  Writing drops the redundant half and orders the
 * points by (kx, ky, kz) with G = 0 moved last (sort_psi_output,
 * util.C:2180-2216), so the files read back into the reference.
 */
#ifndef OA_STATE_H
#define OA_STATE_H

#include "oa_step.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  OA_STATE_TEXT = 0,
  OA_STATE_BINARY = 1,
  OA_STATE_TEXT_GZ = 2,
  OA_STATE_BINARY_GZ = 3
} oa_state_format;

/* One state file into psi[npts] (complex, interleaved) in the stored order
 * kx/ky/kz; grid receives (Nx, Ny, Nz) of the header (may be NULL).
 * OA_EINVAL: unreadable or malformed file, a point outside the stored set,
 * a stored unique point missing from the file, or a header count that
 * differs from the number of records. */
oa_status oa_state_read(const char *path, int format, int npts, const int *kx, const int *ky,
                        const int *kz, double *psi, int grid[3]);

/* One state psi[npts] (stored order kx/ky/kz) to a file in the reference
 * writer's order and encoding. */
oa_status oa_state_write(const char *path, int format, int npts, const int *kx, const int *ky,
                         const int *kz, const double *psi, const int grid[3]);

/* All states of a context from / to the reference's directory layout
 * dir/Spin.<spin>_Kpt.<kpt>_Bead.<bead>_Temper.<temper>/state<i>.out,
 * i = 1 .. nstates (CP_State_GSpacePlane.C:744-760, 2555-2566).  Reading
 * loads the coefficients into the context (oa_psi_set); writing also
 * writes the TimeStamp file ("time step = <iteration>") of that directory,
 * which must exist. */
oa_status oa_states_read(oa_ctx *ctx, const char *dir, int format, int spin, int kpt, int bead,
                         int temper);
oa_status oa_states_write(oa_ctx *ctx, const char *dir, int format, int spin, int kpt, int bead,
                          int temper, int iteration);

#ifdef __cplusplus
}
#endif
#endif
