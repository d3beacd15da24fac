/*
 * pic_oracle.h -- CPU restatement of pbpic's 2D PIC step (TEST INFRASTRUCTURE).
 *
 * This is the oracle that the HIP path is checked against.  It is not part of
 * the product: only tests, the smoke check and the benchmark's CPU-baseline leg
 * load it.  It follows src/programs/pic2d.f90 of xiaocanli/pbpic line by line,
 * keeping the reference's ghosted array layout (emf(0:mx+1,0:my+1),
 * denjr(0:mx+2,0:my+2)), its stored curl components and its array-of-structs
 * particle record.  The MPI exchanges become an exchange of a rank with
 * itself (periodic in y); y-crossing particles stay in place instead of being
 * compacted and appended (pic2d.f90:1072-1146), which only reorders the array.
 *
 * Two instantiations: orc8_* computes in double (the reference built with
 * -fdefault-real-8), orc4_* in float (the reference's default `real`).  Host
 * values cross this interface as double in both.
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_COMPAT_BORIS_OLD_U 0x1
#define ORC_COMPAT_RHO_MIRROR  0x2
#define ORC_COMPAT_JY_DROP     0x4

typedef struct orc_ctx orc_ctx;

#define ORC_DECL(P)                                                            \
orc_ctx *P##_create(int nx, int ny, double dx, double dy, double dt,           \
                    double cva, double mime, double in0, int compat);          \
void P##_destroy(orc_ctx *c);                                                  \
int  P##_set_particles(orc_ctx *c, int64_t n, const int32_t *grid,             \
                       const double *dx, const double *dy, const double *ux,   \
                       const double *uy, const double *uz, const int32_t *q,   \
                       const int32_t *pid);                                    \
int64_t P##_num_particles(orc_ctx *c);                                         \
void P##_get_particles(orc_ctx *c, int32_t *grid, double *dx, double *dy,      \
                       double *ux, double *uy, double *uz, int32_t *q,         \
                       int32_t *pid);                                          \
void P##_set_field(orc_ctx *c, int which, const double *plane);                \
void P##_get_field(orc_ctx *c, int which, double *plane);                      \
void P##_zero_current(orc_ctx *c);                                             \
void P##_pusher(orc_ctx *c);                                                   \
void P##_deposit_batch(orc_ctx *c, int64_t n, const int32_t *i1,               \
                       const int32_t *i2, const int32_t *j1,                   \
                       const int32_t *j2, const double *dx1,                   \
                       const double *dx2, const double *dy1,                   \
                       const double *dy2, const int32_t *q, const double *vz); \
void P##_finish_current(orc_ctx *c);                                           \
void P##_fieldsolver(orc_ctx *c);                                              \
void P##_fieldsenergy(orc_ctx *c, double out[2]);                              \
void P##_countingsort(orc_ctx *c, int32_t *counts, int32_t *perm);             \
void P##_step(orc_ctx *c, int nsteps, int sort_every);                         \
void P##_diag(orc_ctx *c, double rho_bg, double out[16]);                      \
void P##_set_threads(orc_ctx *c, int nthreads);                                \
int  P##_set_ybound(orc_ctx *c, int kind);                                      \
int  P##_set_filters(orc_ctx *c, int marder, int cherenkov);

ORC_DECL(orc8)
ORC_DECL(orc4)

/* Seeded loader shared with the device (Philox4x32-10, Box-Muller): ncells*ppc
 * particles in cell order; grid is 1-based, y fastest.                       */
void orc_init_maxwellian(int nx, int ny, int ppc, double vth, double vdrift_x,
                         int counter_stream, uint64_t seed, uint32_t stream,
                         int32_t *grid, double *dx, double *dy, double *ux,
                         double *uy, double *uz);
/* Seeded Harris current-sheet loader (pic2d.f90:357-648); see pic_oracle.c.
 * orc_harris_count returns the particles per species and the sheet count.   */
int64_t orc_harris_count(int nx, int ny, int ywall, double dx, double dy,
                         int geometry, int ppc, const double *prm, int64_t *ncs);
void orc_init_harris(int nx, int ny, int ywall, double dx, double dy,
                     int geometry, int ppc, const double *prm, int thermal_z,
                     uint64_t seed, uint32_t stream, int32_t *grid_i,
                     double *pi, int32_t *grid_e, double *pe, double *fields);
/* Raw Philox4x32-10 block for counter ctr and key (for known-answer tests).  */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
