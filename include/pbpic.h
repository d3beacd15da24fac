/*
 * pbpic.h -- C ABI of the MI355X-native 2D electromagnetic PIC step.
 *
 * This library replaces the hot path of xiaocanli/pbpic's Fortran program
 * (src/programs/pic2d.f90): the per-step subroutines `pusher`,
 * `currentdeposit`, `fieldsolver`, `fieldsenergy` and `countingsort`, plus the
 * MPI particle / current / field exchange they contain, which becomes an
 * in-device periodic wrap on one GPU.
 *
 * Conventions
 *   - Every entry point returns an int status: 0 = PBPIC_OK, negative = error
 *     (see PBPIC_ERR_*). The library never aborts; pbpic_last_error() gives
 *     a message.  The reference ignores every MPI/HDF5 error code and aborts
 *     through MPI's default handler (pic2d.f90, all `ierr` arguments).
 *   - Only plain C types cross the boundary.  Host arrays are borrowed for the
 *     duration of the call and copied in or out; device buffers are owned by
 *     the context.
 *   - Work is stream-ordered on the context's HIP stream.  Calls that return
 *     data to the host synchronize; pbpic_sync() waits and reports device-side
 *     error flags (CFL violation, non-finite values, bad cell index).
 *   - One context per host thread.  Two contexts never share buffers.
 *
 * Index conventions (shared with the reference)
 *   - Cells: nx x ny, periodic in x and y.  The reference's particle cell id is
 *     `grid = my*(ix-1) + iy` with 1-based ix, iy (pic2d.f90:422, 938-939):
 *     y runs fastest.  The same 1-based `grid` is used at this boundary.
 *   - In-cell offsets dx, dy lie in (-1/2, 1/2] (pic2d.f90:420-421, 1008-1023).
 *   - With params.ybound = PBPIC_YBOUND_CONDUCTING the last two rows of the
 *     nx x ny lattice are guard rows (see PBPIC_YBOUND_CONDUCTING); the
 *     domain is rows 1..ny-2 and particles must lie there.
 *   - Field planes at this boundary are nx*ny arrays, x fastest, i.e. the
 *     owned region emf(1:mx,1:my) of the reference's Fortran arrays
 *     (pic2d.f90:593, 221).  Staggering: Ex,Bx,Jx at (i-1/2, j-1);
 *     Ey,By,Jy at (i-1, j-1/2); Ez,Bz,Jz at (i-1/2, j-1/2); rho and the
 *     node-averaged emfgrid at (i-1, j-1), in units of cells with node (1,1)
 *     at the origin.
 */
#ifndef PBPIC_H
#define PBPIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define PBPIC_OK              0
#define PBPIC_ERR_ARG        -1  /* invalid argument */
#define PBPIC_ERR_HIP        -2  /* HIP runtime error */
#define PBPIC_ERR_CAPACITY   -3  /* particle count exceeds species capacity */
#define PBPIC_ERR_CFL        -4  /* a particle moved more than one cell in a step */
#define PBPIC_ERR_NONFINITE  -5  /* NaN/Inf in a particle momentum */
#define PBPIC_ERR_CELL       -6  /* particle cell index out of range */
#define PBPIC_ERR_STATE      -7  /* call not valid in the current state */
#define PBPIC_ERR_NOMEM      -8  /* device allocation failed */
#define PBPIC_ERR_SPIN       -9  /* sort look-back exceeded its spin limit */

/* ------------------------------------------------------------- precision */
#define PBPIC_PREC_FP64     0  /* FP64 particles, fields, currents */
#define PBPIC_PREC_SPLIT32  1  /* int cell + FP32 offset/momenta, FP32 push,
                                  deposit and FDTD increments; FP64 field
                                  state and FP64 reductions */

/* -------------------------------------------------- reference behaviours
 * Each bit reproduces one behaviour of pic2d.f90 that a textbook scheme would
 * not have.  PBPIC_COMPAT_REFERENCE sets all of them (the parity default).  */
#define PBPIC_COMPAT_BORIS_OLD_U  0x1  /* Boris rotation uses the pre-kick u
                                          (pic2d.f90:985-987) instead of u-  */
#define PBPIC_COMPAT_RHO_MIRROR   0x2  /* rho CIC weights mirrored
                                          (pic2d.f90:1285-1295 vs 948-951)   */
#define PBPIC_COMPAT_JY_DROP      0x4  /* jy deposited at column mx+2 is lost
                                          in the x fold (pic2d.f90:1156,1159)*/
#define PBPIC_COMPAT_REFERENCE    0x7

/* ------------------------------------------------------------------ flags */
#define PBPIC_FLAG_DEPOSIT_RHO  0x1  /* deposit rho at time n during the push
                                        (pic2d.f90:1283-1295)               */
#define PBPIC_FLAG_TRACK_PID    0x2  /* keep a per-particle id array that the
                                        sort permutes (ptl%pid, :47-50)     */
#define PBPIC_FLAG_PUSH_ORDER   0x4  /* sorts order the cells by 8 x 8
                                        supertiles along a Hilbert curve
                                        (cells x-major inside each) and lay
                                        every full 4096-particle tile out in
                                        the push kernel's lane order; the
                                        returned permutation describes that
                                        order, the counts stay per cell.
                                        Clear it for the reference's exact
                                        order (:1322-1327).  Set by
                                        pbpic_params_default               */
#define PBPIC_FLAG_MARDER       0x8  /* divergenceclean after each E update
                                        (pic2d.f90:867-885, commented out at
                                        :754): E += d grad(div E^n -
                                        c^2 rho^n), d = 0.001 dt; needs
                                        PBPIC_FLAG_DEPOSIT_RHO              */
#define PBPIC_FLAG_CHERENKOV    0x10 /* suppresscerenkov after each E update
                                        (:890-907, commented out at :755):
                                        E -= c^2 dt tau curl curl E^n,
                                        tau = dt/8.  Both filters use the
                                        time-n quantities the reference's
                                        calls see, on the whole periodic
                                        lattice (the reference's ghost rows
                                        and columns are stale there)        */

/* ------------------------------------------------------- y boundary kinds
 * The reference is periodic in y (the MPI ring of pic2d.f90:295-298 closes on
 * itself); its perfectly conducting alternative is left commented out at
 * :850-859.  PBPIC_YBOUND_CONDUCTING adds that boundary as new physics:
 * perfectly conducting walls at y = 0 and y = (ny-2) dy with specular
 * particle reflection.  The lattice keeps its periodic storage and the last
 * two rows become guard rows (0-based rows ny-2 and ny-1 = row -1), which
 * hold the mirror images of the fields across the nearer wall: Ey and Bz
 * even, Ez and By odd across half rows; Ex is zero on the walls (rows 0 and
 * ny-2) and Ex, Bx of row ny-1 mirror row 1 (odd, even).  Particles reaching
 * a wall are reflected (offset and u_y change sign) before their deposit;
 * Jz deposited in a guard row is folded back with its image sign.  Field
 * energies weight the two wall rows by 1/2; the Gauss residual covers the
 * rows 1..ny-3 between the walls.  The adjoint covers the walls: its field
 * gradients are with respect to the domain rows (guard rows zero).           */
#define PBPIC_YBOUND_PERIODIC    0
#define PBPIC_YBOUND_CONDUCTING  1

/* ------------------------------------------------------------ field names */
#define PBPIC_EX   0
#define PBPIC_EY   1
#define PBPIC_EZ   2
#define PBPIC_BX   3
#define PBPIC_BY   4
#define PBPIC_BZ   5
#define PBPIC_JX   6   /* normalized current (x in0), as after pusher :1066 */
#define PBPIC_JY   7
#define PBPIC_JZ   8
#define PBPIC_RHO  9   /* normalized charge density (x in0)                 */
#define PBPIC_GEX  10  /* node-averaged emfgrid (pic2d.f90:837-845), get only */
#define PBPIC_GEY  11
#define PBPIC_GEZ  12
#define PBPIC_GBX  13
#define PBPIC_GBY  14
#define PBPIC_GBZ  15

/* ----------------------------------------------------- diagnostics slots */
#define PBPIC_D_KE_E        0  /* electron kinetic energy of the current
                                  momenta, sum (gamma-1)*ms (totene(1,:),
                                  :1061; after a push, that push's tally)   */
#define PBPIC_D_KE_I        1  /* ion kinetic energy (totene(2,:))          */
#define PBPIC_D_WE          2  /* electric energy (totene(3,:), :653-673)    */
#define PBPIC_D_WB          3  /* magnetic energy (totene(4,:))              */
#define PBPIC_D_JE_STEP     4  /* (mime/c^2) sum J.E_avg of the last field
                                  solve; energy per unit time               */
#define PBPIC_D_JE_INT      5  /* time integral of D_JE_STEP since creation */
#define PBPIC_D_GAUSS_L2    6  /* || divE - c^2 in0 (rho_cic + rho_bg) ||_2  */
#define PBPIC_D_GAUSS_LINF  7  /* same, max norm                            */
#define PBPIC_D_GAUSS_DL2   8  /* L2 of the residual's change since the first
                                  pbpic_diag call                           */
#define PBPIC_D_GAUSS_DLINF 9  /* max norm of that change                   */
#define PBPIC_D_STEPS       10 /* steps taken                               */
#define PBPIC_D_TILES_SPILLED 11 /* push tiles (4096 particles) whose window
                                    exceeded even the largest LDS window
                                    (one workgroup per CU) and deposited
                                    with global atomics, since creation: a
                                    performance counter, zero in normal
                                    runs                                    */
#define PBPIC_D_SORTS_RADIX 12 /* species sorts since creation that took the
                                  radix path: every sort not right after a
                                  push (no recorded tile boxes), one whose
                                  per-cell counts are requested, or one with
                                  a tile box above 4096 cells; the others
                                  take the tile-box path.  A performance
                                  counter                                  */
#define PBPIC_NDIAG         16

typedef struct pbpic_ctx pbpic_ctx;

/* Run parameters: the runtime replacement for the compile-time constants of
 * `module parameters` (pic2d.f90:4-142).                                     */
typedef struct pbpic_params {
    int32_t nx, ny;          /* mx, my: cells, both periodic (:12, :206-212) */
    int32_t precision;       /* PBPIC_PREC_*                                 */
    int32_t compat;          /* PBPIC_COMPAT_* bits                          */
    double  dx, dy;          /* cell size (:15)                              */
    double  dt;              /* time step (:42)                              */
    double  cva;             /* c / v_A (:20)                                */
    double  mime;            /* m_i / m_e (:19)                              */
    double  in0;             /* 1/(n0 dx dy) (:31)                           */
    double  rho_background;  /* uniform charge per cell of an immobile
                                background, used only by the Gauss residual */
    int32_t sort_every;      /* countingsort cadence in pbpic_step (:180);
                                0 = never                                    */
    int32_t flags;           /* PBPIC_FLAG_* bits                            */
    int32_t ybound;          /* PBPIC_YBOUND_*; 0 = periodic as the reference */
    int32_t reserved[7];     /* must be zero                                 */
} pbpic_params;

/* Fill *p with the reference's own constants (pic2d.f90:12-42): 256x256,
 * dx=dy=0.04, cva=15, mime=25, in0=1/291, dt=0.045/mime, sort every 50,
 * PBPIC_COMPAT_REFERENCE, FP64.                                              */
void pbpic_params_default(pbpic_params *p);

/* ---------------------------------------------------------- context life
 * Replaces the allocation half of `init` (pic2d.f90:197-308) and
 * `releasememory` (:2557-2582).                                              */
int pbpic_ctx_create(const pbpic_params *p, pbpic_ctx **out);
int pbpic_ctx_destroy(pbpic_ctx *ctx);
const char *pbpic_last_error(const pbpic_ctx *ctx);
int pbpic_sync(pbpic_ctx *ctx);

/* Add a particle species.  q = +1 (ion, mass mime) or -1 (electron, mass 1),
 * the reference's per-particle charge `ptl%q` and mass rule (:964-969).
 * capacity = maximum number of particles.  *species receives its index.
 * Device memory per particle of capacity: 2 x 44 B (FP64) or 2 x 24 B
 * (split) of particle arrays (+2 x 4 B with PBPIC_FLAG_TRACK_PID), and in
 * the split form a further 24 B (the sort's particle records); plus 16 B of
 * sort scratch per particle of the largest species.                        */
int pbpic_add_species(pbpic_ctx *ctx, int32_t q, int64_t capacity,
                      int32_t *species);
int pbpic_num_particles(pbpic_ctx *ctx, int32_t species, int64_t *n);

/* Host particle I/O, structure of arrays, FP64 or FP32 host values (the
 * device form follows params.precision).  grid is the reference's 1-based
 * cell id; pid may be NULL (then ids 1..n are assigned if PID tracking is on).
 * Mirrors the fields of `type particle` (pic2d.f90:47-50).                   */
int pbpic_set_particles(pbpic_ctx *ctx, int32_t species, int64_t n,
                        const int32_t *grid, const double *dx,
                        const double *dy, const double *ux, const double *uy,
                        const double *uz, const int32_t *pid);
int pbpic_get_particles(pbpic_ctx *ctx, int32_t species, int64_t cap,
                        int64_t *n, int32_t *grid, double *dx, double *dy,
                        double *ux, double *uy, double *uz, int32_t *pid);
int pbpic_set_particles_f32(pbpic_ctx *ctx, int32_t species, int64_t n,
                            const int32_t *grid, const float *dx,
                            const float *dy, const float *ux, const float *uy,
                            const float *uz, const int32_t *pid);
int pbpic_get_particles_f32(pbpic_ctx *ctx, int32_t species, int64_t cap,
                            int64_t *n, int32_t *grid, float *dx, float *dy,
                            float *ux, float *uy, float *uz, int32_t *pid);

/* Seeded device-side loader: ppc particles per cell, uniform in each cell,
 * momenta u = gamma v / c drawn from a Maxwellian of thermal speed vth (in
 * units of c, converted as vth/sqrt(1-vth^2) like pic2d.f90:380-384) plus a
 * drift u_d = vd/sqrt(1-vd^2) along x; if counter_stream is non-zero the
 * drift sign alternates between consecutive particles of a cell (two beams).
 * Philox4x32-10 keyed by (seed, stream).  Replaces the /dev/urandom-seeded
 * initparticles (:357-578) for the benchmark configurations; results are in
 * cell order.                                                                */
int pbpic_init_maxwellian(pbpic_ctx *ctx, int32_t species, int32_t ppc,
                          double vth, double vdrift_x, int32_t counter_stream,
                          uint64_t seed, uint32_t stream);

/* Seeded device-side Harris current-sheet loader (initparticles and
 * initfields, pic2d.f90:357-648): fills an ion and an electron species with
 * ppc*nx*(domain rows) particles each and sets E = 0 and the sheet's B.
 * Sheet particles have density n0 sech^2(s/L) across the sheet (s: distance
 * from its centre, drawn by the inverse CDF s = L/2 ln(u/(1-u)), :395-402) and
 * the rest is a uniform background; ion i and electron i share their cell and
 * their offset across the sheet (:418-428); thermal momenta as in
 * pbpic_init_maxwellian, drifts along z.
 *   geometry PBPIC_HARRIS_DOUBLE_X: the reference's two sheets across x at
 *     x = Lx/4 (ions +vdi_z, electrons +vde_z) and x = 3Lx/4 (signs
 *     reversed), By = B0 (tanh((x-Lx/4)/L) - tanh((x-3Lx/4)/L) - 1)
 *     (:606-614); periodic y required.
 *   geometry PBPIC_HARRIS_SINGLE_Y: one sheet across y at y = Ly/2,
 *     Bx = B0 tanh((y-Ly/2)/L), plus the flux perturbation
 *     psi = psi0 cos(2 pi x/Lx) cos(pi (y-Ly/2)/Ly) applied as the discrete
 *     curl of psi at cell centres (so div B = 0 to rounding); requires
 *     PBPIC_YBOUND_CONDUCTING.  Its current is -B0/L sech^2 along z, so an
 *     equilibrium has vdi_z < 0 < vde_z.
 * The sheet fraction of the particles is 2 L nsheets / (2 L nsheets + nb L_n)
 * (L_n: domain length across the sheets), rounded up to an even count as at
 * :375-376; lengths in the units of dx, dy.  Philox4x32-10 keyed by seed,
 * counters (particle, stream, block).                                        */
#define PBPIC_HARRIS_DOUBLE_X  0
#define PBPIC_HARRIS_SINGLE_Y  1
typedef struct pbpic_harris {
    int32_t geometry;        /* PBPIC_HARRIS_*                               */
    int32_t ppc;             /* mean particles per cell and species (ppg :30) */
    double  half_width;      /* L (dcs, :24)                                  */
    double  nb;              /* background / peak density (nbgcs, :29)        */
    double  b0;              /* field amplitude (1 in the reference)          */
    double  psi0;            /* perturbation amplitude (SINGLE_Y only)        */
    double  vthe_cs, vthi_cs;/* sheet thermal speeds / c (:377-379)           */
    double  vthe_bg, vthi_bg;/* background thermal speeds / c (:378-380)      */
    double  vde_z, vdi_z;    /* sheet drift speeds / c along z, first sheet
                                (the reference: -vdrift, +vdrift, :440-447)  */
    int32_t thermal_z;       /* 0: sheet u_z is the drift alone and background
                                u_z = 0, as the reference (:441-447, :523-529);
                                1: Maxwellian u_z about the drift            */
    int32_t reserved0;
    uint64_t seed;
    uint32_t stream;
    int32_t reserved[7];     /* must be zero                                 */
} pbpic_harris;
int pbpic_init_harris(pbpic_ctx *ctx, int32_t ion_species,
                      int32_t electron_species, const pbpic_harris *h);

/* Host field I/O, nx*ny planes, x fastest, FP64 host values.  Setting J or
 * RHO replaces the accumulated current (normalized values, as after the
 * in0 scaling at pic2d.f90:1066-1069).  With conducting y walls, setting a
 * field plane also sets its guard rows to the wall images.                   */
int pbpic_set_field(pbpic_ctx *ctx, int32_t which, const double *plane);
int pbpic_get_field(pbpic_ctx *ctx, int32_t which, double *plane);

/* ------------------------------------------------------- the step kernels */

/* denjr = 0 at the top of the time loop (pic2d.f90:173-176).                */
int pbpic_zero_current(pbpic_ctx *ctx);

/* `pusher` (pic2d.f90:912-1217): for every particle, CIC gather from the
 * node-averaged fields, relativistic Boris push, move (at most one cell),
 * Umeda zigzag deposit of Jx, Jy, CIC Jz (and rho if enabled) through the
 * same arithmetic as `currentdeposit` (:1226-1296), periodic wrap in x and y
 * (the MPI exchange of :1085-1216 with the rank as its own neighbour).
 * Accumulates into J.  The kinetic-energy tally the reference keeps in the
 * pusher (:1061-1062) is computed on demand from the pushed momenta
 * (pbpic_kineticenergy, pbpic_diag): the same values, no per-step cost.      */
int pbpic_push(pbpic_ctx *ctx);

/* `currentdeposit` (pic2d.f90:1226-1296), batched over n host-array entries.
 * i1,j1 are a particle's old 1-based cell, i2,j2 its new unwrapped cell
 * (|i2-i1|<=1, |j2-j1|<=1), dx1..dy2 its old/new offsets, q its charge and
 * vz = uz*c/gamma.  Accumulates into the context's J (and rho if enabled).  */
int pbpic_deposit(pbpic_ctx *ctx, int64_t n, const int32_t *i1,
                  const int32_t *i2, const int32_t *j1, const int32_t *j2,
                  const double *dx1, const double *dx2, const double *dy1,
                  const double *dy2, const int32_t *q, const double *vz);

/* pbpic_deposit without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It has no counterpart in the reference.  The inputs are
 * staged in a fixed per-GPU scratch array of PBPIC_DEPOSIT_NOALLOC_MAX
 * entries (3.9 MB, held until the process ends) instead of a buffer
 * allocated per call; larger n fails with PBPIC_ERR_CAPACITY.  Every context
 * on one GPU shares that array: calls that overlap on one GPU (contexts used
 * from two host threads at once) race.  Results equal pbpic_deposit's.      */
#define PBPIC_DEPOSIT_NOALLOC_MAX 65536
int pbpic_deposit_noalloc(pbpic_ctx *ctx, int64_t n, const int32_t *i1,
                          const int32_t *i2, const int32_t *j1,
                          const int32_t *j2, const double *dx1,
                          const double *dx2, const double *dy1,
                          const double *dy2, const int32_t *q,
                          const double *vz);

/* `fieldsolver` (pic2d.f90:678-860): half-step B, full-step E with -J,
 * half-step B, all on the periodic Yee lattice.  Also accumulates
 * J.E (PBPIC_D_JE_STEP / PBPIC_D_JE_INT).                                    */
int pbpic_fieldsolve(pbpic_ctx *ctx);

/* `fieldsenergy` (pic2d.f90:653-673): out[0] = electric, out[1] = magnetic
 * energy of the node-averaged fields, FP64 reduction.                        */
int pbpic_fieldsenergy(pbpic_ctx *ctx, double out[2]);

/* totene(1:2) of `pusher` (pic2d.f90:1061-1062): out[0] = electron, out[1] =
 * ion kinetic energy, sum (gamma-1)*ms over the current momenta (gamma at
 * n+1/2 after a push), FP64 reduction.                                       */
int pbpic_kineticenergy(pbpic_ctx *ctx, double out[2]);

/* `countingsort` (pic2d.f90:1303-1329) of every species: stable sort of the
 * particles by cell id.                                                      */
int pbpic_countingsort(pbpic_ctx *ctx);

/* The same for one species, optionally returning the per-cell counts
 * (ptlalloc after :1310-1313; nx*ny entries indexed by grid-1) and the
 * permutation (perm[k] = 0-based index before the sort of the particle now at
 * position k).  Either output may be NULL.                                   */
int pbpic_countingsort_species(pbpic_ctx *ctx, int32_t species,
                               int32_t *counts, int32_t *perm);

/* The driver loop body (pic2d.f90:171-188) nsteps times: zero J, push,
 * field solve, and countingsort whenever the step count is a multiple of
 * params.sort_every.                                                         */
int pbpic_step(pbpic_ctx *ctx, int32_t nsteps);

/* FP64 diagnostics, PBPIC_NDIAG slots (see PBPIC_D_*).                       */
int pbpic_diag(pbpic_ctx *ctx, double out[PBPIC_NDIAG]);

/* `energyspectra` (pic2d.f90:1957-1990) of one species: counts[k] (nbins
 * entries, nbins <= 4096) = the particles whose kinetic energy
 * ene = (gamma-1) ms, computed in FP64 as at :1966-1983, has
 * ceil(ene/de) = k+1 and ene < emax.  The reference uses de = 0.005,
 * emax = 5 for electrons and de = 0.025, emax = 25 for ions (:136-140).
 * ene = 0 lands in bin 0 (the reference indexes its array at 0 there) and a
 * quotient rounded past the last bin in the last.                            */
int pbpic_energy_spectrum(pbpic_ctx *ctx, int32_t species, int32_t nbins,
                          double de, double emax, int32_t *counts);

/* `numdensity` (pic2d.f90:1343-1354): particles of all species per cell,
 * nx*ny counts indexed by grid-1.                                            */
int pbpic_number_density(pbpic_ctx *ctx, int32_t *counts);

/* Run counters for checkpoint/restart: the step count (itime; it sets the
 * sort cadence) and the J.E accumulators je[0] = the last field solve's raw
 * sum, je[1] = the time integral before it (PBPIC_D_JE_INT is
 * je[1] + je[0] mime/c^2 dt).  pbpic_set_counters restores them.            */
int pbpic_get_counters(pbpic_ctx *ctx, int64_t *steps, double je[2]);
int pbpic_set_counters(pbpic_ctx *ctx, int64_t steps, const double je[2]);

/* Device buffers of a species (for the PyTorch binding).  ptrs receives 7
 * device pointers: cell (int32, 0-based compact id), dx, dy, ux, uy, uz
 * (double or float by precision), pid (int32 or NULL).  Valid until the next
 * countingsort or destroy.                                                   */
int pbpic_species_device_ptrs(pbpic_ctx *ctx, int32_t species, void **ptrs);

/* Device field planes: ptrs receives 10 pointers: ex, ey, ez, bx, by, bz
 * (double), jx, jy, jz, rho (double or float by precision, unnormalized).
 * The E and B pointers are valid until the next field solve: the solve
 * writes the new fields into a second set of planes and swaps the two.     */
int pbpic_field_device_ptrs(pbpic_ctx *ctx, void **ptrs);

/* The HIP stream (hipStream_t) the context orders its work on.              */
void *pbpic_stream(pbpic_ctx *ctx);

/* Device-to-device variants of the setters, for callers that hold the state
 * on the GPU (the PyTorch binding).  cell is the 0-based compact id
 * (grid - 1); the other arrays have the context's particle precision (double
 * for PBPIC_PREC_FP64, float for PBPIC_PREC_SPLIT32).                        */
int pbpic_set_particles_device(pbpic_ctx *ctx, int32_t species, int64_t n,
                               const int32_t *cell, const void *dx,
                               const void *dy, const void *ux, const void *uy,
                               const void *uz);
int pbpic_set_field_device(pbpic_ctx *ctx, int32_t which, const double *plane);

/* ---------------------------------------------------------------- adjoint
 * Reverse-mode gradient of the time-integrated J.E, the analysis the
 * reference names but leaves empty (python/analysis_jdote.py:1-31); no
 * entry point of the reference corresponds.  pbpic_adjoint_forward runs
 * nsteps steps from the current state (as pbpic_step, sorting at its cadence),
 * keeping particle+field checkpoints every ckpt_every steps and, when device
 * memory allows, the three J planes of every step (the backward then replays
 * without depositing), and returns
 * loss = sum_n dt (mime/c^2) sum_cells J^{n+1/2}.(E^n+E^{n+1})/2.
 * pbpic_adjoint_backward then writes dloss/d(initial state): field_grads holds
 * 6 device pointers (nx*ny doubles: Ex, Ey, Ez, Bx, By, Bz), particle_grads
 * 5 device pointers per species in species order (dx, dy, ux, uy, uz, in the
 * particle precision, in the order the particles had before the forward).
 * Segments between checkpoints are recomputed into a tape of ckpt_every + 1
 * states, which the forward allocates at its end (the backward allocates it
 * if that failed) and the backward frees.  The context ends in the state
 * after the forward.                                                         */
int pbpic_adjoint_forward(pbpic_ctx *ctx, int32_t nsteps, int32_t ckpt_every,
                          double *loss);

/* Restricts the adjoint's loss to a region: weights is an nx*ny host plane
 * (x fastest) multiplying each cell's J.(E^n+E^{n+1})/2 term -- e.g. 1 inside
 * a diffusion region and 0 elsewhere; NULL restores the whole domain.  The
 * diagnostics PBPIC_D_JE_* stay unweighted.  Discards a pending forward.    */
int pbpic_set_loss_region(pbpic_ctx *ctx, const double *weights);
int pbpic_adjoint_backward(pbpic_ctx *ctx, double *const *field_grads,
                           void *const *particle_grads);

/* Per-kernel timing with HIP events on the context's stream (measurement
 * support; the reference has no timers, pic2d.f90:172 only prints itime).
 * enable != 0 starts collecting (and clears earlier samples), 0 stops.
 * pbpic_profile_read waits for the stream and returns, per PBPIC_PROF_* slot,
 * the summed milliseconds and the number of launches timed.                 */
#define PBPIC_PROF_PUSH   0   /* push kernel (gather+Boris+move+deposit)      */
#define PBPIC_PROF_FIELD  1   /* FDTD kernel and the push's node-field build  */
#define PBPIC_PROF_SORT   2   /* counting sort incl. payload gather           */
#define PBPIC_PROF_ZERO   3   /* current clear                                */
#define PBPIC_PROF_ADJ    4   /* adjoint backward (replay + reverse sweep)    */
#define PBPIC_NPROF       5
int pbpic_profile(pbpic_ctx *ctx, int32_t enable);
int pbpic_profile_read(pbpic_ctx *ctx, double ms[PBPIC_NPROF],
                       int64_t count[PBPIC_NPROF]);

#ifdef __cplusplus
}
#endif
#endif /* PBPIC_H */
