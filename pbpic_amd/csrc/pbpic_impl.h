// pbpic_impl.h -- internal declarations shared by the HIP translation units of
// libpbpic.so (not installed; the public interface is include/pbpic.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/pbpic.h"

namespace pbpic {

// Push tiling.  A workgroup of PUSH_THREADS lanes pushes one tile of
// PUSH_TILE = 64 x 64 consecutive particles of a species; each lane moves
// push_vec(precision) consecutive particles (16 bytes per array) per round.
constexpr int PUSH_THREADS = 256;
constexpr int PUSH_TILE = 4096;
// big[]: three deferred-tile counts and a pad, then the three lists (one
// list per tier for the whole device: per-XCD lists left the XCDs that hold
// the fast-spreading species with most of a listed launch's work, 12%
// slower on a Harris sheet)
constexpr int BIG_LISTS = 4;
// Split-form first-launch window at four workgroups per CU: 632 nodes of 48
// bytes and four crossing queues (40 KiB of a CU's 160); at three, 896.
constexpr int PUSH_MAXW_OCC4 = 632;
__host__ __device__ constexpr int push_vec(int precision) { return precision == 0 ? 2 : 4; }

// Push order.  Inside every full tile the sort can lay the particles out so
// that the q-th slot of the tile holds the particle of cell-order rank
// tile_rank(q, V): slot q = (c, l, j) with q = c*64V + l*V + j (round c of
// the tile, lane l, element j) takes rank l*64 + c*V + j.  The tile keeps the
// same particles as in cell order; only their order inside it changes, so
// that the 64 lanes of one wave-wide load hold particles 64 ranks apart (in
// different cells) and their deposits hit different LDS addresses.
__host__ __device__ constexpr int tile_rank(int q, int V)
{
    return ((q % (64 * V)) / V) * 64 + (q / (64 * V)) * V + q % V;
}

// Device-side error flags, OR-ed by kernels into one int and reported by
// pbpic_sync() and every call that returns data.
enum : int {
    ERRF_CFL = 1,        // a particle moved more than one cell
    ERRF_NONFINITE = 2,  // NaN/Inf momentum or Lorentz factor
    ERRF_CELL = 4,       // cell index out of range on input
    ERRF_SPIN = 8,       // sort look-back hit its spin limit
    ERRF_STATE = 16,     // a particle lay outside its tile's recorded bounding box
};

// FP64 device accumulators.
enum : int {
    ACC_KE_E = 0,      // sum (gamma-1)*ms over electrons in the last push
    ACC_KE_I = 1,      // same over ions
    ACC_JE_STEP = 2,   // sum_cells J_raw . E_avg of the last field solve
    ACC_JE_INT = 3,    // time integral of the scaled J.E before the last solve
    ACC_WE = 4,        // sum of squared node-averaged E
    ACC_WB = 5,        // sum of squared node-averaged B
    ACC_G_L2 = 6,      // Gauss residual sum of squares
    ACC_G_LINF = 7,    // Gauss residual max (bit pattern of a double >= 0)
    ACC_G_DL2 = 8,     // sum of squares of the residual change
    ACC_G_DLINF = 9,   // max of the residual change
    ACC_SPILL = 10,    // push tiles whose window did not fit in LDS (global atomics)
    ACC_JEM_STEP = 11, // as ACC_JE_STEP with the loss-region weights
    ACC_JEM_INT = 12,  // as ACC_JE_INT with the loss-region weights
    ACC_N = 16,
};

constexpr int MAX_SPECIES = 4;

// g mod n for any int g (periodic lattice index).
__device__ __forceinline__ int wrap_index(int g, int n)
{
    while (g < 0) g += n;
    while (g >= n) g -= n;
    return g;
}

// FP64 form's fixed point.  A contribution x, times its plane's power-of-two
// scale and 2^FX2_LOBITS (|x| < 2^49), becomes the integer rint(x): x + 1.5
// 2^52 lies in [2^52, 2^53), where doubles are the integers, so the addition
// rounds x to an integer (to nearest even) and the difference of the bit
// patterns is that integer.  Sums of such integers are exact in any order.  A
// global plane holds its sum Q in two 64-bit words, floor(Q / 2^FX2_LOBITS)
// and Q mod 2^FX2_LOBITS (the second plane lo entries on), exact up to 2^33
// contributions a node; the value is (hi + lo 2^-FX2_LOBITS) / scale.
constexpr int FX2_LOBITS = 19;
constexpr double FX2_LO = 1.0 / (double)(1 << FX2_LOBITS);   // exact
__device__ __forceinline__ long long fx2_quant(double x)
{
    constexpr double M = 6755399441055744.0;    // 1.5 * 2^52
    return __double_as_longlong(x + M) - __double_as_longlong(M);
}
__device__ __forceinline__ void fx2_add(long long *hi, int64_t lo, long long q)
{
    const long long h = q >> FX2_LOBITS, l = q & ((1LL << FX2_LOBITS) - 1);
    if (h) atomicAdd(reinterpret_cast<unsigned long long *>(hi), (unsigned long long)h);
    if (l) atomicAdd(reinterpret_cast<unsigned long long *>(hi + lo), (unsigned long long)l);
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Block b of nb is
// given work item xcd_remap(b, nb) so that each XCD takes a contiguous range
// (a bijection of [0, nb)): neighbouring items, which share field halos or
// source cache lines, then meet in one L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// Division by a run-time constant d >= 1 for 32-bit unsigned n
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI'94, round-up variant with a 33-bit magic).
struct FastDiv {
    uint32_t d, m, s;
    void init(uint32_t dd)
    {
        d = dd;
        s = 0;
        while ((1ull << s) < dd) s++;
        m = (uint32_t)((((1ull << 32) * ((1ull << s) - dd)) / dd) + 1);
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const
    {
        uint32_t hi = __umulhi(n, m);
        return (uint32_t)(((uint64_t)hi + n) >> s);
    }
    // n < 2^31 (every cell id: contexts hold fewer than 2^30 cells): hi <= n,
    // so hi + n does not overflow 32 bits
    __host__ __device__ __forceinline__ uint32_t div31(uint32_t n) const
    {
        return ((uint32_t)(((uint64_t)n * m) >> 32) + n) >> s;
    }
};

// One species as the kernels see it.
struct SpeciesDev {
    int32_t *cell;             // 0-based compact id cx*ny + cy (= grid - 1)
    void *ox, *oy, *ux, *uy, *uz;
    int32_t *pid;              // may be null
    // Where the push writes the pushed cell, offsets and momenta: the arrays
    // above (in place) unless a caller keeps the input state (the adjoint's
    // replay writes each step into the next slot of its tape).
    int32_t *cell_o;
    void *ox_o, *oy_o, *ux_o, *uy_o, *uz_o;
    // Split form, on a step whose sort follows: the push writes the pushed
    // particles here as PartRec records (index = slot) and writes cell_o, but
    // not ox_o..uz_o.  Null: the arrays as above.
    void *rec_o;
    int64_t n;
    int32_t q;                 // +1 ion, -1 electron
    int32_t tile0;             // first tile of this species in a push launch
    double ms;                 // mass: mime for q > 0, 1 otherwise (:964-968)
    double c0, c1;             // q*dt*mime/ms/2 and c0/cva (:969-970)
    double qx, qy;             // q*dx/dt, q*dy/dt (:1237-1238)
    // The LDS current window holds 64-bit fixed-point sums of the
    // contributions times these powers of two (jx, jy, jz, rho), each chosen
    // so that one contribution stays below 2^30 in magnitude.
    float jsx, jsy, jsz, jsr;
    // Per push tile: bounding box of the tile's cells as the last push left
    // them, {x0, x extent, y0, y extent} with x0 in [0, nx), y0 in [0, ny)
    // (extent -1: unknown).  tbox_ok: the boxes describe the current state.
    int32_t *tbox;
    int32_t tbox_ok;
};

struct PushParams {
    int nx, ny;
    int nyp;                   // conducting y walls: domain rows [0, nyp); 0 = periodic
    FastDiv dny;
    int compat, deposit_rho;
    int maxw;                  // LDS window capacity in cells
    // Tiles whose window exceeds maxw are deferred by the first launch
    // (pass 1): to a second launch with windows of maxw_mid nodes (pass 2,
    // two workgroups per CU) if they fit it, else to a third with maxw_big
    // (pass 3, one per CU).  big[0] and big[1] count the two lists, which
    // start at big + BIG_LISTS and big + BIG_LISTS + ntiles (big[2] and the
    // list at big + BIG_LISTS + 2 ntiles: pass 4, below).  pass 0: one
    // launch; tiles that do not fit deposit with global atomics, as do those
    // wider than maxw_big.
    int maxw_mid, maxw_big, pass;
    // occ4: the split form's plain push runs its first launch at four
    // workgroups per CU with windows of PUSH_MAXW_OCC4 nodes; the tiles
    // wider than that, up to maxw_occ3 nodes (the context's window), go to a
    // launch at three per CU (pass 4, list big[2]).  maxw_occ3 = 0: no such
    // tier.
    int occ4, maxw_occ3;
    int nodep;                 // 1: push without depositing J (adjoint replay)
    int *big;
    int ntiles;
    int nspecies;
    double dt, cva, idx, idy, dtcva;   // dtcva = dt*cva (:1003)
    const double *ex, *ey, *ez, *bx, *by, *bz;
    // unnormalized current as 64-bit fixed-point planes: the sum of every
    // contribution times the plane's scale jsx..jsr, rounded to an integer
    // (split form), or times 2^FX2_LOBITS as well, rounded, and held in two
    // words (FP64 form, fx2_add; the low-word plane jlo entries after the
    // high one).  Integer sums are exact and independent of the order of the
    // additions; launch_finalize_current converts them.
    void *jx, *jy, *jz, *rho;
    int64_t jlo;               // FP64 form: ncell; split form: 0
    float jsx, jsy, jsz, jsr;  // fixed-point scales (as SpeciesDev's)
    void *nf;                  // node-field array (nodefield_bytes), push precision
    double *acc;
    double *part;              // partial sums of the fixed-order reductions
    int *err;
    SpeciesDev sp[MAX_SPECIES];
};

struct FdtdParams {
    int nx, ny;
    int nyp;                   // conducting y walls: domain rows [0, nyp); 0 = periodic
    double dth, dtc2, idx, idy, in0;   // dt/2, dt*cva^2, 1/dx, 1/dy, in0
    double *ex, *ey, *ez, *bx, *by, *bz;
    // fdtd_kernel writes E^{n+1}, B^{n+1} here, never into the planes above:
    // a workgroup reads its neighbours' cells as halo, and workgroups run in
    // no fixed order, so an in-place update would mix time levels at tile
    // seams.  The caller swaps the two plane sets after the launch.
    double *nex, *ney, *nez, *nbx, *nby, *nbz;
    const void *jx, *jy, *jz;
    const double *fex, *fey, *fez;     // filter increments added to E^{n+1} (or null)
    const double *lmask;               // loss-region weights per cell (or null)
    double *acc;
    double *part;                      // per-workgroup partial sums (launch_reduce_parts)
    double je_pre;                     // dt mime/c^2: integrates the previous J.E step
};

// Reductions in a fixed order.  A reduction kernel writes one partial per
// workgroup and quantity, part[q*nb + block]; launch_reduce_parts then adds
// scale[q] * (the sum of part[q*nb .. q*nb+nb), taken in one fixed order) to
// acc[slot[q]], so that the same inputs give the same bits every run.
struct PartSlots {
    int nq;
    int slot[4];
    double scale[4];
    // ints[q] >= 0: first acc[ints[q]] += acc[slot[q]] * pre (the time
    // integral of the previous step's value), then acc[slot[q]] = the sum
    int ints[4] = {-1, -1, -1, -1};
    double pre = 0;
};
hipError_t launch_reduce_parts(const double *part, int nb, double *acc, const PartSlots &ps,
                               hipStream_t st);

// Marder / Cherenkov increments of E from E^n and rho^n (k_field.hip).
struct FilterParams {
    int nx, ny, marder, cherenkov;
    double idx, idy, cva2in0, mdx, mdy, cdx, cdy;   // mdx = dt 0.001/dx, cdx = dt c^2 tau/dx
    const double *ex, *ey, *ez;
    const void *rho;                               // raw charge, precision of J
    double *fex, *fey, *fez;
};
hipError_t launch_filters(const FilterParams &p, int precision, hipStream_t st);

// Adjoint of the push for one step (k_adjoint.hip).  The particle state of
// step n comes from the tape (cell, ox..uz); xa..uza hold the adjoint of the
// state at n+1 on entry and at n on exit.
struct AdjSpecies {
    const int32_t *cell;
    const void *ox, *oy, *ux, *uy, *uz;
    int64_t n;
    int32_t q, tile0;
    double ms, c0, c1, qx, qy;
};

struct AdjParams {
    int nx, ny, compat, maxw, nspecies;
    int tv;                                      // push order vector width (0: cell order)
    FastDiv dny;                                 // division by ny
    int nyp;                                     // conducting y walls (0: periodic)
    // two-pass launch as PushParams: tiles whose window exceeds maxw but fits
    // maxw_big are deferred to a second launch (big: count + tile list)
    int maxw_big, pass;
    int *big;
    int jmax;                                    // current-adjoint window (nodes)
    double dtcva, idx, idy, cva;
    const double *ex, *ey, *ez, *bx, *by, *bz;   // fields at step n
    // node fields of step n (launch_nodefield's array: per node (cx*ny + cy)
    // 8 floats (split) or 6 doubles (FP64), the first six E, B averaged as
    // from the planes above), read by the LDS window staging
    const void *nf;
    const double *jadj[3];                       // current adjoint (raw J)
    double *nadj;                                // 6 node-adjoint planes
    void *xa[MAX_SPECIES], *ya[MAX_SPECIES], *uxa[MAX_SPECIES], *uya[MAX_SPECIES],
        *uza[MAX_SPECIES];
    AdjSpecies sp[MAX_SPECIES];
};

struct AdjFieldParams {
    int nx, ny;
    int nyp;                                     // conducting y walls (0: periodic)
    double dth, dtc2, idx, idy, in0, w;          // w = dt*mime/cva^2
    double *exa, *eya, *eza, *bxa, *bya, *bza;   // field adjoints (in/out)
    double *etx, *ety, *etz;                     // scratch: Ebar'
    double *jxa, *jya, *jza;                     // current adjoint (out)
    const double *e0x, *e0y, *e0z, *e1x, *e1y, *e1z;   // E^n, E^{n+1}
    const void *jx, *jy, *jz;                    // J^{n+1/2} (raw)
    const double *lmask;                         // loss-region weights (or null)
};

// Launchers (one per translation unit family).
hipError_t launch_push(const PushParams &p, int precision, hipStream_t st);
// Window capacity of the second push pass (one workgroup per CU).
int push_maxw_big(int precision, int deposit_rho, int dep);
int push_maxw_mid(int precision, int deposit_rho, int dep);
int push_maxw_nodep(int precision, int deposit_rho, int maxw);
size_t nodefield_bytes(int precision, int64_t ncell);
// Fills p.nf from the E and B planes; launch_push reads p.nf.
hipError_t launch_nodefield(const PushParams &p, int precision, hipStream_t st);
// Adds species S's kinetic energy to acc[ACC_KE_E] or acc[ACC_KE_I].
hipError_t launch_kinetic(const SpeciesDev &S, int precision, double *acc, double *part,
                          hipStream_t st);
hipError_t launch_deposit_batch(const PushParams &p, int precision, int64_t n,
                                const int32_t *i1, const int32_t *i2,
                                const int32_t *j1, const int32_t *j2,
                                const double *dx1, const double *dx2,
                                const double *dy1, const double *dy2,
                                const int32_t *q, const double *vz,
                                hipStream_t st);
hipError_t launch_fdtd(const FdtdParams &p, int precision, hipStream_t st);
hipError_t launch_energy(const FdtdParams &p, hipStream_t st);
// The charge deposit as 64-bit fixed point (sums exact in any order): split
// form, one word per node with scale rho_scale; FP64 form, two words per node
// (rho_scratch holds 2 ncell), scale 2^30 (rho_scale is not used).
hipError_t launch_gauss(const PushParams &p, int precision, double rho_scale, double *rho_scratch,
                        double *gauss0, int first, double rho_bg,
                        double cva2in0, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t st);
// Up to COPY_MAX device-to-device copies in one launch (copies that are not
// whole 16-byte chunks on 16-byte boundaries fall back to hipMemcpyAsync).
constexpr int COPY_MAX = 8;
struct CopyList {
    const void *src[COPY_MAX];
    void *dst[COPY_MAX];
    size_t bytes[COPY_MAX];
    int64_t first[COPY_MAX];    // first global thread of each copy (set by the launcher)
    int n;
    void add(void *d, const void *s, size_t b) { src[n] = s; dst[n] = d; bytes[n] = b; n++; }
};
hipError_t launch_copy_list(CopyList L, hipStream_t st);
// Conducting y walls: sets the guard rows of the field planes selected by
// mask (bit k: Ex, Ey, Ez, Bx, By, Bz) to their images (the Jz of the guard
// rows is folded back by launch_finalize_current).
hipError_t launch_wall_images(const FdtdParams &p, int mask, hipStream_t st);
// Harris-sheet loader (k_field.hip); see pbpic_init_harris.
struct HarrisDev {
    int geometry, ppc, thermal_z, nn, nt;   // nn, nt: cells across / along the sheets
    int64_t ncs;                            // sheet particles per species
    double L, dn, xc0, xc1;                 // half width, cell size across, centres
    double uic, uec, uib, ueb, udi, ude;    // momenta per unit (u = v/sqrt(1-v^2))
    uint32_t k0, k1, stream;
};
hipError_t launch_init_harris(const SpeciesDev &ion, const SpeciesDev &ele, int precision,
                              int ny, const HarrisDev &h, hipStream_t st);
// J = fixed-point sums / scale, for the 3 current planes and the charge
// plane too when with_rho: float planes from one word (elem_bytes 4, split
// form), double planes from two (elem_bytes 8, FP64 form; low words ncell
// entries after the high ones).
hipError_t launch_finalize_current(void *const *j64, void *const *j, int64_t ncell,
                                   const float *scale, int with_rho, int nx, int nyp,
                                   int elem_bytes, hipStream_t st);
hipError_t launch_zero_current(void *jx, void *jy, void *jz, void *rho,
                               int64_t ncell, int elem_bytes, double *acc,
                               int reset_ke, hipStream_t st);
hipError_t launch_acc_op(double *acc, int op, double a, hipStream_t st);
hipError_t launch_init_maxwellian(const SpeciesDev &s, int precision, int nx,
                                  int ny, int ppc, double vth, double vdrift,
                                  int counter_stream, uint64_t seed,
                                  uint32_t stream, int nyp, hipStream_t st);
hipError_t launch_check_cells(const int32_t *cell, int64_t n, int64_t ncell, int ny,
                              int nyp, int *err, hipStream_t st);

// Diagnostics: energy spectrum of a species (SPEC_BINS_MAX bins at most) and
// per-cell particle counts.
constexpr int SPEC_BINS_MAX = 4096;
hipError_t launch_spectrum(const SpeciesDev &S, int precision, int nbins, double de, double emax,
                           uint32_t *counts, hipStream_t st);
hipError_t launch_cell_counts(const SpeciesDev &sp, int nx, int ny, int32_t *counts, int *err,
                              hipStream_t st);

// Sort (stable LSD radix over the cell id, 8-bit digits).
struct SortScratch {
    uint32_t *keys[2] = {nullptr, nullptr};
    uint32_t *vals[2] = {nullptr, nullptr};
    uint32_t *status = nullptr;     // ntiles * 256
    uint32_t *dhist = nullptr;      // npass * 256 (counts, then offsets)
    uint32_t *ticket = nullptr;     // one counter per pass
    int32_t *counts = nullptr;      // ncell
    const uint32_t *perm = nullptr; // permutation of the last sort (in vals)
    int64_t cap = 0, status_cap = 0;
    // tile-box path (sort_species): per push tile the box area and offset
    // (tiles + 1 entries each, grown with cap), scan block sums, a small
    // control block read back by the host, and the record arrays (7 per
    // record, grown on demand)
    uint32_t *tarea = nullptr, *toff = nullptr, *bsum = nullptr;
    int *ctl = nullptr;
    uint32_t *rbuf = nullptr;
    int64_t rcap = 0;
    // push-order supertile ranks (CellKey), device and host copies
    uint32_t *strank = nullptr;
    std::vector<uint32_t> strank_h;
};
hipError_t launch_push_adj(const AdjParams &p, int precision, hipStream_t st);
int adj_jmax_for(int precision, int maxw, size_t budget);
// Tiles of the adjoint push (2048 particles each) for n particles.
int64_t adj_tiles(int64_t n);
hipError_t launch_fdtd_adj(const AdjFieldParams &p, int precision, hipStream_t st);
// Conducting y walls: transpose of the guard-row imaging applied to the field
// adjoints (exa..bza) in place.
hipError_t launch_wall_images_adj(const AdjFieldParams &p, hipStream_t st);
hipError_t launch_node_transpose(const AdjFieldParams &p, double *nadj, hipStream_t st);
// Reverse of a sort for the five adjoint arrays: pre[perm[k]] = post[k];
// push_order: the sort laid full tiles out in the push kernel's lane order.
hipError_t launch_unpermute(const uint32_t *perm, int64_t n, int precision, int push_order,
                            void *const *post, void *const *pre, hipStream_t st);

// Sort key of a cell.  Reference order: the compact cell id cx*ny + cy.
// Push order: cells grouped in 8 x 8 supertiles, the supertiles in the order
// of a generalized Hilbert curve over the nbx x nby supertile grid
// (supertile_ranks) and the cells x-major inside each.  A run of 4096
// particles (~41 cells at 100 per cell) covers a compact ~6 x 8 block; each
// step of the curve goes to an edge neighbour, so a run that straddles two
// supertiles stays compact too; and a supertile's neighbours in x and in y
// mostly lie a few supertiles away in key order, so the sort's gather, whose
// sources are the destinations' neighbours of a cycle before, finds them
// still in L2.  The rank table (nbx * nby entries, index sx * nby + sy) has a
// device copy for kernels and a host copy for host code.
struct CellKey {
    FastDiv dny;
    uint32_t ny, nby;
    int tiled;
    int64_t nkeys;            // keys lie in [0, nkeys)
    const uint32_t *rank = nullptr;     // device table (tiled keys in kernels)
    const uint32_t *rank_h = nullptr;   // host table (tiled keys in host code)
    void init(int nx, int ny_, int tiled_, const uint32_t *rank_dev = nullptr,
              const uint32_t *rank_host = nullptr)
    {
        dny.init((uint32_t)ny_);
        ny = (uint32_t)ny_;
        nby = (uint32_t)((ny_ + 7) / 8);
        tiled = tiled_;
        nkeys = tiled ? (int64_t)((nx + 7) / 8) * nby * 64 : (int64_t)nx * ny_;
        rank = rank_dev;
        rank_h = rank_host;
    }
    __host__ __device__ uint32_t operator()(uint32_t c) const
    {
        if (!tiled) return c;
        const uint32_t cx = dny.div31(c), cy = c - cx * ny;
        const uint32_t s = (cx >> 3) * nby + (cy >> 3);
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t st = rank[s];
#else
        const uint32_t st = rank_h[s];
#endif
        return (st << 6) | ((cx & 7u) << 3) | (cy & 7u);
    }
};
// Position along the generalized Hilbert curve of every cell of a w x h grid,
// rank[x * h + y] (J. Cerveny's "gilbert" construction for arbitrary
// rectangles: every step goes to an edge neighbour, with at most one diagonal
// step when the dimensions force it).
void supertile_ranks(int w, int h, std::vector<uint32_t> &rank);

// One particle of the split form as the sort's gather reads it: 24 bytes,
// record i holding slot i.
struct PartRec {
    int32_t c;
    float a[5];                // ox, oy, ux, uy, uz
};
static_assert(sizeof(PartRec) == 24, "PartRec is 24 bytes");

// push_order != 0: cells in CellKey's supertile order and full tiles in the
// push order (tile_rank).  counts are indexed by key.(s.counts holds nkeys).
// rec (split form): the species' record buffer (PartRec per particle), or
// null for the array-wise gather; rec_written: the push already wrote sp's
// records there (sp's value arrays are then stale and not read).
// want_counts: fill s.counts (per key); want_perm: leave the permutation at
// s.perm.  Sorts right after a push (sp.tbox_ok) take the tile-box path when
// every tile's box fits; otherwise, or when counts are wanted, the radix path.
hipError_t sort_species(SortScratch &s, const SpeciesDev &sp, int precision, int push_order,
                        const CellKey &key, int64_t ncell, SpeciesDev &dst, int *err,
                        int want_counts, int want_perm, void *rec, bool rec_written,
                        bool *boxed, hipStream_t st);

}  // namespace pbpic
