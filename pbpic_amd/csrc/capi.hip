// capi.hip -- the C ABI of include/pbpic.h: context, buffers, host I/O and
// the step entry points that mirror pic2d.f90's subroutines.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "pbpic_impl.h"

using namespace pbpic;

// Staging area of pbpic_deposit_noalloc, the entry point for callers that
// cannot allocate device memory: PBPIC_DEPOSIT_NOALLOC_MAX entries of 5 int32
// and 5 double inputs (60 B each, plus alignment slack), 3.9 MB per GPU, held
// until the process ends.  In host code the name stands for a host-side
// shadow, not device memory: only its address from hipGetSymbolAddress (taken
// per device, kept in scratch_dev) is handed to copies and kernels.
static __device__ __align__(256) char scratch[PBPIC_DEPOSIT_NOALLOC_MAX * 60 + 256];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_dev[SCRATCH_MAX_DEVICES] = {};

// The device address of `scratch` on the current device dev, looked up once
// per device (each GPU holds its own copy of the array at its own address).
hipError_t scratch_address(int dev, void **out)
{
    *out = nullptr;
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_dev[dev] = ptr;
    }
    *out = scratch_dev[dev];
    return hipSuccess;
}

struct Species {
    int32_t q = 0;
    int64_t cap = 0, n = 0;
    double ms = 1, c0 = 0, c1 = 0, qx = 0, qy = 0;
    int32_t *cell[2] = {nullptr, nullptr};
    void *ox[2] = {nullptr, nullptr}, *oy[2] = {nullptr, nullptr};
    void *ux[2] = {nullptr, nullptr}, *uy[2] = {nullptr, nullptr};
    void *uz[2] = {nullptr, nullptr};
    int32_t *pid[2] = {nullptr, nullptr};
    int cur = 0;
    int32_t *tbox = nullptr;   // per push tile bounding boxes (SpeciesDev::tbox)
    bool tbox_ok = false;      // cleared by anything that moves or rewrites particles
    void *rec = nullptr;       // split form: the sort's PartRec buffer (cap records)
};

}  // namespace

struct pbpic_ctx {
    pbpic_params p{};
    int rsz = 8;
    int dev = 0;
    hipStream_t st = nullptr;
    void *scratch = nullptr;   // this device's address of `scratch` (pbpic_deposit_noalloc)
    double *E[3] = {nullptr, nullptr, nullptr};
    double *B[3] = {nullptr, nullptr, nullptr};
    // the field solve writes E^{n+1}, B^{n+1} into these and swaps them with
    // E, B (an in-place stencil would read neighbours' new values)
    double *E2[3] = {nullptr, nullptr, nullptr};
    double *B2[3] = {nullptr, nullptr, nullptr};
    void *J[4] = {nullptr, nullptr, nullptr, nullptr};
    // split form: the push's 64-bit fixed-point current sums (scales js),
    // converted into J after every deposit
    void *J64[4] = {nullptr, nullptr, nullptr, nullptr};
    float js[4] = {1, 1, 1, 1};
    void *NF = nullptr;        // node fields for the push (rebuilt before each push)
    double *acc = nullptr;
    double *part = nullptr;    // partials of the fixed-order reductions
    int *err = nullptr;
    double *rho_scratch = nullptr, *gauss0 = nullptr;
    int32_t *diag_counts = nullptr;   // spectrum / number-density scratch
    double *filt[3] = {nullptr, nullptr, nullptr};   // filter increments of E
    int32_t *big = nullptr;    // deferred push tiles (counts + lists)
    // The split form's plain push has two builds of its first launch, at
    // four and at three workgroups per CU, which give the same bits.  Which
    // is faster depends on how far the tiles' cells have spread since the
    // last sort: four per CU wins while every tile is compact (a uniform
    // plasma all cycle long, a Harris sheet early in each cycle); three
    // wins once many tiles need wider windows than four per CU leave, and
    // mixing them with the compact ones in one launch overlaps their work.
    // So the choice is kept per tenth of the sort cycle (OCC_BUCKETS): each
    // tenth times one push of each build every OCC_REPROBE cycles (events
    // around the push's launches; the host waits for the second push of the
    // pair, a short stall once per tenth and probing cycle) and runs the
    // faster.  Until a tenth has been timed it runs three per CU.
    static constexpr int OCC_BUCKETS = 10, OCC_REPROBE = 8;
    struct OccBucket {
        hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
        int64_t last_probe = -(1 << 30);   // cycle of the last timed pair
        int stage = 0;                      // 0 idle, 1 the three-per-CU push timed
        bool use4 = false;
    };
    OccBucket occ[OCC_BUCKETS];
    double *lmask = nullptr;   // loss-region weights of the adjoint (nx*ny) or null
    int64_t big_cap = 0;
    bool gauss_first = true;
    std::vector<Species> sp;
    // One-shot push destination per species (cell, dx, dy, ux, uy, uz), set
    // by the adjoint replay: the next push writes there and those arrays
    // become the species' current ones.  Empty: the push works in place.
    std::vector<std::array<void *, 6>> push_out;
    SortScratch sort;
    int64_t steps = 0;
    int64_t sorts_radix = 0;   // sorts that took the radix path (PBPIC_D_SORTS_RADIX)
    int maxw = 0;
    double idx = 0, idy = 0, dth = 0, dtc2 = 0, dtcva = 0;
    std::string last_error;
    // adjoint: checkpoints of a differentiable forward run
    struct {
        bool valid = false;
        int32_t nsteps = 0, ckpt = 0;
        int64_t step0 = 0;
        std::vector<char *> ckpts;     // particle+field state before step k*ckpt
        char *final_state = nullptr;   // state after the forward
        // J after the push of every forward step (3 planes each), so that
        // the backward's replay need not deposit; null if it did not fit
        char *jstore = nullptr;
        // the backward's tape and adjoint buffers, allocated at the end of
        // the forward (or by the backward) and freed by the backward
        char *tape = nullptr;
        double acc_after[ACC_N] = {0};
        double loss = 0;
    } adj;
    // event timing
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> ev_marks;   // (slot, index of start event)
    // host particle I/O: two pinned staging buffers of stage_bytes each
    // (grown on demand up to STAGE_BYTES) and the events of the copies that
    // last used them
    char *stage[2] = {nullptr, nullptr};
    size_t stage_bytes = 0;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_busy[2] = {false, false};
};

namespace {

int fail(pbpic_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->last_error = buf;
    return code;
}

#define HIPCHK(c, x)                                                             \
    do {                                                                         \
        hipError_t _e = (x);                                                     \
        if (_e != hipSuccess)                                                    \
            return fail((c), PBPIC_ERR_HIP, "%s: %s", #x, hipGetErrorString(_e)); \
    } while (0)

int64_t ncell(const pbpic_ctx *c) { return (int64_t)c->p.nx * c->p.ny; }
// Rows of the domain: ny, or ny - 2 with conducting y walls (two guard rows).
int nyp_of(const pbpic_ctx *c) { return c->p.ybound == PBPIC_YBOUND_CONDUCTING ? c->p.ny - 2 : c->p.ny; }
bool walls(const pbpic_ctx *c) { return c->p.ybound == PBPIC_YBOUND_CONDUCTING; }

template <typename T>
hipError_t dalloc(T **p, size_t bytes)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes > 0 ? bytes : 16);
    *p = (T *)q;
    return e;
}

// 2^(29 - e) for M = m 2^e, 1 <= m < 2: the largest power of two s with
// M s < 2^30.
static float fx_scale(double M)
{
    if (!(M > 0) || !std::isfinite(M)) return 1.0f;
    return (float)std::ldexp(1.0, 29 - std::ilogb(M));
}

SpeciesDev species_dev(pbpic_ctx *c, int s, int buf)
{
    Species &S = c->sp[s];
    SpeciesDev d{};
    d.cell = S.cell[buf];
    d.ox = S.ox[buf]; d.oy = S.oy[buf];
    d.ux = S.ux[buf]; d.uy = S.uy[buf]; d.uz = S.uz[buf];
    d.pid = S.pid[buf];
    d.cell_o = d.cell;
    d.ox_o = d.ox; d.oy_o = d.oy; d.ux_o = d.ux; d.uy_o = d.uy; d.uz_o = d.uz;
    d.n = S.n;
    d.q = S.q;
    d.ms = S.ms; d.c0 = S.c0; d.c1 = S.c1; d.qx = S.qx; d.qy = S.qy;
    // bounds of one deposit contribution: |jx| <= |q dx/dt|, |jy| <= |q dy/dt|
    // (segment fractions and weights lie in [-1, 1] and [0, 1], :1244-1257),
    // |jz| <= |q| cva (|v_z| < c = cva, :1277), |rho| <= |q| (:1283-1295)
    d.jsx = fx_scale(fabs(S.qx));
    d.jsy = fx_scale(fabs(S.qy));
    d.jsz = fx_scale(fabs((double)S.q) * c->p.cva);
    d.jsr = fx_scale(fabs((double)S.q));
    d.tbox = S.tbox;
    d.tbox_ok = S.tbox_ok ? 1 : 0;
    return d;
}

// recs: the push writes the split form's particles as the sort's records
// (the sort follows in the same step).
PushParams push_params(pbpic_ctx *c, bool recs = false, bool nodep = false)
{
    PushParams P{};
    P.nx = c->p.nx;
    P.ny = c->p.ny;
    P.nyp = walls(c) ? nyp_of(c) : 0;
    P.dny.init((uint32_t)c->p.ny);
    P.compat = c->p.compat;
    P.deposit_rho = (c->p.flags & PBPIC_FLAG_DEPOSIT_RHO) ? 1 : 0;
    // a push without deposit has no current window: its node-field window
    // takes the LDS of both
    P.maxw = nodep ? push_maxw_nodep(c->p.precision, P.deposit_rho, c->maxw) : c->maxw;
    // the split form's plain push may run its first launch at four
    // workgroups per CU (push_impl decides; launch_push sets the windows)
    P.occ4 = c->rsz == 4 && !nodep && !recs && !P.deposit_rho;
    P.maxw_mid = push_maxw_mid(c->p.precision, P.deposit_rho, nodep ? 0 : 1);
    P.maxw_big = push_maxw_big(c->p.precision, P.deposit_rho, nodep ? 0 : 1);
    P.big = c->big;
    P.pass = 0;
    P.nodep = nodep ? 1 : 0;
    P.dt = c->p.dt;
    P.cva = c->p.cva;
    P.idx = c->idx;
    P.idy = c->idy;
    P.dtcva = c->dtcva;
    P.ex = c->E[0]; P.ey = c->E[1]; P.ez = c->E[2];
    P.bx = c->B[0]; P.by = c->B[1]; P.bz = c->B[2];
    P.jx = c->J64[0]; P.jy = c->J64[1]; P.jz = c->J64[2]; P.rho = c->J64[3];
    P.jlo = c->rsz == 8 ? ncell(c) : 0;
    P.jsx = c->js[0]; P.jsy = c->js[1]; P.jsz = c->js[2]; P.jsr = c->js[3];
    P.nf = c->NF;
    P.acc = c->acc;
    P.part = c->part;
    P.err = c->err;
    P.nspecies = (int)c->sp.size();
    for (int s = 0; s < P.nspecies; s++) P.sp[s] = species_dev(c, s, c->sp[s].cur);
    for (int s = 0; s < (int)c->push_out.size() && s < P.nspecies; s++) {
        SpeciesDev &d = P.sp[s];
        d.cell_o = (int32_t *)c->push_out[s][0];
        d.ox_o = c->push_out[s][1]; d.oy_o = c->push_out[s][2];
        d.ux_o = c->push_out[s][3]; d.uy_o = c->push_out[s][4]; d.uz_o = c->push_out[s][5];
    }
    if (recs && c->rsz == 4)
        for (int s = 0; s < P.nspecies; s++) P.sp[s].rec_o = c->sp[s].rec;
    return P;
}

FdtdParams fdtd_params(pbpic_ctx *c)
{
    FdtdParams F{};
    F.nx = c->p.nx; F.ny = c->p.ny;
    F.nyp = walls(c) ? nyp_of(c) : 0;
    F.dth = c->dth; F.dtc2 = c->dtc2; F.idx = c->idx; F.idy = c->idy; F.in0 = c->p.in0;
    F.ex = c->E[0]; F.ey = c->E[1]; F.ez = c->E[2];
    F.bx = c->B[0]; F.by = c->B[1]; F.bz = c->B[2];
    F.nex = c->E2[0]; F.ney = c->E2[1]; F.nez = c->E2[2];
    F.nbx = c->B2[0]; F.nby = c->B2[1]; F.nbz = c->B2[2];
    F.jx = c->J[0]; F.jy = c->J[1]; F.jz = c->J[2];
    F.fex = c->filt[0]; F.fey = c->filt[1]; F.fez = c->filt[2];
    F.lmask = c->lmask;
    F.acc = c->acc;
    F.part = c->part;
    return F;
}

// Translates and clears the device error flags.
int check_flags(pbpic_ctx *c)
{
    int h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, c->err, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (!h) return PBPIC_OK;
    HIPCHK(c, launch_fill_u32((uint32_t *)c->err, 1, 0u, c->st));
    if (h & ERRF_CELL) return fail(c, PBPIC_ERR_CELL, "particle cell index out of range");
    if (h & ERRF_CFL) return fail(c, PBPIC_ERR_CFL, "a particle moved more than one cell in one step");
    if (h & ERRF_NONFINITE) return fail(c, PBPIC_ERR_NONFINITE, "non-finite particle momentum");
    if (h & ERRF_SPIN) return fail(c, PBPIC_ERR_SPIN, "sort look-back exceeded its spin limit");
    if (h & ERRF_STATE)
        return fail(c, PBPIC_ERR_STATE, "particle cells changed outside the library since the last "
                                        "push (call pbpic_species_device_ptrs or "
                                        "pbpic_set_particles_device after writing them)");
    return fail(c, PBPIC_ERR_STATE, "unknown device error flag %d", h);
}

int valid_species(pbpic_ctx *c, int32_t s)
{
    if (!c) return PBPIC_ERR_ARG;
    if (s < 0 || s >= (int32_t)c->sp.size()) return fail(c, PBPIC_ERR_ARG, "bad species %d", s);
    return PBPIC_OK;
}

hipError_t prof_mark(pbpic_ctx *c, int slot, bool begin);

// rec_written: the push of this step wrote the species' records (step_impl).
int sort_one(pbpic_ctx *c, int s, int32_t *counts, int32_t *perm, uint32_t *perm_dev = nullptr,
             bool rec_written = false)
{
    Species &S = c->sp[s];
    SpeciesDev src = species_dev(c, s, S.cur), dst = species_dev(c, s, S.cur ^ 1);
    HIPCHK(c, prof_mark(c, PBPIC_PROF_SORT, true));
    const bool po = (c->p.flags & PBPIC_FLAG_PUSH_ORDER) != 0;
    CellKey key;
    key.init(c->p.nx, c->p.ny, po ? 1 : 0, c->sort.strank, c->sort.strank_h.data());
    bool boxed = false;
    HIPCHK(c, sort_species(c->sort, src, c->p.precision, po ? 1 : 0, key, ncell(c), dst, c->err,
                           counts != nullptr, perm != nullptr || perm_dev != nullptr, S.rec,
                           rec_written && S.rec, &boxed, c->st));
    if (!boxed && S.n > 0) c->sorts_radix++;
    HIPCHK(c, prof_mark(c, PBPIC_PROF_SORT, false));
    S.cur ^= 1;
    S.tbox_ok = false;
    if (perm_dev && S.n > 0)
        HIPCHK(c, hipMemcpyAsync(perm_dev, c->sort.perm, sizeof(uint32_t) * S.n,
                                 hipMemcpyDeviceToDevice, c->st));
    std::vector<int32_t> kcounts;
    if (counts) {
        // the sort counts per key; the caller gets counts per cell (grid - 1)
        kcounts.resize((size_t)key.nkeys);
        HIPCHK(c, hipMemcpyAsync(kcounts.data(), c->sort.counts, sizeof(int32_t) * key.nkeys,
                                 hipMemcpyDeviceToHost, c->st));
    }
    if (perm && S.n > 0)
        HIPCHK(c, hipMemcpyAsync(perm, c->sort.perm, sizeof(int32_t) * S.n, hipMemcpyDeviceToHost, c->st));
    if (counts || perm) HIPCHK(c, hipStreamSynchronize(c->st));
    if (counts)
        for (int64_t k = 0; k < ncell(c); k++) counts[k] = kcounts[key((uint32_t)k)];
    return PBPIC_OK;
}

// Records a timing event pair around a launch when profiling is on.
hipError_t prof_mark(pbpic_ctx *c, int slot, bool begin)
{
    if (!c->prof) return hipSuccess;
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        c->ev_pool.push_back(e);
    }
    if (begin) c->ev_marks.push_back({slot, c->ev_used});
    return hipEventRecord(c->ev_pool[c->ev_used++], c->st);
}

// Kinetic energy of every species from the current momenta into acc.
int kinetic_impl(pbpic_ctx *c)
{
    HIPCHK(c, launch_acc_op(c->acc, 0, 0.0, c->st));
    for (int s = 0; s < (int)c->sp.size(); s++)
        HIPCHK(c, launch_kinetic(species_dev(c, s, c->sp[s].cur), c->p.precision, c->acc, c->part,
                                 c->st));
    return PBPIC_OK;
}

// The current planes (FP32 split form, FP64) from the fixed-point sums, with
// the Jz wall fold of conducting y walls applied to the sums.
int finalize_current(pbpic_ctx *c)
{
    HIPCHK(c, launch_finalize_current(c->J64, c->J, ncell(c), c->js,
                                      (c->p.flags & PBPIC_FLAG_DEPOSIT_RHO) ? 1 : 0, c->p.nx,
                                      walls(c) ? nyp_of(c) : 0, c->rsz, c->st));
    return PBPIC_OK;
}

// nodep: push without depositing J (the adjoint's replay supplies J).
int push_impl(pbpic_ctx *c, bool recs = false, bool nodep = false)
{
    int64_t ntiles = 0;
    for (auto &S : c->sp) ntiles += (S.n + PUSH_TILE - 1) / PUSH_TILE;
    const int64_t bneed = 3 * ntiles + BIG_LISTS;
    if (bneed > c->big_cap) {
        // the deferred-tile lists of the tiered push: three counts, one entry
        // per tile in each list
        if (c->big) { HIPCHK(c, hipStreamSynchronize(c->st)); (void)hipFree(c->big); c->big = nullptr; }
        HIPCHK(c, dalloc(&c->big, sizeof(int32_t) * (size_t)bneed));
        c->big_cap = bneed;
    }
    PushParams P = push_params(c, recs, nodep);
    // the first launch's build: the faster for this tenth of the sort cycle,
    // or one of a timed pair
    pbpic_ctx::OccBucket *ob = nullptr;
    int probe = -1;
    if (P.occ4 && c->occ[0].ev[0]) {
        const int K = c->p.sort_every > 0 ? c->p.sort_every : 50;
        const int64_t phase = c->steps % K, cycle = c->steps / K;
        ob = &c->occ[phase * pbpic_ctx::OCC_BUCKETS / K];
        // no timing under stream capture (a graph replays without the host)
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        const hipError_t ce = hipStreamIsCapturing(c->st, &cap);
        if (ce != hipSuccess) (void)hipGetLastError();
        const bool capturing = ce != hipSuccess || cap != hipStreamCaptureStatusNone;
        if (capturing) probe = -1;
        else if (ob->stage == 1) probe = 1;            // the pair's second push
        else if (ob->stage == 0 && cycle - ob->last_probe >= pbpic_ctx::OCC_REPROBE) probe = 0;
        P.occ4 = probe >= 0 ? probe == 1 : ob->use4;
    }
    // node fields (emfgrid, pic2d.f90:837-845) are field work: timed as such
    HIPCHK(c, prof_mark(c, PBPIC_PROF_FIELD, true));
    HIPCHK(c, launch_nodefield(P, c->p.precision, c->st));
    HIPCHK(c, prof_mark(c, PBPIC_PROF_FIELD, false));
    HIPCHK(c, prof_mark(c, PBPIC_PROF_PUSH, true));
    if (probe >= 0) HIPCHK(c, hipEventRecord(ob->ev[2 * probe], c->st));
    HIPCHK(c, launch_push(P, c->p.precision, c->st));
    if (probe >= 0) {
        HIPCHK(c, hipEventRecord(ob->ev[2 * probe + 1], c->st));
        ob->stage = probe == 0 ? 1 : 0;
        if (probe == 0) ob->last_probe = c->steps / (c->p.sort_every > 0 ? c->p.sort_every : 50);
        if (probe == 1) {
            // the pair complete: decide now, for the rest of this tenth
            float t3 = 0, t4 = 0;
            if (hipEventSynchronize(ob->ev[3]) == hipSuccess &&
                hipEventElapsedTime(&t3, ob->ev[0], ob->ev[1]) == hipSuccess &&
                hipEventElapsedTime(&t4, ob->ev[2], ob->ev[3]) == hipSuccess)
                ob->use4 = t4 < t3;
            else
                (void)hipGetLastError();
            ob->stage = 0;
        }
    }
    if (!nodep) {
        int rc = finalize_current(c);
        if (rc) return rc;
    }
    for (int s = 0; s < (int)c->push_out.size(); s++) {
        // the pushed state lives where the push wrote it
        Species &S = c->sp[s];
        const int b = S.cur;
        S.cell[b] = (int32_t *)c->push_out[s][0];
        S.ox[b] = c->push_out[s][1]; S.oy[b] = c->push_out[s][2];
        S.ux[b] = c->push_out[s][3]; S.uy[b] = c->push_out[s][4]; S.uz[b] = c->push_out[s][5];
    }
    c->push_out.clear();
    for (auto &S : c->sp) S.tbox_ok = true;
    HIPCHK(c, prof_mark(c, PBPIC_PROF_PUSH, false));
    return PBPIC_OK;
}


int fieldsolve_impl(pbpic_ctx *c)
{
    HIPCHK(c, prof_mark(c, PBPIC_PROF_FIELD, true));
    if (c->filt[0]) {
        const double cva2 = c->p.cva * c->p.cva, dt = c->p.dt;
        const double tau = 0.125 * dt;                       // pic2d.f90:895
        FilterParams q{};
        q.nx = c->p.nx; q.ny = c->p.ny;
        q.marder = (c->p.flags & PBPIC_FLAG_MARDER) ? 1 : 0;
        q.cherenkov = (c->p.flags & PBPIC_FLAG_CHERENKOV) ? 1 : 0;
        q.idx = c->idx; q.idy = c->idy;
        q.cva2in0 = cva2 * c->p.in0;
        q.mdx = dt * 0.001 / c->p.dx; q.mdy = dt * 0.001 / c->p.dy;   // :876-878
        q.cdx = dt * cva2 * tau / c->p.dx; q.cdy = dt * cva2 * tau / c->p.dy;   // :896-897
        q.ex = c->E[0]; q.ey = c->E[1]; q.ez = c->E[2];
        q.rho = c->J[3];
        q.fex = c->filt[0]; q.fey = c->filt[1]; q.fez = c->filt[2];
        HIPCHK(c, launch_filters(q, c->p.precision, c->st));
    }
    FdtdParams F = fdtd_params(c);
    F.je_pre = c->p.dt * c->p.mime / (c->p.cva * c->p.cva);
    HIPCHK(c, launch_fdtd(F, c->p.precision, c->st));
    for (int k = 0; k < 3; k++) {
        std::swap(c->E[k], c->E2[k]);
        std::swap(c->B[k], c->B2[k]);
    }
    if (walls(c)) HIPCHK(c, launch_wall_images(fdtd_params(c), 0x3f, c->st));
    HIPCHK(c, prof_mark(c, PBPIC_PROF_FIELD, false));
    c->steps++;
    return PBPIC_OK;
}

// One time-loop body (pic2d.f90:171-188).  Optionally copies J after the push
// (jcopy: 3 device planes) and the sort permutations (perm_dev: one device
// array per species) for the adjoint replay.  jsrc (the replay of a step
// whose J the checkpointed forward kept): the push does not deposit, and J
// is copied from jsrc before the field solve.
int step_impl(pbpic_ctx *c, void *const *jcopy, uint32_t *const *perm_dev, bool *sorted,
              void *const *jsrc = nullptr)
{
    HIPCHK(c, prof_mark(c, PBPIC_PROF_ZERO, true));
    // the push writes the current planes whole from the fixed-point sums
    // (FP64 form: two words per node)
    void *const *jz = c->J64;
    if (!jsrc)
        HIPCHK(c, launch_zero_current(jz[0], jz[1], jz[2], jz[3], ncell(c) * (c->rsz / 4), 8,
                                      c->acc, 1, c->st));
    HIPCHK(c, prof_mark(c, PBPIC_PROF_ZERO, false));
    // split form: when this step sorts, the push writes the particles as the
    // sort's records, so the sort skips its packing pass.  Until the sort
    // has run, the species' value arrays are stale.
    const bool recs = c->rsz == 4 && c->p.sort_every > 0 &&
                      (c->steps + 1) % c->p.sort_every == 0;
    int rc = push_impl(c, recs, jsrc != nullptr);
    if (rc) return rc;
    if (jsrc || jcopy) {
        CopyList cl{};
        for (int k = 0; k < 3; k++) {
            if (jsrc) cl.add(c->J[k], jsrc[k], (size_t)c->rsz * ncell(c));
            else cl.add(jcopy[k], c->J[k], (size_t)c->rsz * ncell(c));
        }
        HIPCHK(c, launch_copy_list(cl, c->st));
    }
    if (jsrc && jcopy) {
        CopyList cl{};
        for (int k = 0; k < 3; k++) cl.add(jcopy[k], c->J[k], (size_t)c->rsz * ncell(c));
        HIPCHK(c, launch_copy_list(cl, c->st));
    }
    if ((rc = fieldsolve_impl(c))) return rc;
    if (sorted) *sorted = false;
    if (c->p.sort_every > 0 && c->steps % c->p.sort_every == 0) {
        for (int s = 0; s < (int)c->sp.size(); s++)
            if ((rc = sort_one(c, s, nullptr, nullptr, perm_dev ? perm_dev[s] : nullptr, recs)))
                return rc;
        if (sorted) *sorted = true;
    }
    return PBPIC_OK;
}

// Flat device layout of one particle+field state (adjoint checkpoints/tape).
struct StateLayout {
    std::vector<size_t> sp_off;   // per species: start of [cell|ox|oy|ux|uy|uz|pid]
    size_t field_off = 0, total = 0;
};

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

StateLayout state_layout(pbpic_ctx *c)
{
    StateLayout L;
    size_t o = 0;
    for (auto &S : c->sp) {
        L.sp_off.push_back(o);
        const size_t n = (size_t)S.n;
        o += al256(4 * n) + 5 * al256((size_t)c->rsz * n) + (S.pid[0] ? al256(4 * n) : 0);
    }
    L.field_off = o;
    o += 6 * al256(8 * (size_t)ncell(c));
    L.total = o;
    return L;
}

// Pointers into a state buffer: part 0 = cell, 1..5 = dx..uz, 6 = pid;
// fields f0..f5 = Ex..Bz.
void *state_ptr(pbpic_ctx *c, const StateLayout &L, char *buf, int s, int part)
{
    const size_t n = (size_t)c->sp[s].n;
    size_t o = L.sp_off[s];
    if (part >= 1) o += al256(4 * n);
    if (part >= 2) o += (size_t)(part - 1) * al256((size_t)c->rsz * n);
    if (part == 6) o = L.sp_off[s] + al256(4 * n) + 5 * al256((size_t)c->rsz * n);
    return buf + o;
}
double *state_field(pbpic_ctx *c, const StateLayout &L, char *buf, int f)
{
    return (double *)(buf + L.field_off + (size_t)f * al256(8 * (size_t)ncell(c)));
}

enum { STATE_PARTICLES = 1, STATE_FIELDS = 2, STATE_ALL = 3 };

int state_copy(pbpic_ctx *c, const StateLayout &L, char *buf, bool save, int what = STATE_ALL)
{
    const size_t nc = (size_t)ncell(c);
    for (int s = 0; s < (int)c->sp.size() && (what & STATE_PARTICLES); s++) {
        Species &S = c->sp[s];
        const int b = S.cur;
        if (!save) S.tbox_ok = false;
        void *dev[7] = {S.cell[b], S.ox[b], S.oy[b], S.ux[b], S.uy[b], S.uz[b], S.pid[b]};
        CopyList cl{};
        for (int part = 0; part < 7; part++) {
            if (!dev[part] || S.n == 0) continue;
            const size_t bytes = (part == 0 || part == 6 ? 4 : (size_t)c->rsz) * (size_t)S.n;
            void *st = state_ptr(c, L, buf, s, part);
            if (save) cl.add(st, dev[part], bytes);
            else cl.add(dev[part], st, bytes);
        }
        HIPCHK(c, launch_copy_list(cl, c->st));
    }
    CopyList cl{};
    for (int f = 0; f < 6 && (what & STATE_FIELDS); f++) {
        double *dev = f < 3 ? c->E[f] : c->B[f - 3];
        double *st = state_field(c, L, buf, f);
        if (save) cl.add(st, dev, nc * 8);
        else cl.add(dev, st, nc * 8);
    }
    HIPCHK(c, launch_copy_list(cl, c->st));
    return PBPIC_OK;
}

// The adjoint replay points each species' current arrays into its tape (the
// push reads slot k and writes slot k+1, so the tape needs no copy per step).
// TapeGuard keeps the context's own arrays and puts them back: between steps
// (so that a sort writes into the context's spare arrays, never into the tape)
// and when the replay ends or fails.
struct TapeGuard {
    pbpic_ctx *c;
    std::vector<std::array<void *, 12>> own;
    explicit TapeGuard(pbpic_ctx *cc) : c(cc)
    {
        for (auto &S : c->sp)
            own.push_back({S.cell[0], S.cell[1], S.ox[0], S.ox[1], S.oy[0], S.oy[1],
                           S.ux[0], S.ux[1], S.uy[0], S.uy[1], S.uz[0], S.uz[1]});
    }
    void restore()
    {
        for (size_t s = 0; s < own.size(); s++) {
            Species &S = c->sp[s];
            const auto &o = own[s];
            for (int b = 0; b < 2; b++) {
                S.cell[b] = (int32_t *)o[0 + b];
                S.ox[b] = o[2 + b]; S.oy[b] = o[4 + b];
                S.ux[b] = o[6 + b]; S.uy[b] = o[8 + b]; S.uz[b] = o[10 + b];
            }
        }
    }
    // current arrays <- tape slot `from`; the next push writes slot `to`
    void step(const StateLayout &L, char *from, char *to)
    {
        restore();
        c->push_out.assign(c->sp.size(), {});
        for (int s = 0; s < (int)c->sp.size(); s++) {
            Species &S = c->sp[s];
            const int b = S.cur;
            S.cell[b] = (int32_t *)state_ptr(c, L, from, s, 0);
            S.ox[b] = state_ptr(c, L, from, s, 1); S.oy[b] = state_ptr(c, L, from, s, 2);
            S.ux[b] = state_ptr(c, L, from, s, 3); S.uy[b] = state_ptr(c, L, from, s, 4);
            S.uz[b] = state_ptr(c, L, from, s, 5);
            for (int q = 0; q < 6; q++) c->push_out[s][q] = state_ptr(c, L, to, s, q);
        }
    }
    ~TapeGuard()
    {
        restore();
        c->push_out.clear();
        for (auto &S : c->sp) S.tbox_ok = false;
    }
};

double je_total(pbpic_ctx *c, const double *a)
{
    return a[ACC_JE_INT] + a[ACC_JE_STEP] * (c->p.mime / (c->p.cva * c->p.cva)) * c->p.dt;
}

// The adjoint's loss: J.E weighted by the loss region when one is set.
double loss_total(pbpic_ctx *c, const double *a)
{
    if (!c->lmask) return je_total(c, a);
    return a[ACC_JEM_INT] + a[ACC_JEM_STEP] * (c->p.mime / (c->p.cva * c->p.cva)) * c->p.dt;
}

void adj_free(pbpic_ctx *c)
{
    for (auto p : c->adj.ckpts) (void)hipFree(p);
    c->adj.ckpts.clear();
    if (c->adj.final_state) (void)hipFree(c->adj.final_state);
    c->adj.final_state = nullptr;
    if (c->adj.jstore) (void)hipFree(c->adj.jstore);
    c->adj.jstore = nullptr;
    if (c->adj.tape) (void)hipFree(c->adj.tape);
    c->adj.tape = nullptr;
    c->adj.valid = false;
}

// Bytes of the backward's single allocation: the tape of ckpt + 1 states, the
// replayed steps' J when the forward kept none, the sort permutations of a
// segment, the field and particle adjoints, the wide-tile list.
struct AdjTapeSizes {
    size_t states, jtape, perm, fadj, padj, big;
    int nperm;
    size_t total(bool with_j) const { return states + (with_j ? jtape : 0) + perm + fadj + padj + big; }
};
static AdjTapeSizes adj_tape_sizes(pbpic_ctx *c, int K)
{
    const StateLayout L = state_layout(c);
    const size_t nc = (size_t)ncell(c), rz = (size_t)c->rsz;
    const int ns = (int)c->sp.size();
    int64_t ntot = 0, nmax = 0, atiles = 0;
    for (auto &S : c->sp) { ntot += S.n; nmax = std::max(nmax, S.n); atiles += adj_tiles(S.n); }
    AdjTapeSizes z{};
    z.states = (size_t)(K + 1) * L.total;
    z.jtape = (size_t)K * 3 * al256(rz * nc);
    // permutations only for the steps that sort: at most K/sort_every + 1 per segment
    z.nperm = c->p.sort_every > 0 ? std::min(K, K / c->p.sort_every + 1) : 0;
    z.perm = (size_t)z.nperm * al256(4 * (size_t)ntot + 256 * ns);
    z.fadj = 18 * al256(8 * nc);                                 // Ebar, Bbar, E', Jbar, node x6
    z.padj = 5 * al256(rz * (size_t)ntot + 256 * ns) + 5 * al256(rz * (size_t)nmax);
    z.big = al256(sizeof(int32_t) * (size_t)(atiles + 1));
    return z;
}

// Allocates the backward's buffers into c->adj.tape.  If the forward's kept J
// leaves too little room, it is dropped and the replay deposits again,
// keeping J in the tape instead.
static hipError_t adj_tape_alloc(pbpic_ctx *c, int K)
{
    if (c->adj.tape) return hipSuccess;
    const AdjTapeSizes z = adj_tape_sizes(c, K);
    hipError_t e = hipMalloc((void **)&c->adj.tape, z.total(!c->adj.jstore));
    if (e != hipSuccess && c->adj.jstore) {
        (void)hipGetLastError();
        (void)hipFree(c->adj.jstore);
        c->adj.jstore = nullptr;
        e = hipMalloc((void **)&c->adj.tape, z.total(true));
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c->adj.tape = nullptr;
    }
    return e;
}

// Host particle I/O.  The caller's arrays are pageable memory: they are
// moved in chunks through two pinned staging buffers, host threads filling or
// draining one buffer (with the grid <-> cell shift and any precision
// conversion) while the other is on its way over the link.
constexpr size_t STAGE_BYTES = (size_t)64 << 20;

// Staging buffers of at least min(want, STAGE_BYTES) bytes each (1 MiB
// granules), so that a small context pins little host memory.
int stage_init(pbpic_ctx *c, size_t want)
{
    const size_t need = std::min(STAGE_BYTES, ((want + (1u << 20) - 1) >> 20) << 20);
    if (c->stage[0] && c->stage_bytes >= need) return PBPIC_OK;
    for (int k = 0; k < 2; k++) {
        if (c->stage[k]) {
            if (c->stage_busy[k]) HIPCHK(c, hipEventSynchronize(c->stage_ev[k]));
            HIPCHK(c, hipHostFree(c->stage[k]));
            c->stage[k] = nullptr;
        }
        HIPCHK(c, hipHostMalloc((void **)&c->stage[k], need, hipHostMallocDefault));
        if (!c->stage_ev[k]) HIPCHK(c, hipEventCreateWithFlags(&c->stage_ev[k], hipEventDisableTiming));
        c->stage_busy[k] = false;
    }
    c->stage_bytes = need;
    return PBPIC_OK;
}

// Host threads for m elements: up to 8, one per 256 Ki elements.
int64_t host_threads(int64_t m)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    return std::max<int64_t>(1, std::min<int64_t>((int64_t)std::min(8u, hw), m / (1 << 18) + 1));
}

// Runs f(t, i0, i1) over [0, m) split in host_threads(m) parts (t = part).
template <typename F>
void host_parallel(int64_t m, F f)
{
    const int64_t nt = host_threads(m);
    if (nt <= 1) { f(0, (int64_t)0, m); return; }
    std::vector<std::thread> th;
    const int64_t per = (m + nt - 1) / nt;
    for (int64_t t = 0; t < nt; t++) {
        const int64_t i0 = t * per, i1 = std::min(m, i0 + per);
        if (i0 < i1) th.emplace_back(f, (int)t, i0, i1);
    }
    for (auto &x : th) x.join();
}

// n elements of type T to the device array dst; fill(T *stage, e0, i0, i1)
// writes elements e0+i0 .. e0+i1-1 to stage[i0 .. i1).  The copies are
// queued on the context's stream; the caller synchronizes it before the
// caller's arrays may change.
template <typename T, typename Fill>
int stage_h2d(pbpic_ctx *c, T *dst, int64_t n, Fill fill)
{
    if (n <= 0) return PBPIC_OK;
    int rc = stage_init(c, sizeof(T) * (size_t)n);
    if (rc) return rc;
    const int64_t per = (int64_t)(c->stage_bytes / sizeof(T));
    int k = 0;
    for (int64_t e0 = 0; e0 < n; e0 += per, k ^= 1) {
        const int64_t m = std::min(per, n - e0);
        if (c->stage_busy[k]) HIPCHK(c, hipEventSynchronize(c->stage_ev[k]));
        T *st = (T *)c->stage[k];
        host_parallel(m, [&, st, e0](int, int64_t i0, int64_t i1) { fill(st, e0, i0, i1); });
        HIPCHK(c, hipMemcpyAsync(dst + e0, st, sizeof(T) * m, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipEventRecord(c->stage_ev[k], c->st));
        c->stage_busy[k] = true;
    }
    return PBPIC_OK;
}

// n elements of type T from the device array src; drain(const T *stage, e0,
// i0, i1) consumes elements e0+i0 .. e0+i1-1 from stage[i0 .. i1).
template <typename T, typename Drain>
int stage_d2h(pbpic_ctx *c, const T *src, int64_t n, Drain drain)
{
    if (n <= 0) return PBPIC_OK;
    int rc = stage_init(c, sizeof(T) * (size_t)n);
    if (rc) return rc;
    const int64_t per = (int64_t)(c->stage_bytes / sizeof(T));
    int64_t pend0 = -1, pendm = 0;       // chunk waiting in buffer k ^ 1
    int k = 0;
    auto flush = [&](int b, int64_t e0, int64_t m) -> int {
        HIPCHK(c, hipEventSynchronize(c->stage_ev[b]));
        const T *st = (const T *)c->stage[b];
        host_parallel(m, [&, st, e0](int, int64_t i0, int64_t i1) { drain(st, e0, i0, i1); });
        return PBPIC_OK;
    };
    for (int64_t e0 = 0; e0 < n; e0 += per, k ^= 1) {
        const int64_t m = std::min(per, n - e0);
        HIPCHK(c, hipMemcpyAsync(c->stage[k], src + e0, sizeof(T) * m, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipEventRecord(c->stage_ev[k], c->st));
        c->stage_busy[k] = true;
        if (pend0 >= 0 && (rc = flush(k ^ 1, pend0, pendm))) return rc;
        pend0 = e0;
        pendm = m;
    }
    if (pend0 >= 0 && (rc = flush(k ^ 1, pend0, pendm))) return rc;
    return PBPIC_OK;
}

template <typename H>
int set_particles_t(pbpic_ctx *c, int32_t s, int64_t n, const int32_t *grid, const H *dx,
                    const H *dy, const H *ux, const H *uy, const H *uz, const int32_t *pid)
{
    int rc = valid_species(c, s);
    if (rc) return rc;
    Species &S = c->sp[s];
    if (n < 0 || n > S.cap) return fail(c, PBPIC_ERR_CAPACITY, "n=%lld exceeds capacity %lld", (long long)n, (long long)S.cap);
    if (n > 0 && (!grid || !dx || !dy || !ux || !uy || !uz)) return fail(c, PBPIC_ERR_ARG, "null particle array");
    const int64_t nc = ncell(c);
    const int ny = c->p.ny, nyp = nyp_of(c);
    // every grid value checked before any state changes: in 1..nc and, with
    // conducting walls, not in a guard row (first offender per host thread)
    std::vector<int64_t> first((size_t)host_threads(n), -1);
    host_parallel(n, [&](int t, int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; i++) {
            const int32_t g = grid[i];
            if (g < 1 || g > nc || (g - 1) % ny >= nyp) { first[(size_t)t] = i; return; }
        }
    });
    for (int64_t bad : first) {
        if (bad < 0) continue;
        if (grid[bad] >= 1 && grid[bad] <= nc)
            return fail(c, PBPIC_ERR_CELL, "grid[%lld]=%d lies in a guard row of the conducting walls",
                        (long long)bad, grid[bad]);
        return fail(c, PBPIC_ERR_CELL, "grid[%lld]=%d out of 1..%lld", (long long)bad, grid[bad], (long long)nc);
    }
    const int b = S.cur;
    S.tbox_ok = false;
    HIPCHK(c, hipStreamSynchronize(c->st));
    rc = stage_h2d<int32_t>(c, S.cell[b], n, [grid](int32_t *st, int64_t e0, int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; i++) st[i] = grid[e0 + i] - 1;
    });
    if (rc) return rc;
    const H *src[5] = {dx, dy, ux, uy, uz};
    void *dst[5] = {S.ox[b], S.oy[b], S.ux[b], S.uy[b], S.uz[b]};
    for (int k = 0; k < 5; k++) {
        const H *a = src[k];
        if (c->rsz == 8)
            rc = stage_h2d<double>(c, (double *)dst[k], n, [a](double *st, int64_t e0, int64_t i0, int64_t i1) {
                for (int64_t i = i0; i < i1; i++) st[i] = (double)a[e0 + i];
            });
        else
            rc = stage_h2d<float>(c, (float *)dst[k], n, [a](float *st, int64_t e0, int64_t i0, int64_t i1) {
                for (int64_t i = i0; i < i1; i++) st[i] = (float)a[e0 + i];
            });
        if (rc) return rc;
    }
    if (S.pid[b]) {
        rc = stage_h2d<int32_t>(c, S.pid[b], n, [pid](int32_t *st, int64_t e0, int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; i++) st[i] = pid ? pid[e0 + i] : (int32_t)(e0 + i + 1);
        });
        if (rc) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    S.n = n;
    return PBPIC_OK;
}

template <typename H>
int get_particles_t(pbpic_ctx *c, int32_t s, int64_t cap, int64_t *nout, int32_t *grid, H *dx,
                    H *dy, H *ux, H *uy, H *uz, int32_t *pid)
{
    int rc = valid_species(c, s);
    if (rc) return rc;
    if ((rc = check_flags(c))) return rc;
    Species &S = c->sp[s];
    if (nout) *nout = S.n;
    const int64_t n = S.n;
    if (cap < n) return fail(c, PBPIC_ERR_CAPACITY, "host capacity %lld < %lld particles", (long long)cap, (long long)n);
    const int b = S.cur;
    if (grid)
        if ((rc = stage_d2h<int32_t>(c, S.cell[b], n, [grid](const int32_t *st, int64_t e0, int64_t i0, int64_t i1) {
                 for (int64_t i = i0; i < i1; i++) grid[e0 + i] = st[i] + 1;
             })))
            return rc;
    H *dst[5] = {dx, dy, ux, uy, uz};
    void *src[5] = {S.ox[b], S.oy[b], S.ux[b], S.uy[b], S.uz[b]};
    for (int k = 0; k < 5; k++) {
        H *a = dst[k];
        if (!a) continue;
        if (c->rsz == 8)
            rc = stage_d2h<double>(c, (const double *)src[k], n, [a](const double *st, int64_t e0, int64_t i0, int64_t i1) {
                for (int64_t i = i0; i < i1; i++) a[e0 + i] = (H)st[i];
            });
        else
            rc = stage_d2h<float>(c, (const float *)src[k], n, [a](const float *st, int64_t e0, int64_t i0, int64_t i1) {
                for (int64_t i = i0; i < i1; i++) a[e0 + i] = (H)st[i];
            });
        if (rc) return rc;
    }
    if (pid) {
        if (S.pid[b]) {
            if ((rc = stage_d2h<int32_t>(c, S.pid[b], n, [pid](const int32_t *st, int64_t e0, int64_t i0, int64_t i1) {
                     for (int64_t i = i0; i < i1; i++) pid[e0 + i] = st[i];
                 })))
                return rc;
        } else {
            for (int64_t i = 0; i < n; i++) pid[i] = 0;
        }
    }
    return PBPIC_OK;
}

// pbpic_deposit's inputs as staged on the device: 5 int32 and 5 double arrays
// per entry.
constexpr size_t DEPOSIT_ENTRY_BYTES = 5 * 4 + 5 * 8;

int deposit_check(pbpic_ctx *c, int64_t n, const int32_t *i1, const int32_t *i2,
                  const int32_t *j1, const int32_t *j2, const double *dx1, const double *dx2,
                  const double *dy1, const double *dy2, const int32_t *q, const double *vz)
{
    if (!c) return PBPIC_ERR_ARG;
    if (n < 0) return fail(c, PBPIC_ERR_ARG, "deposit count %lld < 0", (long long)n);
    if (n == 0) return PBPIC_OK;
    if (!i1 || !i2 || !j1 || !j2 || !dx1 || !dx2 || !dy1 || !dy2 || !q || !vz)
        return fail(c, PBPIC_ERR_ARG, "null deposit array");
    const int nx = c->p.nx, ny = c->p.ny;
    for (int64_t k = 0; k < n; k++) {
        if (i1[k] < 1 || i1[k] > nx || j1[k] < 1 || j1[k] > ny || std::abs(i2[k] - i1[k]) > 1 ||
            std::abs(j2[k] - j1[k]) > 1 || (q[k] != 1 && q[k] != -1))
            return fail(c, PBPIC_ERR_ARG, "deposit entry %lld out of range", (long long)k);
    }
    return PBPIC_OK;
}

// Uploads n checked entries into buf (n * DEPOSIT_ENTRY_BYTES bytes of device
// memory), deposits them into J (and rho) and waits for the stream.
int deposit_staged(pbpic_ctx *c, void *buf, int64_t n, const int32_t *i1, const int32_t *i2,
                   const int32_t *j1, const int32_t *j2, const double *dx1, const double *dx2,
                   const double *dy1, const double *dy2, const int32_t *q, const double *vz)
{
    const size_t ib = (size_t)n * 4, db = (size_t)n * 8;
    char *b = (char *)buf;
    int32_t *d_i1 = (int32_t *)b, *d_i2 = (int32_t *)(b + ib), *d_j1 = (int32_t *)(b + 2 * ib),
            *d_j2 = (int32_t *)(b + 3 * ib), *d_q = (int32_t *)(b + 4 * ib);
    // the doubles start at an 8-byte boundary: 5 ib = 20 n bytes, padded
    double *d_dx1 = (double *)(b + ((5 * ib + 7) & ~(size_t)7)), *d_dx2 = d_dx1 + n,
           *d_dy1 = d_dx2 + n, *d_dy2 = d_dy1 + n, *d_vz = d_dy2 + n;
    const void *src[10] = {i1, i2, j1, j2, q, dx1, dx2, dy1, dy2, vz};
    void *dst[10] = {d_i1, d_i2, d_j1, d_j2, d_q, d_dx1, d_dx2, d_dy1, d_dy2, d_vz};
    for (int k = 0; k < 10; k++)
        HIPCHK(c, hipMemcpyAsync(dst[k], src[k], k < 5 ? ib : db, hipMemcpyHostToDevice, c->st));
    PushParams P = push_params(c);
    P.sp[0].qx = c->p.dx / c->p.dt;
    P.sp[0].qy = c->p.dy / c->p.dt;
    HIPCHK(c, launch_deposit_batch(P, c->p.precision, n, d_i1, d_i2, d_j1, d_j2, d_dx1, d_dx2,
                                   d_dy1, d_dy2, d_q, d_vz, c->st));
    int rc = finalize_current(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PBPIC_OK;
}

}  // namespace

extern "C" {

void pbpic_params_default(pbpic_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->nx = 256;                 // pic2d.f90:12
    p->ny = 256;
    p->precision = PBPIC_PREC_FP64;
    p->compat = PBPIC_COMPAT_REFERENCE;
    p->dx = 0.04;                // :15
    p->dy = 0.04;
    p->mime = 25.0;              // :19
    p->cva = 15.0;               // :20
    p->in0 = 1.0 / 291.0;        // :31
    p->dt = 0.045 / 25.0;        // :42
    p->sort_every = 50;          // :180
    p->flags = PBPIC_FLAG_DEPOSIT_RHO | PBPIC_FLAG_PUSH_ORDER;
}

int pbpic_ctx_create(const pbpic_params *pp, pbpic_ctx **out)
{
    if (!pp || !out) return PBPIC_ERR_ARG;
    *out = nullptr;
    const pbpic_params &p = *pp;
    if (p.nx < 2 || p.ny < 2 || (int64_t)p.nx * p.ny >= (1ll << 30)) return PBPIC_ERR_ARG;
    if (!(p.dx > 0) || !(p.dy > 0) || !(p.dt > 0) || !(p.cva > 0) || !(p.mime > 0) || !(p.in0 > 0))
        return PBPIC_ERR_ARG;
    if (p.precision != PBPIC_PREC_FP64 && p.precision != PBPIC_PREC_SPLIT32) return PBPIC_ERR_ARG;
    if (p.compat & ~PBPIC_COMPAT_REFERENCE) return PBPIC_ERR_ARG;
    for (int k = 0; k < 7; k++)
        if (p.reserved[k]) return PBPIC_ERR_ARG;
    if (p.ybound != PBPIC_YBOUND_PERIODIC && p.ybound != PBPIC_YBOUND_CONDUCTING) return PBPIC_ERR_ARG;
    if (p.ybound == PBPIC_YBOUND_CONDUCTING && p.ny < 4) return PBPIC_ERR_ARG;
    if ((p.flags & PBPIC_FLAG_MARDER) && !(p.flags & PBPIC_FLAG_DEPOSIT_RHO)) return PBPIC_ERR_ARG;
    pbpic_ctx *c = new pbpic_ctx();
    c->p = p;
    c->rsz = p.precision == PBPIC_PREC_FP64 ? 8 : 4;
    // window capacity in nodes: split form, three push workgroups share a
    // CU's 160 KiB of LDS (896 nodes of 48 bytes and four crossing queues:
    // 52 KiB; push_lds_bytes without the rho plane); FP64 form, two (the
    // FP64 push runs as fast at two workgroups per CU, and the wider window
    // defers fewer tiles once the cells have spread)
    c->maxw = p.precision == PBPIC_PREC_FP64
                  ? push_maxw_mid(p.precision, (p.flags & PBPIC_FLAG_DEPOSIT_RHO) ? 1 : 0, 1)
                  : 896;
    c->idx = 1.0 / p.dx;                   // pic2d.f90:16
    c->idy = 1.0 / p.dy;
    c->dth = p.dt / 2.0;                   // :43
    c->dtc2 = p.dt * (p.cva * p.cva);      // dt*cva2 (:742)
    c->dtcva = p.dt * p.cva;               // (:1003)
    int rc = PBPIC_OK;
    auto chk = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == PBPIC_OK) rc = fail(c, e == hipErrorOutOfMemory ? PBPIC_ERR_NOMEM : PBPIC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    chk(hipGetDevice(&c->dev), "hipGetDevice");
    chk(hipStreamCreate(&c->st), "hipStreamCreate");
    chk(scratch_address(c->dev, &c->scratch), "hipGetSymbolAddress(scratch)");
    const size_t nc = (size_t)p.nx * p.ny;
    for (int k = 0; k < 3; k++) {
        chk(dalloc(&c->E[k], nc * 8), "alloc E");
        chk(dalloc(&c->B[k], nc * 8), "alloc B");
        chk(dalloc(&c->E2[k], nc * 8), "alloc E");
        chk(dalloc(&c->B2[k], nc * 8), "alloc B");
    }
    for (int k = 0; k < 4; k++) chk(dalloc(&c->J[k], nc * c->rsz), "alloc J");
    {
        // the fixed-point sums: one word per node (split form), two (FP64
        // form: high words, then low words)
        const size_t words = c->rsz / 4;
        for (int k = 0; k < 4; k++) chk(dalloc(&c->J64[k], nc * 8 * words), "alloc J sums");
        // one contribution of each plane stays below 2^30 (bounds as in
        // species_dev, |q| = 1): |jx| <= dx/dt, |jy| <= dy/dt, |jz| <= cva, |rho| <= 1
        c->js[0] = fx_scale(p.dx / p.dt);
        c->js[1] = fx_scale(p.dy / p.dt);
        c->js[2] = fx_scale(p.cva);
        c->js[3] = fx_scale(1.0);
    }
    chk(dalloc(&c->NF, nodefield_bytes(c->p.precision, nc)), "alloc node fields");
    if (p.flags & (PBPIC_FLAG_MARDER | PBPIC_FLAG_CHERENKOV))
        for (int k = 0; k < 3; k++) chk(dalloc(&c->filt[k], nc * 8), "alloc filter planes");
    chk(dalloc(&c->acc, ACC_N * sizeof(double)), "alloc acc");
    {
        // two partials per workgroup of the widest reduction launch: the
        // FDTD tiles (32 x 16 cells) or 4096 blocks
        const int64_t nbf = (int64_t)((p.nx + 31) / 32) * ((p.ny + 15) / 16);
        chk(dalloc(&c->part, (size_t)std::max<int64_t>(nbf, 4096) * 2 * sizeof(double)),
            "alloc partials");
    }
    chk(dalloc(&c->err, sizeof(int) * 4), "alloc err");
    for (auto &b : c->occ)
        for (auto &e : b.ev) chk(hipEventCreate(&e), "create push timing events");

    {
        CellKey key;                       // counts are indexed by sort key
        key.init(p.nx, p.ny, 1);
        chk(dalloc(&c->sort.counts, (size_t)std::max<int64_t>(key.nkeys, nc) * sizeof(int32_t)),
            "alloc counts");
        // the push order's supertile ranks (CellKey), host and device copies
        const int nbx = (p.nx + 7) / 8, nby = (p.ny + 7) / 8;
        supertile_ranks(nbx, nby, c->sort.strank_h);
        chk(dalloc(&c->sort.strank, c->sort.strank_h.size() * sizeof(uint32_t)), "alloc ranks");
        if (rc == PBPIC_OK)
            chk(hipMemcpy(c->sort.strank, c->sort.strank_h.data(),
                          c->sort.strank_h.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
                "copy ranks");
    }
    chk(dalloc(&c->sort.dhist, 4 * 256 * sizeof(uint32_t)), "alloc dhist");
    chk(dalloc(&c->sort.ticket, 8 * sizeof(uint32_t)), "alloc ticket");
    if (rc == PBPIC_OK) {
        for (int k = 0; k < 3; k++) {
            chk(hipMemsetAsync(c->E[k], 0, nc * 8, c->st), "zero E");
            chk(hipMemsetAsync(c->B[k], 0, nc * 8, c->st), "zero B");
        }
        for (int k = 0; k < 4; k++) chk(hipMemsetAsync(c->J[k], 0, nc * c->rsz, c->st), "zero J");
        for (int k = 0; k < 4 && c->J64[0]; k++)
            chk(hipMemsetAsync(c->J64[k], 0, nc * 8 * (c->rsz / 4), c->st), "zero J sums");
        chk(hipMemsetAsync(c->acc, 0, ACC_N * sizeof(double), c->st), "zero acc");
        chk(hipMemsetAsync(c->err, 0, sizeof(int) * 4, c->st), "zero err");
        chk(hipStreamSynchronize(c->st), "sync");
    }
    if (rc != PBPIC_OK) {
        std::string msg = c->last_error;
        pbpic_ctx_destroy(c);
        fprintf(stderr, "pbpic_ctx_create: %s\n", msg.c_str());
        return rc;
    }
    *out = c;
    return PBPIC_OK;
}

int pbpic_ctx_destroy(pbpic_ctx *c)
{
    if (!c) return PBPIC_ERR_ARG;
    if (c->st) (void)hipStreamSynchronize(c->st);
    auto f = [](void *p) { if (p) (void)hipFree(p); };
    for (int k = 0; k < 3; k++) { f(c->E[k]); f(c->B[k]); f(c->E2[k]); f(c->B2[k]); }
    for (int k = 0; k < 4; k++) { f(c->J[k]); f(c->J64[k]); }
    f(c->NF);
    f(c->acc); f(c->part); f(c->err); f(c->rho_scratch); f(c->gauss0); f(c->diag_counts);
    f(c->filt[0]); f(c->filt[1]); f(c->filt[2]); f(c->big); f(c->lmask);
    for (auto &S : c->sp)
        for (int b = 0; b < 2; b++) {
            f(S.cell[b]); f(S.ox[b]); f(S.oy[b]); f(S.ux[b]); f(S.uy[b]); f(S.uz[b]); f(S.pid[b]);
        }
    for (auto &S : c->sp) { f(S.tbox); f(S.rec); }
    for (int b = 0; b < 2; b++) { f(c->sort.keys[b]); f(c->sort.vals[b]); }
    f(c->sort.status); f(c->sort.dhist); f(c->sort.ticket); f(c->sort.counts);
    f(c->sort.tarea); f(c->sort.toff); f(c->sort.bsum); f(c->sort.ctl); f(c->sort.rbuf);
    f(c->sort.strank);
    for (auto p : c->adj.ckpts) f(p);
    f(c->adj.final_state); f(c->adj.jstore); f(c->adj.tape);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto &b : c->occ)
        for (auto e : b.ev) if (e) (void)hipEventDestroy(e);
    for (int k = 0; k < 2; k++) {
        if (c->stage[k]) (void)hipHostFree(c->stage[k]);
        if (c->stage_ev[k]) (void)hipEventDestroy(c->stage_ev[k]);
    }
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return PBPIC_OK;
}

const char *pbpic_last_error(const pbpic_ctx *c) { return c ? c->last_error.c_str() : "null context"; }

int pbpic_sync(pbpic_ctx *c)
{
    if (!c) return PBPIC_ERR_ARG;
    return check_flags(c);
}

int pbpic_add_species(pbpic_ctx *c, int32_t q, int64_t capacity, int32_t *species)
{
    if (!c) return PBPIC_ERR_ARG;
    if (q != 1 && q != -1) return fail(c, PBPIC_ERR_ARG, "species charge must be +1 or -1, got %d", q);
    if (capacity < 0 || capacity >= (1ll << 30)) return fail(c, PBPIC_ERR_CAPACITY, "capacity %lld out of range", (long long)capacity);
    if ((int)c->sp.size() >= MAX_SPECIES) return fail(c, PBPIC_ERR_ARG, "at most %d species", MAX_SPECIES);
    Species S;
    S.q = q;
    S.cap = capacity;
    const pbpic_params &p = c->p;
    S.ms = q > 0 ? p.mime : 1.0;                         // pic2d.f90:964-968
    S.c0 = (double)q * p.dt * p.mime / S.ms / 2.0;       // :969
    S.c1 = S.c0 / p.cva;                                 // :970
    S.qx = (double)q * p.dx / p.dt;                      // :1237
    S.qy = (double)q * p.dy / p.dt;
    const size_t n = (size_t)capacity;
    for (int b = 0; b < 2; b++) {
        HIPCHK(c, dalloc(&S.cell[b], n * 4));
        HIPCHK(c, dalloc(&S.ox[b], n * c->rsz));
        HIPCHK(c, dalloc(&S.oy[b], n * c->rsz));
        HIPCHK(c, dalloc(&S.ux[b], n * c->rsz));
        HIPCHK(c, dalloc(&S.uy[b], n * c->rsz));
        HIPCHK(c, dalloc(&S.uz[b], n * c->rsz));
        if (p.flags & PBPIC_FLAG_TRACK_PID) HIPCHK(c, dalloc(&S.pid[b], n * 4));
    }
    HIPCHK(c, dalloc(&S.tbox, (n / PUSH_TILE + 1) * 4 * sizeof(int32_t)));
    // split form: the sort's gather reads 24-byte particle records
    if (c->rsz == 4) HIPCHK(c, dalloc(&S.rec, n * sizeof(PartRec)));
    // sort scratch sized for the largest species
    if (capacity > c->sort.cap) {
        for (int b = 0; b < 2; b++) {
            if (c->sort.keys[b]) (void)hipFree(c->sort.keys[b]);
            if (c->sort.vals[b]) (void)hipFree(c->sort.vals[b]);
            HIPCHK(c, dalloc(&c->sort.keys[b], n * 4));
            HIPCHK(c, dalloc(&c->sort.vals[b], n * 4));
        }
        if (c->sort.status) (void)hipFree(c->sort.status);
        const int64_t ntile = (capacity + 4095) / 4096 + 1;
        HIPCHK(c, dalloc(&c->sort.status, (size_t)ntile * 256 * 4));
        for (uint32_t **q : {&c->sort.tarea, &c->sort.toff, &c->sort.bsum}) {
            if (*q) (void)hipFree(*q);
            HIPCHK(c, dalloc(q, (size_t)(ntile + 1) * 4));
        }
        if (!c->sort.ctl) HIPCHK(c, dalloc(&c->sort.ctl, 8 * sizeof(int)));
        c->sort.cap = capacity;
        c->sort.status_cap = ntile;
    }
    c->sp.push_back(S);
    if (species) *species = (int32_t)c->sp.size() - 1;
    return PBPIC_OK;
}

int pbpic_num_particles(pbpic_ctx *c, int32_t s, int64_t *n)
{
    int rc = valid_species(c, s);
    if (rc) return rc;
    if (n) *n = c->sp[s].n;
    return PBPIC_OK;
}

int pbpic_set_particles(pbpic_ctx *c, int32_t s, int64_t n, const int32_t *grid, const double *dx,
                        const double *dy, const double *ux, const double *uy, const double *uz,
                        const int32_t *pid)
{
    return set_particles_t<double>(c, s, n, grid, dx, dy, ux, uy, uz, pid);
}

int pbpic_set_particles_f32(pbpic_ctx *c, int32_t s, int64_t n, const int32_t *grid,
                            const float *dx, const float *dy, const float *ux, const float *uy,
                            const float *uz, const int32_t *pid)
{
    return set_particles_t<float>(c, s, n, grid, dx, dy, ux, uy, uz, pid);
}

int pbpic_get_particles(pbpic_ctx *c, int32_t s, int64_t cap, int64_t *n, int32_t *grid, double *dx,
                        double *dy, double *ux, double *uy, double *uz, int32_t *pid)
{
    return get_particles_t<double>(c, s, cap, n, grid, dx, dy, ux, uy, uz, pid);
}

int pbpic_get_particles_f32(pbpic_ctx *c, int32_t s, int64_t cap, int64_t *n, int32_t *grid,
                            float *dx, float *dy, float *ux, float *uy, float *uz, int32_t *pid)
{
    return get_particles_t<float>(c, s, cap, n, grid, dx, dy, ux, uy, uz, pid);
}

int pbpic_init_maxwellian(pbpic_ctx *c, int32_t s, int32_t ppc, double vth, double vdrift_x,
                          int32_t counter_stream, uint64_t seed, uint32_t stream)
{
    int rc = valid_species(c, s);
    if (rc) return rc;
    if (ppc < 1 || !(vth >= 0 && vth < 1) || !(fabs(vdrift_x) < 1))
        return fail(c, PBPIC_ERR_ARG, "bad loader arguments");
    Species &S = c->sp[s];
    const int64_t n = (int64_t)c->p.nx * nyp_of(c) * (int64_t)ppc;
    if (n > S.cap) return fail(c, PBPIC_ERR_CAPACITY, "ppc*cells=%lld exceeds capacity %lld", (long long)n, (long long)S.cap);
    S.n = n;
    S.tbox_ok = false;
    SpeciesDev d = species_dev(c, s, S.cur);
    HIPCHK(c, launch_init_maxwellian(d, c->p.precision, c->p.nx, c->p.ny, ppc, vth, vdrift_x,
                                     counter_stream, seed, stream, nyp_of(c), c->st));
    return PBPIC_OK;
}

int pbpic_init_harris(pbpic_ctx *c, int32_t si, int32_t se, const pbpic_harris *h)
{
    int rc = valid_species(c, si);
    if (rc) return rc;
    if ((rc = valid_species(c, se))) return rc;
    if (!h || si == se) return fail(c, PBPIC_ERR_ARG, "bad Harris arguments");
    if (c->sp[si].q != 1 || c->sp[se].q != -1)
        return fail(c, PBPIC_ERR_ARG, "the ion species must have q = +1, the electron species q = -1");
    for (int k = 0; k < 7; k++)
        if (h->reserved[k] || h->reserved0) return fail(c, PBPIC_ERR_ARG, "reserved fields must be zero");
    if (h->geometry == PBPIC_HARRIS_DOUBLE_X) {
        if (walls(c)) return fail(c, PBPIC_ERR_ARG, "the double sheet across x needs periodic y");
    } else if (h->geometry == PBPIC_HARRIS_SINGLE_Y) {
        if (!walls(c)) return fail(c, PBPIC_ERR_ARG, "the sheet across y needs conducting y walls");
    } else {
        return fail(c, PBPIC_ERR_ARG, "unknown Harris geometry %d", h->geometry);
    }
    const double sp[6] = {h->vthe_cs, h->vthi_cs, h->vthe_bg, h->vthi_bg, h->vde_z, h->vdi_z};
    for (double v : sp)
        if (!(fabs(v) < 1)) return fail(c, PBPIC_ERR_ARG, "speeds must lie below c");
    if (h->ppc < 1 || !(h->half_width > 0) || !(h->nb >= 0) || !std::isfinite(h->b0) ||
        !std::isfinite(h->psi0))
        return fail(c, PBPIC_ERR_ARG, "bad Harris parameters");
    const int nx = c->p.nx, ny = c->p.ny, nyp = nyp_of(c);
    const bool dbl = h->geometry == PBPIC_HARRIS_DOUBLE_X;
    const int64_t n = (int64_t)h->ppc * nx * nyp;
    for (int32_t s : {si, se})
        if (n > c->sp[s].cap)
            return fail(c, PBPIC_ERR_CAPACITY, "ppc*cells=%lld exceeds capacity %lld", (long long)n,
                        (long long)c->sp[s].cap);
    // sheet particles: 2 L nsheets / (2 L nsheets + nb L_n) of the total,
    // rounded up and made even (pic2d.f90:375-376)
    const int nsh = dbl ? 2 : 1;
    const double ln = dbl ? nx * c->p.dx : nyp * c->p.dy;
    const double f = 2.0 * h->half_width * nsh / (2.0 * h->half_width * nsh + h->nb * ln);
    int64_t ncs = (int64_t)ceil((double)n * f);
    ncs = (ncs / 2) * 2;
    if (ncs > n) ncs = n - (n & 1);
    auto uc = [](double v) { return v / sqrt(1.0 - v * v); };
    HarrisDev d{};
    d.geometry = h->geometry;
    d.ppc = h->ppc;
    d.thermal_z = h->thermal_z ? 1 : 0;
    d.nn = dbl ? nx : nyp;
    d.nt = dbl ? nyp : nx;
    d.ncs = ncs;
    d.L = h->half_width;
    d.dn = dbl ? c->p.dx : c->p.dy;
    d.xc0 = dbl ? 0.25 * (nx * c->p.dx) : 0.5 * (nyp * c->p.dy);
    d.xc1 = dbl ? 0.75 * (nx * c->p.dx) : d.xc0;
    d.uec = uc(h->vthe_cs); d.uic = uc(h->vthi_cs);
    d.ueb = uc(h->vthe_bg); d.uib = uc(h->vthi_bg);
    d.ude = uc(h->vde_z); d.udi = uc(h->vdi_z);
    d.k0 = (uint32_t)h->seed; d.k1 = (uint32_t)(h->seed >> 32);
    d.stream = h->stream;
    Species &I = c->sp[si], &E = c->sp[se];
    I.n = E.n = n;
    I.tbox_ok = E.tbox_ok = false;
    HIPCHK(c, launch_init_harris(species_dev(c, si, I.cur), species_dev(c, se, E.cur),
                                 c->p.precision, ny, d, c->st));
    // fields (initfields, :583-648): E = 0 and the sheets' B
    const size_t nc = (size_t)ncell(c);
    std::vector<double> bx(nc, 0.0), by(nc, 0.0), zero(nc, 0.0);
    const double L = h->half_width, b0 = h->b0, psi0 = h->psi0, dx = c->p.dx, dy = c->p.dy;
    const double twopi = 6.283185307179586;
    if (dbl) {
        const double lx = nx * dx, xl = 0.25 * lx, xr = 0.75 * lx;
        for (int gx = 0; gx < nx; gx++) {
            const double x = gx * dx;                                       // :608
            const double v = b0 * (tanh((x - xl) / L) - tanh((x - xr) / L) - 1.0);
            for (int gy = 0; gy < ny; gy++) by[(size_t)gy * nx + gx] = v;
        }
    } else {
        // Bx on node rows (walls included), By on half rows, both from the
        // discrete curl of psi taken at cell centres
        const double lx = nx * dx, ly = nyp * dy, yc = 0.5 * ly;
        auto psi = [&](double x, double y) {
            return psi0 * cos(twopi * x / lx) * cos(0.5 * twopi * (y - yc) / ly);
        };
        for (int gy = 0; gy <= nyp; gy++)
            for (int gx = 0; gx < nx; gx++) {
                const double x = (gx + 0.5) * dx, y = gy * dy;
                bx[(size_t)gy * nx + gx] =
                    b0 * tanh((y - yc) / L) - (psi(x, y + 0.5 * dy) - psi(x, y - 0.5 * dy)) / dy;
            }
        for (int gy = 0; gy < nyp; gy++)
            for (int gx = 0; gx < nx; gx++) {
                const double x = gx * dx, y = (gy + 0.5) * dy;
                by[(size_t)gy * nx + gx] = (psi(x + 0.5 * dx, y) - psi(x - 0.5 * dx, y)) / dx;
            }
    }
    for (int k = 0; k < 3; k++)
        HIPCHK(c, hipMemcpyAsync(c->E[k], zero.data(), nc * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->B[0], bx.data(), nc * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->B[1], by.data(), nc * 8, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->B[2], zero.data(), nc * 8, hipMemcpyHostToDevice, c->st));
    if (walls(c)) HIPCHK(c, launch_wall_images(fdtd_params(c), 0x3f, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PBPIC_OK;
}

int pbpic_set_field(pbpic_ctx *c, int32_t which, const double *plane)
{
    if (!c || !plane) return PBPIC_ERR_ARG;
    const size_t nc = (size_t)ncell(c);
    if (which >= PBPIC_EX && which <= PBPIC_EZ) {
        HIPCHK(c, hipMemcpy(c->E[which - PBPIC_EX], plane, nc * 8, hipMemcpyHostToDevice));
        if (walls(c)) HIPCHK(c, launch_wall_images(fdtd_params(c), 1 << which, c->st));
    } else if (which >= PBPIC_BX && which <= PBPIC_BZ) {
        HIPCHK(c, hipMemcpy(c->B[which - PBPIC_BX], plane, nc * 8, hipMemcpyHostToDevice));
        if (walls(c)) HIPCHK(c, launch_wall_images(fdtd_params(c), 1 << which, c->st));
    } else if (which >= PBPIC_JX && which <= PBPIC_RHO) {
        // the device holds the unnormalized sum; the boundary value is x in0
        std::vector<double> vd(nc);
        std::vector<float> vf;
        for (size_t k = 0; k < nc; k++) vd[k] = plane[k] / c->p.in0;
        void *dst = c->J[which - PBPIC_JX];
        const double sc = c->js[which - PBPIC_JX];
        if (c->rsz == 8) {
            HIPCHK(c, hipMemcpy(dst, vd.data(), nc * 8, hipMemcpyHostToDevice));
            // the two-word fixed-point sums a following push adds to: the
            // scaled value x as floor(x) and rint((x - floor(x)) 2^FX2_LOBITS)
            std::vector<long long> vq(2 * nc);
            for (size_t k = 0; k < nc; k++) {
                const double x = vd[k] * sc, h = std::floor(x);
                vq[k] = (long long)h;
                vq[nc + k] = llrint((x - h) * (double)(1 << FX2_LOBITS));
            }
            HIPCHK(c, hipMemcpy(c->J64[which - PBPIC_JX], vq.data(), 2 * nc * 8, hipMemcpyHostToDevice));
        } else {
            vf.assign(vd.begin(), vd.end());
            HIPCHK(c, hipMemcpy(dst, vf.data(), nc * 4, hipMemcpyHostToDevice));
            // the fixed-point sums a following push adds to
            std::vector<long long> vq(nc);
            for (size_t k = 0; k < nc; k++) vq[k] = llrint((double)vf[k] * sc);
            HIPCHK(c, hipMemcpy(c->J64[which - PBPIC_JX], vq.data(), nc * 8, hipMemcpyHostToDevice));
        }
    } else {
        return fail(c, PBPIC_ERR_ARG, "field %d cannot be set", which);
    }
    return PBPIC_OK;
}

int pbpic_get_field(pbpic_ctx *c, int32_t which, double *plane)
{
    if (!c || !plane) return PBPIC_ERR_ARG;
    int rc = check_flags(c);
    if (rc) return rc;
    const int nx = c->p.nx, ny = c->p.ny;
    const size_t nc = (size_t)ncell(c);
    if (which >= PBPIC_EX && which <= PBPIC_EZ) {
        HIPCHK(c, hipMemcpy(plane, c->E[which - PBPIC_EX], nc * 8, hipMemcpyDeviceToHost));
    } else if (which >= PBPIC_BX && which <= PBPIC_BZ) {
        HIPCHK(c, hipMemcpy(plane, c->B[which - PBPIC_BX], nc * 8, hipMemcpyDeviceToHost));
    } else if (which >= PBPIC_JX && which <= PBPIC_RHO) {
        void *src = c->J[which - PBPIC_JX];
        if (c->rsz == 8) {
            HIPCHK(c, hipMemcpy(plane, src, nc * 8, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < nc; k++) plane[k] = plane[k] * c->p.in0;   // :1066
        } else {
            std::vector<float> vf(nc);
            HIPCHK(c, hipMemcpy(vf.data(), src, nc * 4, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < nc; k++) plane[k] = (double)(vf[k] * (float)c->p.in0);
        }
    } else if (which >= PBPIC_GEX && which <= PBPIC_GBZ) {
        // emfgrid node averages (pic2d.f90:837-845)
        const int comp = (which - PBPIC_GEX) % 3;
        const bool isb = which >= PBPIC_GBX;
        std::vector<double> f(nc);
        HIPCHK(c, hipMemcpy(f.data(), isb ? c->B[comp] : c->E[comp], nc * 8, hipMemcpyDeviceToHost));
        for (int gy = 0; gy < ny; gy++)
            for (int gx = 0; gx < nx; gx++) {
                const int gxm = gx == 0 ? nx - 1 : gx - 1, gym = gy == 0 ? ny - 1 : gy - 1;
                const double a = f[(size_t)gy * nx + gx], bx = f[(size_t)gy * nx + gxm];
                const double by = f[(size_t)gym * nx + gx], bxy = f[(size_t)gym * nx + gxm];
                double v;
                if (comp == 0) v = (a + bx) * 0.5;
                else if (comp == 1) v = (a + by) * 0.5;
                else v = (a + bx + by + bxy) * 0.25;
                plane[(size_t)gy * nx + gx] = v;
            }
    } else {
        return fail(c, PBPIC_ERR_ARG, "unknown field %d", which);
    }
    return PBPIC_OK;
}

int pbpic_zero_current(pbpic_ctx *c)
{
    if (!c) return PBPIC_ERR_ARG;
    HIPCHK(c, launch_zero_current(c->J[0], c->J[1], c->J[2], c->J[3], ncell(c), c->rsz, c->acc, 0, c->st));
    if (c->J64[0])
        HIPCHK(c, launch_zero_current(c->J64[0], c->J64[1], c->J64[2], c->J64[3], ncell(c) * (c->rsz / 4), 8, c->acc,
                                      0, c->st));
    return PBPIC_OK;
}

int pbpic_push(pbpic_ctx *c)
{
    if (!c) return PBPIC_ERR_ARG;
    return push_impl(c);
}

int pbpic_deposit(pbpic_ctx *c, int64_t n, const int32_t *i1, const int32_t *i2, const int32_t *j1,
                  const int32_t *j2, const double *dx1, const double *dx2, const double *dy1,
                  const double *dy2, const int32_t *q, const double *vz)
{
    int rc = deposit_check(c, n, i1, i2, j1, j2, dx1, dx2, dy1, dy2, q, vz);
    if (rc || n <= 0) return rc;
    void *buf = nullptr;
    HIPCHK(c, hipMalloc(&buf, (size_t)n * DEPOSIT_ENTRY_BYTES + 8));
    rc = deposit_staged(c, buf, n, i1, i2, j1, j2, dx1, dx2, dy1, dy2, q, vz);
    (void)hipFree(buf);
    return rc;
}

int pbpic_deposit_noalloc(pbpic_ctx *c, int64_t n, const int32_t *i1, const int32_t *i2,
                          const int32_t *j1, const int32_t *j2, const double *dx1,
                          const double *dx2, const double *dy1, const double *dy2,
                          const int32_t *q, const double *vz)
{
    int rc = deposit_check(c, n, i1, i2, j1, j2, dx1, dx2, dy1, dy2, q, vz);
    if (rc || n <= 0) return rc;
    if (n > PBPIC_DEPOSIT_NOALLOC_MAX)
        return fail(c, PBPIC_ERR_CAPACITY,
                    "pbpic_deposit_noalloc: n=%lld exceeds its fixed scratch of %d entries "
                    "(use pbpic_deposit, or split the batch)",
                    (long long)n, PBPIC_DEPOSIT_NOALLOC_MAX);
    if (!c->scratch) return fail(c, PBPIC_ERR_STATE, "no deposit scratch on device %d", c->dev);
    return deposit_staged(c, c->scratch, n, i1, i2, j1, j2, dx1, dx2, dy1, dy2, q, vz);
}

int pbpic_fieldsolve(pbpic_ctx *c)
{
    if (!c) return PBPIC_ERR_ARG;
    return fieldsolve_impl(c);
}

int pbpic_fieldsenergy(pbpic_ctx *c, double out[2])
{
    if (!c || !out) return PBPIC_ERR_ARG;
    HIPCHK(c, launch_acc_op(c->acc, 2, 0.0, c->st));
    HIPCHK(c, launch_energy(fdtd_params(c), c->st));
    double a[ACC_N];
    HIPCHK(c, hipMemcpyAsync(a, c->acc, sizeof a, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const double cva2 = c->p.cva * c->p.cva;
    out[0] = 0.5 * a[ACC_WE] * c->p.mime / (cva2 * cva2 * c->p.in0);   // :671
    out[1] = 0.5 * a[ACC_WB] * c->p.mime / (cva2 * c->p.in0);          // :672
    return PBPIC_OK;
}

int pbpic_kineticenergy(pbpic_ctx *c, double out[2])
{
    if (!c || !out) return PBPIC_ERR_ARG;
    int rc = kinetic_impl(c);
    if (rc) return rc;
    double a[ACC_N];
    HIPCHK(c, hipMemcpyAsync(a, c->acc, sizeof a, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    out[0] = a[ACC_KE_E];
    out[1] = a[ACC_KE_I];
    return PBPIC_OK;
}

int pbpic_countingsort_species(pbpic_ctx *c, int32_t s, int32_t *counts, int32_t *perm)
{
    int rc = valid_species(c, s);
    if (rc) return rc;
    return sort_one(c, s, counts, perm);
}

int pbpic_countingsort(pbpic_ctx *c)
{
    if (!c) return PBPIC_ERR_ARG;
    for (int s = 0; s < (int)c->sp.size(); s++) {
        int rc = sort_one(c, s, nullptr, nullptr);
        if (rc) return rc;
    }
    return PBPIC_OK;
}

int pbpic_step(pbpic_ctx *c, int32_t nsteps)
{
    if (!c || nsteps < 0) return PBPIC_ERR_ARG;
    for (int32_t k = 0; k < nsteps; k++) {
        int rc = step_impl(c, nullptr, nullptr, nullptr);
        if (rc) return rc;
    }
    return PBPIC_OK;
}

int pbpic_diag(pbpic_ctx *c, double out[PBPIC_NDIAG])
{
    if (!c || !out) return PBPIC_ERR_ARG;
    int rc = check_flags(c);
    if (rc) return rc;
    for (int k = 0; k < PBPIC_NDIAG; k++) out[k] = 0;
    const size_t nc = (size_t)ncell(c);
    if (!c->rho_scratch) {
        HIPCHK(c, dalloc(&c->rho_scratch, nc * 8 * (c->rsz / 4)));   // FP64 form: two words
        HIPCHK(c, dalloc(&c->gauss0, nc * 8));
        c->gauss_first = true;
    }
    HIPCHK(c, launch_acc_op(c->acc, 2, 0.0, c->st));
    HIPCHK(c, launch_energy(fdtd_params(c), c->st));
    if ((rc = kinetic_impl(c))) return rc;
    const double cva2 = c->p.cva * c->p.cva;
    // the charge in fixed point, exact in any order.  Split form: one word,
    // a particle adds at most 1 to a node, so ntot * scale stays below 2^62;
    // FP64 form: two words (launch_gauss)
    double rscale = 0;
    if (c->rsz == 4) {
        int64_t ntot = 1;
        for (auto &S : c->sp) ntot += S.n;
        int e = 61;
        while (e > 0 && (double)ntot > std::ldexp(1.0, 62 - e)) e--;
        rscale = std::ldexp(1.0, std::min(e, 52));
    }
    HIPCHK(c, launch_gauss(push_params(c), c->p.precision, rscale, c->rho_scratch, c->gauss0,
                           c->gauss_first ? 1 : 0, c->p.rho_background, cva2 * c->p.in0, c->st));
    double a[ACC_N];
    HIPCHK(c, hipMemcpyAsync(a, c->acc, sizeof a, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->gauss_first = false;
    const double scale = c->p.mime / cva2;
    out[PBPIC_D_KE_E] = a[ACC_KE_E];
    out[PBPIC_D_KE_I] = a[ACC_KE_I];
    out[PBPIC_D_WE] = 0.5 * a[ACC_WE] * c->p.mime / (cva2 * cva2 * c->p.in0);
    out[PBPIC_D_WB] = 0.5 * a[ACC_WB] * c->p.mime / (cva2 * c->p.in0);
    out[PBPIC_D_JE_STEP] = a[ACC_JE_STEP] * scale;
    out[PBPIC_D_JE_INT] = a[ACC_JE_INT] + a[ACC_JE_STEP] * scale * c->p.dt;
    out[PBPIC_D_GAUSS_L2] = sqrt(a[ACC_G_L2]);
    out[PBPIC_D_GAUSS_LINF] = a[ACC_G_LINF];
    out[PBPIC_D_GAUSS_DL2] = sqrt(a[ACC_G_DL2]);
    out[PBPIC_D_GAUSS_DLINF] = a[ACC_G_DLINF];
    out[PBPIC_D_STEPS] = (double)c->steps;
    out[PBPIC_D_TILES_SPILLED] = a[ACC_SPILL];
    out[PBPIC_D_SORTS_RADIX] = (double)c->sorts_radix;
    return PBPIC_OK;
}

static int diag_counts_buf(pbpic_ctx *c)
{
    if (!c->diag_counts)
        HIPCHK(c, dalloc(&c->diag_counts,
                         sizeof(int32_t) * (size_t)std::max<int64_t>(ncell(c), SPEC_BINS_MAX)));
    return PBPIC_OK;
}

int pbpic_energy_spectrum(pbpic_ctx *c, int32_t s, int32_t nbins, double de, double emax,
                          int32_t *counts)
{
    int rc = valid_species(c, s);
    if (rc) return rc;
    if (!counts || nbins < 1 || nbins > SPEC_BINS_MAX || !(de > 0) || !(emax > 0))
        return fail(c, PBPIC_ERR_ARG, "bad spectrum arguments (1 <= nbins <= %d, de > 0)",
                    SPEC_BINS_MAX);
    if ((rc = check_flags(c)) || (rc = diag_counts_buf(c))) return rc;
    HIPCHK(c, launch_fill_u32((uint32_t *)c->diag_counts, nbins, 0u, c->st));
    HIPCHK(c, launch_spectrum(species_dev(c, s, c->sp[s].cur), c->p.precision, nbins, de, emax,
                              (uint32_t *)c->diag_counts, c->st));
    HIPCHK(c, hipMemcpyAsync(counts, c->diag_counts, sizeof(int32_t) * nbins,
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PBPIC_OK;
}

int pbpic_number_density(pbpic_ctx *c, int32_t *counts)
{
    if (!c || !counts) return PBPIC_ERR_ARG;
    int rc = check_flags(c);
    if (rc || (rc = diag_counts_buf(c))) return rc;
    const int64_t nc = ncell(c);
    HIPCHK(c, launch_fill_u32((uint32_t *)c->diag_counts, nc, 0u, c->st));
    for (int s = 0; s < (int)c->sp.size(); s++)
        HIPCHK(c, launch_cell_counts(species_dev(c, s, c->sp[s].cur), c->p.nx, c->p.ny,
                                     c->diag_counts, c->err, c->st));
    HIPCHK(c, hipMemcpyAsync(counts, c->diag_counts, sizeof(int32_t) * nc, hipMemcpyDeviceToHost,
                             c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return check_flags(c);
}

int pbpic_get_counters(pbpic_ctx *c, int64_t *steps, double je[2])
{
    if (!c || !steps || !je) return PBPIC_ERR_ARG;
    double a[ACC_N];
    HIPCHK(c, hipMemcpyAsync(a, c->acc, sizeof a, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    *steps = c->steps;
    je[0] = a[ACC_JE_STEP];
    je[1] = a[ACC_JE_INT];
    return PBPIC_OK;
}

int pbpic_set_counters(pbpic_ctx *c, int64_t steps, const double je[2])
{
    if (!c || !je || steps < 0) return PBPIC_ERR_ARG;
    double a[ACC_N];
    HIPCHK(c, hipMemcpyAsync(a, c->acc, sizeof a, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    a[ACC_JE_STEP] = je[0];
    a[ACC_JE_INT] = je[1];
    HIPCHK(c, hipMemcpyAsync(c->acc, a, sizeof a, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->steps = steps;
    return PBPIC_OK;
}

int pbpic_species_device_ptrs(pbpic_ctx *c, int32_t s, void **ptrs)
{
    int rc = valid_species(c, s);
    if (rc) return rc;
    if (!ptrs) return PBPIC_ERR_ARG;
    Species &S = c->sp[s];
    const int b = S.cur;
    S.tbox_ok = false;          // the caller may write through the pointers
    ptrs[0] = S.cell[b]; ptrs[1] = S.ox[b]; ptrs[2] = S.oy[b];
    ptrs[3] = S.ux[b]; ptrs[4] = S.uy[b]; ptrs[5] = S.uz[b]; ptrs[6] = S.pid[b];
    return PBPIC_OK;
}

int pbpic_field_device_ptrs(pbpic_ctx *c, void **ptrs)
{
    if (!c || !ptrs) return PBPIC_ERR_ARG;
    for (int k = 0; k < 3; k++) { ptrs[k] = c->E[k]; ptrs[3 + k] = c->B[k]; }
    for (int k = 0; k < 4; k++) ptrs[6 + k] = c->J[k];
    return PBPIC_OK;
}

void *pbpic_stream(pbpic_ctx *c) { return c ? (void *)c->st : nullptr; }

int pbpic_set_particles_device(pbpic_ctx *c, int32_t s, int64_t n, const int32_t *cell,
                               const void *dx, const void *dy, const void *ux, const void *uy,
                               const void *uz)
{
    int rc = valid_species(c, s);
    if (rc) return rc;
    Species &S = c->sp[s];
    if (n < 0 || n > S.cap) return fail(c, PBPIC_ERR_CAPACITY, "n=%lld exceeds capacity", (long long)n);
    const int b = S.cur;
    S.tbox_ok = false;
    HIPCHK(c, hipMemcpyAsync(S.cell[b], cell, 4 * n, hipMemcpyDeviceToDevice, c->st));
    const void *src[5] = {dx, dy, ux, uy, uz};
    void *dst[5] = {S.ox[b], S.oy[b], S.ux[b], S.uy[b], S.uz[b]};
    for (int k = 0; k < 5; k++)
        HIPCHK(c, hipMemcpyAsync(dst[k], src[k], (size_t)c->rsz * n, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, launch_check_cells(S.cell[b], n, ncell(c), c->p.ny, nyp_of(c), c->err, c->st));
    S.n = n;
    return PBPIC_OK;
}

int pbpic_set_field_device(pbpic_ctx *c, int32_t which, const double *plane)
{
    if (!c || !plane) return PBPIC_ERR_ARG;
    const size_t nc = (size_t)ncell(c);
    double *dst;
    if (which >= PBPIC_EX && which <= PBPIC_EZ) dst = c->E[which - PBPIC_EX];
    else if (which >= PBPIC_BX && which <= PBPIC_BZ) dst = c->B[which - PBPIC_BX];
    else return fail(c, PBPIC_ERR_ARG, "field %d cannot be set from the device", which);
    HIPCHK(c, hipMemcpyAsync(dst, plane, nc * 8, hipMemcpyDeviceToDevice, c->st));
    if (walls(c)) HIPCHK(c, launch_wall_images(fdtd_params(c), 1 << which, c->st));
    return PBPIC_OK;
}

int pbpic_adjoint_forward(pbpic_ctx *c, int32_t nsteps, int32_t ckpt, double *loss)
{
    if (!c || nsteps < 1 || ckpt < 1) return PBPIC_ERR_ARG;
    if (c->p.flags & (PBPIC_FLAG_MARDER | PBPIC_FLAG_CHERENKOV))
        return fail(c, PBPIC_ERR_STATE, "the adjoint does not support the Marder/Cherenkov filters");
    int rc = check_flags(c);
    if (rc) return rc;
    adj_free(c);
    const StateLayout L = state_layout(c);
    const int nck = (nsteps + ckpt - 1) / ckpt;
    for (int k = 0; k < nck; k++) {
        char *p = nullptr;
        hipError_t e = hipMalloc(&p, L.total);
        if (e != hipSuccess) { adj_free(c); return fail(c, PBPIC_ERR_NOMEM, "checkpoint allocation: %s", hipGetErrorString(e)); }
        c->adj.ckpts.push_back(p);
    }
    HIPCHK(c, hipMalloc(&c->adj.final_state, L.total));
    // J of every step (3 planes each), for a replay without deposit; the
    // replay deposits again when this does not fit
    const size_t jplane = al256((size_t)c->rsz * ncell(c));
    if (hipMalloc(&c->adj.jstore, (size_t)nsteps * 3 * jplane) != hipSuccess) {
        c->adj.jstore = nullptr;
        (void)hipGetLastError();
    }
    double a0[ACC_N], a1[ACC_N];
    HIPCHK(c, hipMemcpyAsync(a0, c->acc, sizeof a0, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->adj.step0 = c->steps;
    for (int k = 0; k < nsteps; k++) {
        if (k % ckpt == 0 && (rc = state_copy(c, L, c->adj.ckpts[k / ckpt], true))) return rc;
        void *jk[3];
        for (int q = 0; q < 3 && c->adj.jstore; q++)
            jk[q] = c->adj.jstore + ((size_t)k * 3 + q) * jplane;
        if ((rc = step_impl(c, c->adj.jstore ? jk : nullptr, nullptr, nullptr))) return rc;
    }
    if ((rc = state_copy(c, L, c->adj.final_state, true))) return rc;
    HIPCHK(c, hipMemcpyAsync(a1, c->acc, sizeof a1, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if ((rc = check_flags(c))) return rc;
    memcpy(c->adj.acc_after, a1, sizeof a1);
    c->adj.nsteps = nsteps;
    c->adj.ckpt = ckpt;
    c->adj.loss = loss_total(c, a1) - loss_total(c, a0);
    c->adj.valid = true;
    if (loss) *loss = c->adj.loss;
    // the backward's buffers, allocated here so that the backward is the
    // sweep alone (an allocation of this size can take seconds while the
    // driver clears memory another process freed); if they do not fit now,
    // the backward tries again and reports the failure
    (void)adj_tape_alloc(c, ckpt);
    return PBPIC_OK;
}

int pbpic_adjoint_backward(pbpic_ctx *c, double *const *fg, void *const *pg)
{
    if (!c || !fg || !pg) return PBPIC_ERR_ARG;
    if (!c->adj.valid) return fail(c, PBPIC_ERR_STATE, "pbpic_adjoint_forward has not been run");
    const StateLayout L = state_layout(c);
    const int K = c->adj.ckpt, N = c->adj.nsteps, ns = (int)c->sp.size();
    const size_t nc = (size_t)ncell(c), rz = (size_t)c->rsz;
    int64_t nmax = 0, atiles = 0;
    for (auto &S : c->sp) { nmax = std::max(nmax, S.n); atiles += adj_tiles(S.n); }
    // one allocation for the tape and the adjoint buffers (normally made at
    // the end of the forward); J of each replayed step is the forward's kept
    // J, else recomputed into the tape
    hipError_t e = adj_tape_alloc(c, K);
    if (e != hipSuccess) return fail(c, PBPIC_ERR_NOMEM, "adjoint tape allocation: %s", hipGetErrorString(e));
    char *base = c->adj.tape;
    c->adj.tape = nullptr;                     // freed at the end of this call
    const int nperm = adj_tape_sizes(c, K).nperm;
    const size_t jplane = al256(rz * nc);
    int rc = PBPIC_OK;
    char *p = base;
    auto take = [&](size_t b) { char *q = p; p += al256(b); return q; };
    std::vector<char *> tstate(K + 1);
    for (int k = 0; k <= K; k++) tstate[k] = take(L.total);
    std::vector<std::array<void *, 3>> tj(K);
    for (int k = 0; k < K && !c->adj.jstore; k++)
        for (int q = 0; q < 3; q++) tj[k][q] = take(rz * nc);
    std::vector<std::vector<uint32_t *>> tperm(nperm, std::vector<uint32_t *>(ns));
    for (int k = 0; k < nperm; k++) for (int s = 0; s < ns; s++) tperm[k][s] = (uint32_t *)take(4 * (size_t)c->sp[s].n);
    std::vector<int> pslot(K, -1);
    double *fa[18];
    for (int q = 0; q < 18; q++) fa[q] = (double *)take(8 * nc);
    std::vector<std::array<void *, 5>> pa(ns), tmp(1);
    for (int s = 0; s < ns; s++) for (int q = 0; q < 5; q++) pa[s][q] = take(rz * (size_t)c->sp[s].n);
    for (int q = 0; q < 5; q++) tmp[0][q] = take(rz * (size_t)nmax);
    int32_t *adj_big = (int32_t *)take(sizeof(int32_t) * (size_t)(atiles + 1));
    std::vector<bool> sorted(K);
    auto run = [&]() -> int {
        HIPCHK(c, hipMemsetAsync(fa[0], 0, (size_t)(p - (char *)fa[0]), c->st));
        HIPCHK(c, prof_mark(c, PBPIC_PROF_ADJ, true));
        const int nseg = (N + K - 1) / K;
        for (int seg = nseg - 1; seg >= 0; seg--) {
            const int s0 = seg * K, kk = std::min(N, s0 + K) - s0;
            int r = state_copy(c, L, c->adj.ckpts[seg], false);
            if (r) return r;
            c->steps = c->adj.step0 + s0;
            if ((r = state_copy(c, L, tstate[0], true))) return r;
            int used = 0;
            TapeGuard guard(c);
            for (int k = 0; k < kk; k++) {
                guard.step(L, tstate[k], tstate[k + 1]);
                bool so = false;
                const bool will = c->p.sort_every > 0 && (c->steps + 1) % c->p.sort_every == 0;
                uint32_t *const *pd = nullptr;
                pslot[k] = -1;
                if (will) {
                    if (used >= nperm) return fail(c, PBPIC_ERR_STATE, "adjoint permutation tape overflow");
                    pslot[k] = used;
                    pd = tperm[used++].data();
                }
                if (c->adj.jstore) {
                    for (int q = 0; q < 3; q++)
                        tj[k][q] = c->adj.jstore + ((size_t)(s0 + k) * 3 + q) * jplane;
                    if ((r = step_impl(c, nullptr, pd, &so, tj[k].data()))) return r;
                } else if ((r = step_impl(c, tj[k].data(), pd, &so))) {
                    return r;
                }
                sorted[k] = so;
                // the push wrote slot k+1; a sort moved the particles into the
                // context's own arrays, and the fields are always the context's
                if ((r = state_copy(c, L, tstate[k + 1], true, so ? STATE_ALL : STATE_FIELDS)))
                    return r;
            }
            guard.restore();
            for (int k = kk - 1; k >= 0; k--) {
                if (sorted[k]) {
                    const int po = (c->p.flags & PBPIC_FLAG_PUSH_ORDER) != 0;
                    for (int s = 0; s < ns; s++) {
                        HIPCHK(c, launch_unpermute(tperm[pslot[k]][s], c->sp[s].n, c->p.precision,
                                                   po, pa[s].data(), tmp[0].data(), c->st));
                        // the unpermuted arrays become the species' adjoint
                        // arrays; a spare of the largest species' size stays
                        // for the next species
                        if (c->sp[s].n == nmax) {
                            std::swap(pa[s], tmp[0]);
                        } else {
                            CopyList cl{};
                            for (int q = 0; q < 5; q++) cl.add(pa[s][q], tmp[0][q], rz * c->sp[s].n);
                            HIPCHK(c, launch_copy_list(cl, c->st));
                        }
                    }
                }
                AdjFieldParams F{};
                F.nx = c->p.nx; F.ny = c->p.ny;
                F.nyp = walls(c) ? nyp_of(c) : 0;
                F.dth = c->dth; F.dtc2 = c->dtc2; F.idx = c->idx; F.idy = c->idy;
                F.in0 = c->p.in0; F.w = c->p.dt * c->p.mime / (c->p.cva * c->p.cva);
                F.exa = fa[0]; F.eya = fa[1]; F.eza = fa[2];
                F.bxa = fa[3]; F.bya = fa[4]; F.bza = fa[5];
                F.etx = fa[6]; F.ety = fa[7]; F.etz = fa[8];
                F.jxa = fa[9]; F.jya = fa[10]; F.jza = fa[11];
                F.e0x = state_field(c, L, tstate[k], 0); F.e0y = state_field(c, L, tstate[k], 1);
                F.e0z = state_field(c, L, tstate[k], 2);
                F.e1x = state_field(c, L, tstate[k + 1], 0); F.e1y = state_field(c, L, tstate[k + 1], 1);
                F.e1z = state_field(c, L, tstate[k + 1], 2);
                F.jx = tj[k][0]; F.jy = tj[k][1]; F.jz = tj[k][2];
                F.lmask = c->lmask;
                HIPCHK(c, launch_fdtd_adj(F, c->p.precision, c->st));
                AdjParams A{};
                A.nx = c->p.nx; A.ny = c->p.ny; A.compat = c->p.compat;
                A.nyp = walls(c) ? nyp_of(c) : 0;
                A.tv = (c->p.flags & PBPIC_FLAG_PUSH_ORDER) ? push_vec(c->p.precision) : 0;
                A.dny.init((uint32_t)c->p.ny);
                // windows of two workgroups per CU: node fields and node
                // adjoints (72 or 96 B of LDS per node), the current adjoint
                // in the rest of the 80 KiB
                A.maxw = c->p.precision == PBPIC_PREC_FP64 ? 768 : 1024;
                A.jmax = adj_jmax_for(c->p.precision, A.maxw, 80 * 1024);
                A.nspecies = ns;
                A.dtcva = c->dtcva; A.idx = c->idx; A.idy = c->idy; A.cva = c->p.cva;
                A.ex = state_field(c, L, tstate[k], 0); A.ey = state_field(c, L, tstate[k], 1);
                A.ez = state_field(c, L, tstate[k], 2); A.bx = state_field(c, L, tstate[k], 3);
                A.by = state_field(c, L, tstate[k], 4); A.bz = state_field(c, L, tstate[k], 5);
                A.jadj[0] = fa[9]; A.jadj[1] = fa[10]; A.jadj[2] = fa[11];
                A.nadj = fa[12];
                {
                    // node fields of step k for the adjoint push's windows
                    PushParams Pn{};
                    Pn.nx = c->p.nx; Pn.ny = c->p.ny;
                    Pn.ex = A.ex; Pn.ey = A.ey; Pn.ez = A.ez;
                    Pn.bx = A.bx; Pn.by = A.by; Pn.bz = A.bz;
                    Pn.nf = c->NF;
                    Pn.big = nullptr;
                    HIPCHK(c, launch_nodefield(Pn, c->p.precision, c->st));
                    A.nf = c->NF;
                }
                A.big = adj_big;
                for (int s = 0; s < ns; s++) {
                    Species &S = c->sp[s];
                    AdjSpecies &T = A.sp[s];
                    T.cell = (const int32_t *)state_ptr(c, L, tstate[k], s, 0);
                    T.ox = state_ptr(c, L, tstate[k], s, 1); T.oy = state_ptr(c, L, tstate[k], s, 2);
                    T.ux = state_ptr(c, L, tstate[k], s, 3); T.uy = state_ptr(c, L, tstate[k], s, 4);
                    T.uz = state_ptr(c, L, tstate[k], s, 5);
                    T.n = S.n; T.q = S.q; T.ms = S.ms; T.c0 = S.c0; T.c1 = S.c1; T.qx = S.qx; T.qy = S.qy;
                    A.xa[s] = pa[s][0]; A.ya[s] = pa[s][1];
                    A.uxa[s] = pa[s][2]; A.uya[s] = pa[s][3]; A.uza[s] = pa[s][4];
                }
                HIPCHK(c, launch_push_adj(A, c->p.precision, c->st));
                HIPCHK(c, launch_node_transpose(F, fa[12], c->st));
            }
        }
        if (walls(c)) {
            // gradients with respect to the domain values: the guard rows of
            // the initial fields are images of them
            AdjFieldParams F{};
            F.nx = c->p.nx; F.ny = c->p.ny; F.nyp = nyp_of(c);
            F.exa = fa[0]; F.eya = fa[1]; F.eza = fa[2];
            F.bxa = fa[3]; F.bya = fa[4]; F.bza = fa[5];
            HIPCHK(c, launch_wall_images_adj(F, c->st));
        }
        HIPCHK(c, prof_mark(c, PBPIC_PROF_ADJ, false));
        for (int q = 0; q < 6; q++)
            HIPCHK(c, hipMemcpyAsync(fg[q], fa[q], 8 * nc, hipMemcpyDeviceToDevice, c->st));
        for (int s = 0; s < ns; s++)
            for (int q = 0; q < 5; q++)
                HIPCHK(c, hipMemcpyAsync(pg[5 * s + q], pa[s][q], rz * c->sp[s].n, hipMemcpyDeviceToDevice, c->st));
        int r = state_copy(c, L, c->adj.final_state, false);
        if (r) return r;
        c->steps = c->adj.step0 + N;
        HIPCHK(c, hipMemcpyAsync(c->acc, c->adj.acc_after, sizeof c->adj.acc_after, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        return check_flags(c);
    };
    rc = run();
    (void)hipStreamSynchronize(c->st);
    (void)hipFree(base);
    return rc;
}

int pbpic_set_loss_region(pbpic_ctx *c, const double *weights)
{
    if (!c) return PBPIC_ERR_ARG;
    adj_free(c);
    if (!weights) {
        if (c->lmask) { HIPCHK(c, hipStreamSynchronize(c->st)); (void)hipFree(c->lmask); }
        c->lmask = nullptr;
        return PBPIC_OK;
    }
    const size_t nc = (size_t)ncell(c);
    if (!c->lmask) HIPCHK(c, dalloc(&c->lmask, nc * 8));
    HIPCHK(c, hipMemcpy(c->lmask, weights, nc * 8, hipMemcpyHostToDevice));
    return PBPIC_OK;
}

int pbpic_profile(pbpic_ctx *c, int32_t enable)
{
    if (!c) return PBPIC_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->prof = enable != 0;
    if (c->prof) {
        c->ev_used = 0;
        c->ev_marks.clear();
    }
    return PBPIC_OK;
}

int pbpic_profile_read(pbpic_ctx *c, double ms[PBPIC_NPROF], int64_t count[PBPIC_NPROF])
{
    if (!c || !ms || !count) return PBPIC_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int k = 0; k < PBPIC_NPROF; k++) { ms[k] = 0; count[k] = 0; }
    for (auto &m : c->ev_marks) {
        float t = 0;
        HIPCHK(c, hipEventElapsedTime(&t, c->ev_pool[m.second], c->ev_pool[m.second + 1]));
        ms[m.first] += t;
        count[m.first] += 1;
    }
    return PBPIC_OK;
}

}  // extern "C"
