// k_sort.hip -- countingsort (pic2d.f90:1303-1329) for gfx950.
//
// The reference is a stable counting sort of the particles by cell id:
// histogram (:1310-1313), exclusive scan (:1315-1320), stable scatter
// (:1322-1327), copy back (:1328).  With millions of cells a per-block x
// per-cell count matrix does not fit.  Right after a push (the sort of the
// time loop, :180-182) the ranks come from the push's tile boxes (the
// tile-box path below); otherwise the stable scatter is done as an LSD radix
// sort of (cell, index) pairs over 8-bit digits, each digit pass itself a
// counting sort (a per-wave ballot histogram with in-order ranks, then a
// decoupled look-back prefix scan across tiles).  Either way a final gather
// moves the particle payload once.  The result (permutation and per-cell
// counts) is identical to the reference's: both are the unique stable sort
// by cell.
#include <cstdlib>

#include "pbpic_impl.h"

namespace pbpic {

constexpr int ST_THREADS = 256;
constexpr int ST_ROUNDS = 16;                    // 16 x 64 items per wave
constexpr int ST_TILE = ST_THREADS * ST_ROUNDS;  // 4096
constexpr uint32_t FLAG_AGG = 1u << 30, FLAG_INC = 2u << 30, CNT_MASK = (1u << 30) - 1;
constexpr int SPIN_LIMIT = 1 << 24;

// Per-cell counts with run-length aggregation inside a wave: consecutive
// equal cell ids (the common case for an almost sorted array) cost one atomic.
__global__ __launch_bounds__(256) void cell_hist_kernel(const int32_t *cell, int64_t n,
                                                        int64_t ncell, CellKey key,
                                                        int32_t *counts, int *err)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool valid = i < n;
    if (valid && (cell[i] < 0 || cell[i] >= ncell)) {
        atomicOr(err, ERRF_CELL);
        valid = false;
    }
    const int k = valid ? (int)key((uint32_t)cell[i]) : -1;
    const int prev = __shfl_up(k, 1);
    // a run ends at the next lane whose key differs (invalid lanes carry -1)
    const bool boundary = lane == 0 || k != prev;
    const uint64_t heads = __ballot(boundary);
    if (valid && boundary) {
        const uint64_t above = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1);
        const int end = above ? __ffsll((unsigned long long)above) - 1 : 64;
        atomicAdd(counts + k, end - lane);
    }
}

// Per-key counts, one push tile (PUSH_TILE particles) per workgroup, using
// the tile's bounding box that the last push recorded (SpeciesDev::tbox):
// counts go to an LDS histogram over the box and are flushed with one global
// atomic per occupied cell.  A tile with no recorded box, a box larger than
// the LDS histogram, or a cell outside its box counts with global atomics.
constexpr int HB_BINS = 2048;
__global__ __launch_bounds__(256) void cell_hist_tiles_kernel(const int32_t *cell, int64_t n,
                                                              int nx, int64_t ncell,
                                                              CellKey key, const int32_t *tbox,
                                                              int32_t *counts, int *err)
{
    __shared__ uint32_t h[HB_BINS];
    const int64_t t0 = (int64_t)blockIdx.x * PUSH_TILE;
    const int4 b = *reinterpret_cast<const int4 *>(tbox + 4 * (size_t)blockIdx.x);
    const int W = b.y + 1, H = b.w + 1;
    const bool boxed = b.y >= 0 && b.w >= 0 && (int64_t)W * H <= HB_BINS;
    const int ny = (int)key.ny;
    if (boxed)
        for (int k = threadIdx.x; k < W * H; k += 256) h[k] = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < PUSH_TILE; q += 256) {
        const int64_t i = t0 + q;
        if (i >= n) break;
        const int c = cell[i];
        if (c < 0 || c >= ncell) { atomicOr(err, ERRF_CELL); continue; }
        if (boxed) {
            const int cx = (int)key.dny.div31((uint32_t)c), cy = c - cx * ny;
            int rx = cx - b.x, ry = cy - b.z;
            if (rx < 0) rx += nx;
            if (ry < 0) ry += ny;
            if (rx < W && ry < H) {
                atomicAdd(&h[ry * W + rx], 1u);
                continue;
            }
        }
        atomicAdd(counts + key((uint32_t)c), 1);
    }
    __syncthreads();
    if (boxed)
        for (int k = threadIdx.x; k < W * H; k += 256) {
            const uint32_t v = h[k];
            if (!v) continue;
            const int ky = k / W, kx = k - ky * W;
            int gx = b.x + kx, gy = b.z + ky;
            if (gx >= nx) gx -= nx;
            if (gy >= ny) gy -= ny;
            atomicAdd(counts + key((uint32_t)(gx * ny + gy)), (int)v);
        }
}

// Digit histograms of all passes from the per-cell counts.
__global__ __launch_bounds__(256) void digit_hist_kernel(const int32_t *counts, int64_t ncell,
                                                         int npass, uint32_t *dhist)
{
    __shared__ uint32_t h[4][256];
    for (int k = threadIdx.x; k < 4 * 256; k += 256) h[k >> 8][k & 255] = 0;
    __syncthreads();
    const int64_t c0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 64;
    const int64_t c1 = c0 + 64 < ncell ? c0 + 64 : ncell;
    for (int64_t c = c0; c < c1; c++) {
        const int32_t v = counts[c];
        if (v) atomicAdd(&h[0][c & 255], (uint32_t)v);
    }
    for (int p = 1; p < npass; p++) {
        uint32_t cur = 0xFFFFFFFFu, acc = 0;
        for (int64_t c = c0; c < c1; c++) {
            const uint32_t d = (uint32_t)(c >> (8 * p)) & 255u;
            if (d != cur) {
                if (acc) atomicAdd(&h[p][cur], acc);
                cur = d;
                acc = 0;
            }
            acc += (uint32_t)counts[c];
        }
        if (acc) atomicAdd(&h[p][cur], acc);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < npass * 256; k += 256)
        if (h[k >> 8][k & 255]) atomicAdd(dhist + k, h[k >> 8][k & 255]);
}

__global__ void digit_scan_kernel(uint32_t *dhist, int npass)
{
    if (threadIdx.x >= npass) return;
    uint32_t *h = dhist + threadIdx.x * 256;
    uint32_t run = 0;
    for (int d = 0; d < 256; d++) {
        const uint32_t v = h[d];
        h[d] = run;
        run += v;
    }
}

template <bool FIRST, bool LAST>
__global__ __launch_bounds__(ST_THREADS) void radix_pass_kernel(
    const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout, int64_t n,
    int shift, const uint32_t *dbase, uint32_t *status, uint32_t *ticket, int *err, CellKey ck)
{
    __shared__ uint32_t whist[4][256];
    __shared__ uint32_t woff[4][256];
    __shared__ uint32_t gbase[256];
    __shared__ int tile_s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < 4 * 256; k += ST_THREADS) whist[k >> 8][k & 255] = 0;
    if (tid == 0) tile_s = (int)atomicAdd(ticket, 1u);
    __syncthreads();
    const int tile = tile_s;
    const int64_t wbase = (int64_t)tile * ST_TILE + (int64_t)wave * (ST_ROUNDS * 64);
    const uint64_t lt = (1ull << lane) - 1;
    uint32_t key[ST_ROUNDS], val[ST_ROUNDS], rank[ST_ROUNDS], dig[ST_ROUNDS];
#pragma unroll
    for (int r = 0; r < ST_ROUNDS; r++) {
        const int64_t i = wbase + r * 64 + lane;
        const bool valid = i < n;
        // the first pass reads cell ids and sorts by their key
        const uint32_t k = valid ? (FIRST ? ck(kin[i]) : kin[i]) : 0u;
        const uint32_t v = FIRST ? (uint32_t)i : (valid ? vin[i] : 0u);
        const uint32_t d = (k >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t base = whist[wave][d];
        rank[r] = base + (uint32_t)__popcll(m & lt);
        if (valid && (m & lt) == 0) whist[wave][d] = base + (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
        key[r] = k;
        val[r] = v;
        dig[r] = valid ? d : 0xFFFFFFFFu;
    }
    __syncthreads();
    {
        const int d = tid;
        const uint32_t c0 = whist[0][d], c1 = whist[1][d], c2 = whist[2][d], c3 = whist[3][d];
        woff[0][d] = 0;
        woff[1][d] = c0;
        woff[2][d] = c0 + c1;
        woff[3][d] = c0 + c1 + c2;
        const uint32_t total = c0 + c1 + c2 + c3;
        uint32_t excl = 0;
        uint32_t *mine = status + (size_t)tile * 256 + d;
        if (tile == 0) {
            __hip_atomic_store(mine, FLAG_INC | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(mine, FLAG_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int j = tile - 1, spins = 0;
            while (true) {
                const uint32_t s = __hip_atomic_load(status + (size_t)j * 256 + d, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                if (s == 0) {
                    if (++spins > SPIN_LIMIT) { atomicOr(err, ERRF_SPIN); break; }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += s & CNT_MASK;
                if ((s & FLAG_INC) || j == 0) break;
                j--;
            }
            __hip_atomic_store(mine, FLAG_INC | (excl + total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        gbase[d] = dbase[d] + excl;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ST_ROUNDS; r++) {
        const uint32_t d = dig[r];
        if (d == 0xFFFFFFFFu) continue;
        const uint32_t pos = gbase[d] + woff[wave][d] + rank[r];
        if ((int64_t)pos >= n) { atomicOr(err, ERRF_SPIN); continue; }
        if (!LAST) kout[pos] = key[r];
        vout[pos] = val[r];
    }
}

// dst[q] = src[perm[rank(q)]]; rank(q) is q in cell order, or the push order
// (tile_rank) inside full tiles when tv = V > 0.  perm_out[q] receives the
// source index of slot q.  One lane moves G = 16 bytes / sizeof(R) consecutive
// slots, whose ranks are consecutive in both orders (G equals the push
// kernel's vector width); when their sources are G consecutive, aligned
// slots -- particles that kept their order since the last sort -- each array
// moves with one vector load, otherwise element by element.
template <typename T, int G> struct Vec;
template <> struct Vec<float, 4> { using t = float4; };
template <> struct Vec<double, 2> { using t = double2; };
template <> struct Vec<int32_t, 4> { using t = int4; };
template <> struct Vec<int32_t, 2> { using t = int2; };
template <> struct Vec<uint32_t, 4> { using t = uint4; };
template <> struct Vec<uint32_t, 2> { using t = uint2; };

template <int G, typename T>
__device__ __forceinline__ void gmove(T *dst, const T *src, const uint32_t (&p)[G], bool contig)
{
    using VT = typename Vec<T, G>::t;
    if (contig) {
        *reinterpret_cast<VT *>(dst) = *reinterpret_cast<const VT *>(src + p[0]);
    } else {
        T v[G];
#pragma unroll
        for (int j = 0; j < G; j++) v[j] = src[p[j]];
        *reinterpret_cast<VT *>(dst) = *reinterpret_cast<const VT *>(v);
    }
}

template <typename R>
__global__ __launch_bounds__(256) void gather_kernel(const uint32_t *perm, int64_t n, int tv,
                                                     uint32_t *perm_out, SpeciesDev src,
                                                     SpeciesDev dst)
{
    constexpr int G = 16 / (int)sizeof(R);
    // XCD-contiguous destination ranges: the sources of neighbouring
    // destinations overlap, and stay in one L2
    const int64_t i0 = ((int64_t)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) * G;
    if (i0 >= n) return;
    if (i0 + G > n) {                        // ragged end: element by element
        for (int64_t i = i0; i < n; i++) {
            const uint32_t p = perm[i];
            if (perm_out) perm_out[i] = p;
            dst.cell[i] = src.cell[p];
            ((R *)dst.ox)[i] = ((const R *)src.ox)[p];
            ((R *)dst.oy)[i] = ((const R *)src.oy)[p];
            ((R *)dst.ux)[i] = ((const R *)src.ux)[p];
            ((R *)dst.uy)[i] = ((const R *)src.uy)[p];
            ((R *)dst.uz)[i] = ((const R *)src.uz)[p];
            if (src.pid) dst.pid[i] = src.pid[p];
        }
        return;
    }
    int64_t k0 = i0;
    if (tv) {
        const int64_t t0 = i0 & ~(int64_t)(PUSH_TILE - 1);
        if (t0 + PUSH_TILE <= n) k0 = t0 + tile_rank((int)(i0 - t0), tv);
    }
    uint32_t p[G];
    {
        using VU = typename Vec<uint32_t, G>::t;
        *reinterpret_cast<VU *>(p) = *reinterpret_cast<const VU *>(perm + k0);
    }
    bool contig = p[0] % G == 0;
#pragma unroll
    for (int j = 1; j < G; j++) contig = contig && p[j] == p[0] + j;
    if (perm_out) {
        using VU = typename Vec<uint32_t, G>::t;
        *reinterpret_cast<VU *>(perm_out + i0) = *reinterpret_cast<const VU *>(p);
    }
    gmove<G>(dst.cell + i0, src.cell, p, contig);
    gmove<G>((R *)dst.ox + i0, (const R *)src.ox, p, contig);
    gmove<G>((R *)dst.oy + i0, (const R *)src.oy, p, contig);
    gmove<G>((R *)dst.ux + i0, (const R *)src.ux, p, contig);
    gmove<G>((R *)dst.uy + i0, (const R *)src.uy, p, contig);
    gmove<G>((R *)dst.uz + i0, (const R *)src.uz, p, contig);
    if (src.pid) gmove<G>(dst.pid + i0, src.pid, p, contig);
}

// Split form: the gather reads each source particle as one 24-byte record
// (cell and the five FP32 values) instead of six scattered 4-byte reads from
// six arrays.  The records come from a coalesced packing pass, or straight
// from the push on a step whose sort follows.  The FP64 form keeps the
// array-wise gather (its 48-byte records were slower).
__global__ __launch_bounds__(256) void pack_records_kernel(SpeciesDev src, PartRec *rec, int64_t n)
{
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (i >= n) return;
    PartRec r;
    r.c = src.cell[i];
    r.a[0] = ((const float *)src.ox)[i];
    r.a[1] = ((const float *)src.oy)[i];
    r.a[2] = ((const float *)src.ux)[i];
    r.a[3] = ((const float *)src.uy)[i];
    r.a[4] = ((const float *)src.uz)[i];
    rec[i] = r;
}

// As gather_kernel<float>, reading the packed records; two 4-slot groups per
// thread (slots i0 and i0 + 256 G), all eight record loads in flight.
__device__ __forceinline__ void gather_rec_one(const uint32_t *perm, int64_t n, int tv,
                                               uint32_t *perm_out, const PartRec *rec,
                                               const SpeciesDev &src, const SpeciesDev &dst,
                                               int64_t i0)
{
    float *out[5] = {(float *)dst.ox, (float *)dst.oy, (float *)dst.ux, (float *)dst.uy,
                     (float *)dst.uz};
    for (int64_t i = i0; i < n && i < i0 + 4; i++) {
        int64_t k = i;
        if (tv) {
            const int64_t t0 = i & ~(int64_t)(PUSH_TILE - 1);
            if (t0 + PUSH_TILE <= n) k = t0 + tile_rank((int)(i - t0), tv);
        }
        const uint32_t p = perm[k];
        if (perm_out) perm_out[i] = p;
        const PartRec r = rec[p];
        dst.cell[i] = r.c;
        for (int q = 0; q < 5; q++) out[q][i] = r.a[q];
        if (src.pid) dst.pid[i] = src.pid[p];
    }
}

__global__ __launch_bounds__(256) void gather_records_kernel(const uint32_t *perm, int64_t n,
                                                             int tv, uint32_t *perm_out,
                                                             const PartRec *rec, SpeciesDev src,
                                                             SpeciesDev dst)
{
    constexpr int G = 4, H = 2;
    const int64_t b0 = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x * H * G;
    int64_t i0[H];
#pragma unroll
    for (int h = 0; h < H; h++) i0[h] = b0 + ((int64_t)h * blockDim.x + threadIdx.x) * G;
    if (i0[H - 1] + G > n) {
#pragma unroll
        for (int h = 0; h < H; h++)
            if (i0[h] < n) gather_rec_one(perm, n, tv, perm_out, rec, src, dst, i0[h]);
        return;
    }
    float *out[5] = {(float *)dst.ox, (float *)dst.oy, (float *)dst.ux, (float *)dst.uy,
                     (float *)dst.uz};
    uint32_t p[H][G];
#pragma unroll
    for (int h = 0; h < H; h++) {
        int64_t k0 = i0[h];
        if (tv) {
            const int64_t t0 = i0[h] & ~(int64_t)(PUSH_TILE - 1);
            if (t0 + PUSH_TILE <= n) k0 = t0 + tile_rank((int)(i0[h] - t0), tv);
        }
        *reinterpret_cast<uint4 *>(p[h]) = *reinterpret_cast<const uint4 *>(perm + k0);
    }
    int c[H][G];
    float a[H][5][G];
#pragma unroll
    for (int h = 0; h < H; h++)
#pragma unroll
        for (int j = 0; j < G; j++) {
            const PartRec r = rec[p[h][j]];
            c[h][j] = r.c;
#pragma unroll
            for (int q = 0; q < 5; q++) a[h][q][j] = r.a[q];
        }
#pragma unroll
    for (int h = 0; h < H; h++) {
        if (perm_out)
            *reinterpret_cast<uint4 *>(perm_out + i0[h]) = *reinterpret_cast<const uint4 *>(p[h]);
        *reinterpret_cast<int4 *>(dst.cell + i0[h]) = make_int4(c[h][0], c[h][1], c[h][2], c[h][3]);
#pragma unroll
        for (int q = 0; q < 5; q++)
            *reinterpret_cast<float4 *>(out[q] + i0[h]) =
                make_float4(a[h][q][0], a[h][q][1], a[h][q][2], a[h][q][3]);
        if (src.pid) gmove<G>(dst.pid + i0[h], src.pid, p[h], false);
    }
}

// ---------------------------------------------------------------------
// Tile-box path.  Right after a push every push tile (PUSH_TILE consecutive
// particles) has the bounding box of its cells recorded (SpeciesDev::tbox),
// a few hundred cells.  The stable rank of particle i is
//   (particles of smaller keys) + (particles of key(i) in earlier tiles)
//   + (particles of key(i) before i in its own tile).
// The first two terms depend only on (tile, box cell): one record per box
// cell of every tile, holding that cell's count in the tile.  Sorting the
// records by key, stably (records are laid out tile by tile), and taking an
// exclusive scan of their counts in that order gives both terms at once.
// Each tile then orders its own particles by box cell (a stable in-tile
// counting sort), one wave per 64 records expands the records into the
// permutation (source of every rank), and the gather moves the payload as
// after the radix path.  The records number sum(box areas), about 6% of the
// particles at 100 per cell, so the three radix passes run over them
// instead of over the particles.
constexpr int BX_MAX = 4096;             // largest box (cells) a tile may have

enum : int { BXF_FALLBACK = 1, BXF_OUTSIDE = 2 };

__global__ __launch_bounds__(256) void box_area_kernel(const int32_t *tbox, int64_t ntiles,
                                                       uint32_t *area, int *ctl)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= ntiles) return;
    const int4 b = *reinterpret_cast<const int4 *>(tbox + 4 * t);
    uint32_t a = 0;
    if (b.y < 0 || b.w < 0 || (int64_t)(b.y + 1) * (b.w + 1) > BX_MAX) {
        atomicOr(ctl, BXF_FALLBACK);
    } else {
        a = (uint32_t)((b.y + 1) * (b.w + 1));
        atomicMax(ctl + 1, (int)a);
    }
    area[t] = a;
}

// Exclusive scan of n u32 values (load(i)) into store(i, v), three kernels:
// 4096 values per block are reduced, the block sums scanned by one block
// (the total goes to *total), then each block scans its values again.
constexpr int SC_T = 256, SC_E = 16, SC_B = SC_T * SC_E;

struct LdArr {
    const uint32_t *a;
    __device__ uint32_t operator()(int64_t i) const { return a[i]; }
};
struct LdGather {           // a[idx[i]]
    const uint32_t *a, *idx;
    __device__ uint32_t operator()(int64_t i) const { return a[idx[i]]; }
};
struct StArr {
    uint32_t *o;
    __device__ void operator()(int64_t i, uint32_t v) const { o[i] = v; }
};

template <class L>
__global__ __launch_bounds__(SC_T) void scan_reduce_kernel(L ld, int64_t n, uint32_t *bsum)
{
    __shared__ uint32_t w[SC_T / 64];
    const int64_t b0 = (int64_t)blockIdx.x * SC_B;
    uint32_t acc = 0;
    for (int e = 0; e < SC_E; e++) {
        const int64_t i = b0 + (int64_t)e * SC_T + threadIdx.x;
        if (i < n) acc += ld(i);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
}

// one block of 1024 threads: exclusive scan of bsum[0, nb) in place
__global__ __launch_bounds__(1024) void scan_top_kernel(uint32_t *bsum, int64_t nb, int *total)
{
    __shared__ uint32_t part[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t a = threadIdx.x * per, e = a + per < nb ? a + per : nb;
    uint32_t sum = 0;
    for (int64_t i = a; i < e; i++) sum += bsum[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int k = 0; k < 1024; k++) {
            const uint32_t v = part[k];
            part[k] = run;
            run += v;
        }
        if (total) *total = (int)run;
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (int64_t i = a; i < e; i++) {
        const uint32_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
}

template <class L, class S>
__global__ __launch_bounds__(SC_T) void scan_down_kernel(L ld, S st, int64_t n,
                                                         const uint32_t *boff)
{
    __shared__ uint32_t w[SC_T / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t b0 = (int64_t)blockIdx.x * SC_B;
    uint32_t run = boff[blockIdx.x];
    for (int e = 0; e < SC_E; e++) {
        const int64_t i = b0 + (int64_t)e * SC_T + threadIdx.x;
        const uint32_t v = i < n ? ld(i) : 0u;
        // inclusive scan across the wave, then across the four waves
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) w[wave] = x;
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (int k = 0; k < SC_T / 64; k++) {
            if (k < wave) before += w[k];
            all += w[k];
        }
        if (i < n) st(i, run + before + x - v);
        run += all;
        __syncthreads();
    }
}

template <class L, class S>
static hipError_t exclusive_scan(L ld, S st, int64_t n, uint32_t *bsum, int *total,
                                 hipStream_t stream)
{
    const int64_t nb = (n + SC_B - 1) / SC_B;
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(scan_reduce_kernel<L>, dim3((unsigned)nb), dim3(SC_T), 0, stream, ld, n, bsum);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(1024), 0, stream, bsum, nb, total);
    hipLaunchKernelGGL((scan_down_kernel<L, S>), dim3((unsigned)nb), dim3(SC_T), 0, stream, ld, st, n,
                       bsum);
    return hipGetLastError();
}

// Box cell of compact cell c in box b = {x0, W-1, y0, H-1} (y fastest), or -1.
__device__ __forceinline__ int box_bin(int c, int4 b, int nx, int ny, const FastDiv &dny)
{
    const int cx = (int)dny.div31((uint32_t)c), cy = c - cx * ny;
    int rx = cx - b.x, ry = cy - b.z;
    if (rx < 0) rx += nx;
    if (ry < 0) ry += ny;
    if (rx > b.y || ry > b.w) return -1;
    return rx * (b.w + 1) + ry;
}

// Digit histograms of all passes over the record keys.
__global__ __launch_bounds__(256) void rec_digit_hist_kernel(const uint32_t *k, int64_t n,
                                                             int npass, uint32_t *dhist)
{
    __shared__ uint32_t h[4][256];
    for (int q = threadIdx.x; q < 4 * 256; q += 256) h[q >> 8][q & 255] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        const uint32_t v = k[i];
        for (int p = 0; p < npass; p++) atomicAdd(&h[p][(v >> (8 * p)) & 255u], 1u);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < npass * 256; q += 256)
        if (h[q >> 8][q & 255]) atomicAdd(dhist + q, h[q >> 8][q & 255]);
}

// One push tile per workgroup: the tile's records and its particles in order
// of box cell, ties in slot order (a stable in-tile counting sort).  Each wave
// counts its quarter of the tile per box cell (an LDS histogram); the sums
// are the records at toff[t] (count, sort key, tile and box cell), and the
// counts become per-wave bases (the cell's offset in the tile, an exclusive
// scan over the box, plus the earlier waves' counts); each wave then walks
// its quarter again in slot order, ranks equal box cells within a round by
// ballots and writes the slot (in the tile) at its position,
// local[t*PUSH_TILE + pos].  loff receives each box cell's offset in the
// tile, at the cell's record.  A particle outside its tile's box sets
// BXF_OUTSIDE in ctl (the caller then takes the radix path).
__global__ __launch_bounds__(256) void tile_order_kernel(const int32_t *cell, int64_t n, int nx,
                                                         int ny, CellKey key, const int32_t *tbox,
                                                         const uint32_t *toff, uint32_t *thist,
                                                         uint32_t *rkey, uint32_t *rinfo,
                                                         uint32_t *loff, uint16_t *local, int *ctl)
{
    extern __shared__ uint32_t cnt[];    // [4][A]
    __shared__ uint32_t part[256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t t = blockIdx.x;
    const int4 b = *reinterpret_cast<const int4 *>(tbox + 4 * t);
    const int A = (b.y + 1) * (b.w + 1);
    for (int k = tid; k < 4 * A; k += 256) cnt[k] = 0;
    __syncthreads();
    const int q0 = wave * (PUSH_TILE / 4);
    uint32_t *mine = cnt + wave * A;
    const FastDiv &dny = key.dny;
    // the tile's cells, kept in registers for the ranking pass
    int bins[PUSH_TILE / 256];
    bool outside = false;
#pragma unroll
    for (int r = 0; r < PUSH_TILE / 256; r++) {
        const int64_t i = t * PUSH_TILE + q0 + r * 64 + lane;
        bins[r] = -1;
        if (i < n) {
            const int c = cell[i];
            const int bin = (unsigned)c < (unsigned)(nx * ny) ? box_bin(c, b, nx, ny, dny) : -1;
            bins[r] = bin;
            if (bin >= 0) atomicAdd(&mine[bin], 1u);
            else outside = true;
        }
    }
    if (outside) atomicOr(ctl, BXF_OUTSIDE);
    __syncthreads();
    // in-tile offset of every box cell: each thread scans a contiguous chunk
    const int per = (A + 255) / 256, k0 = tid * per, k1 = min(A, k0 + per);
    uint32_t sum = 0;
    for (int k = k0; k < k1; k++) sum += cnt[k] + cnt[A + k] + cnt[2 * A + k] + cnt[3 * A + k];
    part[tid] = sum;
    __syncthreads();
    if (tid < 64) {                      // exclusive scan of the 256 chunk sums
        uint32_t v[4], tot = 0;
        for (int j = 0; j < 4; j++) { v[j] = part[4 * tid + j]; tot += v[j]; }
        uint32_t x = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        uint32_t run = x - tot;
        for (int j = 0; j < 4; j++) { part[4 * tid + j] = run; run += v[j]; }
    }
    __syncthreads();
    uint32_t run = part[tid];
    const uint32_t o = toff[t];
    const int H = b.w + 1;
    for (int k = k0; k < k1; k++) {
        const uint32_t c0 = cnt[k], c1 = cnt[A + k], c2 = cnt[2 * A + k], c3 = cnt[3 * A + k];
        int gx = b.x + k / H, gy = b.z + k % H;
        if (gx >= nx) gx -= nx;
        if (gy >= ny) gy -= ny;
        thist[o + k] = c0 + c1 + c2 + c3;
        rkey[o + k] = key((uint32_t)(gx * ny + gy));
        rinfo[o + k] = (uint32_t)(t << 12) | (uint32_t)k;       // tile, box cell (< 2^12)
        loff[o + k] = run;
        cnt[k] = run;
        cnt[A + k] = run + c0;
        cnt[2 * A + k] = run + c0 + c1;
        cnt[3 * A + k] = run + c0 + c1 + c2;
        run += c0 + c1 + c2 + c3;
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1;
    const int nbits = A > 1 ? 32 - __clz(A - 1) : 0;     // bins < 2^nbits <= BX_MAX = 2^12
#pragma unroll
    for (int r = 0; r < PUSH_TILE / 256; r++) {
        const int q = q0 + r * 64 + lane;
        // a particle outside the box is left out (the caller falls back)
        const bool valid = bins[r] >= 0;
        const int bin = valid ? bins[r] : 0;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int j = 0; j < 12; j++) {
            if (j >= nbits) break;                       // block-uniform
            const uint64_t bb = __ballot((bin >> j) & 1);
            m &= ((bin >> j) & 1) ? bb : ~bb;
        }
        const uint32_t base = mine[bin];
        if (valid && (m & lt) == 0) mine[bin] = base + (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
        const uint32_t pos = base + (uint32_t)__popcll(m & lt);
        if (valid && pos < PUSH_TILE) local[t * PUSH_TILE + pos] = (uint16_t)q;
    }
}

// The records of the tile-box path in key order, as expand_kernel reads
// them: order[r] = index of the r-th record, sbase[r] = its first rank.
struct BoxRecs {
    const uint32_t *order, *sbase, *thist, *rinfo, *loff;
    const uint16_t *local;
    int64_t R;
};

// The permutation in rank order from the records in key order: one wave per
// 64 consecutive records, whose ranks form one contiguous range.  Lane l
// holds record r0 + l (first rank B, c particles, tile t, cell offset L in
// the tile); the wave walks the range 64 ranks at a time, each lane finding
// the record of its rank d (the last lane with B <= d) by a binary search
// over the lanes, and writes perm[d] = t*PUSH_TILE + local[t*PUSH_TILE + L +
// d - B]: coalesced writes, contiguous reads of each tile's local order.
__global__ __launch_bounds__(256) void expand_kernel(BoxRecs B, int64_t n, uint32_t *perm,
                                                     int *err)
{
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    if (r0 >= B.R) return;
    const int64_t r = r0 + lane;
    uint32_t b0 = 0xFFFFFFFFu, c = 0, lb = 0, tb = 0;
    if (r < B.R) {
        const uint32_t idx = B.order[r];
        b0 = B.sbase[r];
        c = B.thist[idx];
        tb = (B.rinfo[idx] >> 12) * (uint32_t)PUSH_TILE;
        lb = tb + B.loff[idx] - b0;              // local index = lb + rank (mod 2^32)
    }
    const int last = B.R - r0 < 64 ? (int)(B.R - r0) - 1 : 63;
    const int64_t D0 = (uint32_t)__shfl(b0, 0);
    const int64_t D1 = (int64_t)(uint32_t)__shfl(b0, last) + (uint32_t)__shfl(c, last);
    if (D1 > n) {
        if (lane == 0) atomicOr(err, ERRF_STATE);
        return;
    }
    for (int64_t base = D0; base < D1; base += 64) {     // uniform trip count for the shuffles
        const int64_t d = base + lane;
        int lo = 0, hi = last;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const int mid = (lo + hi + 1) >> 1;
            const uint32_t bm = (uint32_t)__shfl(b0, mid);
            if (lo < hi) {
                if ((int64_t)bm <= d) lo = mid; else hi = mid - 1;
            }
        }
        const uint32_t l = (uint32_t)__shfl(lb, lo), t = (uint32_t)__shfl(tb, lo);
        if (d < D1) perm[d] = t + B.local[(uint32_t)(l + (uint32_t)d)];
    }
}

namespace {
int sgn_i(int v) { return (v > 0) - (v < 0); }
int floor_half(int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }

// One rectangle of the generalized Hilbert curve: origin (x, y), major axis
// (ax, ay), minor axis (bx, by); cells are numbered from *next on.  A long
// rectangle (major more than 1.5 x minor) is cut in two along its major axis;
// otherwise it is walked as up / across / down, the Hilbert curve's three
// parts, with the cuts moved by one cell where a half would be odd so that
// the pieces chain by edge steps.
void gilbert_walk(int x, int y, int ax, int ay, int bx, int by, int h,
                  std::vector<uint32_t> &rank, uint32_t &next)
{
    const int w = std::abs(ax + ay), hh = std::abs(bx + by);
    const int dax = sgn_i(ax), day = sgn_i(ay), dbx = sgn_i(bx), dby = sgn_i(by);
    if (hh == 1) {
        for (int i = 0; i < w; i++, x += dax, y += day) rank[(size_t)x * h + y] = next++;
        return;
    }
    if (w == 1) {
        for (int i = 0; i < hh; i++, x += dbx, y += dby) rank[(size_t)x * h + y] = next++;
        return;
    }
    int ax2 = floor_half(ax), ay2 = floor_half(ay), bx2 = floor_half(bx), by2 = floor_half(by);
    const int w2 = std::abs(ax2 + ay2), h2 = std::abs(bx2 + by2);
    if (2 * w > 3 * hh) {
        if ((w2 & 1) && w > 2) { ax2 += dax; ay2 += day; }
        gilbert_walk(x, y, ax2, ay2, bx, by, h, rank, next);
        gilbert_walk(x + ax2, y + ay2, ax - ax2, ay - ay2, bx, by, h, rank, next);
    } else {
        if ((h2 & 1) && hh > 2) { bx2 += dbx; by2 += dby; }
        gilbert_walk(x, y, bx2, by2, ax2, ay2, h, rank, next);
        gilbert_walk(x + bx2, y + by2, ax, ay, bx - bx2, by - by2, h, rank, next);
        gilbert_walk(x + (ax - dax) + (bx2 - dbx), y + (ay - day) + (by2 - dby), -bx2, -by2,
                     -(ax - ax2), -(ay - ay2), h, rank, next);
    }
}
}  // namespace

void supertile_ranks(int w, int h, std::vector<uint32_t> &rank)
{
    rank.assign((size_t)w * h, 0u);
    uint32_t next = 0;
    if (w <= 0 || h <= 0) return;
    if (w >= h) gilbert_walk(0, 0, w, 0, 0, h, h, rank, next);
    else gilbert_walk(0, 0, 0, h, w, 0, h, rank, next);
}

// The tile-box path of sort_species up to the records in key order (*ok), or
// *ok = false when a tile's box is unknown, too large, or does not hold all
// its particles (the caller then takes the radix path).
static hipError_t box_records(SortScratch &s, const SpeciesDev &sp, const CellKey &key,
                              int64_t ncell, int *err, BoxRecs *out, bool *ok, hipStream_t st)
{
    *ok = false;
    hipError_t e;
    const int64_t n = sp.n, nt = (n + PUSH_TILE - 1) / PUSH_TILE;
    const int nx = (int)(ncell / key.ny), ny = (int)key.ny;
    if ((e = launch_fill_u32((uint32_t *)s.ctl, 8, 0u, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(box_area_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st,
                       (const int32_t *)sp.tbox, nt, s.tarea, s.ctl);
    if ((e = exclusive_scan(LdArr{s.tarea}, StArr{s.toff}, nt, s.bsum, s.ctl + 2, st)) != hipSuccess)
        return e;
    int ctl[4];
    if ((e = hipMemcpyAsync(ctl, s.ctl, sizeof ctl, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    const int64_t R = (uint32_t)ctl[2];
    const int amax = ctl[1];
    // the record radix reuses the particle look-back status (status_cap tiles)
    if (ctl[0] || R == 0 || (R + ST_TILE - 1) / ST_TILE > s.status_cap) return hipSuccess;
    constexpr int NREC = 9;                      // record arrays in rbuf
    if (R > s.rcap) {
        if (s.rbuf) (void)hipFree(s.rbuf);
        s.rbuf = nullptr;
        s.rcap = 0;
        const int64_t cap = R + R / 4 + 4096;
        if ((e = hipMalloc(&s.rbuf, (size_t)cap * NREC * sizeof(uint32_t))) != hipSuccess) {
            (void)hipGetLastError();
            return hipSuccess;                   // no room: the radix path
        }
        s.rcap = cap;
    }
    uint32_t *thist = s.rbuf, *rkey = s.rbuf + s.rcap, *sbase = s.rbuf + 2 * s.rcap;
    uint32_t *rinfo = s.rbuf + 3 * s.rcap, *loff = s.rbuf + 4 * s.rcap;
    uint32_t *kb[2] = {s.rbuf + 5 * s.rcap, s.rbuf + 6 * s.rcap};
    uint32_t *vb[2] = {s.rbuf + 7 * s.rcap, s.rbuf + 8 * s.rcap};
    // the tiles' records, and their particles by box cell (u16 slots, in the
    // first key buffer)
    uint16_t *local = reinterpret_cast<uint16_t *>(s.keys[0]);
    // up to 64 KiB of per-wave counts (BX_MAX cells x 4 waves x 4 bytes) beside
    // the static scan array: above the default dynamic LDS limit for the
    // largest boxes
    (void)hipFuncSetAttribute((const void *)tile_order_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, amax * 16);
    hipLaunchKernelGGL(tile_order_kernel, dim3((unsigned)nt), dim3(256), (size_t)amax * 16, st,
                       sp.cell, n, nx, ny, key, (const int32_t *)sp.tbox, s.toff, thist, rkey, rinfo,
                       loff, local, s.ctl);
    if ((e = hipMemcpyAsync(ctl, s.ctl, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    if (ctl[0]) return hipSuccess;
    // stable radix sort of the records by key (record index = tile-major)
    int bits = 0;
    while ((1ll << bits) < key.nkeys) bits++;
    const int npass = bits > 8 ? (bits + 7) / 8 : 1;
    if ((e = launch_fill_u32(s.dhist, 4 * 256, 0u, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(rec_digit_hist_kernel, dim3(1024), dim3(256), 0, st, rkey, R, npass, s.dhist);
    hipLaunchKernelGGL(digit_scan_kernel, dim3(1), dim3(64), 0, st, s.dhist, npass);
    CellKey ident;
    ident.init(1, 1, 0);                         // key(c) = c: the records carry their keys
    const int64_t rtile = (R + ST_TILE - 1) / ST_TILE;
    const uint32_t *kin = rkey, *vin = nullptr;
    int pb = 0;
    for (int p = 0; p < npass; p++) {
        if ((e = launch_fill_u32(s.status, rtile * 256, 0u, st)) != hipSuccess) return e;
        if ((e = launch_fill_u32(s.ticket + p, 1, 0u, st)) != hipSuccess) return e;
        const bool first = p == 0, last = p == npass - 1;
        const uint32_t *db = s.dhist + p * 256;
        dim3 g((unsigned)rtile), bl(ST_THREADS);
        if (first && last)
            hipLaunchKernelGGL((radix_pass_kernel<true, true>), g, bl, 0, st, kin, vin, kb[pb], vb[pb],
                               R, 8 * p, db, s.status, s.ticket + p, err, ident);
        else if (first)
            hipLaunchKernelGGL((radix_pass_kernel<true, false>), g, bl, 0, st, kin, vin, kb[pb], vb[pb],
                               R, 8 * p, db, s.status, s.ticket + p, err, ident);
        else if (last)
            hipLaunchKernelGGL((radix_pass_kernel<false, true>), g, bl, 0, st, kin, vin, kb[pb], vb[pb],
                               R, 8 * p, db, s.status, s.ticket + p, err, ident);
        else
            hipLaunchKernelGGL((radix_pass_kernel<false, false>), g, bl, 0, st, kin, vin, kb[pb], vb[pb],
                               R, 8 * p, db, s.status, s.ticket + p, err, ident);
        kin = kb[pb];
        vin = vb[pb];
        pb ^= 1;
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // the records' counts scanned in key order: each record's first rank
    if ((e = exclusive_scan(LdGather{thist, vin}, StArr{sbase}, R, s.bsum, nullptr, st)) !=
        hipSuccess)
        return e;
    *out = BoxRecs{vin, sbase, thist, rinfo, loff, local, R};
    *ok = true;
    return hipSuccess;
}

// Adds species sp's particles per cell (compact cell id) to counts, as
// numdensity (pic2d.f90:1343-1354): through the tile boxes of the last push
// when they are current, else with run-aggregated atomics.
hipError_t launch_cell_counts(const SpeciesDev &sp, int nx, int ny, int32_t *counts, int *err,
                              hipStream_t st)
{
    if (sp.n == 0) return hipSuccess;
    CellKey key;
    key.init(nx, ny, 0);
    const int64_t ncell = (int64_t)nx * ny;
    if (sp.tbox_ok)
        hipLaunchKernelGGL(cell_hist_tiles_kernel, dim3((unsigned)((sp.n + PUSH_TILE - 1) / PUSH_TILE)),
                           dim3(256), 0, st, sp.cell, sp.n, nx, ncell, key,
                           (const int32_t *)sp.tbox, counts, err);
    else
        hipLaunchKernelGGL(cell_hist_kernel, dim3((unsigned)((sp.n + 255) / 256)), dim3(256), 0, st,
                           sp.cell, sp.n, ncell, key, counts, err);
    return hipGetLastError();
}

// Sorts species sp into dst (a second buffer set of the same capacity).
hipError_t sort_species(SortScratch &s, const SpeciesDev &sp, int precision, int push_order,
                        const CellKey &key, int64_t ncell,
                        SpeciesDev &dst, int *err, int want_counts, int want_perm, void *rec,
                        bool rec_written, bool *boxed, hipStream_t st)
{
    *boxed = false;
    const int64_t n = sp.n;
    hipError_t e;
    dst.n = n;
    const int64_t nkeys = key.nkeys;
    if (n == 0) return launch_fill_u32((uint32_t *)s.counts, nkeys, 0u, st);
    (void)want_perm;
    const uint32_t *perm = nullptr;
    int pb = 0;
    if (sp.tbox_ok && !want_counts) {
        BoxRecs br{};
        bool ok = false;
        if ((e = box_records(s, sp, key, ncell, err, &br, &ok, st)) != hipSuccess) return e;
        if (ok) {
            *boxed = true;
            // the permutation goes to vals[1]; the slot map (push order) to vals[0]
            hipLaunchKernelGGL(expand_kernel, dim3((unsigned)((br.R + 255) / 256)), dim3(256), 0, st,
                               br, n, s.vals[1], err);
            perm = s.vals[1];
            pb = 0;
        }
    }
    if (!perm) {
    int bits = 0;
    while ((1ll << bits) < nkeys) bits++;
    int npass = (bits + 7) / 8;
    if (npass < 1) npass = 1;
    const int64_t ntile = (n + ST_TILE - 1) / ST_TILE;

    if ((e = launch_fill_u32((uint32_t *)s.counts, nkeys, 0u, st)) != hipSuccess) return e;
    if ((e = launch_fill_u32(s.dhist, 4 * 256, 0u, st)) != hipSuccess) return e;
    if (sp.tbox_ok)
        hipLaunchKernelGGL(cell_hist_tiles_kernel, dim3((unsigned)((n + PUSH_TILE - 1) / PUSH_TILE)),
                           dim3(256), 0, st, sp.cell, n, (int)(ncell / key.ny), ncell, key,
                           (const int32_t *)sp.tbox, s.counts, err);
    else
        hipLaunchKernelGGL(cell_hist_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           sp.cell, n, ncell, key, s.counts, err);
    const int64_t nhb = (nkeys + 256 * 64 - 1) / (256 * 64);
    hipLaunchKernelGGL(digit_hist_kernel, dim3((unsigned)nhb), dim3(256), 0, st, s.counts, nkeys,
                       npass, s.dhist);
    hipLaunchKernelGGL(digit_scan_kernel, dim3(1), dim3(64), 0, st, s.dhist, npass);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    const uint32_t *kin = (const uint32_t *)sp.cell;
    const uint32_t *vin = nullptr;
    for (int p = 0; p < npass; p++) {
        if ((e = launch_fill_u32(s.status, ntile * 256, 0u, st)) != hipSuccess) return e;
        if ((e = launch_fill_u32(s.ticket + p, 1, 0u, st)) != hipSuccess) return e;
        const bool first = p == 0, last = p == npass - 1;
        uint32_t *ko = s.keys[pb], *vo = s.vals[pb];
        const uint32_t *db = s.dhist + p * 256;
        dim3 g((unsigned)ntile), b(ST_THREADS);
        if (first && last)
            hipLaunchKernelGGL((radix_pass_kernel<true, true>), g, b, 0, st, kin, vin, ko, vo, n,
                               8 * p, db, s.status, s.ticket + p, err, key);
        else if (first)
            hipLaunchKernelGGL((radix_pass_kernel<true, false>), g, b, 0, st, kin, vin, ko, vo, n,
                               8 * p, db, s.status, s.ticket + p, err, key);
        else if (last)
            hipLaunchKernelGGL((radix_pass_kernel<false, true>), g, b, 0, st, kin, vin, ko, vo, n,
                               8 * p, db, s.status, s.ticket + p, err, key);
        else
            hipLaunchKernelGGL((radix_pass_kernel<false, false>), g, b, 0, st, kin, vin, ko, vo, n,
                               8 * p, db, s.status, s.ticket + p, err, key);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        kin = ko;
        vin = vo;
        pb ^= 1;
    }
    perm = vin;
    }
    s.perm = perm;
    // the push order needs the slot->source map in a buffer of its own: the
    // value buffer the radix's last pass did not write is free (vals[0] after
    // the tile-box path)
    const int tv = push_order && n >= PUSH_TILE ? push_vec(precision) : 0;
    uint32_t *pout = nullptr;
    if (tv) s.perm = pout = s.vals[pb];
    const int G = precision == PBPIC_PREC_FP64 ? 2 : 4;        // 16 bytes of R per lane
    const unsigned nb = (unsigned)(((n + G - 1) / G + 255) / 256);
    if (precision == PBPIC_PREC_FP64) {
        hipLaunchKernelGGL(gather_kernel<double>, dim3(nb), dim3(256), 0, st, perm, n, tv, pout, sp, dst);
    } else if (rec) {
        if (!rec_written)
            hipLaunchKernelGGL(pack_records_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                               st, sp, (PartRec *)rec, n);
        hipLaunchKernelGGL(gather_records_kernel, dim3((nb + 1) / 2), dim3(256), 0, st, perm, n, tv, pout,
                           (const PartRec *)rec, sp, dst);
    } else {
        hipLaunchKernelGGL(gather_kernel<float>, dim3(nb), dim3(256), 0, st, perm, n, tv, pout, sp, dst);
    }
    return hipGetLastError();
}

}  // namespace pbpic
