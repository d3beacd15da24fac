// k_sort.hip -- countingsort (pic2d.f90:1303-1329) for gfx950.
//
// The reference is a stable counting sort of the particles by cell id:
// histogram (:1310-1313), exclusive scan (:1315-1320), stable scatter
// (:1322-1327), copy back (:1328).  With millions of cells a per-block x
// per-cell count matrix does not fit, so the stable scatter is done as an LSD
// radix sort of (cell, index) pairs over 8-bit digits; each digit pass is
// itself a counting sort (a per-wave ballot histogram with in-order ranks,
// then a decoupled look-back prefix scan across tiles).  A final gather moves
// the particle payload once.  The result (permutation and per-cell counts) is
// identical to the reference's: both are the unique stable sort by cell.
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
                        SpeciesDev &dst, int *err, int want_counts, void *rec,
                        bool rec_written, hipStream_t st)
{
    (void)want_counts;
    const int64_t n = sp.n;
    hipError_t e;
    dst.n = n;
    const int64_t nkeys = key.nkeys;
    if (n == 0) return launch_fill_u32((uint32_t *)s.counts, nkeys, 0u, st);
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
    int pb = 0;
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
    const uint32_t *perm = vin;
    s.perm = perm;
    // the push order needs the slot->source map in a buffer of its own: the
    // value buffer the last pass did not write is free
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
