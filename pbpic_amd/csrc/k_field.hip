// k_field.hip -- Yee/FDTD field solve and field-side reductions for gfx950.
//
// fieldsolver (pic2d.f90:678-860) on the periodic compact lattice: one fused,
// LDS-staged stencil per 32x16 output tile.  E^n is staged with a 3-cell halo
// and B^n with the same halo; the block then computes in LDS
//   B^{n+1/2} on tile+2  (:688-690)
//   E^{n+1}   on tile+1  (:742-744, current normalized by in0 as at :1066)
//   B^{n+1}   on tile    (:801-803)
// and writes E^{n+1}, B^{n+1} once: 9 planes read (E, B, J), 6 written.
// The new fields go to a second set of planes (FdtdParams::nex..nbz): every
// workgroup reads a 3-cell halo of its neighbours' E^n and B^n, and nothing
// orders workgroups, so writing in place would let a late tile read a
// finished neighbour's E^{n+1} as its halo.  The
// curl terms the reference stores between steps are recomputed from E and B,
// which yields the same values (they are formed from the same fields by the
// same expressions, :733-736, :790-793).
//
// Conducting y walls (P.nyp > 0; the reference leaves this boundary commented
// out at :850-859): the last two rows are guards holding mirror images of the
// fields.  Inside the kernel only two of them matter to the domain rows: Ex
// is zero on the wall rows 0 and nyp (set as E^{n+1} is formed, before J.E)
// and B^{n+1} on the wall rows reads Ez^{n+1} in the guard rows, imaged in
// LDS (odd) before that stage.  B^{n+1/2} in the guard rows feeds only Ex on
// the walls, which is overwritten.  wall_images_kernel then rewrites every
// guard row of every plane after the launch.
#include "pbpic_impl.h"

#pragma clang fp contract(off)

namespace pbpic {

constexpr int FT_X = 32, FT_Y = 16, FT_H = 3;
constexpr int FS_X = FT_X + 2 * FT_H, FS_Y = FT_Y + 2 * FT_H;   // 38 x 22
constexpr int FS_N = FS_X * FS_Y;
constexpr int FT_THREADS = 256;

// SPLIT: the stencil differences are taken on the FP64 state, rounded to FP32,
// and the update increments are formed in FP32 and accumulated into FP64.
template <typename RJ, bool SPLIT>
__global__ __launch_bounds__(FT_THREADS) void fdtd_kernel(FdtdParams P)
{
    __shared__ double se[3][FS_N];
    __shared__ double sb[3][FS_N];
    __shared__ double jred[2][FT_THREADS / 64];
    const int tid = threadIdx.x;
    const int nx = P.nx, ny = P.ny;
    const int ntx = (nx + FT_X - 1) / FT_X;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = t % ntx, ty = t / ntx;
    const int x0 = tx * FT_X - FT_H, y0 = ty * FT_Y - FT_H;

    for (int k = tid; k < FS_N; k += FT_THREADS) {
        const int v = k / FS_X, u = k - v * FS_X;
        const size_t g = (size_t)wrap_index(y0 + v, ny) * nx + wrap_index(x0 + u, nx);
        se[0][k] = P.ex[g]; se[1][k] = P.ey[g]; se[2][k] = P.ez[g];
        sb[0][k] = P.bx[g]; sb[1][k] = P.by[g]; sb[2][k] = P.bz[g];
    }
    __syncthreads();

    const double dth = P.dth, idx = P.idx, idy = P.idy;
    const float dthf = (float)P.dth, idxf = (float)P.idx, idyf = (float)P.idy;
    // B^{n+1/2} on u in [1, FS_X-1), v in [1, FS_Y-1)
    for (int k = tid; k < (FS_X - 2) * (FS_Y - 2); k += FT_THREADS) {
        const int v = k / (FS_X - 2) + 1, u = k - (v - 1) * (FS_X - 2) + 1;
        const int c = v * FS_X + u;
        if (!SPLIT) {
            sb[0][c] = sb[0][c] - dth * ((se[2][c] - se[2][c - FS_X]) * idy);
            sb[1][c] = sb[1][c] + dth * ((se[2][c] - se[2][c - 1]) * idx);
            sb[2][c] = sb[2][c] + dth * (((se[0][c + FS_X] - se[0][c]) * idy) -
                                         ((se[1][c + 1] - se[1][c]) * idx));
        } else {
            const float dzy = (float)(se[2][c] - se[2][c - FS_X]);
            const float dzx = (float)(se[2][c] - se[2][c - 1]);
            const float dxy = (float)(se[0][c + FS_X] - se[0][c]);
            const float dyx = (float)(se[1][c + 1] - se[1][c]);
            sb[0][c] += (double)(-(dthf * (dzy * idyf)));
            sb[1][c] += (double)(dthf * (dzx * idxf));
            sb[2][c] += (double)(dthf * ((dxy * idyf) - (dyx * idxf)));
        }
    }
    __syncthreads();

    // E^{n+1} on u in [2, FS_X-2), v in [2, FS_Y-2); J.E on the owned tile
    const double dtc2 = P.dtc2, in0 = P.in0;
    const float dtc2f = (float)P.dtc2, in0f = (float)P.in0;
    const RJ *jx = (const RJ *)P.jx, *jy = (const RJ *)P.jy, *jz = (const RJ *)P.jz;
    double je = 0, jem = 0;
    for (int k = tid; k < (FS_X - 4) * (FS_Y - 4); k += FT_THREADS) {
        const int v = k / (FS_X - 4) + 2, u = k - (v - 2) * (FS_X - 4) + 2;
        const int c = v * FS_X + u;
        const int gx = wrap_index(x0 + u, nx), gy = wrap_index(y0 + v, ny);
        const size_t g = (size_t)gy * nx + gx;
        const double jxr = (double)jx[g], jyr = (double)jy[g], jzr = (double)jz[g];
        const double exo = se[0][c], eyo = se[1][c], ezo = se[2][c];
        double exn, eyn, ezn;
        if (!SPLIT) {
            exn = exo + dtc2 * (((sb[2][c] - sb[2][c - FS_X]) * idy) - jxr * in0);
            eyn = eyo - dtc2 * (((sb[2][c] - sb[2][c - 1]) * idx) + jyr * in0);
            ezn = ezo + dtc2 * ((((sb[1][c + 1] - sb[1][c]) * idx) -
                                 ((sb[0][c + FS_X] - sb[0][c]) * idy)) - jzr * in0);
        } else {
            const float bzy = (float)(sb[2][c] - sb[2][c - FS_X]);
            const float bzx = (float)(sb[2][c] - sb[2][c - 1]);
            const float byx = (float)(sb[1][c + 1] - sb[1][c]);
            const float bxy = (float)(sb[0][c + FS_X] - sb[0][c]);
            exn = exo + (double)(dtc2f * ((bzy * idyf) - (float)jxr * in0f));
            eyn = eyo + (double)(-(dtc2f * ((bzx * idxf) + (float)jyr * in0f)));
            ezn = ezo + (double)(dtc2f * (((byx * idxf) - (bxy * idyf)) - (float)jzr * in0f));
        }
        if (P.nyp && (gy == 0 || gy == P.nyp)) exn = 0.0;   // tangential E on a wall
        const bool own = u >= FT_H && u < FT_H + FT_X && v >= FT_H && v < FT_H + FT_Y &&
                         x0 + u < nx && y0 + v < ny;
        if (own) {
            const double jp = jxr * 0.5 * (exo + exn) + jyr * 0.5 * (eyo + eyn) + jzr * 0.5 * (ezo + ezn);
            je += jp;
            if (P.lmask) jem += P.lmask[g] * jp;
        }
        if (P.fex) {
            // Marder / Cherenkov filters (:754-755), increments from E^n
            exn += P.fex[g]; eyn += P.fey[g]; ezn += P.fez[g];
            if (P.nyp && (gy == 0 || gy == P.nyp)) exn = 0.0;
        }
        se[0][c] = exn; se[1][c] = eyn; se[2][c] = ezn;
    }
    __syncthreads();

    if (P.nyp) {
        // Ez^{n+1} in the guard rows of the E stage's range [2, FS_Y-2): the
        // odd image of the adjacent domain row (row nyp mirrors nyp-1, row
        // nyp+1 mirrors row 0).  A guard at the edge of the range whose
        // mirror lies outside it feeds only guard rows of B^{n+1}.
        bool any = false;
        for (int v = 2; v < FS_Y - 2; v++) {
            const int gy = wrap_index(y0 + v, ny);
            any |= gy >= P.nyp;
        }
        if (any) {
            for (int k = tid; k < (FS_X - 4) * (FS_Y - 4); k += FT_THREADS) {
                const int v = k / (FS_X - 4) + 2, u = k - (v - 2) * (FS_X - 4) + 2;
                const int gy = wrap_index(y0 + v, ny);
                const int c = v * FS_X + u;
                if (gy == P.nyp && v - 1 >= 2) se[2][c] = -se[2][c - FS_X];
                else if (gy == P.nyp + 1 && v + 1 < FS_Y - 2) se[2][c] = -se[2][c + FS_X];
            }
            __syncthreads();
        }
    }

    // B^{n+1} and write-out on the owned tile
    for (int k = tid; k < FT_X * FT_Y; k += FT_THREADS) {
        const int vv = k / FT_X, uu = k - vv * FT_X;
        const int u = uu + FT_H, v = vv + FT_H;
        if (x0 + u >= nx || y0 + v >= ny) continue;
        const int c = v * FS_X + u;
        double bxn, byn, bzn;
        if (!SPLIT) {
            bxn = sb[0][c] - dth * ((se[2][c] - se[2][c - FS_X]) * idy);
            byn = sb[1][c] + dth * ((se[2][c] - se[2][c - 1]) * idx);
            bzn = sb[2][c] + dth * (((se[0][c + FS_X] - se[0][c]) * idy) -
                                    ((se[1][c + 1] - se[1][c]) * idx));
        } else {
            const float dzy = (float)(se[2][c] - se[2][c - FS_X]);
            const float dzx = (float)(se[2][c] - se[2][c - 1]);
            const float dxy = (float)(se[0][c + FS_X] - se[0][c]);
            const float dyx = (float)(se[1][c + 1] - se[1][c]);
            bxn = sb[0][c] + (double)(-(dthf * (dzy * idyf)));
            byn = sb[1][c] + (double)(dthf * (dzx * idxf));
            bzn = sb[2][c] + (double)(dthf * ((dxy * idyf) - (dyx * idxf)));
        }
        const size_t g = (size_t)(y0 + v) * nx + (x0 + u);
        P.nex[g] = se[0][c]; P.ney[g] = se[1][c]; P.nez[g] = se[2][c];
        P.nbx[g] = bxn; P.nby[g] = byn; P.nbz[g] = bzn;
    }

#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        je += __shfl_xor(je, o);
        jem += __shfl_xor(jem, o);
    }
    if ((tid & 63) == 0) { jred[0][tid >> 6] = je; jred[1][tid >> 6] = jem; }
    __syncthreads();
    if (tid == 0) {
        // per-workgroup partials, summed in a fixed order after the launch
        P.part[blockIdx.x] = (jred[0][0] + jred[0][1]) + (jred[0][2] + jred[0][3]);
        P.part[gridDim.x + blockIdx.x] = (jred[1][0] + jred[1][1]) + (jred[1][2] + jred[1][3]);
    }
}

hipError_t launch_fdtd(const FdtdParams &p, int precision, hipStream_t st)
{
    const int nb = ((p.nx + FT_X - 1) / FT_X) * ((p.ny + FT_Y - 1) / FT_Y);
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL((fdtd_kernel<double, false>), dim3(nb), dim3(FT_THREADS), 0, st, p);
    else
        hipLaunchKernelGGL((fdtd_kernel<float, true>), dim3(nb), dim3(FT_THREADS), 0, st, p);
    // the previous step's J.E (and loss) values are integrated as they are
    // replaced: JE_INT += JE_STEP * dt mime/c^2, then JE_STEP = this step's
    PartSlots ps{2, {ACC_JE_STEP, ACC_JEM_STEP, 0, 0}, {1.0, 1.0, 0, 0},
                 {ACC_JE_INT, ACC_JEM_INT, -1, -1}, p.je_pre};
    return launch_reduce_parts(p.part, nb, p.acc, ps, st);
}

// One workgroup per quantity: thread t sums part[q*nb + t + 256 k] for
// k = 0, 1, ... in order, then a fixed tree over the threads.
__global__ __launch_bounds__(256) void reduce_parts_kernel(const double *part, int nb, double *acc,
                                                           PartSlots ps)
{
    __shared__ double w[4];
    const int q = blockIdx.x;
    const double *p = part + (size_t)q * nb;
    double s = 0;
    for (int b = threadIdx.x; b < nb; b += 256) s += p[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double v = ps.scale[q] * ((w[0] + w[1]) + (w[2] + w[3]));
        if (ps.ints[q] >= 0) {
            acc[ps.ints[q]] += acc[ps.slot[q]] * ps.pre;
            acc[ps.slot[q]] = v;
        } else {
            acc[ps.slot[q]] += v;
        }
    }
}

hipError_t launch_reduce_parts(const double *part, int nb, double *acc, const PartSlots &ps,
                               hipStream_t st)
{
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(ps.nq), dim3(256), 0, st, part, nb, acc, ps);
    return hipGetLastError();
}

// fieldsenergy (pic2d.f90:653-673): sums of squared node-averaged E and B
// (emfgrid, :837-845) over all nodes, FP64.  With conducting y walls: over
// the node rows 0..nyp, the two wall rows weighted 1/2.
__global__ __launch_bounds__(256) void energy_kernel(FdtdParams P)
{
    __shared__ double red[2][4];
    const int nx = P.nx, ny = P.ny;
    const int64_t n = (int64_t)nx * ny;
    double se = 0, sbv = 0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256) {
        const int gy = (int)(k / nx), gx = (int)(k - (int64_t)gy * nx);
        const int gxm = gx == 0 ? nx - 1 : gx - 1, gym = gy == 0 ? ny - 1 : gy - 1;
        const size_t c = k, xm = (size_t)gy * nx + gxm, ym = (size_t)gym * nx + gx,
                     xym = (size_t)gym * nx + gxm;
        const double ex = (P.ex[c] + P.ex[xm]) * 0.5, ey = (P.ey[c] + P.ey[ym]) * 0.5;
        const double ez = (P.ez[c] + P.ez[xm] + P.ez[ym] + P.ez[xym]) * 0.25;
        const double bx = (P.bx[c] + P.bx[xm]) * 0.5, by = (P.by[c] + P.by[ym]) * 0.5;
        const double bz = (P.bz[c] + P.bz[xm] + P.bz[ym] + P.bz[xym]) * 0.25;
        double e2 = ex * ex + ey * ey + ez * ez, b2 = bx * bx + by * by + bz * bz;
        if (P.nyp) {
            if (gy > P.nyp) continue;
            if (gy == 0 || gy == P.nyp) { e2 *= 0.5; b2 *= 0.5; }
        }
        se += e2;
        sbv += b2;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        se += __shfl_xor(se, o);
        sbv += __shfl_xor(sbv, o);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = se;
        red[1][threadIdx.x >> 6] = sbv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        P.part[blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        P.part[gridDim.x + blockIdx.x] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

hipError_t launch_energy(const FdtdParams &p, hipStream_t st)
{
    const int64_t n = (int64_t)p.nx * p.ny;
    int nb = (int)((n + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(energy_kernel, dim3(nb), dim3(256), 0, st, p);
    PartSlots ps{2, {ACC_WE, ACC_WB, 0, 0}, {1.0, 1.0, 0, 0}};
    return launch_reduce_parts(p.part, nb, p.acc, ps, st);
}

// Gauss residual: rho by CIC with the gather's weights (:948-951) from the
// current positions, then divE - c^2 in0 (rho + rho_bg) at every node.
// The charge is deposited one push tile (PUSH_TILE particles) per workgroup:
// over the tile's bounding box (the one the last push recorded in
// SpeciesDev::tbox, else a pass over the tile's cells unwrapped around its
// first cell) the charge goes to an LDS node histogram, flushed with one
// global atomic per node; particles outside the box, or all of a tile whose
// box exceeds the histogram, add to rho directly.
constexpr int RB_NODES = 2048;
// The charge as 64-bit fixed point, sums exact in any order.  Split form
// (TW false): each weight times rscale, rounded.  FP64 form (TW): each weight
// times rscale = 2^30 and 2^FX2_LOBITS, rounded (fx2_quant), the global plane
// in two words (fx2_add, the low words lo entries on).  A tile adds at most
// PUSH_TILE weights below 2^49 to an LDS node: below 2^61.
template <bool TW>
__device__ __forceinline__ long long rho_quant(double v, double rscale)
{
    if (TW) return fx2_quant(v * rscale * (double)(1 << FX2_LOBITS));
    return (long long)rint(v * rscale);
}
template <bool TW>
__device__ __forceinline__ void rho_add(long long *p, long long q, int64_t lo)
{
    if (TW) fx2_add(p, lo, q);
    else atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)q);
}
__device__ __forceinline__ void rho_add_lds(long long *p, long long q)
{
    __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(p), (unsigned long long)q,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename R>
__global__ __launch_bounds__(256) void rho_cic_tiles_kernel(PushParams P, SpeciesDev S,
                                                            long long *rho, double rscale)
{
    constexpr bool TW = sizeof(R) == 8;
    const int64_t lo = TW ? (int64_t)P.nx * P.ny : 0;
    __shared__ long long h[RB_NODES];
    __shared__ int bb[4];
    const int nx = P.nx, ny = P.ny;
    const int64_t t0 = (int64_t)blockIdx.x * PUSH_TILE;
    int4 b = make_int4(0, -1, 0, -1);
    if (S.tbox_ok) {
        b = *reinterpret_cast<const int4 *>(S.tbox + 4 * (size_t)blockIdx.x);
    } else {
        const int c0 = S.cell[t0];
        const int x0 = (unsigned)c0 < (unsigned)(nx * ny) ? c0 / ny : 0;
        const int y0 = (unsigned)c0 < (unsigned)(nx * ny) ? c0 - x0 * ny : 0;
        const int nxh = nx >> 1, nyh = ny >> 1;
        int mnx = 1 << 30, mxx = -(1 << 30), mny = 1 << 30, mxy = -(1 << 30);
        if (threadIdx.x == 0) { bb[0] = mnx; bb[1] = mxx; bb[2] = mny; bb[3] = mxy; }
        for (int u = threadIdx.x; u < PUSH_TILE; u += 256) {
            const int64_t i = t0 + u;
            if (i >= S.n) break;
            const int c = S.cell[i];
            if ((unsigned)c >= (unsigned)(nx * ny)) continue;
            const int cx = c / ny, cy = c - cx * ny;
            int rx = cx - x0, ry = cy - y0;
            if (rx > nxh) rx -= nx; else if (rx < -nxh) rx += nx;
            if (ry > nyh) ry -= ny; else if (ry < -nyh) ry += ny;
            mnx = min(mnx, rx); mxx = max(mxx, rx); mny = min(mny, ry); mxy = max(mxy, ry);
        }
        __syncthreads();
        atomicMin(&bb[0], mnx); atomicMax(&bb[1], mxx);
        atomicMin(&bb[2], mny); atomicMax(&bb[3], mxy);
        __syncthreads();
        if (bb[0] <= bb[1] && bb[1] - bb[0] < nx && bb[3] - bb[2] < ny) {
            int bx = x0 + bb[0], by = y0 + bb[2];
            if (bx < 0) bx += nx; else if (bx >= nx) bx -= nx;
            if (by < 0) by += ny; else if (by >= ny) by -= ny;
            b = make_int4(bx, bb[1] - bb[0], by, bb[3] - bb[2]);
        }
    }
    const int W = b.y + 2, H = b.w + 2;                 // nodes of the box's cells
    const bool boxed = b.y >= 0 && b.w >= 0 && (int64_t)W * H <= RB_NODES;
    if (boxed)
        for (int k = threadIdx.x; k < W * H; k += 256) h[k] = 0;
    __syncthreads();
    const double q = (double)S.q;
    for (int u = threadIdx.x; u < PUSH_TILE; u += 256) {
        const int64_t i = t0 + u;
        if (i >= S.n) break;
        const int c = S.cell[i];
        if ((unsigned)c >= (unsigned)(nx * ny)) continue;
        const int cx = c / ny, cy = c - cx * ny;
        const double wx = 0.5 + (double)((const R *)S.ox)[i];
        const double wy = 0.5 + (double)((const R *)S.oy)[i];
        const double w00 = q * (1 - wx) * (1 - wy), w10 = q * wx * (1 - wy);
        const double w01 = q * (1 - wx) * wy, w11 = q * wx * wy;
        if (boxed) {
            int rx = cx - b.x, ry = cy - b.z;
            if (rx < 0) rx += nx;
            if (ry < 0) ry += ny;
            if (rx <= b.y && ry <= b.w) {
                long long *n0 = h + ry * W + rx;
                rho_add_lds(n0, rho_quant<TW>(w00, rscale));
                rho_add_lds(n0 + 1, rho_quant<TW>(w10, rscale));
                rho_add_lds(n0 + W, rho_quant<TW>(w01, rscale));
                rho_add_lds(n0 + W + 1, rho_quant<TW>(w11, rscale));
                continue;
            }
        }
        const int cx1 = cx + 1 == nx ? 0 : cx + 1, cy1 = cy + 1 == ny ? 0 : cy + 1;
        rho_add<TW>(rho + (size_t)cy * nx + cx, rho_quant<TW>(w00, rscale), lo);
        rho_add<TW>(rho + (size_t)cy * nx + cx1, rho_quant<TW>(w10, rscale), lo);
        rho_add<TW>(rho + (size_t)cy1 * nx + cx, rho_quant<TW>(w01, rscale), lo);
        rho_add<TW>(rho + (size_t)cy1 * nx + cx1, rho_quant<TW>(w11, rscale), lo);
    }
    __syncthreads();
    if (boxed)
        for (int k = threadIdx.x; k < W * H; k += 256) {
            const long long v = h[k];
            if (v == 0) continue;
            const int ky = k / W, kx = k - ky * W;
            int gx = b.x + kx, gy = b.z + ky;
            while (gx >= nx) gx -= nx;
            while (gy >= ny) gy -= ny;
            rho_add<TW>(rho + (size_t)gy * nx + gx, v, lo);
        }
}

// rho: the fixed-point charge of rho_cic_tiles_kernel, one word (lo = 0) or
// two (the low words lo entries after the high ones).
__global__ __launch_bounds__(256) void gauss_kernel(PushParams P, const long long *rho,
                                                    double *g0, int first, double rho_bg,
                                                    double cva2in0, double rscale, int64_t lo)
{
    __shared__ double red[2][4];
    __shared__ unsigned long long mx[2][4];
    const int nx = P.nx, ny = P.ny;
    const int64_t n = (int64_t)nx * ny;
    double l2 = 0, dl2 = 0, li = 0, dli = 0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256) {
        const int gy = (int)(k / nx), gx = (int)(k - (int64_t)gy * nx);
        // conducting y walls: the node rows strictly between the walls
        if (P.nyp && (gy == 0 || gy >= P.nyp)) continue;
        const int gxm = gx == 0 ? nx - 1 : gx - 1, gym = gy == 0 ? ny - 1 : gy - 1;
        const double div = (P.ex[k] - P.ex[(size_t)gy * nx + gxm]) * P.idx +
                           (P.ey[k] - P.ey[(size_t)gym * nx + gx]) * P.idy;
        const double rk = lo ? ((double)rho[k] + (double)rho[lo + k] * FX2_LO) / rscale
                             : (double)rho[k] / rscale;
        const double res = div - cva2in0 * (rk + rho_bg);
        if (first) g0[k] = res;
        const double d = res - g0[k];
        l2 += res * res;
        dl2 += d * d;
        li = fmax(li, fabs(res));
        dli = fmax(dli, fabs(d));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        l2 += __shfl_xor(l2, o);
        dl2 += __shfl_xor(dl2, o);
        li = fmax(li, __shfl_xor(li, o));
        dli = fmax(dli, __shfl_xor(dli, o));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = l2; red[1][w] = dl2;
        mx[0][w] = (unsigned long long)__double_as_longlong(li);
        mx[1][w] = (unsigned long long)__double_as_longlong(dli);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        P.part[blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        P.part[gridDim.x + blockIdx.x] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        unsigned long long a = mx[0][0], b = mx[1][0];
        for (int i = 1; i < 4; i++) { a = a > mx[0][i] ? a : mx[0][i]; b = b > mx[1][i] ? b : mx[1][i]; }
        atomicMax((unsigned long long *)(P.acc + ACC_G_LINF), a);
        atomicMax((unsigned long long *)(P.acc + ACC_G_DLINF), b);
    }
}

__global__ void zero_f64_kernel(double *p, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

hipError_t launch_gauss(const PushParams &p, int precision, double rho_scale,
                        double *rho_scratch, double *gauss0, int first, double rho_bg,
                        double cva2in0, hipStream_t st)
{
    const int64_t n = (int64_t)p.nx * p.ny;
    // FP64 form: two words per node (rho_scratch holds 2 n), scale 2^30 (a
    // weight, at most 1, times 2^30 2^FX2_LOBITS stays below 2^51)
    const bool tw = precision == PBPIC_PREC_FP64;
    const int64_t lo = tw ? n : 0;
    if (tw) rho_scale = 1073741824.0;
    long long *rho = reinterpret_cast<long long *>(rho_scratch);
    hipLaunchKernelGGL(zero_f64_kernel, dim3(1024), dim3(256), 0, st, rho_scratch, n + lo);
    for (int s = 0; s < p.nspecies; s++) {
        if (p.sp[s].n == 0) continue;
        const int nt = (int)((p.sp[s].n + PUSH_TILE - 1) / PUSH_TILE);
        if (tw)
            hipLaunchKernelGGL(rho_cic_tiles_kernel<double>, dim3(nt), dim3(256), 0, st, p,
                               p.sp[s], rho, rho_scale);
        else
            hipLaunchKernelGGL(rho_cic_tiles_kernel<float>, dim3(nt), dim3(256), 0, st, p,
                               p.sp[s], rho, rho_scale);
    }
    int nb = (int)((n + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(gauss_kernel, dim3(nb), dim3(256), 0, st, p, rho, gauss0,
                       first, rho_bg, cva2in0, rho_scale, lo);
    PartSlots ps{2, {ACC_G_L2, ACC_G_DL2, 0, 0}, {1.0, 1.0, 0, 0}};
    return launch_reduce_parts(p.part, nb, p.acc, ps, st);
}

// energyspectra (pic2d.f90:1957-1990): counts of particles by kinetic
// energy ene = (gamma - 1) ms in FP64 (:1966-1983), bin k (0-based) holding
// ceil(ene/de) = k + 1 for ene < emax.  ene = 0 (the reference's bin 0, out
// of its array) and a quotient rounded past the last bin go to the end bins.
// LDS histogram per workgroup, flushed with one global atomic per bin.
template <typename R>
__global__ __launch_bounds__(256) void spectrum_kernel(SpeciesDev S, int nbins, double de,
                                                       double emax, uint32_t *counts)
{
    __shared__ uint32_t h[SPEC_BINS_MAX];
    for (int k = threadIdx.x; k < nbins; k += 256) h[k] = 0u;
    __syncthreads();
    const R *ux = (const R *)S.ux, *uy = (const R *)S.uy, *uz = (const R *)S.uz;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < S.n;
         i += (int64_t)gridDim.x * 256) {
        const double a = ux[i], b = uy[i], c = uz[i];
        const double gama = sqrt(1.0 + a * a + b * b + c * c);
        const double ene = (gama - 1.0) * S.ms;
        if (!(ene < emax)) continue;
        double k = ceil(ene / de) - 1.0;
        k = k < 0.0 ? 0.0 : (k > nbins - 1 ? nbins - 1 : k);
        atomicAdd(&h[(int)k], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nbins; k += 256)
        if (h[k]) atomicAdd(counts + k, h[k]);
}

hipError_t launch_spectrum(const SpeciesDev &S, int precision, int nbins, double de, double emax,
                           uint32_t *counts, hipStream_t st)
{
    if (S.n == 0) return hipSuccess;
    int64_t nb = (S.n + 255) / 256;
    if (nb > 1024) nb = 1024;
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL(spectrum_kernel<double>, dim3((unsigned)nb), dim3(256), 0, st, S, nbins,
                           de, emax, counts);
    else
        hipLaunchKernelGGL(spectrum_kernel<float>, dim3((unsigned)nb), dim3(256), 0, st, S, nbins,
                           de, emax, counts);
    return hipGetLastError();
}

// divergenceclean (pic2d.f90:867-885) and suppresscerenkov (:890-907) as
// increments of E computed from E^n and rho^n before the field solve (the
// reference calls them after the E update, but both read only the dive and
// curl terms of E^n and the rho of this step's push), periodic throughout:
//   Marder:    F = div E^n - c^2 in0 rho_raw at the nodes,
//              dEx = dt 0.001/dx (F(x+1) - F(x)), dEy likewise in y
//   Cherenkov: w = dEy/dx - dEx/dy at the cell centres,
//              dEx = -cdy (w - w(y-1)), dEy = cdx (w - w(x-1)),
//              dEz = cdx (d2x Ez) + cdy (d2y Ez) with cdx = dt c^2 (dt/8)/dx
template <typename RJ>
__global__ __launch_bounds__(256) void filter_kernel(FilterParams P)
{
    const int nx = P.nx, ny = P.ny;
    const int64_t n = (int64_t)nx * ny;
    const RJ *rho = (const RJ *)P.rho;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256) {
        const int gy = (int)(k / nx), gx = (int)(k - (int64_t)gy * nx);
        const int xm = gx == 0 ? nx - 1 : gx - 1, xp = gx + 1 == nx ? 0 : gx + 1;
        const int ym = gy == 0 ? ny - 1 : gy - 1, yp = gy + 1 == ny ? 0 : gy + 1;
        auto at = [&](int x, int y) { return (size_t)y * nx + x; };
        const double *ex = P.ex, *ey = P.ey, *ez = P.ez;
        double dex = 0.0, dey = 0.0, dez = 0.0;
        if (P.marder) {
            auto mf = [&](int x, int y) {
                const int xl = x == 0 ? nx - 1 : x - 1, yl = y == 0 ? ny - 1 : y - 1;
                const double div = (ex[at(x, y)] - ex[at(xl, y)]) * P.idx +
                                   (ey[at(x, y)] - ey[at(x, yl)]) * P.idy;
                return div - P.cva2in0 * (double)rho[at(x, y)];
            };
            const double f0 = mf(gx, gy);
            dex = P.mdx * (mf(xp, gy) - f0);
            dey = P.mdy * (mf(gx, yp) - f0);
        }
        if (P.cherenkov) {
            auto w = [&](int x, int y) {
                const int xr = x + 1 == nx ? 0 : x + 1, yr = y + 1 == ny ? 0 : y + 1;
                return (ey[at(xr, y)] - ey[at(x, y)]) * P.idx - (ex[at(x, yr)] - ex[at(x, y)]) * P.idy;
            };
            const double w0 = w(gx, gy);
            dex += -(P.cdy * (w0 - w(gx, ym)));
            dey += P.cdx * (w0 - w(xm, gy));
            const double e0 = ez[at(gx, gy)];
            const double dzx0 = (e0 - ez[at(xm, gy)]) * P.idx, dzx1 = (ez[at(xp, gy)] - e0) * P.idx;
            const double dzy0 = (e0 - ez[at(gx, ym)]) * P.idy, dzy1 = (ez[at(gx, yp)] - e0) * P.idy;
            dez = P.cdx * (dzx1 - dzx0) + P.cdy * (dzy1 - dzy0);
        }
        P.fex[k] = dex;
        P.fey[k] = dey;
        P.fez[k] = dez;
    }
}

hipError_t launch_filters(const FilterParams &p, int precision, hipStream_t st)
{
    const int64_t n = (int64_t)p.nx * p.ny;
    int nb = (int)((n + 255) / 256);
    if (nb > 8192) nb = 8192;
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL(filter_kernel<double>, dim3(nb), dim3(256), 0, st, p);
    else
        hipLaunchKernelGGL(filter_kernel<float>, dim3(nb), dim3(256), 0, st, p);
    return hipGetLastError();
}

// ------------------------------------------------------ conducting y walls
// Guard rows of the field planes: half-row components (Ey, Bz even; Ez, By
// odd) of row nyp mirror row nyp-1 and of row nyp+1 (= row -1) mirror row 0;
// Ex is zero on the wall rows 0 and nyp; Ex, Bx of row nyp+1 (y = -dy) mirror
// row 1 (odd, even).
__global__ void wall_images_kernel(FdtdParams P, int mask)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    if (gx >= P.nx) return;
    const int nx = P.nx, nyp = P.nyp;
    const size_t r0 = gx, r1 = (size_t)nx + gx, rt = (size_t)(nyp - 1) * nx + gx,
                 gt = (size_t)nyp * nx + gx, gb = (size_t)(nyp + 1) * nx + gx;
    double *half[4] = {P.ey, P.ez, P.by, P.bz};
    const int hbit[4] = {1, 2, 4, 5};
    const double hsgn[4] = {1.0, -1.0, -1.0, 1.0};
    for (int k = 0; k < 4; k++) {
        if (!(mask >> hbit[k] & 1)) continue;
        double *a = half[k];
        a[gt] = hsgn[k] * a[rt];
        a[gb] = hsgn[k] * a[r0];
    }
    if (mask & 1) {
        P.ex[r0] = 0.0;
        P.ex[gt] = 0.0;
        P.ex[gb] = -P.ex[r1];
    }
    if (mask >> 3 & 1) P.bx[gb] = P.bx[r1];
}

hipError_t launch_wall_images(const FdtdParams &p, int mask, hipStream_t st)
{
    hipLaunchKernelGGL(wall_images_kernel, dim3((p.nx + 255) / 256), dim3(256), 0, st, p, mask);
    return hipGetLastError();
}

// ---------------------------------------------------------------- utilities
__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        p[k] = v;
}

hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    int64_t nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)nb), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

// Several device-to-device copies in one launch (field planes, current
// planes, particle arrays of a state): each workgroup moves 16-byte chunks
// of one copy, the copies laid end to end over the grid.
__global__ __launch_bounds__(256) void copy_list_kernel(CopyList L)
{
    const int64_t g = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    int k = 0;
    while (k + 1 < L.n && g >= L.first[k + 1]) k++;
    const int64_t i = g - L.first[k];
    if (i < (int64_t)(L.bytes[k] >> 4))
        reinterpret_cast<uint4 *>(L.dst[k])[i] = reinterpret_cast<const uint4 *>(L.src[k])[i];
}

hipError_t launch_copy_list(CopyList L, hipStream_t st)
{
    // chunk offsets of the copies that are whole 16-byte chunks on 16-byte
    // boundaries; the others go through hipMemcpyAsync
    CopyList K{};
    int64_t tot = 0;
    for (int k = 0; k < L.n; k++) {
        if (L.bytes[k] == 0) continue;
        const bool ok = L.bytes[k] % 16 == 0 && ((uintptr_t)L.src[k] | (uintptr_t)L.dst[k]) % 16 == 0;
        if (!ok) {
            hipError_t e = hipMemcpyAsync(L.dst[k], L.src[k], L.bytes[k], hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) return e;
            continue;
        }
        K.src[K.n] = L.src[k]; K.dst[K.n] = L.dst[k]; K.bytes[K.n] = L.bytes[k];
        K.first[K.n] = tot;
        tot += (int64_t)((L.bytes[k] >> 4) + 255) / 256 * 256;   // whole workgroups per copy
        K.n++;
    }
    if (K.n == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_list_kernel, dim3((unsigned)(tot / 256)), dim3(256), 0, st, K);
    return hipGetLastError();
}

// Split form: the FP32 current planes from the fixed-point sums the push
// accumulated (one rounding of an exact sum, the same for any order of the
// contributions).
// With conducting y walls (nyp > 0) Jz is folded on the way: the image of
// charge q at -y carries -q v_z, so the Jz a particle deposits beyond a wall
// folds back with sign -1 into the domain row it mirrors.  The guard rows nyp
// and nyp + 1 are subtracted, as integers, from the domain rows nyp - 1 and 0
// they mirror, and read as zero.
__global__ void finalize_current_kernel(const long long *a, const long long *b,
                                        const long long *c, const long long *d, float *ja,
                                        float *jb, float *jc, float *jd, int64_t n, float4 inv,
                                        int64_t nx, int64_t nyp)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        ja[k] = (float)((double)a[k] * (double)inv.x);
        jb[k] = (float)((double)b[k] * (double)inv.y);
        long long z = c[k];
        if (nyp > 0) {
            if (k < nx) z -= c[k + (nyp + 1) * nx];                        // row 0 - row -1
            else if (k >= (nyp - 1) * nx && k < nyp * nx) z -= c[k + nx];   // row nyp-1 - row nyp
            else if (k >= nyp * nx) z = 0;                                  // the guard rows
        }
        jc[k] = (float)((double)z * (double)inv.z);
        if (d) jd[k] = (float)((double)d[k] * (double)inv.w);
    }
}

// FP64 form: each plane's sums in two words (fx2_add), hi at k and lo at
// n + k, the value (hi + lo 2^-FX2_LOBITS) / scale.  Both conversions to
// double are exact (below 2^53 for fewer than 2^23 full-size contributions a
// node), their sum is one rounding, the scale a power of two: the result
// depends on the integer sums only.  Walls: the guard rows' sums are
// subtracted as integers, as above (the low word of a difference may leave
// [0, 2^FX2_LOBITS); the value is the same).
__global__ void finalize_current2_kernel(const long long *a, const long long *b,
                                         const long long *c, const long long *d, double *ja,
                                         double *jb, double *jc, double *jd, int64_t n,
                                         double4 inv, int64_t nx, int64_t nyp)
{
    constexpr double L = FX2_LO;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        ja[k] = ((double)a[k] + (double)a[n + k] * L) * inv.x;
        jb[k] = ((double)b[k] + (double)b[n + k] * L) * inv.y;
        long long zh = c[k], zl = c[n + k];
        if (nyp > 0) {
            int64_t m = -1;
            if (k < nx) m = k + (nyp + 1) * nx;                              // row 0 - row -1
            else if (k >= (nyp - 1) * nx && k < nyp * nx) m = k + nx;         // row nyp-1 - row nyp
            else if (k >= nyp * nx) zh = zl = 0;                             // the guard rows
            if (m >= 0) { zh -= c[m]; zl -= c[n + m]; }
        }
        jc[k] = ((double)zh + (double)zl * L) * inv.z;
        if (d) jd[k] = ((double)d[k] + (double)d[n + k] * L) * inv.w;
    }
}

hipError_t launch_finalize_current(void *const *j64, void *const *j, int64_t ncell,
                                   const float *scale, int with_rho, int nx, int nyp,
                                   int elem_bytes, hipStream_t st)
{
    int64_t nb = (ncell + 255) / 256;
    if (nb > 8192) nb = 8192;
    if (nb < 1) nb = 1;
    if (elem_bytes == 8) {
        const double4 inv = make_double4(1.0 / scale[0], 1.0 / scale[1], 1.0 / scale[2],
                                         1.0 / scale[3]);
        hipLaunchKernelGGL(finalize_current2_kernel, dim3((unsigned)nb), dim3(256), 0, st,
                           (const long long *)j64[0], (const long long *)j64[1],
                           (const long long *)j64[2],
                           with_rho ? (const long long *)j64[3] : nullptr, (double *)j[0],
                           (double *)j[1], (double *)j[2], (double *)j[3], ncell, inv,
                           (int64_t)nx, (int64_t)nyp);
        return hipGetLastError();
    }
    // 1/scale is a power of two: exact in float
    const float4 inv = make_float4(1.0f / scale[0], 1.0f / scale[1], 1.0f / scale[2],
                                   1.0f / scale[3]);
    hipLaunchKernelGGL(finalize_current_kernel, dim3((unsigned)nb), dim3(256), 0, st,
                       (const long long *)j64[0], (const long long *)j64[1],
                       (const long long *)j64[2], with_rho ? (const long long *)j64[3] : nullptr,
                       (float *)j[0], (float *)j[1], (float *)j[2], (float *)j[3], ncell, inv,
                       (int64_t)nx, (int64_t)nyp);
    return hipGetLastError();
}

// denjr = 0 (pic2d.f90:173-176); optionally resets the kinetic-energy tally
// (totene(1:2,itime) starts from zero each step).
template <typename R>
__global__ void zero_current_kernel(R *a, R *b, R *c, R *d, int64_t n, double *acc, int reset)
{
    const int64_t k0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k0 == 0 && reset) { acc[ACC_KE_E] = 0; acc[ACC_KE_I] = 0; }
    for (int64_t k = k0; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        a[k] = (R)0; b[k] = (R)0; c[k] = (R)0;
        if (d) d[k] = (R)0;
    }
}

hipError_t launch_zero_current(void *jx, void *jy, void *jz, void *rho, int64_t ncell,
                               int elem_bytes, double *acc, int reset_ke, hipStream_t st)
{
    int64_t nb = (ncell + 255) / 256;
    if (nb > 8192) nb = 8192;
    if (nb < 1) nb = 1;
    if (elem_bytes == 8)
        hipLaunchKernelGGL(zero_current_kernel<double>, dim3((unsigned)nb), dim3(256), 0, st,
                           (double *)jx, (double *)jy, (double *)jz, (double *)rho, ncell, acc, reset_ke);
    else
        hipLaunchKernelGGL(zero_current_kernel<float>, dim3((unsigned)nb), dim3(256), 0, st,
                           (float *)jx, (float *)jy, (float *)jz, (float *)rho, ncell, acc, reset_ke);
    return hipGetLastError();
}

// Accumulator maintenance.  op 0: reset the kinetic-energy tally.  op 1: fold
// the last J.E sum into its time integral (scale a = dt*mime/cva^2) and clear
// it.  op 2: clear the energy and Gauss slots.
__global__ void acc_op_kernel(double *acc, int op, double a)
{
    if (threadIdx.x != 0) return;
    if (op == 0) { acc[ACC_KE_E] = 0; acc[ACC_KE_I] = 0; }
    if (op == 1) {
        acc[ACC_JE_INT] += acc[ACC_JE_STEP] * a; acc[ACC_JE_STEP] = 0;
        acc[ACC_JEM_INT] += acc[ACC_JEM_STEP] * a; acc[ACC_JEM_STEP] = 0;
    }
    if (op == 2) {
        acc[ACC_WE] = 0; acc[ACC_WB] = 0; acc[ACC_G_L2] = 0; acc[ACC_G_LINF] = 0;
        acc[ACC_G_DL2] = 0; acc[ACC_G_DLINF] = 0;
    }
}

hipError_t launch_acc_op(double *acc, int op, double a, hipStream_t st)
{
    hipLaunchKernelGGL(acc_op_kernel, dim3(1), dim3(64), 0, st, acc, op, a);
    return hipGetLastError();
}

// Seeded Maxwellian loader: Philox4x32-10 (Salmon et al., SC'11) counter
// (p, stream, 0|1) under key seed, the same words as the oracle's loader.
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1, uint32_t o[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

__device__ __forceinline__ double u01(uint32_t u) { return ((double)u + 0.5) * 2.3283064365386963e-10; }

template <typename R>
__global__ void init_maxwellian_kernel(SpeciesDev S, int ppc, double uth, double ud,
                                       int counter_stream, uint32_t k0, uint32_t k1,
                                       uint32_t stream, int nyp, int ny)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S.n) return;
    uint32_t a[4], b[4];
    philox((uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 0u, k0, k1, a);
    philox((uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 1u, k0, k1, b);
    const double twopi = 6.283185307179586;
    const double r1 = sqrt(-2.0 * log(u01(a[2]))), t1 = twopi * u01(a[3]);
    const double r2 = sqrt(-2.0 * log(u01(b[0]))), t2 = twopi * u01(b[1]);
    const double drift = (counter_stream && ((p % ppc) & 1)) ? -ud : ud;
    // cells of the domain rows [0, nyp) in order (nyp < ny: conducting walls)
    const int64_t c = p / ppc;
    S.cell[p] = (int32_t)(nyp == ny ? c : (c / nyp) * ny + c % nyp);
    ((R *)S.ox)[p] = (R)(u01(a[0]) - 0.5);
    ((R *)S.oy)[p] = (R)(u01(a[1]) - 0.5);
    ((R *)S.ux)[p] = (R)(uth * (r1 * cos(t1)) + drift);
    ((R *)S.uy)[p] = (R)(uth * (r1 * sin(t1)));
    ((R *)S.uz)[p] = (R)(uth * (r2 * cos(t2)));
    if (S.pid) S.pid[p] = (int32_t)(p + 1);
}

hipError_t launch_init_maxwellian(const SpeciesDev &s, int precision, int nx, int ny, int ppc,
                                  double vth, double vdrift, int counter_stream, uint64_t seed,
                                  uint32_t stream, int nyp, hipStream_t st)
{
    (void)nx;
    if (s.n <= 0) return hipSuccess;
    const double uth = vth / sqrt(1.0 - vth * vth);
    const double ud = vdrift / sqrt(1.0 - vdrift * vdrift);
    const int64_t nb = (s.n + 255) / 256;
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL(init_maxwellian_kernel<double>, dim3((unsigned)nb), dim3(256), 0, st,
                           s, ppc, uth, ud, counter_stream, (uint32_t)seed,
                           (uint32_t)(seed >> 32), stream, nyp, ny);
    else
        hipLaunchKernelGGL(init_maxwellian_kernel<float>, dim3((unsigned)nb), dim3(256), 0, st,
                           s, ppc, uth, ud, counter_stream, (uint32_t)seed,
                           (uint32_t)(seed >> 32), stream, nyp, ny);
    return hipGetLastError();
}

// Harris current-sheet loader (initparticles, pic2d.f90:357-578): ion p and
// electron p of species `ion` / `ele` share their cell and their offset across
// the sheet.  Words: counter (p, stream, block) under key (k0, k1); block 0 =
// (ion position along the sheet, electron offset along it, ion Box-Muller pair
// for u_x, u_y), block 1 = (electron pair, pair for the thermal u_z of ion and
// electron), blocks 2..9 = four tries each of the position across the sheet,
// the inverse CDF of sech^2 (:395-402; a background particle takes the first
// word as a uniform position).
__device__ __forceinline__ void harris_cell(double X, int n, int &c, double &off)
{
    c = (int)ceil(X) - 1;                           // ix = ceiling(x/dx) (:403)
    c = c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
    off = X - ((double)c + 0.5);                    // :420
}

template <typename R>
__global__ void init_harris_kernel(SpeciesDev I, SpeciesDev E, int ny, HarrisDev h)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= I.n) return;
    const uint32_t plo = (uint32_t)p, phi = (uint32_t)((uint64_t)p >> 32);
    uint32_t w0[4], w1[4];
    philox(plo, phi, h.stream, 0u, h.k0, h.k1, w0);
    philox(plo, phi, h.stream, 1u, h.k0, h.k1, w1);
    const bool sheet = p < h.ncs;
    const bool second = sheet && h.geometry == PBPIC_HARRIS_DOUBLE_X && p >= h.ncs / 2;
    double Xn;
    if (sheet) {
        const double xc = second ? h.xc1 : h.xc0;
        Xn = -1.0;
        for (int b = 0; b < 8 && Xn < 0; b++) {
            uint32_t w[4];
            philox(plo, phi, h.stream, 2u + b, h.k0, h.k1, w);
            for (int t = 0; t < 4; t++) {
                const double u = u01(w[t]);
                const double s = h.L * log(u / (1.0 - u)) / 2.0;         // :401
                const double X = (s + xc) / h.dn;
                if (ceil(X) - 1 >= 0 && ceil(X) - 1 < h.nn) { Xn = X; break; }
            }
        }
        if (Xn < 0) Xn = xc / h.dn;
    } else {
        uint32_t w[4];
        philox(plo, phi, h.stream, 2u, h.k0, h.k1, w);
        Xn = u01(w[0]) * h.nn;
    }
    int cn, ct;
    double on, ot;
    harris_cell(Xn, h.nn, cn, on);
    harris_cell(u01(w0[0]) * h.nt, h.nt, ct, ot);
    const double ote = u01(w0[1]) - 0.5;                                 // :425
    const double twopi = 6.283185307179586;
    const double r1 = sqrt(-2.0 * log(u01(w0[2]))), t1 = twopi * u01(w0[3]);
    const double r2 = sqrt(-2.0 * log(u01(w1[0]))), t2 = twopi * u01(w1[1]);
    const double r3 = sqrt(-2.0 * log(u01(w1[2]))), t3 = twopi * u01(w1[3]);
    const double ui = sheet ? h.uic : h.uib, ue = sheet ? h.uec : h.ueb;
    double dzi = sheet ? h.udi : 0.0, dze = sheet ? h.ude : 0.0;
    if (second) { dzi = -dzi; dze = -dze; }                             // :491-497
    int cx, cy;
    double oxi, oyi, oxe, oye;
    if (h.geometry == PBPIC_HARRIS_DOUBLE_X) { cx = cn; cy = ct; oxi = oxe = on; oyi = ot; oye = ote; }
    else { cx = ct; cy = cn; oyi = oye = on; oxi = ot; oxe = ote; }
    const int32_t cell = (int32_t)((int64_t)cx * ny + cy);
    I.cell[p] = cell;
    E.cell[p] = cell;
    ((R *)I.ox)[p] = (R)oxi; ((R *)I.oy)[p] = (R)oyi;
    ((R *)E.ox)[p] = (R)oxe; ((R *)E.oy)[p] = (R)oye;
    ((R *)I.ux)[p] = (R)(ui * (r1 * cos(t1)));
    ((R *)I.uy)[p] = (R)(ui * (r1 * sin(t1)));
    ((R *)E.ux)[p] = (R)(ue * (r2 * cos(t2)));
    ((R *)E.uy)[p] = (R)(ue * (r2 * sin(t2)));
    ((R *)I.uz)[p] = (R)(dzi + (h.thermal_z ? ui * (r3 * cos(t3)) : 0.0));
    ((R *)E.uz)[p] = (R)(dze + (h.thermal_z ? ue * (r3 * sin(t3)) : 0.0));
    if (I.pid) I.pid[p] = (int32_t)(p + 1);
    if (E.pid) E.pid[p] = (int32_t)(p + 1);
}

hipError_t launch_init_harris(const SpeciesDev &ion, const SpeciesDev &ele, int precision,
                              int ny, const HarrisDev &h, hipStream_t st)
{
    if (ion.n <= 0) return hipSuccess;
    const int64_t nb = (ion.n + 255) / 256;
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL(init_harris_kernel<double>, dim3((unsigned)nb), dim3(256), 0, st, ion,
                           ele, ny, h);
    else
        hipLaunchKernelGGL(init_harris_kernel<float>, dim3((unsigned)nb), dim3(256), 0, st, ion,
                           ele, ny, h);
    return hipGetLastError();
}

// Cell ids in range, and (conducting y walls, nyp < ny) outside the guard rows.
__global__ void check_cells_kernel(const int32_t *cell, int64_t n, int64_t ncell, int ny,
                                   int nyp, int *err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cell[i];
    if (c < 0 || c >= ncell || (nyp < ny && c % ny >= nyp)) atomicOr(err, ERRF_CELL);
}

hipError_t launch_check_cells(const int32_t *cell, int64_t n, int64_t ncell, int ny, int nyp,
                              int *err, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(check_cells_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       cell, n, ncell, ny, nyp, err);
    return hipGetLastError();
}

}  // namespace pbpic
