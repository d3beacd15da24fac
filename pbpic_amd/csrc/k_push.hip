// k_push.hip -- fused gather + Boris push + move + charge-conserving deposit
// for gfx950 (the `pusher` loop body of pic2d.f90:934-1062 with the
// `currentdeposit` call of :1029 and the wrap/fold/halo of :1032-1216).
//
// Work decomposition.  Particles are kept sorted by cell (countingsort), so a
// contiguous tile of PUSH_TILE particles covers a compact patch of cells.  A
// block of 256 threads takes one tile:
//   A. loads the tile's cell ids and reduces their periodic-unwrapped bounding
//      box;
//   B. stages the node-averaged E and B of that box (+1 node) into LDS and
//      zeroes an LDS current window (box grown by 1 cell below and 2 above,
//      which holds every deposit target of a particle that moves <= 1 cell);
//   C. pushes every particle, gathering from LDS and depositing with LDS
//      atomics;
//   D. flushes the LDS window to the global current with one atomic per
//      non-zero cell value.
// Tiles whose box does not fit the LDS window (unsorted input) fall back to
// global gathers and global atomics.
//
// Lane mapping: the 256 threads form 32 groups of 8 lanes; group G owns the
// run [G*128, G*128+128) of the tile and walks it 8 particles per round.  The
// 8 groups of a wave therefore sit 128 particles apart (usually in different
// cells), which keeps same-address LDS atomic conflicts at about 8-way, while
// each group still reads 8 consecutive elements (64 B in FP64) per load.
//
// Precision.  FP64 reproduces the reference arithmetic operation for
// operation (no FMA contraction), including the absolute-index form of the
// zigzag relay point (:1239-1252), so one step agrees with the oracle to
// rounding of the current sums.  SPLIT32 keeps int cell + float offset and
// evaluates the same scheme in offsets relative to the old cell.
#include "pbpic_impl.h"

#pragma clang fp contract(off)

namespace pbpic {

// Split form, global planes: the integer sum itself, or one contribution
// already multiplied by its plane's scale, rounded as the LDS window rounds
// it (floor(v + 1/2), v_cvt_rpi_i32_f32).
__device__ __forceinline__ void gfix(void *p, long long v)
{
    atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
}
__device__ __forceinline__ void gfixf(void *p, float v)
{
    int qi;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(qi) : "v"(v));
    gfix(p, (long long)qi);
}

// Node-averaged field at compact node (gx, gy): the emfgrid formulas of
// pic2d.f90:837-845, evaluated in FP64 in the reference's order.
__device__ __forceinline__ void node_fields(const PushParams &P, int gx, int gy,
                                            double f[6])
{
    const int nx = P.nx, ny = P.ny;
    const int gxm = gx == 0 ? nx - 1 : gx - 1;
    const int gym = gy == 0 ? ny - 1 : gy - 1;
    const size_t c = (size_t)gy * nx + gx, xm = (size_t)gy * nx + gxm;
    const size_t ym = (size_t)gym * nx + gx, xym = (size_t)gym * nx + gxm;
    f[0] = (P.ex[c] + P.ex[xm]) * 0.5;
    f[1] = (P.ey[c] + P.ey[ym]) * 0.5;
    f[2] = (P.ez[c] + P.ez[xm] + P.ez[ym] + P.ez[xym]) * 0.25;
    f[3] = (P.bx[c] + P.bx[xm]) * 0.5;
    f[4] = (P.by[c] + P.by[ym]) * 0.5;
    f[5] = (P.bz[c] + P.bz[xm] + P.bz[ym] + P.bz[xym]) * 0.25;
}

// Node-field array (emfgrid, pic2d.f90:837-845): the six node-averaged
// components of every node, rounded to the push precision R, stored y fastest
// ((gx*ny + gy) records of NS values) so that a push tile -- a run of
// consecutive cells cx*ny + cy -- stages its window from contiguous column
// runs.  One 32 x 32 node tile per workgroup: the six staggered planes (x
// fastest) are staged with their -x/-y halo in LDS, then written transposed.
constexpr int NF_T = 32;
template <typename R, int NS>
__global__ __launch_bounds__(256) void nodefield_kernel(PushParams P, R *out)
{
    __shared__ double t[6][NF_T + 1][NF_T + 1];     // [plane][gy - Y0 + 1][gx - X0 + 1]
    const int nx = P.nx, ny = P.ny;
    const int X0 = blockIdx.x * NF_T, Y0 = blockIdx.y * NF_T;
    // the push that follows appends deferred tiles to its lists: clear
    // their counts here (the previous push's listed launches have finished
    // reading them)
    if (P.big && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 3) P.big[threadIdx.x] = 0;
    const double *pl[6] = {P.ex, P.ey, P.ez, P.bx, P.by, P.bz};
    for (int k = threadIdx.x; k < (NF_T + 1) * (NF_T + 1); k += 256) {
        const int r = k / (NF_T + 1), cc = k - r * (NF_T + 1);
        const size_t g = (size_t)wrap_index(Y0 - 1 + r, ny) * nx + wrap_index(X0 - 1 + cc, nx);
#pragma unroll
        for (int q = 0; q < 6; q++) t[q][r][cc] = pl[q][g];
    }
    __syncthreads();
    const int j = threadIdx.x & (NF_T - 1);
    for (int i = threadIdx.x / NF_T; i < NF_T; i += 256 / NF_T) {
        const int gx = X0 + i, gy = Y0 + j;
        if (gx >= nx || gy >= ny) continue;
        // the same sums, in the same order, as node_fields()
        R f[NS];
        f[0] = (R)((t[0][j + 1][i + 1] + t[0][j + 1][i]) * 0.5);
        f[1] = (R)((t[1][j + 1][i + 1] + t[1][j][i + 1]) * 0.5);
        f[2] = (R)((t[2][j + 1][i + 1] + t[2][j + 1][i] + t[2][j][i + 1] + t[2][j][i]) * 0.25);
        f[3] = (R)((t[3][j + 1][i + 1] + t[3][j + 1][i]) * 0.5);
        f[4] = (R)((t[4][j + 1][i + 1] + t[4][j][i + 1]) * 0.5);
        f[5] = (R)((t[5][j + 1][i + 1] + t[5][j + 1][i] + t[5][j][i + 1] + t[5][j][i]) * 0.25);
#pragma unroll
        for (int q = 6; q < NS; q++) f[q] = (R)0;
        R *o = out + ((size_t)gx * ny + gy) * NS;
#pragma unroll
        for (int q = 0; q < NS; q += 16 / (int)sizeof(R)) {
            if constexpr (sizeof(R) == 8)
                *reinterpret_cast<double2 *>(o + q) = make_double2(f[q], f[q + 1]);
            else
                *reinterpret_cast<float4 *>(o + q) = make_float4(f[q], f[q + 1], f[q + 2], f[q + 3]);
        }
    }
}

// Deposit contributions of one particle, relative to its old cell (0,0).
template <typename R>
struct Dep {
    R jxa0, jxa1, jya0, jya1;  // segment 1: jx(0,0),(0,1); jy(0,0),(1,0)
    R jxb0, jxb1, jyb0, jyb1;  // segment 2: jx(a,b),(a,b+1); jy(a,b),(a+1,b)
    R jz0, jz1, jz2, jz3;      // jz(0,0),(di,0),(0,dj),(di,dj)
    R r0, r1, r2, r3;          // rho(0,0),(1,0),(0,1),(1,1)
    int a, b, di, dj;
};

__device__ __forceinline__ double sgn1(double x) { return signbit(x) ? -1.0 : 1.0; }
__device__ __forceinline__ float sgn1(float x) { return signbit(x) ? -1.0f : 1.0f; }
template <typename R> __device__ __forceinline__ R rmin(R a, R b) { return b < a ? b : a; }
template <typename R> __device__ __forceinline__ R rmax(R a, R b) { return b > a ? b : a; }

// currentdeposit (pic2d.f90:1226-1296).  i1, j1 = old 1-based cell, a, b =
// cell displacement (i2 = i1 + a).  SPLIT: offsets relative to i1, j1.
template <typename R, bool SPLIT>
__device__ __forceinline__ void deposit_terms(int i1, int j1, int a, int b,
                                              R dx1, R dx2, R dy1, R dy2, R qx,
                                              R qy, R qr, R vz, int compat,
                                              bool want_rho, Dep<R> &d)
{
    const R H = (R)0.5, ONE = (R)1;
    R irr, jrr, fx1, fy1, fx2, fy2, wx1, wy1, wx2, wy2, dimid, djmid;
    if (!SPLIT) {
        const int i2 = i1 + a, j2 = j1 + b;
        const R im = ((R)(i1 + i2) + dx1 + dx2 + ONE) * H;                 // :1239
        const R jm = ((R)(j1 + j2) + dy1 + dy2 + ONE) * H;
        const R ir = rmin((R)(i1 < i2 ? i1 : i2) + ONE,
                          rmax((R)(i1 > i2 ? i1 : i2) + (R)0, im));       // :1241
        const R jr = rmin((R)(j1 < j2 ? j1 : j2) + ONE,
                          rmax((R)(j1 > j2 ? j1 : j2) + (R)0, jm));
        fx1 = qx * (ir - (R)i1 - dx1 - H);                                   // :1244
        fy1 = qy * (jr - (R)j1 - dy1 - H);
        fx2 = qx * ((R)i2 + dx2 + H - ir);
        fy2 = qy * ((R)j2 + dy2 + H - jr);
        wx1 = (ir - (R)i1 + dx1 + H) * H;                                    // :1249
        wy1 = (jr - (R)j1 + dy1 + H) * H;
        wx2 = (ir - (R)i2 + dx2 + H) * H;
        wy2 = (jr - (R)j2 + dy2 + H) * H;
        dimid = im - (R)i1 - H;                                              // :1263
        djmid = jm - (R)j1 - H;
    } else {
        const R af = (R)a, bf = (R)b;
        const R imr = (af + dx1 + dx2 + ONE) * H;        // im - i1
        const R jmr = (bf + dy1 + dy2 + ONE) * H;
        // v_min/v_max: the same values for every finite input
        irr = fmin(fmin(af, (R)0) + ONE, fmax(fmax(af, (R)0), imr));
        jrr = fmin(fmin(bf, (R)0) + ONE, fmax(fmax(bf, (R)0), jmr));
        fx1 = qx * (irr - dx1 - H);
        fy1 = qy * (jrr - dy1 - H);
        fx2 = qx * (af + dx2 + H - irr);
        fy2 = qy * (bf + dy2 + H - jrr);
        wx1 = (irr + dx1 + H) * H;
        wy1 = (jrr + dy1 + H) * H;
        wx2 = (irr - af + dx2 + H) * H;
        wy2 = (jrr - bf + dy2 + H) * H;
        dimid = imr - H;
        djmid = jmr - H;
    }
    d.a = a;
    d.b = b;
    d.jxa0 = fx1 * (ONE - wy1);                                              // :1254
    d.jxa1 = fx1 * wy1;
    d.jya0 = fy1 * (ONE - wx1);
    d.jya1 = fy1 * wx1;
    d.jxb0 = fx2 * (ONE - wy2);
    d.jxb1 = fx2 * wy2;
    d.jyb0 = fy2 * (ONE - wx2);
    d.jyb1 = fy2 * wx2;
    d.di = (int)sgn1(dimid);                                                 // :1266
    d.dj = (int)sgn1(djmid);
    dimid = fabs(dimid);
    djmid = fabs(djmid);
    R s1 = dimid * djmid;                                                    // :1272
    R s2 = dimid - s1;
    R s3 = djmid - s1;
    R s4 = ONE - s1 - s2 - s3;
    const R quz = qr * vz;                                                   // :1277
    d.jz0 = quz * s4;
    d.jz1 = quz * s2;
    d.jz2 = quz * s3;
    d.jz3 = quz * s1;
    if (want_rho) {
        if (compat & PBPIC_COMPAT_RHO_MIRROR) {                              // :1285
            s1 = (H - dx1) * (H - dy1);
            s2 = H - dx1 - s1;
            s3 = H - dy1 - s1;
        } else {                      // the gather's CIC weights (:948-951)
            s1 = (H + dx1) * (H + dy1);
            s2 = H + dx1 - s1;
            s3 = H + dy1 - s1;
        }
        s4 = ONE - s1 - s2 - s3;
        d.r0 = s4 * qr;
        d.r1 = s2 * qr;
        d.r2 = s3 * qr;
        d.r3 = s1 * qr;
    }
}

// Gather + Boris + move for one particle (pic2d.f90:948-1028).  n00..n11 are
// the 6-component node fields at (ix,iy), (ix+1,iy), (ix,iy+1), (ix+1,iy+1).
template <typename R, bool OLDU>
__device__ __forceinline__ void push_math(const PushParams &P, R c0, R c1,
                                          R &ox, R &oy, R &ux, R &uy, R &uz,
                                          const R *n00, const R *n10,
                                          const R *n01, const R *n11, int &dcx,
                                          int &dcy, R &vz, R &gam, int &err)
{
    const R H = (R)0.5, ONE = (R)1;
    const R s1 = (H + ox) * (H + oy);                                        // :948
    const R s2 = H + ox - s1;
    const R s3 = H + oy - s1;
    const R s4 = ONE - s1 - s2 - s3;
    R f[6];
#pragma unroll
    for (int k = 0; k < 6; k++)
        f[k] = n00[k] * s4 + n10[k] * s2 + n01[k] * s3 + n11[k] * s1;        // :952
    const R umx = ux + c1 * f[0];                                            // :971
    const R umy = uy + c1 * f[1];
    const R umz = uz + c1 * f[2];
    R gama = sqrt(ONE + umx * umx + umy * umy + umz * umz);
    const R c2 = c0 / gama;                                                  // :976
    const R tx = c2 * f[3], ty = c2 * f[4], tz = c2 * f[5];
    const R t2 = (R)2 / (ONE + tx * tx + ty * ty + tz * tz);
    const R sx = t2 * tx, sy = t2 * ty, sz = t2 * tz;
    R u0x, u0y, u0z;
    if (OLDU) {                                                              // :985
        u0x = umx + (uy * tz - uz * ty);
        u0y = umy + (uz * tx - ux * tz);
        u0z = umz + (ux * ty - uy * tx);
    } else {
        u0x = umx + (umy * tz - umz * ty);
        u0y = umy + (umz * tx - umx * tz);
        u0z = umz + (umx * ty - umy * tx);
    }
    const R upx = umx + (u0y * sz - u0z * sy);                               // :989
    const R upy = umy + (u0z * sx - u0x * sz);
    const R upz = umz + (u0x * sy - u0y * sx);
    ux = upx + c1 * f[0];                                                    // :993
    uy = upy + c1 * f[1];
    uz = upz + c1 * f[2];
    gama = sqrt(ONE + ux * ux + uy * uy + uz * uz);
    const R c3 = (R)P.dtcva / gama;                                          // :1003
    const R dx1 = ox + c3 * ux * (R)P.idx;
    const R dy1 = oy + c3 * uy * (R)P.idy;
    if (fabs(dx1) > H) {                                                     // :1008
        dcx = (int)sgn1(dx1);
        ox = dx1 + sgn1(-dx1);
    } else {
        dcx = 0;
        ox = dx1;
    }
    if (fabs(dy1) > H) {                                                     // :1017
        dcy = (int)sgn1(dy1);
        oy = dy1 + sgn1(-dy1);
    } else {
        dcy = 0;
        oy = dy1;
    }
    if (fabs(ox) > H || fabs(oy) > H) err |= ERRF_CFL;
    if (!isfinite(gama)) err |= ERRF_NONFINITE;
    vz = uz * (R)P.cva / gama;                                               // :1028
    gam = gama;
}

// Conducting y walls (nyp > 0): a particle that would leave the domain rows
// [0, nyp) is reflected specularly before its deposit -- the offset mirrored
// into the cell it left from, u_y reversed.  Periodic runs pass nyp = 0.
template <typename R>
__device__ __forceinline__ void wall_reflect(int nyp, int cy, int &dcy, R &oy, R &uy)
{
    if (nyp && ((cy == 0 && dcy < 0) || (cy == nyp - 1 && dcy > 0))) {
        dcy = 0;
        oy = -oy;
        uy = -uy;
    }
}

// Node-field array record: 6 field components, padded to 8 floats in FP32
// so that every record starts 16-byte aligned (48-byte FP64 records already
// are).  The LDS window packs the 6 components (24 bytes in FP32): a third
// more nodes in the same LDS, and a node and its +y neighbour, adjacent in
// the y-fastest window, read as six 8-byte pairs.
template <typename R> struct NodeStride { static constexpr int v = 6; };
template <> struct NodeStride<float> { static constexpr int v = 8; };
constexpr int LDS_NS = 6;

// Split-precision form of push_math: the same scheme in FP32 with one
// reciprocal square root shared by every 1/gamma and contracted multiply-adds.
template <bool OLDU>
__device__ __forceinline__ void push_math_split(const PushParams &P, float c0, float c1,
                                                float &ox, float &oy, float &ux, float &uy,
                                                float &uz, const float *n00, const float *n10,
                                                const float *n01, const float *n11, int &dcx,
                                                int &dcy, float &vz, int &err)
{
#pragma clang fp contract(fast)
    const float wx = 0.5f + ox, wy = 0.5f + oy;
    const float s11 = wx * wy, s10 = wx - s11, s01 = wy - s11, s00 = 1.0f - wx - s01;
    float f[6];
#pragma unroll
    for (int k = 0; k < 6; k++) f[k] = n00[k] * s00 + n10[k] * s10 + n01[k] * s01 + n11[k] * s11;
    const float umx = ux + c1 * f[0], umy = uy + c1 * f[1], umz = uz + c1 * f[2];
    // 1 + u^2 >= 1: the raw reciprocal square root needs no denormal fix-up
    const float ig1 = __builtin_amdgcn_rsqf(1.0f + umx * umx + umy * umy + umz * umz);
    const float c2 = c0 * ig1;
    const float tx = c2 * f[3], ty = c2 * f[4], tz = c2 * f[5];
    const float t2 = 2.0f * __builtin_amdgcn_rcpf(1.0f + tx * tx + ty * ty + tz * tz);
    const float sx = t2 * tx, sy = t2 * ty, sz = t2 * tz;
    float qx = umx, qy = umy, qz = umz;
    if (OLDU) { qx = ux; qy = uy; qz = uz; }                                // :985
    const float u0x = umx + (qy * tz - qz * ty);
    const float u0y = umy + (qz * tx - qx * tz);
    const float u0z = umz + (qx * ty - qy * tx);
    ux = umx + (u0y * sz - u0z * sy) + c1 * f[0];
    uy = umy + (u0z * sx - u0x * sz) + c1 * f[1];
    uz = umz + (u0x * sy - u0y * sx) + c1 * f[2];
    const float u2 = ux * ux + uy * uy + uz * uz;
    const float ig = __builtin_amdgcn_rsqf(1.0f + u2);
    const float c3 = (float)P.dtcva * ig;
    const float dx1 = ox + c3 * ux * (float)P.idx;
    const float dy1 = oy + c3 * uy * (float)P.idy;
    dcx = dx1 > 0.5f ? 1 : (dx1 < -0.5f ? -1 : 0);
    dcy = dy1 > 0.5f ? 1 : (dy1 < -0.5f ? -1 : 0);
    ox = dx1 - (float)dcx;
    oy = dy1 - (float)dcy;
    if (fabsf(ox) > 0.5f || fabsf(oy) > 0.5f) err |= ERRF_CFL;
    if (!isfinite(u2)) err |= ERRF_NONFINITE;
    vz = uz * (float)P.cva * ig;
}

// Kinetic energy (gamma - 1) of one particle from its momentum, in the push
// precision: FP64 as the reference (:1062), FP32 as u^2/(gamma+1), which
// equals gamma-1 without the FP32 cancellation of a small gamma-1.
__device__ __forceinline__ double kin_energy(double ux, double uy, double uz)
{
    return sqrt(1.0 + ux * ux + uy * uy + uz * uz) - 1.0;
}
__device__ __forceinline__ float kin_energy(float ux, float uy, float uz)
{
#pragma clang fp contract(fast)
    const float u2 = ux * ux + uy * uy + uz * uz;
    const float gam = __builtin_amdgcn_rcpf(__builtin_amdgcn_rsqf(1.0f + u2));
    return u2 * __builtin_amdgcn_rcpf(gam + 1.0f);
}

// Copies the 6 components of one node record from the node-field array
// (16-byte aligned records) into the LDS window (LDS_NS values per node,
// 8-byte aligned).
template <typename R>
__device__ __forceinline__ void copy_node(R *dst, const R *src)
{
    if constexpr (sizeof(R) == 4) {
        const float4 a = *reinterpret_cast<const float4 *>(src);
        const float2 b = *reinterpret_cast<const float2 *>(src + 4);
        reinterpret_cast<float2 *>(dst)[0] = make_float2(a.x, a.y);
        reinterpret_cast<float2 *>(dst)[1] = make_float2(a.z, a.w);
        reinterpret_cast<float2 *>(dst)[2] = b;
    } else {
#pragma unroll
        for (int q = 0; q < 6; q += 2)
            *reinterpret_cast<double2 *>(dst + q) = *reinterpret_cast<const double2 *>(src + q);
    }
}

// Vector width of the particle streams: each lane moves 16 bytes per array
// per round (2 doubles or 4 floats); a wave-wide load covers 1 KiB.
template <typename R> struct Vw;
template <> struct Vw<double> { static constexpr int V = push_vec(PBPIC_PREC_FP64); };
template <> struct Vw<float> { static constexpr int V = push_vec(PBPIC_PREC_SPLIT32); };

__device__ __forceinline__ void vld(const double *p, int64_t i0, int64_t n, double (&o)[2])
{
    if (i0 + 2 <= n) { const double2 t = *reinterpret_cast<const double2 *>(p + i0); o[0] = t.x; o[1] = t.y; }
    else { o[0] = i0 < n ? p[i0] : 0.0; o[1] = 0.0; }
}
__device__ __forceinline__ void vld(const float *p, int64_t i0, int64_t n, float (&o)[4])
{
    if (i0 + 4 <= n) {
        const float4 t = *reinterpret_cast<const float4 *>(p + i0);
        o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) o[j] = i0 + j < n ? p[i0 + j] : 0.0f;
    }
}
__device__ __forceinline__ void vld(const int32_t *p, int64_t i0, int64_t n, int (&o)[2])
{
    if (i0 + 2 <= n) { const int2 t = *reinterpret_cast<const int2 *>(p + i0); o[0] = t.x; o[1] = t.y; }
    else { o[0] = i0 < n ? p[i0] : -1; o[1] = -1; }
}
__device__ __forceinline__ void vld(const int32_t *p, int64_t i0, int64_t n, int (&o)[4])
{
    if (i0 + 4 <= n) {
        const int4 t = *reinterpret_cast<const int4 *>(p + i0);
        o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) o[j] = i0 + j < n ? p[i0 + j] : -1;
    }
}
__device__ __forceinline__ void vst(double *p, int64_t i0, int64_t n, const double (&v)[2])
{
    if (i0 + 2 <= n) *reinterpret_cast<double2 *>(p + i0) = make_double2(v[0], v[1]);
    else if (i0 < n) p[i0] = v[0];
}
__device__ __forceinline__ void vst(float *p, int64_t i0, int64_t n, const float (&v)[4])
{
    if (i0 + 4 <= n) *reinterpret_cast<float4 *>(p + i0) = make_float4(v[0], v[1], v[2], v[3]);
    else {
#pragma unroll
        for (int j = 0; j < 4; j++) if (i0 + j < n) p[i0 + j] = v[j];
    }
}
__device__ __forceinline__ void vst(int32_t *p, int64_t i0, int64_t n, const int (&v)[2])
{
    if (i0 + 2 <= n) *reinterpret_cast<int2 *>(p + i0) = make_int2(v[0], v[1]);
    else if (i0 < n) p[i0] = v[0];
}
__device__ __forceinline__ void vst(int32_t *p, int64_t i0, int64_t n, const int (&v)[4])
{
    if (i0 + 4 <= n) *reinterpret_cast<int4 *>(p + i0) = make_int4(v[0], v[1], v[2], v[3]);
    else {
#pragma unroll
        for (int j = 0; j < 4; j++) if (i0 + j < n) p[i0 + j] = v[j];
    }
}

// Unchecked 16-byte loads and stores for full tiles (16-byte aligned).  The
// particle streams are read and written once per push and are loaded and
// stored non-temporally (measured on MI355X: a 3-4% faster push in both forms
// than plain loads and stores; non-temporal stores alone were slower).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef int nt_i4 __attribute__((ext_vector_type(4)));
typedef int nt_i2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void ldu(const double *p, double (&o)[2])
{
    const nt_d2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_d2 *>(p));
    o[0] = t.x; o[1] = t.y;
}
__device__ __forceinline__ void ldu(const float *p, float (&o)[4])
{
    const nt_f4 t = __builtin_nontemporal_load(reinterpret_cast<const nt_f4 *>(p));
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
}
__device__ __forceinline__ void ldu(const int32_t *p, int (&o)[2])
{
    const nt_i2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_i2 *>(p));
    o[0] = t.x; o[1] = t.y;
}
__device__ __forceinline__ void ldu(const int32_t *p, int (&o)[4])
{
    const nt_i4 t = __builtin_nontemporal_load(reinterpret_cast<const nt_i4 *>(p));
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
}
__device__ __forceinline__ void stu(double *p, const double (&v)[2])
{
    const nt_d2 t = {v[0], v[1]};
    __builtin_nontemporal_store(t, reinterpret_cast<nt_d2 *>(p));
}
__device__ __forceinline__ void stu(float *p, const float (&v)[4])
{
    const nt_f4 t = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(t, reinterpret_cast<nt_f4 *>(p));
}
__device__ __forceinline__ void stu(int32_t *p, const int (&v)[2])
{
    const nt_i2 t = {v[0], v[1]};
    __builtin_nontemporal_store(t, reinterpret_cast<nt_i2 *>(p));
}
__device__ __forceinline__ void stu(int32_t *p, const int (&v)[4])
{
    const nt_i4 t = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(t, reinterpret_cast<nt_i4 *>(p));
}

// The split form's four particles of a lane as four PartRec records (96
// contiguous bytes at slot i0, a multiple of 4): six 16-byte stores.  One
// store instruction covers 16 of every 96 bytes of the wave's 6 KiB, so the
// stores go through L2, which merges the lines (non-temporal ones ran the
// push at a third of its speed).
__device__ __forceinline__ void st_rec4(void *rec, int64_t i0, const int (&c)[4],
                                        const float (&v)[5][4])
{
    float w[24];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        w[6 * j] = __int_as_float(c[j]);
#pragma unroll
        for (int q = 0; q < 5; q++) w[6 * j + 1 + q] = v[q][j];
    }
    float4 *p = reinterpret_cast<float4 *>((PartRec *)rec + i0);
#pragma unroll
    for (int k = 0; k < 6; k++) p[k] = make_float4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}
// Bounds-checked form for the generic path.
__device__ __forceinline__ void vst_rec4(void *rec, int64_t i0, int64_t n, const int (&c)[4],
                                         const float (&v)[5][4])
{
    if (i0 + 4 <= n) {
        st_rec4(rec, i0, c, v);
        return;
    }
    PartRec *r = (PartRec *)rec;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (i0 + j >= n) break;
        r[i0 + j].c = c[j];
#pragma unroll
        for (int q = 0; q < 5; q++) r[i0 + j].a[q] = v[q][j];
    }
}

// One contribution to the LDS current window.  Split form: the contribution arrives pre-multiplied by its plane's scale
// (a power of two, so the product is exact) and below 2^30 in magnitude; it
// is rounded to the nearest integer and added as a 64-bit integer, which
// gfx950's LDS sustains at a higher rate than floating-point atomics and
// which makes the window sums exact and independent of the lane order.
// FP64 form: the contribution (pre-scaled by 2^FX2_LOBITS as well) rounded
// to an integer (fx2_quant); one tile puts at most 2 PUSH_TILE contributions
// below 2^49 on a node of a plane, so the window sums stay below 2^63.
__device__ __forceinline__ void jadd(double *p, double v)
{
    __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(p),
                           (unsigned long long)fx2_quant(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void jadd(double *p, float v)
{
    // v_cvt_rpi_i32_f32: floor(v + 0.5) in one instruction (the compiler
    // emits round-to-nearest-even plus convert for rintf)
    int qi;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(qi) : "v"(v));
    const long long q = (long long)qi;
    __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(p), (unsigned long long)q,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Adds one particle's deposit to the LDS current window (planes jx, jy, jz
// and optionally rho, WH nodes each, x fastest, row length W).
template <typename R>
__device__ __forceinline__ void deposit_lds(double *jw, int WH, int W, int base,
                                            const Dep<R> &d, bool want_rho)
{
    double *wx = jw, *wy = jw + WH, *wz = jw + 2 * WH;
    jadd(wx + base, d.jxa0);
    jadd(wx + base + W, d.jxa1);
    jadd(wy + base, d.jya0);
    jadd(wy + base + 1, d.jya1);
    if (d.a != 0 || d.b != 0) {
        const int bb = base + d.b * W + d.a;
        jadd(wx + bb, d.jxb0);
        jadd(wx + bb + W, d.jxb1);
        jadd(wy + bb, d.jyb0);
        jadd(wy + bb + 1, d.jyb1);
    }
    jadd(wz + base, d.jz0);
    jadd(wz + base + d.di, d.jz1);
    jadd(wz + base + d.dj * W, d.jz2);
    jadd(wz + base + d.dj * W + d.di, d.jz3);
    if (want_rho) {
        double *wr = jw + 3 * WH;
        jadd(wr + base, d.r0);
        jadd(wr + base + 1, d.r1);
        jadd(wr + base + W, d.r2);
        jadd(wr + base + W + 1, d.r3);
    }
}

// Split form, interior tiles: the jx/jy deposit of particles that cross a
// cell face is deferred to a per-wave LDS queue and done 64 at a time, so that
// the two-segment zigzag (pic2d.f90:1239-1261) runs on full waves; a particle
// that stays in its cell deposits the closed form of its two segments.
constexpr int XQ_N = 128;                    // entries per wave (a power of two)
constexpr size_t XQ_BYTES = (size_t)XQ_N * (2 * sizeof(float2) + sizeof(int));
struct CrossQueue {
    float2 *qx, *qy;                         // (old, new) offsets in x and in y
    int *qc;                                 // window node | (a + 1) << 16 | (b + 1) << 18
    int head, tail, lane;
    __device__ void init(unsigned char *base, int wave, int ln)
    {
        unsigned char *w = base + (size_t)wave * XQ_BYTES;
        qx = (float2 *)w;
        qy = qx + XQ_N;
        qc = (int *)(qy + XQ_N);
        head = tail = 0;
        lane = ln;
    }
    __device__ void push(bool cross, int node, int a, int b, float x1, float x2, float y1,
                         float y2)
    {
        const uint64_t m = __ballot(cross);
        if (cross) {
            const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            const int slot = (tail + rank) & (XQ_N - 1);
            qx[slot] = make_float2(x1, x2);
            qy[slot] = make_float2(y1, y2);
            qc[slot] = node | ((a + 1) << 16) | ((b + 1) << 18);
        }
        tail += __popcll(m);
    }
    // the k oldest entries (k <= 64), one per lane
    __device__ void drain(double *jw, int WH, int W, float qxs, float qys, int compat, int k)
    {
        if (k <= 0) return;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane < k) {
            const int slot = (head + lane) & (XQ_N - 1);
            const float2 x = qx[slot], y = qy[slot];
            const int pk = qc[slot];
            const int node = pk & 0xffff, a = ((pk >> 16) & 3) - 1, b = ((pk >> 18) & 3) - 1;
            Dep<float> d;
            deposit_terms<float, true>(0, 0, a, b, x.x, x.y, y.x, y.y, qxs, qys, 0.0f, 0.0f,
                                       compat, false, d);
            double *wx = jw, *wy = jw + WH;
            jadd(wx + node, d.jxa0);
            jadd(wx + node + W, d.jxa1);
            jadd(wy + node, d.jya0);
            jadd(wy + node + 1, d.jya1);
            const int bb = node + b * W + a;
            jadd(wx + bb, d.jxb0);
            jadd(wx + bb + W, d.jxb1);
            jadd(wy + bb, d.jyb0);
            jadd(wy + bb + 1, d.jyb1);
        }
        head += k;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
};

// LDS layout of a push workgroup: node-field window [maxw][LDS_NS], current
// window [3 or 4 planes][maxw] of 8-byte sums, then (split form) the waves'
// crossing queues.  Split-form offset of the queues:
__host__ __device__ constexpr size_t push_queue_offset(int maxw, int deposit_rho)
{
    return (size_t)maxw * (LDS_NS * 4 + (deposit_rho ? 4 : 3) * 8);
}

// One particle's deposit in the split form on an interior tile (contributions
// pre-scaled to the fixed-point window).  jz: the CIC weights of the move's
// midpoint about the old cell centre (:1263-1281), for every particle.  jx,
// jy: a move inside one cell has its zigzag relay point at the midpoint, and
// the sum of its two segments is q dx/dt * Dx * (1/2 -+ ym) on jx(0,0) and
// jx(0,1) and q dy/dt * Dy * (1/2 -+ xm) on jy(0,0) and jy(1,0) (xm, ym the
// midpoint offsets); a move across a face goes to the queue.
__device__ __forceinline__ void deposit_split_fast(double *jw, int WH, int W, int base, int a,
                                                   int b, float dx1, float dx2, float dy1,
                                                   float dy2, float qxs, float qys, float qzs,
                                                   float rsc, float vz, int compat,
                                                   bool want_rho, CrossQueue &Q)
{
    double *wx = jw, *wy = jw + WH, *wz = jw + 2 * WH;
    const float dimid = ((float)a + dx1 + dx2 + 1.0f) * 0.5f - 0.5f;
    const float djmid = ((float)b + dy1 + dy2 + 1.0f) * 0.5f - 0.5f;
    const int di = signbit(dimid) ? -1 : 1, dj = signbit(djmid) ? -1 : 1;
    const float ai = fabsf(dimid), aj = fabsf(djmid);
    const float s1 = ai * aj, s2 = ai - s1, s3 = aj - s1, s4 = 1.0f - s1 - s2 - s3;
    const float quz = qzs * vz;
    jadd(wz + base, quz * s4);
    jadd(wz + base + di, quz * s2);
    jadd(wz + base + dj * W, quz * s3);
    jadd(wz + base + dj * W + di, quz * s1);
    const bool cross = (a | b) != 0;
    if (!cross) {
        const float xm = (dx1 + dx2) * 0.5f, ym = (dy1 + dy2) * 0.5f;
        const float fx = qxs * (dx2 - dx1), fy = qys * (dy2 - dy1);
        jadd(wx + base, fx * (0.5f - ym));
        jadd(wx + base + W, fx * (0.5f + ym));
        jadd(wy + base, fy * (0.5f - xm));
        jadd(wy + base + 1, fy * (0.5f + xm));
    }
    if (want_rho) {                          // rho at time n (:1283-1295)
        float r1, r2, r3;
        if (compat & PBPIC_COMPAT_RHO_MIRROR) {
            r1 = (0.5f - dx1) * (0.5f - dy1);
            r2 = 0.5f - dx1 - r1;
            r3 = 0.5f - dy1 - r1;
        } else {
            r1 = (0.5f + dx1) * (0.5f + dy1);
            r2 = 0.5f + dx1 - r1;
            r3 = 0.5f + dy1 - r1;
        }
        const float r4 = 1.0f - r1 - r2 - r3, qr = qzs * rsc;
        double *wr = jw + 3 * WH;
        jadd(wr + base, r4 * qr);
        jadd(wr + base + 1, r2 * qr);
        jadd(wr + base + W, r3 * qr);
        jadd(wr + base + W + 1, r1 * qr);
    }
    Q.push(cross, base, a, b, dx1, dx2, dy1, dy2);
}

// One push tile t (one workgroup).  OLDU: the reference's Boris rotation
// with the pre-kick momentum (PBPIC_COMPAT_BORIS_OLD_U); RHO: deposit rho as
// well (PBPIC_FLAG_DEPOSIT_RHO); REC (split form): write the pushed particles
// as the sort's records at S.rec_o in place of the five value arrays; DEP:
// deposit the current (false: the adjoint's replay, whose J comes from the
// checkpointed forward).
// OCC: workgroups per CU the launch is built for.  At four the rounds are
// not prefetched (the registers of the next round's particles would spill;
// the fourth wave per SIMD hides the load latency instead).
template <typename R, bool SPLIT, bool OLDU, bool RHO, bool REC, bool DEP, int NT, int OCC>
__device__ __forceinline__ void push_tile(const PushParams &P, int t)
{
    constexpr bool PF = OCC < 4;
    constexpr int V = Vw<R>::V;
    constexpr int CH = 64 * V;                                   // slots per wave-wide load
    constexpr int NW = NT / 64;                                  // waves per workgroup
    // rounds per wave: 8 (FP64) or 4 (FP32) with four waves, half with eight
    constexpr int ROUNDS = PUSH_TILE / (CH * NW);
    static_assert(ROUNDS * CH * NW == PUSH_TILE, "a tile is whole rounds of every wave");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int red[NW][5];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int s = 0;
    while (s + 1 < P.nspecies && t >= P.sp[s + 1].tile0) s++;
    const SpeciesDev S = P.sp[s];
    const int64_t tstart = (int64_t)(t - S.tile0) * PUSH_TILE;
    // wave w owns rounds w*ROUNDS .. w*ROUNDS+ROUNDS-1 of the tile (tile_rank)
    const int64_t pbase = tstart + (int64_t)wave * ROUNDS * CH + lane * V;
    const int nx = P.nx, ny = P.ny, nxh = P.nx >> 1, nyh = P.ny >> 1;
    const int64_t n = S.n;

    // ---- A: bounding box of the tile's cells, unwrapped around (x0, y0):
    // the box the previous push recorded, or else a pass over the cells
    int x0 = 0, y0 = 0, mnx = 1 << 30, mxx = -(1 << 30), mny = 1 << 30, mxy = -(1 << 30);
    int err = 0, aerr = 0;
    bool boxed = false;
    if (S.tbox_ok) {
        const int4 b = *reinterpret_cast<const int4 *>(S.tbox + 4 * (t - S.tile0));
        if (b.y >= 0 && b.w >= 0) {
            x0 = b.x; y0 = b.z; mnx = 0; mxx = b.y; mny = 0; mxy = b.w;
            boxed = true;
        }
    }
    if (!boxed) {
    const int c00 = S.cell[tstart];
    if ((unsigned)c00 < (unsigned)(nx * ny)) {
        x0 = (int)P.dny.div31((uint32_t)c00);
        y0 = c00 - x0 * ny;
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        int cv[V];
        vld(S.cell, pbase + r * CH, n, cv);
#pragma unroll
        for (int j = 0; j < V; j++) {
            const int c = cv[j];
            if (pbase + r * CH + j >= n) continue;
            if ((unsigned)c >= (unsigned)(nx * ny)) { err |= ERRF_CELL; continue; }
            const int cx = (int)P.dny.div31((uint32_t)c), cy = c - cx * ny;
            int rx = cx - x0, ry = cy - y0;
            if (rx > nxh) rx -= nx; else if (rx < -nxh) rx += nx;
            if (ry > nyh) ry -= ny; else if (ry < -nyh) ry += ny;
            mnx = min(mnx, rx); mxx = max(mxx, rx);
            mny = min(mny, ry); mxy = max(mxy, ry);
        }
    }
    aerr = err;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mnx = min(mnx, __shfl_xor(mnx, o));
        mxx = max(mxx, __shfl_xor(mxx, o));
        mny = min(mny, __shfl_xor(mny, o));
        mxy = max(mxy, __shfl_xor(mxy, o));
        aerr |= __shfl_xor(aerr, o);
    }
    if (lane == 0) {
        red[wave][0] = mnx; red[wave][1] = mxx; red[wave][2] = mny; red[wave][3] = mxy;
        red[wave][4] = aerr;
    }
    __syncthreads();
    aerr = 0;
    mnx = 1 << 30; mxx = -(1 << 30); mny = 1 << 30; mxy = -(1 << 30);
#pragma unroll
    for (int w = 0; w < NW; w++) {
        aerr |= red[w][4];
        mnx = min(mnx, red[w][0]); mxx = max(mxx, red[w][1]);
        mny = min(mny, red[w][2]); mxy = max(mxy, red[w][3]);
    }
    }
    const int wx0 = mnx - 1, wy0 = mny - 1;
    const int W = mxx - mnx + 4, H = mxy - mny + 4;
    const bool inlds = mnx <= mxx && (int64_t)W * H <= P.maxw;
    if (P.pass == 1 && !inlds && mnx <= mxx && (int64_t)W * H <= P.maxw_big) {
        // the box is block-uniform: the whole workgroup leaves together.
        // Lists: 2 the three-per-CU tier (after a four-per-CU first launch),
        // 0 the middle tier, 1 the last
        const int64_t wh = (int64_t)W * H;
        const int which = wh <= P.maxw_occ3 ? 2 : wh <= P.maxw_mid ? 0 : 1;
        if (tid == 0) P.big[BIG_LISTS + which * P.ntiles + atomicAdd(P.big + which, 1)] = t;
        return;
    }
    if (!inlds && mnx <= mxx && tid == 0) unsafeAtomicAdd(P.acc + ACC_SPILL, 1.0);
    const int WH = inlds ? W * H : 0;
    // fast path: a full tile of valid cells, at least one cell away from every
    // domain edge (no periodic unwrap or wrap for any particle or move)
    const bool fast = inlds && aerr == 0 && tstart + PUSH_TILE <= n && x0 + mnx >= 1 &&
                      x0 + mxx <= nx - 2 && y0 + mny >= 1 && y0 + mxy <= ny - 2;
    const int njc = RHO ? 4 : 3;
    constexpr int NS = LDS_NS;
    constexpr int GNS = NodeStride<R>::v;    // node-field array record
    R *nf = (R *)smem;                       // [W][H][NS] node fields, y fastest
    // current window [njc][H][W], x fastest: 64-bit fixed-point sums
    double *jw = (double *)(nf + (size_t)NS * P.maxw);

    // ---- B: stage node fields (contiguous column runs of the y-fastest
    // node-field array), zero the current window
    if (inlds) {
        const R *gnf = (const R *)P.nf;
        for (int k = tid; k < WH; k += NT) {
            const int wi = k / H, wj = k - wi * H;
            const int gx = wrap_index(x0 + wx0 + wi, nx), gy = wrap_index(y0 + wy0 + wj, ny);
            copy_node<R>(nf + (size_t)k * NS, gnf + ((size_t)gx * ny + gy) * GNS);
        }
        if constexpr (DEP)
            for (int k = tid; k < njc * WH; k += NT) jw[k] = 0.0;
    }
    __syncthreads();

    // ---- C: push, V particles per lane per round, next round prefetched
    const R c0 = (R)S.c0, c1 = (R)S.c1, qx = (R)S.qx, qy = (R)S.qy, qr = (R)S.q;
    const R *ox_p = (const R *)S.ox, *oy_p = (const R *)S.oy, *ux_p = (const R *)S.ux,
            *uy_p = (const R *)S.uy, *uz_p = (const R *)S.uz;
    R *ox_q = (R *)S.ox_o, *oy_q = (R *)S.oy_o, *ux_q = (R *)S.ux_o, *uy_q = (R *)S.uy_o,
      *uz_q = (R *)S.uz_o;
    R *gjx = (R *)P.jx, *gjy = (R *)P.jy, *gjz = (R *)P.jz, *grho = (R *)P.rho;
    constexpr bool want_rho = RHO;
    // deposit contributions come out of deposit_terms already multiplied by
    // their plane's fixed-point scale (q*dx/dt, q*dy/dt and q pre-scaled by
    // powers of two, and by 2^FX2_LOBITS in the FP64 form: exact), for the
    // LDS window and for the global planes alike
    const R fxs = SPLIT ? (R)1 : (R)(1 << FX2_LOBITS);
    const R dqx = qx * (R)S.jsx * fxs, dqy = qy * (R)S.jsy * fxs, dqz = qr * (R)S.jsz * fxs,
            rsc = (R)S.jsr / (R)S.jsz;
    // window coordinates of the cells this push reads (o*) and writes (n*):
    // the first check the recorded box, the second become the next one
    int omnx = 1 << 30, omxx = -(1 << 30), omny = 1 << 30, omxy = -(1 << 30);
    int nmnx = 1 << 30, nmxx = -(1 << 30), nmny = 1 << 30, nmxy = -(1 << 30);
    auto track = [&](int lx, int ly, int dcx, int dcy) {
        omnx = min(omnx, lx); omxx = max(omxx, lx);
        omny = min(omny, ly); omxy = max(omxy, ly);
        nmnx = min(nmnx, lx + dcx); nmxx = max(nmxx, lx + dcx);
        nmny = min(nmny, ly + dcy); nmxy = max(nmxy, ly + dcy);
    };

    // one particle: gather, push, move, deposit; returns the new cell
    auto one = [&](int c, R &ox, R &oy, R &ux, R &uy, R &uz) -> int {
        const int cx = (int)P.dny.div31((uint32_t)c), cy = c - cx * ny;
        int rx = cx - x0, ry = cy - y0;
        if (rx > nxh) rx -= nx; else if (rx < -nxh) rx += nx;
        if (ry > nyh) ry -= ny; else if (ry < -nyh) ry += ny;
        const int lx = rx - wx0, ly = ry - wy0;
        R n00[6], n10[6], n01[6], n11[6];
        if (inlds) {
            const R *b = nf + (size_t)(lx * H + ly) * NS;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                n00[q] = b[q];
                n10[q] = b[(size_t)H * NS + q];
                n01[q] = b[NS + q];
                n11[q] = b[(size_t)H * NS + NS + q];
            }
        } else {
            const int cx1 = cx + 1 == nx ? 0 : cx + 1, cy1 = cy + 1 == ny ? 0 : cy + 1;
            double f[6];
            node_fields(P, cx, cy, f);
            for (int q = 0; q < 6; q++) n00[q] = (R)f[q];
            node_fields(P, cx1, cy, f);
            for (int q = 0; q < 6; q++) n10[q] = (R)f[q];
            node_fields(P, cx, cy1, f);
            for (int q = 0; q < 6; q++) n01[q] = (R)f[q];
            node_fields(P, cx1, cy1, f);
            for (int q = 0; q < 6; q++) n11[q] = (R)f[q];
        }
        const R dx0 = ox, dy0 = oy;
        int dcx, dcy;
        R vz, gam;
        if constexpr (SPLIT) {
            push_math_split<OLDU>(P, c0, c1, ox, oy, ux, uy, uz, n00, n10, n01, n11, dcx, dcy, vz,
                            err);
        } else {
            push_math<R, OLDU>(P, c0, c1, ox, oy, ux, uy, uz, n00, n10, n01, n11, dcx, dcy, vz, gam,
                         err);
        }
        wall_reflect(P.nyp, cy, dcy, oy, uy);
        if constexpr (DEP) {
            Dep<R> d;
            deposit_terms<R, SPLIT>(cx + 1, cy + 1, dcx, dcy, dx0, ox, dy0, oy, dqx, dqy,
                                    dqz, vz, P.compat, want_rho, d);
            if (want_rho) { d.r0 *= rsc; d.r1 *= rsc; d.r2 *= rsc; d.r3 *= rsc; }
            // jy at column mx+2 is lost by the reference's fold (:1156, :1159)
            if ((P.compat & PBPIC_COMPAT_JY_DROP) && cx == nx - 1 && dcx == 1) d.jyb1 = (R)0;
            if (d.a == 0 && d.b == 0) {
                d.jxa0 += d.jxb0; d.jxa1 += d.jxb1; d.jya0 += d.jyb0; d.jya1 += d.jyb1;
            }
            if (inlds) {
                deposit_lds(jw, WH, W, ly * W + lx, d, want_rho);
            } else {
                const int xm = cx == 0 ? nx - 1 : cx - 1, xp = cx + 1 == nx ? 0 : cx + 1;
                const int ym = cy == 0 ? ny - 1 : cy - 1, yp = cy + 1 == ny ? 0 : cy + 1;
                const int xs[3] = {xm, cx, xp}, ys[3] = {ym, cy, yp};
                auto at = [&](int ddx, int ddy) -> size_t {
                    int gx = ddx == 2 ? (xp + 1 == nx ? 0 : xp + 1) : xs[ddx + 1];
                    int gy = ddy == 2 ? (yp + 1 == ny ? 0 : yp + 1) : ys[ddy + 1];
                    return (size_t)gy * nx + gx;
                };
                auto put = [&](R *plane, size_t g, R v) {
                    long long *p = reinterpret_cast<long long *>(plane) + g;
                    if constexpr (SPLIT) gfixf(p, v);
                    else fx2_add(p, P.jlo, fx2_quant(v));
                };
                put(gjx, at(0, 0), d.jxa0);
                put(gjx, at(0, 1), d.jxa1);
                put(gjy, at(0, 0), d.jya0);
                put(gjy, at(1, 0), d.jya1);
                if (d.a != 0 || d.b != 0) {
                    put(gjx, at(d.a, d.b), d.jxb0);
                    put(gjx, at(d.a, d.b + 1), d.jxb1);
                    put(gjy, at(d.a, d.b), d.jyb0);
                    put(gjy, at(d.a + 1, d.b), d.jyb1);
                }
                put(gjz, at(0, 0), d.jz0);
                put(gjz, at(d.di, 0), d.jz1);
                put(gjz, at(0, d.dj), d.jz2);
                put(gjz, at(d.di, d.dj), d.jz3);
                if (want_rho) {
                    put(grho, at(0, 0), d.r0);
                    put(grho, at(1, 0), d.r1);
                    put(grho, at(0, 1), d.r2);
                    put(grho, at(1, 1), d.r3);
                }
            }
        }
        track(lx, ly, dcx, dcy);
        // periodic wrap (:1032-1052)
        int ncx = cx + dcx, ncy = cy + dcy;
        ncx = ncx < 0 ? ncx + nx : (ncx >= nx ? ncx - nx : ncx);
        ncy = ncy < 0 ? ncy + ny : (ncy >= ny ? ncy - ny : ncy);
        return ncx * ny + ncy;
    };

    if (fast) {
        const int XW = x0 + wx0, YW = y0 + wy0;        // window origin, absolute
        const int32_t *cp = S.cell + tstart;
        const R *op = ox_p + tstart, *pp = oy_p + tstart, *up = ux_p + tstart,
                *vp = uy_p + tstart, *wp = uz_p + tstart;
        int32_t *cq = S.cell_o + tstart;
        R *oq = ox_q + tstart, *pq = oy_q + tstart, *uq = ux_q + tstart, *vq = uy_q + tstart,
          *wq = uz_q + tstart;
        const int lo = wave * ROUNDS * CH + lane * V;
        CrossQueue Q;
        Q.init(smem + push_queue_offset(P.maxw, RHO ? 1 : 0), wave, lane);
        auto onef = [&](int c, R &ox, R &oy, R &ux, R &uy, R &uz) -> int {
            const int cx = (int)P.dny.div31((uint32_t)c), cy = c - cx * ny;
            const R *b = nf + ((cx - XW) * H + (cy - YW)) * NS;
            R n00[6], n10[6], n01[6], n11[6];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                n00[q] = b[q];
                n10[q] = b[H * NS + q];
                n01[q] = b[NS + q];
                n11[q] = b[H * NS + NS + q];
            }
            const R dx0 = ox, dy0 = oy;
            int dcx, dcy;
            R vz, gam;
            if constexpr (SPLIT) {
                push_math_split<OLDU>(P, c0, c1, ox, oy, ux, uy, uz, n00, n10, n01, n11, dcx, dcy,
                                vz, err);
            } else {
                push_math<R, OLDU>(P, c0, c1, ox, oy, ux, uy, uz, n00, n10, n01, n11, dcx, dcy, vz,
                             gam, err);
            }
            wall_reflect(P.nyp, cy, dcy, oy, uy);
            const int base = (cy - YW) * W + (cx - XW);
            if constexpr (!DEP) {
            } else if constexpr (SPLIT) {
                deposit_split_fast(jw, WH, W, base, dcx, dcy, dx0, ox, dy0, oy, dqx, dqy, dqz,
                                   rsc, vz, P.compat, want_rho, Q);
                if (Q.tail - Q.head >= 64) Q.drain(jw, WH, W, dqx, dqy, P.compat, 64);
            } else {
                Dep<R> d;
                deposit_terms<R, SPLIT>(cx + 1, cy + 1, dcx, dcy, dx0, ox, dy0, oy, dqx, dqy,
                                        dqz, vz, P.compat, want_rho, d);
                if (want_rho) { d.r0 *= rsc; d.r1 *= rsc; d.r2 *= rsc; d.r3 *= rsc; }
                if (d.a == 0 && d.b == 0) {
                    d.jxa0 += d.jxb0; d.jxa1 += d.jxb1; d.jya0 += d.jyb0; d.jya1 += d.jyb1;
                }
                deposit_lds(jw, WH, W, base, d, want_rho);
            }
            track(cx - XW, cy - YW, dcx, dcy);
            return c + dcx * ny + dcy;          // interior tile: no wrap (:1032-1052)
        };
        R cur[5][V], nxt[5][V];
        int ccur[V], cnxt[V];
        ldu(cp + lo, ccur);
        ldu(op + lo, cur[0]); ldu(pp + lo, cur[1]); ldu(up + lo, cur[2]);
        ldu(vp + lo, cur[3]); ldu(wp + lo, cur[4]);
#pragma unroll 1
        for (int r = 0; r < ROUNDS; r++) {
            const int o = lo + r * CH;
            if constexpr (PF) {
                if (r + 1 < ROUNDS) {
                    ldu(cp + o + CH, cnxt);
                    ldu(op + o + CH, nxt[0]); ldu(pp + o + CH, nxt[1]); ldu(up + o + CH, nxt[2]);
                    ldu(vp + o + CH, nxt[3]); ldu(wp + o + CH, nxt[4]);
                }
            } else if (r > 0) {
                ldu(cp + o, ccur);
                ldu(op + o, cur[0]); ldu(pp + o, cur[1]); ldu(up + o, cur[2]);
                ldu(vp + o, cur[3]); ldu(wp + o, cur[4]);
            }
            int nc[V];
#pragma unroll
            for (int j = 0; j < V; j++)
                nc[j] = onef(ccur[j], cur[0][j], cur[1][j], cur[2][j], cur[3][j], cur[4][j]);
            stu(cq + o, nc);
            if constexpr (REC) {
                st_rec4(S.rec_o, tstart + o, nc, cur);
            } else {
                stu(oq + o, cur[0]); stu(pq + o, cur[1]); stu(uq + o, cur[2]);
                stu(vq + o, cur[3]); stu(wq + o, cur[4]);
            }
            if constexpr (PF) {
#pragma unroll
                for (int j = 0; j < V; j++) {
                    ccur[j] = cnxt[j];
#pragma unroll
                    for (int q = 0; q < 5; q++) cur[q][j] = nxt[q][j];
                }
            }
        }
        if constexpr (SPLIT && DEP) Q.drain(jw, WH, W, dqx, dqy, P.compat, Q.tail - Q.head);
    } else {
    // the tile's cells were read in phase A and are re-read here from L2
    R cur[5][V], nxt[5][V];
    int ccur[V], cnxt[V];
    {
        const int64_t i0 = pbase;
        vld(S.cell, i0, n, ccur);
        vld(ox_p, i0, n, cur[0]); vld(oy_p, i0, n, cur[1]); vld(ux_p, i0, n, cur[2]);
        vld(uy_p, i0, n, cur[3]); vld(uz_p, i0, n, cur[4]);
    }
#pragma unroll 1
    for (int r = 0; r < ROUNDS; r++) {
        const int64_t i0 = pbase + r * CH;
        if constexpr (PF) {
            if (r + 1 < ROUNDS) {
                const int64_t i1 = i0 + CH;
                vld(S.cell, i1, n, cnxt);
                vld(ox_p, i1, n, nxt[0]); vld(oy_p, i1, n, nxt[1]); vld(ux_p, i1, n, nxt[2]);
                vld(uy_p, i1, n, nxt[3]); vld(uz_p, i1, n, nxt[4]);
            }
        } else if (r > 0) {
            vld(S.cell, i0, n, ccur);
            vld(ox_p, i0, n, cur[0]); vld(oy_p, i0, n, cur[1]); vld(ux_p, i0, n, cur[2]);
            vld(uy_p, i0, n, cur[3]); vld(uz_p, i0, n, cur[4]);
        }
        if (i0 < n) {
            int nc[V];
#pragma unroll
            for (int j = 0; j < V; j++) {
                const int c = ccur[j];
                nc[j] = c;
                if (i0 + j < n && (unsigned)c < (unsigned)(nx * ny))
                    nc[j] = one(c, cur[0][j], cur[1][j], cur[2][j], cur[3][j], cur[4][j]);
            }
            vst(S.cell_o, i0, n, nc);
            if constexpr (REC) {
                vst_rec4(S.rec_o, i0, n, nc, cur);
            } else {
                vst(ox_q, i0, n, cur[0]); vst(oy_q, i0, n, cur[1]); vst(ux_q, i0, n, cur[2]);
                vst(uy_q, i0, n, cur[3]); vst(uz_q, i0, n, cur[4]);
            }
        }
        if constexpr (PF) {
#pragma unroll
            for (int j = 0; j < V; j++) {
                ccur[j] = cnxt[j];
#pragma unroll
                for (int q = 0; q < 5; q++) cur[q][j] = nxt[q][j];
            }
        }
    }
    }

    // ---- D: flush the window, reduce the kinetic energy and error flags
    __syncthreads();
    if (DEP && inlds) {
        for (int k = tid; k < WH; k += NT) {
            const int wj = k / W, wi = k - wj * W;
            const int gx = wrap_index(x0 + wx0 + wi, nx), gy = wrap_index(y0 + wy0 + wj, ny);
            const size_t g = (size_t)gy * nx + gx;
            // integer sums into the fixed-point planes: exact, any order
            const long long *jb = reinterpret_cast<const long long *>(jw);
            R *const planes[4] = {gjx, gjy, gjz, grho};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (q == 3 && !want_rho) break;
                long long *gp = reinterpret_cast<long long *>(planes[q]) + g;
                const long long v = jb[q * WH + k];
                if (v == 0) continue;
                if constexpr (SPLIT) gfix(gp, v);
                else fx2_add(gp, P.jlo, v);
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) err |= __shfl_xor(err, o);
    if (lane == 0 && err) atomicOr(P.err, err);

    // ---- E: record the box of the new cells for the next push; check that
    // every cell read lay inside the box this push assumed
    int v[8] = {omnx, -omxx, omny, -omxy, nmnx, -nmxx, nmny, -nmxy};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = min(v[q], __shfl_xor(v[q], o));
    __shared__ int bred[NW][8];
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 8; q++) bred[wave][q] = v[q];
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            v[q] = bred[0][q];
#pragma unroll
            for (int w = 1; w < NW; w++) v[q] = min(v[q], bred[w][q]);
        }
        const bool any = v[4] <= -v[5];            // at least one valid particle
        // cells read must lie inside the window's interior [1, W-3] x [1, H-3]
        if (any && (v[0] < 1 || -v[1] > W - 3 || v[2] < 1 || -v[3] > H - 3))
            atomicOr(P.err, ERRF_STATE);
        // the next push unwraps cells into (-n/2, n/2] around the box origin:
        // a box spanning half the domain or more would be ambiguous (for even
        // n the offset n/2 unwraps to -n/2), so it is left unknown and the
        // next push scans the cells
        int4 nb = make_int4(0, -1, 0, -1);
        const int ex = -v[5] - v[4], ey = -v[7] - v[6];
        if (any && aerr == 0 && ex < nxh && ey < nyh) {
            nb.x = wrap_index(x0 + wx0 + v[4], nx);
            nb.y = ex;
            nb.z = wrap_index(y0 + wy0 + v[6], ny);
            nb.w = ey;
        }
        *reinterpret_cast<int4 *>(S.tbox + 4 * (t - S.tile0)) = nb;
    }
}

template <typename R, bool SPLIT, bool OLDU, bool RHO, bool REC, bool DEP, int OCC>
__global__ __launch_bounds__(PUSH_THREADS, OCC) void push_kernel(PushParams P)
{
    push_tile<R, SPLIT, OLDU, RHO, REC, DEP, PUSH_THREADS, OCC>(P, xcd_remap(blockIdx.x, gridDim.x));
}

// Passes 4, 2 and 3: the tiles the first launch deferred, list entry k taken
// by workgroup k mod gridDim.x.  The grid is an eighth of the tiles, so a
// pass with nothing to do costs little, and one with fewer deferred tiles
// than workgroups takes one per workgroup in list order (the hardware's
// dispatch order; taking them from a shared counter instead was 6% slower on
// a Harris sheet).  OCC is the workgroups per CU the pass's LDS window
// allows (its register budget): pass 4 (after a first launch at four per CU)
// three, pass 2 two, pass 3 one.  The last pass, whose window takes a CU's
// LDS, runs workgroups of eight waves (NT = 512, two per SIMD) on the tile's
// rounds, where four would leave one wave per SIMD to hide every memory and
// LDS latency.
template <typename R, bool SPLIT, bool OLDU, bool RHO, bool REC, bool DEP, int NT, int OCC>
__global__ __launch_bounds__(NT, OCC) void push_kernel_listed(PushParams P)
{
    const int which = P.pass == 4 ? 2 : P.pass - 2, cnt = P.big[which];
    for (int k = blockIdx.x; k < cnt; k += gridDim.x) {
        push_tile<R, SPLIT, OLDU, RHO, REC, DEP, NT, OCC>(P, P.big[BIG_LISTS + which * P.ntiles + k]);
        __syncthreads();
    }
}

// Batched currentdeposit with global atomics (unit-test entry point).
template <typename R, bool SPLIT>
__global__ void deposit_batch_kernel(PushParams P, int64_t n, const int32_t *i1,
                                     const int32_t *i2, const int32_t *j1,
                                     const int32_t *j2, const double *dx1,
                                     const double *dx2, const double *dy1,
                                     const double *dy2, const int32_t *q,
                                     const double *vz)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int nx = P.nx, ny = P.ny;
    // qx = q*dx/dt (:1237): sp[0].qx holds dx/dt; q is +-1 (checked on the
    // host), so q*(dx/dt) equals (q*dx)/dt exactly.  The contributions are
    // pre-scaled to the fixed-point planes (powers of two).
    const R fxs = SPLIT ? (R)1 : (R)(1 << FX2_LOBITS);
    const R sx = (R)P.jsx * fxs, sy = (R)P.jsy * fxs, sz = (R)P.jsz * fxs,
            sr = (R)P.jsr / (R)P.jsz;
    const R qr = (R)q[k] * sz;
    const R qxx = (R)((double)q[k] * P.sp[0].qx) * sx, qyy = (R)((double)q[k] * P.sp[0].qy) * sy;
    Dep<R> d;
    const int a = i2[k] - i1[k], b = j2[k] - j1[k];
    deposit_terms<R, SPLIT>(i1[k], j1[k], a, b, (R)dx1[k], (R)dx2[k], (R)dy1[k],
                            (R)dy2[k], qxx, qyy, qr, (R)vz[k], P.compat,
                            P.deposit_rho != 0, d);
    d.r0 *= sr; d.r1 *= sr; d.r2 *= sr; d.r3 *= sr;
    if ((P.compat & PBPIC_COMPAT_JY_DROP) && i2[k] == nx + 1) d.jyb1 = (R)0;
    const int cx = i1[k] - 1, cy = j1[k] - 1;
    auto at = [&](int ddx, int ddy) -> size_t {
        return (size_t)wrap_index(cy + ddy, ny) * nx + wrap_index(cx + ddx, nx);
    };
    auto put = [&](void *plane, size_t g, R v) {
        long long *p = reinterpret_cast<long long *>(plane) + g;
        if constexpr (SPLIT) gfixf(p, v);
        else fx2_add(p, P.jlo, fx2_quant((double)v));
    };
    put(P.jx, at(0, 0), d.jxa0);
    put(P.jx, at(0, 1), d.jxa1);
    put(P.jy, at(0, 0), d.jya0);
    put(P.jy, at(1, 0), d.jya1);
    put(P.jx, at(a, b), d.jxb0);
    put(P.jx, at(a, b + 1), d.jxb1);
    put(P.jy, at(a, b), d.jyb0);
    put(P.jy, at(a + 1, b), d.jyb1);
    put(P.jz, at(0, 0), d.jz0);
    put(P.jz, at(d.di, 0), d.jz1);
    put(P.jz, at(0, d.dj), d.jz2);
    put(P.jz, at(d.di, d.dj), d.jz3);
    if (P.deposit_rho) {
        put(P.rho, at(0, 0), d.r0);
        put(P.rho, at(1, 0), d.r1);
        put(P.rho, at(0, 1), d.r2);
        put(P.rho, at(1, 1), d.r3);
    }
}

static int push_tiles(const PushParams &p, PushParams &q)
{
    q = p;
    int t = 0;
    for (int s = 0; s < p.nspecies; s++) {
        q.sp[s].tile0 = t;
        t += (int)((p.sp[s].n + PUSH_TILE - 1) / PUSH_TILE);
    }
    q.ntiles = t;
    return t;
}

// Kinetic energy tally (pic2d.f90:1061-1062): sum over a species of
// (gamma - 1) * ms from the momenta as the last push left them, FP64 sums.
template <typename R>
__global__ __launch_bounds__(256) void kinetic_kernel(SpeciesDev S, double *part)
{
    __shared__ double red[4];
    double k = 0;
    const R *ux = (const R *)S.ux, *uy = (const R *)S.uy, *uz = (const R *)S.uz;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < S.n;
         i += (int64_t)gridDim.x * 256)
        k += (double)kin_energy(ux[i], uy[i], uz[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = k;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

hipError_t launch_kinetic(const SpeciesDev &S, int precision, double *acc, double *part,
                          hipStream_t st)
{
    if (S.n == 0) return hipSuccess;
    int64_t nb = (S.n + 255) / 256;
    if (nb > 4096) nb = 4096;
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL(kinetic_kernel<double>, dim3((unsigned)nb), dim3(256), 0, st, S, part);
    else
        hipLaunchKernelGGL(kinetic_kernel<float>, dim3((unsigned)nb), dim3(256), 0, st, S, part);
    PartSlots ps{1, {S.q > 0 ? ACC_KE_I : ACC_KE_E, 0, 0, 0}, {S.ms, 0, 0, 0}};
    return launch_reduce_parts(part, (int)nb, acc, ps, st);
}

size_t nodefield_bytes(int precision, int64_t ncell)
{
    return (size_t)ncell * (precision == PBPIC_PREC_FP64 ? 6 * 8 : 8 * 4);
}

hipError_t launch_nodefield(const PushParams &p, int precision, hipStream_t st)
{
    const dim3 g((unsigned)((p.nx + NF_T - 1) / NF_T), (unsigned)((p.ny + NF_T - 1) / NF_T));
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL((nodefield_kernel<double, 6>), g, dim3(256), 0, st, p, (double *)p.nf);
    else
        hipLaunchKernelGGL((nodefield_kernel<float, 8>), g, dim3(256), 0, st, p, (float *)p.nf);
    return hipGetLastError();
}

// dep = 0: a push without deposit (DEP = false) uses the node-field window
// only; nwaves: waves per workgroup (the split form's crossing queues)
size_t push_lds_bytes(int precision, int maxw, int deposit_rho, int dep, int nwaves)
{
    const size_t r = precision == PBPIC_PREC_FP64 ? 8 : 4;
    const size_t ns = LDS_NS;
    if (!dep) return (size_t)maxw * ns * r;
    const size_t windows = (size_t)maxw * (ns * r + (deposit_rho ? 4 : 3) * 8);
    return precision == PBPIC_PREC_FP64 ? windows : windows + (size_t)nwaves * XQ_BYTES;
}

// threads per workgroup of a push pass: eight waves for the last listed pass
constexpr int push_threads_of_pass(int pass) { return pass == 3 ? 2 * PUSH_THREADS : PUSH_THREADS; }

template <typename R, bool SPLIT, bool OLDU, bool RHO, bool REC, bool DEP>
static void launch_push_t(const PushParams &q, int nt, size_t lds, hipStream_t st)
{
    if (q.pass == 3) {
        constexpr int NT = push_threads_of_pass(3);
        (void)hipFuncSetAttribute((const void *)push_kernel_listed<R, SPLIT, OLDU, RHO, REC, DEP, NT, 1>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL((push_kernel_listed<R, SPLIT, OLDU, RHO, REC, DEP, NT, 1>), dim3(nt),
                           dim3(NT), lds, st, q);
        return;
    }
    if (q.pass == 2) {
        constexpr int NT = push_threads_of_pass(2);
        (void)hipFuncSetAttribute((const void *)push_kernel_listed<R, SPLIT, OLDU, RHO, REC, DEP, NT, 2>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL((push_kernel_listed<R, SPLIT, OLDU, RHO, REC, DEP, NT, 2>), dim3(nt),
                           dim3(NT), lds, st, q);
        return;
    }
    if constexpr (SPLIT && DEP && !REC) {
        if (q.pass == 4) {
            (void)hipFuncSetAttribute(
                (const void *)push_kernel_listed<R, SPLIT, OLDU, RHO, REC, DEP, PUSH_THREADS, 3>,
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL((push_kernel_listed<R, SPLIT, OLDU, RHO, REC, DEP, PUSH_THREADS, 3>),
                               dim3(nt), dim3(PUSH_THREADS), lds, st, q);
            return;
        }
    }
    if (q.pass == 4) return;       // only the plain split push has this tier
    if constexpr (SPLIT && DEP && !REC) {
        if (q.occ4) {
            (void)hipFuncSetAttribute((const void *)push_kernel<R, SPLIT, OLDU, RHO, REC, DEP, 4>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL((push_kernel<R, SPLIT, OLDU, RHO, REC, DEP, 4>), dim3(nt),
                               dim3(PUSH_THREADS), lds, st, q);
            return;
        }
    }
    (void)hipFuncSetAttribute((const void *)push_kernel<R, SPLIT, OLDU, RHO, REC, DEP, 3>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((push_kernel<R, SPLIT, OLDU, RHO, REC, DEP, 3>), dim3(nt), dim3(PUSH_THREADS), lds, st,
                       q);
}

template <typename R, bool SPLIT, bool REC, bool DEP>
static void launch_push_o(const PushParams &q, int nt, size_t lds, hipStream_t st)
{
    const bool oldu = (q.compat & PBPIC_COMPAT_BORIS_OLD_U) != 0, rho = q.deposit_rho != 0;
    if constexpr (!DEP) {
        // no deposit: rho is not deposited either
        if (oldu) launch_push_t<R, SPLIT, true, false, REC, false>(q, nt, lds, st);
        else launch_push_t<R, SPLIT, false, false, REC, false>(q, nt, lds, st);
    } else {
        if (oldu && rho) launch_push_t<R, SPLIT, true, true, REC, true>(q, nt, lds, st);
        else if (oldu) launch_push_t<R, SPLIT, true, false, REC, true>(q, nt, lds, st);
        else if (rho) launch_push_t<R, SPLIT, false, true, REC, true>(q, nt, lds, st);
        else launch_push_t<R, SPLIT, false, false, REC, true>(q, nt, lds, st);
    }
}

template <typename R, bool SPLIT>
static void launch_push_f(const PushParams &q, int nt, size_t lds, hipStream_t st)
{
    // the step's sort follows: a separate build of the kernel writes the
    // sort's records, so the other steps' kernel carries none of it
    const bool rec = SPLIT && q.sp[0].rec_o != nullptr;
    if (q.nodep) {
        if (rec) launch_push_o<R, SPLIT, SPLIT, false>(q, nt, lds, st);
        else launch_push_o<R, SPLIT, false, false>(q, nt, lds, st);
        return;
    }
    if (rec) launch_push_o<R, SPLIT, SPLIT, true>(q, nt, lds, st);
    else launch_push_o<R, SPLIT, false, true>(q, nt, lds, st);
}

// Largest window whose LDS fits one workgroup per CU (150 KiB).
// Window capacity (nodes) of a launch whose workgroups take `budget` bytes of LDS.
static int push_maxw_for(int precision, int deposit_rho, int dep, size_t budget, int nwaves = 4)
{
    const size_t per = push_lds_bytes(precision, 1, deposit_rho, dep, nwaves) -
                       push_lds_bytes(precision, 0, deposit_rho, dep, nwaves);
    return (int)((budget - push_lds_bytes(precision, 0, deposit_rho, dep, nwaves)) / per);
}

// one workgroup per CU (150 KiB) and two (78 KiB) of the 160 KiB
int push_maxw_big(int precision, int deposit_rho, int dep)
{
    return push_maxw_for(precision, deposit_rho, dep, 150 * 1024, push_threads_of_pass(3) / 64);
}
int push_maxw_mid(int precision, int deposit_rho, int dep)
{
    return push_maxw_for(precision, deposit_rho, dep, 78 * 1024);
}
// the first launch of a push without deposit: the node-field window that
// fills the LDS the depositing push's window of maxw nodes takes (the same
// workgroups per CU)
int push_maxw_nodep(int precision, int deposit_rho, int maxw)
{
    return push_maxw_for(precision, deposit_rho, 0, push_lds_bytes(precision, maxw, deposit_rho, 1, 4));
}

hipError_t launch_push(const PushParams &p, int precision, hipStream_t st)
{
    PushParams q;
    const int nt = push_tiles(p, q);
    if (nt == 0) return hipSuccess;
    const int dep = p.nodep ? 0 : 1;
    // the split form's plain push: a first launch at four workgroups per CU
    // (PUSH_MAXW_OCC4-node windows), the tiles it cannot hold deferred to a
    // launch at three per CU with the context's window (pass 4), and wider
    // ones to the middle and last tiers as usual
    if (q.occ4) {
        q.maxw_occ3 = p.maxw;
        q.maxw = PUSH_MAXW_OCC4;
    } else {
        q.maxw_occ3 = 0;
    }
    const size_t lds = push_lds_bytes(precision, q.maxw, p.deposit_rho, dep, PUSH_THREADS / 64);
    q.pass = p.big && p.maxw_big > q.maxw ? 1 : 0;
    q.maxw_mid = std::max(q.maxw_occ3 ? q.maxw_occ3 : q.maxw, std::min(p.maxw_mid, p.maxw_big));
    // the list counts were cleared by the node-field launch before this one
    // (launch_nodefield, which every push follows)
    if (precision == PBPIC_PREC_FP64) launch_push_f<double, false>(q, nt, lds, st);
    else launch_push_f<float, true>(q, nt, lds, st);
    if (q.pass) {
        const int ng = std::max(1, (nt + 7) / 8);
        for (int pass : {4, 2, 3}) {
            PushParams b = q;
            b.pass = pass;
            if (pass == 4) {
                if (!q.occ4) continue;
                b.maxw = q.maxw_occ3;
            } else {
                b.maxw = pass == 2 ? q.maxw_mid : p.maxw_big;
                // no middle tier when the window before it is as wide
                if (pass == 2 && b.maxw == (q.occ4 ? q.maxw_occ3 : q.maxw)) continue;
            }
            const size_t lb = push_lds_bytes(precision, b.maxw, p.deposit_rho, dep,
                                             push_threads_of_pass(pass) / 64);
            if (precision == PBPIC_PREC_FP64) launch_push_f<double, false>(b, ng, lb, st);
            else launch_push_f<float, true>(b, ng, lb, st);
        }
    }
    return hipGetLastError();
}

hipError_t launch_deposit_batch(const PushParams &p, int precision, int64_t n,
                                const int32_t *i1, const int32_t *i2,
                                const int32_t *j1, const int32_t *j2,
                                const double *dx1, const double *dx2,
                                const double *dy1, const double *dy2,
                                const int32_t *q, const double *vz,
                                hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int nb = (int)((n + 255) / 256);
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL((deposit_batch_kernel<double, false>), dim3(nb), dim3(256), 0, st,
                           p, n, i1, i2, j1, j2, dx1, dx2, dy1, dy2, q, vz);
    else
        hipLaunchKernelGGL((deposit_batch_kernel<float, true>), dim3(nb), dim3(256), 0, st,
                           p, n, i1, i2, j1, j2, dx1, dx2, dy1, dy2, q, vz);
    return hipGetLastError();
}

}  // namespace pbpic
