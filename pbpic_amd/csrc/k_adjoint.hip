// k_adjoint.hip -- hand-written discrete adjoint of the PIC step for gfx950.
//
// Loss: L = sum_n dt (mime/c^2) sum_cells J^{n+1/2} . (E^n + E^{n+1})/2, the
// time-integrated J.E of the step (the same quantity as PBPIC_D_JE_INT).
// For one step, in reverse:
//   field solve   transposed curl stencils (fdtd_adj_* kernels);
//   deposit       a gather of the current adjoint at the particle's deposit
//                 targets (the transpose of the scatter), differentiated
//                 through the zigzag relay point and the CIC weights of Jz;
//   move          through u/gamma;
//   Boris push    reverse mode through the rotation (both the reference's
//                 pre-kick-u form, pic2d.f90:985-987, and the textbook form);
//   gather        a CIC scatter of the force adjoint onto nodes (LDS window,
//                 like the forward deposit) plus the weight derivative for
//                 the position adjoint, then the transpose of the node
//                 averaging onto the staggered fields;
//   sort          the stored permutation applied in reverse.
// The forward quantities of each particle are recomputed from the stored
// state (checkpointing happens in the host code of capi.hip).
#include "pbpic_impl.h"

// Multiply-adds are contracted.  The per-particle forward recompute keeps
// the forward kernel's expression order; the FP64 forward runs uncontracted,
// so its recompute can differ in the last bit, which matters only for a
// particle within one rounding error of a cell face (it is differentiated on
// the other side of it).
#pragma clang fp contract(fast)

namespace pbpic {

namespace {

constexpr int ADJ_THREADS = 256;
constexpr int ADJ_ROUNDS = 8;
constexpr int ADJ_TILE = ADJ_THREADS * ADJ_ROUNDS;      // 2048, half a push tile
// the second launch (one workgroup per CU) runs eight waves of four ranks each
constexpr int ADJ_THREADS_BIG = 2 * ADJ_THREADS;

// LDS bytes of the node-field window, rounded up to 8-byte alignment.
__host__ __device__ constexpr size_t adj_nf_bytes(size_t r, int maxw)
{
    return ((size_t)maxw * 6 * r + 7) & ~(size_t)7;
}


template <typename R>
__device__ __forceinline__ void a_lds_add(R *p, R v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void a_node_fields(const AdjParams &P, int gx, int gy, double f[6])
{
    const int nx = P.nx, ny = P.ny;
    const int gxm = gx == 0 ? nx - 1 : gx - 1, gym = gy == 0 ? ny - 1 : gy - 1;
    const size_t c = (size_t)gy * nx + gx, xm = (size_t)gy * nx + gxm;
    const size_t ym = (size_t)gym * nx + gx, xym = (size_t)gym * nx + gxm;
    f[0] = (P.ex[c] + P.ex[xm]) * 0.5;
    f[1] = (P.ey[c] + P.ey[ym]) * 0.5;
    f[2] = (P.ez[c] + P.ez[xm] + P.ez[ym] + P.ez[xym]) * 0.25;
    f[3] = (P.bx[c] + P.bx[xm]) * 0.5;
    f[4] = (P.by[c] + P.by[ym]) * 0.5;
    f[5] = (P.bz[c] + P.bz[xm] + P.bz[ym] + P.bz[xym]) * 0.25;
}

template <typename R> __device__ __forceinline__ R a_sgn(R x) { return signbit(x) ? (R)-1 : (R)1; }

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ void cross3(const float a[3], const float b[3], float o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Reverse mode through one zigzag segment pair along one axis.  Given the
// forward inputs (a = cell displacement, o1, o2 = old/new offsets, qf = q*dx/dt)
// and the adjoints of f1, f2, w1, w2, returns the adjoints of o1 and o2.
template <typename R>
__device__ __forceinline__ void zig_axis(int a, R o1, R o2, R qf, R &f1, R &f2, R &w1, R &w2,
                                         R &g, bool fwd_only)
{
    const R H = (R)0.5;
    const R imr = ((R)a + o1 + o2 + (R)1) * H;
    const R lo = (R)(a > 0 ? a : 0), hi = (R)(a < 0 ? a : 0) + (R)1;
    const R irr = imr < lo ? lo : (imr > hi ? hi : imr);
    g = (imr > lo && imr < hi) ? H : (R)0;     // d irr / d o1 = d irr / d o2
    f1 = qf * (irr - o1 - H);
    f2 = qf * ((R)a + o2 + H - irr);
    w1 = (irr + o1 + H) * H;
    w2 = (irr - (R)a + o2 + H) * H;
    (void)fwd_only;
}

}  // namespace

// ---------------------------------------------------------------- push
// One step of the particle adjoint.  State n of the particle (cell, offsets,
// momenta) is read from the tape; the adjoint of the state at n+1 is read
// from, and the adjoint at n written to, the xa..ua arrays (in place).
template <typename R, bool SPLIT, int NT>
__global__ __launch_bounds__(NT, NT == ADJ_THREADS ? 2 : 1) void push_adj_kernel(AdjParams P)
{
    constexpr int NW = NT / 64, ROUNDS = ADJ_TILE / NT;   // waves, ranks per thread
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int red[NW][4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int t;
    if (P.pass == 2) {
        if ((int)blockIdx.x >= P.big[0]) return;     // deferred tiles only
        t = P.big[1 + blockIdx.x];
    } else {
        t = xcd_remap(blockIdx.x, gridDim.x);
    }
    int s = 0;
    while (s + 1 < P.nspecies && t >= P.sp[s + 1].tile0) s++;
    const AdjSpecies S = P.sp[s];
    const int64_t tstart = (int64_t)(t - S.tile0) * ADJ_TILE;
    // Each thread takes 8 consecutive ranks of the sort order -- one cell at
    // the last sort, so while the particles stay there its node adjoints add
    // up in registers and reach LDS once per cell.  The 64 lanes of a wave
    // start 32 or 64 ranks apart, in different cells: their flushes hit
    // different LDS addresses, and their tape loads walk consecutive 16-byte
    // groups.  In push order (full 4096-particle tiles laid out by
    // tile_rank) this tile is half k of push tile T; ranks l*64 + 32k +
    // [0, 32) of it, lane l of wave w taking the 8 ranks from 32k + 8w.
    const int64_t T0 = (tstart / PUSH_TILE) * PUSH_TILE;
    const bool porder = P.tv > 0 && T0 + PUSH_TILE <= S.n;
    // the push's vector width of this precision (P.tv is it or 0): a
    // constant, so the rank -> slot map compiles to shifts
    constexpr int TV = sizeof(R) == 8 ? push_vec(PBPIC_PREC_FP64) : push_vec(PBPIC_PREC_SPLIT32);
    static_assert(ADJ_TILE == 64 * 32 && NW * ROUNDS == 32,
                  "lane/wave rank map: the waves' ranks fill each 32-rank row");
    auto slot_of = [&](int r) -> int64_t {
        if (!porder) return tstart + (int64_t)lane * 32 + wave * ROUNDS + r;
        const int k = (int)((tstart - T0) / ADJ_TILE);
        const int rank = lane * 64 + 32 * k + wave * ROUNDS + r;
        const int l = rank >> 6, c = (rank & 63) / TV, j = rank % TV;
        return T0 + (int64_t)c * 64 * TV + l * TV + j;
    };
    const int nx = P.nx, ny = P.ny, nxh = nx >> 1, nyh = ny >> 1;

    const int c00 = S.cell[tstart];
    const bool c00ok = (unsigned)c00 < (unsigned)(nx * ny);
    const int x0 = c00ok ? (int)P.dny.div31((uint32_t)c00) : 0;
    const int y0 = c00ok ? c00 - x0 * ny : 0;
    int cl[ROUNDS];
    int mnx = 1 << 30, mxx = -(1 << 30), mny = 1 << 30, mxy = -(1 << 30);
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const int64_t i = slot_of(r);
        cl[r] = -1;
        if (i < S.n) {
            const int c = S.cell[i];
            if ((unsigned)c >= (unsigned)(nx * ny)) continue;
            cl[r] = c;
            const int cx = (int)P.dny.div31((uint32_t)c), cy = c - cx * ny;
            int rx = cx - x0, ry = cy - y0;
            if (rx > nxh) rx -= nx; else if (rx < -nxh) rx += nx;
            if (ry > nyh) ry -= ny; else if (ry < -nyh) ry += ny;
            mnx = min(mnx, rx); mxx = max(mxx, rx);
            mny = min(mny, ry); mxy = max(mxy, ry);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mnx = min(mnx, __shfl_xor(mnx, o));
        mxx = max(mxx, __shfl_xor(mxx, o));
        mny = min(mny, __shfl_xor(mny, o));
        mxy = max(mxy, __shfl_xor(mxy, o));
    }
    if (lane == 0) { red[wave][0] = mnx; red[wave][1] = mxx; red[wave][2] = mny; red[wave][3] = mxy; }
    __syncthreads();
    mnx = red[0][0]; mxx = red[0][1]; mny = red[0][2]; mxy = red[0][3];
#pragma unroll
    for (int w = 1; w < NW; w++) {
        mnx = min(mnx, red[w][0]); mxx = max(mxx, red[w][1]);
        mny = min(mny, red[w][2]); mxy = max(mxy, red[w][3]);
    }
    // nodes used: the cell's 4 corners -> window = box + 1 node
    const int wx0 = mnx, wy0 = mny;
    const int W = mxx - mnx + 2, H = mxy - mny + 2;
    const bool inlds = mnx <= mxx && (int64_t)W * H <= P.maxw;
    if (P.pass == 1 && !inlds && mnx <= mxx && (int64_t)W * H <= P.maxw_big) {
        if (tid == 0) P.big[1 + atomicAdd(P.big, 1)] = t;   // block-uniform exit
        return;
    }
    // the current adjoint at the deposit targets, which reach one node
    // below/left and two above/right of a cell: the box grown by one node on
    // each side, staged when it fits the rest of the LDS (else read from L2)
    const int JW = W + 2;
    const bool jlds = inlds && (int64_t)JW * (H + 2) <= P.jmax;
    const int JWH = jlds ? JW * (H + 2) : 0;
    const int WH = inlds ? W * H : 0;
    const bool interior = mnx <= mxx && x0 + mnx >= 1 && x0 + mxx <= nx - 3 && y0 + mny >= 1 &&
                          y0 + mxy <= ny - 3;
    R *nf = (R *)smem;                      // [WH][6] node fields
    // [6][WH] node adjoint accumulators, FP64 in both forms: LDS FP64 atomic
    // adds run at about ten times the rate of FP32 ones on this hardware
    double *na = (double *)(smem + adj_nf_bytes(sizeof(R), P.maxw));
    // [3][jmax] current adjoint at the deposit targets, in the push precision
    R *jw = (R *)(na + (size_t)6 * P.maxw);
    const int jmax = P.jmax;
    if (inlds) {
        // node fields from the node-field array (y fastest: lanes walk a
        // window column, reading consecutive node records); the same values
        // as a_node_fields rounded to R
        for (int k = tid; k < WH; k += NT) {
            const int wi = k / H, wj = k - wi * H;
            const size_t g = (size_t)wrap_index(x0 + wx0 + wi, nx) * ny + wrap_index(y0 + wy0 + wj, ny);
            R *o = nf + (size_t)(wj * W + wi) * 6;
            if constexpr (sizeof(R) == 4) {
                const float4 *src = reinterpret_cast<const float4 *>(P.nf) + 2 * g;
                const float4 a = src[0], b = src[1];
                o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w; o[4] = b.x; o[5] = b.y;
            } else {
                const double2 *src = reinterpret_cast<const double2 *>(P.nf) + 3 * g;
                const double2 a = src[0], b = src[1], cc = src[2];
                o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y; o[4] = cc.x; o[5] = cc.y;
            }
        }
        for (int k = tid; k < 6 * WH; k += NT) na[k] = 0.0;
        for (int k = tid; k < JWH; k += NT) {
            const int wj = k / JW, wi = k - wj * JW;
            const size_t g = (size_t)wrap_index(y0 + wy0 - 1 + wj, ny) * nx +
                             wrap_index(x0 + wx0 - 1 + wi, nx);
            jw[k] = (R)P.jadj[0][g];
            jw[jmax + k] = (R)P.jadj[1][g];
            jw[2 * jmax + k] = (R)P.jadj[2][g];
        }
    }
    __syncthreads();

    const R c0 = (R)S.c0, c1 = (R)S.c1, qx = (R)S.qx, qy = (R)S.qy, qr = (R)S.q;
    const R dtcva = (R)P.dtcva, idx = (R)P.idx, idy = (R)P.idy, cva = (R)P.cva;
    const bool oldu = (P.compat & PBPIC_COMPAT_BORIS_OLD_U) != 0;
    const double *ja = P.jadj[0], *jb = P.jadj[1], *jc = P.jadj[2];
    double *nadj = P.nadj;
    const size_t NC = (size_t)nx * ny;
    // node adjoints of the current cell, flushed when the cell changes
    R nacc[6][4];
    int acell = -1, abase = 0;
    size_t ag[4] = {0, 0, 0, 0};
    auto flush = [&]() {
        if (acell < 0) return;
        if (inlds) {
#pragma unroll
            for (int q = 0; q < 6; q++) {
                double *pq = na + (size_t)q * WH;
                a_lds_add(pq + abase, (double)nacc[q][0]);
                a_lds_add(pq + abase + 1, (double)nacc[q][1]);
                a_lds_add(pq + abase + W, (double)nacc[q][2]);
                a_lds_add(pq + abase + W + 1, (double)nacc[q][3]);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 6; q++) {
                double *pq = nadj + q * NC;
#pragma unroll
                for (int m = 0; m < 4; m++) unsafeAtomicAdd(pq + ag[m], (double)nacc[q][m]);
            }
        }
        acell = -1;
    };
#pragma unroll 1
    for (int r = 0; r < ROUNDS; r++) {
        const int c = cl[r];
        if (c < 0) continue;
        const int64_t i = slot_of(r);
        const R ox = ((const R *)S.ox)[i], oy = ((const R *)S.oy)[i];
        const R u[3] = {((const R *)S.ux)[i], ((const R *)S.uy)[i], ((const R *)S.uz)[i]};
        const int cx = (int)P.dny.div31((uint32_t)c), cy = c - cx * ny;
        int rx = cx - x0, ry = cy - y0;
        if (rx > nxh) rx -= nx; else if (rx < -nxh) rx += nx;
        if (ry > nyh) ry -= ny; else if (ry < -nyh) ry += ny;
        const int lx = rx - wx0, ly = ry - wy0;
        const int cx1 = cx + 1 == nx ? 0 : cx + 1, cy1 = cy + 1 == ny ? 0 : cy + 1;
        R n00[6], n10[6], n01[6], n11[6];
        if (inlds) {
            const R *b = nf + (size_t)(ly * W + lx) * 6;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                n00[q] = b[q]; n10[q] = b[6 + q];
                n01[q] = b[(size_t)W * 6 + q]; n11[q] = b[(size_t)W * 6 + 6 + q];
            }
        } else {
            double f[6];
            a_node_fields(P, cx, cy, f);  for (int q = 0; q < 6; q++) n00[q] = (R)f[q];
            a_node_fields(P, cx1, cy, f); for (int q = 0; q < 6; q++) n10[q] = (R)f[q];
            a_node_fields(P, cx, cy1, f); for (int q = 0; q < 6; q++) n01[q] = (R)f[q];
            a_node_fields(P, cx1, cy1, f); for (int q = 0; q < 6; q++) n11[q] = (R)f[q];
        }
        // ---- forward recompute: gather (:948-963), same arithmetic as the push
        const R Hh = (R)0.5, ONE = (R)1;
        const R wxp = Hh + ox, wxm = Hh - ox, wyp = Hh + oy, wym = Hh - oy;
        const R s11 = (Hh + ox) * (Hh + oy);
        const R s10 = Hh + ox - s11;
        const R s01 = Hh + oy - s11;
        const R s00 = ONE - s11 - s10 - s01;
        R F[6];
#pragma unroll
        for (int q = 0; q < 6; q++) F[q] = n00[q] * s00 + n10[q] * s10 + n01[q] * s01 + n11[q] * s11;
        const R E[3] = {F[0], F[1], F[2]}, Bf[3] = {F[3], F[4], F[5]};
        // Boris (:964-1000)
        R um[3], t[3], sv[3], u0[3], up[3], un[3];
        for (int k = 0; k < 3; k++) um[k] = u[k] + c1 * E[k];
        // 1/gamma and 2/(1+t.t): IEEE in FP64; in the split form the
        // hardware reciprocal (square root) the forward split push uses
        const R q1 = ONE + um[0] * um[0] + um[1] * um[1] + um[2] * um[2];
        R g1, ig1;
        if constexpr (SPLIT) { ig1 = __builtin_amdgcn_rsqf(q1); g1 = q1 * ig1; }
        else { g1 = sqrt(q1); ig1 = ONE / g1; }
        const R c2 = SPLIT ? c0 * ig1 : c0 / g1;
        for (int k = 0; k < 3; k++) t[k] = c2 * Bf[k];
        const R qt = ONE + t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        R t2;
        if constexpr (SPLIT) t2 = (R)2 * __builtin_amdgcn_rcpf(qt);
        else t2 = (R)2 / qt;
        for (int k = 0; k < 3; k++) sv[k] = t2 * t[k];
        R uq[3];
        for (int k = 0; k < 3; k++) uq[k] = oldu ? u[k] : um[k];
        R cr[3];
        cross3(uq, t, cr);
        for (int k = 0; k < 3; k++) u0[k] = um[k] + cr[k];
        cross3(u0, sv, cr);
        for (int k = 0; k < 3; k++) up[k] = um[k] + cr[k];
        for (int k = 0; k < 3; k++) un[k] = up[k] + c1 * E[k];
        const R q2 = ONE + un[0] * un[0] + un[1] * un[1] + un[2] * un[2];
        R gn, ig;
        if constexpr (SPLIT) { ig = __builtin_amdgcn_rsqf(q2); gn = q2 * ig; }
        else { gn = sqrt(q2); ig = ONE / gn; }
        const R v[3] = {un[0] * ig, un[1] * ig, un[2] * ig};
        // move (:1002-1026)
        const R c3 = SPLIT ? dtcva * ig : dtcva / gn;
        const R dx1 = ox + c3 * un[0] * idx;
        const R dy1 = oy + c3 * un[1] * idy;
        const int a = fabs(dx1) > Hh ? (int)a_sgn(dx1) : 0;
        int b = fabs(dy1) > Hh ? (int)a_sgn(dy1) : 0;
        const R ox2 = dx1 + (R)(-a);
        R oy2 = dy1 + (R)(-b);
        // conducting walls: reflected before the deposit (offset mirrored,
        // u_y reversed), so the deposit saw b = 0 and -oy2
        const bool refl = P.nyp && ((cy == 0 && b < 0) || (cy == P.nyp - 1 && b > 0));
        if (refl) { b = 0; oy2 = -oy2; }
        const R vz = SPLIT ? un[2] * cva * ig : un[2] * cva / gn;
        // deposit terms (relative form of :1237-1282)
        R fx1, fx2, wx1, wx2, gx, fy1, fy2, wy1, wy2, gy;
        zig_axis<R>(a, ox, ox2, qx, fx1, fx2, wx1, wx2, gx, false);
        zig_axis<R>(b, oy, oy2, qy, fy1, fy2, wy1, wy2, gy, false);
        const R dimid = ((R)a + ox + ox2) * Hh, djmid = ((R)b + oy + oy2) * Hh;
        const int di = (int)a_sgn(dimid), dj = (int)a_sgn(djmid);
        const R A = fabs(dimid), Bq = fabs(djmid);
        const R quz = qr * vz;
        // ---- reverse: deposit = gather of the current adjoint at the targets
        // deposit targets lie within one cell below/left and two above/right
        // of the cell: no periodic wrap for a tile away from every edge
        const bool drop = (P.compat & PBPIC_COMPAT_JY_DROP) && cx == nx - 1 && a == 1;
        R l1, l2, l3, l4, l5, l6, l7, l8, z0, z1, z2, z3;
        if (jlds) {
            // window node of target (cx + ddx, cy + ddy): (lx + 1 + ddx, ly + 1 + ddy)
            const R *wa = jw + (ly + 1) * JW + lx + 1, *wb = wa + jmax, *wc = wb + jmax;
            const int ab = b * JW + a;
            l1 = wa[0]; l2 = wa[JW]; l3 = wb[0]; l4 = wb[1];
            l5 = wa[ab]; l6 = wa[ab + JW]; l7 = wb[ab];
            l8 = drop ? (R)0 : wb[ab + 1];
            z0 = wc[0]; z1 = wc[di]; z2 = wc[dj * JW]; z3 = wc[dj * JW + di];
        } else {
            const size_t g0 = (size_t)cy * nx + cx;
            auto at = [&](int ddx, int ddy) -> size_t {
                if (interior) return g0 + (ptrdiff_t)ddy * nx + ddx;
                return (size_t)wrap_index(cy + ddy, ny) * nx + wrap_index(cx + ddx, nx);
            };
            l1 = (R)ja[at(0, 0)]; l2 = (R)ja[at(0, 1)];
            l3 = (R)jb[at(0, 0)]; l4 = (R)jb[at(1, 0)];
            l5 = (R)ja[at(a, b)]; l6 = (R)ja[at(a, b + 1)];
            l7 = (R)jb[at(a, b)];
            l8 = drop ? (R)0 : (R)jb[at(a + 1, b)];
            z0 = (R)jc[at(0, 0)]; z1 = (R)jc[at(di, 0)]; z2 = (R)jc[at(0, dj)];
            z3 = (R)jc[at(di, dj)];
        }
        const R bfx1 = l1 * (ONE - wy1) + l2 * wy1, bwy1 = fx1 * (l2 - l1);
        const R bfy1 = l3 * (ONE - wx1) + l4 * wx1, bwx1 = fy1 * (l4 - l3);
        const R bfx2 = l5 * (ONE - wy2) + l6 * wy2, bwy2 = fx2 * (l6 - l5);
        const R bfy2 = l7 * (ONE - wx2) + l8 * wx2, bwx2 = fy2 * (l8 - l7);
        const R birx = qx * bfx1 - qx * bfx2 + Hh * bwx1 + Hh * bwx2;
        const R biry = qy * bfy1 - qy * bfy2 + Hh * bwy1 + Hh * bwy2;
        R bo1x = -qx * bfx1 + Hh * bwx1 + birx * gx;
        R bo2x = qx * bfx2 + Hh * bwx2 + birx * gx;
        R bo1y = -qy * bfy1 + Hh * bwy1 + biry * gy;
        R bo2y = qy * bfy2 + Hh * bwy2 + biry * gy;
        // Jz: s00=(1-A)(1-B) at (0,0), (di,0): A(1-B), (0,dj): (1-A)B, (di,dj): AB
        const R bquz = z0 * (ONE - A) * (ONE - Bq) + z1 * A * (ONE - Bq) + z2 * (ONE - A) * Bq +
                       z3 * A * Bq;
        const R bA = quz * ((z1 - z0) * (ONE - Bq) + (z3 - z2) * Bq);
        const R bB = quz * ((z2 - z0) * (ONE - A) + (z3 - z1) * A);
        const R bdim = bA * (R)di, bdjm = bB * (R)dj;
        bo1x += Hh * bdim; bo2x += Hh * bdim;
        bo1y += Hh * bdjm; bo2y += Hh * bdjm;
        const R bvz = qr * bquz;
        // ---- move: new offset = old + dt c u_x/(gamma dx) - a
        const int64_t ia = i;
        R xa = ((R *)P.xa[s])[ia], ya = ((R *)P.ya[s])[ia];
        R ua[3] = {((R *)P.uxa[s])[ia], ((R *)P.uya[s])[ia], ((R *)P.uza[s])[ia]};
        const R xn = xa + bo2x;                         // adjoint of the new offsets
        const R yn = refl ? -(ya + bo2y) : ya + bo2y;   // (before a reflection)
        R xo = xn + bo1x, yo = yn + bo1y;               // adjoint of the old offsets
        // Q_i = k_i un_i / gamma: dQ_i/dun_j = k_i (delta_ij - v_i v_j)/gamma
        const R kk[3] = {xn * dtcva * idx, yn * dtcva * idy, bvz * cva};
        const R vk = v[0] * kk[0] + v[1] * kk[1] + v[2] * kk[2];
        if (refl) ua[1] = -ua[1];                       // stored u_y = -un_y
        R bun[3];
        for (int k = 0; k < 3; k++) bun[k] = ua[k] + ig * (kk[k] - v[k] * vk);
        // ---- Boris reverse
        R bE[3], bB3[3], bu[3], bum[3], bu0[3], bs[3], bt[3], buq[3];
        for (int k = 0; k < 3; k++) { bE[k] = c1 * bun[k]; bum[k] = bun[k]; }
        // up = um + u0 x s
        cross3(sv, bun, bu0);            // u0_bar = s x up_bar
        cross3(bun, u0, bs);             // s_bar  = up_bar x u0
        // u0 = um + uq x t
        for (int k = 0; k < 3; k++) bum[k] += bu0[k];
        cross3(t, bu0, buq);             // uq_bar = t x u0_bar
        cross3(bu0, uq, bt);             // t_bar  = u0_bar x uq
        // s = t2 t, t2 = 2/(1+t.t)
        const R bt2 = bs[0] * t[0] + bs[1] * t[1] + bs[2] * t[2];
        for (int k = 0; k < 3; k++) bt[k] += t2 * bs[k] - t2 * t2 * bt2 * t[k];
        // t = c2 B, c2 = c0/g1, g1 = sqrt(1+um.um)
        const R bc2 = bt[0] * Bf[0] + bt[1] * Bf[1] + bt[2] * Bf[2];
        for (int k = 0; k < 3; k++) bB3[k] = c2 * bt[k];
        const R bg1 = SPLIT ? -c2 * ig1 * bc2 : -c2 / g1 * bc2;
        for (int k = 0; k < 3; k++) {
            if (oldu) bu[k] = buq[k];
            else { bu[k] = 0; bum[k] += buq[k]; }
        }
        for (int k = 0; k < 3; k++) bum[k] += SPLIT ? bg1 * um[k] * ig1 : bg1 * um[k] / g1;
        for (int k = 0; k < 3; k++) { bu[k] += bum[k]; bE[k] += c1 * bum[k]; }
        // ---- gather reverse: force adjoint -> nodes; weights -> position
        const R bF[6] = {bE[0], bE[1], bE[2], bB3[0], bB3[1], bB3[2]};
        R dFx = 0, dFy = 0;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            dFx += bF[q] * (wym * (n10[q] - n00[q]) + wyp * (n11[q] - n01[q]));
            dFy += bF[q] * (wxm * (n01[q] - n00[q]) + wxp * (n11[q] - n10[q]));
        }
        xo += dFx;
        yo += dFy;
        if constexpr (SPLIT) {
            if (c != acell) {
                flush();
                acell = c;
                abase = ly * W + lx;
                ag[0] = (size_t)cy * nx + cx; ag[1] = (size_t)cy * nx + cx1;
                ag[2] = (size_t)cy1 * nx + cx; ag[3] = (size_t)cy1 * nx + cx1;
#pragma unroll
                for (int q = 0; q < 6; q++)
#pragma unroll
                    for (int m = 0; m < 4; m++) nacc[q][m] = (R)0;
            }
#pragma unroll
            for (int q = 0; q < 6; q++) {
                nacc[q][0] += s00 * bF[q];
                nacc[q][1] += s10 * bF[q];
                nacc[q][2] += s01 * bF[q];
                nacc[q][3] += s11 * bF[q];
            }
        } else if (inlds) {
            // FP64: each particle's node adjoints go straight to the window
            // (register sums per cell would need 48 more registers and spill)
            const int base = ly * W + lx;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                double *pq = na + (size_t)q * WH;
                a_lds_add(pq + base, (double)(s00 * bF[q]));
                a_lds_add(pq + base + 1, (double)(s10 * bF[q]));
                a_lds_add(pq + base + W, (double)(s01 * bF[q]));
                a_lds_add(pq + base + W + 1, (double)(s11 * bF[q]));
            }
        } else {
            const size_t g[4] = {(size_t)cy * nx + cx, (size_t)cy * nx + cx1,
                                 (size_t)cy1 * nx + cx, (size_t)cy1 * nx + cx1};
            const R w[4] = {s00, s10, s01, s11};
#pragma unroll
            for (int q = 0; q < 6; q++)
#pragma unroll
                for (int m = 0; m < 4; m++)
                    unsafeAtomicAdd(nadj + q * NC + g[m], (double)(w[m] * bF[q]));
        }
        ((R *)P.xa[s])[ia] = xo;
        ((R *)P.ya[s])[ia] = yo;
        ((R *)P.uxa[s])[ia] = bu[0];
        ((R *)P.uya[s])[ia] = bu[1];
        ((R *)P.uza[s])[ia] = bu[2];
    }
    flush();
    __syncthreads();
    if (inlds) {
        for (int k = tid; k < WH; k += NT) {
            const int wj = k / W, wi = k - wj * W;
            const size_t g = (size_t)wrap_index(y0 + wy0 + wj, ny) * nx + wrap_index(x0 + wx0 + wi, nx);
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const double vq = na[(size_t)q * WH + k];
                if (vq != 0.0) unsafeAtomicAdd(nadj + q * NC + g, vq);
            }
        }
    }
}

// ---------------------------------------------------------------- fields
// Curl operators of the step, on compact periodic planes:
//   CE(E): Cx = dEz/dy, Cy = -dEz/dx, Cz = -(dEx/dy - dEy/dx)    (B -= dth CE)
//   CB(B): Dx = dBz/dy, Dy = -dBz/dx, Dz = dBy/dx - dBx/dy       (E += dtc2 CB)
// with forward differences toward +1 for dEx/dy, dEy/dx, dBy/dx, dBx/dy and
// backward differences for dEz/d*, dBz/d* (pic2d.f90:733-736, 790-793).
// The transposes below swap the difference direction and the sign.
__device__ __forceinline__ size_t fi(int x, int y, int nx, int ny)
{
    x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
    y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
    return (size_t)y * nx + x;
}

// CE^T(Bbar) -> (ex, ey, ez) contributions at (x, y)
__device__ __forceinline__ void ceT(const double *bx, const double *by, const double *bz, int x,
                                    int y, int nx, int ny, double idx, double idy, double o[3])
{
    const size_t c = fi(x, y, nx, ny);
    o[0] = -idy * (bz[fi(x, y - 1, nx, ny)] - bz[c]);
    o[1] = idx * (bz[fi(x - 1, y, nx, ny)] - bz[c]);
    o[2] = idy * (bx[c] - bx[fi(x, y + 1, nx, ny)]) - idx * (by[c] - by[fi(x + 1, y, nx, ny)]);
}

// CB^T(Ebar) -> (bx, by, bz) contributions at (x, y)
__device__ __forceinline__ void cbT(const double *ex, const double *ey, const double *ez, int x,
                                    int y, int nx, int ny, double idx, double idy, double o[3])
{
    const size_t c = fi(x, y, nx, ny);
    o[0] = -idy * (ez[fi(x, y - 1, nx, ny)] - ez[c]);
    o[1] = idx * (ez[fi(x - 1, y, nx, ny)] - ez[c]);
    o[2] = idy * (ex[c] - ex[fi(x, y + 1, nx, ny)]) - idx * (ey[c] - ey[fi(x + 1, y, nx, ny)]);
}

// Pass A: Ebar' = Ebar^{n+1} + (w/2) J - dth CE^T(Bbar^{n+1});
//         Jbar  = -dtc2 in0 Ebar' + (w/2)(E^n + E^{n+1}).
// With conducting walls (mode 1 then 2) the two halves run apart: Ebar' goes
// through the transpose of the forward's wall projection before Jbar.
template <typename RJ>
__global__ void fdtd_adj_a(AdjFieldParams P, int mode)
{
    const int64_t n = (int64_t)P.nx * P.ny;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(k / P.nx), x = (int)(k - (int64_t)y * P.nx);
        double o[3];
        ceT(P.bxa, P.bya, P.bza, x, y, P.nx, P.ny, P.idx, P.idy, o);
        const RJ *J[3] = {(const RJ *)P.jx, (const RJ *)P.jy, (const RJ *)P.jz};
        double *Ea[3] = {P.exa, P.eya, P.eza};
        const double *E0[3] = {P.e0x, P.e0y, P.e0z}, *E1[3] = {P.e1x, P.e1y, P.e1z};
        double *Jb[3] = {P.jxa, P.jya, P.jza};
        double *Ep[3] = {P.etx, P.ety, P.etz};
        const double wm = P.lmask ? P.w * P.lmask[k] : P.w;   // loss weight of this cell
        for (int q = 0; q < 3; q++) {
            double e;
            if (mode != 2) {
                const double jr = (double)J[q][k];
                e = Ea[q][k] + 0.5 * wm * jr - P.dth * o[q];
                Ep[q][k] = e;
            } else {
                e = Ep[q][k];
            }
            if (mode != 1) Jb[q][k] = -P.dtc2 * P.in0 * e + 0.5 * wm * (E0[q][k] + E1[q][k]);
        }
    }
}

// Conducting walls, per column gx.  Forward maps and their transposes:
//   imaging W (after the solve): guard rows of the half-row planes <- +-
//     mirror (Ey, Bz even; Ez, By odd), Ex <- 0 on the walls and Ex, Bx of
//     row nyp+1 <- -Ex, +Bx of row 1; transpose: the guard adjoint is added
//     (with the sign) to the mirror and cleared, the wall Ex adjoint cleared;
//   projection P (inside the solve, on E'): Ex <- 0 on the walls, Ez guards
//     <- -mirror; transpose likewise on Ebar';
//   Jz fold (after the push): row 0 -= row nyp+1, row nyp-1 -= row nyp,
//     guards <- 0; transpose: guard adjoint <- -(mirror adjoint).
__device__ __forceinline__ void half_T(double *a, size_t rt, size_t gt, size_t r0, size_t gb,
                                       double sgn)
{
    a[rt] += sgn * a[gt];
    a[r0] += sgn * a[gb];
    a[gt] = 0.0;
    a[gb] = 0.0;
}

// which: 0 = W^T on (exa..bza), 1 = P^T on (etx..etz), 2 = fold^T on jza
__global__ void wall_adj_kernel(AdjFieldParams P, int which)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x;
    if (gx >= P.nx) return;
    const int nx = P.nx, nyp = P.nyp;
    const size_t r0 = gx, r1 = (size_t)nx + gx, rt = (size_t)(nyp - 1) * nx + gx,
                 gt = (size_t)nyp * nx + gx, gb = (size_t)(nyp + 1) * nx + gx;
    if (which == 0) {
        half_T(P.eya, rt, gt, r0, gb, 1.0);
        half_T(P.eza, rt, gt, r0, gb, -1.0);
        half_T(P.bya, rt, gt, r0, gb, -1.0);
        half_T(P.bza, rt, gt, r0, gb, 1.0);
        P.exa[r1] -= P.exa[gb];
        P.exa[gb] = 0.0;
        P.exa[r0] = 0.0;
        P.exa[gt] = 0.0;
        P.bxa[r1] += P.bxa[gb];
        P.bxa[gb] = 0.0;
    } else if (which == 1) {
        P.etx[r0] = 0.0;
        P.etx[gt] = 0.0;
        half_T(P.etz, rt, gt, r0, gb, -1.0);
    } else {
        P.jza[gb] = -P.jza[r0];
        P.jza[gt] = -P.jza[rt];
    }
}

// Pass B: Bbar_half = Bbar^{n+1} + dtc2 CB^T(Ebar') (in place on Bbar).
__global__ void fdtd_adj_b(AdjFieldParams P)
{
    const int64_t n = (int64_t)P.nx * P.ny;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(k / P.nx), x = (int)(k - (int64_t)y * P.nx);
        double o[3];
        cbT(P.etx, P.ety, P.etz, x, y, P.nx, P.ny, P.idx, P.idy, o);
        P.bxa[k] += P.dtc2 * o[0];
        P.bya[k] += P.dtc2 * o[1];
        P.bza[k] += P.dtc2 * o[2];
    }
}

// Pass C: Ebar^n = Ebar' - dth CE^T(Bbar_half) + (w/2) J.
template <typename RJ>
__global__ void fdtd_adj_c(AdjFieldParams P)
{
    const int64_t n = (int64_t)P.nx * P.ny;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(k / P.nx), x = (int)(k - (int64_t)y * P.nx);
        double o[3];
        ceT(P.bxa, P.bya, P.bza, x, y, P.nx, P.ny, P.idx, P.idy, o);
        const RJ *J[3] = {(const RJ *)P.jx, (const RJ *)P.jy, (const RJ *)P.jz};
        double *Ea[3] = {P.exa, P.eya, P.eza};
        const double *Ep[3] = {P.etx, P.ety, P.etz};
        const double wm = P.lmask ? P.w * P.lmask[k] : P.w;
        for (int q = 0; q < 3; q++) Ea[q][k] = Ep[q][k] - P.dth * o[q] + 0.5 * wm * (double)J[q][k];
    }
}

// Transpose of the node averaging (emfgrid, pic2d.f90:837-845): node adjoint
// planes -> staggered field adjoints; clears the node planes afterwards.
__global__ void node_transpose_kernel(AdjFieldParams P, double *nadj)
{
    const int nx = P.nx, ny = P.ny;
    const int64_t n = (int64_t)nx * ny;
    double *dst[6] = {P.exa, P.eya, P.eza, P.bxa, P.bya, P.bza};
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(k / nx), x = (int)(k - (int64_t)y * nx);
        const size_t xp = fi(x + 1, y, nx, ny), yp = fi(x, y + 1, nx, ny), xyp = fi(x + 1, y + 1, nx, ny);
        for (int c = 0; c < 2; c++) {
            const double *a0 = nadj + (3 * c + 0) * n, *a1 = nadj + (3 * c + 1) * n,
                         *a2 = nadj + (3 * c + 2) * n;
            dst[3 * c + 0][k] += 0.5 * (a0[k] + a0[xp]);
            dst[3 * c + 1][k] += 0.5 * (a1[k] + a1[yp]);
            dst[3 * c + 2][k] += 0.25 * (a2[k] + a2[xp] + a2[yp] + a2[xyp]);
        }
    }
}

__global__ void zero_f64(double *p, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

// Reverse of a sort: pre[perm[k]] = post[k] for the five adjoint arrays.
struct Ptr5 {
    void *p[5];
};

// Lanes walk the sorted ranks, not the slots: in the push order (tv > 0)
// consecutive slots of a full tile hold ranks 64 apart, in different cells,
// while consecutive ranks of a cell came from nearby sources, so taking the
// ranks in order makes neighbouring lanes' scattered stores land close
// together.
template <typename R>
__global__ void unpermute_kernel(const uint32_t *perm, int64_t n, int tv, Ptr5 post, Ptr5 pre)
{
    const int64_t r = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int64_t k = r;
    if (tv) {
        const int64_t t0 = r & ~(int64_t)(PUSH_TILE - 1);
        if (t0 + PUSH_TILE <= n) {
            // the slot holding rank q of the tile (tile_rank's inverse)
            const int q = (int)(r - t0), l = q >> 6, c = (q & 63) / tv, j = q % tv;
            k = t0 + (int64_t)c * 64 * tv + l * tv + j;
        }
    }
    const uint32_t p = perm[k];
#pragma unroll
    for (int q = 0; q < 5; q++) ((R *)pre.p[q])[p] = ((const R *)post.p[q])[k];
}

size_t adj_lds_bytes(int precision, int maxw, int jmax)
{
    const size_t r = precision == PBPIC_PREC_FP64 ? 8 : 4;
    return adj_nf_bytes(r, maxw) + (size_t)maxw * 6 * 8 + 3 * r * (size_t)jmax;
}

// Current-adjoint window (nodes) that fills the LDS a workgroup may take
// (`budget`, static arrays included) beside node windows of maxw nodes.
int adj_jmax_for(int precision, int maxw, size_t budget)
{
    const size_t r = precision == PBPIC_PREC_FP64 ? 8 : 4, used = adj_lds_bytes(precision, maxw, 0);
    const size_t stat = 128;                         // __shared__ red[8][4] at most
    return used + stat >= budget ? 0 : (int)((budget - used - stat) / (3 * r));
}

int64_t adj_tiles(int64_t n) { return (n + ADJ_TILE - 1) / ADJ_TILE; }

template <int NT>
static void launch_push_adj_nt(const AdjParams &p, int precision, int t, size_t lds, hipStream_t st)
{
    if (precision == PBPIC_PREC_FP64) {
        (void)hipFuncSetAttribute((const void *)push_adj_kernel<double, false, NT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL((push_adj_kernel<double, false, NT>), dim3(t), dim3(NT), lds, st, p);
    } else {
        (void)hipFuncSetAttribute((const void *)push_adj_kernel<float, true, NT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL((push_adj_kernel<float, true, NT>), dim3(t), dim3(NT), lds, st, p);
    }
}

// the first launch at two workgroups of four waves per CU, the second (pass
// 2, one workgroup per CU) with eight waves
static void launch_push_adj_once(const AdjParams &p, int precision, int t, hipStream_t st)
{
    const size_t lds = adj_lds_bytes(precision, p.maxw, p.jmax);
    if (p.pass == 2) launch_push_adj_nt<ADJ_THREADS_BIG>(p, precision, t, lds, st);
    else launch_push_adj_nt<ADJ_THREADS>(p, precision, t, lds, st);
}

hipError_t launch_push_adj(const AdjParams &p0, int precision, hipStream_t st)
{
    AdjParams p = p0;
    int t = 0;
    for (int s = 0; s < p.nspecies; s++) {
        p.sp[s].tile0 = t;
        t += (int)((p.sp[s].n + ADJ_TILE - 1) / ADJ_TILE);
    }
    if (t == 0) return hipSuccess;
    // the second pass gets the window of one workgroup per CU (150 KiB)
    // the second pass: one workgroup per CU (150 KiB), node windows twice
    // the first pass's, the current-adjoint window in the rest
    const size_t cap = 150 * 1024;
    p.maxw_big = 2 * p.maxw;
    while (p.maxw_big > p.maxw && adj_lds_bytes(precision, p.maxw_big, 0) + 64 > cap) p.maxw_big--;
    const int jmax_big = adj_jmax_for(precision, p.maxw_big, cap);
    p.pass = p.big && p.maxw_big > p.maxw ? 1 : 0;
    if (p.pass) {
        hipError_t e = launch_fill_u32((uint32_t *)p.big, 1, 0u, st);
        if (e != hipSuccess) return e;
    }
    launch_push_adj_once(p, precision, t, st);
    if (p.pass) {
        AdjParams b = p;
        b.pass = 2;
        b.maxw = p.maxw_big;
        b.jmax = jmax_big;
        launch_push_adj_once(b, precision, t, st);
    }
    return hipGetLastError();
}

hipError_t launch_wall_images_adj(const AdjFieldParams &p, hipStream_t st)
{
    if (!p.nyp) return hipSuccess;
    hipLaunchKernelGGL(wall_adj_kernel, dim3((p.nx + 255) / 256), dim3(256), 0, st, p, 0);
    return hipGetLastError();
}

hipError_t launch_fdtd_adj(const AdjFieldParams &p, int precision, hipStream_t st)
{
    const int64_t n = (int64_t)p.nx * p.ny;
    int nb = (int)((n + 255) / 256);
    if (nb > 8192) nb = 8192;
    const dim3 wb((p.nx + 255) / 256);
    if (p.nyp) hipLaunchKernelGGL(wall_adj_kernel, wb, dim3(256), 0, st, p, 0);
    for (int mode : {p.nyp ? 1 : 0, p.nyp ? 2 : -1}) {
        if (mode < 0) break;
        if (precision == PBPIC_PREC_FP64)
            hipLaunchKernelGGL(fdtd_adj_a<double>, dim3(nb), dim3(256), 0, st, p, mode);
        else
            hipLaunchKernelGGL(fdtd_adj_a<float>, dim3(nb), dim3(256), 0, st, p, mode);
        if (mode == 1) hipLaunchKernelGGL(wall_adj_kernel, wb, dim3(256), 0, st, p, 1);
    }
    if (p.nyp) hipLaunchKernelGGL(wall_adj_kernel, wb, dim3(256), 0, st, p, 2);
    hipLaunchKernelGGL(fdtd_adj_b, dim3(nb), dim3(256), 0, st, p);
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL(fdtd_adj_c<double>, dim3(nb), dim3(256), 0, st, p);
    else
        hipLaunchKernelGGL(fdtd_adj_c<float>, dim3(nb), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_node_transpose(const AdjFieldParams &p, double *nadj, hipStream_t st)
{
    const int64_t n = (int64_t)p.nx * p.ny;
    int nb = (int)((n + 255) / 256);
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(node_transpose_kernel, dim3(nb), dim3(256), 0, st, p, nadj);
    hipLaunchKernelGGL(zero_f64, dim3(nb), dim3(256), 0, st, nadj, 6 * n);
    return hipGetLastError();
}

hipError_t launch_unpermute(const uint32_t *perm, int64_t n, int precision, int push_order,
                            void *const *post, void *const *pre, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((n + 255) / 256);
    // the sort's layout: full tiles in the push kernel's lane order
    const int tv = push_order && n >= PUSH_TILE ? push_vec(precision) : 0;
    Ptr5 a, b;
    for (int q = 0; q < 5; q++) { a.p[q] = post[q]; b.p[q] = pre[q]; }
    if (precision == PBPIC_PREC_FP64)
        hipLaunchKernelGGL(unpermute_kernel<double>, dim3(nb), dim3(256), 0, st, perm, n, tv, a, b);
    else
        hipLaunchKernelGGL(unpermute_kernel<float>, dim3(nb), dim3(256), 0, st, perm, n, tv, a, b);
    return hipGetLastError();
}

}  // namespace pbpic
