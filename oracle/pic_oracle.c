/*
 * pic_oracle.c -- CPU restatement of pbpic's 2D PIC step (TEST INFRASTRUCTURE).
 *
 * Checker only: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline.  Never part of the product path.  See pic_oracle.h.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pic_oracle.h"

/* ---------------- FP64 instantiation: the reference with real = real(8) */
#define REAL double
#define PFX(name) orc8_##name
#include "pic_oracle_impl.h"
#undef REAL
#undef PFX

/* ---------------- FP32 instantiation: the reference's default real */
#define REAL float
#define PFX(name) orc4_##name
#include "pic_oracle_impl.h"
#undef REAL
#undef PFX

/* ------------------------------------------------ Philox4x32-10 generator
 * Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11: ten rounds of two 32x32->64 multiplies with the published round
 * multipliers and Weyl key increments.  The device loader runs the same
 * integer arithmetic, so both produce identical 32-bit words.               */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* (u + 1/2) 2^-32: a uniform double in (0, 1) from a 32-bit word. */
static double u01(uint32_t u) { return ((double)u + 0.5) * 2.3283064365386963e-10; }

void orc_init_maxwellian(int nx, int ny, int ppc, double vth, double vdrift_x,
                         int counter_stream, uint64_t seed, uint32_t stream,
                         int32_t *grid, double *dx, double *dy, double *ux,
                         double *uy, double *uz)
{
    const int64_t n = (int64_t)nx * ny * ppc;
    const double uth = vth / sqrt(1.0 - vth * vth);
    const double ud = vdrift_x / sqrt(1.0 - vdrift_x * vdrift_x);
    const double twopi = 6.283185307179586;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    /* every particle depends only on its own counter: any thread split gives
       the same arrays */
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        uint32_t a[4], b[4];
        uint32_t ca[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 0u};
        uint32_t cb[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 1u};
        orc_philox4x32_10(ca, key, a);
        orc_philox4x32_10(cb, key, b);
        grid[p] = (int32_t)(p / ppc) + 1;
        dx[p] = u01(a[0]) - 0.5;
        dy[p] = u01(a[1]) - 0.5;
        double r1 = sqrt(-2.0 * log(u01(a[2]))), t1 = twopi * u01(a[3]);
        double r2 = sqrt(-2.0 * log(u01(b[0]))), t2 = twopi * u01(b[1]);
        double drift = ud;
        if (counter_stream && ((p % ppc) & 1)) drift = -ud;
        ux[p] = uth * (r1 * cos(t1)) + drift;
        uy[p] = uth * (r1 * sin(t1));
        uz[p] = uth * (r2 * cos(t2));
    }
}

/* ------------------------------------------------ Harris current sheets
 * Seeded restatement of initparticles / initfields (pic2d.f90:357-648) with
 * the device loader's word assignment.  Counter (p, stream, block) under key
 * seed: block 0 = (ion tangential position, electron tangential offset, ion
 * Box-Muller pair for u_x, u_y), block 1 = (electron pair for u_x, u_y, pair
 * for the thermal u_z of ion and electron), blocks 2..9 = four tries each of
 * the position across the sheet (a background particle takes the first).
 * prm = {L, nb, b0, psi0, vthe_cs, vthi_cs, vthe_bg, vthi_bg, vde_z, vdi_z}.
 * pi, pe: 5 x n arrays (dx, dy, ux, uy, uz) for ions and electrons; fields:
 * 6 planes nx*ny (Ex, Ey, Ez, Bx, By, Bz), x fastest, guard rows imaged.   */
static double uconv(double v) { return v / sqrt(1.0 - v * v); }

static void harris_cell(double X, int n, int *cidx, double *off)
{
    int c = (int)ceil(X) - 1;                     /* ix = ceiling(x/dx), :403 */
    if (c < 0) c = 0;
    if (c > n - 1) c = n - 1;                     /* :408-410 */
    *cidx = c;
    *off = X - ((double)c + 0.5);                 /* :420 */
}

int64_t orc_harris_count(int nx, int ny, int ywall, double dx, double dy, int geometry,
                         int ppc, const double *prm, int64_t *ncs)
{
    const int nyp = ywall ? ny - 2 : ny;
    const int64_t n = (int64_t)ppc * nx * nyp;
    const int nsh = geometry == 0 ? 2 : 1;
    const double ln = geometry == 0 ? nx * dx : nyp * dy;
    const double f = 2.0 * prm[0] * nsh / (2.0 * prm[0] * nsh + prm[1] * ln);
    int64_t k = (int64_t)ceil((double)n * f);     /* :375 */
    k = (k / 2) * 2;                              /* :376 */
    if (k > n) k = n - (n & 1);
    *ncs = k;
    return n;
}

void orc_init_harris(int nx, int ny, int ywall, double dx, double dy, int geometry, int ppc,
                     const double *prm, int thermal_z, uint64_t seed, uint32_t stream,
                     int32_t *grid_i, double *pi, int32_t *grid_e, double *pe,
                     double *fields)
{
    const int nyp = ywall ? ny - 2 : ny;
    int64_t ncs;
    const int64_t n = orc_harris_count(nx, ny, ywall, dx, dy, geometry, ppc, prm, &ncs);
    const double L = prm[0], b0 = prm[2], psi0 = prm[3];
    const double uec = uconv(prm[4]), uic = uconv(prm[5]);
    const double ueb = uconv(prm[6]), uib = uconv(prm[7]);
    const double ude = uconv(prm[8]), udi = uconv(prm[9]);
    const double twopi = 6.283185307179586;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const int nn = geometry == 0 ? nx : nyp, nt = geometry == 0 ? nyp : nx;
    const double dn = geometry == 0 ? dx : dy;
    /* each particle depends only on its own Philox counters */
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        uint32_t w[10][4];
        for (uint32_t b = 0; b < 10; b++) {
            uint32_t ctr[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, b};
            orc_philox4x32_10(ctr, key, w[b]);
        }
        const int sheet = p < ncs;
        const int second = sheet && geometry == 0 && p >= ncs / 2;
        double Xn;
        if (sheet) {
            const double xc = geometry == 0 ? (second ? 0.75 : 0.25) * (nx * dx)
                                            : 0.5 * (nyp * dy);
            Xn = -1.0;
            for (int t = 0; t < 32; t++) {
                const double u = u01(w[2 + t / 4][t % 4]);
                const double s = L * log(u / (1.0 - u)) / 2.0;       /* :401 */
                const double X = (s + xc) / dn;
                if (ceil(X) - 1 >= 0 && ceil(X) - 1 < nn) { Xn = X; break; }
            }
            if (Xn < 0) Xn = xc / dn;
        } else {
            Xn = u01(w[2][0]) * nn;
        }
        int cn, ct;
        double on, ot;
        harris_cell(Xn, nn, &cn, &on);
        harris_cell(u01(w[0][0]) * nt, nt, &ct, &ot);
        const double ote = u01(w[0][1]) - 0.5;                         /* :425 */
        const double r1 = sqrt(-2.0 * log(u01(w[0][2]))), t1 = twopi * u01(w[0][3]);
        const double r2 = sqrt(-2.0 * log(u01(w[1][0]))), t2 = twopi * u01(w[1][1]);
        const double r3 = sqrt(-2.0 * log(u01(w[1][2]))), t3 = twopi * u01(w[1][3]);
        const double ui = sheet ? uic : uib, ue = sheet ? uec : ueb;
        double dzi = sheet ? udi : 0.0, dze = sheet ? ude : 0.0;
        if (second) { dzi = -dzi; dze = -dze; }                      /* :491-497 */
        int cx, cy;
        double oxi, oyi, oxe, oye;
        if (geometry == 0) { cx = cn; cy = ct; oxi = oxe = on; oyi = ot; oye = ote; }
        else { cx = ct; cy = cn; oyi = oye = on; oxi = ot; oxe = ote; }
        const int32_t g = (int32_t)((int64_t)cx * ny + cy + 1);
        grid_i[p] = grid_e[p] = g;
        pi[p] = oxi; pi[n + p] = oyi;
        pe[p] = oxe; pe[n + p] = oye;
        pi[2 * n + p] = ui * (r1 * cos(t1));
        pi[3 * n + p] = ui * (r1 * sin(t1));
        pe[2 * n + p] = ue * (r2 * cos(t2));
        pe[3 * n + p] = ue * (r2 * sin(t2));
        pi[4 * n + p] = dzi + (thermal_z ? ui * (r3 * cos(t3)) : 0.0);
        pe[4 * n + p] = dze + (thermal_z ? ue * (r3 * sin(t3)) : 0.0);
    }
    /* fields: E = 0; B of the sheets (:597-614) */
    const size_t np = (size_t)nx * ny;
    memset(fields, 0, sizeof(double) * 6 * np);
    double *bx = fields + 3 * np, *by = fields + 4 * np;
    if (geometry == 0) {
        const double lx = nx * dx, xl = 0.25 * lx, xr = 0.75 * lx;
        for (int gx = 0; gx < nx; gx++) {
            const double x = gx * dx;
            const double v = b0 * (tanh((x - xl) / L) - tanh((x - xr) / L) - 1.0);
            for (int gy = 0; gy < ny; gy++) by[(size_t)gy * nx + gx] = v;
        }
    } else {
        const double lx = nx * dx, ly = nyp * dy, yc = 0.5 * ly;
#define PSI(x, y) (psi0 * cos(twopi * (x) / lx) * cos(0.5 * twopi * ((y) - yc) / ly))
        for (int gy = 0; gy <= nyp; gy++)           /* Bx on node rows, walls incl. */
            for (int gx = 0; gx < nx; gx++) {
                const double x = (gx + 0.5) * dx, y = gy * dy;
                bx[(size_t)gy * nx + gx] = b0 * tanh((y - yc) / L) -
                                           (PSI(x, y + 0.5 * dy) - PSI(x, y - 0.5 * dy)) / dy;
            }
        for (int gy = 0; gy < nyp; gy++)            /* By on half rows */
            for (int gx = 0; gx < nx; gx++) {
                const double x = gx * dx, y = (gy + 0.5) * dy;
                by[(size_t)gy * nx + gx] = (PSI(x + 0.5 * dx, y) - PSI(x - 0.5 * dx, y)) / dx;
            }
#undef PSI
        if (ywall)
            for (int gx = 0; gx < nx; gx++) {
                by[(size_t)nyp * nx + gx] = -by[(size_t)(nyp - 1) * nx + gx];
                by[(size_t)(nyp + 1) * nx + gx] = -by[gx];
                bx[(size_t)(nyp + 1) * nx + gx] = bx[(size_t)nx + gx];
            }
    }
}
