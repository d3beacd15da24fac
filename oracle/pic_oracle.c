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
