/*
 * pic_oracle_impl.h -- body of the CPU restatement (TEST INFRASTRUCTURE).
 * Included twice by pic_oracle.c with REAL = double / float and PFX(name) =
 * orc8_name / orc4_name.  Every routine cites the lines of
 * src/programs/pic2d.f90 (xiaocanli/pbpic) that it restates.
 *
 * Floating-point evaluation order follows the Fortran source left to right;
 * pic_oracle.c is compiled with -ffp-contract=off so no FMA is introduced.
 */

/* particle record: type particle (pic2d.f90:47-50) */
typedef struct {
    REAL dx, dy, ux, uy, uz;
    int32_t grid, q, pid;
} PFX(ptl_t);

typedef struct {
    int mx, my;
    /* module parameters (pic2d.f90:15-43) */
    REAL dx, dy, idx, idy, dt, dth, cva, cva2, mime, in0;
    int compat;
    int ywall;                 /* conducting walls in y (rows my-1, my: guards) */
    int marder, cherenkov;     /* the filters of :867-907 (off in the reference) */
    int nthreads;
    /* emf(0:mx+1,0:my+1) components (type fields, :52-57) */
    REAL *ex, *dex_dy, *ey, *dey_dx, *ez, *dez_dx, *dez_dy, *dive;
    REAL *bx, *dbx_dy, *by, *dby_dx, *bz, *dbz_dx, *dbz_dy;
    /* emfgrid(0:mx+1,0:my+1) (type fields_grid, :60-63) */
    REAL *gex, *gey, *gez, *gbx, *gby, *gbz;
    /* denjr(0:mx+2,0:my+2) (type density, :65-67, allocated :221) */
    REAL *jx, *jy, *jz, *rho;
    /* per-thread private current copies for the threaded pusher */
    REAL *tj;
    PFX(ptl_t) *ptl, *ptlsort;
    int64_t nptl;
    int32_t *ptlalloc;
    REAL totene[5];            /* totene(1:4,itime) of the current step */
    double ke64[2];            /* FP64 shadow of totene(1:2) */
    double jdote_step, jdote_int;
    double *gauss0;
    long steps;
} PFX(ctx_t);

#define CT PFX(ctx_t)
/* emf / emfgrid element (i,j), 0 <= i <= mx+1, 0 <= j <= my+1 */
#define F(a, i, j) ((a)[(size_t)(i) + (size_t)(c->mx + 2) * (size_t)(j)])
/* denjr element (i,j), 0 <= i <= mx+2, 0 <= j <= my+2 */
#define D(a, i, j) ((a)[(size_t)(i) + (size_t)(c->mx + 3) * (size_t)(j)])

static REAL PFX(sgn1)(REAL x) { return signbit(x) ? (REAL)-1 : (REAL)1; }
static REAL PFX(rmin)(REAL a, REAL b) { return b < a ? b : a; }
static REAL PFX(rmax)(REAL a, REAL b) { return b > a ? b : a; }

orc_ctx *PFX(create)(int nx, int ny, double dx, double dy, double dt,
                     double cva, double mime, double in0, int compat)
{
    CT *c = (CT *)calloc(1, sizeof(CT));
    if (!c) return NULL;
    c->mx = nx; c->my = ny;
    c->dx = (REAL)dx; c->dy = (REAL)dy;
    c->idx = (REAL)1 / c->dx;            /* idx = 1.0/dx (:16) */
    c->idy = (REAL)1 / c->dy;
    c->dt = (REAL)dt;
    c->dth = c->dt / (REAL)2;            /* dth = dt/2.0 (:43) */
    c->cva = (REAL)cva;
    c->cva2 = c->cva * c->cva;           /* cva2 = cva*cva (:21) */
    c->mime = (REAL)mime;
    c->in0 = (REAL)in0;
    c->compat = compat;
    c->nthreads = 1;
    size_t nf = (size_t)(nx + 2) * (size_t)(ny + 2);
    size_t nd = (size_t)(nx + 3) * (size_t)(ny + 3);
    REAL **fp[] = {&c->ex, &c->dex_dy, &c->ey, &c->dey_dx, &c->ez, &c->dez_dx,
                   &c->dez_dy, &c->dive, &c->bx, &c->dbx_dy, &c->by,
                   &c->dby_dx, &c->bz, &c->dbz_dx, &c->dbz_dy, &c->gex,
                   &c->gey, &c->gez, &c->gbx, &c->gby, &c->gbz};
    for (size_t k = 0; k < sizeof(fp) / sizeof(fp[0]); k++)
        *fp[k] = (REAL *)calloc(nf, sizeof(REAL));
    REAL **dp[] = {&c->jx, &c->jy, &c->jz, &c->rho};
    for (size_t k = 0; k < 4; k++) *dp[k] = (REAL *)calloc(nd, sizeof(REAL));
    c->ptlalloc = (int32_t *)calloc((size_t)nx * ny, sizeof(int32_t));
    return (orc_ctx *)c;
}

void PFX(destroy)(orc_ctx *oc)
{
    CT *c = (CT *)oc;
    if (!c) return;
    REAL *fp[] = {c->ex, c->dex_dy, c->ey, c->dey_dx, c->ez, c->dez_dx,
                  c->dez_dy, c->dive, c->bx, c->dbx_dy, c->by, c->dby_dx,
                  c->bz, c->dbz_dx, c->dbz_dy, c->gex, c->gey, c->gez, c->gbx,
                  c->gby, c->gbz, c->jx, c->jy, c->jz, c->rho, c->tj};
    for (size_t k = 0; k < sizeof(fp) / sizeof(fp[0]); k++) free(fp[k]);
    free(c->ptl); free(c->ptlsort); free(c->ptlalloc); free(c->gauss0);
    free(c);
}

/* y boundary: 0 periodic (the reference), 1 conducting walls (see
 * wall_image).  Call before loading fields or particles. */
int PFX(set_ybound)(orc_ctx *oc, int kind)
{
    CT *c = (CT *)oc;
    if (kind != 0 && kind != 1) return -1;
    if (kind == 1 && c->my < 4) return -1;
    c->ywall = kind;
    return 0;
}

void PFX(set_threads)(orc_ctx *oc, int nthreads)
{
    CT *c = (CT *)oc;
    c->nthreads = nthreads < 1 ? 1 : nthreads;
    free(c->tj);
    c->tj = NULL;
    if (c->nthreads > 1) {
        size_t nd = (size_t)(c->mx + 3) * (size_t)(c->my + 3);
        c->tj = (REAL *)calloc(nd * 4 * (size_t)c->nthreads, sizeof(REAL));
    }
}

int PFX(set_particles)(orc_ctx *oc, int64_t n, const int32_t *grid,
                       const double *dx, const double *dy, const double *ux,
                       const double *uy, const double *uz, const int32_t *q,
                       const int32_t *pid)
{
    CT *c = (CT *)oc;
    free(c->ptl); free(c->ptlsort);
    c->ptl = (PFX(ptl_t) *)malloc(sizeof(PFX(ptl_t)) * (size_t)(n > 0 ? n : 1));
    c->ptlsort = (PFX(ptl_t) *)malloc(sizeof(PFX(ptl_t)) * (size_t)(n > 0 ? n : 1));
    if (!c->ptl || !c->ptlsort) return -1;
    c->nptl = n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        PFX(ptl_t) *p = &c->ptl[i];
        p->dx = (REAL)dx[i]; p->dy = (REAL)dy[i];
        p->ux = (REAL)ux[i]; p->uy = (REAL)uy[i]; p->uz = (REAL)uz[i];
        p->grid = grid[i]; p->q = q[i];
        p->pid = pid ? pid[i] : (int32_t)(i + 1);
    }
    return 0;
}

int64_t PFX(num_particles)(orc_ctx *oc) { return ((CT *)oc)->nptl; }

void PFX(get_particles)(orc_ctx *oc, int32_t *grid, double *dx, double *dy,
                        double *ux, double *uy, double *uz, int32_t *q,
                        int32_t *pid)
{
    CT *c = (CT *)oc;
    for (int64_t i = 0; i < c->nptl; i++) {
        const PFX(ptl_t) *p = &c->ptl[i];
        if (grid) grid[i] = p->grid;
        if (dx) dx[i] = p->dx;
        if (dy) dy[i] = p->dy;
        if (ux) ux[i] = p->ux;
        if (uy) uy[i] = p->uy;
        if (uz) uz[i] = p->uz;
        if (q) q[i] = p->q;
        if (pid) pid[i] = p->pid;
    }
}

/* emfgrid from emf, pic2d.f90:837-845 (identical to :636-643). */
static void PFX(emfgrid)(CT *c)
{
    const int mx = c->mx, my = c->my;
    for (int j = 0; j <= my + 1; j++)
        for (int i = 1; i <= mx + 1; i++) {
            F(c->gex, i, j) = (F(c->ex, i, j) + F(c->ex, i - 1, j)) * (REAL)0.5;
            F(c->gbx, i, j) = (F(c->bx, i, j) + F(c->bx, i - 1, j)) * (REAL)0.5;
        }
    for (int j = 1; j <= my + 1; j++)
        for (int i = 0; i <= mx + 1; i++) {
            F(c->gey, i, j) = (F(c->ey, i, j) + F(c->ey, i, j - 1)) * (REAL)0.5;
            F(c->gby, i, j) = (F(c->by, i, j) + F(c->by, i, j - 1)) * (REAL)0.5;
        }
    for (int j = 1; j <= my + 1; j++)
        for (int i = 1; i <= mx + 1; i++) {
            F(c->gez, i, j) = (F(c->ez, i, j) + F(c->ez, i - 1, j) +
                               F(c->ez, i, j - 1) + F(c->ez, i - 1, j - 1)) *
                              (REAL)0.25;
            F(c->gbz, i, j) = (F(c->bz, i, j) + F(c->bz, i - 1, j) +
                               F(c->bz, i, j - 1) + F(c->bz, i - 1, j - 1)) *
                              (REAL)0.25;
        }
}

/* Periodic ghost fill of an emf component: x ghosts as at :693-698 (column 0
 * <- mx, column mx+1 <- 1), y ghosts as the self halo of :705-730 (row 0 <- my,
 * row my+1 <- 1). */
static void PFX(ghost_f)(CT *c, REAL *a)
{
    const int mx = c->mx, my = c->my;
    for (int j = 1; j <= my; j++) {
        F(a, 0, j) = F(a, mx, j);
        F(a, mx + 1, j) = F(a, 1, j);
    }
    for (int i = 0; i <= mx + 1; i++) {
        F(a, i, 0) = F(a, i, my);
        F(a, i, my + 1) = F(a, i, 1);
    }
}

static void PFX(ghost_d)(CT *c, REAL *a)
{
    const int mx = c->mx, my = c->my;
    for (int j = 1; j <= my; j++) {
        D(a, 0, j) = D(a, mx, j);
        D(a, mx + 1, j) = D(a, 1, j);
        D(a, mx + 2, j) = D(a, 2 <= mx ? 2 : 1, j);
    }
    for (int i = 0; i <= mx + 2; i++) {
        D(a, i, 0) = D(a, i, my);
        D(a, i, my + 1) = D(a, i, 1);
        D(a, i, my + 2) = D(a, i, 2 <= my ? 2 : 1);
    }
}

/* Conducting walls in y.  pic2d.f90 is periodic in y; its perfectly
 * conducting boundary is only sketched, commented out, at :850-859.  Here the
 * domain is rows j = 1..my-2, the walls lie on node rows j = 1 (y = 0) and
 * j = my-1, and the rows j = my-1 (cell row above the top wall) and j = my
 * (= row 0, below the bottom wall) are guard rows holding the mirror images
 * of the fields: for the components on half rows (Ey, By, Ez, Bz) row my-1
 * mirrors row my-2 and row my mirrors row 1, even for the normal E and
 * tangential B components (Ey, Bz), odd for Ez and By; of the components on
 * node rows, Ex vanishes on both walls and Ex, Bx of row my (y = -dy) mirror
 * row 2 (y = +dy), odd and even.  The periodic ghost rows follow. */
#ifndef ORC_WALL_KINDS
#define ORC_WALL_KINDS
enum { WI_HALF_EVEN, WI_HALF_ODD, WI_NODE_E, WI_NODE_B };
#endif
static void PFX(wall_image)(CT *c, REAL *a, int kind)
{
    const int mx = c->mx, my = c->my;
    for (int i = 0; i <= mx + 1; i++) {
        switch (kind) {
        case WI_HALF_EVEN:
            F(a, i, my - 1) = F(a, i, my - 2);
            F(a, i, my) = F(a, i, 1);
            break;
        case WI_HALF_ODD:
            F(a, i, my - 1) = -F(a, i, my - 2);
            F(a, i, my) = -F(a, i, 1);
            break;
        case WI_NODE_E:
            F(a, i, 1) = 0;
            F(a, i, my - 1) = 0;
            F(a, i, my) = -F(a, i, 2);
            break;
        default:
            F(a, i, my) = F(a, i, 2);
            break;
        }
        F(a, i, 0) = F(a, i, my);
        F(a, i, my + 1) = F(a, i, 1);
    }
}

static void PFX(wall_images_e)(CT *c)
{
    PFX(wall_image)(c, c->ex, WI_NODE_E);
    PFX(wall_image)(c, c->ey, WI_HALF_EVEN);
    PFX(wall_image)(c, c->ez, WI_HALF_ODD);
}

static void PFX(wall_images_b)(CT *c)
{
    PFX(wall_image)(c, c->bx, WI_NODE_B);
    PFX(wall_image)(c, c->by, WI_HALF_ODD);
    PFX(wall_image)(c, c->bz, WI_HALF_EVEN);
}

/* Curl terms and emfgrid from E and B, as the restart path rebuilds them
 * (pic2d.f90:1906-1926). */
static void PFX(rebuild)(CT *c)
{
    const int mx = c->mx, my = c->my;
    const REAL idx = c->idx, idy = c->idy;
    for (int j = 1; j <= my; j++)
        for (int i = 0; i <= mx + 1; i++) {
            F(c->dbx_dy, i, j) = (F(c->bx, i, j + 1) - F(c->bx, i, j)) * idy;
            F(c->dex_dy, i, j) = (F(c->ex, i, j + 1) - F(c->ex, i, j)) * idy;
        }
    for (int j = 0; j <= my + 1; j++)
        for (int i = 0; i <= mx; i++) {
            F(c->dby_dx, i, j) = (F(c->by, i + 1, j) - F(c->by, i, j)) * idx;
            F(c->dey_dx, i, j) = (F(c->ey, i + 1, j) - F(c->ey, i, j)) * idx;
        }
    for (int j = 0; j <= my + 1; j++)
        for (int i = 1; i <= mx + 1; i++) {
            F(c->dbz_dx, i, j) = (F(c->bz, i, j) - F(c->bz, i - 1, j)) * idx;
            F(c->dez_dx, i, j) = (F(c->ez, i, j) - F(c->ez, i - 1, j)) * idx;
        }
    for (int j = 1; j <= my + 1; j++)
        for (int i = 0; i <= mx + 1; i++) {
            F(c->dbz_dy, i, j) = (F(c->bz, i, j) - F(c->bz, i, j - 1)) * idy;
            F(c->dez_dy, i, j) = (F(c->ez, i, j) - F(c->ez, i, j - 1)) * idy;
        }
    for (int j = 0; j <= my + 1; j++) {
        F(c->dby_dx, mx + 1, j) = F(c->dby_dx, 1, j);
        F(c->dey_dx, mx + 1, j) = F(c->dey_dx, 1, j);
    }
    /* dive as :1915 rebuilds it (y range 0:my; see fieldsolver) */
    for (int j = 1; j <= my + 1; j++)
        for (int i = 1; i <= mx + 1; i++)
            F(c->dive, i, j) = (F(c->ex, i, j) - F(c->ex, i - 1, j)) * idx +
                               (F(c->ey, i, j) - F(c->ey, i, j - 1)) * idy;
    PFX(emfgrid)(c);
}

static REAL *PFX(field_ptr)(CT *c, int which, int *is_d)
{
    *is_d = 0;
    switch (which) {
    case 0: return c->ex;  case 1: return c->ey;  case 2: return c->ez;
    case 3: return c->bx;  case 4: return c->by;  case 5: return c->bz;
    case 6: *is_d = 1; return c->jx;  case 7: *is_d = 1; return c->jy;
    case 8: *is_d = 1; return c->jz;  case 9: *is_d = 1; return c->rho;
    case 10: return c->gex; case 11: return c->gey; case 12: return c->gez;
    case 13: return c->gbx; case 14: return c->gby; case 15: return c->gbz;
    }
    return NULL;
}

/* plane: owned region (1:mx,1:my), x fastest. */
void PFX(set_field)(orc_ctx *oc, int which, const double *plane)
{
    CT *c = (CT *)oc;
    int is_d;
    REAL *a = PFX(field_ptr)(c, which, &is_d);
    if (!a || which >= 10) return;
    for (int j = 1; j <= c->my; j++)
        for (int i = 1; i <= c->mx; i++) {
            REAL v = (REAL)plane[(size_t)(i - 1) + (size_t)c->mx * (j - 1)];
            if (is_d) D(a, i, j) = v; else F(a, i, j) = v;
        }
    if (is_d) PFX(ghost_d)(c, a);
    else {
        static const int kinds[6] = {WI_NODE_E, WI_HALF_EVEN, WI_HALF_ODD,
                                     WI_NODE_B, WI_HALF_ODD, WI_HALF_EVEN};
        PFX(ghost_f)(c, a);
        if (c->ywall) PFX(wall_image)(c, a, kinds[which]);
        PFX(rebuild)(c);
    }
}

void PFX(get_field)(orc_ctx *oc, int which, double *plane)
{
    CT *c = (CT *)oc;
    int is_d;
    REAL *a = PFX(field_ptr)(c, which, &is_d);
    if (!a) return;
    for (int j = 1; j <= c->my; j++)
        for (int i = 1; i <= c->mx; i++)
            plane[(size_t)(i - 1) + (size_t)c->mx * (j - 1)] =
                is_d ? (double)D(a, i, j) : (double)F(a, i, j);
}

/* denjr = 0 (pic2d.f90:173-176) */
void PFX(zero_current)(orc_ctx *oc)
{
    CT *c = (CT *)oc;
    size_t nd = (size_t)(c->mx + 3) * (size_t)(c->my + 3);
    memset(c->jx, 0, nd * sizeof(REAL));
    memset(c->jy, 0, nd * sizeof(REAL));
    memset(c->jz, 0, nd * sizeof(REAL));
    memset(c->rho, 0, nd * sizeof(REAL));
}

/* currentdeposit (pic2d.f90:1226-1296) into the arrays jx..rho (row stride
 * mx+3).  The reference's `save` locals are plain locals here. */
static void PFX(currentdeposit)(const CT *c, REAL *jx, REAL *jy, REAL *jz,
                                REAL *rho, int i1, int i2, int j1, int j2,
                                REAL dx1, REAL dx2, REAL dy1, REAL dy2, int q0,
                                REAL uz)
{
    const REAL H = (REAL)0.5, ONE = (REAL)1;
    REAL qx = (REAL)q0 * c->dx / c->dt;                         /* :1237 */
    REAL qy = (REAL)q0 * c->dy / c->dt;
    REAL im = ((REAL)(i1 + i2) + dx1 + dx2 + ONE) * H;          /* :1239 */
    REAL jm = ((REAL)(j1 + j2) + dy1 + dy2 + ONE) * H;
    REAL ir = PFX(rmin)((REAL)(i1 < i2 ? i1 : i2) + ONE,        /* :1241 */
                        PFX(rmax)((REAL)(i1 > i2 ? i1 : i2) + (REAL)0, im));
    REAL jr = PFX(rmin)((REAL)(j1 < j2 ? j1 : j2) + ONE,
                        PFX(rmax)((REAL)(j1 > j2 ? j1 : j2) + (REAL)0, jm));
    REAL fx1 = qx * (ir - (REAL)i1 - dx1 - H);                  /* :1244 */
    REAL fy1 = qy * (jr - (REAL)j1 - dy1 - H);
    REAL fx2 = qx * ((REAL)i2 + dx2 + H - ir);
    REAL fy2 = qy * ((REAL)j2 + dy2 + H - jr);
    REAL wx1 = (ir - (REAL)i1 + dx1 + H) * H;                   /* :1249 */
    REAL wy1 = (jr - (REAL)j1 + dy1 + H) * H;
    REAL wx2 = (ir - (REAL)i2 + dx2 + H) * H;
    REAL wy2 = (jr - (REAL)j2 + dy2 + H) * H;
    D(jx, i1, j1) += fx1 * (ONE - wy1);                         /* :1254 */
    D(jx, i1, j1 + 1) += fx1 * wy1;
    D(jy, i1, j1) += fy1 * (ONE - wx1);
    D(jy, i1 + 1, j1) += fy1 * wx1;
    D(jx, i2, j2) += fx2 * (ONE - wy2);
    D(jx, i2, j2 + 1) += fx2 * wy2;
    D(jy, i2, j2) += fy2 * (ONE - wx2);
    D(jy, i2 + 1, j2) += fy2 * wx2;
    REAL dimid = im - (REAL)i1 - H;                             /* :1263 */
    REAL djmid = jm - (REAL)j1 - H;
    int di = (int)PFX(sgn1)(dimid);                             /* :1266 */
    int dj = (int)PFX(sgn1)(djmid);
    dimid = (REAL)fabs((double)dimid);
    djmid = (REAL)fabs((double)djmid);
    REAL s1 = dimid * djmid;                                    /* :1272 */
    REAL s2 = dimid - s1;
    REAL s3 = djmid - s1;
    REAL s4 = ONE - s1 - s2 - s3;
    REAL quz = (REAL)q0 * uz;                                   /* :1277 */
    D(jz, i1, j1) += quz * s4;
    D(jz, i1 + di, j1) += quz * s2;
    D(jz, i1, j1 + dj) += quz * s3;
    D(jz, i1 + di, j1 + dj) += quz * s1;
    if (!rho) return;
    if (c->compat & ORC_COMPAT_RHO_MIRROR) {
        /* :1285-1295, weights mirrored through the cell centre */
        s1 = (H - dx1) * (H - dy1);
        s2 = H - dx1 - s1;
        s3 = H - dy1 - s1;
        s4 = ONE - s1 - s2 - s3;
        D(rho, i1, j1) += s4 * (REAL)q0;
        D(rho, i1 + 1, j1) += s2 * (REAL)q0;
        D(rho, i1, j1 + 1) += s3 * (REAL)q0;
        D(rho, i1 + 1, j1 + 1) += s1 * (REAL)q0;
    } else {
        /* CIC weights built like the gather's (:948-951) */
        s1 = (H + dx1) * (H + dy1);
        s2 = H + dx1 - s1;
        s3 = H + dy1 - s1;
        s4 = ONE - s1 - s2 - s3;
        D(rho, i1, j1) += s4 * (REAL)q0;
        D(rho, i1 + 1, j1) += s2 * (REAL)q0;
        D(rho, i1, j1 + 1) += s3 * (REAL)q0;
        D(rho, i1 + 1, j1 + 1) += s1 * (REAL)q0;
    }
}

/* Post-deposit processing of pusher (pic2d.f90:1065-1069, 1148-1216):
 * normalization by in0, x-periodic fold of the guard columns, and the y halo
 * add with the rank as its own neighbour. */
void PFX(finish_current)(orc_ctx *oc)
{
    CT *c = (CT *)oc;
    const int mx = c->mx, my = c->my, mx1 = mx + 2;
    size_t nd = (size_t)(mx + 3) * (size_t)(my + 3);
    for (size_t k = 0; k < nd; k++) {                           /* :1066 */
        c->jx[k] = c->jx[k] * c->in0;
        c->jy[k] = c->jy[k] * c->in0;
        c->jz[k] = c->jz[k] * c->in0;
        c->rho[k] = c->rho[k] * c->in0;
    }
    for (int j = 0; j <= my + 2; j++) {                         /* :1150 */
        D(c->jx, 0, j) = D(c->jx, 0, j) + D(c->jx, mx, j);
        D(c->jx, 1, j) = D(c->jx, 1, j) + D(c->jx, mx + 1, j);
        D(c->jx, mx, j) = D(c->jx, 0, j);
        D(c->jx, mx + 1, j) = D(c->jx, 1, j);
        D(c->jy, 0, j) = D(c->jy, 0, j) + D(c->jy, mx, j);
        D(c->jy, 1, j) = D(c->jy, 1, j) + D(c->jy, mx + 1, j);
        if (!(c->compat & ORC_COMPAT_JY_DROP))                  /* :1156 */
            D(c->jy, 2, j) = D(c->jy, 2, j) + D(c->jy, mx + 2, j);
        D(c->jy, mx, j) = D(c->jy, 0, j);
        D(c->jy, mx + 1, j) = D(c->jy, 1, j);
        if (!(c->compat & ORC_COMPAT_JY_DROP))                  /* :1159 */
            D(c->jy, mx + 2, j) = D(c->jy, 2, j);
        D(c->jz, 0, j) = D(c->jz, 0, j) + D(c->jz, mx, j);
        D(c->jz, 1, j) = D(c->jz, 1, j) + D(c->jz, mx + 1, j);
        D(c->jz, mx, j) = D(c->jz, 0, j);
        D(c->jz, mx + 1, j) = D(c->jz, 1, j);
        D(c->rho, 1, j) = D(c->rho, 1, j) + D(c->rho, mx + 1, j);
        D(c->rho, mx + 1, j) = D(c->rho, 1, j);
        D(c->rho, 0, j) = D(c->rho, mx, j);
    }
    /* :1171-1181 pack; :1182-1203 the exchange (self); :1206-1216 unpack */
    REAL *st = (REAL *)malloc(sizeof(REAL) * (size_t)mx1 * 5);
    REAL *sb = (REAL *)malloc(sizeof(REAL) * (size_t)mx1 * 5);
    for (int i = 0; i <= mx + 1; i++) {
        st[i] = D(c->jx, i, my + 1);
        st[mx1 + i] = D(c->jx, i, my + 2);
        st[2 * mx1 + i] = D(c->jy, i, my + 1);
        st[3 * mx1 + i] = D(c->jz, i, my + 1);
        st[4 * mx1 + i] = D(c->rho, i, my + 1);
        sb[i] = D(c->jx, i, 0);
        sb[mx1 + i] = D(c->jx, i, 1);
        sb[2 * mx1 + i] = D(c->jy, i, 0);
        sb[3 * mx1 + i] = D(c->jz, i, 0);
        sb[4 * mx1 + i] = D(c->rho, i, 1);
    }
    /* recvb = left's sendt = st, recvt = right's sendb = sb */
    for (int i = 0; i <= mx + 1; i++) {
        D(c->jx, i, 1) = D(c->jx, i, 1) + st[i];
        D(c->jx, i, 2) = D(c->jx, i, 2) + st[mx1 + i];
        D(c->jy, i, 1) = D(c->jy, i, 1) + st[2 * mx1 + i];
        D(c->jz, i, 1) = D(c->jz, i, 1) + st[3 * mx1 + i];
        D(c->rho, i, 1) = D(c->rho, i, 1) + st[4 * mx1 + i];
        D(c->jx, i, my) = D(c->jx, i, my) + sb[i];
        D(c->jx, i, my + 1) = D(c->jx, i, my + 1) + sb[mx1 + i];
        D(c->jy, i, my) = D(c->jy, i, my) + sb[2 * mx1 + i];
        D(c->jz, i, my) = D(c->jz, i, my) + sb[3 * mx1 + i];
        D(c->rho, i, my + 1) = D(c->rho, i, my + 1) + sb[4 * mx1 + i];
    }
    free(st); free(sb);
    if (c->ywall) {
        /* Jz deposited beyond a wall (guard rows my-1 and my) belongs to the
         * mirror image of its particle: fold it back with the image's sign
         * (the image of charge q at -y carries -q v_z). */
        for (int i = 0; i <= mx + 2; i++) {
            D(c->jz, i, 1) = D(c->jz, i, 1) - D(c->jz, i, my);
            D(c->jz, i, my - 2) = D(c->jz, i, my - 2) - D(c->jz, i, my - 1);
            D(c->jz, i, my - 1) = 0;
            D(c->jz, i, my) = 0;
            D(c->jz, i, 0) = 0;
        }
    }
}

/* One particle of pusher's loop body (pic2d.f90:934-1062). */
static int PFX(push_one)(const CT *c, PFX(ptl_t) *p, REAL *jx, REAL *jy,
                         REAL *jz, REAL *rho, REAL *dke)
{
    const int mx = c->mx, my = c->my;
    const REAL H = (REAL)0.5, ONE = (REAL)1;
    REAL ux = p->ux, uy = p->uy, uz = p->uz;                    /* :935 */
    int ix = (p->grid - 1) / my + 1;                            /* :938 */
    int iy = (p->grid - 1) % my + 1;
    REAL s1 = (H + p->dx) * (H + p->dy);                        /* :948 */
    REAL s2 = H + p->dx - s1;
    REAL s3 = H + p->dy - s1;
    REAL s4 = ONE - s1 - s2 - s3;
#define GATHER(g) (F(g, ix, iy) * s4 + F(g, ix + 1, iy) * s2 + \
                   F(g, ix, iy + 1) * s3 + F(g, ix + 1, iy + 1) * s1)
    REAL ex = GATHER(c->gex), ey = GATHER(c->gey), ez = GATHER(c->gez);
    REAL bx = GATHER(c->gbx), by = GATHER(c->gby), bz = GATHER(c->gbz);
#undef GATHER
    REAL ms = p->q > 0 ? c->mime : ONE;                         /* :964 */
    REAL c0 = (REAL)p->q * c->dt * c->mime / ms / (REAL)2;      /* :969 */
    REAL c1 = c0 / c->cva;
    REAL umx = ux + c1 * ex;                                    /* :971 */
    REAL umy = uy + c1 * ey;
    REAL umz = uz + c1 * ez;
    REAL gama = (REAL)sqrt((double)(ONE + umx * umx + umy * umy + umz * umz));
    REAL c2 = c0 / gama;                                        /* :976 */
    REAL tx = c2 * bx, ty = c2 * by, tz = c2 * bz;
    REAL t2 = (REAL)2 / (ONE + tx * tx + ty * ty + tz * tz);
    REAL sx = t2 * tx, sy = t2 * ty, sz = t2 * tz;
    REAL u0x, u0y, u0z;
    if (c->compat & ORC_COMPAT_BORIS_OLD_U) {                   /* :985-987 */
        u0x = umx + (uy * tz - uz * ty);
        u0y = umy + (uz * tx - ux * tz);
        u0z = umz + (ux * ty - uy * tx);
    } else {                                                    /* textbook */
        u0x = umx + (umy * tz - umz * ty);
        u0y = umy + (umz * tx - umx * tz);
        u0z = umz + (umx * ty - umy * tx);
    }
    REAL upx = umx + (u0y * sz - u0z * sy);                     /* :989 */
    REAL upy = umy + (u0z * sx - u0x * sz);
    REAL upz = umz + (u0x * sy - u0y * sx);
    ux = upx + c1 * ex;                                         /* :993 */
    uy = upy + c1 * ey;
    uz = upz + c1 * ez;
    gama = (REAL)sqrt((double)(ONE + ux * ux + uy * uy + uz * uz));
    p->ux = ux; p->uy = uy; p->uz = uz;                         /* :998 */
    REAL dx0 = p->dx, dy0 = p->dy;                              /* :1002 */
    REAL c3 = c->dt * c->cva / gama;
    REAL dx1 = dx0 + c3 * ux * c->idx;
    REAL dy1 = dy0 + c3 * uy * c->idy;
    int dcellx, dcelly;
    if ((REAL)fabs((double)dx1) > H) {                          /* :1008 */
        dcellx = (int)PFX(sgn1)(dx1);
        p->dx = dx1 + PFX(sgn1)(-dx1);
    } else {
        dcellx = 0;
        p->dx = dx1;
    }
    if ((REAL)fabs((double)dy1) > H) {                          /* :1017 */
        dcelly = (int)PFX(sgn1)(dy1);
        p->dy = dy1 + PFX(sgn1)(-dy1);
    } else {
        dcelly = 0;
        p->dy = dy1;
    }
    if (c->ywall && ((iy == 1 && dcelly < 0) || (iy == my - 2 && dcelly > 0))) {
        /* specular reflection at a conducting wall: the position mirrored
         * into the cell it left from, u_y reversed, then the deposit */
        dcelly = 0;
        p->dy = -p->dy;
        uy = -uy;
        p->uy = uy;
    }
    int ix1 = ix + dcellx, iy1 = iy + dcelly;                   /* :1026 */
    REAL vz = uz * c->cva / gama;                               /* :1028 */
    PFX(currentdeposit)(c, jx, jy, jz, rho, ix, ix1, iy, iy1, dx0, p->dx, dy0,
                        p->dy, p->q, vz);                       /* :1029 */
    ix = (ix1 + mx - 1) % mx + 1;                               /* :1032 */
    if (iy1 > my) iy = iy1 - my;                                /* :1033 */
    else if (iy1 < 1) iy = iy1 + my;                            /* :1042 */
    else iy = iy1;
    p->grid = my * (ix - 1) + iy;
    *dke = (gama - ONE) * ms;        /* :1062 */
    return p->q > 0 ? 2 : 1;         /* j = ceiling(q*0.5+1.0), :1061 */
}

/* pusher (pic2d.f90:912-1217) incl. the deposit post-processing. */
void PFX(pusher)(orc_ctx *oc)
{
    CT *c = (CT *)oc;
    const int nt = c->nthreads;
    c->totene[1] = c->totene[2] = 0;
    c->ke64[0] = c->ke64[1] = 0;
    REAL *rho = c->rho;
    if (nt <= 1) {
        for (int64_t i = 0; i < c->nptl; i++) {
            REAL dke;
            int j = PFX(push_one)(c, &c->ptl[i], c->jx, c->jy, c->jz, rho, &dke);
            c->totene[j] = c->totene[j] + dke;
            c->ke64[j - 1] += (double)dke;
        }
    } else {
        size_t nd = (size_t)(c->mx + 3) * (size_t)(c->my + 3);
        double k64[2] = {0, 0};
        #pragma omp parallel num_threads(nt) reduction(+ : k64[:2])
        {
            int t = omp_get_thread_num();
            REAL *tj = c->tj + (size_t)t * 4 * nd;
            memset(tj, 0, 4 * nd * sizeof(REAL));
            int64_t lo = c->nptl * t / nt, hi = c->nptl * (t + 1) / nt;
            for (int64_t i = lo; i < hi; i++) {
                REAL dke;
                int j = PFX(push_one)(c, &c->ptl[i], tj, tj + nd, tj + 2 * nd,
                                      tj + 3 * nd, &dke);
                k64[j - 1] += (double)dke;
            }
        }
        /* merge the private copies in thread order */
        #pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t k = 0; k < (int64_t)nd; k++) {
            for (int t = 0; t < nt; t++) {
                const REAL *tj = c->tj + (size_t)t * 4 * nd;
                c->jx[k] += tj[k];
                c->jy[k] += tj[nd + k];
                c->jz[k] += tj[2 * nd + k];
                c->rho[k] += tj[3 * nd + k];
            }
        }
        c->ke64[0] = k64[0];
        c->ke64[1] = k64[1];
        c->totene[1] = (REAL)k64[0];
        c->totene[2] = (REAL)k64[1];
    }
    PFX(finish_current)(oc);
}

void PFX(deposit_batch)(orc_ctx *oc, int64_t n, const int32_t *i1,
                        const int32_t *i2, const int32_t *j1,
                        const int32_t *j2, const double *dx1,
                        const double *dx2, const double *dy1,
                        const double *dy2, const int32_t *q, const double *vz)
{
    CT *c = (CT *)oc;
    for (int64_t k = 0; k < n; k++)
        PFX(currentdeposit)(c, c->jx, c->jy, c->jz, c->rho, i1[k], i2[k],
                            j1[k], j2[k], (REAL)dx1[k], (REAL)dx2[k],
                            (REAL)dy1[k], (REAL)dy2[k], q[k], (REAL)vz[k]);
}

/* y halo of two emf components with the rank as its own neighbour
 * (pic2d.f90:700-730): row my+1 <- row 1, row 0 <- row my. */
static void PFX(halo2)(CT *c, REAL *a, REAL *b)
{
    const int mx = c->mx, my = c->my;
    for (int i = 0; i <= mx + 1; i++) {
        REAL t0 = F(a, i, my), t1 = F(b, i, my);   /* bufsendt */
        REAL b0 = F(a, i, 1), b1 = F(b, i, 1);     /* bufsendb */
        F(a, i, my + 1) = b0; F(b, i, my + 1) = b1;
        F(a, i, 0) = t0; F(b, i, 0) = t1;
    }
}

/* divergenceclean (pic2d.f90:867-885), Marder's pass: E += d grad(dive -
 * c^2 rho) with d = dt*0.001, dive and rho the values held when it is
 * called: div E^n from the end of the previous solve (or from the fields as
 * set) and the rho the push just deposited (normalized).  The source updates
 * ex(1:mx,1:my+1), ey(1:mx+1,1:my) and leaves the ghost copies stale (column
 * 1 is then overwritten from column mx+1 at the next solve, :747-752); here
 * the owned region is updated with periodic neighbours and the ghosts copied. */
static void PFX(divergenceclean)(CT *c)
{
    const int mx = c->mx, my = c->my;
    const REAL dconst = (REAL)0.001;
    const REAL dtdx = c->dt * dconst / c->dx, dtdy = c->dt * dconst / c->dy;
    REAL *mf = (REAL *)calloc((size_t)(mx + 2) * (my + 2), sizeof(REAL));
#define MF(i, j) mf[(size_t)(i) + (size_t)(mx + 2) * (size_t)(j)]
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++)
            MF(i, j) = F(c->dive, i, j) - c->cva2 * D(c->rho, i, j);
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++) {
            const int ip = i == mx ? 1 : i + 1, jp = j == my ? 1 : j + 1;
            F(c->ex, i, j) = F(c->ex, i, j) + dtdx * (MF(ip, j) - MF(i, j));
            F(c->ey, i, j) = F(c->ey, i, j) + dtdy * (MF(i, jp) - MF(i, j));
        }
#undef MF
    free(mf);
    PFX(ghost_f)(c, c->ex);
    PFX(ghost_f)(c, c->ey);
}

/* suppresscerenkov (pic2d.f90:890-907): E -= c^2 dt tau curl curl E with
 * tau = dt/8, built from the curl terms of E^n the previous solve stored:
 * w = dey_dx - dex_dy at the cell centres, and dez_dx, dez_dy.  The source
 * reads rows of dex_dy its curl loop (:790) never writes (0 and my+1) and
 * leaves the ghosts stale; here the owned region is updated with periodic
 * neighbours and the ghosts copied. */
static void PFX(suppresscerenkov)(CT *c)
{
    const int mx = c->mx, my = c->my;
    const REAL tau = (REAL)0.125 * c->dt;
    const REAL dtdx = c->dt * c->cva2 * tau / c->dx, dtdy = c->dt * c->cva2 * tau / c->dy;
#define W(i, j) (F(c->dey_dx, i, j) - F(c->dex_dy, i, j))
    REAL *t = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)mx * my);
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++) {
            const int im = i == 1 ? mx : i - 1, jm = j == 1 ? my : j - 1;
            const int ip = i == mx ? 1 : i + 1, jp = j == my ? 1 : j + 1;
            const size_t k = (size_t)(i - 1) + (size_t)mx * (j - 1);
            t[k] = -(dtdy * (W(i, j) - W(i, jm)));
            t[(size_t)mx * my + k] = dtdx * (W(i, j) - W(im, j));
            t[2 * (size_t)mx * my + k] = dtdx * (F(c->dez_dx, ip, j) - F(c->dez_dx, i, j)) +
                                         dtdy * (F(c->dez_dy, i, jp) - F(c->dez_dy, i, j));
        }
#undef W
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++) {
            const size_t k = (size_t)(i - 1) + (size_t)mx * (j - 1);
            F(c->ex, i, j) = F(c->ex, i, j) + t[k];
            F(c->ey, i, j) = F(c->ey, i, j) + t[(size_t)mx * my + k];
            F(c->ez, i, j) = F(c->ez, i, j) + t[2 * (size_t)mx * my + k];
        }
    free(t);
    PFX(ghost_f)(c, c->ex);
    PFX(ghost_f)(c, c->ey);
    PFX(ghost_f)(c, c->ez);
}

int PFX(set_filters)(orc_ctx *oc, int marder, int cherenkov)
{
    CT *c = (CT *)oc;
    c->marder = marder != 0;
    c->cherenkov = cherenkov != 0;
    return 0;
}

/* fieldsolver (pic2d.f90:678-860). */
void PFX(fieldsolver)(orc_ctx *oc)
{
    CT *c = (CT *)oc;
    const int mx = c->mx, my = c->my;
    const REAL idx = c->idx, idy = c->idy, dth = c->dth;
    const REAL dtc2 = c->dt * c->cva2;
    size_t nf = (size_t)(mx + 2) * (size_t)(my + 2);
    for (size_t k = 0; k < nf; k++) {                           /* :688-690 */
        c->bx[k] = c->bx[k] - dth * c->dez_dy[k];
        c->by[k] = c->by[k] + dth * c->dez_dx[k];
        c->bz[k] = c->bz[k] + dth * (c->dex_dy[k] - c->dey_dx[k]);
    }
    for (int j = 0; j <= my + 1; j++) {                         /* :693-698 */
        F(c->bx, 0, j) = F(c->bx, mx, j); F(c->bx, 1, j) = F(c->bx, mx + 1, j);
        F(c->by, 0, j) = F(c->by, mx, j); F(c->by, 1, j) = F(c->by, mx + 1, j);
        F(c->bz, 0, j) = F(c->bz, mx, j); F(c->bz, 1, j) = F(c->bz, mx + 1, j);
    }
    PFX(halo2)(c, c->by, c->bz);                                /* :700-730 */
    if (c->ywall) PFX(wall_images_b)(c);
    for (int j = 1; j <= my; j++)                               /* :733 */
        for (int i = 0; i <= mx + 1; i++)
            F(c->dbx_dy, i, j) = (F(c->bx, i, j + 1) - F(c->bx, i, j)) * idy;
    for (int j = 0; j <= my + 1; j++)                           /* :734 */
        for (int i = 0; i <= mx; i++)
            F(c->dby_dx, i, j) = (F(c->by, i + 1, j) - F(c->by, i, j)) * idx;
    for (int j = 0; j <= my + 1; j++)                           /* :735 */
        for (int i = 1; i <= mx + 1; i++)
            F(c->dbz_dx, i, j) = (F(c->bz, i, j) - F(c->bz, i - 1, j)) * idx;
    for (int j = 1; j <= my + 1; j++)                           /* :736 */
        for (int i = 0; i <= mx + 1; i++)
            F(c->dbz_dy, i, j) = (F(c->bz, i, j) - F(c->bz, i, j - 1)) * idy;
    for (int j = 0; j <= my + 1; j++)                           /* :739 */
        F(c->dby_dx, mx + 1, j) = F(c->dby_dx, 1, j);
    /* :742-744, emf(i,j) paired with denjr(i,j) (lower-bound alignment) */
    double je = 0;
    for (int j = 0; j <= my + 1; j++)
        for (int i = 0; i <= mx + 1; i++) {
            REAL exo = F(c->ex, i, j), eyo = F(c->ey, i, j), ezo = F(c->ez, i, j);
            F(c->ex, i, j) = exo + dtc2 * (F(c->dbz_dy, i, j) - D(c->jx, i, j));
            F(c->ey, i, j) = eyo - dtc2 * (F(c->dbz_dx, i, j) + D(c->jy, i, j));
            F(c->ez, i, j) = ezo + dtc2 * (F(c->dby_dx, i, j) -
                                           F(c->dbx_dy, i, j) - D(c->jz, i, j));
            if (c->ywall && (j == 1 || j == my - 1)) F(c->ex, i, j) = 0;   /* wall */
            if (i >= 1 && i <= mx && j >= 1 && j <= my)
                je += (double)D(c->jx, i, j) * 0.5 * ((double)exo + (double)F(c->ex, i, j)) +
                      (double)D(c->jy, i, j) * 0.5 * ((double)eyo + (double)F(c->ey, i, j)) +
                      (double)D(c->jz, i, j) * 0.5 * ((double)ezo + (double)F(c->ez, i, j));
        }
    c->jdote_step = je * (double)c->mime / ((double)c->cva2 * (double)c->in0);
    c->jdote_int += c->jdote_step * (double)c->dt;
    for (int j = 0; j <= my + 1; j++) {                         /* :747-752 */
        F(c->ex, 0, j) = F(c->ex, mx, j); F(c->ex, 1, j) = F(c->ex, mx + 1, j);
        F(c->ey, 0, j) = F(c->ey, mx, j); F(c->ey, 1, j) = F(c->ey, mx + 1, j);
        F(c->ez, 0, j) = F(c->ez, mx, j); F(c->ez, 1, j) = F(c->ez, mx + 1, j);
    }
    if (c->marder) PFX(divergenceclean)(c);                     /* :754 */
    if (c->cherenkov) PFX(suppresscerenkov)(c);                 /* :755 */
    PFX(halo2)(c, c->ey, c->ez);                                /* :757-787 */
    if (c->ywall) PFX(wall_images_e)(c);
    for (int j = 1; j <= my; j++)                               /* :790 */
        for (int i = 0; i <= mx + 1; i++)
            F(c->dex_dy, i, j) = (F(c->ex, i, j + 1) - F(c->ex, i, j)) * idy;
    for (int j = 0; j <= my + 1; j++)                           /* :791 */
        for (int i = 0; i <= mx; i++)
            F(c->dey_dx, i, j) = (F(c->ey, i + 1, j) - F(c->ey, i, j)) * idx;
    for (int j = 0; j <= my + 1; j++)                           /* :792 */
        for (int i = 1; i <= mx + 1; i++)
            F(c->dez_dx, i, j) = (F(c->ez, i, j) - F(c->ez, i - 1, j)) * idx;
    for (int j = 1; j <= my + 1; j++)                           /* :793 */
        for (int i = 0; i <= mx + 1; i++)
            F(c->dez_dy, i, j) = (F(c->ez, i, j) - F(c->ez, i, j - 1)) * idy;
    /* :794-795 with the y range 0:my (the source writes 0:mx, a shape bug
       that only differs when mx /= my; dive feeds no part of the step) */
    for (int j = 1; j <= my + 1; j++)
        for (int i = 1; i <= mx + 1; i++)
            F(c->dive, i, j) = (F(c->ex, i, j) - F(c->ex, i - 1, j)) * idx +
                               (F(c->ey, i, j) - F(c->ey, i, j - 1)) * idy;
    for (int j = 0; j <= my + 1; j++)                           /* :798 */
        F(c->dey_dx, mx + 1, j) = F(c->dey_dx, 1, j);
    for (size_t k = 0; k < nf; k++) {                           /* :801-803 */
        c->bx[k] = c->bx[k] - dth * c->dez_dy[k];
        c->by[k] = c->by[k] + dth * c->dez_dx[k];
        c->bz[k] = c->bz[k] + dth * (c->dex_dy[k] - c->dey_dx[k]);
    }
    PFX(halo2)(c, c->by, c->bz);                                /* :805-835 */
    if (c->ywall) PFX(wall_images_b)(c);
    PFX(emfgrid)(c);                                            /* :837-845 */
    double e2[2];
    PFX(fieldsenergy)(oc, e2);                                  /* :848 */
    c->steps++;
}

/* fieldsenergy (pic2d.f90:653-673): the reference's REAL accumulators go to
 * totene(3:4); out[] returns an FP64 sum of the same terms. */
void PFX(fieldsenergy)(orc_ctx *oc, double out[2])
{
    CT *c = (CT *)oc;
    REAL t3 = 0, t4 = 0;
    double d3 = 0, d4 = 0;
    /* with walls: the node rows from wall to wall, the walls weighted 1/2 */
    const int jhi = c->ywall ? c->my - 1 : c->my;
    for (int j = 1; j <= jhi; j++) {
        const int wall = c->ywall && (j == 1 || j == jhi);
        for (int i = 1; i <= c->mx; i++) {
            REAL ene = F(c->gex, i, j) * F(c->gex, i, j) +
                       F(c->gey, i, j) * F(c->gey, i, j) +
                       F(c->gez, i, j) * F(c->gez, i, j);
            double e64 = (double)F(c->gex, i, j) * F(c->gex, i, j) +
                         (double)F(c->gey, i, j) * F(c->gey, i, j) +
                         (double)F(c->gez, i, j) * F(c->gez, i, j);
            if (wall) { ene = ene * (REAL)0.5; e64 *= 0.5; }
            t3 = t3 + ene;
            d3 += e64;
            ene = F(c->gbx, i, j) * F(c->gbx, i, j) +
                  F(c->gby, i, j) * F(c->gby, i, j) +
                  F(c->gbz, i, j) * F(c->gbz, i, j);
            e64 = (double)F(c->gbx, i, j) * F(c->gbx, i, j) +
                  (double)F(c->gby, i, j) * F(c->gby, i, j) +
                  (double)F(c->gbz, i, j) * F(c->gbz, i, j);
            if (wall) { ene = ene * (REAL)0.5; e64 *= 0.5; }
            t4 = t4 + ene;
            d4 += e64;
        }
    }
    c->totene[3] = (REAL)0.5 * t3 * c->mime / (c->cva2 * c->cva2 * c->in0);
    c->totene[4] = (REAL)0.5 * t4 * c->mime / (c->cva2 * c->in0);
    out[0] = 0.5 * d3 * (double)c->mime /
             ((double)c->cva2 * (double)c->cva2 * (double)c->in0);
    out[1] = 0.5 * d4 * (double)c->mime / ((double)c->cva2 * (double)c->in0);
}

/* countingsort (pic2d.f90:1303-1329). counts[] = ptlalloc after the
 * histogram (:1310-1313); perm[k] = 0-based pre-sort index of the particle
 * at 0-based position k. */
void PFX(countingsort)(orc_ctx *oc, int32_t *counts, int32_t *perm)
{
    CT *c = (CT *)oc;
    const int64_t ncell = (int64_t)c->mx * c->my;
    int32_t *pa = c->ptlalloc;
    const int nt = c->nthreads;
    if (nt > 1 && c->nptl >= (int64_t)nt * 4096) {
        /* The same stable counting sort with the particles cut into nt
           contiguous chunks: per-chunk counts, the offset of (chunk t, cell)
           = all particles of lower cells + those of cell in chunks < t, then
           each chunk scatters in index order.  The result is the serial
           loop's, element for element. */
        int32_t *h = (int32_t *)calloc((size_t)nt * (size_t)ncell, sizeof(int32_t));
        if (h) {
            #pragma omp parallel num_threads(nt)
            {
                const int t = omp_get_thread_num();
                int32_t *ht = h + (size_t)t * (size_t)ncell;
                const int64_t lo = c->nptl * t / nt, hi = c->nptl * (t + 1) / nt;
                for (int64_t i = lo; i < hi; i++) ht[c->ptl[i].grid - 1]++;
                #pragma omp barrier
                #pragma omp single
                {
                    int32_t k = 0;
                    for (int64_t j = 0; j < ncell; j++) {
                        int32_t tot = 0;
                        for (int u = 0; u < nt; u++) {
                            int32_t *hu = h + (size_t)u * (size_t)ncell + j;
                            const int32_t v = *hu;
                            *hu = k + tot;
                            tot += v;
                        }
                        if (counts) counts[j] = tot;
                        k += tot;
                    }
                }
                for (int64_t i = lo; i < hi; i++) {
                    const int32_t kk = ht[c->ptl[i].grid - 1]++;
                    c->ptlsort[kk] = c->ptl[i];
                    if (perm) perm[kk] = (int32_t)i;
                }
                #pragma omp barrier
                memcpy(c->ptl + lo, c->ptlsort + lo, sizeof(PFX(ptl_t)) * (size_t)(hi - lo));
            }
            free(h);
            return;
        }
    }
    memset(pa, 0, sizeof(int32_t) * (size_t)ncell);             /* :1309 */
    for (int64_t i = 0; i < c->nptl; i++) pa[c->ptl[i].grid - 1]++;
    if (counts) memcpy(counts, pa, sizeof(int32_t) * (size_t)ncell);
    int32_t k = 0;                                              /* :1315 */
    for (int64_t i = 0; i < ncell; i++) {
        int32_t j = pa[i];
        pa[i] = k;
        k += j;
    }
    for (int64_t i = 0; i < c->nptl; i++) {                     /* :1322 */
        int32_t j = c->ptl[i].grid - 1;
        int32_t kk = pa[j] + 1;
        pa[j] = kk;
        c->ptlsort[kk - 1] = c->ptl[i];
        if (perm) perm[kk - 1] = (int32_t)i;
    }
    memcpy(c->ptl, c->ptlsort, sizeof(PFX(ptl_t)) * (size_t)c->nptl);   /* :1328 */
}

/* The driver loop body (pic2d.f90:171-188). */
void PFX(step)(orc_ctx *oc, int nsteps, int sort_every)
{
    CT *c = (CT *)oc;
    for (int s = 0; s < nsteps; s++) {
        PFX(zero_current)(oc);
        PFX(pusher)(oc);
        PFX(fieldsolver)(oc);
        if (sort_every > 0 && c->steps % sort_every == 0)
            PFX(countingsort)(oc, NULL, NULL);
    }
}

/* FP64 diagnostics: kinetic energies of the last push, field energies, J.E
 * and the Gauss residual divE - c^2 in0 (rho_cic + rho_bg) with correct CIC
 * node weights (the weights of the gather, :948-951). */
void PFX(diag)(orc_ctx *oc, double rho_bg, double out[16])
{
    CT *c = (CT *)oc;
    const int mx = c->mx, my = c->my;
    for (int k = 0; k < 16; k++) out[k] = 0;
    out[0] = c->ke64[0];
    out[1] = c->ke64[1];
    double e2[2];
    PFX(fieldsenergy)(oc, e2);
    out[2] = e2[0];
    out[3] = e2[1];
    out[4] = c->jdote_step;
    out[5] = c->jdote_int;
    size_t nn = (size_t)mx * my;
    double *r = (double *)calloc(nn, sizeof(double));
    for (int64_t i = 0; i < c->nptl; i++) {
        const PFX(ptl_t) *p = &c->ptl[i];
        int ix = (p->grid - 1) / my, iy = (p->grid - 1) % my;  /* 0-based node */
        int ix1 = ix + 1 == mx ? 0 : ix + 1, iy1 = iy + 1 == my ? 0 : iy + 1;
        double wx = 0.5 + (double)p->dx, wy = 0.5 + (double)p->dy, q = p->q;
        r[ix + (size_t)mx * iy] += q * (1 - wx) * (1 - wy);
        r[ix1 + (size_t)mx * iy] += q * wx * (1 - wy);
        r[ix + (size_t)mx * iy1] += q * (1 - wx) * wy;
        r[ix1 + (size_t)mx * iy1] += q * wx * wy;
    }
    double l2 = 0, li = 0, dl2 = 0, dli = 0;
    int first = c->gauss0 == NULL;
    if (first) c->gauss0 = (double *)calloc(nn, sizeof(double));
    /* with walls: the node rows strictly between the walls */
    const int jlo = c->ywall ? 2 : 1, jhi = c->ywall ? my - 2 : my;
    for (int j = jlo; j <= jhi; j++)
        for (int i = 1; i <= mx; i++) {
            double div = ((double)F(c->ex, i, j) - (double)F(c->ex, i - 1, j)) * (double)c->idx +
                         ((double)F(c->ey, i, j) - (double)F(c->ey, i, j - 1)) * (double)c->idy;
            size_t k = (size_t)(i - 1) + (size_t)mx * (j - 1);
            double res = div - (double)c->cva2 * (double)c->in0 * (r[k] + rho_bg);
            if (first) c->gauss0[k] = res;
            double d = res - c->gauss0[k];
            l2 += res * res;
            li = fabs(res) > li ? fabs(res) : li;
            dl2 += d * d;
            dli = fabs(d) > dli ? fabs(d) : dli;
        }
    free(r);
    out[6] = sqrt(l2);
    out[7] = li;
    out[8] = sqrt(dl2);
    out[9] = dli;
    out[10] = (double)c->steps;
}

#undef CT
#undef F
#undef D
