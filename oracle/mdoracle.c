/*
 * mdoracle.c -- CPU FP64 oracle (test infrastructure only; see mdoracle.h).
 *
 * A from-scratch restatement of the MDBas nonbonded path.  Citations are
 * to the reference sources (src/*.c, include/*.h).  Compile with
 * -ffp-contract=off: the list predicates must round like the reference.
 */
#include "mdoracle.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define SQ(x) ((x) * (x))
#define OC_PI 3.141592653589793       /* global.h:52 */
#define OC_TWOPI 6.283185307179586    /* global.h:53 */
#define OC_SQRTPI 1.772453850905516   /* global.h:54 */
#define OC_TOLLIST 0.01               /* global.h:59 */

typedef double complex ocplx;

/* io.c:714-723 */
void oc_param_derive(oc_param *p)
{
    p->rcutOff = 1. / p->cutOff;
    p->cutOff2 = SQ(p->cutOff);
    p->rcutOff2 = 1. / p->cutOff2;
    p->cutOn2 = SQ(p->cutOn);
    p->switch2 = 1. / (SQ(p->cutOff2 - p->cutOn2) * (p->cutOff2 - p->cutOn2));
    p->rTimeStep = 1. / p->timeStep;
}

/* init.c:840-908: norms, reciprocal vectors (divided by det), perpendicular
 * widths pa/pb/pc and the volume. */
void oc_init_box(oc_box *b)
{
    b->u1 = b->b2 * b->c3 - b->c2 * b->b3;
    b->u2 = b->c1 * b->b3 - b->b1 * b->c3;
    b->u3 = b->b1 * b->c2 - b->c1 * b->b2;
    b->v1 = b->c2 * b->a3 - b->a2 * b->c3;
    b->v2 = b->a1 * b->c3 - b->c1 * b->a3;
    b->v3 = b->c1 * b->a2 - b->a1 * b->c2;
    b->w1 = b->a2 * b->b3 - b->b2 * b->a3;
    b->w2 = b->b1 * b->a3 - b->a1 * b->b3;
    b->w3 = b->a1 * b->b2 - b->b1 * b->a2;
    b->det = b->a1 * b->u1 + b->a2 * b->u2 + b->a3 * b->u3;
    b->vol = fabs(b->det);
    b->pa = b->vol / sqrt(SQ(b->u1) + SQ(b->u2) + SQ(b->u3));
    b->pb = b->vol / sqrt(SQ(b->v1) + SQ(b->v2) + SQ(b->v3));
    b->pc = b->vol / sqrt(SQ(b->w1) + SQ(b->w2) + SQ(b->w3));
    b->u = 1.0 / b->pa;
    b->v = 1.0 / b->pb;
    b->w = 1.0 / b->pc;
    if (fabs(b->det) > 0.) {
        b->u1 /= b->det; b->u2 /= b->det; b->u3 /= b->det;
        b->v1 /= b->det; b->v2 /= b->det; b->v3 /= b->det;
        b->w1 /= b->det; b->w2 /= b->det; b->w3 /= b->det;
    } else {
        b->u1 = b->u2 = b->u3 = b->v1 = b->v2 = b->v3 = b->w1 = b->w2 = b->w3 = 0.;
    }
}

/* utils.c:606-609: round half away from zero through a C truncating cast. */
int oc_nint(double x)
{
    return x >= 0. ? (int)(x + 0.5) : (int)(x - 0.5);
}

/* utils.c:38-80: minimum image of d in place, returns |d|^2 summed as
 * (dx^2 + dy^2) + dz^2.  The cubic branch uses a1/u1 on all three axes. */
double oc_dist(const oc_box *b, double d[3])
{
    double xt, yt, zt;
    switch (b->type) {
    case 1:
        d[0] -= b->a1 * oc_nint(d[0] * b->u1);
        d[1] -= b->a1 * oc_nint(d[1] * b->u1);
        d[2] -= b->a1 * oc_nint(d[2] * b->u1);
        break;
    case 2:
        d[0] -= b->a1 * oc_nint(d[0] * b->u1);
        d[1] -= b->b2 * oc_nint(d[1] * b->v2);
        d[2] -= b->c3 * oc_nint(d[2] * b->w3);
        break;
    case 3:
        xt = d[0] * b->u1 + d[1] * b->u2 + d[2] * b->u3;
        yt = d[0] * b->v1 + d[1] * b->v2 + d[2] * b->v3;
        zt = d[0] * b->w1 + d[1] * b->w2 + d[2] * b->w3;
        xt -= oc_nint(xt);
        yt -= oc_nint(yt);
        zt -= oc_nint(zt);
        d[0] = xt * b->a1 + yt * b->b1 + zt * b->c1;
        d[1] = xt * b->a2 + yt * b->b2 + zt * b->c2;
        d[2] = xt * b->a3 + yt * b->b3 + zt * b->c3;
        break;
    default:
        break;
    }
    return SQ(d[0]) + SQ(d[1]) + SQ(d[2]);
}

/* utils.c:162-218: wrap coordinates in place. */
void oc_image_update(const oc_box *b, int n, double *x, double *y, double *z)
{
    for (int i = 0; i < n; i++) {
        double d[3] = {x[i], y[i], z[i]};
        switch (b->type) {
        case 1:
            x[i] = x[i] - b->a1 * oc_nint(x[i] * b->u1);
            y[i] = y[i] - b->a1 * oc_nint(y[i] * b->u1);
            z[i] = z[i] - b->a1 * oc_nint(z[i] * b->u1);
            break;
        case 2:
            x[i] = x[i] - b->a1 * oc_nint(x[i] * b->u1);
            y[i] = y[i] - b->b2 * oc_nint(y[i] * b->v2);
            z[i] = z[i] - b->c3 * oc_nint(z[i] * b->w3);
            break;
        case 3: {
            double xt = d[0] * b->u1 + d[1] * b->u2 + d[2] * b->u3;
            double yt = d[0] * b->v1 + d[1] * b->v2 + d[2] * b->v3;
            double zt = d[0] * b->w1 + d[1] * b->w2 + d[2] * b->w3;
            xt -= oc_nint(xt);
            yt -= oc_nint(yt);
            zt -= oc_nint(zt);
            x[i] = xt * b->a1 + yt * b->b1 + zt * b->c1;
            y[i] = xt * b->a2 + yt * b->b2 + zt * b->c2;
            z[i] = xt * b->a3 + yt * b->b3 + zt * b->c3;
            break;
        }
        default:
            break;
        }
    }
}

/* list.c:38-57: link cells only when every axis has >= 3 cells and there
 * are more than 27 cells. */
int oc_choose_link(const oc_param *p, const oc_box *b)
{
    int ratio = p->linkRatio < 1 ? 1 : (p->linkRatio > 5 ? 5 : p->linkRatio);
    double cutnb = p->cutOff + p->delr;
    int nx = (int)(b->pa * (double)ratio / cutnb);
    int ny = (int)(b->pb * (double)ratio / cutnb);
    int nz = (int)(b->pc * (double)ratio / cutnb);
    if (nx < 3 || ny < 3 || nz < 3) return 0;
    if (nx * ny * nz <= 27) return 0;
    return 1;
}

/* ---------------------------------------------------------------- */
/* Exclusions: list.c:108-894 (ordered) and 1125-1791 (symmetric).   */
/* ---------------------------------------------------------------- */

typedef struct { int *v; int n, cap; } ivec;

static void iv_push(ivec *a, int x)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 8;
        a->v = (int *)realloc(a->v, (size_t)a->cap * sizeof(int));
    }
    a->v[a->n++] = x;
}

static int iv_has(const ivec *a, int x)
{
    for (int k = 0; k < a->n; k++)
        if (a->v[k] == x) return 1;
    return 0;
}

/* Adds pair (i,j) to the exclusion table.  In ordered mode the pair is kept
 * at the smaller index (list.c:141-150); in symmetric mode at both
 * (list.c:1159-1181).  dedupe mirrors the reference: bonds and constraints
 * are appended unconditionally, angle and walk pairs only if absent. */
static int ex_add(ivec *ex, int i, int j, int sym, int dedupe)
{
    int ii = i, jj = j;
    if (!sym && jj < ii) { ii = j; jj = i; }
    if (dedupe && iv_has(&ex[ii], jj)) return 0;
    iv_push(&ex[ii], jj);
    if (sym) iv_push(&ex[jj], ii);
    return 1;
}

/* The exclusion builders with dihedrals, impropers and the 1-4 list
 * (list.c:108-894 ordered, 1125-1791 symmetric):
 *   bonds and constraints appended unconditionally (and recorded as
 *   connections), angle pairs (a,b) (a,c) (b,c) when absent; per dihedral
 *   the pairs (a,b) (a,c) (a,d) (b,c) (b,d) (c,d) when absent, where a new
 *   (a,d) also becomes 1-4 pair (a, d); impropers the same six pairs with
 *   no 1-4 pair; then the connectivity walk a-b-c(-d) over bonds and
 *   constraints: (a,c) when absent, and for each d the check
 *   nPair14 >= 5 nDihedral -> DIHE_NONPARAM_ERROR (110, list.c:835-836)
 *   before (a,d) is added as exclusion and 1-4 pair when absent.
 * Returns the number of exclusion entries (CSR off/lst), -110 / -111 on
 * the reference's errors, or -(needed+1000) when cap is too small.
 * *n14 receives the 1-4 pair count; list14 (room for 5 nDihedral pairs)
 * the pairs as (i, j). */
int oc_exclusions14(int n, int nBond, const int *ba, const int *bb,
                    int nConst, const int *ca, const int *cb,
                    int nAngle, const int *aa, const int *ab, const int *ac,
                    int nDihe, const int *d4, int nImpr, const int *i4,
                    int sym, int *off, int *lst, int cap, int *list14, int *n14)
{
    ivec *ex = (ivec *)calloc((size_t)n, sizeof(ivec));
    ivec *con = (ivec *)calloc((size_t)n, sizeof(ivec));
    int err = 0, np14 = 0;

    for (int i = 0; i < nBond; i++) {                  /* list.c:137-174 */
        ex_add(ex, ba[i], bb[i], sym, 0);
        iv_push(&con[ba[i]], bb[i]);
        iv_push(&con[bb[i]], ba[i]);
    }
    for (int i = 0; i < nConst; i++) {                 /* list.c:176-216 */
        ex_add(ex, ca[i], cb[i], sym, 0);
        iv_push(&con[ca[i]], cb[i]);
        iv_push(&con[cb[i]], ca[i]);
    }
    for (int i = 0; i < nAngle; i++) {                 /* list.c:218-326 */
        ex_add(ex, aa[i], ab[i], sym, 1);
        ex_add(ex, aa[i], ac[i], sym, 1);
        ex_add(ex, ab[i], ac[i], sym, 1);
    }
    for (int i = 0; i < nDihe; i++) {                  /* list.c:334-549 */
        const int a = d4[4 * i], b = d4[4 * i + 1], c = d4[4 * i + 2], d = d4[4 * i + 3];
        ex_add(ex, a, b, sym, 1);
        ex_add(ex, a, c, sym, 1);
        if (ex_add(ex, a, d, sym, 1)) {
            list14[2 * np14] = a;
            list14[2 * np14 + 1] = d;
            np14++;
        }
        ex_add(ex, b, c, sym, 1);
        ex_add(ex, b, d, sym, 1);
        ex_add(ex, c, d, sym, 1);
    }
    for (int i = 0; i < nImpr; i++) {                  /* list.c:551-762 */
        const int a = i4[4 * i], b = i4[4 * i + 1], c = i4[4 * i + 2], d = i4[4 * i + 3];
        ex_add(ex, a, b, sym, 1);
        ex_add(ex, a, c, sym, 1);
        ex_add(ex, a, d, sym, 1);
        ex_add(ex, b, c, sym, 1);
        ex_add(ex, b, d, sym, 1);
        ex_add(ex, c, d, sym, 1);
    }
    for (int ia = 0; ia < n && !err; ia++) {          /* list.c:764-862 */
        for (int j = 0; j < con[ia].n && !err; j++) {
            int ib = con[ia].v[j];
            for (int k = 0; k < con[ib].n && !err; k++) {
                int ic = con[ib].v[k];
                if (ic == ia) continue;
                ex_add(ex, ia, ic, sym, 1);
                for (int l = 0; l < con[ic].n; l++) {
                    int id = con[ic].v[l];
                    if (id == ia || id == ib) continue;
                    if (np14 >= 5 * nDihe) { err = 110; break; }
                    if (ex_add(ex, ia, id, sym, 1)) {
                        list14[2 * np14] = ia;
                        list14[2 * np14 + 1] = id;
                        np14++;
                    }
                }
            }
        }
    }

    int total = 0;
    for (int i = 0; i < n; i++) total += ex[i].n;
    if (!err && !sym && n > 0 && ex[n - 1].n != 0) err = 111;   /* list.c:877 */
    int ret;
    if (err) {
        ret = -err;
    } else if (total > cap) {
        ret = -total - 1000;
    } else {
        int o = 0;
        for (int i = 0; i < n; i++) {
            off[i] = o;
            for (int k = 0; k < ex[i].n; k++) lst[o++] = ex[i].v[k];
        }
        off[n] = o;
        ret = total;
    }
    if (n14) *n14 = np14;
    for (int i = 0; i < n; i++) { free(ex[i].v); free(con[i].v); }
    free(ex);
    free(con);
    return ret;
}

int oc_exclusions(int n, int nBond, const int *ba, const int *bb,
                  int nConst, const int *ca, const int *cb,
                  int nAngle, const int *aa, const int *ab, const int *ac,
                  int sym, int *off, int *lst, int cap)
{
    ivec *ex = (ivec *)calloc((size_t)n, sizeof(ivec));
    ivec *con = (ivec *)calloc((size_t)n, sizeof(ivec));
    int err = 0;

    for (int i = 0; i < nBond; i++) {                  /* list.c:137-174 */
        ex_add(ex, ba[i], bb[i], sym, 0);
        iv_push(&con[ba[i]], bb[i]);
        iv_push(&con[bb[i]], ba[i]);
    }
    for (int i = 0; i < nConst; i++) {                 /* list.c:176-216 */
        ex_add(ex, ca[i], cb[i], sym, 0);
        iv_push(&con[ca[i]], cb[i]);
        iv_push(&con[cb[i]], ca[i]);
    }
    for (int i = 0; i < nAngle; i++) {                 /* list.c:218-326 */
        ex_add(ex, aa[i], ab[i], sym, 1);
        ex_add(ex, aa[i], ac[i], sym, 1);
        ex_add(ex, ab[i], ac[i], sym, 1);
    }
    /* Connectivity walk, list.c:764-862.  A 1-4 path with no dihedral
     * parameters is DIHE_NONPARAM_ERROR (list.c:835-836). */
    for (int ia = 0; ia < n && !err; ia++) {
        for (int j = 0; j < con[ia].n && !err; j++) {
            int ib = con[ia].v[j];
            for (int k = 0; k < con[ib].n && !err; k++) {
                int ic = con[ib].v[k];
                if (ic == ia) continue;
                ex_add(ex, ia, ic, sym, 1);
                for (int l = 0; l < con[ic].n; l++) {
                    int id = con[ic].v[l];
                    if (id == ia || id == ib) continue;
                    err = 110;
                    break;
                }
            }
        }
    }

    int total = 0;
    for (int i = 0; i < n; i++) total += ex[i].n;
    if (!err && !sym && n > 0 && ex[n - 1].n != 0) err = 111;   /* list.c:877 */
    int ret;
    if (err) {
        ret = -err;
    } else if (total > cap) {
        ret = -total - 1000;
    } else {
        int o = 0;
        for (int i = 0; i < n; i++) {
            off[i] = o;
            for (int k = 0; k < ex[i].n; k++) lst[o++] = ex[i].v[k];
        }
        off[n] = o;
        ret = total;
    }
    for (int i = 0; i < n; i++) { free(ex[i].v); free(con[i].v); }
    free(ex);
    free(con);
    return ret;
}

static int is_excluded(const int *exOff, const int *exLst, int i, int j)
{
    for (int k = exOff[i]; k < exOff[i + 1]; k++)
        if (exLst[k] == j) return 1;
    return 0;
}

/* ---------------------------------------------------------------- */
/* O(N^2) Verlet list, list.c:896-1123.                              */
/* ---------------------------------------------------------------- */
long oc_verlet_list(const oc_param *p, const oc_box *b,
                    const double *x, const double *y, const double *z,
                    const int *frozen, const int *exOff, const int *exLst,
                    int *neighList, long cap, int *neighPair, int *neighOrder)
{
    int n = p->nAtom;
    double cutnb2 = p->cutOff + p->delr;
    cutnb2 = SQ(cutnb2);
    long offset = 0;
    for (int i = 0; i < n; i++) { neighPair[i] = 0; neighOrder[i] = i; }
    for (int i = 0; i < n - 1; i++) {
        for (int j = i + 1; j < n; j++) {
            double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
            double r2 = oc_dist(b, d);
            if (r2 <= cutnb2) {
                int excl = (frozen[i] * frozen[j]) ? 1 : is_excluded(exOff, exLst, i, j);
                if (!excl) {
                    if (offset < cap) neighList[offset] = j;
                    neighPair[i]++;
                    offset++;
                }
            }
        }
    }
    return offset;
}

/* ---------------------------------------------------------------- */
/* Link-cell list, list.c:1793-2619.                                 */
/* ---------------------------------------------------------------- */

/* The first 14 entries of the transx/transy/transz tables
 * (list.c:1797-1855): the half shell searched when linkRatio is 1. */
static const int LC_TX[14] = {0, 1, 0, 0, -1, 1, 0, -1, 1, 0, -1, 1, -1, 1};
static const int LC_TY[14] = {0, 0, 1, 0, 1, 1, -1, 0, 0, 1, -1, -1, 1, 1};
static const int LC_TZ[14] = {0, 0, 0, 1, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1};

long oc_link_cell_list(const oc_param *p, const oc_box *b, int build,
                       double *x, double *y, double *z,
                       const int *frozen, const int *exOff, const int *exLst,
                       int *neighList, long cap, int *neighPair, int *neighOrder)
{
    int n = p->nAtom;
    double cutnb = p->cutOff + p->delr;
    double cutnb2 = SQ(cutnb);
    if (p->linkRatio != 1) return -412;   /* only the 14-cell half shell */
    int nlcx = (int)(b->pa * 1.0 / cutnb);
    int nlcy = (int)(b->pb * 1.0 / cutnb);
    int nlcz = (int)(b->pc * 1.0 / cutnb);
    if (nlcx < 3 || nlcy < 3 || nlcz < 3) return -412;   /* list.c:1897-1905 */
    double dnx = (double)nlcx, dny = (double)nlcy, dnz = (double)nlcz;
    int ncell = nlcx * nlcy * nlcz;

    if (build) oc_image_update(b, n, x, y, z);           /* list.c:1927 */

    double *xu = (double *)malloc((size_t)n * sizeof(double));
    double *yu = (double *)malloc((size_t)n * sizeof(double));
    double *zu = (double *)malloc((size_t)n * sizeof(double));
    int *link = (int *)malloc((size_t)n * sizeof(int));
    int *head = (int *)malloc((size_t)ncell * sizeof(int));
    for (int k = 0; k < ncell; k++) head[k] = -1;

    for (int i = 0; i < n; i++) {                          /* list.c:1932-1937 */
        xu[i] = (x[i] * b->u1 + y[i] * b->u2 + z[i] * b->u3) + 0.5;
        yu[i] = (x[i] * b->v1 + y[i] * b->v2 + z[i] * b->v3) + 0.5;
        zu[i] = (x[i] * b->w1 + y[i] * b->w2 + z[i] * b->w3) + 0.5;
    }
    for (int i = 0; i < n; i++) {                          /* list.c:1942-1967 */
        /* (int)dnlcx * xu: the cast binds to dnlcx, the product truncates
         * on assignment. */
        int ix = (int)(dnx * xu[i]);
        int iy = (int)(dny * yu[i]);
        int iz = (int)(dnz * zu[i]);
        if (ix > nlcx - 1) ix = nlcx - 1;
        if (iy > nlcy - 1) iy = nlcy - 1;
        if (iz > nlcz - 1) iz = nlcz - 1;
        int ic = ix + nlcx * (iy + nlcy * iz);
        link[i] = head[ic];
        head[ic] = i;
    }
    int io = 0;                                            /* list.c:1969-1980 */
    for (int k = 0; k < ncell; k++)
        for (int i = head[k]; i != -1; i = link[i]) neighOrder[io++] = i;
    for (int i = 0; i < n; i++) neighPair[i] = 0;

    long offset = 0;
    int ix = 0, iy = 0, iz = 0;
    for (int k = 0; k < ncell; k++) {                      /* list.c:2130-2261 */
        for (int i = head[k]; i != -1; i = link[i]) {
            for (int l = 0; l < 14; l++) {
                double cx = 0., cy = 0., cz = 0.;
                int jx = ix + LC_TX[l], jy = iy + LC_TY[l], jz = iz + LC_TZ[l];
                if (jx > nlcx - 1) { jx -= nlcx; cx = 1.; } else if (jx < 0) { jx += nlcx; cx = -1.; }
                if (jy > nlcy - 1) { jy -= nlcy; cy = 1.; } else if (jy < 0) { jy += nlcy; cy = -1.; }
                if (jz > nlcz - 1) { jz -= nlcz; cz = 1.; } else if (jz < 0) { jz += nlcz; cz = -1.; }
                int ll = jx + nlcx * (jy + nlcy * jz);
                int j = (k == ll) ? link[i] : head[ll];
                for (; j != -1; j = link[j]) {
                    double xt = xu[j] - xu[i] + cx;
                    double yt = yu[j] - yu[i] + cy;
                    double zt = zu[j] - zu[i] + cz;
                    double xd = xt * b->a1 + yt * b->b1 + zt * b->c1;
                    double yd = xt * b->a2 + yt * b->b2 + zt * b->c2;
                    double zd = xt * b->a3 + yt * b->b3 + zt * b->c3;
                    double r2 = SQ(xd) + SQ(yd) + SQ(zd);
                    if (r2 <= cutnb2) {
                        int excl = (frozen[i] * frozen[j]) ? 1 : is_excluded(exOff, exLst, i, j);
                        if (!excl) {
                            if (offset < cap) neighList[offset] = j;
                            neighPair[i]++;
                            offset++;
                        }
                    }
                }
            }
        }
        if (++ix > nlcx - 1) {
            ix = 0;
            if (++iy > nlcy - 1) { iy = 0; iz++; }
        }
    }
    free(xu); free(yu); free(zu); free(link); free(head);
    return offset;
}

/* ---------------------------------------------------------------- */
/* Pair functors.                                                    */
/* ---------------------------------------------------------------- */

/* ewald.c:356-369 */
double oc_ewald_dir(double alpha, double qel, double r, double rt, double *dE)
{
    double ar = alpha * r;
    double e = qel * erfc(ar) * rt;
    double d = 2.0 * qel * alpha * exp(-SQ(ar)) / OC_SQRTPI;
    *dE = -(e + d) * rt;
    return e;
}

/* ewald.c:371-384 */
double oc_ewald_corr(double alpha, double qel, double r, double rt, double *dE)
{
    double ar = alpha * r;
    double e = -qel * erf(ar) * rt;
    double d = 2.0 * qel * alpha * exp(-SQ(ar)) / OC_SQRTPI;
    *dE = -(e + d) * rt;
    return e;
}

/* vdw.c:144-184: 4 eps ((s/r)^12 - (s/r)^6) with the CHARMM switch on
 * (ron, rc]; no rc test inside. */
double oc_vdw_switch(const oc_param *p, double veps, double vsig, double r2,
                     double rt, double *dE)
{
    double s6 = vsig * rt;
    s6 = SQ(s6) * s6 * (SQ(s6) * s6);
    double s12 = SQ(s6);
    if (r2 <= p->cutOn2) {
        *dE = 24. * veps * rt * (s6 - 2. * s12);
        return 4. * veps * (s12 - s6);
    }
    double sw1 = p->cutOff2 - r2;
    double S = SQ(sw1) * (p->cutOff2 + 2. * r2 - 3. * p->cutOn2) * p->switch2;
    double dS = 12. * r2 * sw1 * (p->cutOn2 - r2) * p->switch2;
    double pv = 4. * veps * (s12 - s6);
    double dpv = 24. * veps * (s6 - 2. * s12);
    *dE = (pv * dS + dpv * S) * rt;
    return pv * S;
}

/* elec.c:137-263: SHIFT1, SHIFT2, SWITCH Coulomb functors. */
static double oc_coulomb(const oc_param *p, int type, double qel, double r2,
                         double rt, double *dE)
{
    double pe = p->chargeConst * qel * rt;
    if (type == 2) {
        double s1 = rt * r2 * p->rcutOff, s2 = r2 * p->rcutOff2;
        double f = 1. - 2. * s1 + s2, df = 2. * (s2 - s1);
        *dE = pe * rt * (df - f);
        return pe * f;
    }
    if (type == 3) {
        double s1 = r2 * p->rcutOff2, s2 = SQ(s1);
        double f = 1. - 2. * s1 + s2, df = 4. * (s2 - s1);
        *dE = pe * rt * (df - f);
        return pe * f;
    }
    if (type == 4) {
        if (r2 <= p->cutOn2) {
            *dE = -pe * rt;
            return pe;
        }
        double sw1 = p->cutOff2 - r2;
        double f = SQ(sw1) * (p->cutOff2 + 2. * r2 - 3. * p->cutOn2) * p->switch2;
        double df = 12. * r2 * sw1 * (p->cutOn2 - r2) * p->switch2;
        *dE = pe * rt * (df - f);
        return pe * f;
    }
    *dE = 0.;
    return 0.;
}

/* energy.c:480-627: real-space loop over the list in neighOrder order,
 * then the erf correction over excluded pairs with j > i. */
void oc_ewald_real(const oc_param *p, const oc_box *b,
                   const double *x, const double *y, const double *z,
                   double *fx, double *fy, double *fz,
                   const double *q, const double *eps, const double *sig,
                   const int *neighList, const int *neighPair, const int *neighOrder,
                   const int *exOff, const int *exLst, oc_energy *e)
{
    double eDir = 0., virDir = 0., eVdw = 0., virVdw = 0., eCorr = 0., virCorr = 0.;
    long offset = 0;
    int nin = 0, nlist = 0;
    for (int l = 0; l < p->nAtom; l++) {
        int i = neighOrder[l];
        double fxi = 0., fyi = 0., fzi = 0.;
        for (int k = 0; k < neighPair[i]; k++) {
            int j = neighList[offset++];
            double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
            double r2 = oc_dist(b, d);
            double r = sqrt(r2);
            double rt = 1. / r;
            nlist++;
            if (r2 <= p->cutOff2) {
                double dEl, dVw;
                double qel = p->chargeConst * q[i] * q[j];
                eDir += oc_ewald_dir(p->alpha, qel, r, rt, &dEl);
                double veps = eps[i] * eps[j];
                double vsig = sig[i] + sig[j];
                eVdw += oc_vdw_switch(p, veps, vsig, r2, rt, &dVw);
                virDir += dEl * r;
                virVdw += dVw * r;
                double fxj = (dEl + dVw) * d[0] * rt;
                double fyj = (dEl + dVw) * d[1] * rt;
                double fzj = (dEl + dVw) * d[2] * rt;
                fxi += fxj; fyi += fyj; fzi += fzj;
                fx[j] += -fxj; fy[j] += -fyj; fz[j] += -fzj;
                nin++;
            }
        }
        fx[i] += fxi; fy[i] += fyi; fz[i] += fzi;
    }
    for (int i = 0; i < p->nAtom; i++) {
        double fxi = 0., fyi = 0., fzi = 0.;
        for (int k = exOff[i]; k < exOff[i + 1]; k++) {
            int j = exLst[k];
            if (j > i) {
                double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
                double r2 = oc_dist(b, d);
                double r = sqrt(r2);
                double rt = 1. / r;
                double dC;
                double qel = p->chargeConst * q[i] * q[j];
                eCorr += oc_ewald_corr(p->alpha, qel, r, rt, &dC);
                virCorr += dC * r;
                double fxj = dC * d[0] * rt, fyj = dC * d[1] * rt, fzj = dC * d[2] * rt;
                fxi += fxj; fyi += fyj; fzi += fzj;
                fx[j] += -fxj; fy[j] += -fyj; fz[j] += -fzj;
            }
        }
        fx[i] += fxi; fy[i] += fyi; fz[i] += fzi;
    }
    e->eDir = eDir; e->virDir = virDir;
    e->eVdw = eVdw; e->virVdw = virVdw;
    e->eCorr = eCorr; e->virCorr = virCorr;
    e->nInCut = nin; e->nListed = nlist;
}

/* energy.c:216-345 */
void oc_nonbond(const oc_param *p, const oc_box *b, int elecType,
                const double *x, const double *y, const double *z,
                double *fx, double *fy, double *fz,
                const double *q, const double *eps, const double *sig,
                const int *neighList, const int *neighPair, const int *neighOrder,
                double *eElec, double *eVdw, double *virElec, double *virVdw)
{
    double el = 0., ev = 0., vel = 0., vvw = 0.;
    long offset = 0;
    for (int l = 0; l < p->nAtom; l++) {
        int i = neighOrder[l];
        double fxi = 0., fyi = 0., fzi = 0.;
        for (int k = 0; k < neighPair[i]; k++) {
            int j = neighList[offset++];
            double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
            double r2 = oc_dist(b, d);
            double r = sqrt(r2);
            double rt = 1. / r;
            if (r2 <= p->cutOff2) {
                double dEl, dVw;
                el += oc_coulomb(p, elecType, q[i] * q[j], r2, rt, &dEl);
                ev += oc_vdw_switch(p, eps[i] * eps[j], sig[i] + sig[j], r2, rt, &dVw);
                vel += dEl * r;
                vvw += dVw * r;
                double fxj = (dEl + dVw) * d[0] * rt;
                double fyj = (dEl + dVw) * d[1] * rt;
                double fzj = (dEl + dVw) * d[2] * rt;
                fxi += fxj; fyi += fyj; fzi += fzj;
                fx[j] += -fxj; fy[j] += -fyj; fz[j] += -fzj;
            }
        }
        fx[i] += fxi; fy[i] += fyi; fz[i] += fzi;
    }
    *eElec = el; *eVdw = ev; *virElec = vel; *virVdw = vvw;
}

/* ---------------------------------------------------------------- */
/* FFT: mixed-radix recursive Cooley-Tukey (any length).             */
/* ---------------------------------------------------------------- */

static int small_factor(int n)
{
    for (int f = 2; f * f <= n; f++)
        if (n % f == 0) return f;
    return n;
}

/* out[k] = sum_t in[t*s] w^(t k), w = exp(sign 2 pi i / n); tw is the
 * table exp(sign 2 pi i j / n0) with n dividing n0. */
static void fft_rec(const ocplx *in, ocplx *out, int n, int s, const ocplx *tw,
                    int n0, ocplx *tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = small_factor(n), m = n / p;
    for (int r = 0; r < p; r++) fft_rec(in + r * s, out + r * m, m, s * p, tw, n0, tmp);
    int step = n0 / n;
    for (int k = 0; k < m; k++) {
        for (int qd = 0; qd < p; qd++) {
            ocplx acc = 0;
            int kk = k + qd * m;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * kk) % n;
                acc += out[r * m + k] * tw[e * step];
            }
            tmp[qd] = acc;
        }
        for (int qd = 0; qd < p; qd++) out[k + qd * m] = tmp[qd];
    }
}

static void fft1d_lines(ocplx *a, int n, int count, int stride, int dist, int sign)
{
    ocplx *tw = (ocplx *)malloc((size_t)n * sizeof(ocplx));
    ocplx *in = (ocplx *)malloc((size_t)n * sizeof(ocplx));
    ocplx *out = (ocplx *)malloc((size_t)n * sizeof(ocplx));
    ocplx *tmp = (ocplx *)malloc((size_t)n * sizeof(ocplx));
    for (int j = 0; j < n; j++) {
        double ang = (double)sign * OC_TWOPI * (double)j / (double)n;
        tw[j] = cos(ang) + I * sin(ang);
    }
    for (int c = 0; c < count; c++) {
        ocplx *base = a + (long)c * dist;
        for (int t = 0; t < n; t++) in[t] = base[(long)t * stride];
        fft_rec(in, out, n, 1, tw, n, tmp);
        for (int t = 0; t < n; t++) base[(long)t * stride] = out[t];
    }
    free(tw); free(in); free(out); free(tmp);
}

int oc_fft3d(double *data, int n1, int n2, int n3, int sign)
{
    ocplx *a = (ocplx *)data;
    /* axis 3 (contiguous) */
    fft1d_lines(a, n3, n1 * n2, 1, n3, sign);
    /* axis 2 */
    for (int i = 0; i < n1; i++) fft1d_lines(a + (long)i * n2 * n3, n2, n3, n3, 1, sign);
    /* axis 1 */
    fft1d_lines(a, n1, n2 * n3, n2 * n3, 1, sign);
    return 0;
}

/* ---------------------------------------------------------------- */
/* SPME, spme.c:170-699.                                             */
/* ---------------------------------------------------------------- */

/* spme.c:269-355 for one coordinate, same operation order. */
void oc_bspline(double s, int p, double *bsp, double *bsd)
{
    int is = (int)s;
    bsd[0] = 1.0;
    bsd[1] = -1.0;
    bsp[0] = s - is;
    bsp[1] = 1.0 - bsp[0];
    for (int k = 2; k < p; k++) {
        bsp[k] = 0.0;
        for (int j = k; j > 0; j--) {
            if (k == p - 1) bsd[j] = bsp[j] - bsp[j - 1];
            double ts = s + (double)j - is;
            bsp[j] = (ts * bsp[j] + ((double)(k + 1) - ts) * bsp[j - 1]) / ((double)k);
        }
        if (k == p - 1) bsd[0] = bsp[0];
        double ts = s - is;
        bsp[0] = ts * bsp[0] / ((double)k);
    }
}

/* spme.c:170-267: |bspc(i)|^2 with bspc(i) = e^{2 pi i i (p-1)/m} /
 * sum_{k=0}^{p-2} M_p(k+1) e^{2 pi i i k/m}. */
void oc_bspmod(int m, int p, double *mod2)
{
    double *bsp = (double *)calloc((size_t)(p + 1), sizeof(double));
    ocplx *epl = (ocplx *)malloc((size_t)m * sizeof(ocplx));
    bsp[0] = 0.0;
    bsp[1] = 1.0;
    for (int k = 2; k < p; k++) {
        bsp[k] = 0.0;
        for (int j = k; j > 0; j--)
            bsp[j] = ((double)j * bsp[j] + (double)(k + 1 - j) * bsp[j - 1]) / ((double)k);
    }
    epl[0] = 1.0 + I * 0.0;
    for (int i = 1; i <= m / 2; i++) {
        epl[i] = cexp(I * OC_TWOPI * (double)i / (double)m);
        epl[m - i] = conj(epl[i]);
    }
    for (int i = 0; i < m; i++) {
        ocplx coeff = 0. + I * 0.0;
        for (int k = 0; k < p - 1; k++) coeff += bsp[k + 1] * epl[(i * k) % m];
        ocplx c = epl[(i * (p - 1)) % m] / coeff;
        mod2[i] = creal(c * conj(c));
    }
    free(bsp);
    free(epl);
}

static double spme_core(const oc_param *p, const oc_box *b,
                        const double *x, const double *y, const double *z,
                        double *fx, double *fy, double *fz, const double *q,
                        double stress[6], double *virRec, double *pot);

double oc_spme(const oc_param *p, const oc_box *b,
               const double *x, const double *y, const double *z,
               double *fx, double *fy, double *fz, const double *q,
               double stress[6], double *virRec)
{
    return spme_core(p, b, x, y, z, fx, fy, fz, q, stress, virRec, NULL);
}

/* dE_rec/dq_l of the SPME energy (self and net-charge terms excluded):
 * E_rec = (2pi/V) C sum_g Q(g) (theta*Q)(g) is quadratic in the mesh
 * charges Q, and Q(g) = sum_l q_l M_l(g), so dE/dq_l = 2 (2pi/V) C
 * sum_g (theta*Q)(g) M_l(g): the convolved mesh gathered with the spline
 * weights instead of their derivatives. */
void oc_spme_pot(const oc_param *p, const oc_box *b, const double *x, const double *y,
                 const double *z, const double *q, double *dq)
{
    int n = p->nAtom;
    double *f = (double *)calloc((size_t)3 * n, sizeof(double));
    double st[6], vr;
    spme_core(p, b, x, y, z, f, f + n, f + 2 * n, q, st, &vr, dq);
    free(f);
}

static double spme_core(const oc_param *p, const oc_box *b,
                        const double *x, const double *y, const double *z,
                        double *fx, double *fy, double *fz, const double *q,
                        double stress[6], double *virRec, double *pot)
{
    int n = p->nAtom, m1 = p->m1, m2 = p->m2, m3 = p->m3, P = p->nbsp;
    long G = (long)m1 * m2 * m3;
    double *sx = (double *)malloc((size_t)n * 3 * sizeof(double));
    double *bp = (double *)malloc((size_t)n * 3 * P * sizeof(double));
    double *bd = (double *)malloc((size_t)n * 3 * P * sizeof(double));
    double *qsp = (double *)calloc((size_t)G, sizeof(double));
    ocplx *ft = (ocplx *)malloc((size_t)G * sizeof(ocplx));
    double *b1 = (double *)malloc((size_t)m1 * sizeof(double));
    double *b2 = (double *)malloc((size_t)m2 * sizeof(double));
    double *b3 = (double *)malloc((size_t)m3 * sizeof(double));
    oc_bspmod(m1, P, b1);
    oc_bspmod(m2, P, b2);
    oc_bspmod(m3, P, b3);

    /* self and net-charge terms every call (spme.c:376-389, 680) */
    double systq = 0., eself = 0.;
    for (int i = 0; i < n; i++) { systq += q[i]; eself += SQ(q[i]); }
    eself = -p->chargeConst * p->alpha * eself / OC_SQRTPI;

    double rVol = OC_TWOPI / b->vol;
    double rAlpha2 = -0.25 / SQ(p->alpha);
    double rc = fmin((double)m1 * b->u, (double)m2 * b->v);
    rc = fmin(rc, (double)m3 * b->w);
    rc = rc * 1.05 * OC_TWOPI;
    double rc2 = SQ(rc);

    for (int i = 0; i < n; i++) {                              /* spme.c:434-439 */
        sx[3 * i + 0] = (double)m1 * (x[i] * b->u1 + y[i] * b->u2 + z[i] * b->u3 + 0.5);
        sx[3 * i + 1] = (double)m2 * (x[i] * b->v1 + y[i] * b->v2 + z[i] * b->v3 + 0.5);
        sx[3 * i + 2] = (double)m3 * (x[i] * b->w1 + y[i] * b->w2 + z[i] * b->w3 + 0.5);
        for (int a = 0; a < 3; a++)
            oc_bspline(sx[3 * i + a], P, bp + (3 * i + a) * P, bd + (3 * i + a) * P);
    }
    for (int l = 0; l < n; l++) {                              /* spme.c:449-491 */
        const double *B1 = bp + (3 * l) * P, *B2 = B1 + P, *B3 = B2 + P;
        for (int i = 0; i < P; i++) {
            int ii = (int)sx[3 * l] - i;
            if (ii >= m1) ii = 0;
            if (ii < 0) ii += m1;
            int o1 = m2 * ii;
            for (int j = 0; j < P; j++) {
                int jj = (int)sx[3 * l + 1] - j;
                if (jj >= m2) jj = 0;
                if (jj < 0) jj += m2;
                int o2 = m3 * (jj + o1);
                for (int k = 0; k < P; k++) {
                    int kk = (int)sx[3 * l + 2] - k;
                    if (kk >= m3) kk = 0;
                    if (kk < 0) kk += m3;
                    qsp[kk + o2] += q[l] * B1[i] * B2[j] * B3[k];
                }
            }
        }
    }
    for (long g = 0; g < G; g++) ft[g] = qsp[g] + I * 0.0;
    oc_fft3d((double *)ft, m1, m2, m3, +1);                    /* FFTW_BACKWARD */

    double st[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < m1; i++) {                             /* spme.c:505-583 */
        int ii = i > m1 / 2 ? i - m1 : i;
        double tt = OC_TWOPI * (double)ii;
        double r1x = tt * b->u1, r1y = tt * b->u2, r1z = tt * b->u3;
        double bm1 = b1[i];
        for (int j = 0; j < m2; j++) {
            int jj = j > m2 / 2 ? j - m2 : j;
            tt = OC_TWOPI * (double)jj;
            double r2x = r1x + (tt * b->v1), r2y = r1y + (tt * b->v2), r2z = r1z + (tt * b->v3);
            double bm2 = bm1 * b2[j];
            for (int k = 0; k < m3; k++) {
                int kk = k > m3 / 2 ? k - m3 : k;
                long g = k + (long)m3 * (j + (long)m2 * i);
                tt = OC_TWOPI * (double)kk;
                double rx = r2x + (tt * b->w1), ry = r2y + (tt * b->w2), rz = r2z + (tt * b->w3);
                double bm3 = bm2 * b3[k];
                double rm = SQ(rx) + SQ(ry) + SQ(rz);
                if (rm > DBL_EPSILON && rm <= rc2) {
                    double rrm = 1.0 / rm;
                    ocplx cam = bm3 * exp(rAlpha2 * rm) * rrm * ft[g];
                    double vam = 2.0 * (rrm - rAlpha2) * creal(cam * conj(ft[g]));
                    st[0] -= vam * rx * rx;
                    st[1] -= vam * rx * ry;
                    st[2] -= vam * rx * rz;
                    st[3] -= vam * ry * ry;
                    st[4] -= vam * ry * rz;
                    st[5] -= vam * rz * rz;
                    ft[g] = cam;
                } else {
                    ft[g] = 0.0 + I * 0.0;
                }
            }
        }
    }
    oc_fft3d((double *)ft, m1, m2, m3, -1);                    /* FFTW_FORWARD */

    double fact1 = -2.0 * rVol * p->chargeConst;
    for (int l = 0; l < n; l++) {                              /* spme.c:593-646 */
        const double *B1 = bp + (3 * l) * P, *B2 = B1 + P, *B3 = B2 + P;
        const double *D1 = bd + (3 * l) * P, *D2 = D1 + P, *D3 = D2 + P;
        double phi = 0.;
        for (int i = 0; i < P; i++) {
            int ii = (int)sx[3 * l] - i;
            if (ii >= m1) ii = 0;
            if (ii < 0) ii += m1;
            int o1 = m2 * ii;
            for (int j = 0; j < P; j++) {
                int jj = (int)sx[3 * l + 1] - j;
                if (jj >= m2) jj = 0;
                if (jj < 0) jj += m2;
                int o2 = m3 * (jj + o1);
                for (int k = 0; k < P; k++) {
                    int kk = (int)sx[3 * l + 2] - k;
                    if (kk >= m3) kk = 0;
                    if (kk < 0) kk += m3;
                    double qt = creal(ft[kk + o2]);
                    double fbx = qt * D1[i] * B2[j] * B3[k] * (double)m1;
                    double fby = qt * B1[i] * D2[j] * B3[k] * (double)m2;
                    double fbz = qt * B1[i] * B2[j] * D3[k] * (double)m3;
                    fx[l] += fact1 * q[l] * (fbx * b->u1 + fby * b->v1 + fbz * b->w1);
                    fy[l] += fact1 * q[l] * (fbx * b->u2 + fby * b->v2 + fbz * b->w2);
                    fz[l] += fact1 * q[l] * (fbx * b->u3 + fby * b->v3 + fbz * b->w3);
                    phi += qt * B1[i] * B2[j] * B3[k];
                }
            }
        }
        if (pot) pot[l] = 2.0 * rVol * p->chargeConst * phi;
    }
    double fxm = 0., fym = 0., fzm = 0.;                       /* spme.c:648-670 */
    for (int l = 0; l < n; l++) { fxm += fx[l]; fym += fy[l]; fzm += fz[l]; }
    fxm /= (double)n; fym /= (double)n; fzm /= (double)n;
    for (int l = 0; l < n; l++) { fx[l] -= fxm; fy[l] -= fym; fz[l] -= fzm; }

    ocplx et = 0.0 + I * 0.0;                                  /* spme.c:674-697 */
    for (long g = 0; g < G; g++) et += ft[g] * qsp[g];
    double eNN = -(0.5 * OC_PI * p->chargeConst) * (SQ(systq / p->alpha) / b->vol);
    fact1 = rVol * p->chargeConst;
    double eRec = creal(et);
    stress[0] = fact1 * (st[0] + eRec) + eNN;
    stress[1] = fact1 * st[1];
    stress[2] = fact1 * st[2];
    stress[3] = fact1 * (st[3] + eRec) + eNN;
    stress[4] = fact1 * st[4];
    stress[5] = fact1 * (st[5] + eRec) + eNN;
    eRec = fact1 * eRec + eself + eNN;
    *virRec = -(stress[0] + stress[3] + stress[5]);

    free(sx); free(bp); free(bd); free(qsp); free(ft); free(b1); free(b2); free(b3);
    return eRec;
}

/* Classic Ewald reciprocal sum, E = (2pi/V) C sum_{m != 0} exp(-k^2/4a^2)/k^2
 * |S(k)|^2 with k = 2pi (m1 u + m2 v + m3 w), over the full cube |m_i| <=
 * kmax, plus the same self and net-charge terms as SPME (Essmann et al.
 * 1995, eqs. 2.4-2.5).  Forces are added to fx,fy,fz (no mean removal). */
double oc_ewald_classic(const oc_param *p, const oc_box *b, int kmax,
                        const double *x, const double *y, const double *z,
                        double *fx, double *fy, double *fz, const double *q)
{
    int n = p->nAtom;
    double fac = OC_TWOPI / b->vol * p->chargeConst;
    double e = 0.;
    for (int a = -kmax; a <= kmax; a++)
        for (int c = -kmax; c <= kmax; c++)
            for (int d = -kmax; d <= kmax; d++) {
                if (!a && !c && !d) continue;
                double kx = OC_TWOPI * (a * b->u1 + c * b->v1 + d * b->w1);
                double ky = OC_TWOPI * (a * b->u2 + c * b->v2 + d * b->w2);
                double kz = OC_TWOPI * (a * b->u3 + c * b->v3 + d * b->w3);
                double k2 = kx * kx + ky * ky + kz * kz;
                double f = exp(-k2 / (4. * SQ(p->alpha))) / k2;
                double sr = 0., si = 0.;
                for (int j = 0; j < n; j++) {
                    double ph = kx * x[j] + ky * y[j] + kz * z[j];
                    sr += q[j] * cos(ph);
                    si += q[j] * sin(ph);
                }
                e += f * (sr * sr + si * si);
                for (int j = 0; j < n; j++) {
                    double ph = kx * x[j] + ky * y[j] + kz * z[j];
                    /* Im(conj(S) e^{i k r_j}) = sr sin - si cos */
                    double im = sr * sin(ph) - si * cos(ph);
                    double g = fac * f * 2. * q[j] * im;
                    fx[j] += g * kx;
                    fy[j] += g * ky;
                    fz[j] += g * kz;
                }
            }
    double systq = 0., q2 = 0.;
    for (int j = 0; j < n; j++) { systq += q[j]; q2 += q[j] * q[j]; }
    double eself = -p->chargeConst * p->alpha * q2 / OC_SQRTPI;
    double eNN = -(0.5 * OC_PI * p->chargeConst) * (SQ(systq / p->alpha) / b->vol);
    return fac * e + eself + eNN;
}

/* ---------------------------------------------------------------- */
/* Velocity Verlet NVE + SHAKE, integrate.c:1518-1634, shake.c.      */
/* ---------------------------------------------------------------- */

static double oc_dist2(const oc_box *b, double *dx, double *dy, double *dz)
{
    double d[3] = {*dx, *dy, *dz};
    double r2 = oc_dist(b, d);     /* utils.c:82-122 is dist() on three scalars */
    *dx = d[0]; *dy = d[1]; *dz = d[2];
    return r2;
}

/* shake.c:209-357 */
/* Optional record of each constraint's convergence test at each sweep
 * (conv[sweep * nConst + c]), for the tests of the stopping rule. */
static unsigned char *g_conv;
static int g_conv_cap;

static int oc_shake_r(const oc_param *p, const oc_box *b, int nc, const int *ca,
                      const int *cb, const double *rc2, double *x, double *y,
                      double *z, double *vx, double *vy, double *vz,
                      const double *ddx, const double *ddy, const double *ddz,
                      const double *rmass, const int *nAtConst, double *virshake,
                      int *ncycle)
{
    int n = p->nAtom;
    double *xt = (double *)malloc((size_t)n * 3 * sizeof(double));
    double *yt = xt + n, *zt = yt + n;
    double *rt2 = (double *)malloc((size_t)nc * 4 * sizeof(double));
    double *dtx = rt2 + nc, *dty = dtx + nc, *dtz = dty + nc;
    int icycle = 0, conv = 0;
    double dt = p->timeStep;
    *virshake = 0.;
    do {
        double maxdist = 0.;
        for (int i = 0; i < nc; i++) {
            int ia = ca[i], ib = cb[i];
            dtx[i] = x[ib] - x[ia];
            dty[i] = y[ib] - y[ia];
            dtz[i] = z[ib] - z[ia];
            rt2[i] = oc_dist2(b, &dtx[i], &dty[i], &dtz[i]);
            double dd = fabs(rt2[i] - rc2[i]) / sqrt(rc2[i]);
            maxdist = maxdist >= dd ? maxdist : dd;
            if (g_conv && icycle < g_conv_cap) g_conv[(long)icycle * nc + i] = 0.5 * dd < p->tolShake;
        }
        maxdist = 0.5 * maxdist;
        if (maxdist < p->tolShake) conv = 1;
        if (!conv) {
            for (int i = 0; i < n; i++) { xt[i] = 0.; yt[i] = 0.; zt[i] = 0.; }
            for (int i = 0; i < nc; i++) {
                int ia = ca[i], ib = cb[i];
                double trmi = dt * rmass[ia], trmj = dt * rmass[ib];
                double lambda = -(rc2[i] - rt2[i]) /
                    (dt * (trmi + trmj) * ((ddx[i] * dtx[i]) + (ddy[i] * dty[i]) + (ddz[i] * dtz[i])));
                *virshake += lambda * (SQ(ddx[i]) + SQ(ddy[i]) + SQ(ddz[i]));
                double li = 0.5 * lambda * trmi;
                xt[ia] += ddx[i] * li; yt[ia] += ddy[i] * li; zt[ia] += ddz[i] * li;
                double lj = -0.5 * lambda * trmj;
                xt[ib] += ddx[i] * lj; yt[ib] += ddy[i] * lj; zt[ib] += ddz[i] * lj;
            }
            for (int i = 0; i < nc; i++) {
                int ia = ca[i], ib = cb[i];
                double nia = 1.0 / (double)nAtConst[ia], nib = 1.0 / (double)nAtConst[ib];
                x[ia] += dt * xt[ia] * nia; y[ia] += dt * yt[ia] * nia; z[ia] += dt * zt[ia] * nia;
                x[ib] += dt * xt[ib] * nib; y[ib] += dt * yt[ib] * nib; z[ib] += dt * zt[ib] * nib;
                vx[ia] += xt[ia] * nia; vy[ia] += yt[ia] * nia; vz[ia] += zt[ia] * nia;
                vx[ib] += xt[ib] * nib; vy[ib] += yt[ib] * nib; vz[ib] += zt[ib] * nib;
            }
        }
        icycle++;
    } while (!conv && icycle < p->maxCycle);
    free(xt);
    free(rt2);
    if (ncycle) *ncycle = icycle;
    return conv ? 0 : 311;
}

/* shake.c:359-457 */
static int oc_shake_v(const oc_param *p, int nc, const int *ca, const int *cb,
                      double *vx, double *vy, double *vz, const double *ddx,
                      const double *ddy, const double *ddz, const double *rmass,
                      const int *nAtConst, int *ncycle)
{
    int n = p->nAtom;
    double *xt = (double *)malloc((size_t)n * 3 * sizeof(double));
    double *yt = xt + n, *zt = yt + n;
    int icycle = 0, conv = 0;
    double tolvel = p->tolShake * p->rTimeStep;
    do {
        for (int i = 0; i < n; i++) { xt[i] = 0.; yt[i] = 0.; zt[i] = 0.; }
        double maxdist = 0.;
        for (int i = 0; i < nc; i++) {
            int ia = ca[i], ib = cb[i];
            double trmi = 0.5 * p->timeStep * rmass[ia];
            double trmj = 0.5 * p->timeStep * rmass[ib];
            double lambda = (ddx[i] * (vx[ib] - vx[ia]) + ddy[i] * (vy[ib] - vy[ia]) +
                             ddz[i] * (vz[ib] - vz[ia])) /
                            ((trmi + trmj) * (SQ(ddx[i]) + SQ(ddy[i]) + SQ(ddz[i])));
            double al = fabs(lambda);
            maxdist = maxdist >= al ? maxdist : al;
            if (g_conv && icycle < g_conv_cap) g_conv[(long)icycle * nc + i] = al < tolvel;
            double li = lambda * trmi;
            xt[ia] += ddx[i] * li; yt[ia] += ddy[i] * li; zt[ia] += ddz[i] * li;
            double lj = -lambda * trmj;
            xt[ib] += ddx[i] * lj; yt[ib] += ddy[i] * lj; zt[ib] += ddz[i] * lj;
        }
        if (maxdist < tolvel) conv = 1;
        if (!conv) {
            for (int i = 0; i < nc; i++) {
                int ia = ca[i], ib = cb[i];
                double nia = 1.0 / (double)nAtConst[ia], nib = 1.0 / (double)nAtConst[ib];
                vx[ia] += xt[ia] * nia; vy[ia] += yt[ia] * nia; vz[ia] += zt[ia] * nia;
                vx[ib] += xt[ib] * nib; vy[ib] += yt[ib] * nib; vz[ib] += zt[ib] * nib;
            }
        }
        icycle++;
    } while (!conv && icycle < p->maxCycle);
    free(xt);
    if (ncycle) *ncycle = icycle;
    return conv ? 0 : 311;
}

int oc_vv_nve(const oc_param *p, const oc_box *b, int stage,
              int nConst, const int *ca, const int *cb, const double *rc2,
              double *x, double *y, double *z,
              double *vx, double *vy, double *vz,
              const double *fx, const double *fy, const double *fz,
              const double *mass, const double *rmass, const int *nAtConst,
              double *ekin, double *virshake, int *ncycle)
{
    int n = p->nAtom, err = 0;
    double dt = p->timeStep;
    double *dd = NULL;
    if (nConst > 0) {                                     /* integrate.c:1526-1545 */
        dd = (double *)malloc((size_t)nConst * 3 * sizeof(double));
        for (int i = 0; i < nConst; i++) {
            dd[i] = x[cb[i]] - x[ca[i]];
            dd[nConst + i] = y[cb[i]] - y[ca[i]];
            dd[2 * nConst + i] = z[cb[i]] - z[ca[i]];
            oc_dist2(b, &dd[i], &dd[nConst + i], &dd[2 * nConst + i]);
        }
    }
    for (int i = 0; i < n; i++) {                         /* integrate.c:1554-1561 */
        vx[i] += 0.5 * dt * fx[i] * rmass[i];
        vy[i] += 0.5 * dt * fy[i] * rmass[i];
        vz[i] += 0.5 * dt * fz[i] * rmass[i];
    }
    if (stage == 1) {
        for (int i = 0; i < n; i++) {                     /* integrate.c:1568-1576 */
            x[i] += dt * vx[i];
            y[i] += dt * vy[i];
            z[i] += dt * vz[i];
        }
        if (nConst > 0) {
            double vs = 0.;
            err = oc_shake_r(p, b, nConst, ca, cb, rc2, x, y, z, vx, vy, vz, dd,
                             dd + nConst, dd + 2 * nConst, rmass, nAtConst, &vs, ncycle);
            if (virshake) *virshake = vs;
        }
    } else {
        if (nConst > 0)
            err = oc_shake_v(p, nConst, ca, cb, vx, vy, vz, dd, dd + nConst, dd + 2 * nConst,
                             rmass, nAtConst, ncycle);
        double ek = 0.;                                   /* utils.c:502-515 */
        for (int i = 0; i < n; i++) ek += mass[i] * (SQ(vx[i]) + SQ(vy[i]) + SQ(vz[i]));
        if (ekin) *ekin = ek * 0.5;
        oc_image_update(b, n, x, y, z);                   /* integrate.c:1620-1627 */
    }
    free(dd);
    return err;
}

/* ---------------------------------------------------------------- */
/* Thermostats and barostats (integrate.c:1636-2618).               */
/* ---------------------------------------------------------------- */

#define OC_WATERCOMP 0.007372    /* global.h:56 */
#define OC_RBOLTZUI 0.83144621   /* global.h:47 */
#define OC_HOOVER 5              /* integrate.c:2176 */

/* kinetic (utils.c:502-515) */
static double ens_kin(int n, const double *vx, const double *vy, const double *vz,
                      const double *m)
{
    double ek = 0.;
    for (int i = 0; i < n; i++) ek += m[i] * (SQ(vx[i]) + SQ(vy[i]) + SQ(vz[i]));
    return ek * 0.5;
}

/* stress_kinetic (utils.c:517-542) added to the nine box stresses the way
 * every stage 2 does (integrate.c:1603-1611); the SHAKE stress of stage 2
 * is zero. */
static void ens_stress_kin(int n, const double *vx, const double *vy, const double *vz,
                           const double *m, double st[9])
{
    double s[6] = {0., 0., 0., 0., 0., 0.};
    for (int i = 0; i < n; i++) {
        s[0] += m[i] * SQ(vx[i]);
        s[1] += m[i] * vx[i] * vy[i];
        s[2] += m[i] * vx[i] * vz[i];
        s[3] += m[i] * SQ(vy[i]);
        s[4] += m[i] * vy[i] * vz[i];
        s[5] += m[i] * SQ(vz[i]);
    }
    const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int k = 0; k < 9; k++) st[k] += s[map[k]];
}

static void ens_scale(int n, double *vx, double *vy, double *vz, double f)
{
    for (int i = 0; i < n; i++) {
        vx[i] *= f;
        vy[i] *= f;
        vz[i] *= f;
    }
}

static void ens_kick(int n, double dt, double *vx, double *vy, double *vz, const double *fx,
                     const double *fy, const double *fz, const double *rmass)
{
    for (int i = 0; i < n; i++) {
        vx[i] += 0.5 * dt * fx[i] * rmass[i];
        vy[i] += 0.5 * dt * fy[i] * rmass[i];
        vz[i] += 0.5 * dt * fz[i] * rmass[i];
    }
}

/* scale_box (utils.c:291-374) from saved vectors, vv_scale_box
 * (utils.c:376-459) in place: both rescale the nine vectors and derive the
 * rest exactly as init_box does. */
static void ens_box(oc_box *b, const double cell0[9], double s)
{
    b->a1 = s * cell0[0]; b->a2 = s * cell0[1]; b->a3 = s * cell0[2];
    b->b1 = s * cell0[3]; b->b2 = s * cell0[4]; b->b3 = s * cell0[5];
    b->c1 = s * cell0[6]; b->c2 = s * cell0[7]; b->c3 = s * cell0[8];
    oc_init_box(b);
}

static void ens_cell(const oc_box *b, double c[9])
{
    c[0] = b->a1; c[1] = b->a2; c[2] = b->a3;
    c[3] = b->b1; c[4] = b->b2; c[5] = b->b3;
    c[6] = b->c1; c[7] = b->c2; c[8] = b->c3;
}

/* One Nose-Hoover thermostat-barostat sub-step of vv_npt_h
 * (integrate.c:2296-2382 = 2471-2545): half thermostat, barostat, half
 * thermostat.  vs is the integrator's own virshake local. */
static void ens_hoover(int n, oc_bath *s, double *vx, double *vy, double *vz, const double *m,
                       double qmass, double pmass, double cqts, double chts, double vs,
                       double *volume)
{
    for (int half = 0; half < 3; half++) {
        if (half == 1) {
            s->chiP += 0.5 * chts * (((2.0 * s->kin - s->virpot - vs) -
                                      3.0 * s->press0 * *volume) / pmass - s->chiP * s->chiT);
            const double g = exp(-chts * s->chiP);
            ens_scale(n, vx, vy, vz, g);
            s->kin *= SQ(g);
            *volume *= exp(3.0 * chts * s->chiP);
            s->chiP += 0.5 * chts * (((2.0 * s->kin - s->virpot - vs) -
                                      3.0 * s->press0 * *volume) / pmass - s->chiP * s->chiT);
            continue;
        }
        s->kin = ens_kin(n, vx, vy, vz, m);
        s->chiT += 0.5 * cqts * (2.0 * (s->kin - s->kinTemp0) + pmass * SQ(s->chiP) -
                                 OC_RBOLTZUI * s->temp0) / qmass;
        const double l = exp(-cqts * s->chiT);
        ens_scale(n, vx, vy, vz, l);
        s->kin *= SQ(l);
        s->conint += cqts * s->chiT * (OC_RBOLTZUI * s->temp0 + qmass / SQ(s->tauT));
        s->chiT += 0.5 * cqts * (2.0 * (s->kin - s->kinTemp0) + pmass * SQ(s->chiP) -
                                 OC_RBOLTZUI * s->temp0) / qmass;
    }
}

int oc_vv_ens(const oc_param *p, oc_box *b, oc_bath *s, int stage,
              int nConst, const int *ca, const int *cb, const double *rc2,
              double *x, double *y, double *z,
              double *vx, double *vy, double *vz,
              const double *fx, const double *fy, const double *fz,
              const double *mass, const double *rmass, const int *nAtConst, int *ncycle)
{
    const int n = p->nAtom;
    const double dt = p->timeStep;
    int err = 0;
    /* virshake is a local of each reference integrator; where it is read
     * before a SHAKE call has set it (vv_npt_b without constraints,
     * vv_npt_h's first cycle and its stage 2) this reads zero, as the
     * reference built with zero-initialised locals does */
    double vs = 0.;
    if (s->ens < 1 || s->ens > 4) return 1;
    double *dd = NULL, *ddy = NULL, *ddz = NULL;
    if (nConst > 0) {   /* constraint vectors at entry, entry box */
        dd = (double *)malloc((size_t)nConst * 3 * sizeof(double));
        ddy = dd + nConst;
        ddz = dd + 2 * nConst;
        for (int i = 0; i < nConst; i++) {
            dd[i] = x[cb[i]] - x[ca[i]];
            ddy[i] = y[cb[i]] - y[ca[i]];
            ddz[i] = z[cb[i]] - z[ca[i]];
            oc_dist2(b, &dd[i], &ddy[i], &ddz[i]);
        }
    }
#define SHAKE_R()                                                                           \
    do {                                                                                    \
        if (nConst > 0) {                                                                   \
            int e_ = oc_shake_r(p, b, nConst, ca, cb, rc2, x, y, z, vx, vy, vz, dd, ddy, ddz, \
                                rmass, nAtConst, &vs, ncycle);                              \
            err = err ? err : e_;                                                           \
            s->virshake = vs;                                                               \
        }                                                                                   \
    } while (0)
#define SHAKE_V()                                                                           \
    do {                                                                                    \
        if (nConst > 0) {                                                                   \
            int e_ = oc_shake_v(p, nConst, ca, cb, vx, vy, vz, dd, ddy, ddz, rmass,          \
                                nAtConst, ncycle);                                          \
            err = err ? err : e_;                                                           \
        }                                                                                   \
    } while (0)

    if (s->ens == 1) {                                    /* vv_nvt_b, integrate.c:1636-1768 */
        ens_kick(n, dt, vx, vy, vz, fx, fy, fz, rmass);
        if (stage == 1) {
            for (int i = 0; i < n; i++) {
                x[i] += dt * vx[i];
                y[i] += dt * vy[i];
                z[i] += dt * vz[i];
            }
            SHAKE_R();
        } else {
            SHAKE_V();
            s->kin = ens_kin(n, vx, vy, vz, mass);
            const double l = sqrt(1.0 + dt / s->tauT * (s->kinTemp0 / s->kin - 1.0));
            ens_scale(n, vx, vy, vz, l);
            s->kin *= SQ(l);
            ens_stress_kin(n, vx, vy, vz, mass, s->stress);
        }
    } else if (s->ens == 3) {                             /* vv_nvt_h, integrate.c:1998-2169 */
        const double qmass = 2.0 * s->kinTemp0 * SQ(s->tauT);
        if (stage == 1) {
            s->kin = ens_kin(n, vx, vy, vz, mass);
            s->chiT += 0.5 * dt * (s->kin - s->kinTemp0) / qmass;
            const double l = exp(-0.5 * dt * s->chiT);
            ens_scale(n, vx, vy, vz, l);
            ens_kick(n, dt, vx, vy, vz, fx, fy, fz, rmass);
            s->kin *= SQ(l);
            s->conint += 0.5 * dt * s->chiT * qmass / SQ(s->tauT);
            s->chiT += 0.5 * dt * (s->kin - s->kinTemp0) / qmass;
            for (int i = 0; i < n; i++) {
                x[i] += dt * vx[i];
                y[i] += dt * vy[i];
                z[i] += dt * vz[i];
            }
            SHAKE_R();
        } else {
            ens_kick(n, dt, vx, vy, vz, fx, fy, fz, rmass);
            SHAKE_V();
            s->kin = ens_kin(n, vx, vy, vz, mass);
            s->chiT += 0.5 * dt * (s->kin - s->kinTemp0) / qmass;
            const double l = exp(-0.5 * dt * s->chiT);
            ens_scale(n, vx, vy, vz, l);
            s->kin *= SQ(l);
            s->conint += 0.5 * dt * s->chiT * qmass / SQ(s->tauT);
            s->chiT += 0.5 * dt * (s->kin - s->kinTemp0) / qmass;
            s->consv = s->conint + 0.5 * qmass * SQ(s->chiT);
            ens_stress_kin(n, vx, vy, vz, mass, s->stress);
        }
    } else if (s->ens == 2) {                             /* vv_npt_b, integrate.c:1770-1996 */
        const double volume = b->vol;
        if (stage == 1) {
            s->kin = ens_kin(n, vx, vy, vz, mass);
            ens_kick(n, dt, vx, vy, vz, fx, fy, fz, rmass);
            const int nose = nConst > 0 ? 2 : 1;
            double *old = NULL;
            if (nose > 1) {   /* positions and kicked velocities before the trial cycle */
                old = (double *)malloc((size_t)n * 6 * sizeof(double));
                memcpy(old, x, n * sizeof(double));
                memcpy(old + n, y, n * sizeof(double));
                memcpy(old + 2 * n, z, n * sizeof(double));
                memcpy(old + 3 * n, vx, n * sizeof(double));
                memcpy(old + 4 * n, vy, n * sizeof(double));
                memcpy(old + 5 * n, vz, n * sizeof(double));
            }
            for (int k = 0; k < nose; k++) {
                double c = 1.;
                if (k == nose - 1) {
                    const double pp = (2. * s->kin - s->virpot - vs) / (3. * volume);
                    const double g = 1. + OC_WATERCOMP * dt * (pp - s->press0) / s->tauP;
                    c = cbrt(g);
                    double cell[9];
                    ens_cell(b, cell);
                    ens_box(b, cell, c);
                }
                for (int i = 0; i < n; i++) {
                    x[i] = c * x[i] + dt * vx[i];
                    y[i] = c * y[i] + dt * vy[i];
                    z[i] = c * z[i] + dt * vz[i];
                }
                SHAKE_R();
                if (k < nose - 1) {
                    memcpy(x, old, n * sizeof(double));
                    memcpy(y, old + n, n * sizeof(double));
                    memcpy(z, old + 2 * n, n * sizeof(double));
                    memcpy(vx, old + 3 * n, n * sizeof(double));
                    memcpy(vy, old + 4 * n, n * sizeof(double));
                    memcpy(vz, old + 5 * n, n * sizeof(double));
                }
            }
            free(old);
        } else {
            ens_kick(n, dt, vx, vy, vz, fx, fy, fz, rmass);
            s->kin = ens_kin(n, vx, vy, vz, mass);
            const double l = sqrt(1.0 + dt / s->tauT * (s->kinTemp0 / s->kin - 1.0));
            ens_scale(n, vx, vy, vz, l);
            SHAKE_V();
            s->kin = ens_kin(n, vx, vy, vz, mass);
            ens_stress_kin(n, vx, vy, vz, mass, s->stress);
        }
    } else {                                              /* vv_npt_h, integrate.c:2171-2618 */
        const double hts = 0.5 * dt;
        const double chts = hts / (double)OC_HOOVER;
        const double cqts = 0.25 * dt / (double)OC_HOOVER;
        double cell0[9];
        ens_cell(b, cell0);
        double volume = b->vol;
        const double volume0 = b->vol;
        double masst = 0.;
        for (int i = 0; i < n; i++) masst += mass[i];
        const double qmass = 2.0 * s->kinTemp0 * SQ(s->tauT);
        const double pmass = 2.0 * s->kinTemp0 * SQ(s->tauP);
        if (stage == 1) {
            const double chiT0 = s->chiT, chiP0 = s->chiP, cons0 = s->conint;
            const int nose = nConst > 0 ? 2 : 1;
            for (int k = 0; k < nose; k++) {
                for (int kk = 0; kk < OC_HOOVER; kk++)
                    ens_hoover(n, s, vx, vy, vz, mass, qmass, pmass, cqts, chts, vs, &volume);
                ens_box(b, cell0, cbrt(volume / volume0));
                ens_kick(n, dt, vx, vy, vz, fx, fy, fz, rmass);
                double com[3] = {0., 0., 0.};
                for (int i = 0; i < n; i++) {
                    com[0] += mass[i] * x[i];
                    com[1] += mass[i] * y[i];
                    com[2] += mass[i] * z[i];
                }
                com[0] /= masst;
                com[1] /= masst;
                com[2] /= masst;
                const double c = exp(dt * s->chiP);
                for (int i = 0; i < n; i++) {
                    x[i] = c * (x[i] - com[0]) + dt * vx[i] + com[0];
                    y[i] = c * (y[i] - com[1]) + dt * vy[i] + com[1];
                    z[i] = c * (z[i] - com[2]) + dt * vz[i] + com[2];
                }
                SHAKE_R();
                /* the second cycle restarts the bath from the step's
                 * start but carries on from this cycle's positions and
                 * velocities (integrate.c:2415-2437 saves them, it does
                 * not restore them) */
                if (k < nose - 1) {
                    volume = volume0;
                    s->chiT = chiT0;
                    s->chiP = chiP0;
                    s->conint = cons0;
                }
            }
        } else {
            ens_kick(n, dt, vx, vy, vz, fx, fy, fz, rmass);
            SHAKE_V();
            for (int kk = 0; kk < OC_HOOVER; kk++)
                ens_hoover(n, s, vx, vy, vz, mass, qmass, pmass, cqts, chts, vs, &volume);
            double vom[3] = {0., 0., 0.};
            for (int i = 0; i < n; i++) {
                vom[0] += mass[i] * vx[i];
                vom[1] += mass[i] * vy[i];
                vom[2] += mass[i] * vz[i];
            }
            vom[0] /= masst;
            vom[1] /= masst;
            vom[2] /= masst;
            for (int i = 0; i < n; i++) {
                vx[i] -= vom[0];
                vy[i] -= vom[1];
                vz[i] -= vom[2];
            }
            ens_box(b, cell0, cbrt(volume / volume0));
            s->consv = s->conint + s->press0 * volume +
                       0.5 * (qmass * SQ(s->chiT) + pmass * SQ(s->chiP));
            s->kin = ens_kin(n, vx, vy, vz, mass);
            ens_stress_kin(n, vx, vy, vz, mass, s->stress);
        }
    }
#undef SHAKE_R
#undef SHAKE_V
    if (stage == 2) oc_image_update(b, n, x, y, z);       /* integrate.c:1620-1627 */
    free(dd);
    return err;
}

/* ---------------------------------------------------------------- */
/* Whole MD run (mdbas.c:160-300).                                   */
/* ---------------------------------------------------------------- */

typedef struct {
    int *nl; long cap; int *np; int *no;
} oc_list;

static int md_list(const oc_param *p, const oc_box *b, int keyLink, int build,
                   double *x, double *y, double *z, const int *frozen,
                   const int *exOff, const int *exLst, oc_list *L)
{
    for (;;) {
        long cnt = keyLink
            ? oc_link_cell_list(p, b, build, x, y, z, frozen, exOff, exLst, L->nl, L->cap, L->np, L->no)
            : oc_verlet_list(p, b, x, y, z, frozen, exOff, exLst, L->nl, L->cap, L->np, L->no);
        if (cnt < 0) return (int)(-cnt);
        if (cnt <= L->cap) return 0;
        /* list.c:2110 sizing rule; on overflow the reference reallocs */
        L->cap = (long)((double)cnt * (1. + 2. * OC_TOLLIST)) + 1;
        L->nl = (int *)realloc(L->nl, (size_t)L->cap * sizeof(int));
        build = 0;   /* positions were already wrapped by the first attempt */
    }
}

int oc_md_run(const oc_param *p, const oc_box *b, int keyLink, int nsteps,
              double *x, double *y, double *z,
              double *vx, double *vy, double *vz,
              double *fx, double *fy, double *fz,
              const double *q, const double *eps, const double *sig,
              const double *mass, const double *rmass, const int *frozen,
              const int *nAtConst, int nConst, const int *ca, const int *cb,
              const double *rc2, double *etot, double *epot, double *ekin,
              double *seconds)
{
    int n = p->nAtom, err = 0;
    int *exOff = (int *)malloc((size_t)(n + 1) * sizeof(int));
    int excap = 4 * nConst + 16;
    int *exLst = (int *)malloc((size_t)excap * sizeof(int));
    int ne = oc_exclusions(n, 0, NULL, NULL, nConst, ca, cb, 0, NULL, NULL, NULL,
                           keyLink ? 1 : 0, exOff, exLst, excap);
    if (ne < 0) { free(exOff); free(exLst); return ne > -1000 ? -ne : 1; }
    oc_list L;
    L.cap = 1024;
    L.nl = (int *)malloc((size_t)L.cap * sizeof(int));
    L.np = (int *)malloc((size_t)n * sizeof(int));
    L.no = (int *)malloc((size_t)n * sizeof(int));

    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    err = md_list(p, b, keyLink, 1, x, y, z, frozen, exOff, exLst, &L);
    double ek = 0.;
    for (int i = 0; i < n; i++) ek += mass[i] * (SQ(vx[i]) + SQ(vy[i]) + SQ(vz[i]));
    ek *= 0.5;
    for (int step = 0; step <= nsteps && !err; step++) {
        if (step > 0) {
            double vs;
            err = oc_vv_nve(p, b, 1, nConst, ca, cb, rc2, x, y, z, vx, vy, vz, fx, fy, fz,
                            mass, rmass, nAtConst, NULL, &vs, NULL);
            if (err) break;
            if (step % p->update == 0)
                err = md_list(p, b, keyLink, 0, x, y, z, frozen, exOff, exLst, &L);
            if (err) break;
        }
        for (int i = 0; i < n; i++) { fx[i] = 0.; fy[i] = 0.; fz[i] = 0.; }
        double st[6], vr;
        double eRec = oc_spme(p, b, x, y, z, fx, fy, fz, q, st, &vr);
        oc_energy e;
        oc_ewald_real(p, b, x, y, z, fx, fy, fz, q, eps, sig, L.nl, L.np, L.no, exOff, exLst, &e);
        double pot = (e.eDir + eRec + e.eCorr) + e.eVdw;
        if (step > 0) {
            err = oc_vv_nve(p, b, 2, nConst, ca, cb, rc2, x, y, z, vx, vy, vz, fx, fy, fz,
                            mass, rmass, nAtConst, &ek, NULL, NULL);
        }
        if (epot) epot[step] = pot;
        if (ekin) ekin[step] = ek;
        if (etot) etot[step] = pot + ek;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    free(L.nl); free(L.np); free(L.no); free(exOff); free(exLst);
    return err;
}

/* Gradient of the Ewald-path potential energy (energy.c:450-627 with
 * spme.c:357-699) with respect to the per-atom parameters q, eps, sig,
 * over the same list and exclusions:
 *   real space   E = C q_i q_j erfc(a r)/r
 *                  + 4 eps_i eps_j ((s^12 - s^6) S(r)),  s = (sig_i+sig_j)/r
 *   correction   E = -C q_i q_j erf(a r)/r                    (j > i)
 *   reciprocal   oc_spme_pot
 *   self         E = -C a sum q^2 / sqrt(pi)
 *   net charge   E = -(pi C / 2) (sum q / a)^2 / V.
 * dq, deps, dsig are overwritten. */
void oc_energy_grad(const oc_param *p, const oc_box *b,
                    const double *x, const double *y, const double *z,
                    const double *q, const double *eps, const double *sig,
                    const int *neighList, const int *neighPair, const int *neighOrder,
                    const int *exOff, const int *exLst,
                    double *dq, double *deps, double *dsig)
{
    int n = p->nAtom;
    double C = p->chargeConst, a = p->alpha;
    oc_spme_pot(p, b, x, y, z, q, dq);
    double systq = 0.;
    for (int i = 0; i < n; i++) systq += q[i];
    for (int i = 0; i < n; i++) {
        dq[i] += -2.0 * C * a * q[i] / OC_SQRTPI;
        dq[i] += -(OC_PI * C) * systq / (a * a * b->vol);
        deps[i] = 0.;
        dsig[i] = 0.;
    }
    long offset = 0;
    for (int l = 0; l < n; l++) {
        int i = neighOrder[l];
        for (int k = 0; k < neighPair[i]; k++) {
            int j = neighList[offset++];
            double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
            double r2 = oc_dist(b, d);
            if (r2 > p->cutOff2) continue;
            double r = sqrt(r2), rt = 1. / r;
            double ec = C * erfc(a * r) * rt;
            dq[i] += q[j] * ec;
            dq[j] += q[i] * ec;
            double vsig = sig[i] + sig[j];
            double s = vsig * rt, s2 = s * s, s6 = s2 * s2 * s2, s12 = s6 * s6;
            double s5 = s2 * s2 * s, s11 = s5 * s6;
            double S = 1.;
            if (r2 > p->cutOn2) {
                double sw = p->cutOff2 - r2;
                S = sw * sw * (p->cutOff2 + 2. * r2 - 3. * p->cutOn2) * p->switch2;
            }
            double lj = 4. * (s12 - s6) * S;
            deps[i] += eps[j] * lj;
            deps[j] += eps[i] * lj;
            double gs = 4. * eps[i] * eps[j] * (12. * s11 - 6. * s5) * rt * S;
            dsig[i] += gs;
            dsig[j] += gs;
        }
    }
    for (int i = 0; i < n; i++)
        for (int k = exOff[i]; k < exOff[i + 1]; k++) {
            int j = exLst[k];
            if (j <= i) continue;
            double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
            double r = sqrt(oc_dist(b, d));
            double ec = -C * erf(a * r) / r;
            dq[i] += q[j] * ec;
            dq[j] += q[i] * ec;
        }
}

/* ---------------------------------------------------------------- */
/* Bonded terms: internal.c:21-604.                                  */
/* ---------------------------------------------------------------- */

/* One bond or Urey-Bradley term (internal.c:33-80, 120-150): harmonic
 * 0.5 k (r - r0)^2, or Morse k (e^2 - 2e) + k with e = exp(-beta (r - r0))
 * for bonds of type 1 (BMORSE).  Forces f_a += dE/dr * d/r, f_b -= the
 * same, with d = x_b - x_a under the minimum image. */
static void oc_bond_term(const oc_box *b, const oc_bond *t, int morse_ok,
                         const double *x, const double *y, const double *z,
                         double *fx, double *fy, double *fz, double *e, double *vir)
{
    int i = t->a, j = t->b;
    double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
    double r = sqrt(oc_dist(b, d));
    double db;
    if (morse_ok && t->type == 1) {
        double ma = exp(-t->beta * (r - t->r0));
        double mb = SQ(ma);
        *e += t->k * (mb - 2. * ma) + t->k;
        db = 2. * t->k * t->beta * (ma - mb);
    } else {
        *e += 0.5 * t->k * SQ(r - t->r0);
        db = t->k * (r - t->r0);
    }
    *vir += db * r;
    double tfx = db * d[0] / r, tfy = db * d[1] / r, tfz = db * d[2] / r;
    fx[i] += tfx; fy[i] += tfy; fz[i] += tfz;
    fx[j] += -tfx; fy[j] += -tfy; fz[j] += -tfz;
}

/* angle_energy, internal.c:164-245: harmonic in theta = acos(cos t). */
static void oc_angle_term(const oc_box *b, const oc_angle *t,
                          const double *x, const double *y, const double *z,
                          double *fx, double *fy, double *fz, double *e)
{
    int i = t->a, j = t->b, k = t->c;
    double dab[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
    double rab = sqrt(oc_dist(b, dab)), rabt = 1. / rab;
    double dbc[3] = {x[k] - x[j], y[k] - y[j], z[k] - z[j]};
    double rbc = sqrt(oc_dist(b, dbc)), rbct = 1. / rbc;
    double cost = (dab[0] * dbc[0] + dab[1] * dbc[1] + dab[2] * dbc[2]) / (rab * rbc);
    double sint = sqrt(1. - cost * cost);
    if (sint < 1.e-8) sint = 1.e-8;
    double theta = acos(cost);
    *e += 0.5 * t->k * SQ(theta - t->theta0);
    double da = t->k * (theta - t->theta0) / sint;
    double fa[3], fc[3];
    for (int c = 0; c < 3; c++) {
        fa[c] = da * (dbc[c] * rbct - dab[c] * cost * rabt) * rabt;
        fc[c] = da * (dab[c] * rabt - dbc[c] * cost * rbct) * rbct;
    }
    fx[i] += fa[0]; fy[i] += fa[1]; fz[i] += fa[2];
    fx[j] += -fa[0] - fc[0]; fy[j] += -fa[1] - fc[1]; fz[j] += -fa[2] - fc[2];
    fx[k] += fc[0]; fy[k] += fc[1]; fz[k] += fc[2];
}

/* dihedral_energy / improper_energy, internal.c:247-604 (the same body):
 * phi from the normals pb = dab x dbc, pc = dbc x dcd with the sign of
 * dbc . (pc x pb); cosine k (1 + cos(m phi - phi0)) (type 1 and default)
 * or harmonic 0.5 k dphi^2 with dphi wrapped by nint (type 2). */
static void oc_dihe_term(const oc_box *b, const oc_dihe *t,
                         const double *x, const double *y, const double *z,
                         double *fx, double *fy, double *fz, double *e)
{
    int i = t->a, j = t->b, k = t->c, l = t->d;
    double dab[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
    oc_dist(b, dab);
    double dbc[3] = {x[j] - x[k], y[j] - y[k], z[j] - z[k]};
    double rbc = sqrt(oc_dist(b, dbc));
    double dcd[3] = {x[k] - x[l], y[k] - y[l], z[k] - z[l]};
    oc_dist(b, dcd);
    double pb[3] = {dab[1] * dbc[2] - dab[2] * dbc[1], dab[2] * dbc[0] - dab[0] * dbc[2],
                    dab[0] * dbc[1] - dab[1] * dbc[0]};
    double r2pb = SQ(pb[0]) + SQ(pb[1]) + SQ(pb[2]), rpb = sqrt(r2pb);
    double pc[3] = {dbc[1] * dcd[2] - dbc[2] * dcd[1], dbc[2] * dcd[0] - dbc[0] * dcd[2],
                    dbc[0] * dcd[1] - dbc[1] * dcd[0]};
    double r2pc = SQ(pc[0]) + SQ(pc[1]) + SQ(pc[2]), rpc = sqrt(r2pc);
    double pbpc = pb[0] * pc[0] + pb[1] * pc[1] + pb[2] * pc[2];
    double cosp = pbpc / (rpb * rpc);
    double sinp = (dbc[0] * (pc[1] * pb[2] - pc[2] * pb[1]) + dbc[1] * (pb[0] * pc[2] - pb[2] * pc[0]) +
                   dbc[2] * (pc[0] * pb[1] - pc[1] * pb[0])) / (rpb * rpc * rbc);
    double phi = atan2(sinp, cosp);
    if (sinp >= 0.) sinp = fabs(sinp) > DBL_EPSILON ? fabs(sinp) : DBL_EPSILON;
    else sinp = -(fabs(sinp) > DBL_EPSILON ? fabs(sinp) : DBL_EPSILON);
    double ed, dd;
    if (t->type == 2) {
        phi = phi - t->phi0;
        phi = phi - oc_nint(phi / OC_TWOPI) * OC_TWOPI;
        ed = 0.5 * t->k * (phi * phi);
        dd = t->k * phi / (rpb * rpc * sinp);
    } else {
        ed = t->k * (1. + cos(t->mult * phi - t->phi0));
        dd = -t->k * t->mult * sin(t->mult * phi - t->phi0) / (rpb * rpc * sinp);
    }
    *e += ed;
    double fa[3], fb[3], fc[3], fd[3];
    fa[0] = dd * ((-pc[1] * dbc[2] + pc[2] * dbc[1]) - pbpc * (-pb[1] * dbc[2] + pb[2] * dbc[1]) / r2pb);
    fa[1] = dd * ((pc[0] * dbc[2] - pc[2] * dbc[0]) - pbpc * (pb[0] * dbc[2] - pb[2] * dbc[0]) / r2pb);
    fa[2] = dd * ((-pc[0] * dbc[1] + pc[1] * dbc[0]) - pbpc * (-pb[0] * dbc[1] + pb[1] * dbc[0]) / r2pb);
    fc[0] = dd * ((-pc[1] * dab[2] + pc[2] * dab[1]) - pbpc * (-pb[1] * dab[2] + pb[2] * dab[1]) / r2pb);
    fc[1] = dd * ((pc[0] * dab[2] - pc[2] * dab[0]) - pbpc * (pb[0] * dab[2] - pb[2] * dab[0]) / r2pb);
    fc[2] = dd * ((-pc[0] * dab[1] + pc[1] * dab[0]) - pbpc * (-pb[0] * dab[1] + pb[1] * dab[0]) / r2pb);
    fb[0] = dd * ((-pb[1] * dcd[2] + pb[2] * dcd[1]) - pbpc * (-pc[1] * dcd[2] + pc[2] * dcd[1]) / r2pc);
    fb[1] = dd * ((pb[0] * dcd[2] - pb[2] * dcd[0]) - pbpc * (pc[0] * dcd[2] - pc[2] * dcd[0]) / r2pc);
    fb[2] = dd * ((-pb[0] * dcd[1] + pb[1] * dcd[0]) - pbpc * (-pc[0] * dcd[1] + pc[1] * dcd[0]) / r2pc);
    fd[0] = dd * ((-pb[1] * dbc[2] + pb[2] * dbc[1]) - pbpc * (-pc[1] * dbc[2] + pc[2] * dbc[1]) / r2pc);
    fd[1] = dd * ((pb[0] * dbc[2] - pb[2] * dbc[0]) - pbpc * (pc[0] * dbc[2] - pc[2] * dbc[0]) / r2pc);
    fd[2] = dd * ((-pb[0] * dbc[1] + pb[1] * dbc[0]) - pbpc * (-pc[0] * dbc[1] + pc[1] * dbc[0]) / r2pc);
    fx[i] += fa[0]; fy[i] += fa[1]; fz[i] += fa[2];
    fx[j] += -fa[0] - fc[0] + fb[0]; fy[j] += -fa[1] - fc[1] + fb[1]; fz[j] += -fa[2] - fc[2] + fb[2];
    fx[k] += fc[0] - fb[0] - fd[0]; fy[k] += fc[1] - fb[1] - fd[1]; fz[k] += fc[2] - fb[2] - fd[2];
    fx[l] += fd[0]; fy[l] += fd[1]; fz[l] += fd[2];
}

void oc_bonded(const oc_box *b, int nBond, const oc_bond *bond, int nUb, const oc_bond *ub,
               int nAngle, const oc_angle *angle, int nDihe, const oc_dihe *dihe,
               int nImpr, const oc_dihe *impr,
               const double *x, const double *y, const double *z,
               double *fx, double *fy, double *fz, double out[7])
{
    double eb = 0., eu = 0., ea = 0., ed = 0., ei = 0., vb = 0., vu = 0.;
    for (int l = 0; l < nBond; l++) oc_bond_term(b, &bond[l], 1, x, y, z, fx, fy, fz, &eb, &vb);
    for (int l = 0; l < nAngle; l++) oc_angle_term(b, &angle[l], x, y, z, fx, fy, fz, &ea);
    for (int l = 0; l < nUb; l++) oc_bond_term(b, &ub[l], 0, x, y, z, fx, fy, fz, &eu, &vu);
    for (int l = 0; l < nDihe; l++) oc_dihe_term(b, &dihe[l], x, y, z, fx, fy, fz, &ed);
    for (int l = 0; l < nImpr; l++) oc_dihe_term(b, &impr[l], x, y, z, fx, fy, fz, &ei);
    out[0] = eb; out[1] = eu; out[2] = ea; out[3] = ed; out[4] = ei; out[5] = vb; out[6] = vu;
}

/* ---------------------------------------------------------------- */
/* 1-4 pairs: nonbond14_energy energy.c:356-439, ewald14_energy       */
/* energy.c:654-742, with coulomb14_* elec.c:272-465, vdw14_* vdw.c:  */
/* 202-311, ewald_dir14 / ewald_corr14 ewald.c:386-413.               */
/* ---------------------------------------------------------------- */

/* coulomb14_shift1/shift2/switch (2, 3, 4): scal14 x the plain functor,
 * zero beyond the cutoff; NOELEC and FULL map to coulomb14_none
 * (energy.c:37-54). */
static double oc_coulomb14(const oc_param *p, int type, double scal14, double qel, double r2,
                           double rt, double *de)
{
    double e = 0.;
    *de = 0.;
    if (type == 2 || type == 3) {
        if (r2 <= p->cutOff2) {
            double s1, s2, sf, dsf;
            if (type == 2) {
                s1 = rt * r2 * p->rcutOff;
                s2 = r2 * p->rcutOff2;
                sf = 1. - 2. * s1 + s2;
                dsf = 2. * (s2 - s1);
            } else {
                s1 = r2 * p->rcutOff2;
                s2 = SQ(s1);
                sf = 1. - 2. * s1 + s2;
                dsf = 4. * (s2 - s1);
            }
            double pe = scal14 * p->chargeConst * qel * rt;
            e = pe * sf;
            *de = pe * rt * (dsf - sf);
        }
    } else if (type == 4) {
        if (r2 <= p->cutOn2) {
            double pe = scal14 * p->chargeConst * qel * rt;
            e = pe;
            *de = -pe * rt;
        } else if (r2 <= p->cutOff2) {
            double sw = p->cutOff2 - r2;
            double sf = SQ(sw) * (p->cutOff2 + 2. * r2 - 3. * p->cutOn2) * p->switch2;
            double dsf = 12. * r2 * sw * (p->cutOn2 - r2) * p->switch2;
            double pe = scal14 * p->chargeConst * qel * rt;
            e = pe * sf;
            *de = pe * rt * (dsf - sf);
        }
    }
    return e;
}

/* vdw14_switch (vdw.c:270-311); vdw14_none when vdwOn is 0. */
static double oc_vdw14(const oc_param *p, int vdwOn, double scal14, double veps, double vsig,
                       double r2, double rt, double *dv)
{
    double e = 0.;
    *dv = 0.;
    if (!vdwOn) return 0.;
    double s6 = vsig * rt;
    s6 = s6 * s6 * s6 * s6 * s6 * s6;
    double s12 = SQ(s6);
    if (r2 <= p->cutOn2) {
        e = 4. * scal14 * veps * (s12 - s6);
        *dv = 24. * scal14 * veps * rt * (s6 - 2. * s12);
    } else if (r2 <= p->cutOff2) {
        double sw = p->cutOff2 - r2;
        double sf = SQ(sw) * (p->cutOff2 + 2. * r2 - 3. * p->cutOn2) * p->switch2;
        double dsf = 12. * r2 * sw * (p->cutOn2 - r2) * p->switch2;
        double pv = 4. * scal14 * veps * (s12 - s6);
        double dpv = 24. * scal14 * veps * (s6 - 2. * s12);
        e = pv * sf;
        *dv = (pv * dsf + dpv * sf) * rt;
    }
    return e;
}

void oc_nb14(const oc_param *p, const oc_box *b, int ewald, int elecType, int vdwOn,
             double scal14, int nPair14, const int *list14,
             const double *x, const double *y, const double *z,
             double *fx, double *fy, double *fz,
             const double *q, const double *eps14, const double *sig14, double out[4])
{
    double elec = 0., evdw = 0., virel = 0., virvdw = 0.;
    double eD = 0., eC = 0., vD = 0., vC = 0.;
    for (int k = 0; k < nPair14; k++) {
        int i = list14[2 * k], j = list14[2 * k + 1];
        double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
        double r2 = oc_dist(b, d);
        double r = sqrt(r2), rt = 1. / r;
        double veps = eps14[i] * eps14[j], vsig = sig14[i] + sig14[j];
        double dv;
        double ev = oc_vdw14(p, vdwOn, scal14, veps, vsig, r2, rt, &dv);
        double fs;
        if (ewald) {
            /* ewald_dir14 / ewald_corr14 */
            double qel = p->chargeConst * q[i] * q[j];
            double ar = p->alpha * r;
            double ed = scal14 * qel * erfc(ar) * rt;
            double dd = 2.0 * scal14 * qel * p->alpha * exp(-SQ(ar)) / OC_SQRTPI;
            dd = -(ed + dd) * rt;
            double ec = (scal14 - 1.0) * qel * erf(ar) * rt;
            double dc = 2.0 * (scal14 - 1.0) * qel * p->alpha * exp(-SQ(ar)) / OC_SQRTPI;
            dc = (dc - ec) * rt;
            eD += ed; eC += ec; vD += dd * r; vC += dc * r;
            fs = dd + dc + dv;
        } else {
            double de;
            elec += oc_coulomb14(p, elecType, scal14, q[i] * q[j], r2, rt, &de);
            virel += de * r;
            fs = de + dv;
        }
        evdw += ev;
        virvdw += dv * r;
        double fxj = fs * d[0] * rt, fyj = fs * d[1] * rt, fzj = fs * d[2] * rt;
        fx[i] += fxj; fy[i] += fyj; fz[i] += fzj;
        fx[j] += -fxj; fy[j] += -fyj; fz[j] += -fzj;
    }
    if (ewald) {
        out[0] = eD + eC;
        out[2] = vD + vC;
    } else {
        out[0] = elec;
        out[2] = virel;
    }
    out[1] = evdw;
    out[3] = virvdw;
}

/* ---------------------------------------------------------------- */
/* Second derivatives of the Ewald-path potential and the trajectory */
/* adjoint of the VV + SHAKE dynamics.                               */
/* ---------------------------------------------------------------- */

/* Radial derivatives of one pair energy phi(r): d1 = phi', d2 = phi''.
 * Ewald real space, phi = qel erfc(a r)/r (ewald.c:356-369). */
static void hv_ewald_dir(double a, double qel, double r, double *d1, double *d2)
{
    double rt = 1. / r, ar = a * r;
    double ec = erfc(ar), g = 2. * a * exp(-SQ(ar)) / OC_SQRTPI;
    *d1 = -qel * (ec * rt + g) * rt;
    *d2 = qel * (2. * ec * rt * rt * rt + g * (2. * rt * rt + 2. * a * a));
}

/* Excluded-pair correction, phi = -qel erf(a r)/r (ewald.c:371-384). */
static void hv_ewald_corr(double a, double qel, double r, double *d1, double *d2)
{
    double rt = 1. / r, ar = a * r;
    double ef = erf(ar), g = 2. * a * exp(-SQ(ar)) / OC_SQRTPI;
    *d1 = qel * (ef * rt - g) * rt;
    *d2 = qel * (-2. * ef * rt * rt * rt + g * (2. * rt * rt + 2. * a * a));
}

/* Switched LJ (vdw.c:144-184), V = pv S with pv = 4 veps (s^12 - s^6),
 * s = vsig / r, S the CHARMM switch of r^2 (1 up to cutOn): d1 = V',
 * d2 = V'', dsig = dV'/dvsig. */
static void hv_lj(const oc_param *p, double veps, double vsig, double r, double *d1,
                  double *d2, double *dsig)
{
    double r2 = r * r, rt = 1. / r, s = vsig * rt;
    double s2 = SQ(s), s6 = s2 * s2 * s2, s12 = SQ(s6), s5 = s2 * s2 * s, s11 = s6 * s5;
    double pv = 4. * veps * (s12 - s6);
    double pv1 = 24. * veps * (s6 - 2. * s12) * rt;
    double pv2 = 4. * veps * (156. * s12 - 42. * s6) * rt * rt;
    double pvs = 4. * veps * (12. * s11 - 6. * s5) * rt;            /* d pv / d vsig */
    double pv1s = 24. * veps * (6. * s5 - 24. * s11) * rt * rt;     /* d pv' / d vsig */
    double S = 1., S1 = 0., S2 = 0.;
    if (r2 > p->cutOn2) {
        double sw1 = p->cutOff2 - r2, on = p->cutOn2 - r2;
        S = SQ(sw1) * (p->cutOff2 + 2. * r2 - 3. * p->cutOn2) * p->switch2;
        S1 = 12. * r * sw1 * on * p->switch2;
        S2 = 12. * p->switch2 * (sw1 * on - 2. * r2 * on - 2. * r2 * sw1);
    }
    *d1 = pv1 * S + pv * S1;
    *d2 = pv2 * S + 2. * pv1 * S1 + pv * S2;
    *dsig = pv1s * S + pvs * S1;
}

/* One pair's share of H u: with d = x_j - x_i, g = phi'/r and
 * h = (phi'' - g)/r^2, grad_j E = g d = -grad_i E and
 * (H u)_j = -(H u)_i = g du + h (d . du) d,  du = u_j - u_i. */
static void hv_pair_apply(double d1, double d2, double r, const double d[3], const double du[3],
                          double w[3])
{
    double g = d1 / r, h = (d2 - g) / (r * r);
    double t = d[0] * du[0] + d[1] * du[1] + d[2] * du[2];
    for (int c = 0; c < 3; c++) w[c] = g * du[c] + h * t * d[c];
}

/* B-spline second derivatives in bspgen's layout: bsdd[i] = M_P''(u+i)
 * = M_{P-2}(u+i) - 2 M_{P-2}(u+i-1) + M_{P-2}(u+i-2). */
static void hv_bspline2(double s, int P, double *bsdd)
{
    double b[16], db[16];
    oc_bspline(s, P - 2, b, db);
    for (int i = 0; i < P; i++) {
        double v0 = i < P - 2 ? b[i] : 0.;
        double v1 = (i >= 1 && i - 1 < P - 2) ? b[i - 1] : 0.;
        double v2 = (i >= 2 && i - 2 < P - 2) ? b[i - 2] : 0.;
        bsdd[i] = v0 - 2. * v1 + v2;
    }
}

/* The SPME mesh pieces of the second derivatives.  E_rec = fact sum_g
 * Q(g) phi(g), phi = theta * Q, Q(g) = sum_l q_l W_l(g), fact = (2pi/V) C;
 * with f2 = 2 fact and Q_u(g) = sum_l q_l grad W_l(g) . u_l,
 * phi_u = theta * Q_u:
 *   grad_l E_rec      = f2 q_l sum_g phi grad W_l             (grad, optional)
 *   (H_rec u)_l       = f2 q_l sum_g [phi (Hess W_l) u_l + phi_u grad W_l]
 *   d/dq_l (grad E_rec . u) = f2 sum_g [phi grad W_l . u_l + phi_u W_l].
 * Mesh, splines, index wrap and influence function as spme_core. */
static void hv_spme(const oc_param *p, const oc_box *b, const double *x, const double *y,
                    const double *z, const double *q, const double *u, double *hu, double *mq,
                    double *grad)
{
    int n = p->nAtom, m1 = p->m1, m2 = p->m2, m3 = p->m3, P = p->nbsp;
    long G = (long)m1 * m2 * m3;
    int mm[3] = {m1, m2, m3};
    double U[3][3] = {{b->u1, b->u2, b->u3}, {b->v1, b->v2, b->v3}, {b->w1, b->w2, b->w3}};
    double *sx = (double *)malloc((size_t)n * 3 * sizeof(double));
    double *bp = (double *)malloc((size_t)n * 3 * P * sizeof(double));
    double *bd = (double *)malloc((size_t)n * 3 * P * sizeof(double));
    double *bdd = (double *)malloc((size_t)n * 3 * P * sizeof(double));
    ocplx *fq = (ocplx *)calloc((size_t)G, sizeof(ocplx));
    ocplx *fu = (ocplx *)calloc((size_t)G, sizeof(ocplx));
    double *md[3];
    for (int a = 0; a < 3; a++) {
        md[a] = (double *)malloc((size_t)mm[a] * sizeof(double));
        oc_bspmod(mm[a], P, md[a]);
    }
    for (int i = 0; i < n; i++) {
        sx[3 * i + 0] = (double)m1 * (x[i] * b->u1 + y[i] * b->u2 + z[i] * b->u3 + 0.5);
        sx[3 * i + 1] = (double)m2 * (x[i] * b->v1 + y[i] * b->v2 + z[i] * b->v3 + 0.5);
        sx[3 * i + 2] = (double)m3 * (x[i] * b->w1 + y[i] * b->w2 + z[i] * b->w3 + 0.5);
        for (int a = 0; a < 3; a++) {
            oc_bspline(sx[3 * i + a], P, bp + (3 * i + a) * P, bd + (3 * i + a) * P);
            hv_bspline2(sx[3 * i + a], P, bdd + (3 * i + a) * P);
        }
    }
#define HV_IDX(l, a, i)                                                   \
    ({                                                                    \
        int ii_ = (int)sx[3 * (l) + (a)] - (i);                           \
        if (ii_ >= mm[a]) ii_ = 0;                                        \
        if (ii_ < 0) ii_ += mm[a];                                        \
        ii_;                                                              \
    })
    /* Q and Q_u */
    for (int l = 0; l < n; l++) {
        const double *B[3], *D[3];
        double ua[3] = {0., 0., 0.};
        for (int a = 0; a < 3; a++) {
            B[a] = bp + (3 * l + a) * P;
            D[a] = bd + (3 * l + a) * P;
            if (u) ua[a] = mm[a] * (U[a][0] * u[l] + U[a][1] * u[n + l] + U[a][2] * u[2 * n + l]);
        }
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++)
                for (int k = 0; k < P; k++) {
                    long g = HV_IDX(l, 2, k) + (long)m3 * (HV_IDX(l, 1, j) + (long)m2 * HV_IDX(l, 0, i));
                    fq[g] += q[l] * B[0][i] * B[1][j] * B[2][k];
                    fu[g] += q[l] * (D[0][i] * B[1][j] * B[2][k] * ua[0] +
                                     B[0][i] * D[1][j] * B[2][k] * ua[1] +
                                     B[0][i] * B[1][j] * D[2][k] * ua[2]);
                }
    }
    oc_fft3d((double *)fq, m1, m2, m3, +1);
    oc_fft3d((double *)fu, m1, m2, m3, +1);
    double rAlpha2 = -0.25 / SQ(p->alpha);
    double rc = fmin((double)m1 * b->u, (double)m2 * b->v);
    rc = fmin(rc, (double)m3 * b->w);
    rc = rc * 1.05 * OC_TWOPI;
    double rc2 = SQ(rc);
    for (int i = 0; i < m1; i++) {
        int ii = i > m1 / 2 ? i - m1 : i;
        for (int j = 0; j < m2; j++) {
            int jj = j > m2 / 2 ? j - m2 : j;
            for (int k = 0; k < m3; k++) {
                int kk = k > m3 / 2 ? k - m3 : k;
                long g = k + (long)m3 * (j + (long)m2 * i);
                double rx = OC_TWOPI * (ii * b->u1 + jj * b->v1 + kk * b->w1);
                double ry = OC_TWOPI * (ii * b->u2 + jj * b->v2 + kk * b->w2);
                double rz = OC_TWOPI * (ii * b->u3 + jj * b->v3 + kk * b->w3);
                double rm = SQ(rx) + SQ(ry) + SQ(rz);
                double th = 0.;
                if (rm > DBL_EPSILON && rm <= rc2)
                    th = md[0][i] * md[1][j] * md[2][k] * exp(rAlpha2 * rm) / rm;
                fq[g] *= th;
                fu[g] *= th;
            }
        }
    }
    oc_fft3d((double *)fq, m1, m2, m3, -1);
    oc_fft3d((double *)fu, m1, m2, m3, -1);
    double f2 = 2.0 * (OC_TWOPI / b->vol) * p->chargeConst;
    for (int l = 0; l < n; l++) {
        const double *B[3], *D[3], *DD[3];
        double ul[3] = {0., 0., 0.};
        if (u) { ul[0] = u[l]; ul[1] = u[n + l]; ul[2] = u[2 * n + l]; }
        double ua[3];
        for (int a = 0; a < 3; a++) {
            B[a] = bp + (3 * l + a) * P;
            D[a] = bd + (3 * l + a) * P;
            DD[a] = bdd + (3 * l + a) * P;
            ua[a] = mm[a] * (U[a][0] * ul[0] + U[a][1] * ul[1] + U[a][2] * ul[2]);
        }
        double h[3] = {0., 0., 0.}, gr[3] = {0., 0., 0.}, mql = 0.;
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++)
                for (int k = 0; k < P; k++) {
                    long g = HV_IDX(l, 2, k) + (long)m3 * (HV_IDX(l, 1, j) + (long)m2 * HV_IDX(l, 0, i));
                    double ph = creal(fq[g]), pu = creal(fu[g]);
                    /* dW/ds_a and d2W/ds_a ds_b */
                    double w1[3] = {D[0][i] * B[1][j] * B[2][k], B[0][i] * D[1][j] * B[2][k],
                                    B[0][i] * B[1][j] * D[2][k]};
                    double w2[3][3];
                    w2[0][0] = DD[0][i] * B[1][j] * B[2][k];
                    w2[1][1] = B[0][i] * DD[1][j] * B[2][k];
                    w2[2][2] = B[0][i] * B[1][j] * DD[2][k];
                    w2[0][1] = w2[1][0] = D[0][i] * D[1][j] * B[2][k];
                    w2[0][2] = w2[2][0] = D[0][i] * B[1][j] * D[2][k];
                    w2[1][2] = w2[2][1] = B[0][i] * D[1][j] * D[2][k];
                    double W = B[0][i] * B[1][j] * B[2][k];
                    for (int a = 0; a < 3; a++) {
                        double ca = 0.;   /* (Hess W u_l) along U_a, in s units */
                        for (int bb = 0; bb < 3; bb++) ca += w2[a][bb] * ua[bb];
                        for (int c = 0; c < 3; c++) {
                            h[c] += (ph * ca * mm[a] + pu * w1[a] * mm[a]) * U[a][c];
                            gr[c] += ph * w1[a] * mm[a] * U[a][c];
                        }
                        mql += ph * w1[a] * ua[a];
                    }
                    mql += pu * W;
                }
        for (int c = 0; c < 3; c++) {
            if (hu) hu[c * n + l] += f2 * q[l] * h[c];
            if (grad) grad[c * n + l] = f2 * q[l] * gr[c];
        }
        if (mq) mq[l] += f2 * mql;
    }
#undef HV_IDX
    for (int a = 0; a < 3; a++) free(md[a]);
    free(sx); free(bp); free(bd); free(bdd); free(fq); free(fu);
}

/* The SPME part alone (test access to hv_spme). */
void oc_spme_hvp(const oc_param *p, const oc_box *b, const double *x, const double *y,
                 const double *z, const double *q, const double *u, double *hu, double *mq,
                 double *grad)
{
    hv_spme(p, b, x, y, z, q, u, hu, mq, grad);
}

/* H u and d/d(q, eps, sig) (grad_x E . u) of the Ewald-path potential at
 * (x, y, z): real space over the list (pairs within cutOff), the
 * excluded-pair correction over the symmetric exclusion table exOff/exLst,
 * SPME.  hu (3n, component-major) and mq, me, ms (n) accumulate. */
void oc_hvp(const oc_param *p, const oc_box *b, const double *x, const double *y,
            const double *z, const double *q, const double *eps, const double *sig,
            const int *neighList, const int *neighPair, const int *neighOrder,
            const int *exOff, const int *exLst, const double *u, double *hu, double *mq,
            double *me, double *ms)
{
    int n = p->nAtom;
    double C = p->chargeConst;
    long offset = 0;
    for (int l = 0; l < n; l++) {
        int i = neighOrder[l];
        for (int k = 0; k < neighPair[i]; k++) {
            int j = neighList[offset++];
            double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
            double r2 = oc_dist(b, d);
            if (r2 > p->cutOff2) continue;
            double r = sqrt(r2);
            double du[3] = {u[j] - u[i], u[n + j] - u[n + i], u[2 * n + j] - u[2 * n + i]};
            double t = d[0] * du[0] + d[1] * du[1] + d[2] * du[2];
            double e1, e2, l1, l2, ls, w[3];
            hv_ewald_dir(p->alpha, 1.0, r, &e1, &e2);      /* per unit qel */
            hv_lj(p, 1.0, sig[i] + sig[j], r, &l1, &l2, &ls);   /* per unit veps */
            double qel = C * q[i] * q[j], veps = eps[i] * eps[j];
            hv_pair_apply(qel * e1 + veps * l1, qel * e2 + veps * l2, r, d, du, w);
            for (int c = 0; c < 3; c++) {
                hu[c * n + j] += w[c];
                hu[c * n + i] -= w[c];
            }
            mq[i] += C * q[j] * e1 / r * t;
            mq[j] += C * q[i] * e1 / r * t;
            me[i] += eps[j] * l1 / r * t;
            me[j] += eps[i] * l1 / r * t;
            ms[i] += veps * ls / r * t;
            ms[j] += veps * ls / r * t;
        }
    }
    for (int i = 0; i < n; i++)
        for (int k = exOff[i]; k < exOff[i + 1]; k++) {
            int j = exLst[k];
            double d[3] = {x[j] - x[i], y[j] - y[i], z[j] - z[i]};
            double r = sqrt(oc_dist(b, d));
            double du[3] = {u[j] - u[i], u[n + j] - u[n + i], u[2 * n + j] - u[2 * n + i]};
            double t = d[0] * du[0] + d[1] * du[1] + d[2] * du[2];
            double c1, c2, w[3];
            hv_ewald_corr(p->alpha, 1.0, r, &c1, &c2);
            hv_pair_apply(C * q[i] * q[j] * c1, C * q[i] * q[j] * c2, r, d, du, w);
            for (int c = 0; c < 3; c++) hu[c * n + i] -= w[c];   /* row i of the pair */
            mq[i] += C * q[j] * c1 / r * t;
        }
    hv_spme(p, b, x, y, z, q, u, hu, mq, NULL);
}

/* Rigid three-site molecules for the trajectory adjoint: constraint 3w+c
 * of molecule w joins atoms (a_c, b_c) = (0,1), (0,2), (1,2) of atoms
 * 3w..3w+2, as the water builder lays them out. */
static const int TA_CA[3] = {0, 0, 1}, TA_CB[3] = {1, 2, 2};

static int ta_layout_ok(int n, int nConst, const int *ca, const int *cb)
{
    if (n % 3 || nConst != n) return 0;
    for (int w = 0; w < n / 3; w++)
        for (int c = 0; c < 3; c++)
            if (ca[3 * w + c] != 3 * w + TA_CA[c] || cb[3 * w + c] != 3 * w + TA_CB[c]) return 0;
    return 1;
}

static void ta_solve3(const double M[3][3], const double r[3], double y[3], int trans)
{
    double A[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[i][j] = trans ? M[j][i] : M[i][j];
    /* Gaussian elimination with partial pivoting */
    double bb[3] = {r[0], r[1], r[2]};
    int piv[3] = {0, 1, 2};
    for (int c = 0; c < 3; c++) {
        int best = c;
        for (int k = c + 1; k < 3; k++)
            if (fabs(A[piv[k]][c]) > fabs(A[piv[best]][c])) best = k;
        int t = piv[c]; piv[c] = piv[best]; piv[best] = t;
        for (int k = c + 1; k < 3; k++) {
            double f = A[piv[k]][c] / A[piv[c]][c];
            for (int m = c; m < 3; m++) A[piv[k]][m] -= f * A[piv[c]][m];
            bb[piv[k]] -= f * bb[piv[c]];
        }
    }
    for (int c = 2; c >= 0; c--) {
        double s = bb[piv[c]];
        for (int m = c + 1; m < 3; m++) s -= A[piv[c]][m] * y[m];
        y[c] = s / A[piv[c]][c];
    }
}

/* (R M^-1 D^T)[c][e] for constraint rows r (of R) and d (of D). */
static void ta_gram(const double r[3][3], const double d[3][3], const double rm[3], double G[3][3])
{
    for (int c = 0; c < 3; c++)
        for (int e = 0; e < 3; e++) {
            double s = 0.;
            for (int k = 0; k < 3; k++) {
                double sr = k == TA_CA[c] ? -1. : (k == TA_CB[c] ? 1. : 0.);
                double sd = k == TA_CA[e] ? -1. : (k == TA_CB[e] ? 1. : 0.);
                s += sr * sd * rm[k] * (r[c][0] * d[e][0] + r[c][1] * d[e][1] + r[c][2] * d[e][2]);
            }
            G[c][e] = s;
        }
}

/* (R v)_c = r_c . (w_b v_b - w_a v_a), weights w (NULL: 1). */
static void ta_apply(const double r[3][3], const double v[3][3], const double *w, double out[3])
{
    for (int c = 0; c < 3; c++) {
        double wa = w ? w[TA_CA[c]] : 1., wb = w ? w[TA_CB[c]] : 1., s = 0.;
        for (int k = 0; k < 3; k++) s += r[c][k] * (wb * v[TA_CB[c]][k] - wa * v[TA_CA[c]][k]);
        out[c] = s;
    }
}

static void ta_add_rt(const double r[3][3], const double z[3], double s, double v[3][3])
{
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) {
            v[TA_CA[c]][k] -= s * z[c] * r[c][k];
            v[TA_CB[c]][k] += s * z[c] * r[c][k];
        }
}

static void ta_dirs(const oc_box *b, const double X[3][3], double r[3][3])
{
    for (int c = 0; c < 3; c++) {
        for (int k = 0; k < 3; k++) r[c][k] = X[TA_CB[c]][k] - X[TA_CA[c]][k];
        oc_dist2(b, &r[c][0], &r[c][1], &r[c][2]);
    }
}

/* Saved state of step t: x, y, z, vx, vy, vz, fx, fy, fz (n each). */
typedef struct {
    double xp[3][3], xt[3][3], vt[3][3], dx[3][3], vtil[3][3], rm[3];
    int at[3];
} ta_mol;

/* Molecule w at steps t-1 (S0) and t (S1): v' = v0 + dt/2 F0/m,
 * x~ = x0 + dt v', dx = x1 - x~ (minimum image: the wrap is a lattice
 * shift), v~ = v' + dx/dt + dt/2 F1/m (integrate.c:1518-1634). */
static void ta_load(const oc_box *b, int n, double dt, const double *S0, const double *S1,
                    const double *rmass, int w, ta_mol *M)
{
    for (int k = 0; k < 3; k++) {
        int i = 3 * w + k;
        M->at[k] = i;
        M->rm[k] = rmass[i];
        double d[3];
        for (int c = 0; c < 3; c++) {
            double vp = S0[(3 + c) * n + i] + 0.5 * dt * S0[(6 + c) * n + i] * rmass[i];
            double xtil = S0[c * n + i] + dt * vp;
            M->xp[k][c] = S0[c * n + i];
            M->xt[k][c] = S1[c * n + i];
            M->vt[k][c] = S1[(3 + c) * n + i];
            d[c] = S1[c * n + i] - xtil;
            M->vtil[k][c] = vp;
        }
        oc_dist2(b, &d[0], &d[1], &d[2]);
        for (int c = 0; c < 3; c++) {
            M->dx[k][c] = d[c];
            M->vtil[k][c] += d[c] / dt + 0.5 * dt * S1[(6 + c) * n + i] * rmass[i];
        }
    }
}

/* Stage 2, SHAKE-V as the projection v = v~ + M^-1 B^T mu with B(x_t) v = 0:
 * vb (adjoint of v_t) -> vb2 (adjoint of v~), x_t's part added to xb. */
static void ta_stage2(const oc_box *b, int n, const ta_mol *M, const double *vb, double *xb,
                      double *vb2)
{
    double r[3][3], S[3][3], dv[3][3], vbar[3][3], t[3], mu[3], kap[3];
    ta_dirs(b, M->xt, r);
    ta_gram(r, r, M->rm, S);
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
            dv[k][c] = M->vt[k][c] - M->vtil[k][c];
            vbar[k][c] = vb[c * n + M->at[k]];
        }
    ta_apply(r, dv, NULL, t);
    ta_solve3(S, t, mu, 0);
    ta_apply(r, vbar, M->rm, t);
    ta_solve3(S, t, kap, 0);
    double vtb[3][3];
    memcpy(vtb, vbar, sizeof vtb);
    ta_add_rt(r, kap, -1., vtb);
    for (int c = 0; c < 3; c++) {
        int ia = M->at[TA_CA[c]], ib = M->at[TA_CB[c]];
        for (int k = 0; k < 3; k++) {
            double G = mu[c] * (M->rm[TA_CB[c]] * vtb[TA_CB[c]][k] - M->rm[TA_CA[c]] * vtb[TA_CA[c]][k]) -
                       kap[c] * (M->vt[TA_CB[c]][k] - M->vt[TA_CA[c]][k]);
            xb[k * n + ib] += G;
            xb[k * n + ia] -= G;
        }
    }
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) vb2[c * n + M->at[k]] = vtb[k][c];
}

/* Stage 1, drift + SHAKE-R as x_t = x~ + M^-1 A(x_{t-1})^T gamma with the
 * constraints met at x_t, and v'' = v' + (x_t - x~)/dt: from xb (adjoint
 * of x_t) and vb2 (of v''), the adjoints of x_{t-1} (into xb), v_{t-1}
 * (into vb) and F_{t-1} (into fp). */
static void ta_stage1(const oc_box *b, int n, double dt, const ta_mol *M, double *xb,
                      const double *vb2, double *vb, double *fp)
{
    double dd[3][3], r[3][3], Q[3][3], T[3][3], t[3], gam[3], z[3];
    ta_dirs(b, M->xp, dd);
    ta_dirs(b, M->xt, r);
    ta_gram(dd, dd, M->rm, Q);
    ta_gram(r, dd, M->rm, T);
    ta_apply(dd, M->dx, NULL, t);
    ta_solve3(Q, t, gam, 0);
    double xe[3][3], v2b[3][3];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
            v2b[k][c] = vb2[c * n + M->at[k]];
            xe[k][c] = xb[c * n + M->at[k]] + v2b[k][c] / dt;
        }
    ta_apply(dd, xe, M->rm, t);
    ta_solve3(T, t, z, 1);
    double ub[3][3];
    memcpy(ub, xe, sizeof ub);
    ta_add_rt(r, z, -1., ub);
    double xtb[3][3], xpb[3][3];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
            xtb[k][c] = ub[k][c] - v2b[k][c] / dt;
            xpb[k][c] = xtb[k][c];
        }
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) {
            double coef = gam[c] * (M->rm[TA_CB[c]] * ub[TA_CB[c]][k] - M->rm[TA_CA[c]] * ub[TA_CA[c]][k]);
            xpb[TA_CB[c]][k] += coef;
            xpb[TA_CA[c]][k] -= coef;
        }
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
            int o = c * n + M->at[k];
            double vpb = v2b[k][c] + dt * xtb[k][c];
            xb[o] = xpb[k][c];
            vb[o] = vpb;
            fp[o] = 0.5 * dt * M->rm[k] * vpb;
        }
}

/* Forces F = -grad E + (1/N) sum_k grad_k E_rec (spme.c:648-670 removes
 * the mean reciprocal force):
 *   (dF/dx)^T u     = -H u + (1/N) H_rec (s, ..., s),       s = sum_l u_l
 *   (dF/dtheta)^T u = -d/dtheta (grad E . u) + (1/N) d/dq (grad E_rec . (s, ..., s)).
 * Adds the first to xb (when xb) and the second to gq, ge, gs. */
static void ta_force_vjp(const oc_param *p, const oc_box *b, const double *x, const double *y,
                         const double *z, const double *q, const double *eps, const double *sig,
                         const int *nl, const int *np, const int *no, const int *exOff,
                         const int *exLst, const double *u, double *xb, double *gq, double *ge,
                         double *gs)
{
    int n = p->nAtom;
    double *hu = (double *)calloc((size_t)3 * n, sizeof(double));
    double *mq = (double *)calloc((size_t)n, sizeof(double));
    double *me = (double *)calloc((size_t)n, sizeof(double));
    double *ms = (double *)calloc((size_t)n, sizeof(double));
    double *us = (double *)malloc((size_t)3 * n * sizeof(double));
    oc_hvp(p, b, x, y, z, q, eps, sig, nl, np, no, exOff, exLst, u, hu, mq, me, ms);
    for (int i = 0; i < n; i++) {
        gq[i] -= mq[i];
        ge[i] -= me[i];
        gs[i] -= ms[i];
    }
    if (xb)
        for (long k = 0; k < 3L * n; k++) xb[k] -= hu[k];
    double s[3] = {0., 0., 0.};
    for (int c = 0; c < 3; c++)
        for (int i = 0; i < n; i++) s[c] += u[c * n + i];
    for (int c = 0; c < 3; c++)
        for (int i = 0; i < n; i++) us[c * n + i] = s[c];
    memset(hu, 0, (size_t)3 * n * sizeof(double));
    memset(mq, 0, (size_t)n * sizeof(double));
    hv_spme(p, b, x, y, z, q, us, hu, mq, NULL);
    for (int i = 0; i < n; i++) gq[i] += mq[i] / (double)n;
    if (xb)
        for (long k = 0; k < 3L * n; k++) xb[k] += hu[k] / (double)n;
    free(hu); free(mq); free(me); free(ms); free(us);
}

/* J = (1/T) sum_{t=1..T} E_pot(x_t) over T steps of the MD loop of
 * oc_md_run (NVE, VV + SHAKE, list update every p->update steps) from
 * (x0, v0), and its gradient with respect to q, eps, sig by a reverse
 * sweep over the stored steps: the objective's own terms, SHAKE-V and
 * SHAKE-R by the implicit function theorem on each rigid molecule, and
 * the force vector-Jacobian products by ta_force_vjp, ending with F_0.
 * Rigid three-site molecules only (ta_layout_ok), no frozen atoms.
 * Returns 0, 311 (SHAKE), or 1 for a layout it does not cover. */
int oc_traj_adjoint(const oc_param *p, const oc_box *b, int keyLink, int T,
                    const double *x0, const double *y0, const double *z0,
                    const double *vx0, const double *vy0, const double *vz0,
                    const double *q, const double *eps, const double *sig,
                    const double *mass, const double *rmass, const int *frozen,
                    const int *nAtConst, int nConst, const int *ca, const int *cb,
                    const double *rc2, double *J, double *dq, double *deps, double *dsig)
{
    int n = p->nAtom, err = 0;
    if (T < 1 || !ta_layout_ok(n, nConst, ca, cb)) return 1;
    for (int i = 0; i < n; i++)
        if (frozen && frozen[i]) return 1;
    long sn = 9L * n;
    double *S = (double *)malloc((size_t)(T + 1) * sn * sizeof(double));
    int *exOff = (int *)malloc((size_t)(n + 1) * sizeof(int));
    int *sxOff = (int *)malloc((size_t)(n + 1) * sizeof(int));
    int excap = 8 * nConst + 16;
    int *exLst = (int *)malloc((size_t)excap * sizeof(int));
    int *sxLst = (int *)malloc((size_t)excap * sizeof(int));
    int ne = oc_exclusions(n, 0, NULL, NULL, nConst, ca, cb, 0, NULL, NULL, NULL, keyLink ? 1 : 0,
                           exOff, exLst, excap);
    int ns = oc_exclusions(n, 0, NULL, NULL, nConst, ca, cb, 0, NULL, NULL, NULL, 1, sxOff, sxLst,
                           excap);
    oc_list L;
    L.cap = 1024;
    L.nl = (int *)malloc((size_t)L.cap * sizeof(int));
    L.np = (int *)malloc((size_t)n * sizeof(int));
    L.no = (int *)malloc((size_t)n * sizeof(int));
    double *xb = (double *)calloc((size_t)3 * n, sizeof(double));
    double *vb = (double *)calloc((size_t)3 * n, sizeof(double));
    double *vb2 = (double *)calloc((size_t)3 * n, sizeof(double));
    double *fp = (double *)calloc((size_t)3 * n, sizeof(double));
    double *fbar = (double *)calloc((size_t)3 * n, sizeof(double));
    double *gr = (double *)calloc((size_t)3 * n, sizeof(double));
    double *tq = (double *)calloc((size_t)n, sizeof(double));
    double *te = (double *)calloc((size_t)n, sizeof(double));
    double *ts = (double *)calloc((size_t)n, sizeof(double));
    if (ne < 0 || ns < 0) { err = 1; goto done; }
    for (int i = 0; i < n; i++) { dq[i] = 0.; deps[i] = 0.; dsig[i] = 0.; }
    {
        /* forward, as oc_md_run, keeping (x, v, F) of every step */
        double *x = S, *y = S + n, *z = S + 2 * n;
        memcpy(x, x0, n * sizeof(double)); memcpy(y, y0, n * sizeof(double));
        memcpy(z, z0, n * sizeof(double)); memcpy(S + 3 * n, vx0, n * sizeof(double));
        memcpy(S + 4 * n, vy0, n * sizeof(double)); memcpy(S + 5 * n, vz0, n * sizeof(double));
        err = md_list(p, b, keyLink, 1, x, y, z, frozen, exOff, exLst, &L);
        double Jsum = 0.;
        for (int t = 0; t <= T && !err; t++) {
            double *St = S + sn * t;
            if (t > 0) {
                memcpy(St, St - sn, (size_t)sn * sizeof(double));
                err = oc_vv_nve(p, b, 1, nConst, ca, cb, rc2, St, St + n, St + 2 * n, St + 3 * n,
                                St + 4 * n, St + 5 * n, St + 6 * n, St + 7 * n, St + 8 * n, mass,
                                rmass, nAtConst, NULL, NULL, NULL);
                if (err) break;
                if (t % p->update == 0)
                    err = md_list(p, b, keyLink, 0, St, St + n, St + 2 * n, frozen, exOff, exLst, &L);
                if (err) break;
            }
            double *f = St + 6 * n;
            memset(f, 0, (size_t)3 * n * sizeof(double));
            double st[6], vr;
            double eRec = oc_spme(p, b, St, St + n, St + 2 * n, f, f + n, f + 2 * n, q, st, &vr);
            oc_energy e;
            oc_ewald_real(p, b, St, St + n, St + 2 * n, f, f + n, f + 2 * n, q, eps, sig, L.nl, L.np,
                          L.no, exOff, exLst, &e);
            if (t > 0) {
                Jsum += (e.eDir + eRec + e.eCorr) + e.eVdw;
                err = oc_vv_nve(p, b, 2, nConst, ca, cb, rc2, St, St + n, St + 2 * n, St + 3 * n,
                                St + 4 * n, St + 5 * n, f, f + n, f + 2 * n, mass, rmass, nAtConst,
                                NULL, NULL, NULL);
            }
        }
        if (err) goto done;
        *J = Jsum / (double)T;
    }
    /* reverse sweep */
    for (int t = T; t >= 1 && !err; t--) {
        const double *Sc = S + sn * t, *Sp = S + sn * (t - 1);
        const double *x = Sc, *y = Sc + n, *z = Sc + 2 * n, *f = Sc + 6 * n;
        /* the list of x_t (the in-cutoff set of any valid list) */
        double *xc = (double *)malloc((size_t)3 * n * sizeof(double));
        memcpy(xc, Sc, (size_t)3 * n * sizeof(double));
        err = md_list(p, b, keyLink, 0, xc, xc + n, xc + 2 * n, frozen, exOff, exLst, &L);
        free(xc);
        if (err) break;
        /* objective: (1/T) dE/dtheta and (1/T) grad E = -(1/T)(F + m) */
        oc_energy_grad(p, b, x, y, z, q, eps, sig, L.nl, L.np, L.no, exOff, exLst, tq, te, ts);
        for (int i = 0; i < n; i++) {
            dq[i] += tq[i] / (double)T;
            deps[i] += te[i] / (double)T;
            dsig[i] += ts[i] / (double)T;
        }
        hv_spme(p, b, x, y, z, q, NULL, NULL, NULL, gr);
        double m[3] = {0., 0., 0.};
        for (int c = 0; c < 3; c++) {
            for (int i = 0; i < n; i++) m[c] -= gr[c * n + i];
            m[c] /= (double)n;
        }
        for (int c = 0; c < 3; c++)
            for (int i = 0; i < n; i++) xb[c * n + i] -= (f[c * n + i] + m[c]) / (double)T;
        for (int w = 0; w < n / 3; w++) {
            ta_mol M;
            ta_load(b, n, p->timeStep, Sp, Sc, rmass, w, &M);
            ta_stage2(b, n, &M, vb, xb, vb2);
        }
        for (long k = 0; k < 3L * n; k++)
            fbar[k] = 0.5 * p->timeStep * rmass[k % n] * vb2[k] + fp[k];
        ta_force_vjp(p, b, x, y, z, q, eps, sig, L.nl, L.np, L.no, sxOff, sxLst, fbar, xb, dq,
                     deps, dsig);
        for (int w = 0; w < n / 3; w++) {
            ta_mol M;
            ta_load(b, n, p->timeStep, Sp, Sc, rmass, w, &M);
            ta_stage1(b, n, p->timeStep, &M, xb, vb2, vb, fp);
        }
    }
    if (!err) {
        /* F_0 depends on the parameters too */
        double *xc = (double *)malloc((size_t)3 * n * sizeof(double));
        memcpy(xc, S, (size_t)3 * n * sizeof(double));
        err = md_list(p, b, keyLink, 0, xc, xc + n, xc + 2 * n, frozen, exOff, exLst, &L);
        free(xc);
        if (!err)
            ta_force_vjp(p, b, S, S + n, S + 2 * n, q, eps, sig, L.nl, L.np, L.no, sxOff, sxLst, fp,
                         NULL, dq, deps, dsig);
    }
done:
    free(S); free(exOff); free(sxOff); free(exLst); free(sxLst);
    free(L.nl); free(L.np); free(L.no);
    free(xb); free(vb); free(vb2); free(fp); free(fbar); free(gr); free(tq); free(te); free(ts);
    return err;
}

/* oc_vv_nve with each constraint's convergence test of every SHAKE sweep
 * recorded in conv (cap sweeps of nConst bytes); *ncycle = sweeps run. */
int oc_vv_nve_conv(const oc_param *p, const oc_box *b, int stage,
                   int nConst, const int *ca, const int *cb, const double *rc2,
                   double *x, double *y, double *z, double *vx, double *vy, double *vz,
                   const double *fx, const double *fy, const double *fz,
                   const double *mass, const double *rmass, const int *nAtConst,
                   unsigned char *conv, int cap, int *ncycle)
{
    g_conv = conv;
    g_conv_cap = cap;
    double ek, vs;
    int err = oc_vv_nve(p, b, stage, nConst, ca, cb, rc2, x, y, z, vx, vy, vz, fx, fy, fz, mass,
                        rmass, nAtConst, &ek, &vs, ncycle);
    g_conv = NULL;
    g_conv_cap = 0;
    return err;
}
