/*
 * oracle.c -- CPU restatement of the ems engine step (TEST INFRASTRUCTURE).
 *
 * Parity oracle and CPU baseline only; see oracle.h for what may use it and
 * how each part is pinned. Compiled with -O2 -ffp-contract=off and no
 * fast-math, so every f32/f64 expression rounds exactly as the reference's
 * Rust code (no FMA contraction, SSE arithmetic on x86-64).
 *
 * Reference files cited below are in the ems repository (src/...).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* constants (src/constants.rs)                                        */
/* ------------------------------------------------------------------ */
static const double C0 = 299792458.0;
static const double MU0 = 1.25663706127e-6;
#define EPS0 (1.0 / (MU0 * C0 * C0))
#define Z0 (MU0 * C0)

/* kernel.rs:30-33: subnormal results become 0, everything else unchanged. */
static inline float flush(float x) { return fabsf(x) < FLT_MIN ? 0.0f : x; }

/* kernel.rs:38-42 (openEMS order, every op flushed). */
static inline float oe_h(float da, float h, float db, float a, float b, float c, float d) {
    float curl = flush(flush(flush(a - b) + d) - c);
    return flush(flush(da * h) - flush(db * curl));
}
/* kernel.rs:46-50 */
static inline float oe_e(float ca, float u, float cb, float a, float b, float c, float d) {
    float curl = flush(flush(flush(a - b) - c) + d);
    return flush(flush(ca * u) + flush(cb * curl));
}
/* kernel.rs:76 (Paired order) */
static inline float pair_h(float da, float h, float db, float a, float b, float c, float d) {
    return flush(da * h - db * ((a - b) - (c - d)));
}
/* kernel.rs:270 */
static inline float pair_e(float ca, float u, float cb, float a, float b, float c, float d) {
    return flush(ca * u + cb * ((a - b) - (c - d)));
}
/* kernel.rs:114-117: (1 + a)*p/b, 0 where b = 0. */
static inline double weight(float a, double p, float b) {
    double w = (1.0 + (double)a) * p / (double)b;
    return b != 0.0f ? w : 0.0;
}

/* ------------------------------------------------------------------ */
/* engine state                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    float *rows;
    uint32_t *index; /* NULL: dense */
    uint64_t nrows;
} lane_t;

typedef struct {
    int e_side, comp, other, normal;
    float sign;
    int lo[3], hi[3];
    float *b, *a, *k, *psi;
} cpml_t;

enum {
    OP_SOURCE = 1, OP_PROBE = 2, OP_LUMPED = 3, OP_POLES = 4, OP_PW = 5, OP_MERGED = 6, OP_MUR = 7,
    OP_UPML = 8
};

/* MurBoundary (boundary/mur.rs:94-140): face edges (Edge) and domain edges
 * on two Mur faces (Corner), with the values saved before the E update. */
typedef struct {
    int ne, nc, oe_order, oe_corners;
    int *eax, *cax;
    size_t *enode, *einner, *cnode, *cin0, *cin1;
    float *ek, *ck0, *ck1;
    uint64_t *ehold, *chold;
    float *saved; /* 2 per edge, then 3 per corner */
} mur_t;

/* OpenEmsPml (boundary/upml.rs:541-760): per side (0 E edges, 1 H dual
 * edges) the layer edges in the caller's order with their coefficients
 * ((1-a_p)/(1+a_p), (1-a_q)/(1+a_q), (1+a_a)/((1+a_q)m), (1-a_a)/((1+a_q)m))
 * and state (flux F, start value v), and the edges by x plane for the
 * slabs. */
typedef struct {
    int n[2];
    int *ax[2];
    size_t *idx[2];
    float *co[2];
    float *f[2], *v[2];
    int *perm[2];      /* edge numbers sorted by plane, stable */
    size_t *pstart[2]; /* nx + 1 offsets into perm */
} upml_t;

/* Plane-wave corrections (Correction, planewave.rs:424-431), in list
 * order, and the auxiliary grid (PlaneWave, :458-483). */
typedef struct {
    int n;
    int *axis;
    size_t *idx;
    double *coef, *w;
    uint32_t *m;
} pwcorr_t;
typedef struct {
    int naux;
    double *e, *h, *ca, *cb, *da, *db;
    pwcorr_t ec, hc;
    double *src;
    uint64_t nsrc;
} pw_t;

/* One side of a dispersive medium (Side, dispersive.rs:977-1040): edges in
 * (plane, axis, j, k) order, poles in CSR layout, f64 branch state. */
typedef struct {
    int n;
    int *axis, *plane;
    size_t *idx;
    double *siz;
    uint32_t *pole_end, *cidx;
    uint64_t np;
    double *j, *q;
    float *prev;
} poleside_t;

typedef struct {
    poleside_t side[2]; /* 0 electric, 1 magnetic */
    int nc;
    double *a, *inv_c, *inv_z, *half_l;
    int *kind; /* 0 trapezoidal, 1 leapfrog */
} poles_t;

/* An installed lumped element (lumped.rs:367-406, Installed + Branch +
 * CurrentRecord), optionally a LumpedPort (engine/port.rs:94-104). */
typedef struct {
    int axis, n, series, src_in_branch, branch, nterm, ndrive, columns;
    size_t *idx;
    double *cb, *gsrc;
    double *tw;               /* nterm+ndrive weight rows of n */
    double **tv;              /* per term: value table by step */
    uint64_t *tlen;
    double *v, *vd;           /* scratch: source / drive voltage per edge */
    double *two_l_dt, *inv_c, *z, *j, *q, *javg;
    double *uprev;            /* u^n saved by pre_e */
    int record;
    double *rg, *rc;
    double *rec, *pv, *pi;    /* record values, port V and I */
    uint64_t nrec, cap;
} lumped_t;

typedef struct {
    int type;
    /* source */
    int kind;
    int n;
    int *axis;
    size_t *idx;
    float *wf;
    float *amp;
    uint64_t namp;
    /* probe */
    int quantity, f32sum;
    double *wd;
    uint64_t interval;
    double *vals;
    uint64_t nvals, cap;
    /* lumped element */
    lumped_t *lp;
    /* dispersive medium */
    poles_t *pm;
    /* plane wave */
    pw_t *pw;
    /* reduced currents of cut faces */
    struct merged *mf;
    /* Mur boundary */
    mur_t *mur;
    /* openEMS's PML layers */
    upml_t *upml;
} op_t;

/* MergedFaces (conformal.rs:2313-2340): the tables of orc_merged_desc and
 * the state, Group::i per reduced current and the circulations. */
typedef struct merged {
    int ng, ns, nb, nf;
    double dt;
    double *da, *db, *l, *tw, *ev, *bv, *mv, *fw;
    uint32_t *ts, *bs, *mem, *ef, *brs, *bc, *mrs, *mc, *fs, *fg;
    uint64_t *eo;
    int *tax, *fax;
    size_t *tidx, *fidx;
    double *i, *circ, *next, *x;
    /* CPML terms of driven faces in layers (PmlTerm, conformal.rs:2398-2412) */
    int np;
    int *pax;
    size_t *pi0, *pi1;
    double *pc, *psi;
    uint32_t *pst, *pt;
    double *pw;
} merged_t;

struct orc_engine {
    int n[3];
    size_t N, plane;
    double dt;
    int oe;
    float *u[3], *i[3];
    float *zero;
    lane_t lanes[12];
    cpml_t *cpml;
    int ncpml;
    op_t *ops;
    int nops;
    uint64_t step;
    double *plane_energy;
    /* threads */
    int threads;
    int *slab_start;
    pthread_t *workers;
    pthread_barrier_t bar;
    int cmd; /* 1 H, 2 E, 3 energy (read only), 4/5 fused phases, 9 exit */
    int fuse; /* FusedSteps: 0 Auto, 1 Always, 2 Never (pool.rs:131-164) */
    int energy_flag;
    char err[256];
};

static inline int axis_next(int a) { return (a + 1) % 3; }
static inline int axis_prev(int a) { return (a + 2) % 3; }

/* CoeffRows::row (rows.rs:45-50): row (gi, j) of lane l. */
static inline const float *row(const orc_engine *e, int l, int gi, int j) {
    const lane_t *L = &e->lanes[l];
    size_t r = (size_t)gi * e->n[1] + j;
    size_t id = L->index ? L->index[r] : r;
    return L->rows + id * e->n[2];
}

/* ------------------------------------------------------------------ */
/* H update of one x-plane, with the energy share (kernel.rs:119-173,   */
/* 364-423, 443-511).                                                   */
/* ------------------------------------------------------------------ */

/* One H row; with energy the 4 interleaved f64 accumulators of
 * h_row_energy (kernel.rs:145-172): entry k into acc[k & 3], result
 * (a0 + a1) + (a2 + a3). */
static double hrow(const orc_engine *e, int energy, int commit, float *h, const float *da,
                   const float *db, const float *a, const float *b, const float *c,
                   const float *d, int n) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < n; k++) {
        float old = h[k];
        float nw = e->oe ? oe_h(da[k], old, db[k], a[k], b[k], c[k], d[k])
                         : pair_h(da[k], old, db[k], a[k], b[k], c[k], d[k]);
        if (energy) acc[k & 3] += weight(da[k], (double)old * (double)nw, db[k]);
        if (commit) h[k] = nw;
    }
    return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

/* h_plane (kernel.rs:447-511) + the E part of h_slab's energy
 * (kernel.rs:404-420). Returns the plane value we + wh (energy) or 0. */
static double h_plane(orc_engine *e, int gi, int energy, int commit) {
    const int nx = e->n[0], ny = e->n[1], nz = e->n[2];
    const size_t p = e->plane, off = (size_t)gi * p;
    float *ix = e->i[0] + off, *iy = e->i[1] + off, *iz = e->i[2] + off;
    const float *ux = e->u[0] + off, *uy = e->u[1] + off, *uz = e->u[2] + off;
    const int m = nz - 1;
    double wh = 0.0;
    for (int j = 0; j < ny - 1; j++) {
        size_t r = (size_t)j * nz;
        wh += hrow(e, energy, commit, ix + r, row(e, 6, gi, j), row(e, 9, gi, j), uz + r + nz,
                   uz + r, uy + r + 1, uy + r, m);
    }
    if (gi + 1 < nx) {
        const float *uyn = uy + p, *uzn = uz + p;
        for (int j = 0; j < ny; j++) {
            size_t r = (size_t)j * nz;
            wh += hrow(e, energy, commit, iy + r, row(e, 7, gi, j), row(e, 10, gi, j), ux + r + 1,
                       ux + r, uzn + r, uz + r, m);
        }
        for (int j = 0; j < ny - 1; j++) {
            size_t r = (size_t)j * nz;
            wh += hrow(e, energy, commit, iz + r, row(e, 8, gi, j), row(e, 11, gi, j), uyn + r,
                       uy + r, ux + r + nz, ux + r, nz);
        }
    }
    if (!energy) return 0.0;
    /* e_energy_row (kernel.rs:218-244): per component, 4 accumulators
     * chained over the whole plane by in-plane index & 3. */
    double we = 0.0;
    const float *uc[3] = {ux, uy, uz};
    for (int a = 0; a < 3; a++) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int j = 0; j < ny; j++) {
            const float *ca = row(e, a, gi, j), *cb = row(e, 3 + a, gi, j);
            size_t r = (size_t)j * nz;
            for (int k = 0; k < nz; k++) {
                double x = (double)uc[a][r + k];
                acc[(r + k) & 3] += weight(ca[k], x * x, cb[k]);
            }
        }
        we += (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
    return we + wh;
}

/* update_e_plane / e_rows (kernel.rs:552-697). */
static void e_plane(orc_engine *e, int gi) {
    const int nx = e->n[0], ny = e->n[1], nz = e->n[2];
    const size_t p = e->plane, off = (size_t)gi * p;
    float *ux = e->u[0] + off, *uy = e->u[1] + off, *uz = e->u[2] + off;
    const float *ix = e->i[0] + off, *iy = e->i[1] + off, *iz = e->i[2] + off;
    const float *iyp = gi == 0 ? e->zero : iy - p;
    const float *izp = gi == 0 ? e->zero : iz - p;
    const float *zero = e->zero;
    const int oe = e->oe;
#define EUPD(ca, u, cb, a, b, c, d) (oe ? oe_e(ca, u, cb, a, b, c, d) : pair_e(ca, u, cb, a, b, c, d))
    for (int j = 0; j < ny; j++) {
        size_t r = (size_t)j * nz;
        if (gi + 1 < nx) {
            const float *ca = row(e, 0, gi, j), *cb = row(e, 3, gi, j);
            const float *izm = j == 0 ? zero : iz + r - nz;
            for (int k = 0; k < nz; k++) {
                float dm = k == 0 ? 0.0f : iy[r + k - 1];
                ux[r + k] = EUPD(ca[k], ux[r + k], cb[k], iz[r + k], izm[k], iy[r + k], dm);
            }
        }
        if (j + 1 < ny) {
            const float *ca = row(e, 1, gi, j), *cb = row(e, 4, gi, j);
            for (int k = 0; k < nz; k++) {
                float bm = k == 0 ? 0.0f : ix[r + k - 1];
                uy[r + k] = EUPD(ca[k], uy[r + k], cb[k], ix[r + k], bm, iz[r + k], izp[r + k]);
            }
        }
        {
            const float *ca = row(e, 2, gi, j), *cb = row(e, 5, gi, j);
            const float *ixm = j == 0 ? zero : ix + r - nz;
            for (int k = 0; k < nz - 1; k++)
                uz[r + k] = EUPD(ca[k], uz[r + k], cb[k], iy[r + k], iyp[r + k], ix[r + k], ixm[k]);
        }
    }
#undef EUPD
}

/* ------------------------------------------------------------------ */
/* CPML slab correction (cpml.rs:469-543, 693-736): for the x-planes of  */
/* [x0, x1).                                                             */
/* ------------------------------------------------------------------ */
static void cpml_apply(orc_engine *e, cpml_t *s, int x0, int x1) {
    const int ny = e->n[1], nz = e->n[2];
    const size_t p = e->plane;
    const int es = s->e_side;
    float **tgt = es ? e->u : e->i;
    float **src = es ? e->i : e->u;
    const float scale = es ? 1.0f : -1.0f;
    const int lane = es ? 3 + s->comp : 9 + s->comp;
    const int len = s->hi[2] - s->lo[2], rows = s->hi[1] - s->lo[1];
    int a0 = s->lo[0] > x0 ? s->lo[0] : x0;
    int a1 = s->hi[0] < x1 ? s->hi[0] : x1;
    for (int i = a0; i < a1; i++) {
        const float *cur = src[s->other] + (size_t)i * p;
        const float *oth = cur;
        if (s->normal == 0) oth = es ? (i == 0 ? e->zero : cur - p) : cur + p;
        float *t = tgt[s->comp] + (size_t)i * p;
        for (int j = s->lo[1]; j < s->hi[1]; j++) {
            const float *cf = row(e, lane, i, j);
            size_t q = ((size_t)(i - s->lo[0]) * rows + (j - s->lo[1])) * len;
            for (int k = s->lo[2]; k < s->hi[2]; k++, q++) {
                size_t o = (size_t)j * nz + k;
                float h, l;
                int z;
                switch (s->normal) {
                case 0:
                    h = es ? cur[o] : oth[o];
                    l = es ? oth[o] : cur[o];
                    z = i - s->lo[0];
                    break;
                case 1:
                    h = es ? cur[o] : cur[o + nz];
                    l = es ? (j == 0 ? 0.0f : cur[o - nz]) : cur[o];
                    z = j - s->lo[1];
                    break;
                default:
                    h = es ? cur[o] : cur[o + 1];
                    l = es ? (k == 0 ? 0.0f : cur[o - 1]) : cur[o];
                    z = k - s->lo[2];
                    break;
                }
                float sv = s->sign * (h - l);
                float ps = flush(s->b[z] * s->psi[q] + s->a[z] * sv);
                s->psi[q] = ps;
                t[o] = flush(t[o] + scale * cf[k] * (s->k[z] * sv + ps));
            }
        }
    }
    (void)ny;
}

/* ------------------------------------------------------------------ */
/* slab-parallel phases (pool.rs:314-403)                               */
/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* dispersive media as slab hooks (dispersive.rs:1700-1750)             */
/* ------------------------------------------------------------------ */

/* Side::edges_in (dispersive.rs:1166-1171): edges of the planes [p0, p1). */
static void edges_in(const poleside_t *s, int p0, int p1, int *lo, int *hi) {
    int a = 0, b = s->n;
    while (a < b) {
        int m = (a + b) / 2;
        if (s->plane[m] < p0) a = m + 1; else b = m;
    }
    *lo = a;
    b = s->n;
    while (a < b) {
        int m = (a + b) / 2;
        if (s->plane[m] < p1) a = m + 1; else b = m;
    }
    *hi = a;
}

/* Side::correct through step_edge (dispersive.rs:1183-1343, 1392-1470)
 * for trapezoidal and leapfrog poles; lane 3 + axis (cb) for the electric
 * side on u, 9 + axis (db) for the magnetic side on i. */
static void side_correct(orc_engine *e, const poles_t *m, int f, int p0, int p1) {
    const poleside_t *s = &m->side[f];
    int lo, hi;
    edges_in(s, p0, p1, &lo, &hi);
    const int ny = e->n[1], nz = e->n[2];
    for (int ed = lo; ed < hi; ed++) {
        const int ax = s->axis[ed];
        const size_t idx = s->idx[ed];
        float *fld = f == 0 ? e->u[ax] : e->i[ax];
        const int gi = s->plane[ed];
        const size_t off = idx - (size_t)gi * ny * nz;
        const double cb = (double)row(e, (f == 0 ? 3 : 9) + ax, gi, (int)(off / nz))[off % nz];
        const uint32_t q0 = ed ? s->pole_end[ed - 1] : 0, q1 = s->pole_end[ed];
        const double prev = (double)s->prev[ed];
        const double half_prev = 0.5 * prev;
        const double u_star = (double)fld[idx];
        double sum = 0.0;
        for (uint32_t q = q0; q < q1; q++) {
            const int c = s->cidx[q];
            if (m->kind[c] == 1) /* leapfrog: j^{n+1/2} from u^n */
                sum += m->a[c] * s->j[q] + m->inv_z[c] * (prev - s->q[q] * m->inv_c[c]);
            else
                sum += (m->a[c] * s->j[q] - s->q[q] * m->inv_c[c] + half_prev) * m->inv_z[c];
        }
        const float u_new = (float)((u_star - cb * sum) / (1.0 + 0.5 * cb * s->siz[ed]));
        const double half_new = 0.5 * (double)u_new;
        for (uint32_t q = q0; q < q1; q++) {
            const int c = s->cidx[q];
            const double jo = s->j[q], qo = s->q[q];
            if (m->kind[c] == 1) {
                const double jn = m->a[c] * jo + m->inv_z[c] * (prev - qo * m->inv_c[c]);
                s->j[q] = jn;
                s->q[q] = qo + e->dt * jn;
            } else {
                const double k = m->a[c] * jo - qo * m->inv_c[c] + half_prev;
                const double jb = (k + half_new) * m->inv_z[c];
                s->j[q] = 2.0 * jb - jo;
                s->q[q] = qo + e->dt * jb;
            }
        }
        fld[idx] = u_new;
    }
}

/* Side::capture (dispersive.rs:1173-1179). */
static void side_capture(orc_engine *e, const poles_t *m, int f, int p0, int p1) {
    const poleside_t *s = &m->side[f];
    int lo, hi;
    edges_in(s, p0, p1, &lo, &hi);
    for (int ed = lo; ed < hi; ed++)
        s->prev[ed] = (f == 0 ? e->u : e->i)[s->axis[ed]][s->idx[ed]];
}

/* Parts::post_h_planes / post_e_planes (dispersive.rs:1700-1750). */
static void poles_post_h(orc_engine *e, poles_t *m, int p0, int p1) {
    side_correct(e, m, 1, p0, p1);
    side_capture(e, m, 0, p0, p1);
}
static void poles_post_e(orc_engine *e, poles_t *m, int p0, int p1) {
    side_correct(e, m, 0, p0, p1);
    side_capture(e, m, 1, p0, p1);
}

/* Side::add_energy (dispersive.rs:1489-1517) over one side, from 0.0. */
static double side_energy(const poles_t *m, int f) {
    const poleside_t *s = &m->side[f];
    double w = 0.0;
    for (uint64_t q = 0; q < s->np; q++) {
        const int c = s->cidx[q];
        w += m->half_l[c] * s->j[q] * s->j[q] + 0.5 * m->inv_c[c] * s->q[q] * s->q[q];
    }
    return w;
}

/* Run::refresh (upml.rs:391-400): side s's edges in planes [p0, p1) take
 * their start value from the field. */
static void upml_refresh(orc_engine *e, upml_t *m, int s, int p0, int p1) {
    float *const *fld = s ? e->i : e->u;
    for (size_t r = m->pstart[s][p0]; r < m->pstart[s][p1]; r++) {
        const int q = m->perm[s][r];
        m->v[s][q] = fld[m->ax[s][q]][m->idx[s][q]];
    }
}

/* Run::apply (upml.rs:404-426): the update left the flux increment t in the
 * field; F' = c0 F + t, x = (c1 v - c3 F) + c2 F', every operation flushed. */
static void upml_apply(orc_engine *e, upml_t *m, int s, int p0, int p1) {
    float *const *fld = s ? e->i : e->u;
    for (size_t r = m->pstart[s][p0]; r < m->pstart[s][p1]; r++) {
        const int q = m->perm[s][r];
        float *x = &fld[m->ax[s][q]][m->idx[s][q]];
        const float *c = m->co[s] + 4 * (size_t)q;
        const float f0 = m->f[s][q];
        const float f1 = flush(flush(c[0] * f0) + *x);
        const float v = flush(flush(flush(c[1] * m->v[s][q]) - flush(c[3] * f0)) + flush(c[2] * f1));
        m->f[s][q] = f1, m->v[s][q] = v, *x = v;
    }
}

/* Parts::post_h_slab / post_e_slab (upml.rs:442-470): after the H update
 * the E edges refresh and the H edges complete; after the E update the
 * other way round. */
static void upml_hook(orc_engine *e, int e_half, int p0, int p1) {
    for (int q = 0; q < e->nops; q++) {
        if (e->ops[q].type != OP_UPML) continue;
        upml_refresh(e, e->ops[q].upml, e_half ? 1 : 0, p0, p1);
        upml_apply(e, e->ops[q].upml, e_half ? 0 : 1, p0, p1);
    }
}

static void phase(orc_engine *e, int cmd, int s) {
    int p0 = e->slab_start[s], p1 = e->slab_start[s + 1];
    if (cmd == 1) {
        for (int gi = p0; gi < p1; gi++) {
            double w = h_plane(e, gi, e->energy_flag, 1);
            if (e->energy_flag) e->plane_energy[gi] = w;
        }
        upml_hook(e, 0, p0, p1);
        for (int c = 0; c < e->ncpml; c++)
            if (!e->cpml[c].e_side) cpml_apply(e, &e->cpml[c], p0, p1);
        for (int q = 0; q < e->nops; q++)
            if (e->ops[q].type == OP_POLES) poles_post_h(e, e->ops[q].pm, p0, p1);
    } else if (cmd == 2) {
        for (int gi = p0; gi < p1; gi++) e_plane(e, gi);
        upml_hook(e, 1, p0, p1);
        for (int c = 0; c < e->ncpml; c++)
            if (e->cpml[c].e_side) cpml_apply(e, &e->cpml[c], p0, p1);
        for (int q = 0; q < e->nops; q++)
            if (e->ops[q].type == OP_POLES) poles_post_e(e, e->ops[q].pm, p0, p1);
    } else if (cmd == 3) {
        for (int gi = p0; gi < p1; gi++) e->plane_energy[gi] = h_plane(e, gi, 1, 0);
    } else if (cmd == 4) {
        /* Pool::fused_a (pool.rs:405-473): the H update of the slab's last
         * plane (it reads the next slab's first voltages before that slab
         * updates them) and its CPML H corrections (post_h_planes). */
        const int g = p1 - 1;
        h_plane(e, g, 0, 1);
        for (int c = 0; c < e->ncpml; c++)
            if (!e->cpml[c].e_side) cpml_apply(e, &e->cpml[c], g, g + 1);
    } else if (cmd == 5) {
        /* Pool::fused_b (pool.rs:475-594): plane by plane, H of plane g
         * (all but the last) with its CPML H corrections, then E of plane g
         * with its CPML E corrections. The currents of plane g - 1 are final
         * by then (in this slab, or the previous slab's fused_a). */
        for (int g = p0; g < p1; g++) {
            if (g + 1 < p1) {
                h_plane(e, g, 0, 1);
                for (int c = 0; c < e->ncpml; c++)
                    if (!e->cpml[c].e_side) cpml_apply(e, &e->cpml[c], g, g + 1);
            }
            e_plane(e, g);
            for (int c = 0; c < e->ncpml; c++)
                if (e->cpml[c].e_side) cpml_apply(e, &e->cpml[c], g, g + 1);
        }
    }
}

typedef struct {
    orc_engine *e;
    int s;
} warg_t;

static void *worker(void *arg) {
    warg_t *w = (warg_t *)arg;
    orc_engine *e = w->e;
    int s = w->s;
    free(w);
    for (;;) {
        pthread_barrier_wait(&e->bar);
        int cmd = e->cmd;
        if (cmd == 9) break;
        phase(e, cmd, s);
        pthread_barrier_wait(&e->bar);
    }
    return NULL;
}

static void parallel(orc_engine *e, int cmd) {
    if (e->threads <= 1) {
        phase(e, cmd, 0);
        return;
    }
    e->cmd = cmd;
    pthread_barrier_wait(&e->bar);
    phase(e, cmd, 0);
    pthread_barrier_wait(&e->bar);
}

/* ------------------------------------------------------------------ */
/* serial hooks (engine/mod.rs:833-992 gap schedule; excite.rs,         */
/* probe.rs)                                                            */
/* ------------------------------------------------------------------ */
static float op_amp(const op_t *o, uint64_t n) { return n < o->namp ? o->amp[n] : 0.0f; }

/* EdgeSource::apply (excite.rs:104-115) */
static void source_apply(orc_engine *e, op_t *o, uint64_t n) {
    float a = op_amp(o, n);
    float **f = (o->kind == 2 || o->kind == 3) ? e->i : e->u;
    int hard = o->kind == 1 || o->kind == 3;
    for (int q = 0; q < o->n; q++) {
        float *x = f[o->axis[q]] + o->idx[q];
        if (hard)
            *x = o->wf[q] * a;
        else
            *x += o->wf[q] * a;
    }
}

/* Probe::sample (probe.rs:364-414): f64 sum in entry order. Rust's f64
 * Sum folds from -0.0. */
static void probe_sample(orc_engine *e, op_t *o) {
    float **f = o->quantity == 1 ? e->i : e->u;
    double v = -0.0;
    if (o->f32sum) {
        /* Probe::f32_sum (probe.rs:398-404): (w as f32)*x folded in f32 from 0. */
        float s = 0.0f;
        for (int q = 0; q < o->n; q++) s = s + (float)o->wd[q] * f[o->axis[q]][o->idx[q]];
        v = (double)s;
    } else {
        for (int q = 0; q < o->n; q++) v += o->wd[q] * (double)f[o->axis[q]][o->idx[q]];
    }
    if (o->nvals == o->cap) {
        o->cap = o->cap ? 2 * o->cap : 1024;
        o->vals = (double *)realloc(o->vals, o->cap * sizeof(double));
    }
    o->vals[o->nvals++] = v;
}

/* LumpedElement::pre_e_impl (lumped.rs:1106-1121): u^n of every edge, as
 * f64, for the branch and the current record. pre_e is movable
 * (lumped.rs:1257-1259): it may run anywhere between the previous post_e
 * gap and the E update, where u does not change. */
static void lumped_pre_e(orc_engine *e, op_t *o) {
    lumped_t *L = o->lp;
    if (!L->branch && !L->record) return;
    const float *u = e->u[L->axis];
    for (int q = 0; q < L->n; q++) L->uprev[q] = (double)u[L->idx[q]];
}

static double lumped_value(const lumped_t *L, int t, uint64_t n) {
    return n < L->tlen[t] ? L->tv[t][n] : 0.0;
}

/* LumpedElement::post_e_impl (lumped.rs:1123-1235) without the openEMS
 * order and cylindrical copies, Installed::add_sources (:1240-1249), and
 * LumpedPort::post_e (engine/port.rs:281-297). The source values are
 * Source::value at h_time = (n + 1/2)dt, tabulated by step. */
static void lumped_post_e(orc_engine *e, op_t *o, uint64_t n) {
    lumped_t *L = o->lp;
    float *u = e->u[L->axis];
    const int ne = L->n;
    int active = 0, drive = 0;
    if (L->nterm) {
        for (int q = 0; q < ne; q++) L->v[q] = 0.0;
        for (int t = 0; t < L->nterm; t++) {
            double val = lumped_value(L, t, n);
            if (val != 0.0) {
                active = 1;
                for (int q = 0; q < ne; q++) L->v[q] += L->tw[(size_t)t * ne + q] * val;
            }
        }
    }
    if (L->ndrive) {
        for (int q = 0; q < ne; q++) L->vd[q] = 0.0;
        for (int t = L->nterm; t < L->nterm + L->ndrive; t++) {
            double val = lumped_value(L, t, n);
            if (val != 0.0) {
                drive = 1;
                for (int q = 0; q < ne; q++) L->vd[q] += L->tw[(size_t)t * ne + q] * val;
            }
        }
    }
    if (active && !L->src_in_branch)
        for (int q = 0; q < ne; q++) {
            double add = L->cb[q] * L->gsrc[q] * L->v[q];
            if (add != 0.0) u[L->idx[q]] = (float)((double)u[L->idx[q]] + add);
        }
    if (L->branch) {
        for (int q = 0; q < ne; q++) {
            double v_br = active && L->src_in_branch ? L->v[q] : 0.0;
            if (drive) v_br += L->vd[q];
            double u_star = (double)u[L->idx[q]];
            if (L->branch == 2) {
                /* explicit inductor: dt/L_e = 2/two_l_dt */
                L->j[q] += 2.0 * (L->uprev[q] - v_br) / L->two_l_dt[q];
                L->javg[q] = L->j[q];
                u[L->idx[q]] = (float)(u_star - L->cb[q] * L->j[q]);
                continue;
            }
            double k = L->two_l_dt[q] * L->j[q] - L->q[q] * L->inv_c[q] - v_br + 0.5 * L->uprev[q];
            double z = L->z[q];
            double u_new = (z * u_star - L->cb[q] * k) / (z + 0.5 * L->cb[q]);
            double jb = (k + 0.5 * u_new) / z;
            L->j[q] = 2.0 * jb - L->j[q];
            L->q[q] += e->dt * jb;
            L->javg[q] = jb;
            u[L->idx[q]] = (float)u_new;
        }
    }
    if (L->record) {
        if (L->nrec == L->cap) {
            L->cap = L->cap ? 2 * L->cap : 1024;
            L->rec = (double *)realloc(L->rec, L->cap * sizeof(double));
            L->pv = (double *)realloc(L->pv, L->cap * sizeof(double));
            L->pi = (double *)realloc(L->pi, L->cap * sizeof(double));
        }
        double sum = 0.0;
        for (int q = 0; q < ne; q++) {
            double u0 = L->uprev[q], u1 = (double)u[L->idx[q]];
            double v = active && !L->src_in_branch ? L->v[q] : 0.0;
            sum += L->rg[q] * (0.5 * (u0 + u1) - v) + L->rc[q] * (u1 - u0) / e->dt;
            if (L->branch) sum += L->javg[q];
        }
        double cur = sum / (double)L->series;
        L->rec[L->nrec] = cur;
        if (L->columns) {
            double sv = -0.0; /* Iterator::sum */
            for (int q = 0; q < ne; q++) sv += (double)u[L->idx[q]];
            L->pi[L->nrec] = -cur;
            L->pv[L->nrec] = sv / (double)L->columns;
        }
        L->nrec++;
    }
}

/* LumpedElement::branch_energy (lumped.rs:1075-1087). */
static double lumped_energy(const orc_engine *e, const op_t *o) {
    const lumped_t *L = o->lp;
    if (!L->branch) return 0.0;
    double s = -0.0;
    for (int q = 0; q < L->n; q++) {
        double half_l = 0.25 * L->two_l_dt[q] * e->dt;
        s += half_l * L->j[q] * L->j[q] + 0.5 * L->inv_c[q] * L->q[q] * L->q[q];
    }
    return s;
}

/* interp (planewave.rs:653-655): 4-point Lagrange sum. */
static double pw_interp(const double *v, uint32_t m, const double *w) {
    return w[0] * v[m] + w[1] * v[m + 1] + w[2] * v[m + 2] + w[3] * v[m + 3];
}

/* PlaneWave::post_h, auxiliary-grid branch (planewave.rs:1513-1529): the H
 * corrections from the auxiliary E at n dt, then the auxiliary H update. */
static void pw_post_h(orc_engine *e, pw_t *p) {
    for (int c = 0; c < p->hc.n; c++)
        e->i[p->hc.axis[c]][p->hc.idx[c]] +=
            (float)(p->hc.coef[c] * pw_interp(p->e, p->hc.m[c], p->hc.w + 4 * (size_t)c));
    for (int m = 0; m < p->naux - 1; m++)
        p->h[m] = p->da[m] * p->h[m] - p->db[m] * (p->e[m + 1] - p->e[m]);
}

/* PlaneWave::post_e (planewave.rs:1533-1560): the E corrections from the
 * auxiliary H at (n + 1/2)dt, the auxiliary E update, the hard source. */
static void pw_post_e(orc_engine *e, pw_t *p, uint64_t n) {
    for (int c = 0; c < p->ec.n; c++)
        e->u[p->ec.axis[c]][p->ec.idx[c]] +=
            (float)(p->ec.coef[c] * pw_interp(p->h, p->ec.m[c], p->ec.w + 4 * (size_t)c));
    const int last = p->naux - 1;
    for (int m = 1; m < last; m++)
        p->e[m] = p->ca[m] * p->e[m] - p->cb[m] * (p->h[m] - p->h[m - 1]);
    p->e[0] = n < p->nsrc ? p->src[n] : 0.0;
}

/* MergedFaces::circulations (conformal.rs:2541-2566): per reduced current
 * the weighted sum of its terms' voltages, folding from -0.0. */
static void mf_circulations(const orc_engine *e, merged_t *m) {
    for (int g = 0; g < m->ng; g++) {
        double c = -0.0;
        for (uint32_t t = m->ts[g]; t < m->ts[g + 1]; t++)
            c += m->tw[t] * (double)e->u[m->tax[t]][m->tidx[t]];
        m->circ[g] = c;
    }
}

/* MergedFaces::add_pml (conformal.rs:2582-2638), Cartesian terms: per term
 * in order T = sign (u(hi) - u(p)), psi = b psi + a T (kept if commit), and
 * circ[g] += w (k T + psi) for its targets in order. The terms' targets are
 * held by reduced current; the loop over a current's entries in term order
 * visits them as the term loop does. */
static void mf_add_pml(const orc_engine *e, merged_t *m, int commit) {
    if (!m->np) return;
    double *con = (double *)calloc((size_t)m->np, sizeof(double));
    for (int q = 0; q < m->np; q++) {
        const double *c = m->pc + 4 * (size_t)q;
        const float *u = e->u[m->pax[q]];
        const double d = c[0] * ((double)u[m->pi0[q]] - (double)u[m->pi1[q]]);
        const double psi = c[1] * m->psi[q] + c[2] * d;
        if (commit) m->psi[q] = psi;
        con[q] = c[3] * d + psi;
    }
    for (int g = 0; g < m->ng; g++)
        for (uint32_t k = m->pst[g]; k < m->pst[g + 1]; k++) m->circ[g] += m->pw[k] * con[m->pt[k]];
    free(con);
}

/* MergedFaces::advance (conformal.rs:2569-2576): the merge groups'
 * flush(f32(da i - db c)), then each block (Block::advance :2037-2052):
 * x = B i - dt c, Envelope::solve (:2101-2123), flush(f32(x)). */
static void mf_advance(const merged_t *m, const double *c, double *x, double *out) {
    for (int q = 0; q < m->ns; q++)
        out[q] = (double)flush((float)(m->da[q] * m->i[q] - m->db[q] * c[q]));
    for (int b = 0; b < m->nb; b++) {
        const uint32_t r0 = m->bs[b], nr = m->bs[b + 1] - r0;
        for (uint32_t i = 0; i < nr; i++) {
            double s = -0.0;
            for (uint32_t k = m->brs[r0 + i]; k < m->brs[r0 + i + 1]; k++)
                s += m->bv[k] * m->i[m->mem[r0 + m->bc[k]]];
            x[i] = s - m->dt * c[m->mem[r0 + i]];
        }
        for (uint32_t i = 0; i < nr; i++) {
            const uint32_t f = m->ef[r0 + i];
            const double *r = m->ev + m->eo[r0 + i];
            double s = x[i];
            for (uint32_t j = f; j < i; j++) s -= r[j - f] * x[j];
            x[i] = s / r[i - f];
        }
        for (uint32_t i = nr; i-- > 0;) {
            const uint32_t f = m->ef[r0 + i];
            const double *r = m->ev + m->eo[r0 + i];
            x[i] /= r[i - f];
            const double xi = x[i];
            for (uint32_t j = f; j < i; j++) x[j] -= r[j - f] * xi;
        }
        for (uint32_t i = 0; i < nr; i++) out[m->mem[r0 + i]] = (double)flush((float)x[i]);
    }
}

/* MergedFaces::post_h (conformal.rs:2655-2671): circulations, the update,
 * and every driven face's current sum w i (from -0.0) as f32. */
static void mf_post_h(orc_engine *e, merged_t *m) {
    mf_circulations(e, m);
    mf_add_pml(e, m, 1);
    mf_advance(m, m->circ, m->x, m->next);
    memcpy(m->i, m->next, (size_t)m->ng * sizeof(double));
    for (int q = 0; q < m->nf; q++) {
        double s = -0.0;
        for (uint32_t k = m->fs[q]; k < m->fs[q + 1]; k++) s += m->fw[k] * m->i[m->fg[k]];
        e->i[m->fax[q]][m->fidx[q]] = (float)s;
    }
}

/* Operator::energy of MergedFaces (conformal.rs:2684-2698). */
static double mf_energy(const merged_t *m) {
    double *next = (double *)calloc((size_t)m->ng + 1, sizeof(double));
    double *x = (double *)calloc((size_t)(m->bs[m->nb]) + 1, sizeof(double));
    mf_advance(m, m->circ, x, next);
    double w = 0.0;
    for (int q = 0; q < m->ns; q++) w += 0.5 * m->l[q] * m->i[q] * next[q];
    for (int b = 0; b < m->nb; b++) {
        const uint32_t r0 = m->bs[b];
        for (uint32_t r = r0; r < m->bs[b + 1]; r++) {
            double mn = -0.0;
            for (uint32_t k = m->mrs[r]; k < m->mrs[r + 1]; k++) mn += m->mv[k] * next[m->mem[r0 + m->mc[k]]];
            w += 0.5 * m->i[m->mem[r]] * mn;
        }
    }
    free(next);
    free(x);
    return w;
}

static int is_hsrc(const op_t *o) { return o->type == OP_SOURCE && (o->kind == 2 || o->kind == 3); }
static int is_esrc(const op_t *o) { return o->type == OP_SOURCE && (o->kind == 0 || o->kind == 1); }
static int is_iprobe(const op_t *o) { return o->type == OP_PROBE && o->quantity == 1; }
static int is_vprobe(const op_t *o) { return o->type == OP_PROBE && o->quantity == 0; }

/* hooks_at of each operator (excite.rs:129-140, probe.rs:495-500). */
static int needs(const op_t *o, uint64_t n) {
    if (o->type == OP_SOURCE) {
        if (o->kind == 1 || o->kind == 3) return 1;
        return op_amp(o, n) != 0.0f;
    }
    if (o->type == OP_LUMPED) return 1; /* pre_e / post_e every step */
    if (o->type == OP_POLES) return 0;  /* slab hooks only */
    if (o->type == OP_MUR) return 0;    /* slab hooks only (mur.rs:799-853) */
    if (o->type == OP_UPML) return 0;   /* slab hooks only (upml.rs:442-500) */
    if (o->type == OP_PW || o->type == OP_MERGED) return 1; /* post_h and post_e every step */
    return n % o->interval == 0;
}

/* One step: H phase, mid gap, E phase, post-E gap, with the gap rules of
 * Engine::advance (mod.rs:873-915) for these operator kinds. */
/* mur_1d (mur.rs:313-320) */
static float mur_1d(float i, float k, float x, float b, int oe) {
    return oe ? (i - k * b) + k * x : i + k * (x - b);
}

/* MurSlabs::post_h_slab (mur.rs:799-815): u_b^n, u_1^n of every edge and
 * corner, before the E update. */
static void mur_capture(orc_engine *e, mur_t *m) {
    for (int q = 0; q < m->ne; q++) {
        const float *u = e->u[m->eax[q]];
        m->saved[2 * q] = u[m->enode[q]];
        m->saved[2 * q + 1] = u[m->einner[q]];
    }
    for (int c = 0; c < m->nc; c++) {
        const float *u = e->u[m->cax[c]];
        float *sv = m->saved + 2 * (size_t)m->ne + 3 * (size_t)c;
        sv[0] = u[m->cnode[c]], sv[1] = u[m->cin0[c]], sv[2] = u[m->cin1[c]];
    }
}

/* update (mur.rs:322-355): the face edges, then the corners (whose
 * neighbours are face edges updated here first). */
static void mur_update(orc_engine *e, mur_t *m, uint64_t step) {
    for (int q = 0; q < m->ne; q++) {
        if (step < m->ehold[q]) continue;
        float *u = e->u[m->eax[q]];
        u[m->enode[q]] = mur_1d(m->saved[2 * q + 1], m->ek[q], u[m->einner[q]], m->saved[2 * q],
                                m->oe_order);
    }
    for (int c = 0; c < m->nc; c++) {
        if (step < m->chold[c]) continue;
        float *u = e->u[m->cax[c]];
        const float *sv = m->saved + 2 * (size_t)m->ne + 3 * (size_t)c;
        const float b = sv[0], i0 = sv[1], i1 = sv[2];
        float v;
        if (m->oe_corners) {
            v = i0 - m->ck0[c] * b;
        } else {
            const float v0 = mur_1d(i0, m->ck0[c], u[m->cin0[c]], b, m->oe_order);
            const float v1 = mur_1d(i1, m->ck1[c], u[m->cin1[c]], b, m->oe_order);
            v = 0.5f * (v0 + v1);
        }
        u[m->cnode[c]] = v;
    }
}

static void murs(orc_engine *e, int stage, uint64_t step) {
    for (int q = 0; q < e->nops; q++)
        if (e->ops[q].type == OP_MUR) {
            if (stage == 0) mur_capture(e, e->ops[q].mur);
            else mur_update(e, e->ops[q].mur, step);
        }
}

static void one_step(orc_engine *e, int energy) {
    uint64_t n = e->step;
    int need_post_e = 0, mid = 0, mid_movable = 1;
    for (int q = 0; q < e->nops; q++) {
        op_t *o = &e->ops[q];
        if (!needs(o, n)) continue;
        int own = o->type == OP_PW || o->type == OP_MERGED;
        if (is_esrc(o) || is_vprobe(o) || o->type == OP_LUMPED || own) need_post_e = 1;
        if (is_hsrc(o) || own) mid = 1, mid_movable = 0;
        if (is_iprobe(o) || o->type == OP_LUMPED) mid = 1;
    }
    int gap_b = mid && !mid_movable;
    int moved = mid && mid_movable;
    int gap_a = need_post_e || moved;

    /* The fused schedule (pool.rs:884-940): no energy step, no serial hook
     * between the half-steps and only plane-wise slab hooks (CPML; the
     * dispersive media keep the unfused one here); FusedSteps::Auto when the
     * fields of a slab (24 bytes per node) reach FUSE_MIN_BYTES = 4 MiB. */
    int fused = !energy && !gap_b && e->fuse != 2;
    for (int q = 0; q < e->nops && fused; q++)
        if (e->ops[q].type == OP_POLES || e->ops[q].type == OP_UPML) fused = 0;
    if (fused && e->fuse == 0) fused = 24 * e->N / (size_t)e->threads >= ((size_t)4 << 20);
    if (fused) {
        for (int q = 0; q < e->nops; q++)
            if (e->ops[q].type == OP_LUMPED) lumped_pre_e(e, &e->ops[q]);
        murs(e, 0, n); /* (u is unchanged since the previous post_e) */
        parallel(e, 4);
        parallel(e, 5);
        murs(e, 1, n); /* Fusion::Voltages: after the slab's E hooks */
        goto post_e;
    }
    e->energy_flag = energy;
    parallel(e, 1);
    e->energy_flag = 0;
    murs(e, 0, n);
    if (gap_b) {
        for (int q = 0; q < e->nops; q++) {
            op_t *o = &e->ops[q];
            if (is_hsrc(o)) source_apply(e, o, n);
            if (is_iprobe(o) && n % o->interval == 0) probe_sample(e, o);
            if (o->type == OP_PW) pw_post_h(e, o->pw);
            if (o->type == OP_MERGED) mf_post_h(e, o->mf);
        }
        for (int q = 0; q < e->nops; q++)
            if (e->ops[q].type == OP_LUMPED) lumped_pre_e(e, &e->ops[q]);
    } else {
        /* moved pre_e hooks (mod.rs:966-971): u is unchanged since the
         * previous post_e gap, so they read the same values here. */
        for (int q = 0; q < e->nops; q++)
            if (e->ops[q].type == OP_LUMPED) lumped_pre_e(e, &e->ops[q]);
    }
    parallel(e, 2);
    murs(e, 1, n);
post_e:
    if (gap_a) {
        if (moved)
            for (int q = 0; q < e->nops; q++) {
                op_t *o = &e->ops[q];
                if (is_iprobe(o) && n % o->interval == 0) probe_sample(e, o);
            }
        for (int q = 0; q < e->nops; q++) {
            op_t *o = &e->ops[q];
            if (is_esrc(o)) source_apply(e, o, n);
            if (is_vprobe(o) && n % o->interval == 0) probe_sample(e, o);
            if (o->type == OP_LUMPED) lumped_post_e(e, o, n);
            if (o->type == OP_PW) pw_post_e(e, o->pw, n);
            if (o->type == OP_MERGED) {
                mf_circulations(e, o->mf);
                mf_add_pml(e, o->mf, 0);
            }
        }
    }
    e->step++;
}

/* plane sums in plane order (pool.rs:673-692): Rust's f64 Sum folds from
 * -0.0; the invariant is scaled by dt/4. */
static double plane_sum(orc_engine *e) {
    double s = -0.0;
    for (int gi = 0; gi < e->n[0]; gi++) s += e->plane_energy[gi];
    return 0.25 * e->dt * s;
}

/* ------------------------------------------------------------------ */
/* public API                                                           */
/* ------------------------------------------------------------------ */
orc_engine *orc_create(int nx, int ny, int nz, double dt, int curl_order, int threads) {
    if (nx < 2 || ny < 2 || nz < 2 || !(dt > 0.0)) return NULL;
    orc_engine *e = (orc_engine *)calloc(1, sizeof(orc_engine));
    e->n[0] = nx, e->n[1] = ny, e->n[2] = nz;
    e->plane = (size_t)ny * nz;
    e->N = (size_t)nx * e->plane;
    e->dt = dt;
    e->oe = curl_order == 1;
    for (int a = 0; a < 3; a++) {
        e->u[a] = (float *)calloc(e->N, sizeof(float));
        e->i[a] = (float *)calloc(e->N, sizeof(float));
    }
    e->zero = (float *)calloc(e->plane, sizeof(float));
    e->plane_energy = (double *)calloc(nx, sizeof(double));
    /* Layout::new (fields.rs:38-59) */
    if (threads < 1) threads = 1;
    if (threads > nx) threads = nx;
    e->threads = threads;
    e->slab_start = (int *)calloc(threads + 1, sizeof(int));
    int q = nx / threads, r = nx % threads, p = 0;
    for (int s = 0; s < threads; s++) {
        e->slab_start[s] = p;
        p += q + (s < r ? 1 : 0);
    }
    e->slab_start[threads] = nx;
    if (threads > 1) {
        pthread_barrier_init(&e->bar, NULL, threads);
        e->workers = (pthread_t *)calloc(threads, sizeof(pthread_t));
        for (int s = 1; s < threads; s++) {
            warg_t *w = (warg_t *)malloc(sizeof(warg_t));
            w->e = e, w->s = s;
            pthread_create(&e->workers[s], NULL, worker, w);
        }
    }
    return e;
}

void orc_destroy(orc_engine *e) {
    if (!e) return;
    if (e->threads > 1) {
        e->cmd = 9;
        pthread_barrier_wait(&e->bar);
        for (int s = 1; s < e->threads; s++) pthread_join(e->workers[s], NULL);
        pthread_barrier_destroy(&e->bar);
        free(e->workers);
    }
    for (int a = 0; a < 3; a++) free(e->u[a]), free(e->i[a]);
    for (int l = 0; l < 12; l++) free(e->lanes[l].rows), free(e->lanes[l].index);
    for (int c = 0; c < e->ncpml; c++) {
        cpml_t *s = &e->cpml[c];
        free(s->b), free(s->a), free(s->k), free(s->psi);
    }
    free(e->cpml);
    for (int q = 0; q < e->nops; q++) {
        op_t *o = &e->ops[q];
        free(o->axis), free(o->idx), free(o->wf), free(o->amp), free(o->wd), free(o->vals);
        pw_t *pwp = o->pw;
        if (pwp) {
            pwcorr_t *cc[2] = {&pwp->ec, &pwp->hc};
            for (int t = 0; t < 2; t++)
                free(cc[t]->axis), free(cc[t]->idx), free(cc[t]->coef), free(cc[t]->w), free(cc[t]->m);
            free(pwp->e), free(pwp->h), free(pwp->ca), free(pwp->cb), free(pwp->da), free(pwp->db);
            free(pwp->src), free(pwp);
        }
        poles_t *m = o->pm;
        if (m) {
            for (int f = 0; f < 2; f++) {
                poleside_t *ps = &m->side[f];
                free(ps->axis), free(ps->plane), free(ps->idx), free(ps->siz);
                free(ps->pole_end), free(ps->cidx), free(ps->j), free(ps->q), free(ps->prev);
            }
            free(m->a), free(m->inv_c), free(m->inv_z), free(m->half_l), free(m->kind), free(m);
        }
        mur_t *mu = o->mur;
        if (mu) {
            free(mu->eax), free(mu->cax), free(mu->enode), free(mu->einner), free(mu->cnode);
            free(mu->cin0), free(mu->cin1), free(mu->ek), free(mu->ck0), free(mu->ck1);
            free(mu->ehold), free(mu->chold), free(mu->saved), free(mu);
        }
        upml_t *up = o->upml;
        if (up) {
            for (int t = 0; t < 2; t++)
                free(up->ax[t]), free(up->idx[t]), free(up->co[t]), free(up->f[t]), free(up->v[t]),
                    free(up->perm[t]), free(up->pstart[t]);
            free(up);
        }
        lumped_t *L = o->lp;
        if (L) {
            for (int t = 0; t < L->nterm + L->ndrive; t++) free(L->tv[t]);
            free(L->tv), free(L->tlen), free(L->idx), free(L->cb), free(L->gsrc), free(L->tw);
            free(L->v), free(L->vd), free(L->two_l_dt), free(L->inv_c), free(L->z), free(L->j);
            free(L->q), free(L->javg), free(L->uprev), free(L->rg), free(L->rc), free(L->rec);
            free(L->pv), free(L->pi), free(L);
        }
    }
    free(e->ops);
    free(e->zero), free(e->plane_energy), free(e->slab_start);
    free(e);
}

const char *orc_last_error(orc_engine *e) { return e ? e->err : "null engine"; }

int orc_set_lane(orc_engine *e, int lane, const float *rows, uint64_t nrows,
                 const uint32_t *index) {
    if (lane < 0 || lane >= 12) return -1;
    size_t nr = (size_t)e->n[0] * e->n[1];
    if (!index && nrows != nr) {
        snprintf(e->err, sizeof e->err, "dense lane needs %zu rows", nr);
        return -1;
    }
    lane_t *L = &e->lanes[lane];
    free(L->rows), free(L->index);
    L->rows = (float *)malloc(nrows * e->n[2] * sizeof(float));
    memcpy(L->rows, rows, nrows * e->n[2] * sizeof(float));
    L->nrows = nrows;
    L->index = NULL;
    if (index) {
        L->index = (uint32_t *)malloc(nr * sizeof(uint32_t));
        memcpy(L->index, index, nr * sizeof(uint32_t));
        for (size_t q = 0; q < nr; q++)
            if (index[q] >= nrows) {
                snprintf(e->err, sizeof e->err, "row index out of range");
                return -1;
            }
    }
    return 0;
}

int orc_add_cpml(orc_engine *e, int e_side, int comp, int normal, const int lo[3],
                 const int hi[3], const float *b, const float *a, const float *k) {
    if (comp == normal || comp < 0 || comp > 2 || normal < 0 || normal > 2) return -1;
    for (int d = 0; d < 3; d++)
        if (lo[d] < 0 || hi[d] > e->n[d] || lo[d] > hi[d]) return -1;
    for (int q = 0; q < e->nops; q++)
        if (e->ops[q].type == OP_UPML) return -1;
    e->cpml = (cpml_t *)realloc(e->cpml, (e->ncpml + 1) * sizeof(cpml_t));
    cpml_t *s = &e->cpml[e->ncpml++];
    memset(s, 0, sizeof *s);
    s->e_side = e_side;
    s->comp = comp;
    s->normal = normal;
    /* Slab::new (cpml.rs:415-437) */
    if (normal == axis_next(comp)) {
        s->sign = 1.0f;
        s->other = axis_prev(comp);
    } else {
        s->sign = -1.0f;
        s->other = axis_next(comp);
    }
    memcpy(s->lo, lo, sizeof s->lo);
    memcpy(s->hi, hi, sizeof s->hi);
    int depth = hi[normal] - lo[normal];
    s->b = (float *)malloc((depth + 1) * sizeof(float));
    s->a = (float *)malloc((depth + 1) * sizeof(float));
    s->k = (float *)malloc((depth + 1) * sizeof(float));
    memcpy(s->b, b, depth * sizeof(float));
    memcpy(s->a, a, depth * sizeof(float));
    memcpy(s->k, k, depth * sizeof(float));
    size_t len = (size_t)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    s->psi = (float *)calloc(len ? len : 1, sizeof(float));
    return e->ncpml - 1;
}

static int add_op(orc_engine *e) {
    e->ops = (op_t *)realloc(e->ops, (e->nops + 1) * sizeof(op_t));
    memset(&e->ops[e->nops], 0, sizeof(op_t));
    return e->nops++;
}

static int node_ok(const orc_engine *e, const int *nd) {
    return nd[0] >= 0 && nd[0] < e->n[0] && nd[1] >= 0 && nd[1] < e->n[1] && nd[2] >= 0 &&
           nd[2] < e->n[2];
}

int orc_add_source(orc_engine *e, int kind, int n_edges, const int *axis, const int *nodes,
                   const float *w, const float *amp, uint64_t n_amp) {
    if (kind < 0 || kind > 3) return -1;
    for (int q = 0; q < n_edges; q++)
        if (!node_ok(e, nodes + 3 * q) || axis[q] < 0 || axis[q] > 2) return -1;
    int id = add_op(e);
    op_t *o = &e->ops[id];
    o->type = OP_SOURCE;
    o->kind = kind;
    o->n = n_edges;
    o->axis = (int *)malloc(n_edges * sizeof(int) + 1);
    o->idx = (size_t *)malloc(n_edges * sizeof(size_t) + 1);
    o->wf = (float *)malloc(n_edges * sizeof(float) + 1);
    for (int q = 0; q < n_edges; q++) {
        o->axis[q] = axis[q];
        o->idx[q] = ((size_t)nodes[3 * q] * e->n[1] + nodes[3 * q + 1]) * e->n[2] + nodes[3 * q + 2];
        o->wf[q] = w[q];
    }
    o->amp = (float *)malloc(n_amp * sizeof(float) + 1);
    memcpy(o->amp, amp, n_amp * sizeof(float));
    o->namp = n_amp;
    return id;
}

int orc_add_probe(orc_engine *e, int quantity, int n_entries, const int *axis, const int *nodes,
                  const double *w, uint64_t interval) {
    for (int q = 0; q < n_entries; q++)
        if (!node_ok(e, nodes + 3 * q) || axis[q] < 0 || axis[q] > 2) return -1;
    int id = add_op(e);
    op_t *o = &e->ops[id];
    o->type = OP_PROBE;
    o->f32sum = (quantity & 4) != 0;
    o->quantity = quantity & ~4;
    o->n = n_entries;
    o->interval = interval ? interval : 1;
    o->axis = (int *)malloc(n_entries * sizeof(int) + 1);
    o->idx = (size_t *)malloc(n_entries * sizeof(size_t) + 1);
    o->wd = (double *)malloc(n_entries * sizeof(double) + 1);
    for (int q = 0; q < n_entries; q++) {
        o->axis[q] = axis[q];
        o->idx[q] = ((size_t)nodes[3 * q] * e->n[1] + nodes[3 * q + 1]) * e->n[2] + nodes[3 * q + 2];
        o->wd[q] = w[q];
    }
    return id;
}

static double *dup_d(const double *src, size_t n) {
    double *d = (double *)calloc(n ? n : 1, sizeof(double));
    if (src) memcpy(d, src, n * sizeof(double));
    return d;
}

int orc_add_lumped(orc_engine *e, const orc_lumped_desc *d) {
    if (d->axis < 0 || d->axis > 2 || d->n_edges < 1 || d->series < 1 || !d->cb || !d->g_src)
        return -1;
    if (d->branch < 0 || d->branch > 2 || (d->branch && (!d->two_l_dt || !d->inv_c || !d->z)))
        return -1;
    if (d->port_columns && !d->rec_g) return -1;
    for (int q = 0; q < d->n_edges; q++) {
        const int *nd = d->nodes + 3 * q;
        if (!node_ok(e, nd) || nd[d->axis] + 1 >= e->n[d->axis]) return -1;
    }
    int id = add_op(e);
    op_t *o = &e->ops[id];
    o->type = OP_LUMPED;
    lumped_t *L = (lumped_t *)calloc(1, sizeof(lumped_t));
    o->lp = L;
    const int ne = d->n_edges;
    L->axis = d->axis, L->n = ne, L->series = d->series, L->src_in_branch = d->src_in_branch;
    L->branch = d->branch, L->nterm = d->n_terms, L->ndrive = d->n_drive;
    L->columns = d->port_columns;
    L->idx = (size_t *)malloc(ne * sizeof(size_t));
    for (int q = 0; q < ne; q++) {
        const int *nd = d->nodes + 3 * q;
        L->idx[q] = ((size_t)nd[0] * e->n[1] + nd[1]) * e->n[2] + nd[2];
    }
    L->cb = dup_d(d->cb, ne), L->gsrc = dup_d(d->g_src, ne);
    int nt = d->n_terms + d->n_drive;
    L->tw = (double *)calloc((size_t)(nt ? nt : 1) * ne, sizeof(double));
    if (d->n_terms) memcpy(L->tw, d->term_weight, (size_t)d->n_terms * ne * sizeof(double));
    if (d->n_drive)
        memcpy(L->tw + (size_t)d->n_terms * ne, d->drive_weight,
               (size_t)d->n_drive * ne * sizeof(double));
    L->tv = (double **)calloc(nt ? nt : 1, sizeof(double *));
    L->tlen = (uint64_t *)calloc(nt ? nt : 1, sizeof(uint64_t));
    L->v = dup_d(NULL, ne), L->vd = dup_d(NULL, ne);
    if (d->branch) {
        L->two_l_dt = dup_d(d->two_l_dt, ne), L->inv_c = dup_d(d->inv_c, ne);
        L->z = dup_d(d->z, ne);
    }
    L->j = dup_d(NULL, ne), L->q = dup_d(NULL, ne), L->javg = dup_d(NULL, ne);
    L->uprev = dup_d(NULL, ne);
    L->record = d->rec_g != NULL;
    if (L->record) L->rg = dup_d(d->rec_g, ne), L->rc = dup_d(d->rec_c, ne);
    return id;
}

int orc_lumped_table(orc_engine *e, int id, int term, const double *values, uint64_t count) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_LUMPED) return -1;
    lumped_t *L = e->ops[id].lp;
    if (term < 0 || term >= L->nterm + L->ndrive) return -1;
    free(L->tv[term]);
    L->tv[term] = dup_d(values, count);
    L->tlen[term] = count;
    return 0;
}

uint64_t orc_lumped_values(orc_engine *e, int id, int which, double *out, uint64_t cap) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_LUMPED) return 0;
    lumped_t *L = e->ops[id].lp;
    const double *src = which == 0 ? L->rec : which == 1 ? L->pv : L->pi;
    if (!src || (which && !L->columns)) return 0;
    uint64_t n = L->nrec < cap ? L->nrec : cap;
    if (out) memcpy(out, src, n * sizeof(double));
    return L->nrec;
}

int orc_lumped_branch(orc_engine *e, int id, double *j, double *q, double *javg) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_LUMPED) return -1;
    lumped_t *L = e->ops[id].lp;
    memcpy(j, L->j, L->n * sizeof(double));
    memcpy(q, L->q, L->n * sizeof(double));
    memcpy(javg, L->javg, L->n * sizeof(double));
    return 0;
}

int orc_add_dispersive(orc_engine *e, const orc_pole_side *el, const orc_pole_side *mg,
                       const orc_pole_consts *cs) {
    const orc_pole_side *sd[2] = {el, mg};
    for (int f = 0; f < 2; f++)
        for (int q = 0; q < sd[f]->n_edges; q++)
            if (!node_ok(e, sd[f]->nodes + 3 * q)) return -1;
    int id = add_op(e);
    op_t *o = &e->ops[id];
    o->type = OP_POLES;
    poles_t *m = (poles_t *)calloc(1, sizeof(poles_t));
    o->pm = m;
    m->nc = cs->n;
    m->a = dup_d(cs->a, cs->n), m->inv_c = dup_d(cs->inv_c, cs->n);
    m->inv_z = dup_d(cs->inv_z, cs->n), m->half_l = dup_d(cs->half_l, cs->n);
    m->kind = (int *)calloc(cs->n ? cs->n : 1, sizeof(int));
    memcpy(m->kind, cs->kind, cs->n * sizeof(int));
    for (int f = 0; f < 2; f++) {
        const orc_pole_side *d = sd[f];
        poleside_t *s = &m->side[f];
        const int ne = d->n_edges;
        s->n = ne;
        s->np = d->n_poles;
        s->axis = (int *)calloc(ne ? ne : 1, sizeof(int));
        s->plane = (int *)calloc(ne ? ne : 1, sizeof(int));
        s->idx = (size_t *)calloc(ne ? ne : 1, sizeof(size_t));
        s->siz = dup_d(d->sum_inv_z, ne);
        s->pole_end = (uint32_t *)calloc(ne ? ne : 1, sizeof(uint32_t));
        s->cidx = (uint32_t *)calloc(d->n_poles ? d->n_poles : 1, sizeof(uint32_t));
        s->j = dup_d(NULL, d->n_poles), s->q = dup_d(NULL, d->n_poles);
        s->prev = (float *)calloc(ne ? ne : 1, sizeof(float));
        for (int q = 0; q < ne; q++) {
            const int *nd = d->nodes + 3 * q;
            s->axis[q] = d->axis[q];
            s->plane[q] = nd[0];
            s->idx[q] = ((size_t)nd[0] * e->n[1] + nd[1]) * e->n[2] + nd[2];
            s->pole_end[q] = d->pole_end[q];
        }
        memcpy(s->cidx, d->cidx, d->n_poles * sizeof(uint32_t));
    }
    return id;
}

int orc_dispersive_state(orc_engine *e, int id, int field, double *j, double *q) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_POLES || field < 0 || field > 1)
        return -1;
    const poleside_t *s = &e->ops[id].pm->side[field];
    memcpy(j, s->j, s->np * sizeof(double));
    memcpy(q, s->q, s->np * sizeof(double));
    return 0;
}

static void pwcorr_copy(orc_engine *e, pwcorr_t *d, const orc_pw_corrections *c) {
    d->n = c->n;
    d->axis = (int *)calloc(c->n ? c->n : 1, sizeof(int));
    d->idx = (size_t *)calloc(c->n ? c->n : 1, sizeof(size_t));
    d->coef = dup_d(c->coef, c->n);
    d->w = dup_d(c->w, 4 * (size_t)c->n);
    d->m = (uint32_t *)calloc(c->n ? c->n : 1, sizeof(uint32_t));
    for (int q = 0; q < c->n; q++) {
        const int *nd = c->nodes + 3 * q;
        d->axis[q] = c->axis[q];
        d->idx[q] = ((size_t)nd[0] * e->n[1] + nd[1]) * e->n[2] + nd[2];
        d->m[q] = c->m[q];
    }
}

int orc_add_planewave(orc_engine *e, const orc_pw_desc *d) {
    if (d->n_aux < 3) return -1;
    const orc_pw_corrections *cs[2] = {&d->e_corr, &d->h_corr};
    for (int t = 0; t < 2; t++)
        for (int q = 0; q < cs[t]->n; q++)
            if (!node_ok(e, cs[t]->nodes + 3 * q) || cs[t]->m[q] + 3 >= (uint32_t)d->n_aux - t)
                return -1;
    int id = add_op(e);
    op_t *o = &e->ops[id];
    o->type = OP_PW;
    pw_t *p = (pw_t *)calloc(1, sizeof(pw_t));
    o->pw = p;
    p->naux = d->n_aux;
    p->e = dup_d(NULL, d->n_aux), p->h = dup_d(NULL, d->n_aux - 1);
    p->ca = dup_d(d->ca, d->n_aux), p->cb = dup_d(d->cb, d->n_aux);
    p->da = dup_d(d->da, d->n_aux - 1), p->db = dup_d(d->db, d->n_aux - 1);
    pwcorr_copy(e, &p->ec, &d->e_corr);
    pwcorr_copy(e, &p->hc, &d->h_corr);
    return id;
}

int orc_planewave_table(orc_engine *e, int id, const double *values, uint64_t count) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_PW) return -1;
    pw_t *p = e->ops[id].pw;
    free(p->src);
    p->src = dup_d(values, count);
    p->nsrc = count;
    return 0;
}

int orc_planewave_aux(orc_engine *e, int id, double *ev, double *hv) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_PW) return -1;
    pw_t *p = e->ops[id].pw;
    memcpy(ev, p->e, p->naux * sizeof(double));
    memcpy(hv, p->h, (p->naux - 1) * sizeof(double));
    return 0;
}

static void *dup_n(const void *src, size_t n, size_t sz) {
    void *p = calloc(n + 1, sz);
    if (src && n) memcpy(p, src, n * sz);
    return p;
}

int orc_add_merged(orc_engine *e, const orc_merged_desc *d) {
    if (d->n_groups < 0 || d->n_scalar < 0 || d->n_scalar > d->n_groups || d->n_blocks < 0 ||
        d->n_faces < 0)
        return -1;
    const uint32_t ng = d->n_groups, nb = d->n_blocks, nf = d->n_faces;
    const uint32_t nt = d->term_start[ng], nr = d->blk_start[nb], nw = d->face_start[nf];
    for (uint32_t t = 0; t < nt; t++)
        if (d->term_axis[t] < 0 || d->term_axis[t] > 2 || d->term_idx[t] >= e->N) return -1;
    for (uint32_t q = 0; q < nf; q++)
        if (d->face_axis[q] < 0 || d->face_axis[q] > 2 || d->face_idx[q] >= e->N) return -1;
    for (uint32_t k = 0; k < nw; k++)
        if (d->face_group[k] >= ng) return -1;
    for (uint32_t r = 0; r < nr; r++)
        if (d->members[r] >= ng) return -1;
    int id = add_op(e);
    op_t *o = &e->ops[id];
    o->type = OP_MERGED;
    merged_t *m = (merged_t *)calloc(1, sizeof(merged_t));
    o->mf = m;
    m->ng = ng, m->ns = d->n_scalar, m->nb = nb, m->nf = nf, m->dt = d->dt;
    const uint64_t nev = d->env_off[nr];
    const uint32_t nbv = d->brow_start[nr], nmv = d->mrow_start[nr];
    m->da = dup_n(d->da, ng, 8), m->db = dup_n(d->db, ng, 8), m->l = dup_n(d->l, ng, 8);
    m->ts = dup_n(d->term_start, ng + 1, 4), m->tw = dup_n(d->term_w, nt, 8);
    m->tax = (int *)calloc(nt + 1, sizeof(int)), m->tidx = (size_t *)calloc(nt + 1, sizeof(size_t));
    for (uint32_t t = 0; t < nt; t++) m->tax[t] = d->term_axis[t], m->tidx[t] = d->term_idx[t];
    m->bs = dup_n(d->blk_start, nb + 1, 4), m->mem = dup_n(d->members, nr, 4);
    m->ef = dup_n(d->env_first, nr, 4), m->eo = dup_n(d->env_off, nr + 1, 8);
    m->ev = dup_n(d->env_val, nev, 8);
    m->brs = dup_n(d->brow_start, nr + 1, 4), m->bc = dup_n(d->bcol, nbv, 4);
    m->bv = dup_n(d->bval, nbv, 8);
    m->mrs = dup_n(d->mrow_start, nr + 1, 4), m->mc = dup_n(d->mcol, nmv, 4);
    m->mv = dup_n(d->mval, nmv, 8);
    m->fax = (int *)calloc(nf + 1, sizeof(int)), m->fidx = (size_t *)calloc(nf + 1, sizeof(size_t));
    for (uint32_t q = 0; q < nf; q++) m->fax[q] = d->face_axis[q], m->fidx[q] = d->face_idx[q];
    m->fs = dup_n(d->face_start, nf + 1, 4), m->fg = dup_n(d->face_group, nw, 4);
    m->fw = dup_n(d->face_w, nw, 8);
    const uint32_t npml = d->n_pml < 0 ? 0 : (uint32_t)d->n_pml, npw = d->pml_start[ng];
    for (uint32_t q = 0; q < npml; q++)
        if (d->pml_axis[q] < 0 || d->pml_axis[q] > 2 || d->pml_idx0[q] >= e->N ||
            d->pml_idx1[q] >= e->N)
            return -1;
    m->np = npml;
    m->pax = (int *)calloc(npml + 1, sizeof(int));
    m->pi0 = (size_t *)calloc(npml + 1, sizeof(size_t));
    m->pi1 = (size_t *)calloc(npml + 1, sizeof(size_t));
    for (uint32_t q = 0; q < npml; q++)
        m->pax[q] = d->pml_axis[q], m->pi0[q] = d->pml_idx0[q], m->pi1[q] = d->pml_idx1[q];
    m->pc = dup_n(d->pml_coef, 4 * (size_t)npml, 8), m->psi = dup_n(NULL, npml, 8);
    m->pst = dup_n(d->pml_start, ng + 1, 4), m->pt = dup_n(d->pml_term, npw, 4);
    m->pw = dup_n(d->pml_w, npw, 8);
    m->i = dup_n(NULL, ng, 8), m->circ = dup_n(NULL, ng, 8), m->next = dup_n(NULL, ng, 8);
    m->x = dup_n(NULL, nr, 8);
    return id;
}

int orc_merged_state(orc_engine *e, int id, double *cur, double *circ) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_MERGED) return -1;
    const merged_t *m = e->ops[id].mf;
    if (cur) memcpy(cur, m->i, (size_t)m->ng * sizeof(double));
    if (circ) memcpy(circ, m->circ, (size_t)m->ng * sizeof(double));
    return 0;
}

void orc_run(orc_engine *e, uint64_t n) {
    for (uint64_t t = 0; t < n; t++) one_step(e, 0);
}

/* Engine::operator_energy (mod.rs:753-755): Operator::energy in
 * registration order, Sum folding from -0.0; 0 for every kind but a lumped
 * element's branch. The CPML boundary counts as one operator. */
static double ops_energy(const orc_engine *e) {
    double w = -0.0;
    if (e->ncpml) w += 0.0;
    for (int q = 0; q < e->nops; q++) {
        const op_t *o = &e->ops[q];
        if (o->type == OP_POLES) /* Dispersive::energy (dispersive.rs:2789-2794) */
            w += side_energy(o->pm, 0) + side_energy(o->pm, 1);
        else if (o->type == OP_MERGED)
            w += mf_energy(o->mf);
        else
            w += o->type == OP_LUMPED ? lumped_energy(e, o) : 0.0;
    }
    return w;
}

/* Engine::step_with_energy (mod.rs:819-823): operator energies before the
 * step plus the in-step invariant. */
double orc_step_with_energy(orc_engine *e) {
    double w_ops = ops_energy(e);
    one_step(e, 1);
    return plane_sum(e) + w_ops;
}

/* Engine::energy (mod.rs:722-726): read-only evaluation of the state. */
double orc_energy(orc_engine *e) {
    parallel(e, 3);
    double w_ops = ops_energy(e);
    return plane_sum(e) + w_ops;
}

/* ------------------------------------------------------------------ */
/* openEMS's energy measure (EnergyMeasure::OpenEms, run.rs:66-75)      */
/* ------------------------------------------------------------------ */

/* oe_square (kernel.rs:198-210): x^2 in f64, infinite where its f32 would
 * be (round to nearest: from the midpoint between FLT_MAX and 2^128). */
static inline double oe_square(double x) {
    const double s = x * x;
    return s >= 0x1.ffffffp+127 ? INFINITY : s;
}

/* h_row_energy with OE (kernel.rs:132-173), read-only: the new values
 * squared, the row's two ends weighted by `ends`, 4 accumulators. */
static double hrow_oe(const orc_engine *e, const float *h, const float *da, const float *db,
                      const float *a, const float *b, const float *c, const float *d, int n,
                      double ends) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < n; k++) {
        float nw = e->oe ? oe_h(da[k], h[k], db[k], a[k], b[k], c[k], d[k])
                         : pair_h(da[k], h[k], db[k], a[k], b[k], c[k], d[k]);
        const double w = (k == 0 || k + 1 == n) ? ends : 1.0;
        acc[k & 3] += w * oe_square((double)nw);
    }
    return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

/* The plane value of h_slab with OE on a Cartesian grid (kernel.rs:364-423,
 * 447-511): eps0 sum u^2 + mu0 sum w i'^2; the dual edges of the outer
 * planes normal to their own axis weigh 4, the E of the last node plane
 * along any axis is left out. */
static double plane_openems(const orc_engine *e, int gi) {
    const int nx = e->n[0], ny = e->n[1], nz = e->n[2];
    const size_t p = e->plane, off = (size_t)gi * p;
    const float *ix = e->i[0] + off, *iy = e->i[1] + off, *iz = e->i[2] + off;
    const float *ux = e->u[0] + off, *uy = e->u[1] + off, *uz = e->u[2] + off;
    const int m = nz - 1;
    const double wx = (gi == 0 || gi + 1 == nx) ? 4.0 : 1.0;
    double wh = 0.0;
    for (int j = 0; j < ny - 1; j++) {
        size_t r = (size_t)j * nz;
        wh += wx * hrow_oe(e, ix + r, row(e, 6, gi, j), row(e, 9, gi, j), uz + r + nz, uz + r,
                           uy + r + 1, uy + r, m, 1.0);
    }
    if (gi + 1 < nx) {
        const float *uyn = uy + p, *uzn = uz + p;
        for (int j = 0; j < ny; j++) {
            size_t r = (size_t)j * nz;
            const double wy = (j == 0 || j + 1 == ny) ? 4.0 : 1.0;
            wh += wy * hrow_oe(e, iy + r, row(e, 7, gi, j), row(e, 10, gi, j), ux + r + 1, ux + r,
                               uzn + r, uz + r, m, 1.0);
        }
        for (int j = 0; j < ny - 1; j++) {
            size_t r = (size_t)j * nz;
            wh += hrow_oe(e, iz + r, row(e, 8, gi, j), row(e, 11, gi, j), uyn + r, uy + r,
                          ux + r + nz, ux + r, nz, 4.0);
        }
    }
    double we = 0.0;
    if (gi + 1 < nx) {
        const float *uc[3] = {ux, uy, uz};
        for (int a = 0; a < 3; a++) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            for (int j = 0; j < ny - 1; j++)
                for (int k = 0; k < nz - 1; k++) {
                    const size_t q = (size_t)j * nz + k;
                    acc[q & 3] += oe_square((double)uc[a][q]);
                }
            we += (acc[0] + acc[1]) + (acc[2] + acc[3]);
        }
    }
    return EPS0 * we + MU0 * wh;
}

/* The layer step of an H edge from the kernel's value k (Run::apply and
 * OpenEmsPml::lookahead_current, upml.rs:404-426, 719-731), read-only. */
static float upml_complete(const upml_t *m, int q, float k) {
    const float *c = m->co[1] + 4 * (size_t)q;
    const float f0 = m->f[1][q];
    const float f1 = flush(flush(c[0] * f0) + k);
    return flush(flush(flush(c[1] * m->v[1][q]) - flush(c[3] * f0)) + flush(c[2] * f1));
}

/* Parts::energy_openems_slab (upml.rs:505-537): per H run (component,
 * plane) mu0 sum (v^2 - inc^2) over its edges in order, inc the kernel's
 * flux increment evaluated in the Paired order as written there, v the
 * completed layer value; added to the plane in component order. */
static void upml_energy_openems(const orc_engine *e, const upml_t *m, double *planes) {
    const int nz = e->n[2];
    const size_t p = e->plane;
    for (int gi = 0; gi < e->n[0]; gi++)
        for (int comp = 0; comp < 3; comp++) {
            int any = 0;
            double sum = 0.0;
            for (size_t r = m->pstart[1][gi]; r < m->pstart[1][gi + 1]; r++) {
                const int q = m->perm[1][r];
                if (m->ax[1][q] != comp) continue;
                any = 1;
                const size_t idx = m->idx[1][q], o = idx - (size_t)gi * p;
                const int j = (int)(o / nz), k = (int)(o % nz);
                const float *ux = e->u[0], *uy = e->u[1], *uz = e->u[2];
                float t;
                if (comp == 0)
                    t = (uz[idx + nz] - uz[idx]) - (uy[idx + 1] - uy[idx]);
                else if (comp == 1)
                    t = (ux[idx + 1] - ux[idx]) - (uz[idx + p] - uz[idx]);
                else
                    t = (uy[idx + p] - uy[idx]) - (ux[idx + nz] - ux[idx]);
                const float db = row(e, 9 + comp, gi, j)[k];
                const float inc = flush(0.0f * m->v[1][q] - db * t);
                const float v = upml_complete(m, q, inc);
                sum += (double)v * (double)v - (double)inc * (double)inc;
            }
            if (any) planes[gi] += MU0 * sum;
        }
}

/* The kernel's look-ahead i^{n+1/2} of one dual edge (dump.rs:1830-1845). */
static float lookahead_h(const orc_engine *e, int c, const int p[3]) {
    const int a = axis_next(c), b = axis_prev(c);
    const size_t idx = ((size_t)p[0] * e->n[1] + p[1]) * e->n[2] + p[2];
    int pa[3] = {p[0], p[1], p[2]}, pb[3] = {p[0], p[1], p[2]};
    pa[a]++, pb[b]++;
    const float t0 = pa[a] < e->n[a]
                         ? e->u[b][((size_t)pa[0] * e->n[1] + pa[1]) * e->n[2] + pa[2]] : 0.0f;
    const float t2 = pb[b] < e->n[b]
                         ? e->u[a][((size_t)pb[0] * e->n[1] + pb[1]) * e->n[2] + pb[2]] : 0.0f;
    const float t1 = e->u[b][idx], t3 = e->u[a][idx];
    const float da = row(e, 6 + c, p[0], p[1])[p[2]], db = row(e, 9 + c, p[0], p[1])[p[2]];
    const float h = e->i[c][idx];
    return e->oe ? oe_h(da, h, db, t0, t1, t2, t3)
                 : flush(da * h - db * ((t0 - t1) - (t2 - t3)));
}

typedef struct {
    int axis;
    int node[3];
} hedge_t;

static int hedge_cmp(const void *x, const void *y) {
    const hedge_t *a = (const hedge_t *)x, *b = (const hedge_t *)y;
    if (a->axis != b->axis) return a->axis < b->axis ? -1 : 1;
    for (int d = 0; d < 3; d++)
        if (a->node[d] != b->node[d]) return a->node[d] < b->node[d] ? -1 : 1;
    return 0;
}

/* lookahead_source_energy (dump.rs:1735-1782): mu0 sum w ((k + s)^2 - k^2)
 * over the H sources' edges (sorted by axis and node, without repeats), k
 * the look-ahead (completed by the layers on their edges), k + s with the
 * next step's H sources applied in registration order. */
static double hsource_energy(const orc_engine *e) {
    size_t cnt = 0;
    for (int q = 0; q < e->nops; q++)
        if (is_hsrc(&e->ops[q])) cnt += e->ops[q].n;
    if (!cnt) return 0.0;
    hedge_t *ed = (hedge_t *)malloc(cnt * sizeof(hedge_t));
    size_t w = 0;
    for (int q = 0; q < e->nops; q++) {
        const op_t *o = &e->ops[q];
        if (!is_hsrc(o)) continue;
        for (int r = 0; r < o->n; r++) {
            size_t idx = o->idx[r];
            ed[w].axis = o->axis[r];
            ed[w].node[2] = (int)(idx % e->n[2]);
            idx /= e->n[2];
            ed[w].node[1] = (int)(idx % e->n[1]);
            ed[w].node[0] = (int)(idx / e->n[1]);
            w++;
        }
    }
    qsort(ed, cnt, sizeof(hedge_t), hedge_cmp);
    const upml_t *pm = NULL;
    for (int q = 0; q < e->nops; q++)
        if (e->ops[q].type == OP_UPML) pm = e->ops[q].upml;
    double sum = -0.0;
    for (size_t r = 0; r < cnt; r++) {
        if (r && hedge_cmp(&ed[r], &ed[r - 1]) == 0) continue;
        const int c = ed[r].axis;
        const int *p = ed[r].node;
        const size_t idx = ((size_t)p[0] * e->n[1] + p[1]) * e->n[2] + p[2];
        float k = lookahead_h(e, c, p);
        if (pm)
            for (int q = 0; q < pm->n[1]; q++)
                if (pm->ax[1][q] == c && pm->idx[1][q] == idx) {
                    k = upml_complete(pm, q, k);
                    break;
                }
        float v = k;
        for (int q = 0; q < e->nops; q++) {
            const op_t *o = &e->ops[q];
            if (!is_hsrc(o)) continue;
            const float a = op_amp(o, e->step);
            for (int s = 0; s < o->n; s++)
                if (o->axis[s] == c && o->idx[s] == idx) {
                    const float x = o->wf[s] * a;
                    v = (o->kind == 3) ? x : v + x;
                }
        }
        const double wt = (p[c] == 0 || p[c] + 1 == e->n[c]) ? 4.0 : 1.0;
        const double with = (double)v * (double)v * wt, without = (double)k * (double)k * wt;
        sum += with - without;
    }
    free(ed);
    return MU0 * sum;
}

double orc_energy_openems_ex(orc_engine *e, int layers) {
    const int nx = e->n[0];
    double *planes = (double *)calloc(nx, sizeof(double));
    for (int gi = 0; gi < nx; gi++) planes[gi] = plane_openems(e, gi);
    for (int q = 0; q < e->nops && layers; q++)
        if (e->ops[q].type == OP_UPML) upml_energy_openems(e, e->ops[q].upml, planes);
    double s = -0.0;
    for (int gi = 0; gi < nx; gi++) s += planes[gi];
    free(planes);
    return s + hsource_energy(e);
}

double orc_energy_openems(orc_engine *e) { return orc_energy_openems_ex(e, 1); }

uint64_t orc_timestep(orc_engine *e) { return e->step; }

int orc_add_mur(orc_engine *e, int ne, const int *eax, const int *enode, const int *einner,
                const float *ek, const uint64_t *ehold, int nc, const int *cax, const int *cnode,
                const int *cin0, const int *cin1, const float *ck0, const float *ck1,
                const uint64_t *chold, int oe_order, int oe_corners) {
    if (ne < 0 || nc < 0) return -1;
    for (int q = 0; q < ne; q++)
        if (eax[q] < 0 || eax[q] > 2 || !node_ok(e, enode + 3 * q) || !node_ok(e, einner + 3 * q))
            return -1;
    for (int q = 0; q < nc; q++)
        if (cax[q] < 0 || cax[q] > 2 || !node_ok(e, cnode + 3 * q) || !node_ok(e, cin0 + 3 * q) ||
            !node_ok(e, cin1 + 3 * q))
            return -1;
    int id = add_op(e);
    op_t *o = &e->ops[id];
    o->type = OP_MUR;
    mur_t *m = (mur_t *)calloc(1, sizeof(mur_t));
    o->mur = m;
    m->ne = ne, m->nc = nc, m->oe_order = oe_order, m->oe_corners = oe_corners;
    m->eax = (int *)calloc(ne + 1, sizeof(int));
    m->enode = (size_t *)calloc(ne + 1, sizeof(size_t));
    m->einner = (size_t *)calloc(ne + 1, sizeof(size_t));
    m->ek = (float *)calloc(ne + 1, sizeof(float));
    m->ehold = (uint64_t *)calloc(ne + 1, sizeof(uint64_t));
    m->cax = (int *)calloc(nc + 1, sizeof(int));
    m->cnode = (size_t *)calloc(nc + 1, sizeof(size_t));
    m->cin0 = (size_t *)calloc(nc + 1, sizeof(size_t));
    m->cin1 = (size_t *)calloc(nc + 1, sizeof(size_t));
    m->ck0 = (float *)calloc(nc + 1, sizeof(float));
    m->ck1 = (float *)calloc(nc + 1, sizeof(float));
    m->chold = (uint64_t *)calloc(nc + 1, sizeof(uint64_t));
    m->saved = (float *)calloc(2 * (size_t)ne + 3 * (size_t)nc + 1, sizeof(float));
#define FLAT(p) (((size_t)(p)[0] * e->n[1] + (p)[1]) * e->n[2] + (p)[2])
    for (int q = 0; q < ne; q++) {
        m->eax[q] = eax[q], m->enode[q] = FLAT(enode + 3 * q), m->einner[q] = FLAT(einner + 3 * q);
        m->ek[q] = ek[q], m->ehold[q] = ehold ? ehold[q] : 0;
    }
    for (int q = 0; q < nc; q++) {
        m->cax[q] = cax[q], m->cnode[q] = FLAT(cnode + 3 * q);
        m->cin0[q] = FLAT(cin0 + 3 * q), m->cin1[q] = FLAT(cin1 + 3 * q);
        m->ck0[q] = ck0[q], m->ck1[q] = ck1[q], m->chold[q] = chold ? chold[q] : 0;
    }
#undef FLAT
    return id;
}

int orc_add_upml(orc_engine *e, int ne, const int *eax, const int *enode, const float *eco,
                 int nh, const int *hax, const int *hnode, const float *hco) {
    if (ne < 0 || nh < 0) return -1;
    if (e->ncpml) return -1;
    for (int q = 0; q < e->nops; q++) {
        const int t = e->ops[q].type;
        if (t == OP_UPML || t == OP_POLES || t == OP_MUR || t == OP_PW || t == OP_MERGED) return -1;
    }
    const int cnt[2] = {ne, nh};
    const int *ax[2] = {eax, hax}, *nd[2] = {enode, hnode};
    const float *co[2] = {eco, hco};
    for (int s = 0; s < 2; s++)
        for (int q = 0; q < cnt[s]; q++)
            if (ax[s][q] < 0 || ax[s][q] > 2 || !node_ok(e, nd[s] + 3 * q)) return -1;
    int id = add_op(e);
    op_t *o = &e->ops[id];
    o->type = OP_UPML;
    upml_t *m = (upml_t *)calloc(1, sizeof(upml_t));
    o->upml = m;
    for (int s = 0; s < 2; s++) {
        const int n = cnt[s];
        m->n[s] = n;
        m->ax[s] = (int *)calloc(n + 1, sizeof(int));
        m->idx[s] = (size_t *)calloc(n + 1, sizeof(size_t));
        m->co[s] = (float *)calloc(4 * (size_t)n + 1, sizeof(float));
        m->f[s] = (float *)calloc(n + 1, sizeof(float));
        m->v[s] = (float *)calloc(n + 1, sizeof(float));
        m->perm[s] = (int *)calloc(n + 1, sizeof(int));
        m->pstart[s] = (size_t *)calloc(e->n[0] + 1, sizeof(size_t));
        for (int q = 0; q < n; q++) {
            const int *p = nd[s] + 3 * q;
            m->ax[s][q] = ax[s][q];
            m->idx[s][q] = ((size_t)p[0] * e->n[1] + p[1]) * e->n[2] + p[2];
            memcpy(m->co[s] + 4 * (size_t)q, co[s] + 4 * (size_t)q, 4 * sizeof(float));
            m->pstart[s][p[0] + 1]++;
        }
        for (int g = 0; g < e->n[0]; g++) m->pstart[s][g + 1] += m->pstart[s][g];
        size_t *fill = (size_t *)calloc(e->n[0] + 1, sizeof(size_t));
        memcpy(fill, m->pstart[s], e->n[0] * sizeof(size_t));
        for (int q = 0; q < n; q++) m->perm[s][fill[nd[s][3 * q]]++] = q;
        free(fill);
    }
    return id;
}

int orc_upml_state(orc_engine *e, int id, int side, float *flux, float *value) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_UPML || side < 0 || side > 1) return -1;
    const upml_t *m = e->ops[id].upml;
    memcpy(flux, m->f[side], m->n[side] * sizeof(float));
    memcpy(value, m->v[side], m->n[side] * sizeof(float));
    return 0;
}

void orc_set_fused(orc_engine *e, int mode) { e->fuse = mode < 0 || mode > 2 ? 0 : mode; }

int orc_read(orc_engine *e, int field, int axis, float *out) {
    if (axis < 0 || axis > 2) return -1;
    memcpy(out, (field ? e->i : e->u)[axis], e->N * sizeof(float));
    return 0;
}

int orc_write(orc_engine *e, int field, int axis, const float *in) {
    if (axis < 0 || axis > 2) return -1;
    memcpy((field ? e->i : e->u)[axis], in, e->N * sizeof(float));
    return 0;
}

uint64_t orc_probe_count(orc_engine *e, int id) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_PROBE) return 0;
    return e->ops[id].nvals;
}

uint64_t orc_probe_values(orc_engine *e, int id, double *out, uint64_t cap) {
    if (id < 0 || id >= e->nops || e->ops[id].type != OP_PROBE) return 0;
    op_t *o = &e->ops[id];
    uint64_t n = o->nvals < cap ? o->nvals : cap;
    memcpy(out, o->vals, n * sizeof(double));
    return n;
}

/* ------------------------------------------------------------------ */
/* setup restatements                                                   */
/* ------------------------------------------------------------------ */

/* Grid::delta / dual_delta (grid.rs:562-597) */
static double delta(const double *l, int i) { return l[i + 1] - l[i]; }
static double dual_delta(const double *l, int n, int i) {
    double lo = i > 0 ? delta(l, i - 1) : 0.0;
    double hi = i + 1 < n ? delta(l, i) : 0.0;
    return 0.5 * (lo + hi);
}
static double min_delta(const double *l, int n) {
    double m = INFINITY;
    for (int i = 0; i + 1 < n; i++) {
        double d = l[i + 1] - l[i];
        m = d < m ? d : m; /* f64::min */
    }
    return m;
}

double orc_courant_limit(const double *x, int nx, const double *y, int ny, const double *z,
                         int nz) {
    double d[3] = {min_delta(x, nx), min_delta(y, ny), min_delta(z, nz)};
    /* Iterator::sum over the three terms: folds from -0.0 */
    double s = -0.0;
    for (int a = 0; a < 3; a++) s += 1.0 / (d[a] * d[a]);
    return 1.0 / (C0 * sqrt(s));
}

int orc_coefficients(const double *x, int nx, const double *y, int ny, const double *z,
                     int nz, const double *eps_r, const double *kappa, const double *mu_r,
                     const double *sigma_m, const uint8_t *pec, const int pec_face[6],
                     double dt, float *lanes) {
    const double *L[3] = {x, y, z};
    const int n[3] = {nx, ny, nz};
    const size_t N = (size_t)nx * ny * nz;
    memset(lanes, 0, 12 * N * sizeof(float));
    for (int a = 0; a < 3; a++) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++) {
                    int p[3] = {i, j, k};
                    size_t idx = ((size_t)i * ny + j) * nz + k;
                    size_t ai = (size_t)a * N + idx;
                    if (p[a] + 1 < n[a]) {
                        int on_face = (p[b] == 0 && pec_face[2 * b]) ||
                                      (p[b] == n[b] - 1 && pec_face[2 * b + 1]) ||
                                      (p[c] == 0 && pec_face[2 * c]) ||
                                      (p[c] == n[c] - 1 && pec_face[2 * c + 1]);
                        int pm = pec ? pec[ai] : 0;
                        if (!(on_face || pm)) {
                            /* e_edge_geometry (coeffs.rs:766-774) */
                            double geo = dual_delta(L[b], n[b], p[b]) * dual_delta(L[c], n[c], p[c]) /
                                         delta(L[a], p[a]);
                            double cap = EPS0 * eps_r[ai] * geo;
                            double g = kappa[ai] * geo;
                            /* e_coefficients (coeffs.rs:817-820) */
                            double den = cap + 0.5 * g * dt;
                            lanes[(size_t)a * N + idx] = (float)((cap - 0.5 * g * dt) / den);
                            lanes[(size_t)(3 + a) * N + idx] = (float)(dt / den);
                        }
                    }
                    if (p[b] + 1 < n[b] && p[c] + 1 < n[c]) {
                        double geo = delta(L[b], p[b]) * delta(L[c], p[c]) / dual_delta(L[a], n[a], p[a]);
                        double l = MU0 * mu_r[ai] * geo;
                        double r = sigma_m[ai] * geo;
                        double den = l + 0.5 * r * dt;
                        lanes[(size_t)(6 + a) * N + idx] = (float)((l - 0.5 * r * dt) / den);
                        lanes[(size_t)(9 + a) * N + idx] = (float)(dt / den);
                    }
                }
    }
    return 0;
}

void orc_cpml_coefficients(double m, double sigma_factor, double kappa_max, double alpha_max,
                           int convolution, double rho, double delta_, double dt, double n,
                           double out[7]) {
    double g = pow(rho, m);
    /* Cpml::sigma_max (fdtd.rs:266-271) with eta = n*Z0 */
    double eta = n * Z0;
    double sigma = sigma_factor * 0.8 * (m + 1.0) / (eta * delta_) * g;
    double kappa = 1.0 + (kappa_max - 1.0) * g;
    double alpha = alpha_max * (1.0 - rho);
    double b = exp(-(sigma / kappa + alpha) * dt / EPS0);
    double den = sigma * kappa + kappa * kappa * alpha;
    double x = (sigma / kappa + alpha) * dt / EPS0;
    double a, shift;
    if (!(sigma > 0.0 && den > 0.0)) {
        a = 0.0, shift = 0.0;
    } else if (convolution == 1 && x <= 8.0) {
        double cdt = sigma * dt / (EPS0 * kappa * kappa);
        double gg = x < 1e-3 ? 0.5 - x / 3.0 + x * x / 8.0 : (-expm1(-x) - x * b) / (x * x);
        double a1 = -cdt * gg;
        double a0 = cdt * expm1(-x) / x - a1;
        a = a0 + a1 / b, shift = a1 / b;
    } else {
        a = sigma * (b - 1.0) / den, shift = 0.0;
    }
    out[0] = rho, out[1] = sigma, out[2] = kappa, out[3] = alpha, out[4] = b, out[5] = a;
    out[6] = 1.0 / kappa - 1.0 - shift;
}

/* run::drive (run.rs:424-540) without settle, over a scripted stepper
 * (run.rs:548-582: dt = 1, energy w[min(n, len-1)] at timestep n). */
int orc_drive_script(const double *w, uint64_t nw, uint64_t max, int has_criteria,
                     double energy_ratio, uint64_t interval, int has_excitation_end,
                     double excitation_end, uint64_t *steps_out) {
    uint64_t steps = 0, n = 0;
    double wv = 0.0, w_max = 0.0;
    int reason;
    for (;;) {
        if (steps >= max) {
            reason = 1;
            break;
        }
        uint64_t tn = n;
        wv = w[tn < nw ? tn : nw - 1];
        n += 1;
        steps += 1;
        w_max = fmax(w_max, wv);
        double ratio = w_max > 0.0 ? wv / w_max : 1.0;
        double t = (double)tn * 1.0;
        if (!isfinite(wv) || wv < -1e-12 * w_max) {
            reason = 3;
            break;
        }
        if (has_criteria && (!has_excitation_end || t > excitation_end) && w_max > 0.0 &&
            ratio < energy_ratio) {
            reason = 0;
            break;
        }
        uint64_t chunk = interval - 1 < max - steps ? interval - 1 : max - steps;
        n += chunk;
        steps += chunk;
    }
    *steps_out = steps;
    return reason;
}
