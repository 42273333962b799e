/*
 * o2f_oracle.c -- CPU restatement of the reference solver's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it (as oracle/liboracle.so).  Nothing in the product
 * library links, calls or falls back to it.
 *
 * It restates, line by line in operation order, the algorithm of the
 * reference C++/OpenMP solver (ThibaultMaurelOujia/Optimisation2Forme).
 * Every function cites the reference file:line it follows.  It is compiled
 * with -ffp-contract=off so that no multiply-add is fused, exactly like the
 * HIP product, which lets the parity tests compare bit for bit.
 *
 * Places where the reference's result depends on std::unordered_set
 * iteration order (implementation-defined; libc++ on the author's machine,
 * libstdc++ here) use one fixed canonical order instead; each is a
 * rounding-level choice and is named where it is made:
 *   - a cell's edges are visited in ascending edge id;
 *   - a vertex's incident cells (1-3 cell cases) in ascending cell id;
 *   - equal atan2 angles in the vertex fan are broken by ascending cell id.
 *
 * Parity pins (tests/test_oracle_*.py): closed forms (Euler flux
 * consistency and upwinding of HLLC, exact gradients of linear fields,
 * freestream preservation, SSP-RK3 amplification polynomial, uniform
 * pressure gives C_L=C_D=0), brute force (dense J^T on a tiny mesh vs
 * numpy.linalg.solve; BiCGSTAB vs scipy), transpose identity of the
 * assembled blocks, and the shipped fixtures (quad mesh statistics, the
 * shipped flow state is a near-steady state of the shipped mesh, the sign
 * of the shape gradient against the shipped optimisation trajectory).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* std::min / std::max / std::clamp semantics (NaN handling included). */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double sclamp(double v, double lo, double hi) {
    return (v < lo) ? lo : (hi < v) ? hi : v;
}

typedef struct {
    double gamma, viscosity, Pr, cp, cfl, rho_ref, p_ref, mach;
    double venkat_k; /* k_eps of computeVenkatAlpha, reference 1e-2 */
    int limiter;     /* 0 Venkatakrishnan, 1 Barth-Jespersen */
    int interp_mode; /* 0 reference (valence >= 5 never written), 1 mean-value */
    int bc_known;    /* boundary_conditions == "inflow_outflow_x" */
    int heat_flux;   /* 0: reference (gamma-1) grad e; 1: grad T, T = (gamma-1) e / R (extension) */
} orc_params;

typedef struct {
    int nv, nc, ne;
    /* vertices */
    double *vx, *vy;
    char *vbc;
    int *vptr, *vcell; /* incident cells, ascending id */
    int *vord;         /* incident cells, ascending atan2 angle (same vptr) */
    /* cells */
    int *cnv, *cvtx; /* 4 per cell, -1 pad */
    double *cvol, *ccx, *ccy;
    int *cedge; /* 4 per cell, ascending edge id, -1 pad */
    int *cnbr;  /* 4 per cell, distinct face neighbours, -1 pad */
    /* edges */
    int *eva, *evb, *ecl, *ecr;
    char *ebc;
    double *elen, *enx, *eny, *emx, *emy;
    /* state */
    double *U[4], *Uold[4], *G[8], *R[4];
    double *fL[4], *fR[4];
    double *etau[3]; /* wall tau_xx, tau_xy_yx, tau_yy */
    /* per-call scratch (the reference keeps these in function statics) */
    double *vval[5], *eval_[5], *flux[4], *ctq[5];
} orc_mesh;

/* ------------------------------------------------------------------ */
/* boundary.cpp                                                        */
/* ------------------------------------------------------------------ */

/* src/boundary.cpp:7-45 */
static void bc_inflow_outflow(const orc_params *p, char bc, double *rho, double *ru,
                              double *rv, double *E) {
    double rho0 = p->rho_ref;
    double p0 = p->p_ref;
    double M = p->mach;
    double a0 = sqrt(p->gamma * p0 / rho0);
    double u0 = M * a0;
    double v0 = 0.0;
    double kinetic = 0.5 * rho0 * (u0 * u0 + v0 * v0);
    double E0 = p->p_ref / (p->gamma - 1.0) + kinetic;
    if (bc == 'i' || bc == 'f') {
        *rho = rho0;
        *ru = rho0 * u0;
        *rv = rho0 * v0;
        *E = E0;
    } else if (bc == 'w') {
        const double kinetic_old = 0.5 * (*ru * *ru + *rv * *rv) / *rho;
        const double e_internal = *E - kinetic_old;
        *ru = 0.0;
        *rv = 0.0;
        *E = e_internal;
    }
    /* 'o' and anything else: unchanged */
}

/* src/boundary.cpp:48-80 (the ghost state written into the R side) */
static void bc_interface(const orc_params *p, char bc, double rL, double ruL, double rvL,
                         double EL, double *rR, double *ruR, double *rvR, double *ER) {
    *rR = rL;
    *ruR = ruL;
    *rvR = rvL;
    *ER = EL;
    if (bc == 'i' || bc == 'o' || bc == 'f') {
        bc_inflow_outflow(p, bc, rR, ruR, rvR, ER);
    } else if (bc == 'w') {
        *ruR = -ruL;
        *rvR = -rvL;
    }
}

/* src/boundary.cpp:101-127 */
static void bc_diffusion(char bc, double *txx, double *tyy, double *txy, double *qx,
                         double *qy) {
    if (bc == 'f' || bc == 'w') {
        *txx = 0.0;
        *tyy = 0.0;
        *txy = 0.0;
        *qx = 0.0;
        *qy = 0.0;
    }
}

static inline int is_tagged(char bc) {
    return bc == 'i' || bc == 'o' || bc == 'f' || bc == 'w';
}

/* ------------------------------------------------------------------ */
/* mesh.cpp: topology and geometry                                     */
/* ------------------------------------------------------------------ */

typedef struct {
    int a, b, cell;
    long seq;
} halfedge;

static int cmp_halfedge(const void *x, const void *y) {
    const halfedge *p = (const halfedge *)x, *q = (const halfedge *)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    return p->seq < q->seq ? -1 : (p->seq > q->seq);
}

typedef struct {
    double ang;
    int cell;
} angcell;

static int cmp_angcell(const void *x, const void *y) {
    const angcell *p = (const angcell *)x, *q = (const angcell *)y;
    if (p->ang < q->ang) return -1;
    if (q->ang < p->ang) return 1;
    return p->cell < q->cell ? -1 : (p->cell > q->cell);
}

static int cmp_int(const void *x, const void *y) {
    int a = *(const int *)x, b = *(const int *)y;
    return a < b ? -1 : (a > b);
}

/* Cell area and centre: src/mesh.cpp:123-130 (triangle), :153-163 (quad). */
static void cell_geometry(const orc_mesh *m, int c, double *vol, double *cx, double *cy) {
    const int *v = m->cvtx + 4 * c;
    double x0 = m->vx[v[0]], y0 = m->vy[v[0]], x1 = m->vx[v[1]], y1 = m->vy[v[1]];
    double x2 = m->vx[v[2]], y2 = m->vy[v[2]];
    if (m->cnv[c] == 3) {
        double ax = x1 - x0, ay = y1 - y0, bx = x2 - x0, by = y2 - y0;
        *vol = fabs(0.5 * (ax * by - bx * ay));
        *cx = 1.0 / 3.0 * (x0 + x1 + x2);
        *cy = 1.0 / 3.0 * (y0 + y1 + y2);
    } else {
        double x3 = m->vx[v[3]], y3 = m->vy[v[3]];
        double ax = x1 - x0, ay = y1 - y0, bx = x2 - x0, by = y2 - y0;
        double cx_ = x1 - x2, cy_ = y1 - y2, dx = x3 - x2, dy = y3 - y2;
        *vol = 0.5 * (fabs(ax * by - bx * ay) + fabs(cx_ * dy - dx * cy_));
        *cx = 0.25 * (x0 + x1 + x2 + x3);
        *cy = 0.25 * (y0 + y1 + y2 + y3);
    }
}

/* Edge length, unit normal into the left cell, midpoint:
 * src/mesh.cpp:199-211 (construction), src/shape_optimization.cpp:760-801 (update). */
static void edge_geometry(orc_mesh *m, int e) {
    int a = m->eva[e], b = m->evb[e];
    double tx = m->vx[a] - m->vx[b], ty = m->vy[a] - m->vy[b], tz = 0.0;
    double L = sqrt(tx * tx + ty * ty + tz * tz);
    double nx = ty / L, ny = -tx / L;
    double mx = 0.5 * (m->vx[a] + m->vx[b]), my = 0.5 * (m->vy[a] + m->vy[b]);
    int cl = m->ecl[e];
    double dx = m->ccx[cl] - mx, dy = m->ccy[cl] - my;
    if (dx * nx + dy * ny < 0) {
        nx = -nx;
        ny = -ny;
    }
    m->elen[e] = L;
    m->enx[e] = nx;
    m->eny[e] = ny;
    m->emx[e] = mx;
    m->emy[e] = my;
}

static void vertex_order(orc_mesh *m, int v) {
    int n = m->vptr[v + 1] - m->vptr[v];
    angcell *tmp = (angcell *)malloc(sizeof(angcell) * (n > 0 ? n : 1));
    for (int k = 0; k < n; ++k) {
        int c = m->vcell[m->vptr[v] + k];
        tmp[k].ang = atan2(m->ccy[c] - m->vy[v], m->ccx[c] - m->vx[v]);
        tmp[k].cell = c;
    }
    qsort(tmp, n, sizeof(angcell), cmp_angcell);
    for (int k = 0; k < n; ++k) m->vord[m->vptr[v] + k] = tmp[k].cell;
    free(tmp);
}

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* Mesh::Mesh, src/mesh.cpp:7-297.  Inputs are the arrays the Gmsh API hands
 * over: node coordinates in node order, cells (3 or 4 vertex indices) in
 * element order, boundary line elements with their tag ('i','o','f','w') in
 * physical-group order (the last assignment to a vertex or an edge wins,
 * src/mesh.cpp:76-110). */
orc_mesh *orc_mesh_create(int nv, const double *xy, int nc, const int *cell_nv,
                          const int *cell_vtx, int nb, const int *bnd_vtx,
                          const char *bnd_tag) {
    orc_mesh *m = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    m->nv = nv;
    m->nc = nc;
    m->vx = dalloc(nv);
    m->vy = dalloc(nv);
    m->vbc = (char *)calloc(nv ? nv : 1, 1);
    for (int i = 0; i < nv; ++i) {
        m->vx[i] = xy[2 * i];
        m->vy[i] = xy[2 * i + 1];
    }
    /* boundary tags, last wins */
    halfedge *bl = (halfedge *)malloc(sizeof(halfedge) * (nb ? nb : 1));
    for (int j = 0; j < nb; ++j) {
        int a = bnd_vtx[2 * j], b = bnd_vtx[2 * j + 1];
        m->vbc[a] = bnd_tag[j];
        m->vbc[b] = bnd_tag[j];
        bl[j].a = a < b ? a : b;
        bl[j].b = a < b ? b : a;
        bl[j].cell = (unsigned char)bnd_tag[j];
        bl[j].seq = j;
    }
    qsort(bl, nb, sizeof(halfedge), cmp_halfedge);

    m->cnv = (int *)malloc(sizeof(int) * nc);
    m->cvtx = (int *)malloc(sizeof(int) * 4 * nc);
    m->cvol = dalloc(nc);
    m->ccx = dalloc(nc);
    m->ccy = dalloc(nc);
    long nh = 0;
    for (int c = 0; c < nc; ++c) {
        m->cnv[c] = cell_nv[c];
        for (int k = 0; k < 4; ++k) m->cvtx[4 * c + k] = (k < cell_nv[c]) ? cell_vtx[4 * c + k] : -1;
        nh += cell_nv[c];
    }
    for (int c = 0; c < nc; ++c) cell_geometry(m, c, &m->cvol[c], &m->ccx[c], &m->ccy[c]);

    /* vertex -> cells (ascending cell id) */
    m->vptr = (int *)calloc(nv + 1, sizeof(int));
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < m->cnv[c]; ++k) m->vptr[m->cvtx[4 * c + k] + 1]++;
    for (int i = 0; i < nv; ++i) m->vptr[i + 1] += m->vptr[i];
    m->vcell = (int *)malloc(sizeof(int) * (m->vptr[nv] ? m->vptr[nv] : 1));
    m->vord = (int *)malloc(sizeof(int) * (m->vptr[nv] ? m->vptr[nv] : 1));
    {
        int *fill = (int *)calloc(nv ? nv : 1, sizeof(int));
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < m->cnv[c]; ++k) {
                int v = m->cvtx[4 * c + k];
                m->vcell[m->vptr[v] + fill[v]++] = c;
            }
        free(fill);
    }

    /* edge map (std::map keyed by the sorted vertex pair): recordEdge,
     * src/mesh.cpp:288-297 -- first recorder is left, the last later one right */
    halfedge *he = (halfedge *)malloc(sizeof(halfedge) * (nh ? nh : 1));
    long h = 0;
    for (int c = 0; c < nc; ++c) {
        int n = m->cnv[c];
        for (int k = 0; k < n; ++k) {
            int a = m->cvtx[4 * c + k], b = m->cvtx[4 * c + (k + 1) % n];
            he[h].a = a < b ? a : b;
            he[h].b = a < b ? b : a;
            he[h].cell = c;
            he[h].seq = h;
            ++h;
        }
    }
    qsort(he, nh, sizeof(halfedge), cmp_halfedge);
    int ne = 0;
    for (long i = 0; i < nh; ++i)
        if (i == 0 || he[i].a != he[i - 1].a || he[i].b != he[i - 1].b) ++ne;
    m->ne = ne;
    m->eva = (int *)malloc(sizeof(int) * ne);
    m->evb = (int *)malloc(sizeof(int) * ne);
    m->ecl = (int *)malloc(sizeof(int) * ne);
    m->ecr = (int *)malloc(sizeof(int) * ne);
    m->ebc = (char *)calloc(ne, 1);
    m->elen = dalloc(ne);
    m->enx = dalloc(ne);
    m->eny = dalloc(ne);
    m->emx = dalloc(ne);
    m->emy = dalloc(ne);
    m->cedge = (int *)malloc(sizeof(int) * 4 * nc);
    m->cnbr = (int *)malloc(sizeof(int) * 4 * nc);
    for (int i = 0; i < 4 * nc; ++i) m->cedge[i] = m->cnbr[i] = -1;
    int *cfill = (int *)calloc(nc ? nc : 1, sizeof(int));
    int *nfill = (int *)calloc(nc ? nc : 1, sizeof(int));
    long bi = 0;
    int e = -1;
    for (long i = 0; i < nh; ++i) {
        if (i == 0 || he[i].a != he[i - 1].a || he[i].b != he[i - 1].b) {
            ++e;
            m->eva[e] = he[i].a;
            m->evb[e] = he[i].b;
            m->ecl[e] = he[i].cell;
            m->ecr[e] = -1;
        } else {
            m->ecr[e] = he[i].cell;
        }
    }
    for (e = 0; e < ne; ++e) {
        /* boundary tag lookup (last line with this key wins) */
        while (bi < nb && (bl[bi].a < m->eva[e] || (bl[bi].a == m->eva[e] && bl[bi].b < m->evb[e]))) ++bi;
        long bj = bi;
        while (bj < nb && bl[bj].a == m->eva[e] && bl[bj].b == m->evb[e]) {
            m->ebc[e] = (char)bl[bj].cell;
            ++bj;
        }
        edge_geometry(m, e);
        int cl = m->ecl[e], cr = m->ecr[e];
        m->cedge[4 * cl + cfill[cl]++] = e;
        if (cr >= 0) {
            m->cedge[4 * cr + cfill[cr]++] = e;
            int dup = 0;
            for (int k = 0; k < nfill[cl]; ++k) dup |= m->cnbr[4 * cl + k] == cr;
            if (!dup) m->cnbr[4 * cl + nfill[cl]++] = cr;
            dup = 0;
            for (int k = 0; k < nfill[cr]; ++k) dup |= m->cnbr[4 * cr + k] == cl;
            if (!dup) m->cnbr[4 * cr + nfill[cr]++] = cl;
        }
    }
    /* edges arrive in ascending id already; keep the canonical order explicit */
    for (int c = 0; c < nc; ++c) qsort(m->cedge + 4 * c, cfill[c], sizeof(int), cmp_int);
    free(cfill);
    free(nfill);
    free(he);
    free(bl);

    for (int v = 0; v < nv; ++v) vertex_order(m, v);

    for (int q = 0; q < 4; ++q) {
        m->U[q] = dalloc(nc);
        m->Uold[q] = dalloc(nc);
        m->R[q] = dalloc(nc);
        m->fL[q] = dalloc(ne);
        m->fR[q] = dalloc(ne);
        m->flux[q] = dalloc(ne);
    }
    for (int q = 0; q < 8; ++q) m->G[q] = dalloc(nc);
    for (int q = 0; q < 3; ++q) m->etau[q] = dalloc(ne);
    for (int q = 0; q < 5; ++q) {
        m->vval[q] = dalloc(nv);
        m->eval_[q] = dalloc(ne);
        m->ctq[q] = dalloc(nc);
    }
    return m;
}

void orc_mesh_destroy(orc_mesh *m) {
    if (!m) return;
    void *ptrs[] = {m->vx, m->vy, m->vbc, m->vptr, m->vcell, m->vord, m->cnv, m->cvtx,
                    m->cvol, m->ccx, m->ccy, m->cedge, m->cnbr, m->eva, m->evb, m->ecl,
                    m->ecr, m->ebc, m->elen, m->enx, m->eny, m->emx, m->emy};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    for (int q = 0; q < 4; ++q) {
        free(m->U[q]);
        free(m->Uold[q]);
        free(m->R[q]);
        free(m->fL[q]);
        free(m->fR[q]);
        free(m->flux[q]);
    }
    for (int q = 0; q < 8; ++q) free(m->G[q]);
    for (int q = 0; q < 3; ++q) free(m->etau[q]);
    for (int q = 0; q < 5; ++q) {
        free(m->vval[q]);
        free(m->eval_[q]);
        free(m->ctq[q]);
    }
    free(m);
}

static void copy_d(double **dst, double *src, size_t n) {
    *dst = dalloc(n);
    memcpy(*dst, src, n * sizeof(double));
}
static void copy_i(int **dst, int *src, size_t n) {
    *dst = (int *)malloc((n ? n : 1) * sizeof(int));
    memcpy(*dst, src, n * sizeof(int));
}
static void copy_c(char **dst, char *src, size_t n) {
    *dst = (char *)malloc(n ? n : 1);
    memcpy(*dst, src, n);
}

/* Deep copy, as `Mesh mesh_plus = mesh;` in src/shape_optimization.cpp:989. */
orc_mesh *orc_mesh_clone(const orc_mesh *s) {
    orc_mesh *m = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    int nv = s->nv, nc = s->nc, ne = s->ne, nvc = s->vptr[nv];
    m->nv = nv;
    m->nc = nc;
    m->ne = ne;
    copy_d(&m->vx, s->vx, nv);
    copy_d(&m->vy, s->vy, nv);
    copy_c(&m->vbc, s->vbc, nv);
    copy_i(&m->vptr, s->vptr, nv + 1);
    copy_i(&m->vcell, s->vcell, nvc);
    copy_i(&m->vord, s->vord, nvc);
    copy_i(&m->cnv, s->cnv, nc);
    copy_i(&m->cvtx, s->cvtx, 4 * nc);
    copy_d(&m->cvol, s->cvol, nc);
    copy_d(&m->ccx, s->ccx, nc);
    copy_d(&m->ccy, s->ccy, nc);
    copy_i(&m->cedge, s->cedge, 4 * nc);
    copy_i(&m->cnbr, s->cnbr, 4 * nc);
    copy_i(&m->eva, s->eva, ne);
    copy_i(&m->evb, s->evb, ne);
    copy_i(&m->ecl, s->ecl, ne);
    copy_i(&m->ecr, s->ecr, ne);
    copy_c(&m->ebc, s->ebc, ne);
    copy_d(&m->elen, s->elen, ne);
    copy_d(&m->enx, s->enx, ne);
    copy_d(&m->eny, s->eny, ne);
    copy_d(&m->emx, s->emx, ne);
    copy_d(&m->emy, s->emy, ne);
    for (int q = 0; q < 4; ++q) {
        copy_d(&m->U[q], s->U[q], nc);
        copy_d(&m->Uold[q], s->Uold[q], nc);
        copy_d(&m->R[q], s->R[q], nc);
        copy_d(&m->fL[q], s->fL[q], ne);
        copy_d(&m->fR[q], s->fR[q], ne);
        copy_d(&m->flux[q], s->flux[q], ne);
    }
    for (int q = 0; q < 8; ++q) copy_d(&m->G[q], s->G[q], nc);
    for (int q = 0; q < 3; ++q) copy_d(&m->etau[q], s->etau[q], ne);
    for (int q = 0; q < 5; ++q) {
        copy_d(&m->vval[q], s->vval[q], nv);
        copy_d(&m->eval_[q], s->eval_[q], ne);
        copy_d(&m->ctq[q], s->ctq[q], nc);
    }
    return m;
}

void orc_counts(const orc_mesh *m, int *nv, int *ne, int *nc) {
    *nv = m->nv;
    *ne = m->ne;
    *nc = m->nc;
}

/* Topology / geometry export for the tests (reference order). */
void orc_get_edges(const orc_mesh *m, int *va, int *vb, int *cl, int *cr, char *bc,
                   double *geo /* 5*ne: len nx ny mx my */) {
    for (int e = 0; e < m->ne; ++e) {
        va[e] = m->eva[e];
        vb[e] = m->evb[e];
        cl[e] = m->ecl[e];
        cr[e] = m->ecr[e];
        bc[e] = m->ebc[e];
        geo[e] = m->elen[e];
        geo[m->ne + e] = m->enx[e];
        geo[2 * m->ne + e] = m->eny[e];
        geo[3 * m->ne + e] = m->emx[e];
        geo[4 * m->ne + e] = m->emy[e];
    }
}

void orc_get_cells(const orc_mesh *m, double *geo /* 3*nc: vol cx cy */, int *cedge) {
    for (int c = 0; c < m->nc; ++c) {
        geo[c] = m->cvol[c];
        geo[m->nc + c] = m->ccx[c];
        geo[2 * m->nc + c] = m->ccy[c];
        for (int k = 0; k < 4; ++k) cedge[4 * c + k] = m->cedge[4 * c + k];
    }
}

void orc_get_vertex_bc(const orc_mesh *m, char *bc, double *xy) {
    for (int v = 0; v < m->nv; ++v) {
        bc[v] = m->vbc[v];
        xy[2 * v] = m->vx[v];
        xy[2 * v + 1] = m->vy[v];
    }
}

/* ------------------------------------------------------------------ */
/* state I/O, initial condition                                        */
/* ------------------------------------------------------------------ */

void orc_set_state(orc_mesh *m, const double *U) {
    for (int q = 0; q < 4; ++q) memcpy(m->U[q], U + (size_t)q * m->nc, sizeof(double) * m->nc);
}
void orc_get_state(const orc_mesh *m, double *U) {
    for (int q = 0; q < 4; ++q) memcpy(U + (size_t)q * m->nc, m->U[q], sizeof(double) * m->nc);
}
void orc_get_residual(const orc_mesh *m, double *R) {
    for (int q = 0; q < 4; ++q) memcpy(R + (size_t)q * m->nc, m->R[q], sizeof(double) * m->nc);
}
void orc_get_gradients(const orc_mesh *m, double *G) {
    for (int q = 0; q < 8; ++q) memcpy(G + (size_t)q * m->nc, m->G[q], sizeof(double) * m->nc);
}
void orc_get_faces(const orc_mesh *m, double *L, double *R) {
    for (int q = 0; q < 4; ++q) {
        memcpy(L + (size_t)q * m->ne, m->fL[q], sizeof(double) * m->ne);
        memcpy(R + (size_t)q * m->ne, m->fR[q], sizeof(double) * m->ne);
    }
}
void orc_set_faces(orc_mesh *m, const double *L, const double *R) {
    for (int q = 0; q < 4; ++q) {
        memcpy(m->fL[q], L + (size_t)q * m->ne, sizeof(double) * m->ne);
        memcpy(m->fR[q], R + (size_t)q * m->ne, sizeof(double) * m->ne);
    }
}

/* init_one_x + primitives_to_conserved, src/init.cpp:8-16, :59-81 */
void orc_init_freestream(orc_mesh *m, const orc_params *p) {
    const double gamma = p->gamma;
    double rho0 = p->rho_ref, p0 = p->p_ref;
    const double a0 = sqrt(gamma * p0 / rho0);
    double u0 = p->mach * a0, v0 = 0.0;
    double ru = rho0 * u0, rv = rho0 * v0;
    double kinetic = 0.5 * rho0 * (u0 * u0 + v0 * v0);
    double E = p0 / (gamma - 1.0) + kinetic;
    for (int c = 0; c < m->nc; ++c) {
        m->U[0][c] = rho0;
        m->U[1][c] = ru;
        m->U[2][c] = rv;
        m->U[3][c] = E;
    }
}

/* ------------------------------------------------------------------ */
/* advection.cpp                                                       */
/* ------------------------------------------------------------------ */

/* Cell -> vertex interpolation of `nvar` cell arrays into out[q][v]:
 * src/advection.cpp:26-128 (same weights in src/diffusion.cpp:28-136).
 * Valence >= 5 is not handled by the reference (no branch writes the value);
 * interp_mode 0 reproduces that (value stays 0), mode 1 uses the n = 4
 * mean-value formula generalised to n angularly ordered cells. */
static void interp_vertex(const orc_mesh *m, const orc_params *p, int v, int nvar,
                          double *const *src, double *const *out) {
    int base = m->vptr[v];
    int n = m->vptr[v + 1] - base;
    double px = m->vx[v], py = m->vy[v];
    if (n == 3) {
        int id0 = m->vcell[base], id1 = m->vcell[base + 1], id2 = m->vcell[base + 2];
        double c0x = m->ccx[id0], c0y = m->ccy[id0], c1x = m->ccx[id1], c1y = m->ccy[id1];
        double c2x = m->ccx[id2], c2y = m->ccy[id2];
        double w0_num = (c1x - px) * (c2y - py) - (c2x - px) * (c1y - py);
        double w0_den = (c1x - c0x) * (c2y - c0y) - (c2x - c0x) * (c1y - c0y);
        double w0 = (w0_num / w0_den);
        double w1_num = (c2x - px) * (c0y - py) - (c0x - px) * (c2y - py);
        double w1_den = (c1x - c0x) * (c2y - c0y) - (c2x - c0x) * (c1y - c0y);
        double w1 = (w1_num / w1_den);
        double w2 = 1 - w0 - w1;
        for (int q = 0; q < nvar; ++q)
            out[q][v] = w0 * src[q][id0] + w1 * src[q][id1] + w2 * src[q][id2];
    } else if (n == 4 || (n >= 5 && p->interp_mode == 1)) {
        double wst[16], dst[16], ux[16], uy[16];
        double *w = n <= 16 ? wst : dalloc(n);
        double *d = n <= 16 ? dst : dalloc(n);
        double *Ux = n <= 16 ? ux : dalloc(n);
        double *Uy = n <= 16 ? uy : dalloc(n);
        const int *ord = m->vord + base;
        for (int k = 0; k < n; ++k) {
            double dx = m->ccx[ord[k]] - px;
            double dy = m->ccy[ord[k]] - py;
            d[k] = hypot(dx, dy);
            Ux[k] = dx / d[k];
            Uy[k] = dy / d[k];
        }
        for (int k = 0; k < n; ++k) {
            int km1 = (k + n - 1) % n, kp1 = (k + 1) % n;
            double cos_prev = sclamp(Ux[km1] * Ux[k] + Uy[km1] * Uy[k], -1.0, 1.0);
            double cos_next = sclamp(Ux[k] * Ux[kp1] + Uy[k] * Uy[kp1], -1.0, 1.0);
            double theta_prev = acos(cos_prev);
            double theta_next = acos(cos_next);
            w[k] = (tan(0.5 * theta_prev) + tan(0.5 * theta_next)) / d[k];
        }
        double wsum = 0.0;
        for (int k = 0; k < n; ++k) wsum = wsum + w[k];
        for (int k = 0; k < n; ++k) w[k] /= wsum;
        for (int q = 0; q < nvar; ++q) {
            double s = w[0] * src[q][ord[0]];
            for (int k = 1; k < n; ++k) s = s + w[k] * src[q][ord[k]];
            out[q][v] = s;
        }
        if (n > 16) {
            free(w);
            free(d);
            free(Ux);
            free(Uy);
        }
    } else if (n == 2) {
        int id0 = m->vcell[base], id1 = m->vcell[base + 1];
        for (int q = 0; q < nvar; ++q) out[q][v] = 0.5 * (src[q][id0] + src[q][id1]);
    } else if (n == 1) {
        int id0 = m->vcell[base];
        for (int q = 0; q < nvar; ++q) out[q][v] = src[q][id0];
    } else {
        /* valence 0 or (reference mode) >= 5: the reference never writes the
         * entry; its zero-initialised buffer keeps 0 before the BC override */
        for (int q = 0; q < nvar; ++q) out[q][v] = 0.0;
    }
}

/* computeScalarGradientAdvection, src/advection.cpp:6-191 */
static void gradient_advection(orc_mesh *m, const orc_params *p) {
    double *vv[4] = {m->vval[0], m->vval[1], m->vval[2], m->vval[3]};
    double *ev[4] = {m->eval_[0], m->eval_[1], m->eval_[2], m->eval_[3]};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m->nv; ++i) {
        interp_vertex(m, p, i, 4, m->U, vv);
        if (is_tagged(m->vbc[i]))
            bc_inflow_outflow(p, m->vbc[i], &vv[0][i], &vv[1][i], &vv[2][i], &vv[3][i]);
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < m->ne; ++e) {
        int a = m->eva[e], b = m->evb[e];
        double rho = 0.5 * (vv[0][a] + vv[0][b]);
        double ru = 0.5 * (vv[1][a] + vv[1][b]);
        double rv = 0.5 * (vv[2][a] + vv[2][b]);
        double E = 0.5 * (vv[3][a] + vv[3][b]);
        if (is_tagged(m->ebc[e])) bc_inflow_outflow(p, m->ebc[e], &rho, &ru, &rv, &E);
        ev[0][e] = rho;
        ev[1][e] = ru;
        ev[2][e] = rv;
        ev[3][e] = E;
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < m->nc; ++c) {
        double gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0};
        double cv = m->cvol[c];
        for (int k = 0; k < 4; ++k) {
            int e = m->cedge[4 * c + k];
            if (e < 0) break;
            double el = m->elen[e], enx = m->enx[e], eny = m->eny[e];
            if (m->ecl[e] == c) {
                enx = -enx;
                eny = -eny;
            }
            for (int q = 0; q < 4; ++q) {
                gx[q] += ev[q][e] * el * enx / cv;
                gy[q] += ev[q][e] * el * eny / cv;
            }
        }
        for (int q = 0; q < 4; ++q) {
            m->G[q][c] = gx[q];
            m->G[4 + q][c] = gy[q];
        }
    }
}

/* computeBarthJespersenAlpha, src/advection.cpp:195-207 */
static inline double alpha_bj(double phi_i, double phi_pred, double phi_min, double phi_max) {
    double alpha = 1.0;
    if (fabs(phi_pred - phi_i) < DBL_EPSILON) return alpha;
    if (phi_pred > phi_i)
        alpha = smin(1.0, (phi_max - phi_i) / (phi_pred - phi_i));
    else if (phi_pred < phi_i)
        alpha = smin(1.0, (phi_min - phi_i) / (phi_pred - phi_i));
    return alpha;
}

/* computeVenkatAlpha, src/advection.cpp:209-230 */
static inline double alpha_venkat(double phi_i, double phi_pred, double phi_min,
                                  double phi_max, double cellArea, double k_eps) {
    const double delta = phi_pred - phi_i;
    if (fabs(delta) < DBL_EPSILON) return 1.0;
    const double h = sqrt(cellArea / M_PI);
    const double eps = k_eps * k_eps * k_eps * h * h * h;
    double a = (delta > 0.0) ? (phi_max - phi_i) : (phi_min - phi_i);
    double alpha = (a * a + eps + 2.0 * a * delta) / (a * a + eps + 2.0 * delta * delta);
    return sclamp(alpha, 0.0, 1.0);
}

/* computeInterfaceStates, src/advection.cpp:232-321 */
static void interface_states(orc_mesh *m, const orc_params *p) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < m->nc; ++c) {
        double u[4], mn[4], mx[4], gx[4], gy[4], al[4];
        for (int q = 0; q < 4; ++q) {
            u[q] = mn[q] = mx[q] = m->U[q][c];
            gx[q] = m->G[q][c];
            gy[q] = m->G[4 + q][c];
            al[q] = 1.0;
        }
        for (int k = 0; k < 4; ++k) {
            int nb = m->cnbr[4 * c + k];
            if (nb < 0) break;
            for (int q = 0; q < 4; ++q) {
                double v = m->U[q][nb];
                if (v < mn[q]) mn[q] = v;
                if (v > mx[q]) mx[q] = v;
            }
        }
        double cx = m->ccx[c], cy = m->ccy[c], vol = m->cvol[c];
        for (int k = 0; k < 4; ++k) {
            int e = m->cedge[4 * c + k];
            if (e < 0) break;
            double dx = m->emx[e] - cx, dy = m->emy[e] - cy;
            for (int q = 0; q < 4; ++q) {
                double phi_pred = u[q] + gx[q] * dx + gy[q] * dy;
                double a = (p->limiter == 1) ? alpha_bj(u[q], phi_pred, mn[q], mx[q])
                                             : alpha_venkat(u[q], phi_pred, mn[q], mx[q], vol, p->venkat_k);
                al[q] = smin(al[q], a);
            }
        }
        for (int k = 0; k < 4; ++k) {
            int e = m->cedge[4 * c + k];
            if (e < 0) break;
            double dx = m->emx[e] - cx, dy = m->emy[e] - cy;
            double f[4];
            for (int q = 0; q < 4; ++q) f[q] = u[q] + al[q] * (gx[q] * dx + gy[q] * dy);
            if (m->ecl[e] == c) {
                for (int q = 0; q < 4; ++q) m->fL[q][e] = f[q];
                if (is_tagged(m->ebc[e]))
                    bc_interface(p, m->ebc[e], f[0], f[1], f[2], f[3], &m->fR[0][e], &m->fR[1][e],
                                 &m->fR[2][e], &m->fR[3][e]);
            } else if (m->ecr[e] == c) {
                for (int q = 0; q < 4; ++q) m->fR[q][e] = f[q];
            }
        }
    }
}

/* compute_hllc_flux, src/advection.cpp:325-412.  (enx, eny) is the stored
 * edge normal (pointing into the left cell); the flux is along -n. */
void orc_hllc(const double *qL, const double *qR, double enx, double eny, double gamma,
              double *F) {
    double nx = -enx, ny = -eny;
    double rho_L = qL[0];
    double u_L = (qL[1] * (nx) + qL[2] * ny) / rho_L;
    double v_L = (qL[1] * (-ny) + qL[2] * nx) / rho_L;
    double E_L = qL[3];
    double kin_L = 0.5 * rho_L * (u_L * u_L + v_L * v_L);
    double p_L = (gamma - 1.0) * (E_L - kin_L);
    double c_L = sqrt(gamma * p_L / rho_L);

    double rho_R = qR[0];
    double u_R = (qR[1] * (nx) + qR[2] * ny) / rho_R;
    double v_R = (qR[1] * (-ny) + qR[2] * nx) / rho_R;
    double E_R = qR[3];
    double kin_R = 0.5 * rho_R * (u_R * u_R + v_R * v_R);
    double p_R = (gamma - 1.0) * (E_R - kin_R);
    double c_R = sqrt(gamma * p_R / rho_R);

    double S_L = smin(u_L - c_L, u_R - c_R);
    double S_R = smax(u_L + c_L, u_R + c_R);
    double numer = p_R - p_L + rho_L * u_L * (S_L - u_L) - rho_R * u_R * (S_R - u_R);
    double denom = rho_L * (S_L - u_L) - rho_R * (S_R - u_R);
    double S_star = numer / denom;
    double rho_L_star = rho_L * (S_L - u_L) / (S_L - S_star);
    double rho_R_star = rho_R * (S_R - u_R) / (S_R - S_star);
    double rho_v_L_star = rho_L_star * v_L;
    double rho_v_R_star = rho_R_star * v_R;
    double p_star = p_L + rho_L * (S_L - u_L) * (S_star - u_L);
    double E_L_star = ((S_L - u_L) * E_L - p_L * u_L + p_star * S_star) / (S_L - S_star);
    double E_R_star = ((S_R - u_R) * E_R - p_R * u_R + p_star * S_star) / (S_R - S_star);
    double rho_u_L_star = rho_L_star * S_star;
    double rho_u_R_star = rho_R_star * S_star;
    double f0, fu, fv, fE;
    if (S_L >= 0.0) {
        f0 = rho_L * u_L;
        fu = rho_L * u_L * u_L + p_L;
        fv = rho_L * u_L * v_L;
        fE = (E_L + p_L) * u_L;
    } else if (S_L <= 0.0 && S_star >= 0.0) {
        f0 = rho_u_L_star;
        fu = rho_u_L_star * rho_u_L_star / rho_L_star + p_star;
        fv = rho_v_L_star * rho_u_L_star / rho_L_star;
        fE = (E_L_star + p_star) * rho_u_L_star / rho_L_star;
    } else if (S_star <= 0.0 && S_R >= 0.0) {
        f0 = rho_u_R_star;
        fu = rho_u_R_star * rho_u_R_star / rho_R_star + p_star;
        fv = rho_v_R_star * rho_u_R_star / rho_R_star;
        fE = (E_R_star + p_star) * rho_u_R_star / rho_R_star;
    } else {
        f0 = rho_R * u_R;
        fu = rho_R * u_R * u_R + p_R;
        fv = rho_R * u_R * v_R;
        fE = (E_R + p_R) * u_R;
    }
    F[0] = f0;
    F[1] = (fu * nx + fv * (-ny));
    F[2] = (fu * ny + fv * (nx));
    F[3] = fE;
}

void orc_hllc_batch(const double *qL, const double *qR, const double *n, double gamma,
                    double *F, int count) {
    for (int i = 0; i < count; ++i) {
        double a[4], b[4], f[4];
        for (int q = 0; q < 4; ++q) {
            a[q] = qL[(size_t)q * count + i];
            b[q] = qR[(size_t)q * count + i];
        }
        orc_hllc(a, b, n[i], n[(size_t)count + i], gamma, f);
        for (int q = 0; q < 4; ++q) F[(size_t)q * count + i] = f[q];
    }
}

static inline void edge_hllc(const orc_mesh *m, const orc_params *p, int e, double *F) {
    double a[4], b[4];
    for (int q = 0; q < 4; ++q) {
        a[q] = m->fL[q][e];
        b[q] = m->fR[q][e];
    }
    orc_hllc(a, b, m->enx[e], m->eny[e], p->gamma, F);
}

/* computeAdvectionTerm, src/advection.cpp:415-466 */
int orc_advection(orc_mesh *m, const orc_params *p) {
    if (!p->bc_known) return -1; /* apply_boundary_conditions throws, src/boundary.cpp:89-91 */
    gradient_advection(m, p);
    interface_states(m, p);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < m->ne; ++e) {
        double F[4];
        edge_hllc(m, p, e, F);
        for (int q = 0; q < 4; ++q) m->flux[q][e] = F[q];
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < m->nc; ++c) {
        double r[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < 4; ++k) {
            int e = m->cedge[4 * c + k];
            if (e < 0) break;
            double sign = (m->ecl[e] == c ? +1.0 : -1.0);
            double L = m->elen[e];
            for (int q = 0; q < 4; ++q) r[q] += sign * m->flux[q][e] * L;
        }
        double invV = 1.0 / m->cvol[c];
        for (int q = 0; q < 4; ++q) m->R[q][c] = r[q] * -invV;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* diffusion.cpp                                                       */
/* ------------------------------------------------------------------ */

/* computeDiffusionTerm + computeScalarDivergenceDiffusion,
 * src/diffusion.cpp:206-243 and :6-203.  Adds into R; reads the cell
 * gradients left by orc_advection. */
int orc_diffusion(orc_mesh *m, const orc_params *p) {
    const double mu = p->viscosity;
    const double gamma = p->gamma;
    double kappa = p->viscosity * p->cp / p->Pr;
    if (p->heat_flux == 1) kappa = kappa / (p->cp - p->cp / p->gamma);
    double **tq = m->ctq; /* tau_xx, tau_yy, tau_xy, q_x, q_y */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < m->nc; ++c) {
        double rho = m->U[0][c], E = m->U[3][c];
        double rx = m->G[0][c], rux = m->G[1][c], rvx = m->G[2][c], Ex = m->G[3][c];
        double ry = m->G[4][c], ruy = m->G[5][c], rvy = m->G[6][c], Ey = m->G[7][c];
        double u = m->U[1][c] / rho;
        double v = m->U[2][c] / rho;
        double u_x = (rux - u * rx) / rho;
        double v_y = (rvy - v * ry) / rho;
        double u_y = (ruy - u * ry) / rho;
        double v_x = (rvx - v * rx) / rho;
        tq[0][c] = 2 * mu * u_x - 2.0 / 3.0 * mu * (u_x + v_y);
        tq[1][c] = 2 * mu * v_y - 2.0 / 3.0 * mu * (u_x + v_y);
        tq[2][c] = mu * (u_y + v_x);
        double T_x = (gamma - 1.0) * ((Ex * rho - E * rx) / (rho * rho) - (u * u_x + v * v_x));
        double T_y = (gamma - 1.0) * ((Ey * rho - E * ry) / (rho * rho) - (u * u_y + v * v_y));
        tq[3][c] = -kappa * T_x;
        tq[4][c] = -kappa * T_y;
    }
    double **vv = m->vval, **ev = m->eval_;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m->nv; ++i) {
        interp_vertex(m, p, i, 5, tq, vv);
        if (is_tagged(m->vbc[i])) bc_diffusion(m->vbc[i], &vv[0][i], &vv[1][i], &vv[2][i], &vv[3][i], &vv[4][i]);
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < m->ne; ++e) {
        int a = m->eva[e], b = m->evb[e];
        double t[5];
        for (int q = 0; q < 5; ++q) t[q] = 0.5 * (vv[q][a] + vv[q][b]);
        if (is_tagged(m->ebc[e])) {
            bc_diffusion(m->ebc[e], &t[0], &t[1], &t[2], &t[3], &t[4]);
            if (m->ebc[e] == 'w') {
                m->etau[0][e] = t[0];
                m->etau[1][e] = t[2];
                m->etau[2][e] = t[1];
            }
        }
        for (int q = 0; q < 5; ++q) ev[q][e] = t[q];
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < m->nc; ++c) {
        double cv = m->cvol[c];
        double u = m->U[1][c] / m->U[0][c];
        double v = m->U[2][c] / m->U[0][c];
        for (int k = 0; k < 4; ++k) {
            int e = m->cedge[4 * c + k];
            if (e < 0) break;
            double el = m->elen[e], enx = m->enx[e], eny = m->eny[e];
            if (m->ecl[e] == c) {
                enx = -enx;
                eny = -eny;
            }
            double txx = ev[0][e], tyy = ev[1][e], txy = ev[2][e], qx = ev[3][e], qy = ev[4][e];
            m->R[1][c] += txx * el * enx / cv + txy * el * eny / cv;
            m->R[2][c] += txy * el * enx / cv + tyy * el * eny / cv;
            m->R[3][c] += (u * txx + v * txy - qx) * el * enx / cv + (u * txy + v * tyy - qy) * el * eny / cv;
        }
    }
    return 0;
}

/* computeResidual, src/shape_optimization.cpp:171-189 */
int orc_residual(orc_mesh *m, const orc_params *p) {
    if (orc_advection(m, p)) return -1;
    return orc_diffusion(m, p);
}

/* ------------------------------------------------------------------ */
/* simulation.cpp: dt, SSP-RK3                                         */
/* ------------------------------------------------------------------ */

/* Time-step reduction, src/simulation.cpp:77-110 */
double orc_dt(const orc_mesh *m, const orc_params *p) {
    double max_wave = 1e-8;
    double min_h = INFINITY;
    for (int c = 0; c < m->nc; ++c) {
        double rho = m->U[0][c], ru = m->U[1][c], rv = m->U[2][c], E = m->U[3][c];
        double u = ru / rho, v = rv / rho;
        double kinetic = 0.5 * (ru * ru + rv * rv) / rho;
        double pr = (p->gamma - 1.0) * (E - kinetic);
        double cs = sqrt(p->gamma * pr / rho);
        double speed = hypot(u, v) + cs;
        max_wave = smax(max_wave, speed);
        double perimeter = 0.0;
        for (int k = 0; k < 4; ++k) {
            int e = m->cedge[4 * c + k];
            if (e < 0) break;
            perimeter += m->elen[e];
        }
        double h = 2.0 * m->cvol[c] / perimeter;
        min_h = smin(min_h, h);
    }
    double dt_adv = p->cfl * min_h / max_wave;
    double dt_diff = 0.5 * (min_h * min_h) / p->viscosity;
    return smin(dt_adv, dt_diff);
}

/* rk3_substep, src/simulation.cpp:145-171 */
static int rk3_substep(orc_mesh *m, const orc_params *p, double c0, double c1, double dt) {
    if (orc_residual(m, p)) return -1;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < m->nc; ++c)
        for (int q = 0; q < 4; ++q)
            m->U[q][c] = c0 * m->Uold[q][c] + c1 * (m->U[q][c] + dt * (m->R[q][c]));
    return 0;
}

/* The two halves of step_SSP_RK3 on their own (src/simulation.cpp:130-133
 * and Simulator::rk3_substep, :145-171). */
void orc_save_old(orc_mesh *m) {
    for (int q = 0; q < 4; ++q) memcpy(m->Uold[q], m->U[q], sizeof(double) * m->nc);
}
int orc_rk3_substep(orc_mesh *m, const orc_params *p, double c0, double c1, double dt) {
    return rk3_substep(m, p, c0, c1, dt);
}

/* step_SSP_RK3, src/simulation.cpp:130-142 */
int orc_rk3_step(orc_mesh *m, const orc_params *p, double dt) {
    for (int q = 0; q < 4; ++q) memcpy(m->Uold[q], m->U[q], sizeof(double) * m->nc);
    if (rk3_substep(m, p, 0.0, 1.0, dt)) return -1;
    if (rk3_substep(m, p, 3.0 / 4.0, 1.0 / 4.0, dt)) return -1;
    return rk3_substep(m, p, 1.0 / 3.0, 2.0 / 3.0, dt);
}

/* n steps of the time loop body (dt reduction then one SSP-RK3 step),
 * src/simulation.cpp:75-124; dt of each step is written to dts if non-NULL. */
int orc_run_steps(orc_mesh *m, const orc_params *p, int nsteps, double *dts) {
    for (int s = 0; s < nsteps; ++s) {
        double dt = orc_dt(m, p);
        if (dts) dts[s] = dt;
        if (orc_rk3_step(m, p, dt)) return -1;
    }
    return 0;
}

/* Simulator::time_loop, src/simulation.cpp:68-126, without its output and
 * hot-reload branches (they change no state when the config file does not
 * change): t = 0; while (t < final_time) { dt; step_SSP_RK3(dt); t += dt },
 * the last step may overshoot.  max_steps bounds the loop.  *steps and *t
 * receive the step count and the final t. */
int orc_time_loop(orc_mesh *m, const orc_params *p, double final_time, int max_steps, int *steps,
                  double *t_out) {
    double t = 0.0;
    int it = 0;
    while (t < final_time && it < max_steps) {
        double dt = orc_dt(m, p);
        if (orc_rk3_step(m, p, dt)) return -1;
        t += dt;
        ++it;
    }
    *steps = it;
    *t_out = t;
    return 0;
}

/* Density residual norm sqrt(mean_c R_rho^2) / (rho_inf U_inf / c), c = 1. */
double orc_residual_norm(const orc_mesh *m, const orc_params *p) {
    double s = 0.0;
    for (int c = 0; c < m->nc; ++c) s += m->R[0][c] * m->R[0][c];
    double a_inf = sqrt(p->gamma * p->p_ref / p->rho_ref);
    double scale = p->rho_ref * (p->mach * a_inf) / 1.0;
    return sqrt(s / m->nc) / scale;
}

/* The time loop run to a steady state instead of a final time (no
 * counterpart in the reference, whose loop stops at final_time,
 * src/simulation.cpp:75): steps n = 0, k, 2k, ... evaluate the density
 * residual norm of R(U^n) (the first stage's residual); a step whose norm
 * is <= tol completes and the loop stops after it.  tol <= 0: max_steps. */
int orc_run_to_residual(orc_mesh *m, const orc_params *p, double tol, int max_steps, int check_every,
                        int *steps, double *res) {
    double r = -1.0;
    int n = 0;
    while (n < max_steps) {
        const int check = tol > 0.0 && n % check_every == 0;
        const double dt = orc_dt(m, p);
        for (int q = 0; q < 4; ++q) memcpy(m->Uold[q], m->U[q], sizeof(double) * m->nc);
        if (rk3_substep(m, p, 0.0, 1.0, dt)) return -1;
        if (check) r = orc_residual_norm(m, p);
        if (rk3_substep(m, p, 3.0 / 4.0, 1.0 / 4.0, dt)) return -1;
        if (rk3_substep(m, p, 1.0 / 3.0, 2.0 / 3.0, dt)) return -1;
        ++n;
        if (check && r <= tol) break;
    }
    *steps = n;
    *res = r;
    return 0;
}

/* ------------------------------------------------------------------ */
/* compute.cpp: C_L, C_D; shape_optimization.cpp objective             */
/* ------------------------------------------------------------------ */

static void wall_forces(const orc_mesh *m, const orc_params *p, int skip_x_gt_1_3, double *CL,
                        double *CD) {
    double gamma = p->gamma;
    double rho_inf = p->rho_ref, p_inf = p->p_ref;
    const double a_inf = sqrt(gamma * p_inf / rho_inf);
    double U_inf = p->mach * a_inf;
    double q_inf = 0.5 * (rho_inf * U_inf * U_inf);
    double corde = 1;
    double L = 0.0, D = 0.0;
    for (int e = 0; e < m->ne; ++e) {
        if (m->ebc[e] != 'w') continue;
        if (skip_x_gt_1_3 && m->emx[e] > 1.3) continue;
        double nx = m->enx[e], ny = m->eny[e], A = m->elen[e];
        double pr = (p->gamma - 1.0) *
                    (m->fL[3][e] - 0.5 * (m->fL[1][e] * m->fL[1][e] + m->fL[2][e] * m->fL[2][e]) / m->fL[0][e]);
        double txx = m->etau[0][e], txy = m->etau[1][e], tyy = m->etau[2][e];
        L += (-pr * ny + txy * nx + tyy * ny) * A;
        D += (-pr * nx + txx * nx + txy * ny) * A;
    }
    *CL = L / (q_inf * corde);
    *CD = D / (q_inf * corde);
}

/* computeLiftDragCoefficients, src/compute.cpp:92-132 */
void orc_lift_drag(const orc_mesh *m, const orc_params *p, double *CL, double *CD) {
    wall_forces(m, p, 0, CL, CD);
}

/* computeDirectObjectiveTerm, src/shape_optimization.cpp:893-939 */
double orc_objective(const orc_mesh *m, const orc_params *p, double beta) {
    double CL, CD;
    wall_forces(m, p, 1, &CL, &CD);
    return -CL + beta * CD;
}

/* ------------------------------------------------------------------ */
/* shape_optimization.cpp: adjoint                                     */
/* ------------------------------------------------------------------ */

/* assembleAdjointRHS, src/shape_optimization.cpp:491-554 (b in SoA). */
void orc_adjoint_rhs(const orc_mesh *m, const orc_params *p, double beta, double *b) {
    int nc = m->nc;
    double gamma = p->gamma;
    double rho_inf = p->rho_ref, p_inf = p->p_ref;
    const double a_inf = sqrt(gamma * p_inf / rho_inf);
    double U_inf = p->mach * a_inf;
    double q_inf = 0.5 * (rho_inf * U_inf * U_inf);
    double corde = 1;
    memset(b, 0, sizeof(double) * 4 * nc);
    for (int e = 0; e < m->ne; ++e) {
        if (m->ebc[e] != 'w') continue;
        int iL = m->ecl[e];
        double A_f = m->elen[e], nx = m->enx[e], ny = m->eny[e];
        double rL = m->fL[0][e], ruL = m->fL[1][e], rvL = m->fL[2][e];
        double dp[4];
        dp[0] = 0.5 * (p->gamma - 1.0) * ((ruL * ruL + rvL * rvL) / (rL * rL));
        dp[1] = -(p->gamma - 1.0) * (ruL / rL);
        dp[2] = -(p->gamma - 1.0) * (rvL / rL);
        dp[3] = (p->gamma - 1.0);
        double coeff_CL = +ny * A_f / (q_inf * corde);
        for (int q = 0; q < 4; ++q) b[(size_t)q * nc + iL] += coeff_CL * dp[q];
        double coeff_CD = -nx * A_f / (q_inf * corde) * beta;
        for (int q = 0; q < 4; ++q) b[(size_t)q * nc + iL] += coeff_CD * dp[q];
    }
}

/* FD flux-Jacobian blocks of one edge, src/shape_optimization.cpp:245-429:
 * FUL[r][c] = dF_r/dU_L,c and FUR likewise by central differences with an
 * absolute step eps on the stored face states.  A boundary edge rebuilds its
 * ghost state from each perturbed left state.  (The wall branch for the
 * right-state perturbation, :348-350, only feeds FUR, which the boundary
 * scatter never reads, so FUR is left zero for boundary edges.) */
void orc_edge_blocks(const orc_mesh *m, const orc_params *p, int e, double eps,
                     double FUL[4][4], double FUR[4][4]) {
    char bc = m->ebc[e];
    double base_L[4], base_R[4];
    for (int q = 0; q < 4; ++q) {
        base_L[q] = m->fL[q][e];
        base_R[q] = m->fR[q][e];
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) FUL[r][c] = FUR[r][c] = 0.0;
    for (int c = 0; c < 4; ++c) {
        double Lp[4], Lm[4], Rp[4], Rm[4], Fp[4], Fm[4];
        for (int q = 0; q < 4; ++q) {
            Lp[q] = Lm[q] = base_L[q];
            Rp[q] = Rm[q] = base_R[q];
        }
        Lp[c] = base_L[c] + eps;
        Lm[c] = base_L[c] - eps;
        if (bc != '\0') {
            bc_interface(p, bc, Lp[0], Lp[1], Lp[2], Lp[3], &Rp[0], &Rp[1], &Rp[2], &Rp[3]);
            bc_interface(p, bc, Lm[0], Lm[1], Lm[2], Lm[3], &Rm[0], &Rm[1], &Rm[2], &Rm[3]);
        }
        orc_hllc(Lp, Rp, m->enx[e], m->eny[e], p->gamma, Fp);
        orc_hllc(Lm, Rm, m->enx[e], m->eny[e], p->gamma, Fm);
        for (int r = 0; r < 4; ++r) FUL[r][c] = (Fp[r] - Fm[r]) / (2 * eps);
    }
    if (bc != '\0') return;
    for (int c = 0; c < 4; ++c) {
        double Rp[4], Rm[4], Fp[4], Fm[4];
        for (int q = 0; q < 4; ++q) Rp[q] = Rm[q] = base_R[q];
        Rp[c] = base_R[c] + eps;
        Rm[c] = base_R[c] - eps;
        orc_hllc(base_L, Rp, m->enx[e], m->eny[e], p->gamma, Fp);
        orc_hllc(base_L, Rm, m->enx[e], m->eny[e], p->gamma, Fm);
        for (int r = 0; r < 4; ++r) FUR[r][c] = (Fp[r] - Fm[r]) / (2 * eps);
    }
}

/* applyJacobianTranspose, src/shape_optimization.cpp:237-484 (x, y SoA).
 * The reference scatters with `omp atomic`; here edges are scattered in
 * ascending edge order, one valid serialisation of those atomics. */
void orc_apply_jt(const orc_mesh *m, const orc_params *p, const double *x, double *y,
                  double eps) {
    int nc = m->nc, ne = m->ne;
    double *blk = dalloc((size_t)32 * ne);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e)
        orc_edge_blocks(m, p, e, eps, (double(*)[4])(blk + 32 * (size_t)e),
                        (double(*)[4])(blk + 32 * (size_t)e + 16));
    memset(y, 0, sizeof(double) * 4 * nc);
    const double *xr = x, *xru = x + nc, *xrv = x + 2 * (size_t)nc, *xE = x + 3 * (size_t)nc;
    double *yq[4] = {y, y + nc, y + 2 * (size_t)nc, y + 3 * (size_t)nc};
    for (int e = 0; e < ne; ++e) {
        double(*FUL)[4] = (double(*)[4])(blk + 32 * (size_t)e);
        double(*FUR)[4] = (double(*)[4])(blk + 32 * (size_t)e + 16);
        double L = m->elen[e];
        int iL = m->ecl[e];
        double invV_L = 1.0 / m->cvol[iL];
        double D[4];
        if (m->ebc[e] != '\0') {
            D[0] = (-xr[iL] * invV_L) * L;
            D[1] = (-xru[iL] * invV_L) * L;
            D[2] = (-xrv[iL] * invV_L) * L;
            D[3] = (-xE[iL] * invV_L) * L;
            for (int k = 0; k < 4; ++k)
                yq[k][iL] += (FUL[0][k] * D[0] + FUL[1][k] * D[1] + FUL[2][k] * D[2] + FUL[3][k] * D[3]);
            continue;
        }
        int iR = m->ecr[e];
        double invV_R = 1.0 / m->cvol[iR];
        D[0] = (-xr[iL] * invV_L + xr[iR] * invV_R) * L;
        D[1] = (-xru[iL] * invV_L + xru[iR] * invV_R) * L;
        D[2] = (-xrv[iL] * invV_L + xrv[iR] * invV_R) * L;
        D[3] = (-xE[iL] * invV_L + xE[iR] * invV_R) * L;
        for (int k = 0; k < 4; ++k)
            yq[k][iL] += (FUL[0][k] * D[0] + FUL[1][k] * D[1] + FUL[2][k] * D[2] + FUL[3][k] * D[3]);
        for (int k = 0; k < 4; ++k)
            yq[k][iR] += (FUR[0][k] * D[0] + FUR[1][k] * D[1] + FUR[2][k] * D[2] + FUR[3][k] * D[3]);
    }
    free(blk);
}

/* Packed (4i+k) matvec glue of JacobianOpT::_mul, src/shape_optimization.cpp:612-631 */
static void jt_packed(const orc_mesh *m, const orc_params *p, double eps, const double *xp,
                      double *yp, double *sx, double *sy) {
    int nc = m->nc;
    for (int i = 0; i < nc; ++i)
        for (int k = 0; k < 4; ++k) sx[(size_t)k * nc + i] = xp[4 * (size_t)i + k];
    orc_apply_jt(m, p, sx, sy, eps);
    for (int i = 0; i < nc; ++i)
        for (int k = 0; k < 4; ++k) yp[4 * (size_t)i + k] = sy[(size_t)k * nc + i];
}

/* Dot products of the BiCGSTAB restatement.  Sequential by default (the
 * reference's Eigen sums in its own packet order; the difference is
 * rounding).  In device order the sum follows the gfx950 reduction exactly,
 * so a device solve can be compared bit for bit: the vectors are taken in
 * the device's SoA element order (element j = variable j / nc of cell
 * j % nc), thread t of block b of a grid of G = min(512, ceil(n / 256))
 * blocks of 256 sums elements b*256 + t, b*256 + t + 256 G, ... in order,
 * each block adds its 256 sums by a halving tree (s[t] += s[t + h] for
 * h = 128 .. 1), and the G block sums are added the same way (thread t
 * sums blocks t, t + 256 in order, then the tree). */
static int g_dot_device_order = 0;
static double g_restart_eps = DBL_EPSILON; /* Eigen's NumTraits<double>::epsilon() */
static int g_restarts = 0;                 /* restarts taken by the last orc_bicgstab */

int orc_bicgstab_restarts(void) { return g_restarts; }

void orc_bicgstab_config(int device_order, double restart_eps) {
    g_dot_device_order = device_order;
    g_restart_eps = restart_eps;
}

static void halving_tree(double *s) {
    for (int h = 128; h > 0; h >>= 1)
        for (int t = 0; t < h; ++t) s[t] += s[t + h];
}

static double dot_device_order(const double *a, const double *b, size_t n) {
    const size_t nc = n / 4;
    size_t blocks = (n + 255) / 256;
    const int G = (int)(blocks < 1 ? 1 : blocks > 512 ? 512 : blocks);
    const size_t stride = (size_t)G * 256;
    double part[512], s[256];
    for (int blk = 0; blk < G; ++blk) {
        for (int t = 0; t < 256; ++t) {
            double x = 0.0;
            for (size_t j = (size_t)blk * 256 + t; j < n; j += stride) {
                const size_t k = 4 * (j % nc) + j / nc; /* packed index of SoA element j */
                x += a[k] * b[k];
            }
            s[t] = x;
        }
        halving_tree(s);
        part[blk] = s[0];
    }
    for (int t = 0; t < 256; ++t) {
        double x = 0.0;
        for (int bb = t; bb < G; bb += 256) x += part[bb];
        s[t] = x;
    }
    halving_tree(s);
    return s[0];
}

static double dotp(const double *a, const double *b, size_t n) {
    if (g_dot_device_order) return dot_device_order(a, b, n);
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Eigen::BiCGSTAB with IdentityPreconditioner (Eigen 3.4,
 * Eigen/src/IterativeLinearSolvers/BiCGSTAB.h, function internal::bicgstab),
 * restated: the van der Vorst (1992) recurrence, stop when
 * |r|^2 <= tol^2 |b|^2 or maxit, restart with r0 = r when |r0.r| <
 * eps_mach^2 |r0|^2 (the first restart also resets the iteration count).
 * Sums are sequential here; Eigen's packet reductions reorder them at the
 * rounding level. Returns the iteration count. */
static int bicgstab_chunk(const orc_mesh *m, const orc_params *p, double eps, const double *b,
                          double *x, int maxit, double tol, double *work) {
    size_t n = 4 * (size_t)m->nc;
    double *r = work, *r0 = work + n, *v = work + 2 * n, *pp = work + 3 * n, *s = work + 4 * n,
           *t = work + 5 * n, *sx = work + 6 * n, *sy = work + 7 * n;
    jt_packed(m, p, eps, x, r, sx, sy);
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    memcpy(r0, r, n * sizeof(double));
    double r0_sqnorm = dotp(r0, r0, n);
    double rhs_sqnorm = dotp(b, b, n);
    if (rhs_sqnorm == 0) {
        memset(x, 0, n * sizeof(double));
        return 0;
    }
    double rho = 1, alpha = 1, w = 1;
    memset(v, 0, n * sizeof(double));
    memset(pp, 0, n * sizeof(double));
    double tol2 = tol * tol * rhs_sqnorm;
    double eps2 = g_restart_eps * g_restart_eps;
    int i = 0, restarts = 0;
    while (dotp(r, r, n) > tol2 && i < maxit) {
        double rho_old = rho;
        rho = dotp(r0, r, n);
        if (fabs(rho) < eps2 * r0_sqnorm) {
            jt_packed(m, p, eps, x, r, sx, sy);
            for (size_t k = 0; k < n; ++k) r[k] = b[k] - r[k];
            memcpy(r0, r, n * sizeof(double));
            rho = r0_sqnorm = dotp(r, r, n);
            ++g_restarts;
            if (restarts++ == 0) i = 0;
        }
        double beta = (rho / rho_old) * (alpha / w);
        for (size_t k = 0; k < n; ++k) pp[k] = r[k] + beta * (pp[k] - w * v[k]);
        jt_packed(m, p, eps, pp, v, sx, sy);
        alpha = rho / dotp(r0, v, n);
        for (size_t k = 0; k < n; ++k) s[k] = r[k] - alpha * v[k];
        jt_packed(m, p, eps, s, t, sx, sy);
        double tmp = dotp(t, t, n);
        w = (tmp > 0) ? dotp(t, s, n) / tmp : 0.0;
        for (size_t k = 0; k < n; ++k) x[k] += alpha * pp[k] + w * s[k];
        for (size_t k = 0; k < n; ++k) r[k] = s[k] - w * t[k];
        ++i;
    }
    return i;
}

/* solveAdjointSystem, src/shape_optimization.cpp:658-740.  lambda SoA.
 * Returns 0, -2 "BiCGSTAB diverged", -3 "BiCGSTAB did not converge". */
int orc_bicgstab(const orc_mesh *m, const orc_params *p, double beta, double eps, double tol,
                 int chunk, int max_total, double *lambda, int *iters, double *relres_out) {
    int nc = m->nc;
    size_t n = 4 * (size_t)nc;
    double *bs = dalloc(n), *b = dalloc(n), *u = dalloc(n), *r = dalloc(n);
    double *work = dalloc(8 * n);
    orc_adjoint_rhs(m, p, beta, bs);
    for (int i = 0; i < nc; ++i)
        for (int k = 0; k < 4; ++k) b[4 * (size_t)i + k] = bs[(size_t)k * nc + i];
    g_restarts = 0;
    double bnorm = sqrt(dotp(b, b, n));
    double relres = INFINITY;
    int it_total = 0, status = 0;
    while (it_total < max_total) {
        const int it = bicgstab_chunk(m, p, eps, b, u, chunk, tol, work);
        it_total += it;
        jt_packed(m, p, eps, u, r, work + 6 * n, work + 7 * n);
        for (size_t k = 0; k < n; ++k) r[k] = b[k] - r[k];
        relres = sqrt(dotp(r, r, n)) / (bnorm > 0 ? bnorm : 1.0);
        if (relres <= tol) break;
        if (!isfinite(relres)) {
            status = -2;
            break;
        }
        /* a chunk that ran no iteration leaves x as it was: stop instead of
         * repeating it forever (the reference's loop would not end) */
        if (it == 0) break;
    }
    if (status == 0 && relres > tol) status = -3;
    for (int i = 0; i < nc; ++i)
        for (int k = 0; k < 4; ++k) lambda[(size_t)k * nc + i] = u[4 * (size_t)i + k];
    *iters = it_total;
    *relres_out = relres;
    free(bs);
    free(b);
    free(u);
    free(r);
    free(work);
    return status;
}

/* ------------------------------------------------------------------ */
/* shape_optimization.cpp: NACA perturbation and shape gradient        */
/* ------------------------------------------------------------------ */

/* yc_naca4 / yt_naca4, src/shape_optimization.cpp:807-826 */
static double yc_naca4(double m, double p, double xr) {
    xr = sclamp(xr, 0.0, 1.0);
    if (m == 0.0) return 0.0;
    if (xr <= p) return m / (p * p) * (2.0 * p * xr - xr * xr);
    return m / ((1.0 - p) * (1.0 - p)) * ((1.0 - 2.0 * p) + 2.0 * p * xr - xr * xr);
}
static double yt_naca4(double t, double c, double xr) {
    xr = sclamp(xr, 0.0, 1.0);
    return 5.0 * t * c *
           (0.2969 * sqrt(xr) - 0.1260 * xr - 0.3516 * xr * xr + 0.2843 * xr * xr * xr -
            0.1015 * xr * xr * xr * xr);
}

/* updateDerivedGeometry, src/shape_optimization.cpp:748-805 */
static void update_derived_geometry(orc_mesh *m) {
    for (int e = 0; e < m->ne; ++e) {
        if (m->ebc[e] != 'w' && m->vbc[m->eva[e]] != 'w' && m->vbc[m->evb[e]] != 'w') continue;
        int cl = m->ecl[e];
        /* length and normal first, then the left cell, then the flip test */
        cell_geometry(m, cl, &m->cvol[cl], &m->ccx[cl], &m->ccy[cl]);
        edge_geometry(m, e);
    }
}

/* applyShapePerturbationNACA12, src/shape_optimization.cpp:828-890.
 * which: 0 = "m"/"cambrure", 1 = "p"/"position_cambrure", other = no-op. */
int orc_perturb_shape(orc_mesh *m, int which, double pos_x, double c, double m0, double p0,
                      double t, double eps) {
    const double freeze_start = 0.9;
    double m1 = m0, p1 = p0;
    if (which == 0)
        m1 = m0 + eps;
    else if (which == 1)
        p1 = p0 + eps;
    else
        return 0;
    for (int v = 0; v < m->nv; ++v) {
        if (m->vbc[v] != 'w') continue;
        const double x_local = m->vx[v] - pos_x;
        const double xr = x_local / c;
        if (xr >= freeze_start) continue;
        const double yc0 = yc_naca4(m0, p0, xr);
        const double yt0 = yt_naca4(t, c, xr);
        const double s = (m->vy[v] >= yc0) ? +1.0 : -1.0;
        const double yc1 = yc_naca4(m1, p1, xr);
        m->vx[v] = pos_x + xr * c;
        m->vy[v] = yc1 + s * yt0;
    }
    update_derived_geometry(m);
    return 1;
}

/* computeShapeGradient, src/shape_optimization.cpp:942-1016, for one
 * parameter.  lambda is SoA; the dot product runs in the reference's packed
 * order.  Leaves the face states of m at the current U (computeResidual at
 * :956).  Returns 0 or -1 (unknown BC mode). */
int orc_shape_gradient(orc_mesh *m, const orc_params *p, const double *lambda, int which,
                       double pos_x, double c, double m0, double p0, double t, double beta,
                       double eps, double *grad) {
    int nc = m->nc;
    if (orc_residual(m, p)) return -1;
    orc_mesh *mp = orc_mesh_clone(m);
    orc_perturb_shape(mp, which, pos_x, c, m0, p0, t, eps);
    orc_residual(mp, p);
    orc_mesh *mm = orc_mesh_clone(m);
    orc_perturb_shape(mm, which, pos_x, c, m0, p0, t, -eps);
    orc_residual(mm, p);
    double s = 0.0;
    for (int i = 0; i < nc; ++i)
        for (int k = 0; k < 4; ++k) {
            double d = (mp->R[k][i] - mm->R[k][i]) / (2.0 * eps);
            s += lambda[(size_t)k * nc + i] * d;
        }
    double Jp = orc_objective(mp, p, beta);
    double Jm = orc_objective(mm, p, beta);
    double dJ = (Jp - Jm) / (2.0 * eps);
    *grad = dJ - s;
    orc_mesh_destroy(mp);
    orc_mesh_destroy(mm);
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
