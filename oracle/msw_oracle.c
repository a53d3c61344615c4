/*
 * CPU oracle: serial C restatement of the reference model (test
 * infrastructure only; see msw_oracle.h).  Loop order, scatter form and
 * expression association follow the reference Fortran line by line so that
 * the two agree to round-off; each function names the routine it restates.
 *
 * Parity pins: the reference Fortran itself (oracle/_ref, compiled from
 * the src/ Fortran) through tests/test_oracle_vs_reference.py and the committed
 * golden vectors; closed forms (RK4 on a linear ODE, mass conservation) and
 * the TLM/adjoint identities in tests/test_oracle_math.py.
 */
#define _GNU_SOURCE
#include "msw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GRAVITY 9.80616      /* src/mpas_sw_dyncore.f90:414 */
#define RHO_REF 1000.0       /* src/mpas_sw_dyncore.f90:460 */

typedef struct diagset {
    double *v, *h_edge, *h_vertex, *circulation, *vorticity, *divergence, *ke;
    double *pv_edge, *pv_vertex, *pv_cell, *vorticity_cell, *gradPVn, *gradPVt;
} diagset;

typedef struct reconset {
    double *X, *Y, *Z, *Zonal, *Meridional;
} reconset;

struct mswo_ctx {
    int C, E, V, mE, mE2, vdeg;
    /* mesh (owned copies) */
    int *coe, *voe, *nEoC, *eoc, *coc, *nEoE, *eoe, *eov, *cov;
    double *woe, *kite, *dc, *dv, *areaC, *areaT, *fV, *hs, *usrc, *msd2, *msd4;
    double *latC, *lonC, *coeffs;
    mswo_config cfg;
    /* state: level 1 and 2 (the reference's u(2,1,E), h(2,1,C)) */
    double *h[2], *u[2], *h_tl[2], *u_tl[2], *h_ad[2], *u_ad[2];
    diagset d, dtl, dad;
    double *tend_h, *tend_u, *tend_h_tl, *tend_u_tl, *tend_h_ad, *tend_u_ad;
    reconset r, rtl, rad;
};

/* accessors for (slot, entity) arrays with 1-based values -> 0-based index */
#define CELL1(c, e) ((c)->coe[2 * (e)] - 1)
#define CELL2(c, e) ((c)->coe[2 * (e) + 1] - 1)
#define VTX1(c, e)  ((c)->voe[2 * (e)] - 1)
#define VTX2(c, e)  ((c)->voe[2 * (e) + 1] - 1)
#define EOC(c, i, ic) ((c)->eoc[(ic) * (c)->mE + (i)] - 1)
#define EOE(c, j, ie) ((c)->eoe[(ie) * (c)->mE2 + (j)] - 1)
#define WOE(c, j, ie) ((c)->woe[(ie) * (c)->mE2 + (j)])
#define EOV(c, i, iv) ((c)->eov[(iv) * (c)->vdeg + (i)] - 1)
#define COV(c, i, iv) ((c)->cov[(iv) * (c)->vdeg + (i)] - 1)
#define KITE(c, i, iv) ((c)->kite[(iv) * (c)->vdeg + (i)])

static void *dupe(const void *src, size_t bytes)
{
    void *p = malloc(bytes ? bytes : 1);
    if (p && src) memcpy(p, src, bytes);
    else if (p) memset(p, 0, bytes);
    return p;
}

static double *zeros(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static int alloc_diag(diagset *d, int C, int E, int V)
{
    d->v = zeros(E); d->h_edge = zeros(E); d->h_vertex = zeros(V);
    d->circulation = zeros(V); d->vorticity = zeros(V); d->divergence = zeros(C);
    d->ke = zeros(C); d->pv_edge = zeros(E); d->pv_vertex = zeros(V);
    d->pv_cell = zeros(C); d->vorticity_cell = zeros(C); d->gradPVn = zeros(E);
    d->gradPVt = zeros(E);
    return d->v && d->h_edge && d->h_vertex && d->circulation && d->vorticity &&
           d->divergence && d->ke && d->pv_edge && d->pv_vertex && d->pv_cell &&
           d->vorticity_cell && d->gradPVn && d->gradPVt;
}

static void free_diag(diagset *d)
{
    free(d->v); free(d->h_edge); free(d->h_vertex); free(d->circulation);
    free(d->vorticity); free(d->divergence); free(d->ke); free(d->pv_edge);
    free(d->pv_vertex); free(d->pv_cell); free(d->vorticity_cell);
    free(d->gradPVn); free(d->gradPVt);
}

static int alloc_recon(reconset *r, int C)
{
    r->X = zeros(C); r->Y = zeros(C); r->Z = zeros(C); r->Zonal = zeros(C);
    r->Meridional = zeros(C);
    return r->X && r->Y && r->Z && r->Zonal && r->Meridional;
}

static void free_recon(reconset *r)
{
    free(r->X); free(r->Y); free(r->Z); free(r->Zonal); free(r->Meridional);
}

mswo_ctx *mswo_create(const mswo_mesh *m, const mswo_config *cfg)
{
    mswo_ctx *c = (mswo_ctx *)calloc(1, sizeof(mswo_ctx));
    if (!c) return NULL;
    int C = m->nCells, E = m->nEdges, V = m->nVertices;
    c->C = C; c->E = E; c->V = V; c->mE = m->maxEdges; c->mE2 = m->maxEdges2;
    c->vdeg = m->vertexDegree;
    size_t I = sizeof(int), D = sizeof(double);
    c->coe = dupe(m->cellsOnEdge, 2 * E * I);
    c->voe = dupe(m->verticesOnEdge, 2 * E * I);
    c->nEoC = dupe(m->nEdgesOnCell, C * I);
    c->eoc = dupe(m->edgesOnCell, (size_t)c->mE * C * I);
    c->coc = dupe(m->cellsOnCell, (size_t)c->mE * C * I);
    c->nEoE = dupe(m->nEdgesOnEdge, E * I);
    c->eoe = dupe(m->edgesOnEdge, (size_t)c->mE2 * E * I);
    c->woe = dupe(m->weightsOnEdge, (size_t)c->mE2 * E * D);
    c->eov = dupe(m->edgesOnVertex, (size_t)c->vdeg * V * I);
    c->cov = dupe(m->cellsOnVertex, (size_t)c->vdeg * V * I);
    c->kite = dupe(m->kiteAreasOnVertex, (size_t)c->vdeg * V * D);
    c->dc = dupe(m->dcEdge, E * D);
    c->dv = dupe(m->dvEdge, E * D);
    c->areaC = dupe(m->areaCell, C * D);
    c->areaT = dupe(m->areaTriangle, V * D);
    c->fV = dupe(m->fVertex, V * D);
    c->hs = dupe(m->h_s, C * D);
    c->usrc = dupe(m->u_src, E * D);
    c->msd2 = dupe(m->meshScalingDel2, E * D);
    c->msd4 = dupe(m->meshScalingDel4, E * D);
    c->latC = dupe(m->latCell, C * D);
    c->lonC = dupe(m->lonCell, C * D);
    c->coeffs = dupe(m->coeffs_reconstruct, (size_t)3 * c->mE * C * D);
    if (!m->meshScalingDel2) for (int e = 0; e < E; ++e) c->msd2[e] = 1.0;
    if (!m->meshScalingDel4) for (int e = 0; e < E; ++e) c->msd4[e] = 1.0;
    c->cfg = *cfg;
    int ok = c->coe && c->voe && c->nEoC && c->eoc && c->coc && c->nEoE && c->eoe &&
             c->woe && c->eov && c->cov && c->kite && c->dc && c->dv && c->areaC &&
             c->areaT && c->fV && c->hs && c->usrc && c->msd2 && c->msd4 && c->latC &&
             c->lonC && c->coeffs;
    for (int l = 0; l < 2; ++l) {
        c->h[l] = zeros(C); c->u[l] = zeros(E);
        c->h_tl[l] = zeros(C); c->u_tl[l] = zeros(E);
        c->h_ad[l] = zeros(C); c->u_ad[l] = zeros(E);
        ok = ok && c->h[l] && c->u[l] && c->h_tl[l] && c->u_tl[l] && c->h_ad[l] && c->u_ad[l];
    }
    ok = ok && alloc_diag(&c->d, C, E, V) && alloc_diag(&c->dtl, C, E, V) &&
         alloc_diag(&c->dad, C, E, V);
    c->tend_h = zeros(C); c->tend_u = zeros(E);
    c->tend_h_tl = zeros(C); c->tend_u_tl = zeros(E);
    c->tend_h_ad = zeros(C); c->tend_u_ad = zeros(E);
    ok = ok && c->tend_h && c->tend_u && c->tend_h_tl && c->tend_u_tl && c->tend_h_ad &&
         c->tend_u_ad;
    ok = ok && alloc_recon(&c->r, C) && alloc_recon(&c->rtl, C) && alloc_recon(&c->rad, C);
    if (!ok) { mswo_destroy(c); return NULL; }
    return c;
}

void mswo_destroy(mswo_ctx *c)
{
    if (!c) return;
    free(c->coe); free(c->voe); free(c->nEoC); free(c->eoc); free(c->coc);
    free(c->nEoE); free(c->eoe); free(c->woe); free(c->eov); free(c->cov);
    free(c->kite); free(c->dc); free(c->dv); free(c->areaC); free(c->areaT);
    free(c->fV); free(c->hs); free(c->usrc); free(c->msd2); free(c->msd4);
    free(c->latC); free(c->lonC); free(c->coeffs);
    for (int l = 0; l < 2; ++l) {
        free(c->h[l]); free(c->u[l]); free(c->h_tl[l]); free(c->u_tl[l]);
        free(c->h_ad[l]); free(c->u_ad[l]);
    }
    free_diag(&c->d); free_diag(&c->dtl); free_diag(&c->dad);
    free(c->tend_h); free(c->tend_u); free(c->tend_h_tl); free(c->tend_u_tl);
    free(c->tend_h_ad); free(c->tend_u_ad);
    free_recon(&c->r); free_recon(&c->rtl); free_recon(&c->rad);
    free(c);
}

void mswo_set_config(mswo_ctx *c, const mswo_config *cfg) { c->cfg = *cfg; }

static double *diag_field(diagset *d, const char *n, int C, int E, int V, long *len)
{
    struct { const char *name; double *p; int kind; } t[] = {
        {"v", d->v, 1}, {"h_edge", d->h_edge, 1}, {"h_vertex", d->h_vertex, 2},
        {"circulation", d->circulation, 2}, {"vorticity", d->vorticity, 2},
        {"divergence", d->divergence, 0}, {"ke", d->ke, 0}, {"pv_edge", d->pv_edge, 1},
        {"pv_vertex", d->pv_vertex, 2}, {"pv_cell", d->pv_cell, 0},
        {"vorticity_cell", d->vorticity_cell, 0}, {"gradPVn", d->gradPVn, 1},
        {"gradPVt", d->gradPVt, 1}};
    for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i)
        if (!strcmp(n, t[i].name)) {
            *len = t[i].kind == 0 ? C : (t[i].kind == 1 ? E : V);
            return t[i].p;
        }
    return NULL;
}

double *mswo_array(mswo_ctx *c, const char *name, long *n)
{
    char base[64];
    size_t L = strlen(name);
    if (L >= sizeof(base)) return NULL;
    strcpy(base, name);
    diagset *ds = &c->d;
    reconset *rs = &c->r;
    if (L > 3 && !strcmp(name + L - 3, "_tl")) { ds = &c->dtl; rs = &c->rtl; base[L - 3] = 0; }
    else if (L > 3 && !strcmp(name + L - 3, "_ad")) { ds = &c->dad; rs = &c->rad; base[L - 3] = 0; }
    struct { const char *name; double *p; long n; } st[] = {
        {"h1", c->h[0], c->C}, {"h2", c->h[1], c->C}, {"u1", c->u[0], c->E}, {"u2", c->u[1], c->E},
        {"h_tl1", c->h_tl[0], c->C}, {"h_tl2", c->h_tl[1], c->C},
        {"u_tl1", c->u_tl[0], c->E}, {"u_tl2", c->u_tl[1], c->E},
        {"h_ad1", c->h_ad[0], c->C}, {"h_ad2", c->h_ad[1], c->C},
        {"u_ad1", c->u_ad[0], c->E}, {"u_ad2", c->u_ad[1], c->E},
        {"tend_h", c->tend_h, c->C}, {"tend_u", c->tend_u, c->E},
        {"tend_h_tl", c->tend_h_tl, c->C}, {"tend_u_tl", c->tend_u_tl, c->E},
        {"tend_h_ad", c->tend_h_ad, c->C}, {"tend_u_ad", c->tend_u_ad, c->E},
        {"fVertex", c->fV, c->V}, {"h_s", c->hs, c->C}};
    for (size_t i = 0; i < sizeof(st) / sizeof(st[0]); ++i)
        if (!strcmp(name, st[i].name)) { *n = st[i].n; return st[i].p; }
    if (!strcmp(base, "uReconstructX")) { *n = c->C; return rs->X; }
    if (!strcmp(base, "uReconstructY")) { *n = c->C; return rs->Y; }
    if (!strcmp(base, "uReconstructZ")) { *n = c->C; return rs->Z; }
    if (!strcmp(base, "uReconstructZonal")) { *n = c->C; return rs->Zonal; }
    if (!strcmp(base, "uReconstructMeridional")) { *n = c->C; return rs->Meridional; }
    return diag_field(ds, base, c->C, c->E, c->V, n);
}

/* ---------------------------------------------------------------------- */
/* sw_compute_solve_diagnostics: src/mpas_sw_dyncore.f90:1-361 (order 2;
 * orders 3/4 give the same numbers because deriv_two is never filled,
 * src/mpas_sw_module.f90:285-286)                                          */
void mswo_diag(mswo_ctx *c, const double *h, const double *u)
{
    const int C = c->C, E = c->E, V = c->V;
    const double dt = c->cfg.dt, apvm = c->cfg.apvm;
    diagset *d = &c->d;
    for (int e = 0; e < E; ++e)                                   /* :95-105 */
        d->h_edge[e] = 0.5 * (h[CELL1(c, e)] + h[CELL2(c, e)]);
    for (int v = 0; v < V; ++v) d->circulation[v] = 0.0;          /* :208-219 */
    for (int e = 0; e < E; ++e) {
        d->circulation[VTX1(c, e)] = d->circulation[VTX1(c, e)] - c->dc[e] * u[e];
        d->circulation[VTX2(c, e)] = d->circulation[VTX2(c, e)] + c->dc[e] * u[e];
    }
    for (int v = 0; v < V; ++v) d->vorticity[v] = d->circulation[v] / c->areaT[v];
    for (int i = 0; i < C; ++i) d->divergence[i] = 0.0;           /* :222-242 */
    for (int e = 0; e < E; ++e) {
        d->divergence[CELL1(c, e)] = d->divergence[CELL1(c, e)] + u[e] * c->dv[e];
        d->divergence[CELL2(c, e)] = d->divergence[CELL2(c, e)] - u[e] * c->dv[e];
    }
    for (int i = 0; i < C; ++i) {
        double r = 1.0 / c->areaC[i];
        d->divergence[i] = d->divergence[i] * r;
    }
    for (int i = 0; i < C; ++i) {                                 /* :245-256 */
        d->ke[i] = 0.0;
        for (int k = 0; k < c->nEoC[i]; ++k) {
            int e = EOC(c, k, i);
            d->ke[i] = d->ke[i] + 0.25 * c->dc[e] * c->dv[e] * (u[e] * u[e]);
        }
        d->ke[i] = d->ke[i] / c->areaC[i];
    }
    for (int e = 0; e < E; ++e) {                                 /* :259-267 */
        d->v[e] = 0.0;
        for (int j = 0; j < c->nEoE[e]; ++j)
            d->v[e] = d->v[e] + WOE(c, j, e) * u[EOE(c, j, e)];
    }
    for (int v = 0; v < V; ++v) {                                 /* :282-292 */
        d->h_vertex[v] = 0.0;
        for (int i = 0; i < c->vdeg; ++i)
            d->h_vertex[v] = d->h_vertex[v] + h[COV(c, i, v)] * KITE(c, i, v);
        d->h_vertex[v] = d->h_vertex[v] / c->areaT[v];
        d->pv_vertex[v] = (c->fV[v] + d->vorticity[v]) / d->h_vertex[v];
    }
    for (int e = 0; e < E; ++e)                                   /* :296-301 */
        d->gradPVt[e] = (d->pv_vertex[VTX2(c, e)] - d->pv_vertex[VTX1(c, e)]) / c->dv[e];
    for (int e = 0; e < E; ++e) d->pv_edge[e] = 0.0;              /* :305-313 */
    for (int v = 0; v < V; ++v)
        for (int i = 0; i < c->vdeg; ++i) {
            int e = EOV(c, i, v);
            d->pv_edge[e] = d->pv_edge[e] + 0.5 * d->pv_vertex[v];
        }
    for (int e = 0; e < E; ++e)                                   /* :316-320 */
        d->pv_edge[e] = d->pv_edge[e] - apvm * d->v[e] * dt * d->gradPVt[e];
    for (int i = 0; i < C; ++i) { d->pv_cell[i] = 0.0; d->vorticity_cell[i] = 0.0; }
    for (int v = 0; v < V; ++v)                                   /* :324-340 */
        for (int i = 0; i < c->vdeg; ++i) {
            int ic = COV(c, i, v);
            d->pv_cell[ic] = d->pv_cell[ic] + KITE(c, i, v) * d->pv_vertex[v] / c->areaC[ic];
            d->vorticity_cell[ic] =
                d->vorticity_cell[ic] + KITE(c, i, v) * d->vorticity[v] / c->areaC[ic];
        }
    for (int e = 0; e < E; ++e)                                   /* :344-352 */
        d->gradPVn[e] = (d->pv_cell[CELL2(c, e)] - d->pv_cell[CELL1(c, e)]) / c->dc[e];
    for (int e = 0; e < E; ++e)                                   /* :355-359 */
        d->pv_edge[e] = d->pv_edge[e] - apvm * u[e] * dt * d->gradPVn[e];
}

/* del4 helper: delsq_u, delsq_vorticity, delsq_divergence of a field pair
 * (src/mpas_sw_dyncore.f90:536-589)                                        */
static void del4_parts(mswo_ctx *c, const double *div, const double *vort, double *dsu,
                       double *dsv, double *dsd)
{
    const int C = c->C, E = c->E, V = c->V;
    for (int e = 0; e < E; ++e)
        dsu[e] = (div[CELL2(c, e)] - div[CELL1(c, e)]) / c->dc[e] -
                 (vort[VTX2(c, e)] - vort[VTX1(c, e)]) / c->dv[e];
    for (int v = 0; v < V; ++v) dsv[v] = 0.0;
    for (int e = 0; e < E; ++e) {
        dsv[VTX1(c, e)] = dsv[VTX1(c, e)] - c->dc[e] * dsu[e];
        dsv[VTX2(c, e)] = dsv[VTX2(c, e)] + c->dc[e] * dsu[e];
    }
    for (int v = 0; v < V; ++v) {
        double r = 1.0 / c->areaT[v];
        dsv[v] = dsv[v] * r;
    }
    for (int i = 0; i < C; ++i) dsd[i] = 0.0;
    for (int e = 0; e < E; ++e) {
        dsd[CELL1(c, e)] = dsd[CELL1(c, e)] + dsu[e] * c->dv[e];
        dsd[CELL2(c, e)] = dsd[CELL2(c, e)] - dsu[e] * c->dv[e];
    }
    for (int i = 0; i < C; ++i) {
        double r = 1.0 / c->areaC[i];
        dsd[i] = dsd[i] * r;
    }
}

/* sw_compute_tend: src/mpas_sw_dyncore.f90:395-639 */
void mswo_tend(mswo_ctx *c, const double *h, const double *u)
{
    const int C = c->C, E = c->E, V = c->V;
    diagset *d = &c->d;
    double *th = c->tend_h, *tu = c->tend_u;
    for (int i = 0; i < C; ++i) th[i] = 0.0;                      /* :468-482 */
    for (int e = 0; e < E; ++e) {
        double flux = u[e] * c->dv[e] * d->h_edge[e];
        th[CELL1(c, e)] = th[CELL1(c, e)] - flux;
        th[CELL2(c, e)] = th[CELL2(c, e)] + flux;
    }
    for (int i = 0; i < C; ++i) th[i] = th[i] / c->areaC[i];
    for (int e = 0; e < E; ++e) {                                 /* :485-505 */
        int c1 = CELL1(c, e), c2 = CELL2(c, e);
        double q = 0.0;
        for (int j = 0; j < c->nEoE[e]; ++j) {
            int eo = EOE(c, j, e);
            double workpv = 0.5 * (d->pv_edge[e] + d->pv_edge[eo]);
            q = q + WOE(c, j, e) * u[eo] * workpv * d->h_edge[eo];
        }
        tu[e] = q - (d->ke[c2] - d->ke[c1] +
                     GRAVITY * (h[c2] + c->hs[c2] - h[c1] - c->hs[c1])) / c->dc[e];
    }
    if (c->cfg.visc2 > 0.0) {                                     /* :510-524 */
        for (int e = 0; e < E; ++e) {
            double ud = (d->divergence[CELL2(c, e)] - d->divergence[CELL1(c, e)]) / c->dc[e] -
                        (d->vorticity[VTX2(c, e)] - d->vorticity[VTX1(c, e)]) / c->dv[e];
            ud = c->msd2[e] * c->cfg.visc2 * ud;
            tu[e] = tu[e] + ud;
        }
    }
    if (c->cfg.visc4 > 0.0) {                                     /* :530-617 */
        double *dsu = zeros(E), *dsv = zeros(V), *dsd = zeros(C);
        if (dsu && dsv && dsd) {
            del4_parts(c, d->divergence, d->vorticity, dsu, dsv, dsd);
            for (int e = 0; e < E; ++e) {
                double ud = (dsd[CELL2(c, e)] - dsd[CELL1(c, e)]) / c->dc[e] -
                            (dsv[VTX2(c, e)] - dsv[VTX1(c, e)]) / c->dv[e];
                ud = c->msd4[e] * c->cfg.visc4 * ud;
                tu[e] = tu[e] - ud;
            }
        }
        free(dsu); free(dsv); free(dsd);
    }
    if (c->cfg.wind_stress)                                       /* :620-625 */
        for (int e = 0; e < E; ++e) tu[e] = tu[e] + c->usrc[e] / RHO_REF / d->h_edge[e];
    if (c->cfg.bottom_drag)                                       /* :627-637 */
        for (int e = 0; e < E; ++e) {
            double ke_edge = 0.5 * (d->ke[CELL1(c, e)] + d->ke[CELL2(c, e)]);
            tu[e] = tu[e] - 1.0e-3 * u[e] * sqrt(2.0 * ke_edge) / d->h_edge[e];
        }
}

/* ---------------------------------------------------------------------- */
/* sw_compute_solve_diagnostics_tlm: src/mpas_sw_dyncore_tlm.f90:1-447 */
void mswo_diag_tlm(mswo_ctx *c, const double *h, const double *h_tl, const double *u,
                   const double *u_tl)
{
    const int C = c->C, E = c->E, V = c->V;
    const double dt = c->cfg.dt, apvm = c->cfg.apvm;
    diagset *d = &c->d, *t = &c->dtl;
    for (int e = 0; e < E; ++e) {                                 /* :117-126 */
        t->h_edge[e] = 0.5 * (h_tl[CELL1(c, e)] + h_tl[CELL2(c, e)]);
        d->h_edge[e] = 0.5 * (h[CELL1(c, e)] + h[CELL2(c, e)]);
    }
    for (int v = 0; v < V; ++v) { t->circulation[v] = 0.0; d->circulation[v] = 0.0; }
    for (int e = 0; e < E; ++e) {                                 /* :255-271 */
        int v1 = VTX1(c, e), v2 = VTX2(c, e);
        t->circulation[v1] = t->circulation[v1] - c->dc[e] * u_tl[e];
        d->circulation[v1] = d->circulation[v1] - c->dc[e] * u[e];
        t->circulation[v2] = t->circulation[v2] + c->dc[e] * u_tl[e];
        d->circulation[v2] = d->circulation[v2] + c->dc[e] * u[e];
    }
    for (int v = 0; v < V; ++v) {
        t->vorticity[v] = t->circulation[v] / c->areaT[v];
        d->vorticity[v] = d->circulation[v] / c->areaT[v];
    }
    for (int i = 0; i < C; ++i) { t->divergence[i] = 0.0; d->divergence[i] = 0.0; }
    for (int e = 0; e < E; ++e) {                                 /* :274-298 */
        int c1 = CELL1(c, e), c2 = CELL2(c, e);
        t->divergence[c1] = t->divergence[c1] + u_tl[e] * c->dv[e];
        d->divergence[c1] = d->divergence[c1] + u[e] * c->dv[e];
        t->divergence[c2] = t->divergence[c2] - u_tl[e] * c->dv[e];
        d->divergence[c2] = d->divergence[c2] - u[e] * c->dv[e];
    }
    for (int i = 0; i < C; ++i) {
        double r = 1.0 / c->areaC[i];
        t->divergence[i] = t->divergence[i] * r;
        d->divergence[i] = d->divergence[i] * r;
    }
    for (int i = 0; i < C; ++i) {                                 /* :301-316 */
        t->ke[i] = 0.0;
        d->ke[i] = 0.0;
        for (int k = 0; k < c->nEoC[i]; ++k) {
            int e = EOC(c, k, i);
            t->ke[i] = t->ke[i] + 0.25 * c->dc[e] * c->dv[e] * (2.0 * u[e] * u_tl[e]);
            d->ke[i] = d->ke[i] + 0.25 * c->dc[e] * c->dv[e] * (u[e] * u[e]);
        }
        t->ke[i] = t->ke[i] / c->areaC[i];
        d->ke[i] = d->ke[i] / c->areaC[i];
    }
    for (int e = 0; e < E; ++e) {                                 /* :319-329 */
        t->v[e] = 0.0;
        d->v[e] = 0.0;
        for (int j = 0; j < c->nEoE[e]; ++j) {
            int eo = EOE(c, j, e);
            t->v[e] = t->v[e] + WOE(c, j, e) * u_tl[eo];
            d->v[e] = d->v[e] + WOE(c, j, e) * u[eo];
        }
    }
    for (int v = 0; v < V; ++v) {                                 /* :344-360 */
        t->h_vertex[v] = 0.0;
        d->h_vertex[v] = 0.0;
        for (int i = 0; i < c->vdeg; ++i) {
            t->h_vertex[v] = t->h_vertex[v] + h_tl[COV(c, i, v)] * KITE(c, i, v);
            d->h_vertex[v] = d->h_vertex[v] + h[COV(c, i, v)] * KITE(c, i, v);
        }
        t->h_vertex[v] = t->h_vertex[v] / c->areaT[v];
        d->h_vertex[v] = d->h_vertex[v] / c->areaT[v];
        double hv = d->h_vertex[v];
        t->pv_vertex[v] = t->vorticity[v] / hv -
                          t->h_vertex[v] * (c->fV[v] + d->vorticity[v]) / (hv * hv);
        d->pv_vertex[v] = (c->fV[v] + d->vorticity[v]) / hv;
    }
    for (int e = 0; e < E; ++e) {                                 /* :364-369 */
        t->gradPVt[e] = (t->pv_vertex[VTX2(c, e)] - t->pv_vertex[VTX1(c, e)]) / c->dv[e];
        d->gradPVt[e] = (d->pv_vertex[VTX2(c, e)] - d->pv_vertex[VTX1(c, e)]) / c->dv[e];
    }
    for (int e = 0; e < E; ++e) { t->pv_edge[e] = 0.0; d->pv_edge[e] = 0.0; }
    for (int v = 0; v < V; ++v)                                   /* :375-383 */
        for (int i = 0; i < c->vdeg; ++i) {
            int e = EOV(c, i, v);
            t->pv_edge[e] = t->pv_edge[e] + 0.5 * t->pv_vertex[v];
            d->pv_edge[e] = d->pv_edge[e] + 0.5 * d->pv_vertex[v];
        }
    for (int e = 0; e < E; ++e) {                                 /* :386-393 */
        t->pv_edge[e] = t->pv_edge[e] - apvm * t->v[e] * dt * d->gradPVt[e] -
                        apvm * d->v[e] * dt * t->gradPVt[e];
        d->pv_edge[e] = d->pv_edge[e] - apvm * d->v[e] * dt * d->gradPVt[e];
    }
    for (int i = 0; i < C; ++i) {
        t->pv_cell[i] = 0.0; d->pv_cell[i] = 0.0;
        t->vorticity_cell[i] = 0.0; d->vorticity_cell[i] = 0.0;
    }
    for (int v = 0; v < V; ++v)                                   /* :401-421 */
        for (int i = 0; i < c->vdeg; ++i) {
            int ic = COV(c, i, v);
            double k = KITE(c, i, v), A = c->areaC[ic];
            t->pv_cell[ic] = t->pv_cell[ic] + k * t->pv_vertex[v] / A;
            d->pv_cell[ic] = d->pv_cell[ic] + k * d->pv_vertex[v] / A;
            t->vorticity_cell[ic] = t->vorticity_cell[ic] + k * t->vorticity[v] / A;
            d->vorticity_cell[ic] = d->vorticity_cell[ic] + k * d->vorticity[v] / A;
        }
    for (int e = 0; e < E; ++e) {                                 /* :425-434 */
        t->gradPVn[e] = (t->pv_cell[CELL2(c, e)] - t->pv_cell[CELL1(c, e)]) / c->dc[e];
        d->gradPVn[e] = (d->pv_cell[CELL2(c, e)] - d->pv_cell[CELL1(c, e)]) / c->dc[e];
    }
    for (int e = 0; e < E; ++e) {                                 /* :437-445 */
        t->pv_edge[e] = t->pv_edge[e] - apvm * u_tl[e] * dt * d->gradPVn[e] -
                        apvm * u[e] * dt * t->gradPVn[e];
        d->pv_edge[e] = d->pv_edge[e] - apvm * u[e] * dt * d->gradPVn[e];
    }
}

/* sw_compute_tend_tlm: src/mpas_sw_dyncore_tlm.f90:449-770 */
void mswo_tend_tlm(mswo_ctx *c, const double *h, const double *h_tl, const double *u,
                   const double *u_tl)
{
    const int C = c->C, E = c->E, V = c->V;
    diagset *d = &c->d, *t = &c->dtl;
    double *th = c->tend_h, *tu = c->tend_u, *tht = c->tend_h_tl, *tut = c->tend_u_tl;
    for (int i = 0; i < C; ++i) { tht[i] = 0.0; th[i] = 0.0; }    /* :541-561 */
    for (int e = 0; e < E; ++e) {
        int c1 = CELL1(c, e), c2 = CELL2(c, e);
        double flux_tl = u_tl[e] * c->dv[e] * d->h_edge[e] + u[e] * c->dv[e] * t->h_edge[e];
        double flux = u[e] * c->dv[e] * d->h_edge[e];
        tht[c1] = tht[c1] - flux_tl;
        th[c1] = th[c1] - flux;
        tht[c2] = tht[c2] + flux_tl;
        th[c2] = th[c2] + flux;
    }
    for (int i = 0; i < C; ++i) {
        tht[i] = tht[i] / c->areaC[i];
        th[i] = th[i] / c->areaC[i];
    }
    for (int e = 0; e < E; ++e) {                                 /* :564-595 */
        int c1 = CELL1(c, e), c2 = CELL2(c, e);
        double q_tl = 0.0, q = 0.0;
        for (int j = 0; j < c->nEoE[e]; ++j) {
            int eo = EOE(c, j, e);
            double w = WOE(c, j, e);
            double workpv_tl = 0.5 * (t->pv_edge[e] + t->pv_edge[eo]);
            double workpv = 0.5 * (d->pv_edge[e] + d->pv_edge[eo]);
            q_tl = q_tl + w * u_tl[eo] * workpv * d->h_edge[eo] +
                   w * u[eo] * workpv_tl * d->h_edge[eo] + w * u[eo] * workpv * t->h_edge[eo];
            q = q + w * u[eo] * workpv * d->h_edge[eo];
        }
        tut[e] = q_tl - (t->ke[c2] - t->ke[c1] + GRAVITY * (h_tl[c2] - h_tl[c1])) / c->dc[e];
        tu[e] = q - (d->ke[c2] - d->ke[c1] +
                     GRAVITY * (h[c2] + c->hs[c2] - h[c1] - c->hs[c1])) / c->dc[e];
    }
    if (c->cfg.visc2 > 0.0) {                                     /* :600-618 */
        for (int e = 0; e < E; ++e) {
            int c1 = CELL1(c, e), c2 = CELL2(c, e), v1 = VTX1(c, e), v2 = VTX2(c, e);
            double ud_tl = (t->divergence[c2] - t->divergence[c1]) / c->dc[e] -
                           (t->vorticity[v2] - t->vorticity[v1]) / c->dv[e];
            double ud = (d->divergence[c2] - d->divergence[c1]) / c->dc[e] -
                        (d->vorticity[v2] - d->vorticity[v1]) / c->dv[e];
            ud_tl = c->msd2[e] * c->cfg.visc2 * ud_tl;
            ud = c->msd2[e] * c->cfg.visc2 * ud;
            tut[e] = tut[e] + ud_tl;
            tu[e] = tu[e] + ud;
        }
    }
    if (c->cfg.visc4 > 0.0) {                                     /* :624-737 */
        double *a = zeros(E), *b = zeros(V), *cc = zeros(C);
        double *at = zeros(E), *bt = zeros(V), *cct = zeros(C);
        if (a && b && cc && at && bt && cct) {
            del4_parts(c, t->divergence, t->vorticity, at, bt, cct);
            del4_parts(c, d->divergence, d->vorticity, a, b, cc);
            for (int e = 0; e < E; ++e) {
                int c1 = CELL1(c, e), c2 = CELL2(c, e), v1 = VTX1(c, e), v2 = VTX2(c, e);
                double ud_tl = (cct[c2] - cct[c1]) / c->dc[e] - (bt[v2] - bt[v1]) / c->dv[e];
                double ud = (cc[c2] - cc[c1]) / c->dc[e] - (b[v2] - b[v1]) / c->dv[e];
                ud_tl = c->msd4[e] * c->cfg.visc4 * ud_tl;
                ud = c->msd4[e] * c->cfg.visc4 * ud;
                tut[e] = tut[e] - ud_tl;
                tu[e] = tu[e] - ud;
            }
        }
        free(a); free(b); free(cc); free(at); free(bt); free(cct);
    }
    if (c->cfg.wind_stress)                                       /* :740-748 */
        for (int e = 0; e < E; ++e) {
            double he = d->h_edge[e];
            tut[e] = tut[e] - t->h_edge[e] * c->usrc[e] / RHO_REF / (he * he);
            tu[e] = tu[e] + c->usrc[e] / RHO_REF / he;
        }
    if (c->cfg.bottom_drag)                                       /* :750-767 */
        for (int e = 0; e < E; ++e) {
            int c1 = CELL1(c, e), c2 = CELL2(c, e);
            double ke_edge_tl = 0.5 * (t->ke[c1] + t->ke[c2]);
            double ke_edge = 0.5 * (d->ke[c1] + d->ke[c2]);
            double he = d->h_edge[e];
            tut[e] = tut[e] - 1.0e-3 * u_tl[e] * sqrt(2.0 * ke_edge) / he -
                     0.5 * 1.0e-3 * u[e] * (2.0 * ke_edge_tl) / sqrt(2.0 * ke_edge) / he +
                     1.0e-3 * u[e] * sqrt(2.0 * ke_edge) * t->h_edge[e] / (he * he);
            tu[e] = tu[e] - 1.0e-3 * u[e] * sqrt(2.0 * ke_edge) / he;
        }
}

/* ---------------------------------------------------------------------- */
/* sw_compute_solve_diagnostics_adj: src/mpas_sw_dyncore_adj.f90:1-453.
 * Recomputes the nonlinear fields it needs into locals (:126-190), then the
 * reverse sweep; every diagnostic adjoint it consumes is zeroed.           */
void mswo_diag_adj(mswo_ctx *c, const double *h, double *h_ad, const double *u, double *u_ad)
{
    const int C = c->C, E = c->E, V = c->V;
    const double dt = c->cfg.dt, apvm = c->cfg.apvm;
    diagset *a = &c->dad;
    double *circ = zeros(V), *vort = zeros(V), *vv = zeros(E), *hv = zeros(V);
    double *pvv = zeros(V), *gpt = zeros(E), *pc = zeros(C), *gpn = zeros(E);
    if (!(circ && vort && vv && hv && pvv && gpt && pc && gpn)) goto out;
    for (int e = 0; e < E; ++e) {                                 /* :127-138 */
        circ[VTX1(c, e)] = circ[VTX1(c, e)] - c->dc[e] * u[e];
        circ[VTX2(c, e)] = circ[VTX2(c, e)] + c->dc[e] * u[e];
    }
    for (int v = 0; v < V; ++v) vort[v] = circ[v] / c->areaT[v];
    for (int e = 0; e < E; ++e) {                                 /* :140-148 */
        vv[e] = 0.0;
        for (int j = 0; j < c->nEoE[e]; ++j) vv[e] = vv[e] + WOE(c, j, e) * u[EOE(c, j, e)];
    }
    for (int v = 0; v < V; ++v) {                                 /* :150-159 */
        hv[v] = 0.0;
        for (int i = 0; i < c->vdeg; ++i) hv[v] = hv[v] + h[COV(c, i, v)] * KITE(c, i, v);
        hv[v] = hv[v] / c->areaT[v];
        pvv[v] = (c->fV[v] + vort[v]) / hv[v];
    }
    for (int e = 0; e < E; ++e)                                   /* :161-165 */
        gpt[e] = (pvv[VTX2(c, e)] - pvv[VTX1(c, e)]) / c->dv[e];
    for (int v = 0; v < V; ++v)                                   /* :167-179 */
        for (int i = 0; i < c->vdeg; ++i) {
            int ic = COV(c, i, v);
            pc[ic] = pc[ic] + KITE(c, i, v) * pvv[v] / c->areaC[ic];
        }
    for (int e = 0; e < E; ++e)                                   /* :181-188 */
        gpn[e] = (pc[CELL2(c, e)] - pc[CELL1(c, e)]) / c->dc[e];

    for (int e = E - 1; e >= 0; --e) {                            /* :193-200 */
        a->gradPVn[e] = a->gradPVn[e] - apvm * u[e] * dt * a->pv_edge[e];
        u_ad[e] = u_ad[e] - apvm * dt * gpn[e] * a->pv_edge[e];
    }
    for (int e = E - 1; e >= 0; --e) {                            /* :202-213 */
        int c1 = CELL1(c, e), c2 = CELL2(c, e);
        a->pv_cell[c1] = a->pv_cell[c1] - a->gradPVn[e] / c->dc[e];
        a->pv_cell[c2] = a->pv_cell[c2] + a->gradPVn[e] / c->dc[e];
        a->gradPVn[e] = 0.0;
    }
    for (int v = V - 1; v >= 0; --v)                              /* :215-229 */
        for (int i = c->vdeg - 1; i >= 0; --i) {
            int ic = COV(c, i, v);
            a->vorticity[v] =
                a->vorticity[v] + KITE(c, i, v) / c->areaC[ic] * a->vorticity_cell[ic];
            a->pv_vertex[v] = a->pv_vertex[v] + KITE(c, i, v) / c->areaC[ic] * a->pv_cell[ic];
        }
    for (int i = 0; i < C; ++i) { a->vorticity_cell[i] = 0.0; a->pv_cell[i] = 0.0; }
    for (int e = E - 1; e >= 0; --e) {                            /* :235-242 */
        a->gradPVt[e] = a->gradPVt[e] - apvm * vv[e] * dt * a->pv_edge[e];
        a->v[e] = a->v[e] - apvm * dt * gpt[e] * a->pv_edge[e];
    }
    for (int v = V - 1; v >= 0; --v)                              /* :244-252 */
        for (int i = c->vdeg - 1; i >= 0; --i) {
            int e = EOV(c, i, v);
            a->pv_vertex[v] = a->pv_vertex[v] + 0.5 * a->pv_edge[e];
        }
    for (int e = 0; e < E; ++e) a->pv_edge[e] = 0.0;
    for (int e = E - 1; e >= 0; --e) {                            /* :254-262 */
        int v1 = VTX1(c, e), v2 = VTX2(c, e);
        a->pv_vertex[v1] = a->pv_vertex[v1] - a->gradPVt[e] / c->dv[e];
        a->pv_vertex[v2] = a->pv_vertex[v2] + a->gradPVt[e] / c->dv[e];
        a->gradPVt[e] = 0.0;
    }
    for (int v = V - 1; v >= 0; --v) {                            /* :268-292 */
        a->h_vertex[v] = a->h_vertex[v] -
                         (c->fV[v] + vort[v]) / (hv[v] * hv[v]) * a->pv_vertex[v];
        a->vorticity[v] = a->vorticity[v] + a->pv_vertex[v] / hv[v];
        a->pv_vertex[v] = 0.0;
        a->h_vertex[v] = a->h_vertex[v] / c->areaT[v];
        for (int i = c->vdeg - 1; i >= 0; --i) {
            int ic = COV(c, i, v);
            h_ad[ic] = h_ad[ic] + KITE(c, i, v) * a->h_vertex[v];
        }
        a->h_vertex[v] = 0.0;
    }
    for (int e = E - 1; e >= 0; --e)                              /* :296-304 */
        for (int j = c->nEoE[e] - 1; j >= 0; --j) {
            int eo = EOE(c, j, e);
            u_ad[eo] = u_ad[eo] + WOE(c, j, e) * a->v[e];
        }
    for (int e = 0; e < E; ++e) a->v[e] = 0.0;
    for (int i = C - 1; i >= 0; --i) {                            /* :306-318 */
        a->ke[i] = a->ke[i] / c->areaC[i];
        for (int k = c->nEoC[i] - 1; k >= 0; --k) {
            int e = EOC(c, k, i);
            u_ad[e] = u_ad[e] + 0.25 * c->dc[e] * c->dv[e] * 2.0 * u[e] * a->ke[i];
        }
    }
    for (int i = 0; i < C; ++i) a->ke[i] = 0.0;
    for (int i = C - 1; i >= 0; --i) {                            /* :321-341 */
        double r = 1.0 / c->areaC[i];
        a->divergence[i] = a->divergence[i] * r;
    }
    for (int e = E - 1; e >= 0; --e) {
        int c1 = CELL1(c, e), c2 = CELL2(c, e);
        u_ad[e] = u_ad[e] - c->dv[e] * a->divergence[c2];
        u_ad[e] = u_ad[e] + c->dv[e] * a->divergence[c1];
    }
    for (int i = 0; i < C; ++i) a->divergence[i] = 0.0;
    for (int v = V - 1; v >= 0; --v) {                            /* :343-356 */
        a->circulation[v] = a->circulation[v] + a->vorticity[v] / c->areaT[v];
        a->vorticity[v] = 0.0;
    }
    for (int e = E - 1; e >= 0; --e) {
        u_ad[e] = u_ad[e] + c->dc[e] * a->circulation[VTX2(c, e)];
        u_ad[e] = u_ad[e] - c->dc[e] * a->circulation[VTX1(c, e)];
    }
    for (int v = 0; v < V; ++v) a->circulation[v] = 0.0;
    for (int e = E - 1; e >= 0; --e) {                            /* :358-369 */
        int c1 = CELL1(c, e), c2 = CELL2(c, e);
        h_ad[c2] = h_ad[c2] + 0.5 * a->h_edge[e];
        h_ad[c1] = h_ad[c1] + 0.5 * a->h_edge[e];
        a->h_edge[e] = 0.0;
    }
out:
    free(circ); free(vort); free(vv); free(hv); free(pvv); free(gpt); free(pc); free(gpn);
}

/* sw_compute_tend_adj: src/mpas_sw_dyncore_adj.f90:455-738.  Needs the
 * forward h_edge, pv_edge and ke of the same state in the context's diag. */
void mswo_tend_adj(mswo_ctx *c, const double *h, double *h_ad, const double *u, double *u_ad)
{
    const int C = c->C, E = c->E, V = c->V;
    diagset *d = &c->d, *a = &c->dad;
    double *tha = c->tend_h_ad, *tua = c->tend_u_ad;
    if (c->cfg.bottom_drag)                                       /* :549-565 */
        for (int e = E - 1; e >= 0; --e) {
            int c1 = CELL1(c, e), c2 = CELL2(c, e);
            double ke_edge = 0.5 * (d->ke[c1] + d->ke[c2]);
            double he = d->h_edge[e];
            a->h_edge[e] = a->h_edge[e] + 1.0e-3 * u[e] * sqrt(2.0 * ke_edge) / (he * he) * tua[e];
            double ke_edge_ad = -0.5 * 1.0e-3 * u[e] * 2.0 / sqrt(2.0 * ke_edge) / he * tua[e];
            u_ad[e] = u_ad[e] - 1.0e-3 * sqrt(2.0 * ke_edge) / he * tua[e];
            a->ke[c2] = a->ke[c2] + 0.5 * ke_edge_ad;
            a->ke[c1] = a->ke[c1] + 0.5 * ke_edge_ad;
        }
    if (c->cfg.wind_stress)                                       /* :567-572 */
        for (int e = E - 1; e >= 0; --e) {
            double he = d->h_edge[e];
            a->h_edge[e] = a->h_edge[e] - c->usrc[e] / RHO_REF / (he * he) * tua[e];
        }
    if (c->cfg.visc4 > 0.0) {                                     /* :574-669 */
        double *dda = zeros(C), *dua = zeros(E), *dca = zeros(V), *dva = zeros(V);
        if (dda && dua && dca && dva) {
            for (int e = E - 1; e >= 0; --e) {
                int c1 = CELL1(c, e), c2 = CELL2(c, e), v1 = VTX1(c, e), v2 = VTX2(c, e);
                double uda = -tua[e];
                uda = c->msd4[e] * c->cfg.visc4 * uda;
                dva[v1] = dva[v1] + uda / c->dv[e];
                dva[v2] = dva[v2] - uda / c->dv[e];
                dda[c1] = dda[c1] - uda / c->dc[e];
                dda[c2] = dda[c2] + uda / c->dc[e];
            }
            for (int i = C - 1; i >= 0; --i) {
                double r = 1.0 / c->areaC[i];
                dda[i] = dda[i] * r;
            }
            for (int e = E - 1; e >= 0; --e) {
                int c1 = CELL1(c, e), c2 = CELL2(c, e);
                dua[e] = dua[e] - c->dv[e] * dda[c2];
                dua[e] = dua[e] + c->dv[e] * dda[c1];
            }
            for (int v = V - 1; v >= 0; --v) {
                double r = 1.0 / c->areaT[v];
                dca[v] = dca[v] + dva[v] * r;
                dva[v] = 0.0;
            }
            for (int e = E - 1; e >= 0; --e) {
                dua[e] = dua[e] + c->dc[e] * dca[VTX2(c, e)];
                dua[e] = dua[e] - c->dc[e] * dca[VTX1(c, e)];
            }
            for (int e = E - 1; e >= 0; --e) {
                int c1 = CELL1(c, e), c2 = CELL2(c, e), v1 = VTX1(c, e), v2 = VTX2(c, e);
                a->vorticity[v1] = a->vorticity[v1] + dua[e] / c->dv[e];
                a->vorticity[v2] = a->vorticity[v2] - dua[e] / c->dv[e];
                a->divergence[c1] = a->divergence[c1] - dua[e] / c->dc[e];
                a->divergence[c2] = a->divergence[c2] + dua[e] / c->dc[e];
            }
        }
        free(dda); free(dua); free(dca); free(dva);
    }
    if (c->cfg.visc2 > 0.0)                                       /* :671-687 */
        for (int e = 0; e < E; ++e) {
            int c1 = CELL1(c, e), c2 = CELL2(c, e), v1 = VTX1(c, e), v2 = VTX2(c, e);
            double uda = tua[e];
            uda = c->msd2[e] * c->cfg.visc2 * uda;
            a->vorticity[v1] = a->vorticity[v1] + uda / c->dv[e];
            a->vorticity[v2] = a->vorticity[v2] - uda / c->dv[e];
            a->divergence[c1] = a->divergence[c1] - uda / c->dc[e];
            a->divergence[c2] = a->divergence[c2] + uda / c->dc[e];
        }
    for (int e = E - 1; e >= 0; --e) {                            /* :689-717 */
        int c1 = CELL1(c, e), c2 = CELL2(c, e);
        h_ad[c1] = h_ad[c1] + GRAVITY / c->dc[e] * tua[e];
        h_ad[c2] = h_ad[c2] - GRAVITY / c->dc[e] * tua[e];
        a->ke[c1] = a->ke[c1] + tua[e] / c->dc[e];
        a->ke[c2] = a->ke[c2] - tua[e] / c->dc[e];
        double q_ad = tua[e];
        tua[e] = 0.0;
        for (int j = c->nEoE[e] - 1; j >= 0; --j) {
            int eo = EOE(c, j, e);
            double w = WOE(c, j, e);
            double workpv = 0.5 * (d->pv_edge[e] + d->pv_edge[eo]);
            a->h_edge[eo] = a->h_edge[eo] + w * u[eo] * workpv * q_ad;
            double workpv_ad = w * u[eo] * d->h_edge[eo] * q_ad;
            u_ad[eo] = u_ad[eo] + w * workpv * d->h_edge[eo] * q_ad;
            a->pv_edge[eo] = a->pv_edge[eo] + 0.5 * workpv_ad;
            a->pv_edge[e] = a->pv_edge[e] + 0.5 * workpv_ad;
        }
    }
    for (int e = 0; e < E; ++e) tua[e] = 0.0;
    for (int i = C - 1; i >= 0; --i) tha[i] = tha[i] / c->areaC[i]; /* :720-736 */
    for (int e = 0; e < E; ++e) {
        int c1 = CELL1(c, e), c2 = CELL2(c, e);
        double flux_ad = tha[c2];
        flux_ad = flux_ad - tha[c1];
        a->h_edge[e] = a->h_edge[e] + u[e] * c->dv[e] * flux_ad;
        u_ad[e] = u_ad[e] + c->dv[e] * d->h_edge[e] * flux_ad;
    }
    for (int i = 0; i < C; ++i) tha[i] = 0.0;
    (void)h;
}

/* ---------------------------------------------------------------------- */
/* mpas_reconstruct_2d: src/mpas_sw_operators.f90:283-370 (on a sphere) */
static void recon_core(mswo_ctx *c, const double *u, reconset *r)
{
    for (int i = 0; i < c->C; ++i) {
        r->X[i] = 0.0; r->Y[i] = 0.0; r->Z[i] = 0.0;
        for (int k = 0; k < c->nEoC[i]; ++k) {
            int e = EOC(c, k, i);
            const double *cf = c->coeffs + ((size_t)i * c->mE + k) * 3;
            r->X[i] = r->X[i] + cf[0] * u[e];
            r->Y[i] = r->Y[i] + cf[1] * u[e];
            r->Z[i] = r->Z[i] + cf[2] * u[e];
        }
    }
    for (int i = 0; i < c->C; ++i) {
        double clat, slat, clon, slon;  /* the reference's compiler emits sincos */
        sincos(c->latC[i], &slat, &clat);
        sincos(c->lonC[i], &slon, &clon);
        r->Zonal[i] = -r->X[i] * slon + r->Y[i] * clon;
        r->Meridional[i] = -(r->X[i] * clon + r->Y[i] * slon) * slat + r->Z[i] * clat;
    }
}

void mswo_reconstruct(mswo_ctx *c, const double *u) { recon_core(c, u, &c->r); }

/* mpas_reconstruct_2d_tlm: src/mpas_sw_dyncore_tlm.f90:772-870 */
void mswo_reconstruct_tlm(mswo_ctx *c, const double *u, const double *u_tl)
{
    recon_core(c, u_tl, &c->rtl);
    recon_core(c, u, &c->r);
}

/* mpas_reconstruct_2d_adj: src/mpas_sw_dyncore_adj.f90:740-831 */
void mswo_reconstruct_adj(mswo_ctx *c, const double *u, double *u_ad)
{
    reconset *a = &c->rad;
    for (int i = c->C - 1; i >= 0; --i) {
        double clat, slat, clon, slon;  /* the reference's compiler emits sincos */
        sincos(c->latC[i], &slat, &clat);
        sincos(c->lonC[i], &slon, &clon);
        a->Z[i] = a->Z[i] + clat * a->Meridional[i];
        a->Y[i] = a->Y[i] - slon * slat * a->Meridional[i];
        a->X[i] = a->X[i] - clon * slat * a->Meridional[i];
        a->Meridional[i] = 0.0;
        a->Y[i] = a->Y[i] + clon * a->Zonal[i];
        a->X[i] = a->X[i] - slon * a->Zonal[i];
        a->Zonal[i] = 0.0;
    }
    for (int i = c->C - 1; i >= 0; --i) {
        for (int k = c->nEoC[i] - 1; k >= 0; --k) {
            int e = EOC(c, k, i);
            const double *cf = c->coeffs + ((size_t)i * c->mE + k) * 3;
            u_ad[e] = u_ad[e] + cf[2] * a->Z[i];
            u_ad[e] = u_ad[e] + cf[1] * a->Y[i];
            u_ad[e] = u_ad[e] + cf[0] * a->X[i];
        }
        a->X[i] = 0.0; a->Y[i] = 0.0; a->Z[i] = 0.0;
    }
    (void)u;
}

/* ---------------------------------------------------------------------- */
/* sw_mpas_init_block, the parts on the step path: src/mpas_sw_module.f90:495-526 */
void mswo_init_block(mswo_ctx *c)
{
    mswo_diag(c, c->h[0], c->u[0]);
    mswo_reconstruct(c, c->u[0]);
}

static void rk_coeffs(double dt, double w[4], double sw[4])
{
    w[0] = dt / 6.0; w[1] = dt / 3.0; w[2] = dt / 3.0; w[3] = dt / 6.0;
    sw[0] = dt / 2.0; sw[1] = dt / 2.0; sw[2] = dt; sw[3] = 0.0;
}

/* sw_rk4: src/mpas_sw_module.f90:528-635 */
void mswo_rk4(mswo_ctx *c)
{
    const int C = c->C, E = c->E;
    double w[4], sw[4];
    double *up = (double *)malloc(E * sizeof(double)), *hp = (double *)malloc(C * sizeof(double));
    if (!up || !hp) { free(up); free(hp); return; }
    mswo_diag(c, c->h[0], c->u[0]);                               /* :543-556 */
    mswo_reconstruct(c, c->u[0]);
    memcpy(up, c->u[0], E * sizeof(double));
    memcpy(hp, c->h[0], C * sizeof(double));
    memcpy(c->u[1], c->u[0], E * sizeof(double));
    memcpy(c->h[1], c->h[0], C * sizeof(double));
    rk_coeffs(c->cfg.dt, w, sw);
    for (int s = 0; s < 4; ++s) {                                 /* :577-613 */
        mswo_tend(c, hp, up);
        if (s < 3) {
            for (int e = 0; e < E; ++e) up[e] = c->u[0][e] + sw[s] * c->tend_u[e];
            for (int i = 0; i < C; ++i) hp[i] = c->h[0][i] + sw[s] * c->tend_h[i];
        }
        if (c->cfg.test_case == 1) memcpy(up, c->u[0], E * sizeof(double));
        mswo_diag(c, hp, up);
        for (int e = 0; e < E; ++e) c->u[1][e] = c->u[1][e] + w[s] * c->tend_u[e];
        for (int i = 0; i < C; ++i) c->h[1][i] = c->h[1][i] + w[s] * c->tend_h[i];
    }
    if (c->cfg.test_case == 1) memcpy(c->u[1], c->u[0], E * sizeof(double));
    mswo_diag(c, c->h[1], c->u[1]);                               /* :616-629 */
    mswo_reconstruct(c, c->u[1]);
    memcpy(c->u[0], c->u[1], E * sizeof(double));
    memcpy(c->h[0], c->h[1], C * sizeof(double));
    free(up); free(hp);
}

/* sw_rk4_tlm: src/mpas_sw_module.f90:637-781 (also advances the nonlinear
 * state, :774-777) */
void mswo_rk4_tlm(mswo_ctx *c)
{
    const int C = c->C, E = c->E;
    double w[4], sw[4];
    size_t eb = E * sizeof(double), cb = C * sizeof(double);
    double *up = malloc(eb), *upt = malloc(eb), *hp = malloc(cb), *hpt = malloc(cb);
    if (!up || !upt || !hp || !hpt) goto out;
    mswo_diag_tlm(c, c->h[0], c->h_tl[0], c->u[0], c->u_tl[0]);   /* :655-673 */
    mswo_reconstruct_tlm(c, c->u[0], c->u_tl[0]);
    memcpy(upt, c->u_tl[0], eb); memcpy(up, c->u[0], eb);
    memcpy(hpt, c->h_tl[0], cb); memcpy(hp, c->h[0], cb);
    memcpy(c->u_tl[1], c->u_tl[0], eb); memcpy(c->u[1], c->u[0], eb);
    memcpy(c->h_tl[1], c->h_tl[0], cb); memcpy(c->h[1], c->h[0], cb);
    rk_coeffs(c->cfg.dt, w, sw);
    for (int s = 0; s < 4; ++s) {                                 /* :700-747 */
        mswo_tend_tlm(c, hp, hpt, up, upt);
        if (s < 3) {
            for (int e = 0; e < E; ++e) upt[e] = c->u_tl[0][e] + sw[s] * c->tend_u_tl[e];
            for (int e = 0; e < E; ++e) up[e] = c->u[0][e] + sw[s] * c->tend_u[e];
            for (int i = 0; i < C; ++i) hpt[i] = c->h_tl[0][i] + sw[s] * c->tend_h_tl[i];
            for (int i = 0; i < C; ++i) hp[i] = c->h[0][i] + sw[s] * c->tend_h[i];
        }
        if (c->cfg.test_case == 1) { memcpy(upt, c->u_tl[0], eb); memcpy(up, c->u[0], eb); }
        mswo_diag_tlm(c, hp, hpt, up, upt);
        for (int e = 0; e < E; ++e) c->u_tl[1][e] = c->u_tl[1][e] + w[s] * c->tend_u_tl[e];
        for (int e = 0; e < E; ++e) c->u[1][e] = c->u[1][e] + w[s] * c->tend_u[e];
        for (int i = 0; i < C; ++i) c->h_tl[1][i] = c->h_tl[1][i] + w[s] * c->tend_h_tl[i];
        for (int i = 0; i < C; ++i) c->h[1][i] = c->h[1][i] + w[s] * c->tend_h[i];
    }
    if (c->cfg.test_case == 1) { memcpy(c->u_tl[1], c->u_tl[0], eb); memcpy(c->u[1], c->u[0], eb); }
    mswo_diag_tlm(c, c->h[1], c->h_tl[1], c->u[1], c->u_tl[1]);   /* :754-772 */
    mswo_reconstruct_tlm(c, c->u[1], c->u_tl[1]);
    memcpy(c->u_tl[0], c->u_tl[1], eb); memcpy(c->u[0], c->u[1], eb);
    memcpy(c->h_tl[0], c->h_tl[1], cb); memcpy(c->h[0], c->h[1], cb);
out:
    free(up); free(upt); free(hp); free(hpt);
}

/* sw_rk4_adj: src/mpas_sw_module.f90:783-1002.  On entry h1,u1 hold the
 * nonlinear state at the step start and h_ad1,u_ad1 the adjoint at the step
 * end; on exit h_ad1,u_ad1 hold the adjoint at the step start, h2,u2 the
 * nonlinear end state, h1,u1 are unchanged. */
int mswo_rk4_adj(mswo_ctx *c)
{
    const int C = c->C, E = c->E;
    double w[4], sw[4];
    size_t eb = E * sizeof(double), cb = C * sizeof(double);
    double *up = malloc(5 * eb), *hp = malloc(5 * cb);
    double *upa = calloc(E, sizeof(double)), *hpa = calloc(C, sizeof(double));
    if (!up || !hp || !upa || !hpa) { free(up); free(hp); free(upa); free(hpa); return -1; }
#define UP(k) (up + (size_t)(k) * E)
#define HP(k) (hp + (size_t)(k) * C)
    rk_coeffs(c->cfg.dt, w, sw);
    mswo_diag(c, c->h[0], c->u[0]);                               /* :815-828 */
    mswo_reconstruct(c, c->u[0]);
    memcpy(UP(0), c->u[0], eb); memcpy(HP(0), c->h[0], cb);
    memcpy(c->u[1], c->u[0], eb); memcpy(c->h[1], c->h[0], cb);
    for (int s = 0; s < 4; ++s) {                                 /* :838-878 */
        mswo_tend(c, HP(s), UP(s));
        if (s < 3) {
            for (int e = 0; e < E; ++e) UP(s + 1)[e] = c->u[0][e] + sw[s] * c->tend_u[e];
            for (int i = 0; i < C; ++i) HP(s + 1)[i] = c->h[0][i] + sw[s] * c->tend_h[i];
        } else {
            memcpy(UP(s + 1), UP(s), eb); memcpy(HP(s + 1), HP(s), cb);
        }
        if (c->cfg.test_case == 1) memcpy(UP(s + 1), c->u[0], eb);
        mswo_diag(c, HP(s + 1), UP(s + 1));
        for (int e = 0; e < E; ++e) c->u[1][e] = c->u[1][e] + w[s] * c->tend_u[e];
        for (int i = 0; i < C; ++i) c->h[1][i] = c->h[1][i] + w[s] * c->tend_h[i];
    }
    memcpy(c->h_ad[1], c->h_ad[0], cb);                           /* :882-885 */
    memset(c->h_ad[0], 0, cb);
    memcpy(c->u_ad[1], c->u_ad[0], eb);
    memset(c->u_ad[0], 0, eb);
    mswo_reconstruct_adj(c, c->u[1], c->u_ad[1]);                 /* :887-905 */
    mswo_diag_adj(c, c->h[1], c->h_ad[1], c->u[1], c->u_ad[1]);
    if (c->cfg.test_case == 1) {                                  /* :907-910 */
        for (int e = 0; e < E; ++e) c->u_ad[0][e] = c->u_ad[0][e] + c->u_ad[1][e];
        memset(c->u_ad[1], 0, eb);
    }
    for (int s = 3; s >= 0; --s) {                                /* :912-970 */
        for (int i = 0; i < C; ++i) c->tend_h_ad[i] = c->tend_h_ad[i] + w[s] * c->h_ad[1][i];
        for (int e = 0; e < E; ++e) c->tend_u_ad[e] = c->tend_u_ad[e] + w[s] * c->u_ad[1][e];
        mswo_diag_adj(c, HP(s + 1), hpa, UP(s + 1), upa);
        if (c->cfg.test_case == 1) {
            for (int e = 0; e < E; ++e) { c->u_ad[0][e] = c->u_ad[0][e] + upa[e]; upa[e] = 0.0; }
        }
        if (s < 3) {
            for (int i = 0; i < C; ++i) {
                c->tend_h_ad[i] = c->tend_h_ad[i] + sw[s] * hpa[i];
                c->h_ad[0][i] = c->h_ad[0][i] + hpa[i];
                hpa[i] = 0.0;
            }
            for (int e = 0; e < E; ++e) {
                c->tend_u_ad[e] = c->tend_u_ad[e] + sw[s] * upa[e];
                c->u_ad[0][e] = c->u_ad[0][e] + upa[e];
                upa[e] = 0.0;
            }
        }
        mswo_diag(c, HP(s), UP(s));
        mswo_tend_adj(c, HP(s), hpa, UP(s), upa);
    }
    for (int i = 0; i < C; ++i) { c->h_ad[0][i] = c->h_ad[0][i] + c->h_ad[1][i]; c->h_ad[1][i] = 0.0; }
    for (int e = 0; e < E; ++e) { c->u_ad[0][e] = c->u_ad[0][e] + c->u_ad[1][e]; c->u_ad[1][e] = 0.0; }
    for (int i = 0; i < C; ++i) { c->h_ad[0][i] = c->h_ad[0][i] + hpa[i]; hpa[i] = 0.0; }
    for (int e = 0; e < E; ++e) { c->u_ad[0][e] = c->u_ad[0][e] + upa[e]; upa[e] = 0.0; }
    mswo_reconstruct_adj(c, c->u[0], c->u_ad[0]);                 /* :982-1000 */
    mswo_diag_adj(c, c->h[0], c->h_ad[0], c->u[0], c->u_ad[0]);
#undef UP
#undef HP
    free(up); free(hp); free(upa); free(hpa);
    return 0;
}
