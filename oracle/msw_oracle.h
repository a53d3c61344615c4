/*
 * CPU oracle: a serial C restatement of the reference shallow-water model,
 * its tangent-linear model and its adjoint (xtian15/MPAS-SW-TL-AD, the src/ Fortran files).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only to check results.  The
 * product path (the HIP library) never calls it.
 *
 * Every routine follows the reference loop order and scatter form so that
 * results agree with the reference Fortran to the last bit or nearly so; each
 * function cites the reference routine it restates.  Arrays use the
 * reference's Fortran layout flattened column-major with nVertLevels = 1:
 * a (slot, entity) array element (s, e) lives at [e*nslot + s]; connectivity
 * values stay 1-based.
 */
#ifndef MSW_ORACLE_H
#define MSW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mswo_mesh {
    int nCells, nEdges, nVertices, maxEdges, maxEdges2, vertexDegree;
    const int *cellsOnEdge, *verticesOnEdge;                  /* (2,E)         */
    const int *nEdgesOnCell, *edgesOnCell, *cellsOnCell;      /* (C), (mE,C)   */
    const int *nEdgesOnEdge, *edgesOnEdge;                    /* (E), (mE2,E)  */
    const double *weightsOnEdge;                              /* (mE2,E)       */
    const int *edgesOnVertex, *cellsOnVertex;                 /* (3,V)         */
    const double *kiteAreasOnVertex;                          /* (3,V)         */
    const double *dcEdge, *dvEdge, *areaCell, *areaTriangle;  /* scaled by a   */
    const double *fVertex, *h_s, *u_src;                      /* (V),(C),(E)   */
    const double *meshScalingDel2, *meshScalingDel4;          /* (E)           */
    const double *latCell, *lonCell;                          /* (C)           */
    const double *coeffs_reconstruct;                         /* (3,mE,C)      */
} mswo_mesh;

typedef struct mswo_config {
    double dt, apvm, visc2, visc4;
    int test_case, adv_order, monotonic, wind_stress, bottom_drag;
} mswo_config;

typedef struct mswo_ctx mswo_ctx;

/* create a model instance (copies the mesh); NULL on allocation failure */
mswo_ctx *mswo_create(const mswo_mesh *mesh, const mswo_config *cfg);
void mswo_destroy(mswo_ctx *ctx);
void mswo_set_config(mswo_ctx *ctx, const mswo_config *cfg);

/* Pointer to a named module array and its length, NULL if the name is
 * unknown.  Names: h1 h2 u1 u2 h_tl1 h_tl2 u_tl1 u_tl2 h_ad1 h_ad2 u_ad1
 * u_ad2, the diagnostic names of the reference (v h_edge h_vertex
 * circulation vorticity divergence ke pv_edge pv_vertex pv_cell
 * vorticity_cell gradPVn gradPVt) with optional _tl/_ad suffix, tend_h tend_u
 * (+_tl/_ad), uReconstructX/Y/Z/Zonal/Meridional (+_tl/_ad). */
double *mswo_array(mswo_ctx *ctx, const char *name, long *n);

/* L1 kernels on explicit state arrays; diagnostic and tendency outputs go
 * to the context's module arrays, adjoint inputs are consumed from them. */
void mswo_diag(mswo_ctx *ctx, const double *h, const double *u);
void mswo_tend(mswo_ctx *ctx, const double *h, const double *u);
void mswo_diag_tlm(mswo_ctx *ctx, const double *h, const double *h_tl,
                   const double *u, const double *u_tl);
void mswo_tend_tlm(mswo_ctx *ctx, const double *h, const double *h_tl,
                   const double *u, const double *u_tl);
void mswo_diag_adj(mswo_ctx *ctx, const double *h, double *h_ad,
                   const double *u, double *u_ad);
void mswo_tend_adj(mswo_ctx *ctx, const double *h, double *h_ad,
                   const double *u, double *u_ad);
void mswo_reconstruct(mswo_ctx *ctx, const double *u);
void mswo_reconstruct_tlm(mswo_ctx *ctx, const double *u, const double *u_tl);
void mswo_reconstruct_adj(mswo_ctx *ctx, const double *u, double *u_ad);

/* L2 drivers acting on the context's state, with the reference semantics */
void mswo_init_block(mswo_ctx *ctx);
void mswo_rk4(mswo_ctx *ctx);
void mswo_rk4_tlm(mswo_ctx *ctx);
int  mswo_rk4_adj(mswo_ctx *ctx);   /* 0, or -1 on allocation failure */

#ifdef __cplusplus
}
#endif
#endif
