/*
 * mpas_sw_amd: C ABI of the MI355X (gfx950) MPAS shallow-water RK4 core,
 * its tangent-linear model (TLM) and its adjoint.
 *
 * This is the drop-in boundary for the reference's compiled module
 * `module_sw_mpas` (xtian15/MPAS-SW-TL-AD, built by f2py from the src/ Fortran at
 * src/Makefile:7-8).  The reference exposes argument-less Fortran module
 * procedures acting on module globals; this library exposes the same
 * computations on explicit device buffers.  The Python facade
 * `module_sw_mpas.mpas_sw_module` (package mpas-sw-tl-ad_amd) restores the
 * reference's attribute interface on top of it.
 *
 * Conventions
 *  - All state pointers are DEVICE pointers owned by the caller, in the
 *    library's internal entity order (see msw_mesh_desc perms), with batch
 *    members stored innermost: element (entity i, member b) at [i*B + b].
 *  - `prec` selects the member arithmetic: MSW_FP64 (double*) or MSW_FP32
 *    (float*).  Nonlinear base states of the TLM/adjoint are always double.
 *  - FP64 member arrays of a context with B >= 64 passed to msw_fwd_steps,
 *    msw_tlm_steps and msw_adj_step(s) must be 16-byte aligned (MSW_EINVAL
 *    otherwise); device allocations and torch tensors are.
 *  - Every call is ordered on the stream given at creation (e.g. the caller's
 *    torch stream) and returns 0 or a negative MSW_E* code; the library never
 *    aborts.  msw_last_error() describes the last failure of this thread.
 *  - A context is not thread-safe.  Use one per (mesh, batch, precision).
 */
#ifndef MPAS_SW_AMD_H
#define MPAS_SW_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSW_ABI_VERSION 1

enum msw_status {
    MSW_OK = 0,
    MSW_EINVAL = -1,      /* bad argument or inconsistent mesh      */
    MSW_ENOMEM = -2,      /* device or host allocation failed       */
    MSW_EHIP = -3,        /* HIP runtime error                      */
    MSW_EUNSUPPORTED = -4 /* option the GPU path does not implement */
};

enum msw_precision { MSW_FP64 = 0, MSW_FP32 = 1 };

/* Mesh in the reference's layout: the arrays src/mpas_sw_driver.py:87-119
 * loads, already scaled to the sphere of radius a by the test-case routine
 * (src/mpas_sw_testcases.f90:91-104).  Connectivity is 1-based;
 * (slot, entity) arrays are flattened entity-major (Fortran column-major of
 * the reference's (maxEdges, nCells) etc.). */
typedef struct msw_mesh_desc {
    int32_t nCells, nEdges, nVertices;
    int32_t maxEdges, maxEdges2, vertexDegree;
    const int32_t *cellsOnEdge;      /* (2, E)          */
    const int32_t *verticesOnEdge;   /* (2, E)          */
    const int32_t *nEdgesOnCell;     /* (C)             */
    const int32_t *edgesOnCell;      /* (maxEdges, C)   */
    const int32_t *nEdgesOnEdge;     /* (E)             */
    const int32_t *edgesOnEdge;      /* (maxEdges2, E)  */
    const double  *weightsOnEdge;    /* (maxEdges2, E)  */
    const int32_t *cellsOnVertex;    /* (3, V)          */
    const double  *kiteAreasOnVertex;/* (3, V)          */
    const double  *dcEdge, *dvEdge;  /* (E)             */
    const double  *areaCell;         /* (C)             */
    const double  *areaTriangle;     /* (V)             */
    const double  *fVertex;          /* (V)             */
    const double  *h_s;              /* (C) bottom topography */
    const double  *latCell, *lonCell;/* (C)             */
    const double  *coeffs_reconstruct; /* (3, maxEdges, C), may be NULL */
    /* Internal order: perm[k] = reference index (0-based) of internal entity
     * k.  NULL means identity.  State buffers passed to the step functions
     * use the internal order; msw_permute converts. */
    const int32_t *cellPerm, *edgePerm, *vertexPerm;
} msw_mesh_desc;

/* Configuration scalars of src/mpas_sw_module.f90:23-42 that the stepped
 * path reads. */
typedef struct msw_config {
    double dt;                 /* config_dt                   */
    double apvm_upwinding;     /* config_apvm_upwinding       */
    double visc2, visc4;       /* config_h_mom_eddy_visc2/4   */
    int32_t test_case;         /* config_test_case (1 freezes u) */
    int32_t thickness_adv_order;
    int32_t monotonic, wind_stress, bottom_drag;
} msw_config;

typedef struct msw_ctx msw_ctx;

/* Replaces: var_allocation/var_allocation_tlm/var_allocation_adj and the
 * mesh part of sw_mpas_init_block (src/mpas_sw_module.f90:211-526).
 * batch: members per entity, a power of two in [1, 256]. */
int msw_create(const msw_mesh_desc *mesh, const msw_config *cfg, int prec, int batch,
               void *hip_stream, msw_ctx **out);
int msw_destroy(msw_ctx *ctx);
/* Update dt/apvm/visc/drag/wind between calls (same mesh). */
int msw_set_config(msw_ctx *ctx, const msw_config *cfg);
/* Replaces: the module array u_src (src/mpas_sw_module.f90:78,223-224), the
 * wind-stress source of sw_compute_tend (src/mpas_sw_dyncore.f90:619-625).
 * u_src: device FP64 [nEdges] in the internal edge order, copied on the
 * creation stream; NULL resets it to zero (its value after var_allocation).
 * Read only when config wind_stress is set. */
int msw_set_forcing(msw_ctx *ctx, const double *u_src);
/* Replaces: the module arrays meshScalingDel2/meshScalingDel4 that
 * compute_mesh_scaling fills (src/mpas_sw_dyncore.f90:362-392) and the del2
 * and del4 terms scale per edge (:520, :606, and their TLM and adjoint
 * twins).  del2, del4: device FP64 [nEdges] in the internal edge order,
 * copied on the creation stream; both NULL (the default) means all factors
 * are 1, their value unless config_h_ScaleWithMesh is set. */
int msw_set_mesh_scaling(msw_ctx *ctx, const double *del2, const double *del4);
/* Use a captured hipGraph per step for the step functions (default off). */
int msw_set_graphs(msw_ctx *ctx, int enable);
/* Replaces: the module arrays tend_h, tend_u, tend_h_tl, tend_u_tl
 * (src/mpas_sw_module.f90:188-195), which the last RK stage leaves holding
 * its tendencies after sw_rk4 (:577-586), sw_rk4_tlm (:700-712) and
 * sw_rk4_adj (its recomputed forward step, :838-848).  When set, the step
 * entry points write the last stage's tendencies here at every step, so
 * after a call they hold those of the last step run (for msw_adj_steps the
 * earliest step of the window, which its reverse sweep runs last):
 *   msw_fwd_steps: tend_h [nCells][B], tend_u [nEdges][B], the members'
 *     tendencies in the member precision;
 *   msw_base_steps, msw_tlm_steps, msw_adj_step, msw_adj_steps: tend_h
 *     [nCells], tend_u [nEdges], FP64, of the base state;
 *   msw_tlm_steps: tend_h_tl [nCells][B], tend_u_tl [nEdges][B], the
 *     members' TL tendencies in the member precision.
 * Device pointers in the internal order; NULL leaves that array unwritten
 * (the default for all four).  While any is set the steps run as stage
 * launches (MSW_OPT_PERSISTENT does not apply).  Waits for the creation
 * stream; captured step graphs are dropped. */
int msw_set_tend_output(msw_ctx *ctx, void *tend_h, void *tend_u, void *tend_h_tl, void *tend_u_tl);

/* Options without a counterpart in the reference.
 * MSW_OPT_TL_FUSED_FLUX (default 0): 1 forms the TL tendency of the TLM step
 * (src/mpas_sw_dyncore_tlm.f90:541-595) from the TL thickness flux
 * u_tl*h_edge + u*h_edge_tl at each edge, regrouping the reference's sums so
 * each member gathers two fields over edgesOnEdge instead of three; results
 * then differ from the reference in rounding only.  0 keeps the reference's
 * term order (bit-identical FP64 TLM).
 *
 * MSW_OPT_XCD_STRIPE (default 2): log2 of the number of consecutive blocks of
 * a launch that go to one XCD before the next XCD's turn (0..6), or 7: each
 * XCD takes one contiguous range of the launch's blocks.  Results do
 * not depend on it; the fastest value depends on mesh, batch and precision
 * (the Python layer's SWModel.autotune times the candidates).
 *
 * MSW_OPT_BLOCK (default 256): threads per block of the stage kernels, 64,
 * 128, 256 or 512 (with 64 or more members a wavefront is one entity, so a
 * block covers block/64 consecutive entities).  Results do not depend on it;
 * SWModel.autotune times it after the stripe.
 *
 * MSW_OPT_PERSISTENT (default 0): 1 runs msw_fwd_steps and msw_tlm_steps of a
 * batch-1 context as ONE cooperative launch for the whole window: every pass
 * of every RK stage takes (instance, patch) work items from a queue (patches
 * of up to 256 entities inside one instance, see msw_set_instances) and the
 * blocks meet at a grid barrier between passes.  Same results, bit for bit;
 * the call waits for the launch.  No viscosity, drag or wind stress
 * (MSW_EUNSUPPORTED); graphs are not used.
 *
 * MSW_OPT_BARRIER_POLLS (default 2^20): how many times a block polls a grid
 * barrier of the persistent launch before every block leaves the kernel and
 * the call returns MSW_EHIP. */
enum msw_option {
    MSW_OPT_TL_FUSED_FLUX = 1, MSW_OPT_XCD_STRIPE = 2, MSW_OPT_BLOCK = 3,
    MSW_OPT_PERSISTENT = 4, MSW_OPT_BARRIER_POLLS = 5
};
int msw_set_option(msw_ctx *ctx, int option, int value);

/* No counterpart in the reference.  The mesh of the context is a disjoint
 * union of n instances (an uneven batch packed into one index space):
 * instance i's cells, edges and vertices start at cell_start[i],
 * edge_start[i], vertex_start[i] (internal order, increasing, the first 0).
 * The persistent path cuts its work items at these boundaries; results do
 * not depend on them.  Host arrays of n entries each. */
int msw_set_instances(msw_ctx *ctx, int n, const int32_t *cell_start, const int32_t *edge_start,
                      const int32_t *vertex_start);

/* Replaces: sw_rk4 (src/mpas_sw_module.f90:528-635), nsteps times, on a
 * batch of B independent states h[C][B], u[E][B] of type `prec`. */
int msw_fwd_steps(msw_ctx *ctx, void *h, void *u, int nsteps);

/* FP64 single-member forward (the nonlinear trajectory of the TLM/adjoint).
 * If traj_h/traj_u are non-NULL the state at the START of step k is stored
 * at traj_h + k*C, traj_u + k*E. */
int msw_base_steps(msw_ctx *ctx, double *h, double *u, int nsteps, double *traj_h,
                   double *traj_u);

/* Replaces: sw_rk4_tlm (src/mpas_sw_module.f90:637-781), nsteps times.
 * Advances the FP64 nonlinear state (h,u) [B=1] and the batched TL state
 * (h_tl,u_tl) [B members, `prec`] together, as the reference does
 * (:774-777).  Optional trajectory storage as in msw_base_steps.  With no
 * tendency output set, the nonlinear state runs a step ahead of the
 * members within a call (the next step's base stages are formed while the
 * current step's member stages run); after an error return, (h,u) and
 * (h_tl,u_tl) are unspecified. */
int msw_tlm_steps(msw_ctx *ctx, double *h, double *u, void *h_tl, void *u_tl, int nsteps,
                  double *traj_h, double *traj_u);

/* Replaces: sw_rk4_adj (src/mpas_sw_module.f90:783-1002), one step.
 * (h,u): FP64 nonlinear state at the step start (unchanged).
 * (h_ad,u_ad): batched adjoint at the step end on entry, at the step start
 * on return.  (h_end,u_end): optional FP64 nonlinear end state (the
 * reference's h(2),u(2)). */
int msw_adj_step(msw_ctx *ctx, const double *h, const double *u, void *h_ad, void *u_ad,
                 double *h_end, double *u_end);

/* nsteps adjoint steps backwards over a stored trajectory: step k (from
 * nsteps-1 down to 0) uses traj_h + k*C, traj_u + k*E as its start state. */
int msw_adj_steps(msw_ctx *ctx, const double *traj_h, const double *traj_u, void *h_ad,
                  void *u_ad, int nsteps);

/* Replaces: sw_compute_solve_diagnostics (src/mpas_sw_dyncore.f90:1-361) for
 * one FP64 state; writes the 13 diagnostic fields into `out` (double,
 * packed in the order of msw_diag_names, each of length C, E or V). */
int msw_diagnostics(msw_ctx *ctx, const double *h, const double *u, double *out);
/* Replaces: sw_compute_solve_diagnostics_adj (src/mpas_sw_dyncore_adj.f90:1-453)
 * as sw_rk4_adj applies it to the seeds left in the module's diagnostic
 * adjoint arrays at the end state of its step (src/mpas_sw_module.f90:894-905):
 * h_ad += and u_ad += the adjoint of the diagnostics of the FP64 state (h,u),
 * applied to `seeds`: 13 batched arrays of `prec` in the order and sizes of
 * msw_diag_names, each [n][B], packed back to back.  Seeds are not modified
 * (the reference zeroes them after use). */
int msw_diagnostics_adj(msw_ctx *ctx, const double *h, const double *u, const void *seeds, void *h_ad,
                        void *u_ad);
/* comma-separated names and entity kinds ('c','e','v') of msw_diagnostics */
const char *msw_diag_names(void);
const char *msw_diag_kinds(void);

/* Global FP64 invariants of one FP64 state (B = 1), deterministic
 * reductions; no counterpart in the reference's step (its drivers print
 * numpy sums only).  out: device double[2] = { mass sum_c A_c h_c,
 * total energy sum_c A_c (g h_c (h_c/2 + h_s,c) + h_c ke_c) } with ke the
 * kinetic energy of src/mpas_sw_dyncore.f90:245-256. */
int msw_invariants(msw_ctx *ctx, const double *h, const double *u, double *out);

/* Replaces: mpas_reconstruct_2d (src/mpas_sw_operators.f90:283-370) for the
 * batch: out5 = [X, Y, Z, zonal, meridional], each [C][B] of `prec`. */
int msw_reconstruct(msw_ctx *ctx, const void *u, void *out5);
/* Replaces: mpas_reconstruct_2d_adj (src/mpas_sw_dyncore_adj.f90:740-831) as
 * sw_rk4_adj applies it to the end-state adjoint (src/mpas_sw_module.f90:887-892):
 * u_ad += R^T seeds5, seeds5 = [X, Y, Z, zonal, meridional] adjoint seeds, each
 * [C][B] of `prec` (not modified; the reference zeroes them after use). */
int msw_reconstruct_adj(msw_ctx *ctx, const void *seeds5, void *u_ad);

/* Deterministic FP64 inner products per member:
 *   out[b] = sum_c wh[c] xh[c][b] yh[c][b] + sum_e wu[e] xu[e][b] yu[e][b]
 * Any of wh/wu may be NULL (weight 1); yh/yu NULL means 1; xh or xu NULL
 * drops that part.  out: device double[B]. */
int msw_dot(msw_ctx *ctx, const void *xh, const void *xu, const void *yh, const void *yu,
            const double *wh, const double *wu, double *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): msw_dot for
 * members of a mesh with n_cells cells and n_edges edges, without a context,
 * ordered on hip_stream (NULL: the null stream) and bitwise equal to msw_dot
 * on a context of that mesh, batch and precision.  It allocates nothing: the
 * partial sums live in a 4 MiB device array of the library (one per GPU,
 * held until the process ends), so batch is at most 256 (MSW_EINVAL above).
 * Calls that overlap on one GPU share that array and race. */
int msw_dot_noalloc(int n_cells, int n_edges, int batch, int prec, const void *xh, const void *xu,
                    const void *yh, const void *yu, const double *wh, const double *wu, double *out,
                    void *hip_stream);

/* Per-member non-finite detector: flags[b] = 1 if any NaN/Inf in member b
 * of (h,u); flags is a device int[B] the caller zeroes. */
int msw_nonfinite(msw_ctx *ctx, const void *h, const void *u, int *flags);

/* Reorder rows between the reference order and the internal order.
 * kind: 'c','e','v'; to_internal: 1 (reference -> internal) or 0.
 * Rows hold `batch` values of precision `prec`. */
int msw_permute(msw_ctx *ctx, int kind, int to_internal, int prec, int batch, const void *src,
                void *dst);

/* Device memory the context holds, in bytes. */
int64_t msw_device_bytes(const msw_ctx *ctx);
int msw_abi_version(void);
const char *msw_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
