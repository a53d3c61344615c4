// Internal declarations shared by the host runtime (msw_runtime.cpp) and the
// HIP kernels (msw_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace msw {

constexpr int kMaxEdges = 6;
constexpr int kMaxEdges2 = 12;
// edgesOnEdge of an edge = the other edges of its two cells, so with cells of
// at most 6 edges no edge has more than 10; the member kernels gather only
// these 10 slots (slots 10 and 11 of the 12-wide tables are always padding)
constexpr int kEdgesOnEdge = 2 * (kMaxEdges - 1);
constexpr int kVertexDegree = 3;
constexpr double kGravity = 9.80616;   // src/mpas_sw_dyncore.f90:414
constexpr double kRhoRef = 1000.0;     // src/mpas_sw_dyncore.f90:458
constexpr double kDragCoef = 1.0e-3;   // bottom drag, src/mpas_sw_dyncore.f90:633

// Device-resident mesh tables, entities in the library's internal order
// (space-filling curve), 0-based.  "R" tables list neighbours in the order
// the reference's scatter loops visit them (ascending reference index), so a
// gather over them adds the terms in the reference's order.  Lists are padded
// to their fixed width with a valid index (the entity's own first neighbour,
// or the edge itself) and a zero sign/weight/kite, so padded terms are 0.
struct DMesh {
    int nC, nE, nV;
    int stripe_log2;      // blocks are dealt to the 8 XCDs in stripes of 2^stripe_log2
    int block;            // threads per block of those launches: 64, 128, 256 or 512
    // cells
    const int* nEoC;      // [C]
    const int* eoc;       // [C][6]  edgesOnCell, slot order
    const double* sgnC;   // [C][6]  +1 if the cell is cell1 of eoc[slot]
    const int* eocR;      // [C][6]  edges in ascending reference edge index
    const double* sgnCR;  // [C][6]  +1 if the cell is cell1 of eocR[k]
    const double* dvCR;   // [C][6]  dvEdge of eocR[k]
    const int* vocR;      // [C][6]  vertices in ascending reference vertex index
    const double* kiteCR; // [C][6]  kite area of the cell at vocR[k]
    const double* keC;    // [C][6]  0.25*dcEdge*dvEdge of eoc[k] (0 on padded slots)
    const double* areaC;  // [C]
    const double* hs;     // [C]
    // edges
    const int* coe;       // [E][2]
    const int* voe;       // [E][2]
    const int* nEoE;      // [E]
    const int* eoe;       // [E][12]
    const double* woe;    // [E][12]
    const int* eoeT;      // [E][12] slot of this edge in eoe[eoe[e][j]]
    const int* eSlot;     // [E][2]  slot of the edge in eoc of its cell 1 and cell 2
    const double* woeT;   // [E][12] woe[eoe[e][j]][eoeT[e][j]]: this edge's weight in
                          //         the list of its j-th neighbour (0 when padded)
    const double* dc;     // [E]
    const double* dv;     // [E]
    // vertices
    const int* eovR;      // [V][3]  edges in ascending reference edge index
    const double* sgnV;   // [V][3]  +1 if the vertex is vertex2 of eovR[k]
    const double* sdcV;   // [V][3]  sgnV * dcEdge(eovR[k]) (a sign flip: exact)
    const int* cov;       // [V][3]  cellsOnVertex, slot order
    const double* kite;   // [V][3]
    const double* areaT;  // [V]
    const double* fV;     // [V]
};

// Physics configuration the stencil needs.
struct DConfig {
    double dt;
    double apvm;
    int frozen_u;   // test case 1: wind field held fixed (src/mpas_sw_module.f90:596-598)
    int tl_fused;   // TL tendency from the fused TL thickness flux (fewer gathers,
                    // sums regrouped); 0 keeps the reference's term order
    double visc2;   // config_h_mom_eddy_visc2 (del2 on u when > 0)
    double visc4;   // config_h_mom_eddy_visc4 (del4 on u when > 0)
    int wind;       // config_wind_stress
    int drag;       // config_bottom_drag
    const double* usrc;   // [E] wind-stress source u_src (internal order), FP64
    const double* msd2;   // [E] meshScalingDel2 of the del2 term, nullptr: all 1
    const double* msd4;   // [E] meshScalingDel4 of the del4 term, nullptr: all 1
};

// Fields the del2/del4 terms of tend_u read (src/mpas_sw_dyncore.f90:508-617):
// divergence [C] and vorticity [V] of the stage state, and for del4 the
// divergence [C] and vorticity [V] of nabla^2 u.  Null when not needed.
template <typename T>
struct Dissip {
    const T* div;
    const T* vort;
    const T* dsd;
    const T* dsv;
};

// Per-stage base (nonlinear) fields kept for the tangent-linear and adjoint
// kernels.  All double, B = 1.
struct BaseStage {
    double* h;      // [C]  provisional state
    double* u;      // [E]
    double* he;     // [E]  h_edge
    double* pe;     // [E]  pv_edge
    double* v;      // [E]  tangential velocity
    double* gpt;    // [E]  gradPVt
    double* gpn;    // [E]  gradPVn
    double* ke;     // [C]
    double* pc;     // [C]  pv_cell
    double* pvv;    // [V]  pv_vertex
    double* hv;     // [V]  h_vertex
    double* vort;   // [V]  relative vorticity
    double* rec;    // [E][4] pv_edge, u, h_edge, gradPVt packed (32-byte
                    //        aligned) for gathers over edgesOnEdge
};

// Coefficients of a TL or adjoint stage operator, formed once per stage
// from the base fields (B = 1, FP64) and shared by every member.  Each
// member kernel then reads contiguous coefficient rows instead of gathering
// base values through the connectivity.  The TL stage uses e48 and c32 with
// its own layout (k_tl_coef); TL and adjoint steps never overlap.
// The rows are FP64 in FP64 contexts and FP32 in FP32 contexts (the same
// row strides in elements).
struct AdjCoef {
    double* e48;   // adjoint [E][24]: woeT*workpv, woeT*(-apvm*dt*gradPVt(eo)) per
                   //   edgesOnEdge slot; the TL stage uses [E][48]
    double* e8;    // [E][16]: u*dv, dv*he, sum_j 0.5*w*u*he, apvm*dt*gradPVn, drag/wind
                   //   h_edge and u terms, ke coefficient, u, he, 0.5*u*he,
                   //   1/areaCell of cell 1 and cell 2, 0...
    double* v8;    // [V][8]:  kite/areaCell[3], pv_edge terms[3], h_vertex, circulation
    double* c32;   // [C][32]: pv_cell[6], gravity[6], ke[6], h_edge[6], 1/areaCell, 0...
};

// RK4 coefficients of src/mpas_sw_module.f90:567-575.
struct RKCoeffs {
    double w[4];
    double sw[4];
};
RKCoeffs rk_coeffs(double dt);

// ---- launchers (msw_kernels.hip).  B = 1 << lb members per entity, stored
// innermost.  T is the member arithmetic type.

// Nonlinear stage: diagnostics at state (h,u) into `st` (which may alias
// `st.h/st.u` = (h,u)), then tendency and the RK4 update of stage `s`:
//   acc = (s==0 ? x : acc) + w[s]*tend ;  pnext = x + sw[s]*tend (s<3)
//   s==3: x = acc + w[3]*tend, and tend into (toh, tou) when they are set
template <typename T>
void fwd_diag(const DMesh& m, const DConfig& c, int lb, const T* h, const T* u,
              T* pvv, T* hv, T* vort, T* ke, T* pc, T* he, T* pe, T* v, T* gpt, T* gpn,
              double* rec, hipStream_t s);
template <typename T>
void fwd_tend_update_he(const DMesh& m, const DConfig& c, int lb, int stage, const RKCoeffs& k,
                        const T* h, const T* u, const T* he, const T* pe, const T* ke,
                        const Dissip<T>& dp, T* xh, T* xu, T* acch, T* accu, T* ph, T* pu,
                        const double* rec, T* toh, T* tou, hipStream_t s);

// Scratch of the del2/del4 terms for one field set (members or base):
// div [C], dsu [E], dsd [C], dsv [V]; vort [V] comes from the vertex pass.
template <typename T>
struct DissipScratch {
    T* div;
    T* dsu;
    T* dsd;
    T* dsv;
};
// divergence of u and, for del4, the nabla^2 u fields; returns what the
// tendency kernels read (all null when no viscosity is on)
template <typename T>
Dissip<T> dissip_fields(const DMesh& m, const DConfig& c, int lb, const T* u, const T* vort,
                        const DissipScratch<T>& ds, hipStream_t s);

// Coefficient rows of the TL / adjoint stage operator from base fields `b`
// (B = 1); the member launchers below read them from `q`.
// f32: rows stored in FP32 (contexts with FP32 members, which round every
// coefficient to FP32 where they read it, so the results are the same)
void tl_coef(const DMesh& m, const DConfig& c, const BaseStage& b, const AdjCoef& q, bool f32, hipStream_t s);
void adj_coef(const DMesh& m, const DConfig& c, const BaseStage& b, const AdjCoef& q, bool f32, hipStream_t s);
// the four stages' rows in one pass over the mesh tables (b[s], q[s] for s = 0..3)
void adj_coef4(const DMesh& m, const DConfig& c, const BaseStage* b, const AdjCoef* q, bool f32, hipStream_t s);

// Tangent-linear stage on top of base fields `b` of the same stage.
// vort_tl [V] and ds are used only when a viscosity is on; the last stage
// (stage 3) leaves its TL tendencies in (toh, tou) when they are set.
template <typename T>
void tl_stage(const DMesh& m, const DConfig& c, int lb, int stage, const RKCoeffs& k,
              const BaseStage& b, const T* h_tl, const T* u_tl,
              T* pvv_tl, T* ke_tl, T* pc_tl, T* pe_tl, T* he_tl, T* vort_tl,
              const DissipScratch<T>& ds, const AdjCoef& q,
              T* xh, T* xu, T* acch, T* accu, T* ph, T* pu, T* toh, T* tou, hipStream_t s);

// Adjoint of one RK4 stage: pa = F'(p_s)^T g, then the epilogue
//   s>0: accx = (first ? pa : accx + pa) ; g = w[s-1]*lam + sw[s-1]*pa
//   s==0: lam = (accx + lam) + pa
// The first stage of a step (first = true) reads g = w[3]*lam directly from
// lam, scaling on load; later stages read g from (gh, gu).
template <typename T>
void adj_stage(const DMesh& m, const DConfig& c, int lb, int stage, bool first, const RKCoeffs& k,
               const BaseStage& b, T* gh, T* gu, T* pe_ad, T* he_ad, T* ua_p, T* pc_ad, T* ha_p,
               T* ke_ad, T* hv_ad, T* circ_ad, const DissipScratch<T>& ds, const AdjCoef& q,
               T* lamh, T* lamu, T* acch, T* accu, hipStream_t s);
// g = w[3] * lam  (seed of the reverse sweep)
template <typename T>
void adj_seed(int n, T* g, const T* lam, double w3, hipStream_t s);

// FP64 fields of the state a diagnostic adjoint is linearised about (B = 1)
struct DiagBase {
    const double *u, *v, *gpn, *gpt;   // [E]
    const double *hv, *vort;           // [V]
};
// member seeds of the 13 diagnostic adjoints, [n][B] each
template <typename T>
struct DiagSeeds {
    const T *v, *he, *hv, *circ, *vort, *div, *ke, *pe, *pvv, *pc, *vc, *gpn, *gpt;
};
// h_ad += and u_ad += the adjoint of the diagnostics applied to the seeds
// (src/mpas_sw_dyncore_adj.f90:1-453); e1..e3 [E], c1 [C], v1, v2 [V]: member scratch
template <typename T>
void diag_adj(const DMesh& m, const DConfig& c, int lb, const DiagBase& bd, const DiagSeeds<T>& sd,
              T* e1, T* e2, T* e3, T* c1, T* v1, T* v2, T* h_ad, T* u_ad, hipStream_t s);

// Reconstruction of the cell-centre velocity (src/mpas_sw_operators.f90:283-370)
template <typename T>
void reconstruct(int nC, int lb, const int* nEoC, const int* eoc, const double* coeffs,
                 const double* trig, const T* u, T* out5, hipStream_t s);

// Adjoint of the reconstruction (src/mpas_sw_dyncore_adj.f90:740-831):
// u_ad += R^T seeds5, seeds5 = [X, Y, Z, zonal, meridional] adjoints, each
// [C][B]; xyz is [3][C][B] scratch.
template <typename T>
void reconstruct_adj(const DMesh& m, int lb, const double* coeffs, const double* trig,
                     const T* seeds5, T* xyz, T* u_ad, hipStream_t s);

// Deterministic per-member weighted dot products: out[b] = sum_i w_i x[i][b] y[i][b]
// (w may be null).  out is double[B].  `partial` holds nblk*B doubles.
template <typename T>
void dot_batched(long n, int lb, const T* x, const T* y, const double* w, double* partial,
                 int nblk, double* out, int accumulate, hipStream_t s);
int dot_blocks(long n);
// Partials of the allocation-free inner product (msw_dot_noalloc) live in a
// file-scope device array of msw_kernels.hip: dot_blocks(n) <= 2048 blocks of
// up to 256 members.  dot_scratch returns this device's address of it
// (looked up once per device, with that device current): 0 or a hipError_t.
constexpr long kDotScratchDoubles = 2048L * 256L;
int dot_scratch(double** out);

// copies and conversions
template <typename TD, typename TS>
void convert(long n, TD* dst, const TS* src, hipStream_t s);
template <typename T>
void fill(long n, T* dst, T val, hipStream_t s);
// dst[i][b] = src[perm[i]][b]   (gather rows)
template <typename T>
void gather_rows(long n, int lb, const int* perm, const T* src, T* dst, hipStream_t s);
// dst[perm[i]][b] = src[i][b]   (scatter rows)
template <typename T>
void scatter_rows(long n, int lb, const int* perm, const T* src, T* dst, hipStream_t s);
// out[b] = 1 if any non-finite value in member b of x (n rows)
template <typename T>
void nonfinite_flags(long n, int lb, const T* x, int* flags, hipStream_t s);

// Arguments of the persistent path (one member per entity): a window of
// nsteps forward (h, u: the member state) or TLM steps (bh, bu: the FP64
// base state; h, u: the TL state) in one cooperative launch.  Work items are
// patches (first entity, count <= 256) of vertices [np_v], then cells
// [np_c], then edges [np_e]; queue (2 x 8 counters), bar and err start at zero.
struct PersistArgs {
    DMesh m;
    DConfig c;
    RKCoeffs k;
    int nsteps;
    const int2* patch;
    int np_v, np_c, np_e;
    BaseStage st[4];          // base stage fields (st[0].h/u = bh/bu)
    double* bh;
    double* bu;
    double* bacc;             // [C+E]
    double* bend;             // [C+E]
    double* traj_h;           // optional [nsteps][C], [nsteps][E]
    double* traj_u;
    void* h;                  // member (forward) or TL state [C], [E]
    void* u;
    void* acc;                // [C+E]
    void* pA;                 // [C+E]
    void* pB;                 // [C+E]
    void* pvv;                // [V]
    void* ke;                 // [C]
    void* pc;                 // [C]
    void* he;                 // [E]
    void* pe;                 // [E]
    unsigned long long* queue;
    unsigned long long* bar;
    int* err;
    unsigned spin_limit;      // polls of a grid barrier before the launch gives up
};
// Launch it (cooperatively, every block resident): 0 or a hipError_t.
template <typename T>
int persistent_steps(const PersistArgs& a, bool tlm, hipStream_t s, int* grid);

}  // namespace msw

namespace msw {
// circulation (V), divergence (C) and vorticity_cell (C) for the diagnostic
// output (src/mpas_sw_dyncore.f90:208-242, :324-340); FP64, B = 1.
void diag_extra(const DMesh& m, const double* u, const double* vort, double* circ, double* div,
                double* vc, hipStream_t s);
}  // namespace msw
