// HIP kernels for gfx950: the MPAS shallow-water stencil (nonlinear,
// tangent-linear and adjoint) with ensemble members stored innermost.
//
// Thread mapping: thread t handles member (t & (B-1)) of entity (t >> lb),
// B = 1 << lb.  For B >= 64 a whole wavefront works on one entity, so every
// connectivity index and every base-state value is wave-uniform (scalar
// loads, one SGPR) and each neighbour gather is one contiguous B*sizeof(T)
// row.  For B = 1 (the plain path) each lane is one entity.
//
// All gathers visit neighbours in the order the reference's scatter loops
// add them (the "R" tables of DMesh) and every expression keeps the
// reference's association, so with FP contraction off the FP64 forward and
// tangent-linear results are the reference's to the bit.  Citations name the
// reference loop each kernel restates.
#include "msw_internal.hpp"

namespace msw {

namespace {

constexpr int kBlock = 256;

template <bool WU>
__device__ __forceinline__ int entity_of(long t, int lb)
{
    int e = static_cast<int>(t >> lb);
    if constexpr (WU) e = __builtin_amdgcn_readfirstlane(e);
    return e;
}

inline int grid_for(long threads)
{
    long blocks = (threads + kBlock - 1) / kBlock;
    const long cap = 1L << 20;   // grid-stride beyond ~268M threads
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    return static_cast<int>(blocks);
}

#define GRID_LOOP(total)                                                               \
    for (long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x; t < (total); \
         t += static_cast<long>(gridDim.x) * kBlock)

template <typename T>
__device__ __forceinline__ T ld(const T* p, long row, int B, int b)
{
    return p[row * B + b];
}

// ---------------------------------------------------------------------------
// Nonlinear diagnostics (src/mpas_sw_dyncore.f90:1-361)

// circulation/vorticity (:208-219), h_vertex and pv_vertex (:282-292)
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_fwd_vertex(DMesh m, int lb, const T* __restrict__ h,
                                                       const T* __restrict__ u, T* __restrict__ pvv,
                                                       T* __restrict__ hv, T* __restrict__ vort)
{
    const int B = 1 << lb;
    GRID_LOOP(static_cast<long>(m.nV) << lb)
    {
        const int iv = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        T circ = T(0);
#pragma unroll
        for (int i = 0; i < kVertexDegree; ++i) {
            const int e = m.eovR[3 * iv + i];
            circ = circ + T(m.sgnV[3 * iv + i]) * (T(m.dc[e]) * ld(u, e, B, b));
        }
        const T at = T(m.areaT[iv]);
        const T vo = circ / at;
        T hh = T(0);
#pragma unroll
        for (int i = 0; i < kVertexDegree; ++i)
            hh = hh + ld(h, m.cov[3 * iv + i], B, b) * T(m.kite[3 * iv + i]);
        hh = hh / at;
        const long o = static_cast<long>(iv) * B + b;
        pvv[o] = (T(m.fV[iv]) + vo) / hh;
        if (hv) hv[o] = hh;
        if (vort) vort[o] = vo;
    }
}

// kinetic energy (:245-256), pv_cell (:324-340)
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_fwd_cell(DMesh m, int lb, const T* __restrict__ u,
                                                     const T* __restrict__ pvv, T* __restrict__ ke,
                                                     T* __restrict__ pc)
{
    const int B = 1 << lb;
    GRID_LOOP(static_cast<long>(m.nC) << lb)
    {
        const int ic = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        const int n = m.nEoC[ic];
        const T area = T(m.areaC[ic]);
        T k = T(0);
        for (int i = 0; i < n; ++i) {
            const int e = m.eoc[kMaxEdges * ic + i];
            const T uu = ld(u, e, B, b);
            k = k + T(0.25) * T(m.dc[e]) * T(m.dv[e]) * (uu * uu);
        }
        k = k / area;
        T p = T(0);
        for (int i = 0; i < n; ++i) {
            const int iv = m.vocR[kMaxEdges * ic + i];
            p = p + T(m.kiteCR[kMaxEdges * ic + i]) * ld(pvv, iv, B, b) / area;
        }
        const long o = static_cast<long>(ic) * B + b;
        ke[o] = k;
        pc[o] = p;
    }
}

// h_edge (:95-105), v (:259-267), gradPVt (:296-301), pv_edge (:305-320),
// gradPVn and the second upwinding (:344-359)
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_fwd_edge(DMesh m, DConfig c, int lb,
                                                     const T* __restrict__ h,
                                                     const T* __restrict__ u,
                                                     const T* __restrict__ pvv,
                                                     const T* __restrict__ pc, T* __restrict__ he,
                                                     T* __restrict__ pe, T* __restrict__ vout,
                                                     T* __restrict__ gpt_out,
                                                     T* __restrict__ gpn_out)
{
    const int B = 1 << lb;
    const T apvm = T(c.apvm), dt = T(c.dt);
    GRID_LOOP(static_cast<long>(m.nE) << lb)
    {
        const int e = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        const int c1 = m.coe[2 * e], c2 = m.coe[2 * e + 1];
        const int v1 = m.voe[2 * e], v2 = m.voe[2 * e + 1];
        const T hh = T(0.5) * (ld(h, c1, B, b) + ld(h, c2, B, b));
        const int n = m.nEoE[e];
        T vv = T(0);
        for (int j = 0; j < n; ++j)
            vv = vv + T(m.woe[kMaxEdges2 * e + j]) * ld(u, m.eoe[kMaxEdges2 * e + j], B, b);
        const T p1 = ld(pvv, v1, B, b), p2 = ld(pvv, v2, B, b);
        const T gt = (p2 - p1) / T(m.dv[e]);
        const T gn = (ld(pc, c2, B, b) - ld(pc, c1, B, b)) / T(m.dc[e]);
        T p = T(0.5) * p1 + T(0.5) * p2;
        p = p - apvm * vv * dt * gt;
        p = p - apvm * ld(u, e, B, b) * dt * gn;
        const long o = static_cast<long>(e) * B + b;
        he[o] = hh;
        pe[o] = p;
        if (vout) vout[o] = vv;
        if (gpt_out) gpt_out[o] = gt;
        if (gpn_out) gpn_out[o] = gn;
    }
}

// RK4 bookkeeping of one stage for one value (src/mpas_sw_module.f90:590-611)
template <typename T>
__device__ __forceinline__ void rk_update(int stage, T w, T sw, T tend, T* x, T* acc, T* pnext,
                                          long o, bool frozen)
{
    const T xv = x[o];
    if (stage == 0) {
        acc[o] = xv + w * tend;
        pnext[o] = frozen ? xv : xv + sw * tend;
    } else if (stage < 3) {
        acc[o] = acc[o] + w * tend;
        pnext[o] = frozen ? xv : xv + sw * tend;
    } else {
        if (!frozen) x[o] = acc[o] + w * tend;
    }
}

// tend_u (:485-505) on edges and tend_h (:468-482) on cells, then RK update
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_fwd_tend(DMesh m, DConfig c, int lb, int stage,
                                                     T w, T sw, const T* __restrict__ h,
                                                     const T* __restrict__ u,
                                                     const T* __restrict__ he,
                                                     const T* __restrict__ pe,
                                                     const T* __restrict__ ke, T* xh, T* xu,
                                                     T* acch, T* accu, T* ph, T* pu)
{
    const int B = 1 << lb;
    const long nEt = static_cast<long>(m.nE) << lb;
    GRID_LOOP(nEt + (static_cast<long>(m.nC) << lb))
    {
        const int b = static_cast<int>(t & (B - 1));
        if (t < nEt) {
            const int e = entity_of<WU>(t, lb);
            const int c1 = m.coe[2 * e], c2 = m.coe[2 * e + 1];
            const T pee = ld(pe, e, B, b);
            const int n = m.nEoE[e];
            T q = T(0);
            for (int j = 0; j < n; ++j) {
                const int eo = m.eoe[kMaxEdges2 * e + j];
                const T workpv = T(0.5) * (pee + ld(pe, eo, B, b));
                q = q + T(m.woe[kMaxEdges2 * e + j]) * ld(u, eo, B, b) * workpv * ld(he, eo, B, b);
            }
            const T tend = q - (ld(ke, c2, B, b) - ld(ke, c1, B, b) +
                                T(kGravity) * (ld(h, c2, B, b) + T(m.hs[c2]) - ld(h, c1, B, b) -
                                               T(m.hs[c1]))) /
                                   T(m.dc[e]);
            rk_update(stage, w, sw, tend, xu, accu, pu, static_cast<long>(e) * B + b,
                      c.frozen_u != 0);
        } else {
            const int ic = entity_of<WU>(t - nEt, lb);
            const int n = m.nEoC[ic];
            T th = T(0);
            for (int k = 0; k < n; ++k) {
                const int e = m.eocR[kMaxEdges * ic + k];
                const T flux = ld(u, e, B, b) * T(m.dv[e]) * ld(he, e, B, b);
                th = th + T(-m.sgnCR[kMaxEdges * ic + k]) * flux;
            }
            th = th / T(m.areaC[ic]);
            rk_update(stage, w, sw, th, xh, acch, ph, static_cast<long>(ic) * B + b, false);
        }
    }
}

// ---------------------------------------------------------------------------
// Tangent-linear stage (src/mpas_sw_dyncore_tlm.f90).  Base fields are FP64,
// B = 1, shared by every member.

// :255-271, :344-360
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_tl_vertex(DMesh m, int lb, const double* __restrict__ hvb,
                                                      const double* __restrict__ vortb,
                                                      const T* __restrict__ h_tl,
                                                      const T* __restrict__ u_tl,
                                                      T* __restrict__ pvv_tl)
{
    const int B = 1 << lb;
    GRID_LOOP(static_cast<long>(m.nV) << lb)
    {
        const int iv = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        T ct = T(0);
#pragma unroll
        for (int i = 0; i < kVertexDegree; ++i) {
            const int e = m.eovR[3 * iv + i];
            ct = ct + T(m.sgnV[3 * iv + i]) * (T(m.dc[e]) * ld(u_tl, e, B, b));
        }
        const T at = T(m.areaT[iv]);
        const T vt = ct / at;
        T ht = T(0);
#pragma unroll
        for (int i = 0; i < kVertexDegree; ++i)
            ht = ht + ld(h_tl, m.cov[3 * iv + i], B, b) * T(m.kite[3 * iv + i]);
        ht = ht / at;
        const T hv = T(hvb[iv]);
        pvv_tl[static_cast<long>(iv) * B + b] =
            vt / hv - ht * (T(m.fV[iv]) + T(vortb[iv])) / (hv * hv);
    }
}

// :301-316, :401-421
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_tl_cell(DMesh m, int lb, const double* __restrict__ ub,
                                                    const T* __restrict__ u_tl,
                                                    const T* __restrict__ pvv_tl,
                                                    T* __restrict__ ke_tl, T* __restrict__ pc_tl)
{
    const int B = 1 << lb;
    GRID_LOOP(static_cast<long>(m.nC) << lb)
    {
        const int ic = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        const int n = m.nEoC[ic];
        const T area = T(m.areaC[ic]);
        T k = T(0);
        for (int i = 0; i < n; ++i) {
            const int e = m.eoc[kMaxEdges * ic + i];
            k = k + T(0.25) * T(m.dc[e]) * T(m.dv[e]) * (T(2.0) * T(ub[e]) * ld(u_tl, e, B, b));
        }
        k = k / area;
        T p = T(0);
        for (int i = 0; i < n; ++i) {
            const int iv = m.vocR[kMaxEdges * ic + i];
            p = p + T(m.kiteCR[kMaxEdges * ic + i]) * ld(pvv_tl, iv, B, b) / area;
        }
        const long o = static_cast<long>(ic) * B + b;
        ke_tl[o] = k;
        pc_tl[o] = p;
    }
}

// :319-329, :364-393, :425-445
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_tl_edge(DMesh m, DConfig c, int lb, BaseStage bs,
                                                    const T* __restrict__ u_tl,
                                                    const T* __restrict__ pvv_tl,
                                                    const T* __restrict__ pc_tl,
                                                    T* __restrict__ pe_tl)
{
    const int B = 1 << lb;
    const T apvm = T(c.apvm), dt = T(c.dt);
    GRID_LOOP(static_cast<long>(m.nE) << lb)
    {
        const int e = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        const int c1 = m.coe[2 * e], c2 = m.coe[2 * e + 1];
        const int v1 = m.voe[2 * e], v2 = m.voe[2 * e + 1];
        const int n = m.nEoE[e];
        T vt = T(0);
        for (int j = 0; j < n; ++j)
            vt = vt + T(m.woe[kMaxEdges2 * e + j]) * ld(u_tl, m.eoe[kMaxEdges2 * e + j], B, b);
        const T p1 = ld(pvv_tl, v1, B, b), p2 = ld(pvv_tl, v2, B, b);
        const T gtt = (p2 - p1) / T(m.dv[e]);
        const T gnt = (ld(pc_tl, c2, B, b) - ld(pc_tl, c1, B, b)) / T(m.dc[e]);
        T p = T(0.5) * p1 + T(0.5) * p2;
        p = p - apvm * vt * dt * T(bs.gpt[e]) - apvm * T(bs.v[e]) * dt * gtt;
        p = p - apvm * ld(u_tl, e, B, b) * dt * T(bs.gpn[e]) - apvm * T(bs.u[e]) * dt * gnt;
        pe_tl[static_cast<long>(e) * B + b] = p;
    }
}

// tend_u_tl (:564-595) and tend_h_tl (:541-561), then RK update
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_tl_tend(DMesh m, DConfig c, int lb, int stage, T w,
                                                    T sw, BaseStage bs,
                                                    const T* __restrict__ h_tl,
                                                    const T* __restrict__ u_tl,
                                                    const T* __restrict__ pe_tl,
                                                    const T* __restrict__ ke_tl, T* xh, T* xu,
                                                    T* acch, T* accu, T* ph, T* pu)
{
    const int B = 1 << lb;
    const long nEt = static_cast<long>(m.nE) << lb;
    GRID_LOOP(nEt + (static_cast<long>(m.nC) << lb))
    {
        const int b = static_cast<int>(t & (B - 1));
        if (t < nEt) {
            const int e = entity_of<WU>(t, lb);
            const int c1 = m.coe[2 * e], c2 = m.coe[2 * e + 1];
            const T pte = ld(pe_tl, e, B, b);
            const T pbe = T(bs.pe[e]);
            const int n = m.nEoE[e];
            T qt = T(0);
            for (int j = 0; j < n; ++j) {
                const int eo = m.eoe[kMaxEdges2 * e + j];
                const T wgt = T(m.woe[kMaxEdges2 * e + j]);
                const T wpvt = T(0.5) * (pte + ld(pe_tl, eo, B, b));
                const T wpv = T(0.5) * (pbe + T(bs.pe[eo]));
                const T heb = T(bs.he[eo]);
                const T het = T(0.5) * (ld(h_tl, m.coe[2 * eo], B, b) +
                                        ld(h_tl, m.coe[2 * eo + 1], B, b));
                const T ubo = T(bs.u[eo]);
                qt = qt + wgt * ld(u_tl, eo, B, b) * wpv * heb + wgt * ubo * wpvt * heb +
                     wgt * ubo * wpv * het;
            }
            const T tend = qt - (ld(ke_tl, c2, B, b) - ld(ke_tl, c1, B, b) +
                                 T(kGravity) * (ld(h_tl, c2, B, b) - ld(h_tl, c1, B, b))) /
                                    T(m.dc[e]);
            rk_update(stage, w, sw, tend, xu, accu, pu, static_cast<long>(e) * B + b,
                      c.frozen_u != 0);
        } else {
            const int ic = entity_of<WU>(t - nEt, lb);
            const int n = m.nEoC[ic];
            T th = T(0);
            for (int k = 0; k < n; ++k) {
                const int e = m.eocR[kMaxEdges * ic + k];
                const T het = T(0.5) * (ld(h_tl, m.coe[2 * e], B, b) +
                                        ld(h_tl, m.coe[2 * e + 1], B, b));
                const T flux = ld(u_tl, e, B, b) * T(m.dv[e]) * T(bs.he[e]) +
                               T(bs.u[e]) * T(m.dv[e]) * het;
                th = th + T(-m.sgnCR[kMaxEdges * ic + k]) * flux;
            }
            th = th / T(m.areaC[ic]);
            rk_update(stage, w, sw, th, xh, acch, ph, static_cast<long>(ic) * B + b, false);
        }
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the tendency operator (src/mpas_sw_dyncore_adj.f90), written as
// gathers over the transposed connectivity: every scatter-add of the
// reference becomes a sum at the receiving entity, so no atomics are needed
// and results are deterministic.

// tend_u/q adjoint (:689-717), tend_h adjoint (:720-736), ke -> u (:306-318),
// upwinding in the normal direction (:193-200)
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_adj_edge1(DMesh m, DConfig c, int lb, BaseStage bs,
                                                      const T* __restrict__ gh,
                                                      const T* __restrict__ gu,
                                                      T* __restrict__ pe_ad,
                                                      T* __restrict__ he_ad,
                                                      T* __restrict__ ua_p)
{
    const int B = 1 << lb;
    const T apvm = T(c.apvm), dt = T(c.dt);
    GRID_LOOP(static_cast<long>(m.nE) << lb)
    {
        const int e = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        const int c1 = m.coe[2 * e], c2 = m.coe[2 * e + 1];
        const T a1 = T(m.areaC[c1]), a2 = T(m.areaC[c2]);
        const T dv = T(m.dv[e]), dc = T(m.dc[e]);
        const T ub = T(bs.u[e]), heb = T(bs.he[e]), peb = T(bs.pe[e]);
        const T flux_ad = ld(gh, c2, B, b) / a2 - ld(gh, c1, B, b) / a1;
        const T ge = ld(gu, e, B, b);
        T pad = T(0);
        T had = ub * dv * flux_ad;
        T uad = dv * heb * flux_ad;
        const int n = m.nEoE[e];
        for (int j = 0; j < n; ++j) {
            const int eo = m.eoe[kMaxEdges2 * e + j];
            const int jt = m.eoeT[kMaxEdges2 * e + j];
            const T wgt = T(m.woe[kMaxEdges2 * e + j]);
            const T wt = T(m.woe[kMaxEdges2 * eo + jt]);
            const T wpv = T(0.5) * (peb + T(bs.pe[eo]));
            // this edge's own q: pv_edge_ad(e) += 0.5*w*u(eo)*he(eo)*q_ad(e)
            pad = pad + T(0.5) * (wgt * T(bs.u[eo]) * T(bs.he[eo]) * ge);
            // edge eo's q, in which this edge is neighbour jt
            const T go = ld(gu, eo, B, b);
            had = had + wt * ub * wpv * go;
            uad = uad + wt * wpv * heb * go;
            pad = pad + T(0.5) * (wt * ub * heb * go);
        }
        // ke_ad of both cells from tend_u's kinetic-energy gradient
        T k1 = T(0), k2 = T(0);
        const int n1 = m.nEoC[c1], n2 = m.nEoC[c2];
        for (int i = 0; i < n1; ++i) {
            const int ee = m.eoc[kMaxEdges * c1 + i];
            k1 = k1 + T(m.sgnC[kMaxEdges * c1 + i]) * ld(gu, ee, B, b) / T(m.dc[ee]);
        }
        for (int i = 0; i < n2; ++i) {
            const int ee = m.eoc[kMaxEdges * c2 + i];
            k2 = k2 + T(m.sgnC[kMaxEdges * c2 + i]) * ld(gu, ee, B, b) / T(m.dc[ee]);
        }
        const T kc = T(0.25) * dc * dv * T(2.0) * ub;
        uad = uad + kc * (k1 / a1) + kc * (k2 / a2);
        uad = uad - apvm * dt * T(bs.gpn[e]) * pad;
        const long o = static_cast<long>(e) * B + b;
        pe_ad[o] = pad;
        he_ad[o] = had;
        ua_p[o] = uad;
    }
}

// gradPVn -> pv_cell (:202-213), tend_u's gravity term -> h (:696-697),
// h_edge -> h (:358-369)
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_adj_cell1(DMesh m, DConfig c, int lb, BaseStage bs,
                                                      const T* __restrict__ gu,
                                                      const T* __restrict__ pe_ad,
                                                      const T* __restrict__ he_ad,
                                                      T* __restrict__ pc_ad,
                                                      T* __restrict__ ha_p)
{
    const int B = 1 << lb;
    const T apvm = T(c.apvm), dt = T(c.dt);
    GRID_LOOP(static_cast<long>(m.nC) << lb)
    {
        const int ic = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        const int n = m.nEoC[ic];
        T pcad = T(0), had = T(0);
        for (int i = 0; i < n; ++i) {
            const int e = m.eoc[kMaxEdges * ic + i];
            const T s = T(m.sgnC[kMaxEdges * ic + i]);
            const T dc = T(m.dc[e]);
            const T gpn_ad = -apvm * T(bs.u[e]) * dt * ld(pe_ad, e, B, b);
            pcad = pcad - s * gpn_ad / dc;
            had = had + s * (T(kGravity) / dc * ld(gu, e, B, b)) + T(0.5) * ld(he_ad, e, B, b);
        }
        const long o = static_cast<long>(ic) * B + b;
        pc_ad[o] = pcad;
        ha_p[o] = had;
    }
}

// pv_cell -> pv_vertex (:215-229), pv_edge -> pv_vertex (:244-252),
// gradPVt -> pv_vertex (:235-262), pv_vertex -> h_vertex, vorticity
// (:268-284), vorticity -> circulation (:343-349)
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_adj_vertex(DMesh m, DConfig c, int lb, BaseStage bs,
                                                       const T* __restrict__ pe_ad,
                                                       const T* __restrict__ pc_ad,
                                                       T* __restrict__ hv_ad,
                                                       T* __restrict__ circ_ad)
{
    const int B = 1 << lb;
    const T apvm = T(c.apvm), dt = T(c.dt);
    GRID_LOOP(static_cast<long>(m.nV) << lb)
    {
        const int iv = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        T pa = T(0);
#pragma unroll
        for (int i = 0; i < kVertexDegree; ++i) {
            const int ic = m.cov[3 * iv + i];
            pa = pa + T(m.kite[3 * iv + i]) / T(m.areaC[ic]) * ld(pc_ad, ic, B, b);
        }
#pragma unroll
        for (int i = 0; i < kVertexDegree; ++i) {
            const int e = m.eovR[3 * iv + i];
            const T pad = ld(pe_ad, e, B, b);
            const T gpt_ad = -apvm * T(bs.v[e]) * dt * pad;
            pa = pa + T(0.5) * pad + T(m.sgnV[3 * iv + i]) * gpt_ad / T(m.dv[e]);
        }
        const T hv = T(bs.hv[iv]);
        const T at = T(m.areaT[iv]);
        T hva = -(T(m.fV[iv]) + T(bs.vort[iv])) / (hv * hv) * pa;
        hva = hva / at;
        const T vora = pa / hv;
        const long o = static_cast<long>(iv) * B + b;
        hv_ad[o] = hva;
        circ_ad[o] = vora / at;
    }
}

// Epilogue of an adjoint stage for one value
template <typename T>
__device__ __forceinline__ void adj_update(int stage, bool first, T w, T sw, T pa, T* lam, T* acc,
                                           T* g, long o)
{
    if (stage > 0) {
        acc[o] = first ? pa : acc[o] + pa;
        g[o] = w * lam[o] + sw * pa;
    } else {
        lam[o] = (acc[o] + lam[o]) + pa;
    }
}

// v -> u through the transposed edgesOnEdge (:296-304), circulation -> u
// (:350-355) on edges; h_vertex -> h (:284-290) on cells; then the RK4
// adjoint bookkeeping (src/mpas_sw_module.f90:912-980)
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_adj_final(DMesh m, DConfig c, int lb, int stage,
                                                      bool first, T w, T sw, BaseStage bs,
                                                      const T* __restrict__ pe_ad,
                                                      const T* __restrict__ ua_p,
                                                      const T* __restrict__ ha_p,
                                                      const T* __restrict__ hv_ad,
                                                      const T* __restrict__ circ_ad, T* lamh,
                                                      T* lamu, T* acch, T* accu, T* gh, T* gu)
{
    const int B = 1 << lb;
    const T apvm = T(c.apvm), dt = T(c.dt);
    const long nEt = static_cast<long>(m.nE) << lb;
    GRID_LOOP(nEt + (static_cast<long>(m.nC) << lb))
    {
        const int b = static_cast<int>(t & (B - 1));
        if (t < nEt) {
            const int e = entity_of<WU>(t, lb);
            T ua = ld(ua_p, e, B, b);
            const int n = m.nEoE[e];
            for (int j = 0; j < n; ++j) {
                const int eo = m.eoe[kMaxEdges2 * e + j];
                const int jt = m.eoeT[kMaxEdges2 * e + j];
                const T vad = -apvm * dt * T(bs.gpt[eo]) * ld(pe_ad, eo, B, b);
                ua = ua + T(m.woe[kMaxEdges2 * eo + jt]) * vad;
            }
            ua = ua + T(m.dc[e]) * (ld(circ_ad, m.voe[2 * e + 1], B, b) -
                                    ld(circ_ad, m.voe[2 * e], B, b));
            adj_update(stage, first, w, sw, ua, lamu, accu, gu, static_cast<long>(e) * B + b);
        } else {
            const int ic = entity_of<WU>(t - nEt, lb);
            const int n = m.nEoC[ic];
            T ha = ld(ha_p, ic, B, b);
            for (int k = 0; k < n; ++k)
                ha = ha + T(m.kiteCR[kMaxEdges * ic + k]) * ld(hv_ad, m.vocR[kMaxEdges * ic + k], B, b);
            adj_update(stage, first, w, sw, ha, lamh, acch, gh, static_cast<long>(ic) * B + b);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(long n, T* __restrict__ g,
                                                  const T* __restrict__ lam, T w)
{
    GRID_LOOP(n) { g[t] = w * lam[t]; }
}

// ---------------------------------------------------------------------------
// Reconstruction (src/mpas_sw_operators.f90:283-370); trig = (cos lat, sin
// lat, cos lon, sin lon) per cell, evaluated on the host with the same libm
// as the reference.
template <typename T, bool WU>
__global__ __launch_bounds__(kBlock) void k_reconstruct(int nC, int lb, const int* __restrict__ nEoC,
                                                        const int* __restrict__ eoc,
                                                        const double* __restrict__ coeffs,
                                                        const double* __restrict__ trig,
                                                        const T* __restrict__ u, T* __restrict__ out)
{
    const int B = 1 << lb;
    const long plane = static_cast<long>(nC) * B;
    GRID_LOOP(static_cast<long>(nC) << lb)
    {
        const int ic = entity_of<WU>(t, lb);
        const int b = static_cast<int>(t & (B - 1));
        T X = T(0), Y = T(0), Z = T(0);
        const int n = nEoC[ic];
        for (int i = 0; i < n; ++i) {
            const T uu = ld(u, eoc[kMaxEdges * ic + i], B, b);
            const double* cf = coeffs + (static_cast<long>(ic) * kMaxEdges + i) * 3;
            X = X + T(cf[0]) * uu;
            Y = Y + T(cf[1]) * uu;
            Z = Z + T(cf[2]) * uu;
        }
        const T clat = T(trig[4 * ic]), slat = T(trig[4 * ic + 1]);
        const T clon = T(trig[4 * ic + 2]), slon = T(trig[4 * ic + 3]);
        const long o = static_cast<long>(ic) * B + b;
        out[o] = X;
        out[plane + o] = Y;
        out[2 * plane + o] = Z;
        out[3 * plane + o] = -X * slon + Y * clon;
        out[4 * plane + o] = -(X * clon + Y * slon) * slat + Z * clat;
    }
}

// ---------------------------------------------------------------------------
// Deterministic batched reductions in FP64.  Block k sums rows
// [k*chunk, (k+1)*chunk) for every member; the host-side order of the final
// sum over blocks is fixed, so results do not depend on scheduling.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot_partial(long n, int lb, long chunk,
                                                        const T* __restrict__ x,
                                                        const T* __restrict__ y,
                                                        const double* __restrict__ w,
                                                        double* __restrict__ partial)
{
    __shared__ double red[kBlock];
    const int B = 1 << lb;
    const int b = threadIdx.x & (B - 1);
    const int lane_row = threadIdx.x >> lb;
    const int rows_per_pass = kBlock >> lb;
    const long r0 = static_cast<long>(blockIdx.x) * chunk;
    long r1 = r0 + chunk;
    if (r1 > n) r1 = n;
    double s = 0.0;
    for (long r = r0 + lane_row; r < r1; r += rows_per_pass) {
        const double xv = static_cast<double>(x[r * B + b]);
        const double yv = y ? static_cast<double>(y[r * B + b]) : 1.0;
        s += w ? w[r] * (xv * yv) : xv * yv;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kBlock / 2; off >= B; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (static_cast<int>(threadIdx.x) < B) partial[static_cast<long>(blockIdx.x) * B + threadIdx.x] = red[threadIdx.x];
}

__global__ void k_dot_final(int nblk, int B, int accumulate, const double* __restrict__ partial,
                            double* __restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s = 0.0;
    for (int k = 0; k < nblk; ++k) s += partial[static_cast<long>(k) * B + b];
    out[b] = accumulate ? out[b] + s : s;
}

template <typename TD, typename TS>
__global__ __launch_bounds__(kBlock) void k_convert(long n, TD* __restrict__ d,
                                                    const TS* __restrict__ s)
{
    GRID_LOOP(n) { d[t] = static_cast<TD>(s[t]); }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(long n, T* __restrict__ d, T v)
{
    GRID_LOOP(n) { d[t] = v; }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_rows(long n, int lb, const int* __restrict__ perm,
                                                        const T* __restrict__ s, T* __restrict__ d)
{
    const int B = 1 << lb;
    GRID_LOOP(n << lb) { d[t] = s[static_cast<long>(perm[t >> lb]) * B + (t & (B - 1))]; }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter_rows(long n, int lb, const int* __restrict__ perm,
                                                         const T* __restrict__ s, T* __restrict__ d)
{
    const int B = 1 << lb;
    GRID_LOOP(n << lb) { d[static_cast<long>(perm[t >> lb]) * B + (t & (B - 1))] = s[t]; }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_nonfinite(long n, int lb, const T* __restrict__ x,
                                                      int* __restrict__ flags)
{
    const int B = 1 << lb;
    GRID_LOOP(n << lb)
    {
        const T v = x[t];
        if (!(v == v) || v - v != T(0)) flags[t & (B - 1)] = 1;   // NaN or Inf
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers

RKCoeffs rk_coeffs(double dt)
{
    RKCoeffs k;
    k.w[0] = dt / 6.0; k.w[1] = dt / 3.0; k.w[2] = dt / 3.0; k.w[3] = dt / 6.0;
    k.sw[0] = dt / 2.0; k.sw[1] = dt / 2.0; k.sw[2] = dt; k.sw[3] = 0.0;
    return k;
}

#define LAUNCH_WU(kernel, T, lb, threads, stream, ...)                                       \
    do {                                                                                     \
        const int _g = grid_for(threads);                                                    \
        if ((lb) >= 6)                                                                       \
            hipLaunchKernelGGL((kernel<T, true>), dim3(_g), dim3(kBlock), 0, stream, __VA_ARGS__); \
        else                                                                                 \
            hipLaunchKernelGGL((kernel<T, false>), dim3(_g), dim3(kBlock), 0, stream, __VA_ARGS__); \
    } while (0)

template <typename T>
void fwd_diag(const DMesh& m, const DConfig& c, int lb, const T* h, const T* u, T* pvv, T* hv,
              T* vort, T* ke, T* pc, T* he, T* pe, T* v, T* gpt, T* gpn, hipStream_t s)
{
    LAUNCH_WU(k_fwd_vertex, T, lb, static_cast<long>(m.nV) << lb, s, m, lb, h, u, pvv, hv, vort);
    LAUNCH_WU(k_fwd_cell, T, lb, static_cast<long>(m.nC) << lb, s, m, lb, u, pvv, ke, pc);
    LAUNCH_WU(k_fwd_edge, T, lb, static_cast<long>(m.nE) << lb, s, m, c, lb, h, u, pvv, pc, he, pe, v,
              gpt, gpn);
}

template <typename T>
void fwd_tend_update_he(const DMesh& m, const DConfig& c, int lb, int stage, const RKCoeffs& k,
                        const T* h, const T* u, const T* he, const T* pe, const T* ke, T* xh,
                        T* xu, T* acch, T* accu, T* ph, T* pu, hipStream_t s)
{
    const long threads = (static_cast<long>(m.nE) + m.nC) << lb;
    LAUNCH_WU(k_fwd_tend, T, lb, threads, s, m, c, lb, stage, T(k.w[stage]), T(k.sw[stage]), h, u,
              he, pe, ke, xh, xu, acch, accu, ph, pu);
}

template <typename T>
void tl_stage(const DMesh& m, const DConfig& c, int lb, int stage, const RKCoeffs& k,
              const BaseStage& b, const T* h_tl, const T* u_tl, T* pvv_tl, T* ke_tl, T* pc_tl,
              T* pe_tl, T* xh, T* xu, T* acch, T* accu, T* ph, T* pu, hipStream_t s)
{
    LAUNCH_WU(k_tl_vertex, T, lb, static_cast<long>(m.nV) << lb, s, m, lb, b.hv, b.vort, h_tl, u_tl,
              pvv_tl);
    LAUNCH_WU(k_tl_cell, T, lb, static_cast<long>(m.nC) << lb, s, m, lb, b.u, u_tl, pvv_tl, ke_tl,
              pc_tl);
    LAUNCH_WU(k_tl_edge, T, lb, static_cast<long>(m.nE) << lb, s, m, c, lb, b, u_tl, pvv_tl, pc_tl,
              pe_tl);
    const long threads = (static_cast<long>(m.nE) + m.nC) << lb;
    LAUNCH_WU(k_tl_tend, T, lb, threads, s, m, c, lb, stage, T(k.w[stage]), T(k.sw[stage]), b, h_tl,
              u_tl, pe_tl, ke_tl, xh, xu, acch, accu, ph, pu);
}

template <typename T>
void adj_stage(const DMesh& m, const DConfig& c, int lb, int stage, bool first, const RKCoeffs& k,
               const BaseStage& b, T* gh, T* gu, T* pe_ad, T* he_ad, T* ua_p, T* pc_ad, T* ha_p,
               T* hv_ad, T* circ_ad, T* lamh, T* lamu, T* acch, T* accu, hipStream_t s)
{
    LAUNCH_WU(k_adj_edge1, T, lb, static_cast<long>(m.nE) << lb, s, m, c, lb, b, gh, gu, pe_ad, he_ad,
              ua_p);
    LAUNCH_WU(k_adj_cell1, T, lb, static_cast<long>(m.nC) << lb, s, m, c, lb, b, gu, pe_ad, he_ad,
              pc_ad, ha_p);
    LAUNCH_WU(k_adj_vertex, T, lb, static_cast<long>(m.nV) << lb, s, m, c, lb, b, pe_ad, pc_ad, hv_ad,
              circ_ad);
    const double w = stage > 0 ? k.w[stage - 1] : 0.0;
    const double sw = stage > 0 ? k.sw[stage - 1] : 0.0;
    const long threads = (static_cast<long>(m.nE) + m.nC) << lb;
    LAUNCH_WU(k_adj_final, T, lb, threads, s, m, c, lb, stage, first, T(w), T(sw), b, pe_ad, ua_p,
              ha_p, hv_ad, circ_ad, lamh, lamu, acch, accu, gh, gu);
}

template <typename T>
void adj_seed(int n, T* g, const T* lam, double w3, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, static_cast<long>(n), g,
                       lam, T(w3));
}

template <typename T>
void reconstruct(int nC, int lb, const int* nEoC, const int* eoc, const double* coeffs,
                 const double* trig, const T* u, T* out5, hipStream_t s)
{
    LAUNCH_WU(k_reconstruct, T, lb, static_cast<long>(nC) << lb, s, nC, lb, nEoC, eoc, coeffs, trig,
              u, out5);
}

int dot_blocks(long n)
{
    long nb = (n + 4095) / 4096;
    if (nb < 1) nb = 1;
    if (nb > 2048) nb = 2048;
    return static_cast<int>(nb);
}

template <typename T>
void dot_batched(long n, int lb, const T* x, const T* y, const double* w, double* partial, int nblk,
                 double* out, int accumulate, hipStream_t s)
{
    const long chunk = (n + nblk - 1) / nblk;
    hipLaunchKernelGGL(k_dot_partial<T>, dim3(nblk), dim3(kBlock), 0, s, n, lb, chunk, x, y, w,
                       partial);
    const int B = 1 << lb;
    hipLaunchKernelGGL(k_dot_final, dim3((B + 63) / 64), dim3(64), 0, s, nblk, B, accumulate,
                       partial, out);
}

template <typename TD, typename TS>
void convert(long n, TD* dst, const TS* src, hipStream_t s)
{
    hipLaunchKernelGGL((k_convert<TD, TS>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, dst, src);
}

template <typename T>
void fill(long n, T* dst, T val, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, dst, val);
}

template <typename T>
void gather_rows(long n, int lb, const int* perm, const T* src, T* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_rows<T>, dim3(grid_for(n << lb)), dim3(kBlock), 0, s, n, lb, perm,
                       src, dst);
}

template <typename T>
void scatter_rows(long n, int lb, const int* perm, const T* src, T* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter_rows<T>, dim3(grid_for(n << lb)), dim3(kBlock), 0, s, n, lb, perm,
                       src, dst);
}

template <typename T>
void nonfinite_flags(long n, int lb, const T* x, int* flags, hipStream_t s)
{
    hipLaunchKernelGGL(k_nonfinite<T>, dim3(grid_for(n << lb)), dim3(kBlock), 0, s, n, lb, x, flags);
}

#define INST(T)                                                                                     \
    template void fwd_diag<T>(const DMesh&, const DConfig&, int, const T*, const T*, T*, T*, T*, T*, \
                              T*, T*, T*, T*, T*, T*, hipStream_t);                                 \
    template void fwd_tend_update_he<T>(const DMesh&, const DConfig&, int, int, const RKCoeffs&,    \
                                        const T*, const T*, const T*, const T*, const T*, T*, T*,   \
                                        T*, T*, T*, T*, hipStream_t);                               \
    template void tl_stage<T>(const DMesh&, const DConfig&, int, int, const RKCoeffs&,              \
                              const BaseStage&, const T*, const T*, T*, T*, T*, T*, T*, T*, T*, T*, \
                              T*, T*, hipStream_t);                                                 \
    template void adj_stage<T>(const DMesh&, const DConfig&, int, int, bool, const RKCoeffs&,       \
                               const BaseStage&, T*, T*, T*, T*, T*, T*, T*, T*, T*, T*, T*, T*, T*, \
                               hipStream_t);                                                        \
    template void adj_seed<T>(int, T*, const T*, double, hipStream_t);                              \
    template void reconstruct<T>(int, int, const int*, const int*, const double*, const double*,    \
                                 const T*, T*, hipStream_t);                                        \
    template void dot_batched<T>(long, int, const T*, const T*, const double*, double*, int,         \
                                 double*, int, hipStream_t);                                        \
    template void fill<T>(long, T*, T, hipStream_t);                                                \
    template void gather_rows<T>(long, int, const int*, const T*, T*, hipStream_t);                 \
    template void scatter_rows<T>(long, int, const int*, const T*, T*, hipStream_t);                \
    template void nonfinite_flags<T>(long, int, const T*, int*, hipStream_t);

INST(double)
INST(float)
template void convert<float, double>(long, float*, const double*, hipStream_t);
template void convert<double, float>(long, double*, const float*, hipStream_t);
template void convert<double, double>(long, double*, const double*, hipStream_t);
template void convert<float, float>(long, float*, const float*, hipStream_t);

}  // namespace msw

namespace msw {
namespace {
__global__ __launch_bounds__(kBlock) void k_diag_extra(DMesh m, const double* __restrict__ u,
                                                       const double* __restrict__ vort,
                                                       double* __restrict__ circ,
                                                       double* __restrict__ div,
                                                       double* __restrict__ vc)
{
    GRID_LOOP(static_cast<long>(m.nV) + m.nC)
    {
        if (t < m.nV) {
            const int iv = static_cast<int>(t);
            double s = 0.0;
            for (int i = 0; i < kVertexDegree; ++i) {
                const int e = m.eovR[3 * iv + i];
                s = s + m.sgnV[3 * iv + i] * (m.dc[e] * u[e]);
            }
            circ[iv] = s;
        } else {
            const int ic = static_cast<int>(t - m.nV);
            const int n = m.nEoC[ic];
            double d = 0.0;
            for (int k = 0; k < n; ++k) {
                const int e = m.eocR[kMaxEdges * ic + k];
                d = d + m.sgnCR[kMaxEdges * ic + k] * (u[e] * m.dv[e]);
            }
            const double r = 1.0 / m.areaC[ic];
            div[ic] = d * r;
            double w = 0.0;
            for (int k = 0; k < n; ++k)
                w = w + m.kiteCR[kMaxEdges * ic + k] * vort[m.vocR[kMaxEdges * ic + k]] / m.areaC[ic];
            vc[ic] = w;
        }
    }
}
}  // namespace

void diag_extra(const DMesh& m, const double* u, const double* vort, double* circ, double* div,
                double* vc, hipStream_t s)
{
    hipLaunchKernelGGL(k_diag_extra, dim3(grid_for(static_cast<long>(m.nV) + m.nC)), dim3(kBlock),
                       0, s, m, u, vort, circ, div, vc);
}
}  // namespace msw
