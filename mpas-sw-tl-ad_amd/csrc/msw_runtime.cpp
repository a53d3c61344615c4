// Host runtime of libmpas_sw_amd: mesh tables, device buffers, the RK4
// forward / tangent-linear / adjoint drivers and the extern "C" boundary
// declared in include/mpas_sw_amd.h.
//
// The drivers restate the reference's RK4 routines
// (src/mpas_sw_module.f90:528-1002) as sequences of stage kernels on device
// buffers; the per-stage kernels live in msw_kernels.hip.
#include "../../include/mpas_sw_amd.h"
#include "msw_internal.hpp"

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <tuple>
#include <vector>

using namespace msw;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return fail(MSW_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
    } while (0)

// Device allocation owned by a context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct msw_ctx {
    int prec = MSW_FP64;
    int B = 1, lb = 0;
    int nC = 0, nE = 0, nV = 0;
    hipStream_t stream = nullptr;
    msw_config cfg{};
    DConfig dcfg{};
    DMesh dm{};
    std::vector<DevBuf> bufs;
    int64_t bytes = 0;

    // reconstruction tables
    const double* coeffs = nullptr;   // [C][6][3]
    const double* trig = nullptr;     // [C][4]
    // permutations (device): perm[k] = reference index of internal entity k
    const int* permC = nullptr;
    const int* permE = nullptr;
    const int* permV = nullptr;

    // FP64 base (B = 1) buffers
    BaseStage st[4]{};
    double* bacc = nullptr;           // [C+E] accumulator of the base RK4
    double* bx = nullptr;             // [C+E] step-start copy for the adjoint
    double* bend = nullptr;           // [C+E] scratch end state
    double* dextra = nullptr;         // [V + 2C] circulation, divergence, vorticity_cell

    // member scratch ([entity][B] of prec)
    void* macc = nullptr;             // [C+E]
    void* mpA = nullptr;              // [C+E]
    double* dgb = nullptr;            // 13 FP64 diagnostics for msw_diagnostics_adj
    void* mpB = nullptr;              // [C+E]
    void* mv1 = nullptr;              // [V]
    void* mv2 = nullptr;              // [V]
    void* mc1 = nullptr;              // [C]
    void* mc2 = nullptr;              // [C]
    void* mc3 = nullptr;              // [C]
    // del2/del4 scratch, allocated when a viscosity is first switched on
    void* xdiv = nullptr;             // [C]  member precision
    void* xdsu = nullptr;             // [E]
    void* xdsd = nullptr;             // [C]
    void* xdsv = nullptr;             // [V]
    void* xvort = nullptr;            // [V]
    double* bdiv = nullptr;           // [C]  base (B = 1, FP64)
    double* bdsu = nullptr;           // [E]
    double* bdsd = nullptr;           // [C]
    double* bdsv = nullptr;           // [V]
    double* usrc = nullptr;           // [E]  wind-stress source, internal order
    double* msd = nullptr;            // [2E] meshScalingDel2, meshScalingDel4 (internal order)
    bool msd_on = false;              // false: the factors are all 1
    AdjCoef qs[4]{};                  // per-stage coefficients of the two-stream steps
    // second base-stage / coefficient set and step-start state: a multi-step
    // adjoint sweep forms the next step's set while the current one is used
    BaseStage stB[4]{};
    AdjCoef qsB[4]{};
    double* bxB = nullptr;            // [C+E]
    double* st0 = nullptr;            // [C+E] copy of the TLM step-start state (stage 0)
    hipStream_t side = nullptr;       // base trajectory stream (default priority)
    hipEvent_t ev[8]{};               // fork, stage-ready x4, join
    bool tl_fused = false;            // MSW_OPT_TL_FUSED_FLUX
    void* me1 = nullptr;              // [E]
    void* me2 = nullptr;              // [E]
    void* me3 = nullptr;              // [E]
    double* inv = nullptr;            // [2C + 3V + 2E + 2C] invariants scratch, made on first use
    // persistent path (MSW_OPT_PERSISTENT): work items and counters
    bool persistent = false;
    unsigned spin_limit = 1u << 20;
    std::vector<int> inst_c{0}, inst_e{0}, inst_v{0};   // instance starts (internal order)
    const int2* patches = nullptr;
    int np_v = 0, np_c = 0, np_e = 0;
    unsigned long long* pcount = nullptr;   // queue counters [2][8], barrier, error flag
    int* hflag = nullptr;                   // pinned host copy of the error flag
    double* dpartial = nullptr;       // dot partials
    int dot_nblk = 0;
    // msw_set_tend_output: where the last RK stage leaves its tendencies
    // (h, u, h_tl, u_tl); null: not kept
    void* tout[4]{};

    // graphs, keyed by (kind * 16 + variant, the pointers they bake in);
    // at most kMaxGraphs are kept, the least recently launched is dropped
    struct GraphEntry {
        hipGraphExec_t exec;
        unsigned long last;
    };
    static constexpr size_t kMaxGraphs = 32;
    bool use_graphs = false;
    hipStream_t cap_stream = nullptr;
    std::map<std::tuple<int, const void*, const void*, const void*, const void*>, GraphEntry> graphs;
    unsigned long graph_clock = 0;

    void clear_graphs()
    {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second.exec);
        graphs.clear();
    }

    ~msw_ctx()
    {
        clear_graphs();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        if (side) (void)hipStreamDestroy(side);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& b : bufs)
            if (b.p) (void)hipFree(b.p);
        if (hflag) (void)hipHostFree(hflag);
    }

    template <typename T>
    int alloc(T** out, size_t n)
    {
        void* p = nullptr;
        size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            *out = nullptr;
            return fail(MSW_ENOMEM, std::string("hipMalloc(") + std::to_string(bytes) +
                                        "): " + hipGetErrorString(e));
        }
        bufs.push_back({p, bytes});
        this->bytes += static_cast<int64_t>(bytes);
        *out = static_cast<T*>(p);
        return MSW_OK;
    }

    template <typename T>
    int upload(const T** out, const std::vector<T>& host)
    {
        T* p = nullptr;
        int rc = alloc(&p, host.size());
        if (rc) return rc;
        HIP_TRY(hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
        *out = p;
        return MSW_OK;
    }

    size_t esz() const { return prec == MSW_FP64 ? sizeof(double) : sizeof(float); }
};

namespace {

bool is_pow2(int b) { return b > 0 && (b & (b - 1)) == 0; }

// Build the internal-order mesh tables on the host and upload them.
int build_mesh(msw_ctx* c, const msw_mesh_desc* d)
{
    const int C = d->nCells, E = d->nEdges, V = d->nVertices;
    const int mE = d->maxEdges, mE2 = d->maxEdges2, vd = d->vertexDegree;
    if (C <= 0 || E <= 0 || V <= 0) return fail(MSW_EINVAL, "empty mesh");
    if (vd != kVertexDegree) return fail(MSW_EUNSUPPORTED, "vertexDegree must be 3");
    if (mE < 1 || mE2 < 1) return fail(MSW_EINVAL, "maxEdges/maxEdges2 must be positive");
    const void* req[] = {d->cellsOnEdge, d->verticesOnEdge, d->nEdgesOnCell, d->edgesOnCell,
                         d->nEdgesOnEdge, d->edgesOnEdge, d->weightsOnEdge, d->cellsOnVertex,
                         d->kiteAreasOnVertex, d->dcEdge, d->dvEdge, d->areaCell,
                         d->areaTriangle, d->fVertex, d->h_s, d->latCell, d->lonCell};
    for (const void* p : req)
        if (!p) return fail(MSW_EINVAL, "a required mesh array is NULL");

    // permutations: new -> old, and inverse
    auto mkperm = [](const int32_t* p, int n, std::vector<int>& fwd, std::vector<int>& inv) {
        fwd.resize(n);
        inv.assign(n, -1);
        for (int k = 0; k < n; ++k) {
            int o = p ? p[k] : k;
            if (o < 0 || o >= n || inv[o] != -1) return false;
            fwd[k] = o;
            inv[o] = k;
        }
        return true;
    };
    std::vector<int> pc, ic, pe, ie, pv, iv;
    if (!mkperm(d->cellPerm, C, pc, ic) || !mkperm(d->edgePerm, E, pe, ie) ||
        !mkperm(d->vertexPerm, V, pv, iv))
        return fail(MSW_EINVAL, "a permutation is not a bijection");

    auto chk = [](int32_t x, int n) { return x >= 1 && x <= n; };
    // ---- validate reference connectivity
    for (int e = 0; e < E; ++e) {
        for (int k = 0; k < 2; ++k) {
            if (!chk(d->cellsOnEdge[2 * e + k], C))
                return fail(MSW_EINVAL, "cellsOnEdge out of range at edge " + std::to_string(e + 1));
            if (!chk(d->verticesOnEdge[2 * e + k], V))
                return fail(MSW_EINVAL, "verticesOnEdge out of range at edge " + std::to_string(e + 1));
        }
        int n = d->nEdgesOnEdge[e];
        if (n < 0 || n > std::min(mE2, kMaxEdges2))
            return fail(n > kMaxEdges2 ? MSW_EUNSUPPORTED : MSW_EINVAL,
                        "nEdgesOnEdge out of range at edge " + std::to_string(e + 1));
        if (n > kEdgesOnEdge)
            return fail(MSW_EINVAL, "edge " + std::to_string(e + 1) + " has " + std::to_string(n) +
                                        " edgesOnEdge; cells of at most 6 edges give at most 10");
        for (int j = 0; j < n; ++j)
            if (!chk(d->edgesOnEdge[static_cast<long>(mE2) * e + j], E))
                return fail(MSW_EINVAL, "edgesOnEdge out of range at edge " + std::to_string(e + 1));
        if (!(d->dcEdge[e] > 0.0) || !(d->dvEdge[e] > 0.0))
            return fail(MSW_EINVAL, "non-positive dcEdge/dvEdge at edge " + std::to_string(e + 1));
    }
    for (int i = 0; i < C; ++i) {
        int n = d->nEdgesOnCell[i];
        if (n < 1 || n > std::min(mE, kMaxEdges))
            return fail(n > kMaxEdges ? MSW_EUNSUPPORTED : MSW_EINVAL,
                        "nEdgesOnCell out of range at cell " + std::to_string(i + 1));
        for (int k = 0; k < n; ++k)
            if (!chk(d->edgesOnCell[static_cast<long>(mE) * i + k], E))
                return fail(MSW_EINVAL, "edgesOnCell out of range at cell " + std::to_string(i + 1));
        if (!(d->areaCell[i] > 0.0))
            return fail(MSW_EINVAL, "non-positive areaCell at cell " + std::to_string(i + 1));
    }
    for (int v = 0; v < V; ++v) {
        for (int k = 0; k < 3; ++k)
            if (!chk(d->cellsOnVertex[3 * v + k], C))
                return fail(MSW_EINVAL, "cellsOnVertex out of range at vertex " + std::to_string(v + 1));
        if (!(d->areaTriangle[v] > 0.0))
            return fail(MSW_EINVAL, "non-positive areaTriangle at vertex " + std::to_string(v + 1));
    }

    // ---- cells
    std::vector<int> nEoC(C), eoc(static_cast<size_t>(C) * kMaxEdges, 0);
    std::vector<double> sgnC(static_cast<size_t>(C) * kMaxEdges, 0.0);
    std::vector<int> eocR(static_cast<size_t>(C) * kMaxEdges, 0), vocR(static_cast<size_t>(C) * kMaxEdges, 0);
    std::vector<double> sgnCR(static_cast<size_t>(C) * kMaxEdges, 0.0), kiteCR(static_cast<size_t>(C) * kMaxEdges, 0.0),
        keC(static_cast<size_t>(C) * kMaxEdges, 0.0), dvCR(static_cast<size_t>(C) * kMaxEdges, 0.0);
    std::vector<double> areaC(C), hs(C), trig(static_cast<size_t>(C) * 4),
        coeffs(static_cast<size_t>(C) * kMaxEdges * 3, 0.0);
    // edges of each reference cell in ascending reference edge index
    std::vector<std::vector<int>> cellEdges(C), cellVerts(C);
    for (int e = 0; e < E; ++e) {
        int c1 = d->cellsOnEdge[2 * e] - 1, c2 = d->cellsOnEdge[2 * e + 1] - 1;
        cellEdges[c1].push_back(e);
        if (c2 != c1) cellEdges[c2].push_back(e);
    }
    for (int v = 0; v < V; ++v)
        for (int k = 0; k < 3; ++k) {
            int cc = d->cellsOnVertex[3 * v + k] - 1;
            if (cellVerts[cc].empty() || cellVerts[cc].back() != v) cellVerts[cc].push_back(v);
        }
    for (int k = 0; k < C; ++k) {
        const int o = pc[k];
        const int n = d->nEdgesOnCell[o];
        nEoC[k] = n;
        if (static_cast<int>(cellEdges[o].size()) != n || static_cast<int>(cellVerts[o].size()) != n)
            return fail(MSW_EINVAL, "cell " + std::to_string(o + 1) +
                                        ": edgesOnCell, cellsOnEdge and cellsOnVertex disagree");
        for (int i = n; i < kMaxEdges; ++i) {
            // padding: a valid neighbour with a zero sign / kite
            eoc[kMaxEdges * k + i] = ie[d->edgesOnCell[static_cast<long>(mE) * o] - 1];
            eocR[kMaxEdges * k + i] = ie[cellEdges[o][0]];
            dvCR[kMaxEdges * k + i] = d->dvEdge[cellEdges[o][0]];
            vocR[kMaxEdges * k + i] = iv[cellVerts[o][0]];
        }
        for (int i = 0; i < n; ++i) {
            int e = d->edgesOnCell[static_cast<long>(mE) * o + i] - 1;
            eoc[kMaxEdges * k + i] = ie[e];
            sgnC[kMaxEdges * k + i] = (d->cellsOnEdge[2 * e] - 1 == o) ? 1.0 : -1.0;
            keC[kMaxEdges * k + i] = 0.25 * d->dcEdge[e] * d->dvEdge[e];
            int er = cellEdges[o][i];
            eocR[kMaxEdges * k + i] = ie[er];
            dvCR[kMaxEdges * k + i] = d->dvEdge[er];
            sgnCR[kMaxEdges * k + i] = (d->cellsOnEdge[2 * er] - 1 == o) ? 1.0 : -1.0;
            int vr = cellVerts[o][i];
            vocR[kMaxEdges * k + i] = iv[vr];
            double kite = 0.0;
            for (int s = 0; s < 3; ++s)
                if (d->cellsOnVertex[3 * vr + s] - 1 == o) kite = d->kiteAreasOnVertex[3 * vr + s];
            kiteCR[kMaxEdges * k + i] = kite;
            if (d->coeffs_reconstruct)
                for (int r = 0; r < 3; ++r)
                    coeffs[(static_cast<size_t>(k) * kMaxEdges + i) * 3 + r] =
                        d->coeffs_reconstruct[(static_cast<size_t>(o) * mE + i) * 3 + r];
        }
        areaC[k] = d->areaCell[o];
        hs[k] = d->h_s[o];
        // one sincos per angle, as the reference's compiled reconstruction
        // evaluates them (src/mpas_sw_operators.f90:350-353)
        sincos(d->latCell[o], &trig[4 * k + 1], &trig[4 * k]);
        sincos(d->lonCell[o], &trig[4 * k + 3], &trig[4 * k + 2]);
    }
    // ---- edges
    std::vector<int> coe(2 * static_cast<size_t>(E)), voe(2 * static_cast<size_t>(E)), nEoE(E),
        eoe(static_cast<size_t>(E) * kMaxEdges2, 0), eoeT(static_cast<size_t>(E) * kMaxEdges2, 0);
    std::vector<double> woe(static_cast<size_t>(E) * kMaxEdges2, 0.0), dc(E), dv(E);
    for (int k = 0; k < E; ++k) {
        const int o = pe[k];
        coe[2 * k] = ic[d->cellsOnEdge[2 * o] - 1];
        coe[2 * k + 1] = ic[d->cellsOnEdge[2 * o + 1] - 1];
        voe[2 * k] = iv[d->verticesOnEdge[2 * o] - 1];
        voe[2 * k + 1] = iv[d->verticesOnEdge[2 * o + 1] - 1];
        const int n = d->nEdgesOnEdge[o];
        nEoE[k] = n;
        for (int j = n; j < kMaxEdges2; ++j) {
            // padding: the edge itself with a zero weight; the transposed
            // slot points at this edge's first padded (zero-weight) slot
            eoe[kMaxEdges2 * k + j] = k;
            eoeT[kMaxEdges2 * k + j] = n;
        }
        for (int j = 0; j < n; ++j) {
            const int eo = d->edgesOnEdge[static_cast<long>(mE2) * o + j] - 1;
            eoe[kMaxEdges2 * k + j] = ie[eo];
            woe[kMaxEdges2 * k + j] = d->weightsOnEdge[static_cast<long>(mE2) * o + j];
            // slot of o in eo's list
            int jt = -1;
            const int n2 = d->nEdgesOnEdge[eo];
            for (int jj = 0; jj < n2; ++jj)
                if (d->edgesOnEdge[static_cast<long>(mE2) * eo + jj] - 1 == o) { jt = jj; break; }
            if (jt < 0)
                return fail(MSW_EINVAL, "edgesOnEdge is not symmetric at edge " + std::to_string(o + 1));
            eoeT[kMaxEdges2 * k + j] = jt;
        }
        dc[k] = d->dcEdge[o];
        dv[k] = d->dvEdge[o];
    }
    // slot of each edge in the edgesOnCell list of its two cells
    std::vector<int> eSlot(2 * static_cast<size_t>(E), 0);
    for (int k = 0; k < C; ++k)
        for (int i = 0; i < nEoC[k]; ++i) {
            const int e = eoc[kMaxEdges * static_cast<size_t>(k) + i];
            eSlot[2 * static_cast<size_t>(e) + (coe[2 * static_cast<size_t>(e)] == k ? 0 : 1)] = i;
        }
    std::vector<double> woeT(static_cast<size_t>(E) * kMaxEdges2, 0.0);
    for (size_t k = 0; k < static_cast<size_t>(E); ++k)
        for (int j = 0; j < kMaxEdges2; ++j)
            woeT[kMaxEdges2 * k + j] =
                woe[kMaxEdges2 * static_cast<size_t>(eoe[kMaxEdges2 * k + j]) + eoeT[kMaxEdges2 * k + j]];
    // ---- vertices
    std::vector<std::vector<int>> vertEdges(V);
    for (int e = 0; e < E; ++e) {
        int v1 = d->verticesOnEdge[2 * e] - 1, v2 = d->verticesOnEdge[2 * e + 1] - 1;
        vertEdges[v1].push_back(e);
        if (v2 != v1) vertEdges[v2].push_back(e);
    }
    std::vector<int> eovR(3 * static_cast<size_t>(V)), cov(3 * static_cast<size_t>(V));
    std::vector<double> sgnV(3 * static_cast<size_t>(V)), sdcV(3 * static_cast<size_t>(V)),
        kite(3 * static_cast<size_t>(V)), areaT(V), fV(V);
    for (int k = 0; k < V; ++k) {
        const int o = pv[k];
        if (vertEdges[o].size() != 3)
            return fail(MSW_EINVAL, "vertex " + std::to_string(o + 1) + " is not on exactly 3 edges");
        for (int i = 0; i < 3; ++i) {
            const int e = vertEdges[o][i];
            eovR[3 * k + i] = ie[e];
            sgnV[3 * k + i] = (d->verticesOnEdge[2 * e + 1] - 1 == o) ? 1.0 : -1.0;
            sdcV[3 * k + i] = sgnV[3 * k + i] * d->dcEdge[e];
            cov[3 * k + i] = ic[d->cellsOnVertex[3 * o + i] - 1];
            kite[3 * k + i] = d->kiteAreasOnVertex[3 * o + i];
        }
        areaT[k] = d->areaTriangle[o];
        fV[k] = d->fVertex[o];
    }

    DMesh& m = c->dm;
    m.nC = C; m.nE = E; m.nV = V;
    m.stripe_log2 = 2;
    m.block = 256;
    int rc = 0;
#define UP(dst, vec) if ((rc = c->upload(&(dst), vec))) return rc
    UP(m.nEoC, nEoC); UP(m.eoc, eoc); UP(m.sgnC, sgnC); UP(m.eocR, eocR); UP(m.sgnCR, sgnCR); UP(m.dvCR, dvCR);
    UP(m.vocR, vocR); UP(m.kiteCR, kiteCR); UP(m.keC, keC); UP(m.areaC, areaC); UP(m.hs, hs);
    UP(m.coe, coe); UP(m.voe, voe); UP(m.nEoE, nEoE); UP(m.eoe, eoe); UP(m.woe, woe);
    UP(m.eoeT, eoeT); UP(m.woeT, woeT); UP(m.eSlot, eSlot); UP(m.dc, dc); UP(m.dv, dv);
    UP(m.eovR, eovR); UP(m.sgnV, sgnV); UP(m.sdcV, sdcV); UP(m.cov, cov); UP(m.kite, kite); UP(m.areaT, areaT);
    UP(m.fV, fV);
    UP(c->coeffs, coeffs); UP(c->trig, trig);
    std::vector<int> vpc(pc.begin(), pc.end()), vpe(pe.begin(), pe.end()), vpv(pv.begin(), pv.end());
    UP(c->permC, vpc); UP(c->permE, vpe); UP(c->permV, vpv);
#undef UP
    return MSW_OK;
}

int alloc_buffers(msw_ctx* c)
{
    const size_t C = c->nC, E = c->nE, V = c->nV, B = c->B;
    int rc;
    for (BaseStage* set : {c->st, c->stB})
        for (int s = 0; s < 4; ++s) {
            BaseStage& b = set[s];
            if (s > 0) {
                if ((rc = c->alloc(&b.h, C + E))) return rc;
                b.u = b.h + C;
            }
            double* eb = nullptr;
            if ((rc = c->alloc(&eb, 5 * E + 2 * C + 3 * V))) return rc;
            b.he = eb; b.pe = eb + E; b.v = eb + 2 * E; b.gpt = eb + 3 * E; b.gpn = eb + 4 * E;
            b.ke = eb + 5 * E; b.pc = b.ke + C;
            b.pvv = b.pc + C; b.hv = b.pvv + V; b.vort = b.hv + V;
            if ((rc = c->alloc(&b.rec, 4 * E))) return rc;
        }
    if ((rc = c->alloc(&c->bacc, C + E))) return rc;
    if ((rc = c->alloc(&c->bx, C + E))) return rc;
    if ((rc = c->alloc(&c->bxB, C + E))) return rc;
    if ((rc = c->alloc(&c->bend, C + E))) return rc;
    if ((rc = c->alloc(&c->dextra, V + 2 * C))) return rc;
    const size_t es = c->esz();
    char* mem = nullptr;
    const size_t total = (3 * (C + E) + 2 * V + 3 * C + 3 * E) * B * es;
    if ((rc = c->alloc(&mem, total))) return rc;
    size_t off = 0;
    auto take = [&](void** p, size_t rows) { *p = mem + off; off += rows * B * es; };
    take(&c->macc, C + E); take(&c->mpA, C + E); take(&c->mpB, C + E);
    take(&c->mv1, V); take(&c->mv2, V); take(&c->mc1, C); take(&c->mc2, C); take(&c->mc3, C);
    take(&c->me1, E); take(&c->me2, E); take(&c->me3, E);
    for (AdjCoef* set : {c->qs, c->qsB})
        for (int st = 0; st < 4; ++st) {
            double* qq = nullptr;
            if ((rc = c->alloc(&qq, 64 * E + 8 * V + 32 * C))) return rc;
            set[st].e48 = qq;
            set[st].e8 = qq + 48 * E;
            set[st].v8 = qq + 64 * E;
            set[st].c32 = qq + 64 * E + 8 * V;
        }
    if ((rc = c->alloc(&c->st0, C + E))) return rc;
    c->dot_nblk = dot_blocks(std::max<long>(C, E));
    if ((rc = c->alloc(&c->dpartial, static_cast<size_t>(c->dot_nblk) * B + 2 * B))) return rc;
    return MSW_OK;
}

// Scratch of the del2/del4 terms, made when a viscosity is first on.
int ensure_dissip(msw_ctx* c)
{
    if (!(c->cfg.visc2 > 0.0 || c->cfg.visc4 > 0.0) || c->xdiv) return MSW_OK;
    const size_t C = c->nC, E = c->nE, V = c->nV, B = c->B;
    int rc;
    char* mem = nullptr;
    if ((rc = c->alloc(&mem, (2 * C + E + 2 * V) * B * c->esz()))) return rc;
    size_t off = 0;
    auto take = [&](void** p, size_t rows) { *p = mem + off; off += rows * B * c->esz(); };
    take(&c->xdiv, C); take(&c->xdsu, E); take(&c->xdsd, C); take(&c->xdsv, V); take(&c->xvort, V);
    double* bm = nullptr;
    if ((rc = c->alloc(&bm, 2 * C + E + V))) return rc;
    c->bdiv = bm; c->bdsu = bm + C; c->bdsd = bm + C + E; c->bdsv = bm + 2 * C + E;
    return MSW_OK;
}

void set_dcfg(msw_ctx* c)
{
    const msw_config& g = c->cfg;
    c->dcfg.dt = g.dt;
    c->dcfg.apvm = g.apvm_upwinding;
    c->dcfg.frozen_u = g.test_case == 1 ? 1 : 0;
    c->dcfg.visc2 = g.visc2;
    c->dcfg.visc4 = g.visc4;
    c->dcfg.wind = g.wind_stress ? 1 : 0;
    c->dcfg.drag = g.bottom_drag ? 1 : 0;
    c->dcfg.usrc = c->usrc;
    c->dcfg.msd2 = c->msd_on ? c->msd : nullptr;
    c->dcfg.msd4 = c->msd_on ? c->msd + c->nE : nullptr;
    c->dcfg.tl_fused = c->tl_fused ? 1 : 0;
}

template <typename T>
DissipScratch<T> member_ds(msw_ctx* c)
{
    return {static_cast<T*>(c->xdiv), static_cast<T*>(c->xdsu), static_cast<T*>(c->xdsd),
            static_cast<T*>(c->xdsv)};
}

DissipScratch<double> base_ds(msw_ctx* c) { return {c->bdiv, c->bdsu, c->bdsd, c->bdsv}; }

// base (B = 1) tendency + RK update of one stage, with the dissipation terms
// (the last stage leaves its tendency in tout[0..1] when they are set)
void base_tend(msw_ctx* c, const BaseStage& b, int s, const RKCoeffs& k, double* xh, double* xu,
               double* nh, hipStream_t st)
{
    const Dissip<double> dp = dissip_fields<double>(c->dm, c->dcfg, 0, b.u, b.vort, base_ds(c), st);
    const size_t C = c->nC;
    fwd_tend_update_he<double>(c->dm, c->dcfg, 0, s, k, b.h, b.u, b.he, b.pe, b.ke, dp, xh, xu,
                               c->bacc, c->bacc + C, nh, nh + C, b.rec, static_cast<double*>(c->tout[0]),
                               static_cast<double*>(c->tout[1]), st);
}

// The side stream joins the main stream's work (and a capture in progress).
int fork_side(msw_ctx* c)
{
    HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    HIP_TRY(hipStreamWaitEvent(c->side, c->ev[0], 0));
    return MSW_OK;
}

int join_side(msw_ctx* c)
{
    HIP_TRY(hipEventRecord(c->ev[5], c->side));
    HIP_TRY(hipStreamWaitEvent(c->stream, c->ev[5], 0));
    return MSW_OK;
}

// Run `work` between fork_side and join_side.  The side stream is joined back
// also when `work` fails, so a graph capture in progress ends well formed and
// the error is what the caller sees.
template <typename F>
int forked(msw_ctx* c, F&& work)
{
    int rc = fork_side(c);
    if (rc) return rc;
    rc = work();
    const int jrc = join_side(c);
    return rc ? rc : jrc;
}

// ---------------------------------------------------------------------------
// drivers

template <typename T>
void enqueue_fwd_step(msw_ctx* c, T* h, T* u)
{
    const DMesh& m = c->dm;
    const RKCoeffs k = rk_coeffs(c->cfg.dt);
    const size_t C = c->nC, E = c->nE, B = c->B;
    T* acc = static_cast<T*>(c->macc);
    T* pA = static_cast<T*>(c->mpA);
    T* pB = static_cast<T*>(c->mpB);
    T* pvv = static_cast<T*>(c->mv1);
    T* ke = static_cast<T*>(c->mc1);
    T* pc = static_cast<T*>(c->mc2);
    T* he = static_cast<T*>(c->me1);
    T* pe = static_cast<T*>(c->me2);
    const T* rh = h;
    const T* ru = u;
    T* vort = static_cast<T*>(c->xvort);   // null unless a viscosity is on
    for (int s = 0; s < 4; ++s) {
        fwd_diag<T>(m, c->dcfg, c->lb, rh, ru, pvv, nullptr, vort, ke, pc, he, pe, nullptr,
                    nullptr, nullptr, nullptr, c->stream);
        const Dissip<T> dp = dissip_fields<T>(m, c->dcfg, c->lb, ru, vort, member_ds<T>(c), c->stream);
        T* nb = (s % 2 == 0) ? pA : pB;
        fwd_tend_update_he<T>(m, c->dcfg, c->lb, s, k, rh, ru, he, pe, ke, dp, h, u, acc,
                              acc + C * B, nb, nb + C * B, nullptr, static_cast<T*>(c->tout[0]),
                              static_cast<T*>(c->tout[1]), c->stream);
        rh = nb;
        ru = nb + C * B;
    }
    (void)E;
}

// FP64 B=1 base step storing per-stage fields in c->st (stage 0 reads x).
void enqueue_base_step(msw_ctx* c, double* h, double* u)
{
    const DMesh& m = c->dm;
    const RKCoeffs k = rk_coeffs(c->cfg.dt);
    const size_t C = c->nC;
    c->st[0].h = h;
    c->st[0].u = u;
    for (int s = 0; s < 4; ++s) {
        BaseStage& b = c->st[s];
        fwd_diag<double>(m, c->dcfg, 0, b.h, b.u, b.pvv, b.hv, b.vort, b.ke, b.pc, b.he, b.pe, b.v,
                         b.gpt, b.gpn, b.rec, c->stream);
        double* nh = s < 3 ? c->st[s + 1].h : c->bend;
        base_tend(c, b, s, k, h, u, nh, c->stream);
    }
    (void)C;
}

// One TLM step.  The FP64 base trajectory (diagnostics, coefficient rows and
// RK update of each stage, B = 1) runs on the side stream, a stage ahead of
// the member kernels on the main stream, which wait only for the base
// fields of the stage they are about to use.
template <typename T>
int enqueue_tlm_step(msw_ctx* c, double* h, double* u, T* htl, T* utl)
{
    const DMesh& m = c->dm;
    const RKCoeffs k = rk_coeffs(c->cfg.dt);
    const size_t C = c->nC, E = c->nE, B = c->B;
    T* acc = static_cast<T*>(c->macc);
    T* pA = static_cast<T*>(c->mpA);
    T* pB = static_cast<T*>(c->mpB);
    T* pvv_tl = static_cast<T*>(c->mv1);
    T* ke_tl = static_cast<T*>(c->mc1);
    T* pc_tl = static_cast<T*>(c->mc2);
    T* pe_tl = static_cast<T*>(c->me1);
    T* he_tl = static_cast<T*>(c->me2);
    hipStream_t ss = c->side;
    return forked(c, [&]() -> int {
    // stage 0 reads a copy, so the base RK update may overwrite (h, u) while
    // the member kernels of earlier stages still run
    HIP_TRY(hipMemcpyAsync(c->st0, h, C * sizeof(double), hipMemcpyDeviceToDevice, ss));
    HIP_TRY(hipMemcpyAsync(c->st0 + C, u, E * sizeof(double), hipMemcpyDeviceToDevice, ss));
    c->st[0].h = c->st0;
    c->st[0].u = c->st0 + C;
    for (int s = 0; s < 4; ++s) {
        BaseStage& b = c->st[s];
        fwd_diag<double>(m, c->dcfg, 0, b.h, b.u, b.pvv, b.hv, b.vort, b.ke, b.pc, b.he, b.pe, b.v,
                         b.gpt, b.gpn, b.rec, ss);
        tl_coef(m, c->dcfg, b, c->qs[s], c->prec == MSW_FP32, ss);
        HIP_TRY(hipEventRecord(c->ev[1 + s], ss));
        double* nh = s < 3 ? c->st[s + 1].h : c->bend;
        base_tend(c, b, s, k, h, u, nh, ss);
    }
    const T* th = htl;
    const T* tu = utl;
    for (int s = 0; s < 4; ++s) {
        HIP_TRY(hipStreamWaitEvent(c->stream, c->ev[1 + s], 0));
        T* nb = (s % 2 == 0) ? pA : pB;
        tl_stage<T>(m, c->dcfg, c->lb, s, k, c->st[s], th, tu, pvv_tl, ke_tl, pc_tl, pe_tl, he_tl,
                    static_cast<T*>(c->xvort), member_ds<T>(c), c->qs[s], htl, utl, acc,
                    acc + C * B, nb, nb + C * B, static_cast<T*>(c->tout[2]), static_cast<T*>(c->tout[3]),
                    c->stream);
        th = nb;
        tu = nb + C * B;
    }
    return MSW_OK;
    });
}

// Side stream: the FP64 base trajectory of one TLM step (B = 1) into set
// (st, qs): the step-start state (h, u) is copied into x, the stage-0 input
// the member kernels read, then each stage's diagnostics, TL coefficient
// rows and RK update run; stage 3 writes the step's end state into (h, u).
int side_tlm_base(msw_ctx* c, BaseStage* st, AdjCoef* qs, double* x, double* h, double* u)
{
    const DMesh& m = c->dm;
    const RKCoeffs k = rk_coeffs(c->cfg.dt);
    const size_t C = c->nC, E = c->nE;
    hipStream_t ss = c->side;
    HIP_TRY(hipMemcpyAsync(x, h, C * sizeof(double), hipMemcpyDeviceToDevice, ss));
    HIP_TRY(hipMemcpyAsync(x + C, u, E * sizeof(double), hipMemcpyDeviceToDevice, ss));
    st[0].h = x;
    st[0].u = x + C;
    for (int s = 0; s < 4; ++s) {
        BaseStage& b = st[s];
        fwd_diag<double>(m, c->dcfg, 0, b.h, b.u, b.pvv, b.hv, b.vort, b.ke, b.pc, b.he, b.pe, b.v,
                         b.gpt, b.gpn, b.rec, ss);
        tl_coef(m, c->dcfg, b, qs[s], c->prec == MSW_FP32, ss);
        double* nh = s < 3 ? st[s + 1].h : c->bend;
        base_tend(c, b, s, k, h, u, nh, ss);
    }
    return MSW_OK;
}

// One step of a multi-step TLM window, pipelined like the adjoint sweep: the
// member stages use base set `par`, formed while the previous step ran (or
// before the first), while the side stream forms the other set for the next
// step (when `next`) from the base state, which it then advances.  With no
// serial base prologue per step, every base kernel overlaps member work.
template <typename T>
int enqueue_tlm_step_piped(msw_ctx* c, int par, bool next, double* h, double* u, T* htl, T* utl)
{
    const DMesh& m = c->dm;
    const RKCoeffs k = rk_coeffs(c->cfg.dt);
    const size_t C = c->nC, B = c->B;
    T* acc = static_cast<T*>(c->macc);
    T* pA = static_cast<T*>(c->mpA);
    T* pB = static_cast<T*>(c->mpB);
    BaseStage* cur = par ? c->stB : c->st;
    AdjCoef* qcur = par ? c->qsB : c->qs;
    return forked(c, [&]() -> int {
        if (next) {
            const int rc = par ? side_tlm_base(c, c->st, c->qs, c->bx, h, u)
                               : side_tlm_base(c, c->stB, c->qsB, c->bxB, h, u);
            if (rc) return rc;
        }
        const T* th = htl;
        const T* tu = utl;
        for (int s = 0; s < 4; ++s) {
            T* nb = (s % 2 == 0) ? pA : pB;
            tl_stage<T>(m, c->dcfg, c->lb, s, k, cur[s], th, tu, static_cast<T*>(c->mv1), static_cast<T*>(c->mc1),
                        static_cast<T*>(c->mc2), static_cast<T*>(c->me1), static_cast<T*>(c->me2),
                        static_cast<T*>(c->xvort), member_ds<T>(c), qcur[s], htl, utl, acc, acc + C * B, nb,
                        nb + C * B, static_cast<T*>(c->tout[2]), static_cast<T*>(c->tout[3]), c->stream);
            th = nb;
            tu = nb + C * B;
        }
        return MSW_OK;
    });
}

// Side stream: the four base stages of an adjoint step whose start state is
// in x (FP64, B = 1) into `st`, then the coefficient rows of stages 3, 2, 1, 0
// into `qs`; with `events`, ev[1+s] is recorded once stage s's rows exist.
int side_adj_base(msw_ctx* c, BaseStage* st, AdjCoef* qs, double* x, bool want_end, bool events)
{
    const DMesh& m = c->dm;
    const RKCoeffs k = rk_coeffs(c->cfg.dt);
    const size_t C = c->nC;
    hipStream_t ss = c->side;
    st[0].h = x;
    st[0].u = x + C;
    for (int s = 0; s < 4; ++s) {
        BaseStage& b = st[s];
        fwd_diag<double>(m, c->dcfg, 0, b.h, b.u, b.pvv, b.hv, b.vort, b.ke, b.pc, b.he, b.pe, b.v,
                         b.gpt, b.gpn, b.rec, ss);
        if (s < 3 || want_end || c->tout[0] || c->tout[1]) {
            // stage 3 writes x_new into its "x" argument: give it the end buffer
            double* nh = s < 3 ? st[s + 1].h : c->bend;
            double* wxh = s < 3 ? x : c->bend;
            double* wxu = s < 3 ? x + C : c->bend + C;
            base_tend(c, b, s, k, wxh, wxu, nh, ss);
        }
    }
    if (events) {
        // stage by stage, so each member stage starts once its rows exist
        for (int s = 3; s >= 0; --s) {
            adj_coef(m, c->dcfg, st[s], qs[s], c->prec == MSW_FP32, ss);
            HIP_TRY(hipEventRecord(c->ev[1 + s], ss));
        }
    } else {
        adj_coef4(m, c->dcfg, st, qs, c->prec == MSW_FP32, ss);
    }
    return MSW_OK;
}

// Main stream: the member adjoint stages 3, 2, 1, 0 with coefficient set
// `qs`; with `wait`, stage s first waits for ev[1+s].  The first stage reads
// the tendency adjoint w[3]*lam from lam itself.
template <typename T>
int main_adj_stages(msw_ctx* c, const BaseStage* st, const AdjCoef* qs, T* lamh, T* lamu, bool wait)
{
    const DMesh& m = c->dm;
    const RKCoeffs k = rk_coeffs(c->cfg.dt);
    const size_t C = c->nC, B = c->B;
    T* g = static_cast<T*>(c->mpA);
    T* acc = static_cast<T*>(c->macc);
    // frozen wind: the u tendency adjoint (g + C*B) is read as zero by every
    // stage and never written (a fill kernel, so graph capture stays valid)
    if (c->dcfg.frozen_u) fill<T>(static_cast<long>(c->nE) * static_cast<long>(B), g + C * B, T(0), c->stream);
    for (int s = 3; s >= 0; --s) {
        if (wait) HIP_TRY(hipStreamWaitEvent(c->stream, c->ev[1 + s], 0));
        adj_stage<T>(m, c->dcfg, c->lb, s, s == 3, k, st[s], g, g + C * B,
                     static_cast<T*>(c->me1), static_cast<T*>(c->me2), static_cast<T*>(c->me3),
                     static_cast<T*>(c->mc1), static_cast<T*>(c->mc2), static_cast<T*>(c->mc3),
                     static_cast<T*>(c->mv1),
                     static_cast<T*>(c->mv2), member_ds<T>(c), qs[s], lamh, lamu, acc,
                     acc + C * B, c->stream);
    }
    return MSW_OK;
}

// One adjoint step; the base step-start state must already be in c->bx.
// The side stream recomputes the four base stages, then forms the
// coefficient rows of stages 3, 2, 1, 0 while the main stream runs the
// member adjoint stages in that order.
template <typename T>
int enqueue_adj_step(msw_ctx* c, T* lamh, T* lamu, bool want_end)
{
    return forked(c, [&]() -> int {
        int rc = side_adj_base(c, c->st, c->qs, c->bx, want_end, true);
        return rc ? rc : main_adj_stages<T>(c, c->st, c->qs, lamh, lamu, true);
    });
}

// One step of a multi-step adjoint sweep: the member stages use base set
// `par` (formed during the previous step), while the side stream forms the
// other set for the next step from the start state already copied into that
// set's buffer (when `next`).
template <typename T>
int enqueue_adj_step_piped(msw_ctx* c, int par, bool next, T* lamh, T* lamu)
{
    BaseStage* cur = par ? c->stB : c->st;
    AdjCoef* qcur = par ? c->qsB : c->qs;
    return forked(c, [&]() -> int {
        if (next) {
            BaseStage* nst = par ? c->st : c->stB;
            AdjCoef* nq = par ? c->qs : c->qsB;
            double* nx = par ? c->bx : c->bxB;
            const int rc = side_adj_base(c, nst, nq, nx, false, false);
            if (rc) return rc;
        }
        return main_adj_stages<T>(c, cur, qcur, lamh, lamu, false);
    });
}

// Patches of up to 256 consecutive entities of each kind that never cross an
// instance boundary, uploaded as the persistent path's work items.
int build_patches(msw_ctx* c)
{
    std::vector<int2> p;
    const std::vector<int>* starts[3] = {&c->inst_v, &c->inst_c, &c->inst_e};
    const int totals[3] = {c->nV, c->nC, c->nE};
    int counts[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        const std::vector<int>& st = *starts[k];
        for (size_t i = 0; i < st.size(); ++i) {
            const int a = st[i], b = i + 1 < st.size() ? st[i + 1] : totals[k];
            for (int q = a; q < b; q += 256) {
                p.push_back(make_int2(q, std::min(256, b - q)));
                ++counts[k];
            }
        }
    }
    const int2* dev = nullptr;
    int rc = c->upload(&dev, p);
    if (rc) return rc;
    c->patches = dev;
    c->np_v = counts[0];
    c->np_c = counts[1];
    c->np_e = counts[2];
    return MSW_OK;
}

int check_cfg(msw_ctx* c)
{
    const msw_config& g = c->cfg;
    if (g.visc2 < 0.0 || g.visc4 < 0.0) return fail(MSW_EINVAL, "negative viscosity");
    if (!(g.dt > 0.0)) return fail(MSW_EINVAL, "dt must be positive");
    return MSW_OK;
}

// Copy an FP64 state (h [C], u [E]) on the context's stream.
int store_state(msw_ctx* c, double* dh, double* du, const double* h, const double* u)
{
    HIP_TRY(hipMemcpyAsync(dh, h, static_cast<size_t>(c->nC) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(du, u, static_cast<size_t>(c->nE) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return MSW_OK;
}

// Run `body(v)` once per step, either directly or through a captured graph
// keyed by (kind, variant, pointers); step i uses variant `variant(i)`
// (0 when no variant function is given).
// `pre(i)` enqueues step i's copies outside the graphs and returns a status.
template <typename F>
int run_steps_var(msw_ctx* c, int kind, const void* p0, const void* p1, const void* p2, const void* p3,
                  int nsteps, F&& body, const std::function<int(int)>* pre,
                  const std::function<int(int)>* variant)
{
    for (int i = 0; i < nsteps; ++i) {
        if (pre) {
            const int prc = (*pre)(i);
            if (prc) return prc;
        }
        const int v = variant ? (*variant)(i) : 0;
        if (!c->use_graphs) {
            const int brc = body(v);
            if (brc) return brc;
            continue;
        }
        auto key = std::make_tuple(kind * 16 + v, p0, p1, p2, p3);
        auto it = c->graphs.find(key);
        if (it == c->graphs.end()) {
            if (c->graphs.size() >= msw_ctx::kMaxGraphs) {
                // drop the least recently launched graph once it has run
                auto lru = c->graphs.begin();
                for (auto g = c->graphs.begin(); g != c->graphs.end(); ++g)
                    if (g->second.last < lru->second.last) lru = g;
                HIP_TRY(hipStreamSynchronize(c->stream));
                (void)hipGraphExecDestroy(lru->second.exec);
                c->graphs.erase(lru);
            }
            if (!c->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
            hipStream_t user = c->stream;
            c->stream = c->cap_stream;
            hipGraph_t graph = nullptr;
            int brc = MSW_OK;
            hipError_t e = hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                brc = body(v);
                e = hipStreamEndCapture(c->cap_stream, &graph);
            }
            c->stream = user;
            if (brc) {
                if (graph) (void)hipGraphDestroy(graph);
                return brc;
            }
            if (e != hipSuccess) return fail(MSW_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
            hipGraphExec_t exec = nullptr;
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return fail(MSW_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            it = c->graphs.emplace(key, msw_ctx::GraphEntry{exec, 0}).first;
        }
        it->second.last = ++c->graph_clock;
        HIP_TRY(hipGraphLaunch(it->second.exec, c->stream));
    }
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

template <typename F>
int run_steps(msw_ctx* c, int kind, const void* p0, const void* p1, const void* p2, const void* p3,
              int nsteps, F&& body, const std::function<int(int)>* pre = nullptr)
{
    return run_steps_var(c, kind, p0, p1, p2, p3, nsteps, [&](int) { return body(); }, pre, nullptr);
}

// A window of forward (tlm = false; h, u: member state) or TLM steps (bh, bu:
// base state; h, u: TL state) through the persistent kernel, then a check of
// its error flag (this waits for the launch).
int run_persistent(msw_ctx* c, bool tlm, double* bh, double* bu, void* h, void* u, int nsteps, double* traj_h,
                   double* traj_u)
{
    if (c->B != 1) return fail(MSW_EUNSUPPORTED, "the persistent path runs one member per entity (batch 1)");
    const DConfig& d = c->dcfg;
    if (d.visc2 > 0.0 || d.visc4 > 0.0 || d.wind || d.drag)
        return fail(MSW_EUNSUPPORTED, "the persistent path has no viscosity, drag or wind-stress terms");
    if (nsteps == 0) return MSW_OK;
    int rc = MSW_OK;
    if (!c->patches && (rc = build_patches(c))) return rc;
    if (!c->pcount && (rc = c->alloc(&c->pcount, 18))) return rc;
    if (!c->hflag) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&c->hflag), sizeof(int), hipHostMallocDefault));
    const size_t C = c->nC;
    PersistArgs a{};
    a.m = c->dm;
    a.c = c->dcfg;
    a.k = rk_coeffs(c->cfg.dt);
    a.nsteps = nsteps;
    a.patch = c->patches;
    a.np_v = c->np_v;
    a.np_c = c->np_c;
    a.np_e = c->np_e;
    for (int s = 0; s < 4; ++s) a.st[s] = c->st[s];
    a.st[0].h = bh;
    a.st[0].u = bu;
    a.bh = bh;
    a.bu = bu;
    a.bacc = c->bacc;
    a.bend = c->bend;
    a.traj_h = traj_h;
    a.traj_u = traj_u;
    a.h = h;
    a.u = u;
    a.acc = c->macc;
    a.pA = c->mpA;
    a.pB = c->mpB;
    a.pvv = c->mv1;
    a.ke = c->mc1;
    a.pc = c->mc2;
    a.he = c->me2;
    a.pe = c->me1;
    a.queue = c->pcount;   // 2 x 8 queue counters
    a.bar = c->pcount + 16;
    a.err = reinterpret_cast<int*>(c->pcount + 17);
    a.spin_limit = c->spin_limit;
    (void)C;
    HIP_TRY(hipMemsetAsync(c->pcount, 0, 18 * sizeof(unsigned long long), c->stream));
    int grid = 0;
    const int e = c->prec == MSW_FP64 ? persistent_steps<double>(a, tlm, c->stream, &grid)
                                      : persistent_steps<float>(a, tlm, c->stream, &grid);
    if (e) return fail(MSW_EHIP, std::string("persistent launch: ") + hipGetErrorString(static_cast<hipError_t>(e)));
    HIP_TRY(hipMemcpyAsync(c->hflag, a.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (*c->hflag)
        return fail(MSW_EHIP, "persistent launch: a grid barrier waited longer than its limit (" +
                                  std::to_string(c->spin_limit) + " polls) and every block left the kernel");
    return MSW_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

template <typename T>
DiagSeeds<T> diag_seeds(const msw_ctx* c, const void* p)
{
    const size_t C = c->nC, E = c->nE, V = c->nV, B = c->B;
    const T* q = static_cast<const T*>(p);
    DiagSeeds<T> s;
    const T** f[13] = {&s.v, &s.he, &s.hv, &s.circ, &s.vort, &s.div, &s.ke,
                       &s.pe, &s.pvv, &s.pc, &s.vc, &s.gpn, &s.gpt};
    const size_t n[13] = {E, E, V, V, V, C, C, E, V, C, C, E, E};
    for (int k = 0; k < 13; ++k) {
        *f[k] = q;
        q += n[k] * B;
    }
    return s;
}

// A named range around each step and operator entry point, recorded by
// rocprofv3 --marker-trace (a no-op when no profiler is attached); the range
// covers the host call, i.e. the enqueueing of the asynchronous work.
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};
#define MSW_TRACE() TraceRange msw_trace_range_(__func__)

extern "C" {

int msw_abi_version(void) { return MSW_ABI_VERSION; }

const char* msw_last_error(void) { return g_err.c_str(); }

int msw_create(const msw_mesh_desc* mesh, const msw_config* cfg, int prec, int batch,
               void* hip_stream, msw_ctx** out)
{
    MSW_TRACE();
    if (!out) return fail(MSW_EINVAL, "out is NULL");
    *out = nullptr;
    if (!mesh || !cfg) return fail(MSW_EINVAL, "mesh or cfg is NULL");
    if (prec != MSW_FP64 && prec != MSW_FP32) return fail(MSW_EINVAL, "prec must be MSW_FP64 or MSW_FP32");
    if (!is_pow2(batch) || batch > 256) return fail(MSW_EINVAL, "batch must be a power of two in [1,256]");
    std::unique_ptr<msw_ctx> c(new (std::nothrow) msw_ctx());
    if (!c) return fail(MSW_ENOMEM, "host allocation failed");
    c->prec = prec;
    c->B = batch;
    c->lb = 0;
    while ((1 << c->lb) < batch) ++c->lb;
    c->nC = mesh->nCells;
    c->nE = mesh->nEdges;
    c->nV = mesh->nVertices;
    // kernels index (entity, member) items with 32-bit arithmetic
    const long max_rows = std::max<long>({static_cast<long>(c->nC), static_cast<long>(c->nE),
                                          static_cast<long>(c->nV)});
    if ((max_rows << c->lb) >= (1L << 31) || 68L * c->nE >= (1L << 31))
        return fail(MSW_EINVAL, "mesh entities times batch (and 68 x edges) must stay below 2^31");
    c->stream = static_cast<hipStream_t>(hip_stream);
    c->cfg = *cfg;
    int rc = build_mesh(c.get(), mesh);
    if (rc) return rc;
    if ((rc = alloc_buffers(c.get()))) return rc;
    if ((rc = c->alloc(&c->usrc, c->nE))) return rc;
    HIP_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    for (auto& e : c->ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_TRY(hipMemset(c->usrc, 0, static_cast<size_t>(c->nE) * sizeof(double)));
    if ((rc = ensure_dissip(c.get()))) return rc;
    set_dcfg(c.get());
    HIP_TRY(hipDeviceSynchronize());
    *out = c.release();
    return MSW_OK;
}

int msw_destroy(msw_ctx* ctx)
{
    if (!ctx) return MSW_OK;
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return MSW_OK;
}

int msw_set_config(msw_ctx* ctx, const msw_config* cfg)
{
    if (!ctx || !cfg) return fail(MSW_EINVAL, "NULL argument");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->cfg = *cfg;
    int rc = ensure_dissip(ctx);
    if (rc) return rc;
    set_dcfg(ctx);
    ctx->clear_graphs();
    return MSW_OK;
}

int msw_set_forcing(msw_ctx* ctx, const double* u_src)
{
    if (!ctx) return fail(MSW_EINVAL, "NULL ctx");
    if (u_src)
        HIP_TRY(hipMemcpyAsync(ctx->usrc, u_src, static_cast<size_t>(ctx->nE) * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream));
    else
        HIP_TRY(hipMemsetAsync(ctx->usrc, 0, static_cast<size_t>(ctx->nE) * sizeof(double), ctx->stream));
    return MSW_OK;
}

int msw_set_mesh_scaling(msw_ctx* ctx, const double* del2, const double* del4)
{
    if (!ctx) return fail(MSW_EINVAL, "NULL ctx");
    if ((!del2) != (!del4)) return fail(MSW_EINVAL, "del2 and del4 go together");
    const size_t E = ctx->nE;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (del2) {
        if (!ctx->msd) {
            int rc = ctx->alloc(&ctx->msd, 2 * E);
            if (rc) return rc;
        }
        HIP_TRY(hipMemcpyAsync(ctx->msd, del2, E * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->msd + E, del4, E * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    ctx->msd_on = del2 != nullptr;
    set_dcfg(ctx);
    ctx->clear_graphs();
    return MSW_OK;
}

int msw_set_option(msw_ctx* ctx, int option, int value)
{
    if (!ctx) return fail(MSW_EINVAL, "NULL ctx");
    if (option == MSW_OPT_XCD_STRIPE && (value < 0 || value > 7))
        return fail(MSW_EINVAL, "MSW_OPT_XCD_STRIPE takes log2 of the stripe length, 0..6, or 7 "
                                "(one contiguous range per XCD)");
    if (option == MSW_OPT_BLOCK && value != 64 && value != 128 && value != 256 && value != 512)
        return fail(MSW_EINVAL, "MSW_OPT_BLOCK takes 64, 128, 256 or 512 threads");
    if (option == MSW_OPT_BARRIER_POLLS && value < 1) return fail(MSW_EINVAL, "MSW_OPT_BARRIER_POLLS must be positive");
    if (option != MSW_OPT_TL_FUSED_FLUX && option != MSW_OPT_XCD_STRIPE && option != MSW_OPT_BLOCK &&
        option != MSW_OPT_PERSISTENT && option != MSW_OPT_BARRIER_POLLS)
        return fail(MSW_EINVAL, "unknown option");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (option == MSW_OPT_TL_FUSED_FLUX) ctx->tl_fused = value != 0;
    else if (option == MSW_OPT_XCD_STRIPE) ctx->dm.stripe_log2 = value;
    else if (option == MSW_OPT_BLOCK) ctx->dm.block = value;
    else if (option == MSW_OPT_PERSISTENT) ctx->persistent = value != 0;
    else ctx->spin_limit = static_cast<unsigned>(value);
    set_dcfg(ctx);
    ctx->clear_graphs();
    return MSW_OK;
}

int msw_set_instances(msw_ctx* ctx, int n, const int32_t* cell_start, const int32_t* edge_start,
                      const int32_t* vertex_start)
{
    if (!ctx || n < 1 || !cell_start || !edge_start || !vertex_start) return fail(MSW_EINVAL, "bad argument");
    const int32_t* starts[3] = {cell_start, edge_start, vertex_start};
    const int totals[3] = {ctx->nC, ctx->nE, ctx->nV};
    for (int k = 0; k < 3; ++k) {
        if (starts[k][0] != 0) return fail(MSW_EINVAL, "the first instance must start at entity 0");
        for (int i = 1; i < n; ++i)
            if (starts[k][i] <= starts[k][i - 1] || starts[k][i] >= totals[k])
                return fail(MSW_EINVAL, "instance starts must increase and stay inside the mesh");
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->inst_c.assign(cell_start, cell_start + n);
    ctx->inst_e.assign(edge_start, edge_start + n);
    ctx->inst_v.assign(vertex_start, vertex_start + n);
    ctx->patches = nullptr;   // rebuilt at the next persistent call (the old table stays allocated)
    return MSW_OK;
}

int msw_set_tend_output(msw_ctx* ctx, void* tend_h, void* tend_u, void* tend_h_tl, void* tend_u_tl)
{
    if (!ctx) return fail(MSW_EINVAL, "NULL ctx");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->tout[0] = tend_h;
    ctx->tout[1] = tend_u;
    ctx->tout[2] = tend_h_tl;
    ctx->tout[3] = tend_u_tl;
    ctx->clear_graphs();   // the pointers are launch arguments of the captured steps
    return MSW_OK;
}

int msw_set_graphs(msw_ctx* ctx, int enable)
{
    if (!ctx) return fail(MSW_EINVAL, "NULL ctx");
    ctx->use_graphs = enable != 0;
    return MSW_OK;
}

// a tendency output is set: the steps run as stage launches, which write it
static bool tend_kept(const msw_ctx* ctx)
{
    return ctx->tout[0] || ctx->tout[1] || ctx->tout[2] || ctx->tout[3];
}

// FP64 member arrays of a 64-member-or-larger context move as 16-byte pairs
static bool members_aligned(const msw_ctx* ctx, const void* a, const void* b)
{
    if (ctx->prec != MSW_FP64 || ctx->B < 64) return true;
    return (reinterpret_cast<uintptr_t>(a) % 16 == 0) && (reinterpret_cast<uintptr_t>(b) % 16 == 0);
}

int msw_fwd_steps(msw_ctx* ctx, void* h, void* u, int nsteps)
{
    MSW_TRACE();
    if (!ctx || !h || !u || nsteps < 0) return fail(MSW_EINVAL, "bad argument");
    if (!members_aligned(ctx, h, u)) return fail(MSW_EINVAL, "h/u must be 16-byte aligned");
    int rc = check_cfg(ctx);
    if (rc) return rc;
    if (ctx->persistent && !tend_kept(ctx))
        return run_persistent(ctx, false, nullptr, nullptr, h, u, nsteps, nullptr, nullptr);
    if (ctx->prec == MSW_FP64)
        return run_steps(ctx, 1, h, u, nullptr, nullptr, nsteps,
                         [&] { enqueue_fwd_step<double>(ctx, static_cast<double*>(h), static_cast<double*>(u)); return MSW_OK; });
    return run_steps(ctx, 2, h, u, nullptr, nullptr, nsteps,
                     [&] { enqueue_fwd_step<float>(ctx, static_cast<float*>(h), static_cast<float*>(u)); return MSW_OK; });
}

int msw_base_steps(msw_ctx* ctx, double* h, double* u, int nsteps, double* traj_h, double* traj_u)
{
    MSW_TRACE();
    if (!ctx || !h || !u || nsteps < 0 || (!traj_h) != (!traj_u)) return fail(MSW_EINVAL, "bad argument");
    int rc = check_cfg(ctx);
    if (rc) return rc;
    const size_t C = ctx->nC, E = ctx->nE;
    std::function<int(int)> pre = [&](int i) { return store_state(ctx, traj_h + i * C, traj_u + i * E, h, u); };
    return run_steps(ctx, 3, h, u, nullptr, nullptr, nsteps, [&] { enqueue_base_step(ctx, h, u); return MSW_OK; },
                     traj_h ? &pre : nullptr);
}

int msw_tlm_steps(msw_ctx* ctx, double* h, double* u, void* h_tl, void* u_tl, int nsteps,
                  double* traj_h, double* traj_u)
{
    MSW_TRACE();
    if (!ctx || !h || !u || !h_tl || !u_tl || nsteps < 0 || (!traj_h) != (!traj_u))
        return fail(MSW_EINVAL, "bad argument");
    if (!members_aligned(ctx, h_tl, u_tl)) return fail(MSW_EINVAL, "h_tl/u_tl must be 16-byte aligned");
    int rc = check_cfg(ctx);
    if (rc) return rc;
    if (ctx->persistent && !tend_kept(ctx))
        return run_persistent(ctx, true, h, u, h_tl, u_tl, nsteps, traj_h, traj_u);
    const size_t C = ctx->nC, E = ctx->nE;
    if (!tend_kept(ctx)) {
        if (nsteps == 0) return MSW_OK;
        // step i runs with base set i % 2, formed while step i-1 ran; the
        // first set is formed here.  (h, u) run a step ahead of the members:
        // before step i they hold the start state of step i+1.
        if (traj_h && (rc = store_state(ctx, traj_h, traj_u, h, u))) return rc;
        if ((rc = forked(ctx, [&] { return side_tlm_base(ctx, ctx->st, ctx->qs, ctx->bx, h, u); }))) return rc;
        std::function<int(int)> pre = [&](int i) {
            return i + 1 < nsteps ? store_state(ctx, traj_h + (i + 1) * C, traj_u + (i + 1) * E, h, u)
                                  : static_cast<int>(MSW_OK);
        };
        std::function<int(int)> variant = [&](int i) { return (i % 2) * 2 + (i + 1 < nsteps ? 1 : 0); };
        if (ctx->prec == MSW_FP64)
            return run_steps_var(ctx, 12, h, u, h_tl, u_tl, nsteps, [&](int v) {
                return enqueue_tlm_step_piped<double>(ctx, v >> 1, (v & 1) != 0, h, u, static_cast<double*>(h_tl),
                                                      static_cast<double*>(u_tl));
            }, traj_h ? &pre : nullptr, &variant);
        return run_steps_var(ctx, 13, h, u, h_tl, u_tl, nsteps, [&](int v) {
            return enqueue_tlm_step_piped<float>(ctx, v >> 1, (v & 1) != 0, h, u, static_cast<float*>(h_tl),
                                                 static_cast<float*>(u_tl));
        }, traj_h ? &pre : nullptr, &variant);
    }
    // with tendency outputs, each step forms its own base set first (the
    // outputs then hold that step's tendencies)
    std::function<int(int)> pre = [&](int i) { return store_state(ctx, traj_h + i * C, traj_u + i * E, h, u); };
    const std::function<int(int)>* pp = traj_h ? &pre : nullptr;
    if (ctx->prec == MSW_FP64)
        return run_steps(ctx, 4, h, u, h_tl, u_tl, nsteps, [&] {
            return enqueue_tlm_step<double>(ctx, h, u, static_cast<double*>(h_tl), static_cast<double*>(u_tl));
        }, pp);
    return run_steps(ctx, 5, h, u, h_tl, u_tl, nsteps, [&] {
        return enqueue_tlm_step<float>(ctx, h, u, static_cast<float*>(h_tl), static_cast<float*>(u_tl));
    }, pp);
}

static int adj_common(msw_ctx* ctx, void* h_ad, void* u_ad, int nsteps, bool want_end,
                      const std::function<int(int)>& pre)
{
    if (!members_aligned(ctx, h_ad, u_ad)) return fail(MSW_EINVAL, "h_ad/u_ad must be 16-byte aligned");
    if (ctx->prec == MSW_FP64)
        return run_steps(ctx, want_end ? 6 : 7, h_ad, u_ad, nullptr, nullptr, nsteps, [&] {
            return enqueue_adj_step<double>(ctx, static_cast<double*>(h_ad), static_cast<double*>(u_ad), want_end);
        }, &pre);
    return run_steps(ctx, want_end ? 8 : 9, h_ad, u_ad, nullptr, nullptr, nsteps, [&] {
        return enqueue_adj_step<float>(ctx, static_cast<float*>(h_ad), static_cast<float*>(u_ad), want_end);
    }, &pre);
}

int msw_adj_step(msw_ctx* ctx, const double* h, const double* u, void* h_ad, void* u_ad,
                 double* h_end, double* u_end)
{
    MSW_TRACE();
    if (!ctx || !h || !u || !h_ad || !u_ad) return fail(MSW_EINVAL, "bad argument");
    int rc = check_cfg(ctx);
    if (rc) return rc;
    const size_t C = ctx->nC, E = ctx->nE;
    std::function<int(int)> pre = [&](int) { return store_state(ctx, ctx->bx, ctx->bx + C, h, u); };
    const bool want_end = h_end || u_end;
    rc = adj_common(ctx, h_ad, u_ad, 1, want_end, pre);
    if (rc) return rc;
    if (h_end) HIP_TRY(hipMemcpyAsync(h_end, ctx->bend, C * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if (u_end) HIP_TRY(hipMemcpyAsync(u_end, ctx->bend + C, E * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return MSW_OK;
}

int msw_adj_steps(msw_ctx* ctx, const double* traj_h, const double* traj_u, void* h_ad, void* u_ad,
                  int nsteps)
{
    MSW_TRACE();
    if (!ctx || !traj_h || !traj_u || !h_ad || !u_ad || nsteps < 0) return fail(MSW_EINVAL, "bad argument");
    int rc = check_cfg(ctx);
    if (rc) return rc;
    if (!members_aligned(ctx, h_ad, u_ad)) return fail(MSW_EINVAL, "h_ad/u_ad must be 16-byte aligned");
    if (nsteps == 0) return MSW_OK;
    const size_t C = ctx->nC, E = ctx->nE;
    // step i (reverse order) runs the adjoint of step k = nsteps-1-i with base
    // set i % 2, formed while step i-1 ran; before it, the start state of step
    // k-1 goes into the other set's buffer for the side stream to work on
    auto copy_start = [&](size_t k, double* dst) {
        return store_state(ctx, dst, dst + C, traj_h + k * C, traj_u + k * E);
    };
    if ((rc = copy_start(static_cast<size_t>(nsteps - 1), ctx->bx))) return rc;
    if ((rc = forked(ctx, [&] { return side_adj_base(ctx, ctx->st, ctx->qs, ctx->bx, false, false); }))) return rc;
    std::function<int(int)> pre = [&](int i) {
        return i + 1 < nsteps ? copy_start(static_cast<size_t>(nsteps - 2 - i), (i % 2) ? ctx->bx : ctx->bxB)
                              : static_cast<int>(MSW_OK);
    };
    std::function<int(int)> variant = [&](int i) { return (i % 2) * 2 + (i + 1 < nsteps ? 1 : 0); };
    if (ctx->prec == MSW_FP64)
        return run_steps_var(ctx, 10, h_ad, u_ad, nullptr, nullptr, nsteps, [&](int v) {
            return enqueue_adj_step_piped<double>(ctx, v >> 1, (v & 1) != 0, static_cast<double*>(h_ad),
                                                  static_cast<double*>(u_ad));
        }, &pre, &variant);
    return run_steps_var(ctx, 11, h_ad, u_ad, nullptr, nullptr, nsteps, [&](int v) {
        return enqueue_adj_step_piped<float>(ctx, v >> 1, (v & 1) != 0, static_cast<float*>(h_ad),
                                             static_cast<float*>(u_ad));
    }, &pre, &variant);
}

static const char* kDiagNames =
    "v,h_edge,h_vertex,circulation,vorticity,divergence,ke,pv_edge,pv_vertex,pv_cell,"
    "vorticity_cell,gradPVn,gradPVt";
static const char* kDiagKinds = "e,e,v,v,v,c,c,e,v,c,c,e,e";

const char* msw_diag_names(void) { return kDiagNames; }
const char* msw_diag_kinds(void) { return kDiagKinds; }

int msw_diagnostics(msw_ctx* ctx, const double* h, const double* u, double* out)
{
    MSW_TRACE();
    if (!ctx || !h || !u || !out) return fail(MSW_EINVAL, "bad argument");
    const size_t C = ctx->nC, E = ctx->nE, V = ctx->nV;
    double* o = out;
    double* v = o; o += E;
    double* he = o; o += E;
    double* hv = o; o += V;
    double* circ = o; o += V;
    double* vort = o; o += V;
    double* div = o; o += C;
    double* ke = o; o += C;
    double* pe = o; o += E;
    double* pvv = o; o += V;
    double* pc = o; o += C;
    double* vc = o; o += C;
    double* gpn = o; o += E;
    double* gpt = o; o += E;
    fwd_diag<double>(ctx->dm, ctx->dcfg, 0, h, u, pvv, hv, vort, ke, pc, he, pe, v, gpt, gpn, nullptr, ctx->stream);
    diag_extra(ctx->dm, u, vort, circ, div, vc, ctx->stream);
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

int msw_diagnostics_adj(msw_ctx* ctx, const double* h, const double* u, const void* seeds, void* h_ad,
                        void* u_ad)
{
    MSW_TRACE();
    if (!ctx || !h || !u || !seeds || !h_ad || !u_ad) return fail(MSW_EINVAL, "bad argument");
    int rc = check_cfg(ctx);
    if (rc) return rc;
    const size_t C = ctx->nC, E = ctx->nE, V = ctx->nV;
    if (!ctx->dgb && (rc = ctx->alloc(&ctx->dgb, 5 * E + 4 * V + 4 * C))) return rc;
    // the 13 diagnostics of the state, packed as msw_diagnostics writes them
    if ((rc = msw_diagnostics(ctx, h, u, ctx->dgb))) return rc;
    const double* d = ctx->dgb;
    DiagBase bd;
    bd.u = u;
    bd.v = d;
    bd.hv = d + 2 * E;
    bd.vort = d + 2 * E + 2 * V;
    bd.gpn = d + 3 * E + 4 * V + 4 * C;
    bd.gpt = bd.gpn + E;
    const DMesh& m = ctx->dm;
    if (ctx->prec == MSW_FP64)
        diag_adj<double>(m, ctx->dcfg, ctx->lb, bd, diag_seeds<double>(ctx, seeds), static_cast<double*>(ctx->me1),
                         static_cast<double*>(ctx->me2), static_cast<double*>(ctx->me3),
                         static_cast<double*>(ctx->mc1), static_cast<double*>(ctx->mv1),
                         static_cast<double*>(ctx->mv2), static_cast<double*>(h_ad), static_cast<double*>(u_ad),
                         ctx->stream);
    else
        diag_adj<float>(m, ctx->dcfg, ctx->lb, bd, diag_seeds<float>(ctx, seeds), static_cast<float*>(ctx->me1),
                        static_cast<float*>(ctx->me2), static_cast<float*>(ctx->me3),
                        static_cast<float*>(ctx->mc1), static_cast<float*>(ctx->mv1),
                        static_cast<float*>(ctx->mv2), static_cast<float*>(h_ad), static_cast<float*>(u_ad),
                        ctx->stream);
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

int msw_invariants(msw_ctx* ctx, const double* h, const double* u, double* out)
{
    MSW_TRACE();
    if (!ctx || !h || !u || !out) return fail(MSW_EINVAL, "bad argument");
    const size_t C = ctx->nC, E = ctx->nE, V = ctx->nV;
    const DMesh& m = ctx->dm;
    hipStream_t s = ctx->stream;
    if (!ctx->inv) {
        int rc = ctx->alloc(&ctx->inv, 4 * C + 3 * V + 2 * E);
        if (rc) return rc;
        // weights 0.5*g*A and g*A of the potential-energy terms
        adj_seed<double>(static_cast<int>(C), ctx->inv + 2 * C + 3 * V + 2 * E, m.areaC, 0.5 * kGravity, s);
        adj_seed<double>(static_cast<int>(C), ctx->inv + 3 * C + 3 * V + 2 * E, m.areaC, kGravity, s);
    }
    double* ke = ctx->inv;
    double* pc = ke + C;
    double* pvv = pc + C;
    double* hv = pvv + V;
    double* vort = hv + V;
    double* he = vort + V;
    double* pe = he + E;
    const double* wE = pe + E;
    const double* wP = wE + C;
    fwd_diag<double>(m, ctx->dcfg, 0, h, u, pvv, hv, vort, ke, pc, he, pe, nullptr, nullptr, nullptr,
                     nullptr, s);
    const int nb = ctx->dot_nblk;
    dot_batched<double>(static_cast<long>(C), 0, h, nullptr, m.areaC, ctx->dpartial, nb, out, 0, s);
    dot_batched<double>(static_cast<long>(C), 0, h, h, wE, ctx->dpartial, nb, out + 1, 0, s);
    dot_batched<double>(static_cast<long>(C), 0, h, m.hs, wP, ctx->dpartial, nb, out + 1, 1, s);
    dot_batched<double>(static_cast<long>(C), 0, h, ke, m.areaC, ctx->dpartial, nb, out + 1, 1, s);
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

int msw_reconstruct(msw_ctx* ctx, const void* u, void* out5)
{
    MSW_TRACE();
    if (!ctx || !u || !out5) return fail(MSW_EINVAL, "bad argument");
    if (ctx->prec == MSW_FP64)
        reconstruct<double>(ctx->nC, ctx->lb, ctx->dm.nEoC, ctx->dm.eoc, ctx->coeffs, ctx->trig,
                            static_cast<const double*>(u), static_cast<double*>(out5), ctx->stream);
    else
        reconstruct<float>(ctx->nC, ctx->lb, ctx->dm.nEoC, ctx->dm.eoc, ctx->coeffs, ctx->trig,
                           static_cast<const float*>(u), static_cast<float*>(out5), ctx->stream);
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

int msw_reconstruct_adj(msw_ctx* ctx, const void* seeds5, void* u_ad)
{
    MSW_TRACE();
    if (!ctx || !seeds5 || !u_ad) return fail(MSW_EINVAL, "bad argument");
    // scratch: the member C-fields mc1..mc3 are contiguous in the member block
    if (ctx->prec == MSW_FP64)
        reconstruct_adj<double>(ctx->dm, ctx->lb, ctx->coeffs, ctx->trig,
                                static_cast<const double*>(seeds5), static_cast<double*>(ctx->mc1),
                                static_cast<double*>(u_ad), ctx->stream);
    else
        reconstruct_adj<float>(ctx->dm, ctx->lb, ctx->coeffs, ctx->trig,
                               static_cast<const float*>(seeds5), static_cast<float*>(ctx->mc1),
                               static_cast<float*>(u_ad), ctx->stream);
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

int msw_dot(msw_ctx* ctx, const void* xh, const void* xu, const void* yh, const void* yu,
            const double* wh, const double* wu, double* out)
{
    MSW_TRACE();
    if (!ctx || !out || (!xh && !xu)) return fail(MSW_EINVAL, "bad argument");
    bool acc = false;
    auto run = [&](long n, const void* x, const void* y, const double* w) {
        if (ctx->prec == MSW_FP64)
            dot_batched<double>(n, ctx->lb, static_cast<const double*>(x), static_cast<const double*>(y), w,
                                ctx->dpartial, ctx->dot_nblk, out, acc ? 1 : 0, ctx->stream);
        else
            dot_batched<float>(n, ctx->lb, static_cast<const float*>(x), static_cast<const float*>(y), w,
                               ctx->dpartial, ctx->dot_nblk, out, acc ? 1 : 0, ctx->stream);
        acc = true;
    };
    if (xh) run(ctx->nC, xh, yh, wh);
    if (xu) run(ctx->nE, xu, yu, wu);
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

int msw_dot_noalloc(int n_cells, int n_edges, int batch, int prec, const void* xh, const void* xu,
                    const void* yh, const void* yu, const double* wh, const double* wu, double* out,
                    void* hip_stream)
{
    MSW_TRACE();
    if (!out || (!xh && !xu) || n_cells < 0 || n_edges < 0) return fail(MSW_EINVAL, "bad argument");
    if (prec != MSW_FP64 && prec != MSW_FP32) return fail(MSW_EINVAL, "prec must be MSW_FP64 or MSW_FP32");
    if (!is_pow2(batch) || batch > 256)
        return fail(MSW_EINVAL, "batch must be a power of two in [1,256] (the scratch holds 2048 x 256 "
                                "FP64 partials)");
    int lb = 0;
    while ((1 << lb) < batch) ++lb;
    // the block split msw_dot uses for a mesh of these sizes
    const int nblk = dot_blocks(std::max<long>(n_cells, n_edges));
    if (static_cast<long>(nblk) * batch > kDotScratchDoubles) return fail(MSW_EINVAL, "too many partials");
    double* partial = nullptr;
    const int e = dot_scratch(&partial);
    if (e) return fail(MSW_EHIP, std::string("scratch address: ") + hipGetErrorString(static_cast<hipError_t>(e)));
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    bool acc = false;
    auto run = [&](long n, const void* x, const void* y, const double* w) {
        if (prec == MSW_FP64)
            dot_batched<double>(n, lb, static_cast<const double*>(x), static_cast<const double*>(y), w, partial,
                                nblk, out, acc ? 1 : 0, s);
        else
            dot_batched<float>(n, lb, static_cast<const float*>(x), static_cast<const float*>(y), w, partial,
                               nblk, out, acc ? 1 : 0, s);
        acc = true;
    };
    if (xh) run(n_cells, xh, yh, wh);
    if (xu) run(n_edges, xu, yu, wu);
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

int msw_nonfinite(msw_ctx* ctx, const void* h, const void* u, int* flags)
{
    MSW_TRACE();
    if (!ctx || !flags) return fail(MSW_EINVAL, "bad argument");
    if (ctx->prec == MSW_FP64) {
        if (h) nonfinite_flags<double>(ctx->nC, ctx->lb, static_cast<const double*>(h), flags, ctx->stream);
        if (u) nonfinite_flags<double>(ctx->nE, ctx->lb, static_cast<const double*>(u), flags, ctx->stream);
    } else {
        if (h) nonfinite_flags<float>(ctx->nC, ctx->lb, static_cast<const float*>(h), flags, ctx->stream);
        if (u) nonfinite_flags<float>(ctx->nE, ctx->lb, static_cast<const float*>(u), flags, ctx->stream);
    }
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

int msw_permute(msw_ctx* ctx, int kind, int to_internal, int prec, int batch, const void* src, void* dst)
{
    MSW_TRACE();
    if (!ctx || !src || !dst || src == dst) return fail(MSW_EINVAL, "bad argument");
    if (!is_pow2(batch) || batch > 1024) return fail(MSW_EINVAL, "batch must be a power of two <= 1024");
    int lb = 0;
    while ((1 << lb) < batch) ++lb;
    const int* perm = nullptr;
    long n = 0;
    if (kind == 'c') { perm = ctx->permC; n = ctx->nC; }
    else if (kind == 'e') { perm = ctx->permE; n = ctx->nE; }
    else if (kind == 'v') { perm = ctx->permV; n = ctx->nV; }
    else return fail(MSW_EINVAL, "kind must be 'c', 'e' or 'v'");
    if (prec == MSW_FP64) {
        if (to_internal) gather_rows<double>(n, lb, perm, static_cast<const double*>(src), static_cast<double*>(dst), ctx->stream);
        else scatter_rows<double>(n, lb, perm, static_cast<const double*>(src), static_cast<double*>(dst), ctx->stream);
    } else if (prec == MSW_FP32) {
        if (to_internal) gather_rows<float>(n, lb, perm, static_cast<const float*>(src), static_cast<float*>(dst), ctx->stream);
        else scatter_rows<float>(n, lb, perm, static_cast<const float*>(src), static_cast<float*>(dst), ctx->stream);
    } else {
        return fail(MSW_EINVAL, "bad precision");
    }
    HIP_TRY(hipGetLastError());
    return MSW_OK;
}

int64_t msw_device_bytes(const msw_ctx* ctx) { return ctx ? ctx->bytes : 0; }

}  // extern "C"
