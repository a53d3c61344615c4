// CPU stand-in for the HIP runtime and for the kernel launchers of
// libmpas_sw_amd, linked with the library's real host runtime
// (msw_runtime.cpp) into a host-only build that runs under AddressSanitizer.
// It checks the runtime's host logic, not the kernels' arithmetic:
//
//  * device memory is host memory in a registry; every buffer a launcher's
//    kernels would read or write (extents as in msw_kernels.hip) and every
//    copy must lie inside a live allocation;
//  * streams, events and graph capture follow HIP's rules (capture on the
//    null stream, unjoined side streams and waits on events outside the
//    capture are errors), and captured graphs replay the recorded work;
//  * a vector clock per stream (and per stream of each graph replay) orders
//    the work; two accesses to overlapping bytes from different streams, at
//    least one a write, with neither ordered before the other by events or
//    host synchronisation, are reported as a race;
//  * a copy, a launch or an allocation can be made to fail on request, to
//    check that errors reach the caller.
// Nothing here runs on a GPU.
#include <hip/hip_runtime.h>

#include "../../mpas-sw-tl-ad_amd/csrc/msw_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

namespace mock {

using VC = std::vector<unsigned long>;

unsigned long vat(const VC& v, int i) { return i < static_cast<int>(v.size()) ? v[i] : 0; }

void vjoin(VC& a, const VC& b)
{
    if (a.size() < b.size()) a.resize(b.size(), 0);
    for (size_t i = 0; i < b.size(); ++i) a[i] = std::max(a[i], b[i]);
}

// an execution context: a stream, a stream of a graph replay, or the host
struct Lane {
    int id = 0;
    VC vc;
    void tick()
    {
        if (static_cast<int>(vc.size()) <= id) vc.resize(id + 1, 0);
        ++vc[id];
    }
};

struct Access {
    uintptr_t lo, hi;
    int lane;
    unsigned long clk;
    bool write;
    std::string what;
};

struct Alloc {
    size_t size;
    std::vector<Access> acc;
};

std::map<uintptr_t, Alloc> allocs;
std::vector<Lane*> live;   // lanes that may still run work
int next_id = 1;
Lane host;
Lane* cur = nullptr;   // lane of the operation being executed
std::string cur_op;

long nviol = 0;
std::vector<std::string> msgs;
hipError_t sticky = hipSuccess;
long fail_at[3] = {-1, -1, -1};   // copy, launch, malloc: index of the call that fails
long ncalls[3] = {0, 0, 0};
long graphs_live = 0;
long graphs_made = 0;

void violation(const std::string& m)
{
    ++nviol;
    if (msgs.size() < 200) msgs.push_back(m);
}

bool fail_now(int kind)
{
    const bool f = fail_at[kind] >= 0 && ncalls[kind] == fail_at[kind];
    ++ncalls[kind];
    if (f) fail_at[kind] = -1;
    return f;
}

void prune(Alloc& a)
{
    if (a.acc.size() < 256) return;
    auto seen_by_all = [](const Access& x) {
        for (const Lane* l : live)
            if (vat(l->vc, x.lane) < x.clk) return false;
        return true;
    };
    a.acc.erase(std::remove_if(a.acc.begin(), a.acc.end(), seen_by_all), a.acc.end());
}

std::map<uintptr_t, Alloc>::iterator find_alloc(uintptr_t p)
{
    auto it = allocs.upper_bound(p);
    if (it == allocs.begin()) return allocs.end();
    --it;
    if (p >= it->first + it->second.size) return allocs.end();
    return it;
}

bool valid(const void* p, size_t bytes)
{
    if (!bytes) return true;
    auto it = find_alloc(reinterpret_cast<uintptr_t>(p));
    return it != allocs.end() && reinterpret_cast<uintptr_t>(p) + bytes <= it->first + it->second.size;
}

// an access of `bytes` bytes at p by the current operation
void touch(const void* p, size_t bytes, bool write, const std::string& what)
{
    if (!bytes) return;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    auto it = find_alloc(a);
    if (it == allocs.end()) {
        violation(cur_op + ": " + what + " is not live device memory");
        return;
    }
    const uintptr_t end = it->first + it->second.size;
    if (a + bytes > end) {
        violation(cur_op + ": " + what + " overruns its allocation by " + std::to_string(a + bytes - end) +
                  " bytes");
        return;
    }
    Lane& l = *cur;
    for (const Access& x : it->second.acc)
        if (x.lane != l.id && (x.write || write) && x.lo < a + bytes && a < x.hi && vat(l.vc, x.lane) < x.clk)
            violation(cur_op + ": " + what + (write ? " (write)" : " (read)") + " races with " + x.what +
                      (x.write ? " (write)" : " (read)") + " on another stream");
    it->second.acc.push_back({a, a + bytes, l.id, vat(l.vc, l.id), write, cur_op + ": " + what});
    prune(it->second);
}

// a buffer a kernel of a launcher reads or writes
struct Buf {
    const void* p;
    size_t bytes;
    bool write;
    std::string what;
    bool optional;   // a null pointer means "not used"
    size_t align;
};

}  // namespace mock

struct ihipStream_t {
    mock::Lane lane;
    int cap = 0;             // id of the capture this stream takes part in (0: none)
    long ops_in_cap = 0;
    long joined_at = -1;
    bool destroyed = false;
};

struct ihipEvent_t {
    mock::VC vc;
    bool recorded = false;
    int cap = 0;             // capture the record happened in
    ihipStream_t* cap_src = nullptr;
    long cap_op = 0;
};

namespace mock {

struct Capture {
    ihipStream_t* origin = nullptr;
    std::set<ihipStream_t*> members;
    bool ended = false;
};

std::map<int, Capture> captures;
int next_cap = 1;

struct Rec {
    int kind;
    ihipStream_t* s;
    ihipEvent_t* e;
    std::function<void()> fn;
    std::string name;
};

}  // namespace mock

struct ihipGraph_t {
    std::vector<mock::Rec> ops;
};
struct ihipGraphExec_t {
    std::vector<mock::Rec> ops;
};

namespace mock {

ihipStream_t null_stream;
bool null_init = false;

ihipStream_t* stream_of(hipStream_t s)
{
    if (s) return s;
    if (!null_init) {
        null_init = true;
        null_stream.lane.id = next_id++;
        null_stream.lane.vc = host.vc;
        live.push_back(&null_stream.lane);
    }
    return &null_stream;
}

std::map<int, std::vector<Rec>> cap_ops;

void run_on(Lane& l, const std::string& name, const std::function<void()>& fn)
{
    l.tick();
    Lane* save = cur;
    const std::string save_op = cur_op;
    cur = &l;
    cur_op = name;
    fn();
    cur = save;
    cur_op = save_op;
}

// enqueue work on a stream: recorded into its capture, or run now
void enqueue(ihipStream_t* s, const std::string& name, std::function<void()> fn)
{
    if (s->cap) {
        cap_ops[s->cap].push_back({0, s, nullptr, std::move(fn), name});
        ++s->ops_in_cap;
        return;
    }
    vjoin(s->lane.vc, host.vc);   // work enqueued after what the host has waited for
    run_on(s->lane, name, fn);
}

void launch(hipStream_t st, const std::string& name, std::vector<Buf> bufs)
{
    if (fail_now(1)) sticky = hipErrorLaunchFailure;
    for (const Buf& b : bufs) {
        if (!b.p) {
            if (!b.optional && b.bytes) violation(name + ": required buffer " + b.what + " is NULL");
            continue;
        }
        if (b.align && reinterpret_cast<uintptr_t>(b.p) % b.align)
            violation(name + ": " + b.what + " is not " + std::to_string(b.align) + "-byte aligned");
    }
    enqueue(stream_of(st), name, [bufs]() {
        for (const Buf& b : bufs)
            if (b.p) touch(b.p, b.bytes, b.write, b.what);
    });
}

}  // namespace mock

using namespace mock;

// ---------------------------------------------------------------------------
// HIP runtime API

hipError_t hipMalloc(void** p, size_t bytes)
{
    if (fail_now(2)) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    void* q = nullptr;
    if (posix_memalign(&q, 256, std::max<size_t>(bytes, 1))) return hipErrorOutOfMemory;
    std::memset(q, 0, bytes);
    allocs[reinterpret_cast<uintptr_t>(q)] = Alloc{std::max<size_t>(bytes, 1), {}};
    *p = q;
    return hipSuccess;
}

hipError_t hipFree(void* p)
{
    if (!p) return hipSuccess;
    auto it = allocs.find(reinterpret_cast<uintptr_t>(p));
    if (it == allocs.end()) {
        violation("hipFree of a pointer that is not a live allocation");
        return hipErrorInvalidValue;
    }
    for (Lane* l : live) vjoin(host.vc, l->vc);   // hipFree waits for the device
    allocs.erase(it);
    std::free(p);
    return hipSuccess;
}

static hipError_t copy_async(void* dst, const void* src, size_t n, hipMemcpyKind kind, hipStream_t st,
                             const char* name)
{
    if (fail_now(0)) return hipErrorInvalidValue;
    const bool dsrc = kind == hipMemcpyDeviceToDevice || kind == hipMemcpyDeviceToHost;
    const bool ddst = kind == hipMemcpyDeviceToDevice || kind == hipMemcpyHostToDevice;
    std::vector<unsigned char> staged;
    if (!dsrc) staged.assign(static_cast<const unsigned char*>(src), static_cast<const unsigned char*>(src) + n);
    enqueue(stream_of(st), name, [=]() {
        if (dsrc) touch(src, n, false, "copy source");
        if (ddst) touch(dst, n, true, "copy destination");
        if ((!dsrc || valid(src, n)) && (!ddst || valid(dst, n)))
            std::memmove(dst, dsrc ? src : staged.data(), n);
    });
    return hipSuccess;
}

hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind kind, hipStream_t st)
{
    return copy_async(dst, src, n, kind, st, "hipMemcpyAsync");
}

hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind kind)
{
    ihipStream_t* s = stream_of(nullptr);
    if (s->cap) return hipErrorStreamCaptureUnsupported;
    hipError_t e = copy_async(dst, src, n, kind, nullptr, "hipMemcpy");
    vjoin(host.vc, s->lane.vc);
    return e;
}

hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t st)
{
    if (fail_now(0)) return hipErrorInvalidValue;
    enqueue(stream_of(st), "hipMemsetAsync", [=]() {
        touch(p, n, true, "memset destination");
        if (valid(p, n)) std::memset(p, v, n);
    });
    return hipSuccess;
}

hipError_t hipMemset(void* p, int v, size_t n)
{
    hipError_t e = hipMemsetAsync(p, v, n, nullptr);
    vjoin(host.vc, stream_of(nullptr)->lane.vc);
    return e;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned)
{
    auto* q = new ihipStream_t();
    q->lane.id = next_id++;
    q->lane.vc = host.vc;
    live.push_back(&q->lane);
    *s = q;
    return hipSuccess;
}

hipError_t hipStreamDestroy(hipStream_t s)
{
    if (!s || s->destroyed) return hipErrorInvalidResourceHandle;
    if (s->cap) violation("hipStreamDestroy of a stream in a capture");
    s->destroyed = true;
    live.erase(std::remove(live.begin(), live.end(), &s->lane), live.end());
    return hipSuccess;   // kept allocated: graphs may still name it
}

hipError_t hipStreamSynchronize(hipStream_t st)
{
    ihipStream_t* s = stream_of(st);
    if (s->cap) {
        violation("hipStreamSynchronize on a capturing stream");
        return hipErrorStreamCaptureUnsupported;
    }
    vjoin(host.vc, s->lane.vc);
    return hipSuccess;
}

hipError_t hipDeviceSynchronize(void)
{
    for (Lane* l : live) vjoin(host.vc, l->vc);
    return hipSuccess;
}

hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned)
{
    *e = new ihipEvent_t();
    return hipSuccess;
}

hipError_t hipEventDestroy(hipEvent_t e)
{
    if (!e) return hipErrorInvalidResourceHandle;
    return hipSuccess;   // kept allocated: graphs may still name it
}

hipError_t hipEventRecord(hipEvent_t e, hipStream_t st)
{
    ihipStream_t* s = stream_of(st);
    if (s->cap) {
        cap_ops[s->cap].push_back({1, s, e, nullptr, "record"});
        ++s->ops_in_cap;
        e->cap = s->cap;
        e->cap_src = s;
        e->cap_op = s->ops_in_cap;
        return hipSuccess;
    }
    vjoin(s->lane.vc, host.vc);
    s->lane.tick();
    e->vc = s->lane.vc;
    e->recorded = true;
    e->cap = 0;
    return hipSuccess;
}

hipError_t hipStreamWaitEvent(hipStream_t st, hipEvent_t e, unsigned)
{
    ihipStream_t* s = stream_of(st);
    if (e->cap) {
        Capture& c = captures[e->cap];
        if (c.ended) {
            violation("wait on an event recorded in a finished capture");
            return hipErrorStreamCaptureIsolation;
        }
        if (!s->cap) {   // the stream joins the capture
            s->cap = e->cap;
            s->ops_in_cap = 0;
            s->joined_at = -1;
            c.members.insert(s);
        } else if (s->cap != e->cap) {
            violation("wait across two captures");
            return hipErrorStreamCaptureIsolation;
        }
        cap_ops[s->cap].push_back({2, s, e, nullptr, "wait"});
        ++s->ops_in_cap;
        if (s == c.origin && e->cap_src != s) e->cap_src->joined_at = e->cap_op;
        return hipSuccess;
    }
    if (s->cap) {
        violation("a capturing stream waits on an event recorded outside the capture");
        return hipErrorStreamCaptureIsolation;
    }
    if (e->recorded) vjoin(s->lane.vc, e->vc);
    return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t st, hipStreamCaptureMode)
{
    if (!st) {
        violation("capture on the null stream");
        return hipErrorStreamCaptureUnsupported;
    }
    if (st->cap) return hipErrorStreamCaptureUnsupported;
    const int id = next_cap++;
    captures[id].origin = st;
    captures[id].members.insert(st);
    st->cap = id;
    st->ops_in_cap = 0;
    st->joined_at = -1;
    return hipSuccess;
}

hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t* g)
{
    if (!st || !st->cap || captures[st->cap].origin != st) return hipErrorInvalidValue;
    const int id = st->cap;
    Capture& c = captures[id];
    hipError_t rc = hipSuccess;
    for (ihipStream_t* m : c.members) {
        if (m != st && m->joined_at != m->ops_in_cap) {
            violation("a stream that joined the capture was not joined back before its end");
            rc = hipErrorStreamCaptureUnjoined;
        }
        m->cap = 0;
    }
    c.ended = true;
    if (rc != hipSuccess) {
        *g = nullptr;
        cap_ops.erase(id);
        return rc;
    }
    *g = new ihipGraph_t{std::move(cap_ops[id])};
    cap_ops.erase(id);
    return hipSuccess;
}

hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t)
{
    if (!g) return hipErrorInvalidValue;
    *x = new ihipGraphExec_t{g->ops};
    ++graphs_live;
    ++graphs_made;
    return hipSuccess;
}

hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}

hipError_t hipGraphExecDestroy(hipGraphExec_t x)
{
    if (!x) return hipErrorInvalidValue;
    delete x;
    --graphs_live;
    return hipSuccess;
}

hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t st)
{
    if (fail_now(1)) return hipErrorLaunchFailure;
    ihipStream_t* s = stream_of(st);
    if (s->cap) {
        violation("graph launch inside a capture");
        return hipErrorStreamCaptureUnsupported;
    }
    vjoin(s->lane.vc, host.vc);
    s->lane.tick();
    // each captured stream becomes a lane of this replay, forked from s
    std::map<ihipStream_t*, std::unique_ptr<Lane>> lanes;
    std::map<ihipEvent_t*, VC> evs;
    auto lane_of = [&](ihipStream_t* q) -> Lane& {
        auto& p = lanes[q];
        if (!p) {
            p.reset(new Lane());
            p->id = next_id++;
            p->vc = s->lane.vc;
            live.push_back(p.get());
        }
        return *p;
    };
    for (const Rec& r : x->ops) {
        Lane& l = lane_of(r.s);
        if (r.kind == 0) {
            run_on(l, "graph: " + r.name, r.fn);
        } else if (r.kind == 1) {
            l.tick();
            evs[r.e] = l.vc;
        } else {
            auto it = evs.find(r.e);
            if (it == evs.end()) violation("graph waits on an event not recorded earlier in it");
            else vjoin(l.vc, it->second);
        }
    }
    for (auto& kv : lanes) {
        vjoin(s->lane.vc, kv.second->vc);
        live.erase(std::remove(live.begin(), live.end(), kv.second.get()), live.end());
    }
    s->lane.tick();
    return hipSuccess;
}

hipError_t hipHostMalloc(void** p, size_t bytes, unsigned)
{
    if (fail_now(2)) return hipErrorOutOfMemory;
    *p = std::calloc(1, std::max<size_t>(bytes, 1));   // host memory: not in the device registry
    return *p ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipHostFree(void* p)
{
    std::free(p);
    return hipSuccess;
}

hipError_t hipGetLastError(void)
{
    const hipError_t e = sticky;
    sticky = hipSuccess;
    return e;
}

const char* hipGetErrorString(hipError_t e)
{
    switch (e) {
    case hipSuccess: return "no error";
    case hipErrorInvalidValue: return "invalid argument (injected or detected by the host check)";
    case hipErrorOutOfMemory: return "out of memory";
    case hipErrorLaunchFailure: return "unspecified launch failure";
    default: return "hip error";
    }
}

// ---------------------------------------------------------------------------
// launchers (extents as the kernels of msw_kernels.hip touch them)

namespace msw {

RKCoeffs rk_coeffs(double dt)
{
    RKCoeffs k;
    k.w[0] = dt / 6.0; k.w[1] = dt / 3.0; k.w[2] = dt / 3.0; k.w[3] = dt / 6.0;
    k.sw[0] = dt / 2.0; k.sw[1] = dt / 2.0; k.sw[2] = dt; k.sw[3] = 0.0;
    return k;
}

int dot_blocks(long n)
{
    long nb = (n + 4095) / 4096;
    if (nb < 1) nb = 1;
    if (nb > 2048) nb = 2048;
    return static_cast<int>(nb);
}

int dot_scratch(double** out)
{
    static void* p = nullptr;   // the library's file-scope array, registered once
    if (!p && hipMalloc(&p, kDotScratchDoubles * sizeof(double)) != hipSuccess) return hipErrorOutOfMemory;
    *out = static_cast<double*>(p);
    return 0;
}

namespace {

const size_t D = sizeof(double), I = sizeof(int);

Buf R(const void* p, size_t n, const char* w) { return {p, n, false, w, false, 0}; }
Buf W(const void* p, size_t n, const char* w) { return {p, n, true, w, false, 0}; }
Buf RO(const void* p, size_t n, const char* w) { return {p, n, false, w, true, 0}; }

// mesh tables every stage kernel may read
void mesh_reads(const DMesh& m, std::vector<Buf>& b)
{
    const size_t C = m.nC, E = m.nE, V = m.nV;
    const Buf t[] = {R(m.nEoC, C * I, "nEoC"), R(m.eoc, 6 * C * I, "eoc"), R(m.sgnC, 6 * C * D, "sgnC"),
                     R(m.eocR, 6 * C * I, "eocR"), R(m.sgnCR, 6 * C * D, "sgnCR"), R(m.dvCR, 6 * C * D, "dvCR"),
                     R(m.vocR, 6 * C * I, "vocR"), R(m.kiteCR, 6 * C * D, "kiteCR"), R(m.keC, 6 * C * D, "keC"),
                     R(m.areaC, C * D, "areaC"), R(m.hs, C * D, "hs"), R(m.coe, 2 * E * I, "coe"),
                     R(m.voe, 2 * E * I, "voe"), R(m.nEoE, E * I, "nEoE"), R(m.eoe, 12 * E * I, "eoe"),
                     R(m.woe, 12 * E * D, "woe"), R(m.eoeT, 12 * E * I, "eoeT"), R(m.eSlot, 2 * E * I, "eSlot"),
                     R(m.woeT, 12 * E * D, "woeT"), R(m.dc, E * D, "dc"), R(m.dv, E * D, "dv"),
                     R(m.eovR, 3 * V * I, "eovR"), R(m.sgnV, 3 * V * D, "sgnV"), R(m.sdcV, 3 * V * D, "sdcV"),
                     R(m.cov, 3 * V * I, "cov"), R(m.kite, 3 * V * D, "kite"), R(m.areaT, V * D, "areaT"),
                     R(m.fV, V * D, "fV")};
    b.insert(b.end(), std::begin(t), std::end(t));
}

void check_shape(const DMesh& m, const char* name)
{
    if (m.block != 64 && m.block != 128 && m.block != 256 && m.block != 512)
        violation(std::string(name) + ": block size " + std::to_string(m.block));
    if (m.stripe_log2 < 0 || m.stripe_log2 > 7)
        violation(std::string(name) + ": stripe option " + std::to_string(m.stripe_log2));
}

void config_reads(const DMesh& m, const DConfig& c, std::vector<Buf>& b)
{
    const size_t E = m.nE;
    if (c.wind) b.push_back(R(c.usrc, E * D, "u_src"));
    if (c.msd2) b.push_back(R(c.msd2, E * D, "meshScalingDel2"));
    if (c.msd4) b.push_back(R(c.msd4, E * D, "meshScalingDel4"));
}

// member array of n rows of B values; FP64 arrays of a B >= 64 launch are
// read in 16-byte pairs
template <typename T>
Buf MR(const T* p, size_t rows, int lb, const std::string& w, bool opt = false)
{
    return {p, (rows << lb) * sizeof(T), false, w, opt, (sizeof(T) == 8 && lb >= 6) ? 16u : 0u};
}
template <typename T>
Buf MW(T* p, size_t rows, int lb, const std::string& w, bool opt = false)
{
    return {p, (rows << lb) * sizeof(T), true, w, opt, (sizeof(T) == 8 && lb >= 6) ? 16u : 0u};
}

// RK4 bookkeeping rows of one stage (rk_update)
template <typename T>
void rk_bufs(std::vector<Buf>& b, int stage, T* x, T* acc, T* pn, size_t rows, int lb, bool frozen,
             const char* what)
{
    const std::string nx = std::string(what) + " x", na = std::string(what) + " acc",
                      nn = std::string(what) + " next";
    if (stage == 0) {
        b.push_back(MR(x, rows, lb, nx));
        b.push_back(MW(acc, rows, lb, na));
        b.push_back(MW(pn, rows, lb, nn));
    } else if (stage < 3) {
        b.push_back(MR(x, rows, lb, nx));
        b.push_back(MR(acc, rows, lb, na));
        b.push_back(MW(acc, rows, lb, na));
        b.push_back(MW(pn, rows, lb, nn));
    } else {
        b.push_back(MR(acc, rows, lb, na));
        if (!frozen) b.push_back(MW(x, rows, lb, nx));
    }
}

template <typename T>
void dissip_reads(const DMesh& m, const DConfig& c, int lb, const Dissip<T>& dp, std::vector<Buf>& b)
{
    if (c.visc2 > 0.0 || c.visc4 > 0.0) {
        b.push_back(MR(dp.div, m.nC, lb, "divergence"));
        b.push_back(MR(dp.vort, m.nV, lb, "vorticity"));
    }
    if (c.visc4 > 0.0) {
        b.push_back(MR(dp.dsd, m.nC, lb, "del2 u divergence"));
        b.push_back(MR(dp.dsv, m.nV, lb, "del2 u vorticity"));
    }
}

void base_reads(const BaseStage& s, size_t C, size_t E, size_t V, std::vector<Buf>& b, bool all)
{
    b.push_back(R(s.he, E * D, "base h_edge"));
    b.push_back(R(s.u, E * D, "base u"));
    if (!all) return;
    b.push_back(R(s.pe, E * D, "base pv_edge"));
    b.push_back(R(s.v, E * D, "base v"));
    b.push_back(R(s.gpt, E * D, "base gradPVt"));
    b.push_back(R(s.gpn, E * D, "base gradPVn"));
    b.push_back(R(s.ke, C * D, "base ke"));
    b.push_back(R(s.hv, V * D, "base h_vertex"));
    b.push_back(R(s.vort, V * D, "base vorticity"));
    b.push_back({s.rec, 4 * E * D, false, "base edge records", false, 32});
}

}  // namespace

template <typename T>
void fwd_diag(const DMesh& m, const DConfig& c, int lb, const T* h, const T* u, T* pvv, T* hv, T* vort,
              T* ke, T* pc, T* he, T* pe, T* v, T* gpt, T* gpn, double* rec, hipStream_t s)
{
    check_shape(m, "fwd_diag");
    const size_t C = m.nC, E = m.nE, V = m.nV;
    std::vector<Buf> b;
    mesh_reads(m, b);
    std::vector<Buf> bv = b, bc = b, be = b;
    bv.insert(bv.end(), {MR(h, C, lb, "h"), MR(u, E, lb, "u"), MW(pvv, V, lb, "pv_vertex"),
                         MW(hv, V, lb, "h_vertex", true), MW(vort, V, lb, "vorticity", true)});
    launch(s, "k_fwd_vertex", bv);
    bc.insert(bc.end(), {MR(u, E, lb, "u"), MR(pvv, V, lb, "pv_vertex"), MW(ke, C, lb, "ke"), MW(pc, C, lb, "pv_cell")});
    launch(s, "k_fwd_cell", bc);
    be.insert(be.end(), {MR(h, C, lb, "h"), MR(u, E, lb, "u"), MR(pvv, V, lb, "pv_vertex"), MR(pc, C, lb, "pv_cell"),
                         MW(he, E, lb, "h_edge"), MW(pe, E, lb, "pv_edge"), MW(v, E, lb, "v", true),
                         MW(gpt, E, lb, "gradPVt", true), MW(gpn, E, lb, "gradPVn", true),
                         {rec, 4 * E * D, true, "edge records", true, 32}});
    launch(s, "k_fwd_edge", be);
    (void)c;
}

template <typename T>
void fwd_tend_update_he(const DMesh& m, const DConfig& c, int lb, int stage, const RKCoeffs& k, const T* h,
                        const T* u, const T* he, const T* pe, const T* ke, const Dissip<T>& dp, T* xh, T* xu,
                        T* acch, T* accu, T* ph, T* pu, const double* rec, T* toh, T* tou, hipStream_t s)
{
    check_shape(m, "fwd_tend");
    const size_t C = m.nC, E = m.nE;
    std::vector<Buf> b;
    mesh_reads(m, b);
    config_reads(m, c, b);
    dissip_reads(m, c, lb, dp, b);
    if (lb == 0 && rec) b.push_back({rec, 4 * E * D, false, "edge records", false, 32});
    b.insert(b.end(), {MR(h, C, lb, "h"), MR(u, E, lb, "u"), MR(he, E, lb, "h_edge"), MR(pe, E, lb, "pv_edge"),
                       MR(ke, C, lb, "ke")});
    rk_bufs(b, stage, xu, accu, pu, E, lb, c.frozen_u != 0, "u");
    rk_bufs(b, stage, xh, acch, ph, C, lb, false, "h");
    // the last stage's tendencies, when kept (one value per lane, any alignment)
    if (stage == 3 && toh) b.push_back(Buf{toh, (C << lb) * sizeof(T), true, "tend_h", false, 0u});
    if (stage == 3 && tou) b.push_back(Buf{tou, (E << lb) * sizeof(T), true, "tend_u", false, 0u});
    launch(s, "k_fwd_tend", b);
    (void)k;
}

template <typename T>
Dissip<T> dissip_fields(const DMesh& m, const DConfig& c, int lb, const T* u, const T* vort,
                        const DissipScratch<T>& ds, hipStream_t s)
{
    Dissip<T> dp{nullptr, nullptr, nullptr, nullptr};
    if (!(c.visc2 > 0.0 || c.visc4 > 0.0)) return dp;
    const size_t C = m.nC, E = m.nE, V = m.nV;
    std::vector<Buf> b;
    mesh_reads(m, b);
    std::vector<Buf> b1 = b;
    b1.insert(b1.end(), {MR(u, E, lb, "u"), MW(ds.div, C, lb, "divergence")});
    launch(s, "k_div", b1);
    dp.div = ds.div;
    dp.vort = vort;
    if (c.visc4 > 0.0) {
        std::vector<Buf> b2 = b, b3 = b, b4 = b;
        b2.insert(b2.end(), {MR(ds.div, C, lb, "divergence"), MR(vort, V, lb, "vorticity"), MW(ds.dsu, E, lb, "del2 u")});
        launch(s, "k_del4_edge", b2);
        b3.insert(b3.end(), {MR(ds.dsu, E, lb, "del2 u"), MW(ds.dsv, V, lb, "del2 u vorticity")});
        launch(s, "k_del4_vertex", b3);
        b4.insert(b4.end(), {MR(ds.dsu, E, lb, "del2 u"), MW(ds.dsd, C, lb, "del2 u divergence")});
        launch(s, "k_del4_cell", b4);
        dp.dsd = ds.dsd;
        dp.dsv = ds.dsv;
    }
    return dp;
}

void tl_coef(const DMesh& m, const DConfig& c, const BaseStage& bs, const AdjCoef& q, bool f32, hipStream_t s)
{
    check_shape(m, "tl_coef");
    const size_t C = m.nC, E = m.nE, V = m.nV;
    const size_t cs = f32 ? sizeof(float) : sizeof(double);   // coefficient rows
    std::vector<Buf> b;
    mesh_reads(m, b);
    base_reads(bs, C, E, V, b, false);
    b.push_back({bs.rec, 4 * E * D, false, "base edge records", false, 32});
    b.push_back(R(bs.pe, E * D, "base pv_edge"));
    b.push_back(W(q.e48, 48 * E * cs, "TL edge coefficients"));
    b.push_back(W(q.c32, 32 * C * cs, "TL cell coefficients"));
    launch(s, "k_tl_coef", b);
    (void)c;
}

void adj_coef(const DMesh& m, const DConfig& c, const BaseStage& bs, const AdjCoef& q, bool f32, hipStream_t s)
{
    check_shape(m, "adj_coef");
    const size_t C = m.nC, E = m.nE, V = m.nV;
    const size_t cs = f32 ? sizeof(float) : sizeof(double);   // coefficient rows
    std::vector<Buf> b;
    mesh_reads(m, b);
    config_reads(m, c, b);
    base_reads(bs, C, E, V, b, true);
    std::vector<Buf> be = b, bv = b, bc = b;
    be.insert(be.end(), {W(q.e48, 24 * E * cs, "adjoint edge slot coefficients"),
                         W(q.e8, 16 * E * cs, "adjoint edge coefficients")});
    launch(s, "k_adj_coef_edge", be);
    bv.push_back(W(q.v8, 8 * V * cs, "adjoint vertex coefficients"));
    launch(s, "k_adj_coef_vertex", bv);
    bc.push_back(W(q.c32, 32 * C * cs, "adjoint cell coefficients"));
    launch(s, "k_adj_coef_cell", bc);
}

void adj_coef4(const DMesh& m, const DConfig& c, const BaseStage* bs, const AdjCoef* q, bool f32, hipStream_t s)
{
    // one launch per kernel over the four stages: the same buffers as four
    // single-stage launches
    for (int i = 0; i < 4; ++i) adj_coef(m, c, bs[i], q[i], f32, s);
}

template <typename T>
void tl_stage(const DMesh& m, const DConfig& c, int lb, int stage, const RKCoeffs& k, const BaseStage& bs,
              const T* h_tl, const T* u_tl, T* pvv_tl, T* ke_tl, T* pc_tl, T* pe_tl, T* he_tl, T* vort_tl,
              const DissipScratch<T>& ds, const AdjCoef& q, T* xh, T* xu, T* acch, T* accu, T* ph, T* pu,
              T* toh, T* tou, hipStream_t s)
{
    check_shape(m, "tl_stage");
    const size_t C = m.nC, E = m.nE, V = m.nV;
    const bool visc = c.visc2 > 0.0 || c.visc4 > 0.0;
    std::vector<Buf> b;
    mesh_reads(m, b);
    std::vector<Buf> b1 = b, b2 = b, b3 = b, b4 = b;
    b1.insert(b1.end(), {R(bs.hv, V * D, "base h_vertex"), R(bs.vort, V * D, "base vorticity"), MR(h_tl, C, lb, "h_tl"),
                         MR(u_tl, E, lb, "u_tl"), MW(pvv_tl, V, lb, "pv_vertex_tl")});
    if (visc) b1.push_back(MW(vort_tl, V, lb, "vorticity_tl"));
    launch(s, "k_tl_vertex", b1);
    b2.insert(b2.end(), {R(q.c32, 32 * C * sizeof(T), "TL cell coefficients"), MR(u_tl, E, lb, "u_tl"),
                         MR(pvv_tl, V, lb, "pv_vertex_tl"), MW(ke_tl, C, lb, "ke_tl"), MW(pc_tl, C, lb, "pv_cell_tl")});
    launch(s, "k_tl_cell", b2);
    b3.insert(b3.end(), {R(bs.gpt, E * D, "base gradPVt"), R(bs.v, E * D, "base v"), R(bs.gpn, E * D, "base gradPVn"),
                         R(bs.u, E * D, "base u"), R(bs.he, E * D, "base h_edge"), MR(h_tl, C, lb, "h_tl"),
                         MR(u_tl, E, lb, "u_tl"), MR(pvv_tl, V, lb, "pv_vertex_tl"), MR(pc_tl, C, lb, "pv_cell_tl"),
                         MW(pe_tl, E, lb, "pv_edge_tl"), MW(he_tl, E, lb, "h_edge_tl")});
    launch(s, "k_tl_edge", b3);
    const Dissip<T> dp = dissip_fields<T>(m, c, lb, u_tl, vort_tl, ds, s);
    config_reads(m, c, b4);
    dissip_reads(m, c, lb, dp, b4);
    if (c.drag || c.wind) {
        b4.push_back(R(bs.he, E * D, "base h_edge"));
        b4.push_back(R(bs.u, E * D, "base u"));
        b4.push_back(R(bs.ke, C * D, "base ke"));
    }
    b4.insert(b4.end(), {R(q.e48, 48 * E * sizeof(T), "TL edge coefficients"), R(q.c32, 32 * C * sizeof(T), "TL cell coefficients"),
                         MR(h_tl, C, lb, "h_tl"), MR(u_tl, E, lb, "u_tl"), MR(pe_tl, E, lb, "pv_edge_tl"),
                         MR(he_tl, E, lb, "h_edge_tl"), MR(ke_tl, C, lb, "ke_tl")});
    rk_bufs(b4, stage, xu, accu, pu, E, lb, c.frozen_u != 0, "u_tl");
    rk_bufs(b4, stage, xh, acch, ph, C, lb, false, "h_tl");
    if (stage == 3 && toh) b4.push_back(Buf{toh, (C << lb) * sizeof(T), true, "tend_h_tl", false, 0u});
    if (stage == 3 && tou) b4.push_back(Buf{tou, (E << lb) * sizeof(T), true, "tend_u_tl", false, 0u});
    launch(s, "k_tl_tend", b4);
    (void)k;
}

template <typename T>
void adj_stage(const DMesh& m, const DConfig& c, int lb, int stage, bool first, const RKCoeffs& k,
               const BaseStage& bs, T* gh, T* gu, T* pe_ad, T* he_ad, T* ua_p, T* pc_ad, T* ha_p, T* ke_ad,
               T* hv_ad, T* circ_ad, const DissipScratch<T>& ds, const AdjCoef& q, T* lamh, T* lamu, T* acch,
               T* accu, hipStream_t s)
{
    check_shape(m, "adj_stage");
    const size_t C = m.nC, E = m.nE, V = m.nV;
    const bool visc = c.visc2 > 0.0 || c.visc4 > 0.0;
    const T* rgh = first ? lamh : gh;
    const T* rgu = (first && !c.frozen_u) ? lamu : gu;
    std::vector<Buf> b;
    mesh_reads(m, b);
    config_reads(m, c, b);
    if (c.visc4 > 0.0) {
        std::vector<Buf> d1 = b, d2 = b, d3 = b;
        d1.insert(d1.end(), {MR(rgu, E, lb, "u tendency adjoint"), MW(ds.dsv, V, lb, "del2 u circulation adjoint")});
        launch(s, "k_adj_del4_vertex", d1);
        d2.insert(d2.end(), {MR(rgu, E, lb, "u tendency adjoint"), MW(ds.dsd, C, lb, "del2 u divergence adjoint")});
        launch(s, "k_adj_del4_cell", d2);
        d3.insert(d3.end(), {MR(ds.dsd, C, lb, "del2 u divergence adjoint"), MR(ds.dsv, V, lb, "del2 u circulation adjoint"),
                             MW(ds.dsu, E, lb, "del2 u adjoint")});
        launch(s, "k_adj_del4_edge", d3);
    }
    std::vector<Buf> b1 = b, b2 = b, b3 = b, b4 = b;
    b1.insert(b1.end(), {R(q.e48, 24 * E * sizeof(T), "adjoint edge slot coefficients"), R(q.e8, 16 * E * sizeof(T), "adjoint edge coefficients"),
                         MR(rgh, C, lb, "h tendency adjoint"), MR(rgu, E, lb, "u tendency adjoint"),
                         MW(pe_ad, E, lb, "pv_edge_ad"), MW(he_ad, E, lb, "h_edge_ad"), MW(ua_p, E, lb, "partial u_ad")});
    launch(s, "k_adj_edge1", b1);
    b2.insert(b2.end(), {R(q.c32, 32 * C * sizeof(T), "adjoint cell coefficients"), MR(rgu, E, lb, "u tendency adjoint"),
                         MR(pe_ad, E, lb, "pv_edge_ad"), MR(he_ad, E, lb, "h_edge_ad"), MW(pc_ad, C, lb, "pv_cell_ad"),
                         MW(ha_p, C, lb, "partial h_ad"), MW(ke_ad, C, lb, "ke_ad")});
    if (c.visc4 > 0.0) b2.push_back(MR(ds.dsu, E, lb, "del2 u adjoint"));
    if (visc) b2.push_back(MW(ds.div, C, lb, "divergence adjoint"));
    launch(s, "k_adj_cell1", b2);
    b3.insert(b3.end(), {R(q.v8, 8 * V * sizeof(T), "adjoint vertex coefficients"), MR(pe_ad, E, lb, "pv_edge_ad"),
                         MR(pc_ad, C, lb, "pv_cell_ad"), MW(hv_ad, V, lb, "h_vertex_ad"), MW(circ_ad, V, lb, "circulation_ad")});
    if (c.visc2 > 0.0) b3.push_back(MR(rgu, E, lb, "u tendency adjoint"));
    if (c.visc4 > 0.0) b3.push_back(MR(ds.dsu, E, lb, "del2 u adjoint"));
    launch(s, "k_adj_vertex", b3);
    b4.insert(b4.end(), {R(q.e48, 24 * E * sizeof(T), "adjoint edge slot coefficients"), R(q.e8, 16 * E * sizeof(T), "adjoint edge coefficients"),
                         MR(pe_ad, E, lb, "pv_edge_ad"), MR(ua_p, E, lb, "partial u_ad"), MR(ha_p, C, lb, "partial h_ad"),
                         MR(hv_ad, V, lb, "h_vertex_ad"), MR(circ_ad, V, lb, "circulation_ad"), MR(ke_ad, C, lb, "ke_ad")});
    if (visc) b4.push_back(MR(ds.div, C, lb, "divergence adjoint"));
    // adj_update: stage > 0: acc (first: written, else read and written),
    // g written (u: not with the wind frozen), lam read; stage 0: acc and lam
    // read, lam written
    for (int kind = 0; kind < 2; ++kind) {
        const size_t rows = kind ? C : E;
        T* lam = kind ? lamh : lamu;
        T* acc = kind ? acch : accu;
        T* g = kind ? gh : gu;
        const bool frozen = !kind && c.frozen_u;
        if (stage > 0) {
            if (!first) b4.push_back(MR(acc, rows, lb, kind ? "h adjoint accumulator" : "u adjoint accumulator"));
            b4.push_back(MW(acc, rows, lb, kind ? "h adjoint accumulator" : "u adjoint accumulator"));
            b4.push_back(MR(lam, rows, lb, kind ? "h adjoint" : "u adjoint"));
            if (!frozen) b4.push_back(MW(g, rows, lb, kind ? "h tendency adjoint" : "u tendency adjoint"));
        } else {
            b4.push_back(MR(acc, rows, lb, kind ? "h adjoint accumulator" : "u adjoint accumulator"));
            b4.push_back(MR(lam, rows, lb, kind ? "h adjoint" : "u adjoint"));
            b4.push_back(MW(lam, rows, lb, kind ? "h adjoint" : "u adjoint"));
        }
    }
    launch(s, "k_adj_final", b4);
    (void)k;
    (void)bs;
}

template <typename T>
void adj_seed(int n, T* g, const T* lam, double w3, hipStream_t s)
{
    launch(s, "k_scale", {R(lam, static_cast<size_t>(n) * sizeof(T), "lam"), W(g, static_cast<size_t>(n) * sizeof(T), "g")});
    (void)w3;
}

template <typename T>
void diag_adj(const DMesh& m, const DConfig& c, int lb, const DiagBase& bd, const DiagSeeds<T>& sd, T* e1, T* e2,
              T* e3, T* c1, T* v1, T* v2, T* h_ad, T* u_ad, hipStream_t s)
{
    check_shape(m, "diag_adj");
    const size_t C = m.nC, E = m.nE, V = m.nV;
    std::vector<Buf> b;
    mesh_reads(m, b);
    std::vector<Buf> b1 = b, b2 = b, b3 = b, b4 = b;
    b1.insert(b1.end(), {R(bd.u, E * D, "base u"), R(bd.v, E * D, "base v"), R(bd.gpt, E * D, "base gradPVt"),
                         MR(sd.pe, E, lb, "pv_edge seed"), MR(sd.gpn, E, lb, "gradPVn seed"), MR(sd.gpt, E, lb, "gradPVt seed"),
                         MR(sd.v, E, lb, "v seed"), MW(e1, E, lb, "gradPVn adjoint"), MW(e2, E, lb, "gradPVt adjoint"),
                         MW(e3, E, lb, "v adjoint")});
    launch(s, "k_dadj_edge", b1);
    b2.insert(b2.end(), {MR(sd.pc, C, lb, "pv_cell seed"), MR(e1, E, lb, "gradPVn adjoint"), MW(c1, C, lb, "pv_cell adjoint")});
    launch(s, "k_dadj_cell", b2);
    b3.insert(b3.end(), {R(bd.hv, V * D, "base h_vertex"), R(bd.vort, V * D, "base vorticity"), MR(sd.vort, V, lb, "vorticity seed"),
                         MR(sd.pvv, V, lb, "pv_vertex seed"), MR(sd.vc, C, lb, "vorticity_cell seed"), MR(c1, C, lb, "pv_cell adjoint"),
                         MR(sd.pe, E, lb, "pv_edge seed"), MR(e2, E, lb, "gradPVt adjoint"), MR(sd.hv, V, lb, "h_vertex seed"),
                         MR(sd.circ, V, lb, "circulation seed"), MW(v1, V, lb, "h_vertex adjoint"), MW(v2, V, lb, "circulation adjoint")});
    launch(s, "k_dadj_vertex", b3);
    b4.insert(b4.end(), {R(bd.gpn, E * D, "base gradPVn"), R(bd.u, E * D, "base u"), MR(sd.pe, E, lb, "pv_edge seed"),
                         MR(e3, E, lb, "v adjoint"), MR(sd.ke, C, lb, "ke seed"), MR(sd.div, C, lb, "divergence seed"),
                         MR(v2, V, lb, "circulation adjoint"), MR(v1, V, lb, "h_vertex adjoint"), MR(sd.he, E, lb, "h_edge seed"),
                         MR(h_ad, C, lb, "h_ad"), MW(h_ad, C, lb, "h_ad"), MR(u_ad, E, lb, "u_ad"), MW(u_ad, E, lb, "u_ad")});
    launch(s, "k_dadj_final", b4);
    (void)c;
}

template <typename T>
void reconstruct(int nC, int lb, const int* nEoC, const int* eoc, const double* coeffs, const double* trig, const T* u,
                 T* out5, hipStream_t s)
{
    const size_t C = nC;
    launch(s, "k_reconstruct", {R(nEoC, C * I, "nEoC"), R(eoc, 6 * C * I, "eoc"), R(coeffs, 18 * C * D, "coefficients"),
                                R(trig, 4 * C * D, "trig"), R(u, sizeof(T), "u (extent not known here)"),
                                MW(out5, 5 * C, lb, "reconstructed velocity")});
}

template <typename T>
void reconstruct_adj(const DMesh& m, int lb, const double* coeffs, const double* trig, const T* seeds5, T* xyz,
                     T* u_ad, hipStream_t s)
{
    check_shape(m, "reconstruct_adj");
    const size_t C = m.nC, E = m.nE;
    std::vector<Buf> b;
    mesh_reads(m, b);
    std::vector<Buf> b1 = b, b2 = b;
    b1.insert(b1.end(), {R(trig, 4 * C * D, "trig"), MR(seeds5, 5 * C, lb, "seeds"), MW(xyz, 3 * C, lb, "xyz adjoint")});
    launch(s, "k_recon_adj_cell", b1);
    b2.insert(b2.end(), {R(coeffs, 18 * C * D, "coefficients"), MR(xyz, 3 * C, lb, "xyz adjoint"), MR(u_ad, E, lb, "u_ad"),
                         MW(u_ad, E, lb, "u_ad")});
    launch(s, "k_recon_adj_edge", b2);
}

template <typename T>
void dot_batched(long n, int lb, const T* x, const T* y, const double* w, double* partial, int nblk, double* out,
                 int accumulate, hipStream_t s)
{
    const size_t rows = static_cast<size_t>(n), B = size_t(1) << lb;
    launch(s, "k_dot_partial", {MR(x, rows, lb, "x"), MR(y, rows, lb, "y", true), RO(w, rows * D, "weights"),
                                W(partial, static_cast<size_t>(nblk) * B * D, "partials")});
    std::vector<Buf> b = {R(partial, static_cast<size_t>(nblk) * B * D, "partials"), W(out, B * D, "out")};
    if (accumulate) b.push_back(R(out, B * D, "out"));
    launch(s, "k_dot_final", b);
}

template <typename TD, typename TS>
void convert(long n, TD* dst, const TS* src, hipStream_t s)
{
    launch(s, "k_convert", {R(src, n * sizeof(TS), "src"), W(dst, n * sizeof(TD), "dst")});
}

template <typename T>
void fill(long n, T* dst, T val, hipStream_t s)
{
    launch(s, "k_fill", {W(dst, n * sizeof(T), "dst")});
    (void)val;
}

template <typename T>
void gather_rows(long n, int lb, const int* perm, const T* src, T* dst, hipStream_t s)
{
    launch(s, "k_gather_rows", {R(perm, n * I, "perm"), MR(src, n, lb, "src"), MW(dst, n, lb, "dst")});
}

template <typename T>
void scatter_rows(long n, int lb, const int* perm, const T* src, T* dst, hipStream_t s)
{
    launch(s, "k_scatter_rows", {R(perm, n * I, "perm"), MR(src, n, lb, "src"), MW(dst, n, lb, "dst")});
}

template <typename T>
void nonfinite_flags(long n, int lb, const T* x, int* flags, hipStream_t s)
{
    const size_t B = size_t(1) << lb;
    launch(s, "k_nonfinite", {MR(x, n, lb, "x"), R(flags, B * I, "flags"), W(flags, B * I, "flags")});
}

void diag_extra(const DMesh& m, const double* u, const double* vort, double* circ, double* div, double* vc, hipStream_t s)
{
    const size_t C = m.nC, E = m.nE, V = m.nV;
    std::vector<Buf> b;
    mesh_reads(m, b);
    b.insert(b.end(), {R(u, E * D, "u"), R(vort, V * D, "vorticity"), W(circ, V * D, "circulation"),
                       W(div, C * D, "divergence"), W(vc, C * D, "vorticity_cell")});
    launch(s, "k_diag_extra", b);
}

template <typename T>
int persistent_steps(const PersistArgs& a, bool tlm, hipStream_t s, int* grid)
{
    const DMesh& m = a.m;
    check_shape(m, "persistent_steps");
    const size_t C = m.nC, E = m.nE, V = m.nV, n = static_cast<size_t>(a.nsteps);
    std::vector<Buf> b;
    mesh_reads(m, b);
    b.insert(b.end(), {R(a.patch, static_cast<size_t>(a.np_v + a.np_c + a.np_e) * sizeof(int2), "patches"),
                       R(a.queue, 16 * 8, "queue"), W(a.queue, 16 * 8, "queue"), R(a.bar, 8, "barrier"), W(a.bar, 8, "barrier"),
                       R(a.err, 4, "error flag"), W(a.err, 4, "error flag"),
                       MR(static_cast<const T*>(a.h), C, 0, "h"), MW(static_cast<T*>(a.h), C, 0, "h"),
                       MR(static_cast<const T*>(a.u), E, 0, "u"), MW(static_cast<T*>(a.u), E, 0, "u"),
                       MW(static_cast<T*>(a.acc), C + E, 0, "acc"), MW(static_cast<T*>(a.pA), C + E, 0, "stage A"),
                       MW(static_cast<T*>(a.pB), C + E, 0, "stage B"), MW(static_cast<T*>(a.pvv), V, 0, "pv_vertex"),
                       MW(static_cast<T*>(a.ke), C, 0, "ke"), MW(static_cast<T*>(a.pc), C, 0, "pv_cell"),
                       MW(static_cast<T*>(a.he), E, 0, "h_edge"), MW(static_cast<T*>(a.pe), E, 0, "pv_edge")});
    if (a.np_v < 1 || a.np_c < 1 || a.np_e < 1) violation("persistent_steps: an entity kind has no work items");
    if (tlm) {
        b.insert(b.end(), {R(a.bh, C * D, "base h"), W(a.bh, C * D, "base h"), R(a.bu, E * D, "base u"),
                           W(a.bu, E * D, "base u"), W(a.bacc, (C + E) * D, "base acc"), W(a.bend, (C + E) * D, "base end")});
        for (int st = 0; st < 4; ++st) {
            const BaseStage& q = a.st[st];
            b.insert(b.end(), {W(q.h, C * D, "base stage h"), W(q.u, E * D, "base stage u"), W(q.he, E * D, "base h_edge"),
                               W(q.pe, E * D, "base pv_edge"), W(q.v, E * D, "base v"), W(q.gpt, E * D, "base gradPVt"),
                               W(q.gpn, E * D, "base gradPVn"), W(q.ke, C * D, "base ke"), W(q.pc, C * D, "base pv_cell"),
                               W(q.pvv, V * D, "base pv_vertex"), W(q.hv, V * D, "base h_vertex"),
                               W(q.vort, V * D, "base vorticity"), {q.rec, 4 * E * D, true, "base records", false, 32}});
        }
        if (a.traj_h) {
            b.push_back(W(a.traj_h, n * C * D, "trajectory h"));
            b.push_back(W(a.traj_u, n * E * D, "trajectory u"));
        }
    }
    if (grid) *grid = 1024;
    launch(s, "k_persist", b);
    return 0;
}

template int persistent_steps<double>(const PersistArgs&, bool, hipStream_t, int*);
template int persistent_steps<float>(const PersistArgs&, bool, hipStream_t, int*);

#define INST(T)                                                                                                  \
    template void fwd_diag<T>(const DMesh&, const DConfig&, int, const T*, const T*, T*, T*, T*, T*, T*, T*, T*,  \
                              T*, T*, T*, double*, hipStream_t);                                                 \
    template void fwd_tend_update_he<T>(const DMesh&, const DConfig&, int, int, const RKCoeffs&, const T*,       \
                                        const T*, const T*, const T*, const T*, const Dissip<T>&, T*, T*, T*, T*, \
                                        T*, T*, const double*, T*, T*, hipStream_t);                             \
    template Dissip<T> dissip_fields<T>(const DMesh&, const DConfig&, int, const T*, const T*,                    \
                                        const DissipScratch<T>&, hipStream_t);                                   \
    template void tl_stage<T>(const DMesh&, const DConfig&, int, int, const RKCoeffs&, const BaseStage&,         \
                              const T*, const T*, T*, T*, T*, T*, T*, T*, const DissipScratch<T>&,               \
                              const AdjCoef&, T*, T*, T*, T*, T*, T*, T*, T*, hipStream_t);                      \
    template void adj_stage<T>(const DMesh&, const DConfig&, int, int, bool, const RKCoeffs&, const BaseStage&,  \
                               T*, T*, T*, T*, T*, T*, T*, T*, T*, T*, const DissipScratch<T>&, const AdjCoef&,  \
                               T*, T*, T*, T*, hipStream_t);                                                     \
    template void adj_seed<T>(int, T*, const T*, double, hipStream_t);                                           \
    template void diag_adj<T>(const DMesh&, const DConfig&, int, const DiagBase&, const DiagSeeds<T>&, T*, T*,   \
                              T*, T*, T*, T*, T*, T*, hipStream_t);                                              \
    template void reconstruct_adj<T>(const DMesh&, int, const double*, const double*, const T*, T*, T*,          \
                                     hipStream_t);                                                               \
    template void reconstruct<T>(int, int, const int*, const int*, const double*, const double*, const T*, T*,   \
                                 hipStream_t);                                                                   \
    template void dot_batched<T>(long, int, const T*, const T*, const double*, double*, int, double*, int,       \
                                 hipStream_t);                                                                   \
    template void fill<T>(long, T*, T, hipStream_t);                                                             \
    template void gather_rows<T>(long, int, const int*, const T*, T*, hipStream_t);                              \
    template void scatter_rows<T>(long, int, const int*, const T*, T*, hipStream_t);                             \
    template void nonfinite_flags<T>(long, int, const T*, int*, hipStream_t);

INST(double)
INST(float)
template void convert<float, double>(long, float*, const double*, hipStream_t);
template void convert<double, float>(long, double*, const float*, hipStream_t);
template void convert<double, double>(long, double*, const double*, hipStream_t);
template void convert<float, float>(long, float*, const float*, hipStream_t);

}  // namespace msw

// ---------------------------------------------------------------------------
// controls for the test driver

extern "C" {

void* mock_dev_alloc(size_t bytes)
{
    void* p = nullptr;
    return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}

int mock_dev_free(void* p) { return hipFree(p) == hipSuccess ? 0 : -1; }

// the n-th call (0 = the next one) of kind 0 (copy/memset), 1 (launch) or
// 2 (allocation) fails
void mock_fail(int kind, long n)
{
    ncalls[kind] = 0;
    fail_at[kind] = n;
}

void mock_clear_fail(void)
{
    for (long& f : fail_at) f = -1;
}

long mock_calls(int kind) { return ncalls[kind]; }

long mock_violations(void) { return nviol; }

const char* mock_message(int i) { return i >= 0 && i < static_cast<int>(msgs.size()) ? msgs[i].c_str() : ""; }

long mock_graphs_live(void) { return graphs_live; }

long mock_graphs_made(void) { return graphs_made; }

long mock_allocations(void) { return static_cast<long>(allocs.size()); }

}  // extern "C"
