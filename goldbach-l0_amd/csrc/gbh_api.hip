// C ABI of the device pipeline (include/gbh.h): context, memory, streams,
// events and one launch function per stage. Replaces the reference's Level
// Zero layer (src/l0_context.cpp:75-356) and the positional kernel argument
// lists of src/verifier.cpp; the kernels themselves are in gbh_kernels.hip,
// compiled into this translation unit so the library is one code object.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gbh.h"
#include "gbh_kernels.hip"

// The allocation-free counter's workspace (gbh_launch_count_noalloc): one
// copy per device, at the device address gbh_open looks up. In host code
// this name is a host-side stand-in, never a device address: launches take
// the looked-up pointer.
static __device__ __align__(256) uint32_t scratch[GBH_NOALLOC_SCRATCH_WORDS];

struct gbh_ctx {
    int device = 0;
    hipStream_t stream[GBH_STREAMS] = {};
    gbh_device_info_t info{};
    uint32_t* scratch_dev = nullptr;   // `scratch` on this context's device
};
struct gbh_event {
    hipEvent_t ev = nullptr;
};
struct gbh_graph {
    hipGraphExec_t exec = nullptr;
    uint64_t instantiated = 0, updated = 0;
};

namespace {

thread_local std::string g_err;
uint32_t g_min_useful = 0;   // fewest useful candidates of any even class

int fail(int code, const std::string& what)
{
    g_err = what;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(GBH_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define GBH_HIP(call)                                              \
    do {                                                           \
        const hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(e_, #call);          \
    } while (0)

hipStream_t S(gbh_ctx* c, int s) { return c->stream[unsigned(s) % GBH_STREAMS]; }

// Candidate tables of the checker (static: they depend on the even class
// only). Per even class r = 2 cr: the odd primes p <= 9781 with r - p a unit
// mod 30, ascending (the "useful" candidates of make_stats_plan,
// gstats.hpp:163-180; the futile ones can only give q = 3 or 5, which the
// class map does not hold and only evens below 9786 would need).
int upload_candidates()
{
    std::vector<uint32_t> pr;
    std::vector<char> comp(9782, 0);
    for (uint32_t p = 3; p * p <= 9781; p += 2)
        if (!comp[p]) for (uint32_t m = p * p; m <= 9781; m += 2 * p) comp[m] = 1;
    for (uint32_t p = 3; p <= 9781; p += 2) if (!comp[p]) pr.push_back(p);
    if (pr.size() != GBH_NCAND) return fail(GBH_E_ARG, "candidate list size");
    static const int cls_of[30] = {-1, 0, -1, -1, -1, -1, -1, 1, -1, -1, -1, 2, -1, 3, -1,
                                   -1, -1, 4, -1, 5, -1, -1, -1, 6, -1, -1, -1, -1, -1, 7};
    std::vector<uint16_t> cp(GBH_EVEN_CLASSES * GBH_NCAND, 0), cpi(GBH_EVEN_CLASSES * GBH_NCAND, 0);
    std::vector<uint8_t> cc(GBH_EVEN_CLASSES * GBH_NCAND, 0);
    uint32_t cn[GBH_EVEN_CLASSES] = {};
    for (uint32_t cr = 0; cr < GBH_EVEN_CLASSES; ++cr) {
        for (uint32_t i = 0; i < pr.size(); ++i) {
            const uint32_t res = (2 * cr + 30 - pr[i] % 30) % 30;
            if (cls_of[res] < 0) continue;
            const uint32_t j = cn[cr]++;
            cp[cr * GBH_NCAND + j] = uint16_t(pr[i]);
            cpi[cr * GBH_NCAND + j] = uint16_t(i);
            cc[cr * GBH_NCAND + j] = uint8_t(cls_of[res]);
        }
    }
    GBH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(gbh_dev::c_cand_p), cp.data(), cp.size() * 2));
    GBH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(gbh_dev::c_cand_pi), cpi.data(), cpi.size() * 2));
    GBH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(gbh_dev::c_cand_c), cc.data(), cc.size()));
    GBH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(gbh_dev::c_cand_n), cn, sizeof cn));
    g_min_useful = *std::min_element(cn, cn + GBH_EVEN_CLASSES);
    return GBH_OK;
}

unsigned grid_of(uint64_t n, unsigned block, unsigned cap)
{
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return unsigned(g);
}

}  // namespace

extern "C" {

const char* gbh_last_error(void) { return g_err.c_str(); }

int gbh_open(int device, gbh_ctx** out)
{
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
        return fail(GBH_E_NODEV, "no HIP device " + std::to_string(device));
    GBH_HIP(hipSetDevice(device));
    hipDeviceProp_t pr;
    GBH_HIP(hipGetDeviceProperties(&pr, device));
    if (std::strncmp(pr.gcnArchName, "gfx950", 6) != 0)
        return fail(GBH_E_NODEV, std::string("device is ") + pr.gcnArchName + ", this library is built for gfx950");
    auto* c = new gbh_ctx;
    c->device = device;
    for (auto& s : c->stream)
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            gbh_close(c);
            return fail(GBH_E_HIP, "hipStreamCreate");
        }
    std::snprintf(c->info.name, sizeof c->info.name, "%s", pr.name);
    std::snprintf(c->info.arch, sizeof c->info.arch, "%s", pr.gcnArchName);
    c->info.cus = pr.multiProcessorCount;
    c->info.xcds_hint = 8;
    c->info.lds_per_block = pr.sharedMemPerBlock;
    c->info.hbm_bytes = pr.totalGlobalMem;
    c->info.clock_khz = pr.clockRate;
    c->info.mem_clock_khz = pr.memoryClockRate;
    if (const int rc = upload_candidates()) { gbh_close(c); return rc; }
    // the allocation-free counter's workspace on this device (the device is current)
    const hipError_t es = hipGetSymbolAddress(reinterpret_cast<void**>(&c->scratch_dev), HIP_SYMBOL(scratch));
    if (es != hipSuccess || !c->scratch_dev) {
        gbh_close(c);
        return hip_fail(es != hipSuccess ? es : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    }
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&gbh_dev::k_window),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) (void)hipGetLastError();   // optional: windows up to the default limit still work
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&gbh_dev::k_harvest<false, false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 144 * 1024) != hipSuccess)
        (void)hipGetLastError();
    *out = c;
    return GBH_OK;
}

void gbh_close(gbh_ctx* c)
{
    if (!c) return;
    for (auto& s : c->stream)
        if (s) (void)hipStreamDestroy(s);
    delete c;
}

int gbh_device_info(gbh_ctx* c, gbh_device_info_t* info)
{
    *info = c->info;
    return GBH_OK;
}

int gbh_alloc(gbh_ctx*, size_t bytes, int kind, void** out)
{
    *out = nullptr;
    if (kind == 1) GBH_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
    else GBH_HIP(hipMalloc(out, bytes));
    return GBH_OK;
}

int gbh_free(gbh_ctx*, void* p, int kind)
{
    if (!p) return GBH_OK;
    if (kind == 1) GBH_HIP(hipHostFree(p));
    else GBH_HIP(hipFree(p));
    return GBH_OK;
}

int gbh_copy_h2d(gbh_ctx* c, int s, void* dst, const void* src, size_t bytes)
{
    GBH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(c, s)));
    return GBH_OK;
}

int gbh_copy_d2h(gbh_ctx* c, int s, void* dst, const void* src, size_t bytes)
{
    GBH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(c, s)));
    return GBH_OK;
}

int gbh_copy_d2h_rows(gbh_ctx* c, int s, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                      size_t rows)
{
    if (width > dpitch || width > spitch) return fail(GBH_E_ARG, "gbh_copy_d2h_rows: width above a pitch");
    if (!rows || !width) return GBH_OK;
    GBH_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDeviceToHost, S(c, s)));
    return GBH_OK;
}

int gbh_zero(gbh_ctx* c, int s, void* dst, size_t bytes)
{
    if (bytes % 4) return fail(GBH_E_ARG, "gbh_zero: size not a multiple of 4");
    const size_t n = bytes / 4;
    hipLaunchKernelGGL(gbh_dev::k_fill, dim3(grid_of(n, 256, 4096)), dim3(256), 0, S(c, s),
                       static_cast<uint32_t*>(dst), n, 0u);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_fill32(gbh_ctx* c, int s, void* dst, size_t n, uint32_t v)
{
    if (!n) return GBH_OK;
    hipLaunchKernelGGL(gbh_dev::k_fill, dim3(grid_of(n, 256, 4096)), dim3(256), 0, S(c, s),
                       static_cast<uint32_t*>(dst), n, v);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_stream_sync(gbh_ctx* c, int s)
{
    GBH_HIP(hipStreamSynchronize(S(c, s)));
    return GBH_OK;
}

int gbh_event_create(gbh_ctx*, gbh_event** ev)
{
    auto* e = new gbh_event;
    const hipError_t r = hipEventCreate(&e->ev);
    if (r != hipSuccess) { delete e; return hip_fail(r, "hipEventCreate"); }
    *ev = e;
    return GBH_OK;
}

int gbh_event_destroy(gbh_ctx*, gbh_event* ev)
{
    if (!ev) return GBH_OK;
    (void)hipEventDestroy(ev->ev);
    delete ev;
    return GBH_OK;
}

int gbh_event_record(gbh_ctx* c, int s, gbh_event* ev)
{
    GBH_HIP(hipEventRecord(ev->ev, S(c, s)));
    return GBH_OK;
}

int gbh_event_sync(gbh_ctx*, gbh_event* ev)
{
    GBH_HIP(hipEventSynchronize(ev->ev));
    return GBH_OK;
}

int gbh_event_elapsed_ms(gbh_ctx*, gbh_event* a, gbh_event* b, float* ms)
{
    GBH_HIP(hipEventElapsedTime(ms, a->ev, b->ev));
    return GBH_OK;
}

int gbh_stream_wait(gbh_ctx* c, int s, gbh_event* ev)
{
    GBH_HIP(hipStreamWaitEvent(S(c, s), ev->ev, 0));
    return GBH_OK;
}

int gbh_graph_create(gbh_ctx*, gbh_graph** g)
{
    *g = new gbh_graph;
    return GBH_OK;
}

int gbh_graph_destroy(gbh_ctx*, gbh_graph* g)
{
    if (!g) return GBH_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    delete g;
    return GBH_OK;
}

int gbh_graph_begin(gbh_ctx* c, int s)
{
    GBH_HIP(hipStreamBeginCapture(S(c, s), hipStreamCaptureModeRelaxed));
    return GBH_OK;
}

int gbh_graph_end(gbh_ctx* c, int s, gbh_graph* g)
{
    hipGraph_t graph = nullptr;
    GBH_HIP(hipStreamEndCapture(S(c, s), &graph));
    bool ok = false;
    if (g->exec) {
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        hipGraphNode_t bad = nullptr;
        ok = hipGraphExecUpdate(g->exec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (ok) {
            ++g->updated;
        } else {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(g->exec);
            g->exec = nullptr;
        }
    }
    if (!ok) {
        const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            g->exec = nullptr;
            (void)hipGraphDestroy(graph);
            return hip_fail(e, "hipGraphInstantiate");
        }
        ++g->instantiated;
    }
    const hipError_t e = hipGraphLaunch(g->exec, S(c, s));
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return GBH_OK;
}

int gbh_graph_stats(gbh_ctx*, const gbh_graph* g, uint64_t* instantiated, uint64_t* updated)
{
    *instantiated = g->instantiated;
    *updated = g->updated;
    return GBH_OK;
}

namespace {

int check_batch(const gbh_batch& b)
{
    if (!b.nseg || b.nseg > GBH_MAX_BATCH) return fail(GBH_E_ARG, "batch size outside [1, GBH_MAX_BATCH]");
    if (b.wbits < (1u << 12) || b.wbits > (1u << 20) || (b.wbits & (b.wbits - 1)))
        return fail(GBH_E_ARG, "window bits must be a power of two in [2^12, 2^20]");
    if (uint64_t(b.nwin) * (b.wbits / 64) + 16 > b.cw) return fail(GBH_E_ARG, "windows exceed the class stride");
    if (!b.geoms || !b.map) return fail(GBH_E_ARG, "batch without geometry or map");
    return GBH_OK;
}

}  // namespace

int gbh_reset_outputs(gbh_ctx* c, int s, gbh_segout* out, uint32_t nseg)
{
    if (!nseg) return GBH_OK;
    hipLaunchKernelGGL(gbh_dev::k_reset_out, dim3(64, nseg), dim3(256), 0, S(c, s), out);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_setup(gbh_ctx* c, int s, const gbh_setup_args* a)
{
    if (const int rc = check_batch(a->b)) return rc;
    if (!a->nprimes) return GBH_OK;
    hipLaunchKernelGGL(gbh_dev::k_setup, dim3(grid_of(a->nprimes, 256, 1u << 30), a->b.nseg), dim3(256), 0, S(c, s),
                       a->b, a->primes, a->nprimes, a->i0);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_harvest(gbh_ctx* c, int s, const gbh_harvest_args* a)
{
    if (const int rc = check_batch(a->b)) return rc;
    if (!a->nchunks) return GBH_OK;
    if (a->runs && (!a->nruns || !a->run_nseg_max || a->run_nseg_max > a->b.nseg))
        return fail(GBH_E_ARG, "harvest: run list without runs or with a run longer than the batch");
    const bool direct = a->rdesc != nullptr;
    if (direct && (!a->region || !a->fills || a->stage))
        return fail(GBH_E_ARG, "harvest: the direct form needs regions and fill counts and no stage");
    const uint64_t nb = uint64_t(a->runs ? a->run_nseg_max : a->b.nseg) * 8 * (direct ? 1 : a->b.nwin);
    if (nb > GBH_HL_BUCKETS) return fail(GBH_E_ARG, "harvest: more window lists than the work-group counts");
    if (a->m_base % 240) return fail(GBH_E_ARG, "harvest: bitmap base not a multiple of 240");
    if (!a->primes.base || !a->primes.gap || !a->primes.n)
        return fail(GBH_E_ARG, "harvest: no packed prime stream");
    if (a->stage > GBH_HL_STAGE) return fail(GBH_E_ARG, "harvest: stage above GBH_HL_STAGE");
    size_t lds = a->stage ? size_t(nb) * 8 + size_t(a->stage) * 6 + 16 : size_t(nb) * 4;
    uint32_t q_off = 0;
    const bool queues = direct && (a->few || a->hit_queue);
    if (queues) {                             // the per-wave queues after the counters
        q_off = uint32_t((lds + 15) & ~size_t(15));
        lds = q_off + std::max(a->few ? sizeof(gbh_dev::FewLds) : size_t(0),
                               a->hit_queue ? sizeof(gbh_dev::HitQLds) : size_t(0));
    }
    if (lds > 144 * 1024) return fail(GBH_E_ARG, "harvest: lists and stage exceed the work-group's LDS");
    const bool few = a->few && direct;
    const void* fn = few ? reinterpret_cast<const void*>(&gbh_dev::k_harvest<true, true>)
                   : direct ? reinterpret_cast<const void*>(&gbh_dev::k_harvest<true, false>)
                            : reinterpret_cast<const void*>(&gbh_dev::k_harvest<false, false>);
    uint32_t grid = a->nchunks;
    if (a->queue) {                           // persistent: the work-groups the GPU holds at once
        int per_cu = 0;
        GBH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds));
        if (a->queue_wg) per_cu = std::min(per_cu, int(a->queue_wg));
        grid = std::min<uint32_t>(a->nchunks, uint32_t(std::max(per_cu, 1)) * uint32_t(std::max(c->info.cus, 1)));
        if (!a->queue_zeroed) {
            hipLaunchKernelGGL(gbh_dev::k_fill, dim3(1), dim3(256), 0, S(c, s), a->queue, size_t(1), 0u);
            GBH_HIP(hipGetLastError());
        }
    }
#define GBH_HARVEST_ARGS                                                                                           \
    a->b, a->primes, a->chunks, a->runs, a->runs ? a->nruns : 1u, uint32_t(nb), a->t_hl, a->t_hi, a->t_hi_end,      \
        a->mbits, a->m_base, a->m_span, a->cap, a->stage, a->hits, a->hit_n, a->rdesc, a->region, a->fills, a->err, \
        a->nchunks, a->queue, q_off, a->hit_queue
    if (few)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(gbh_dev::k_harvest<true, true>), dim3(grid), dim3(256), lds, S(c, s),
                           GBH_HARVEST_ARGS);
    else if (direct)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(gbh_dev::k_harvest<true, false>), dim3(grid), dim3(256), lds, S(c, s),
                           GBH_HARVEST_ARGS);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(gbh_dev::k_harvest<false, false>), dim3(grid), dim3(256), lds, S(c, s),
                           GBH_HARVEST_ARGS);
#undef GBH_HARVEST_ARGS
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_partition(gbh_ctx* c, int s, const gbh_partition_args* a)
{
    if (const int rc = check_batch(a->b)) return rc;
    if (!a->ngroups) return GBH_OK;
    if (!a->rdesc || !a->region || !a->fills || !a->groups || !a->hits || !a->hit_n)
        return fail(GBH_E_ARG, "partition: null argument");
    if (a->b.nwin > 256) return fail(GBH_E_ARG, "partition: more than 256 windows per class");
    if (a->part_cap == GBH_PART_CAP_SMALL)
        hipLaunchKernelGGL(gbh_dev::k_partition<GBH_PART_CAP_SMALL>, dim3(a->b.nseg * 8, a->ngroups), dim3(512), 0,
                           S(c, s), a->rdesc, a->region, a->fills, a->groups, a->b.nseg * 8, a->b.nwin, a->b.wbits,
                           a->cap, a->hits, a->hit_n, a->err);
    else if (a->part_cap == GBH_PART_CAP)
        hipLaunchKernelGGL(gbh_dev::k_partition<GBH_PART_CAP>, dim3(a->b.nseg * 8, a->ngroups), dim3(512), 0, S(c, s),
                           a->rdesc, a->region, a->fills, a->groups, a->b.nseg * 8, a->b.nwin, a->b.wbits, a->cap,
                           a->hits, a->hit_n, a->err);
    else
        return fail(GBH_E_ARG, "partition: part_cap must be GBH_PART_CAP_SMALL or GBH_PART_CAP");
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_mtrans(gbh_ctx* c, int s, const gbh_mtrans_args* a)
{
    if (const int rc = check_batch(a->b)) return rc;
    if (a->nc == 0 || a->nc % 64 || uint64_t(a->nc) / 64 > a->b.cw) return fail(GBH_E_ARG, "mtrans: class array size");
    hipLaunchKernelGGL(gbh_dev::k_mtrans, dim3((a->nc / 64 + 255) / 256, a->b.nseg), dim3(256), 0, S(c, s), a->b,
                       a->nc, a->mbits, a->word0, a->mwords);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_window(gbh_ctx* c, int s, const gbh_window_args* a)
{
    if (const int rc = check_batch(a->b)) return rc;
    if (a->ncoop > a->nprimes) return fail(GBH_E_ARG, "ncoop > nprimes");
    const size_t lds = a->b.wbits / 8;
    const uint32_t phases = a->phases ? a->phases : 3u;
    if (phases > 3) return fail(GBH_E_ARG, "window: phases beyond 3");
    if (phases == 2 && !a->hits) return GBH_OK;    // the hit-list phase alone, without lists: nothing to do
    hipLaunchKernelGGL(gbh_dev::k_window, dim3(a->b.nwin, GBH_CLASSES, a->b.nseg), dim3(256), lds, S(c, s), a->b,
                       a->primes, a->nprimes, a->ncoop, a->i0, a->hits, a->hit_n, a->cap, phases);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_check_a(gbh_ctx* c, int s, const gbh_check_args* a)
{
    if (const int rc = check_batch(a->b)) return rc;
    if (a->ka > GBH_KA_MAX || a->ka % 16 || a->ka > g_min_useful)
        return fail(GBH_E_ARG, "ka must be a multiple of 16, at most GBH_KA_MAX and every class's useful candidates");
    // every class array holds nc <= nwin * wbits bits; the checker's words
    // stay below that (host contract: n_gpu >= q_first + 9780)
    const uint32_t words = uint32_t((uint64_t(a->b.nwin) * a->b.wbits + 63) / 64);
    const uint32_t cg = a->class_groups ? a->class_groups : 1u;
    if (GBH_EVEN_CLASSES % cg) return fail(GBH_E_ARG, "check: class_groups must divide the 15 even classes");
    if (!a->coff || !a->coffa || !a->pending || !a->out) return fail(GBH_E_ARG, "check: null argument");
    if (!(a->kinds & 4u)) {
        hipLaunchKernelGGL(gbh_dev::k_coff, dim3((GBH_NCAND_PAD + 255) / 256, GBH_EVEN_CLASSES, a->b.nseg),
                           dim3(256), 0, S(c, s), a->b, a->coff, a->coffa);
        GBH_HIP(hipGetLastError());
    }
    // sampled segments (histogram on) and the others: one launch each, every
    // work-group leaves at once on the other kind of segment
    const dim3 grid(grid_of(words, gbh_dev::kBlockWords, 1u << 30), a->b.nseg, cg);
    if (a->kinds & 1u) {
        hipLaunchKernelGGL(gbh_dev::k_check_a<false>, grid, dim3(GBH_CHECK_WG), 0, S(c, s), a->b, a->coffa, a->ka,
                           a->pending, a->pcap, a->out);
        GBH_HIP(hipGetLastError());
    }
    if (a->kinds & 2u) {
        hipLaunchKernelGGL(gbh_dev::k_check_a<true>, grid, dim3(GBH_CHECK_WG), 0, S(c, s), a->b, a->coffa, a->ka,
                           a->pending, a->pcap, a->out);
        GBH_HIP(hipGetLastError());
    }
    return GBH_OK;
}

int gbh_launch_check_b(gbh_ctx* c, int s, const gbh_checkb_args* a)
{
    if (const int rc = check_batch(a->b)) return rc;
    if (a->kend && a->kend <= a->ka) return fail(GBH_E_ARG, "check_b: kend must lie above ka (0: every slot)");
    hipLaunchKernelGGL(gbh_dev::k_check_b, dim3(grid_of(a->pcap, 256, 512), a->b.nseg), dim3(256), 0, S(c, s), a->b,
                       a->coff, a->ka, a->kend ? a->kend : 0xFFFFFFFFu, a->tail_t, a->pending, a->pcap, a->out,
                       a->ridge_more);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_mapcmp(gbh_ctx* c, int s, const gbh_mapcmp_args* a)
{
    if (!a->map || !a->ref || !a->geom || !a->diff) return fail(GBH_E_ARG, "mapcmp: null argument");
    hipLaunchKernelGGL(gbh_dev::k_mapcmp, dim3(256, GBH_CLASSES), dim3(256), 0, S(c, s), a->map, a->ref, a->geom,
                       a->cw, a->diff);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

}  // extern "C"

namespace {

// The counter and its gap walk across words, with the per-word scratch at
// `sc` (the caller's workspace or the library's static array).
int launch_count(gbh_ctx* c, int s, const gbh_count_args* a, uint32_t* sc)
{
    if (const int rc = check_batch(a->b)) return rc;
    if (a->swords < 2ull * GBH_COUNT_SUB * a->nwords) return fail(GBH_E_ARG, "count scratch too small");
    if (uint64_t(a->nwords) + 1 > a->b.cw) return fail(GBH_E_ARG, "count words exceed the class stride");
    if (!a->cs || !a->out) return fail(GBH_E_ARG, "count: null argument");
    if (!a->nwords) return GBH_OK;
    if (!sc) return fail(GBH_E_ARG, "count: null scratch");
    // about eight work-groups per CU over the batch, each over several word
    // pieces (one histogram flush per work-group, not per 256 pieces)
    const unsigned grid = std::max(1u, std::min(grid_of(uint64_t(a->nwords) * GBH_COUNT_SUB, 256, 1u << 30),
                                                (8u * uint32_t(std::max(c->info.cus, 1)) + a->b.nseg - 1) / a->b.nseg));
    hipLaunchKernelGGL(gbh_dev::k_count_a, dim3(grid, a->b.nseg), dim3(256), 0, S(c, s), a->b, a->cs, sc,
                       a->swords, a->out);
    GBH_HIP(hipGetLastError());
    hipLaunchKernelGGL(gbh_dev::k_count_b, dim3(grid, a->b.nseg), dim3(256), 0, S(c, s), a->cs, sc, a->swords,
                       a->out);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

}  // namespace

extern "C" {

int gbh_launch_count(gbh_ctx* c, int s, const gbh_count_args* a)
{
    return launch_count(c, s, a, a->scratch);
}

int gbh_launch_count_noalloc(gbh_ctx* c, int s, const gbh_count_args* a)
{
    if (a->scratch) return fail(GBH_E_ARG, "count_noalloc: scratch must be NULL (the library's array is used)");
    if (uint64_t(a->swords) * a->b.nseg > GBH_NOALLOC_SCRATCH_WORDS)
        return fail(GBH_E_ARG, "count_noalloc: nseg * swords above GBH_NOALLOC_SCRATCH_WORDS");
    return launch_count(c, s, a, c->scratch_dev);
}

}  // extern "C"
