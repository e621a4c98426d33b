// C ABI of the device pipeline (include/gbh.h): context, memory, streams,
// events and one launch function per stage. Replaces the reference's Level
// Zero layer (src/l0_context.cpp:75-356) and the positional kernel argument
// lists of src/verifier.cpp; the kernels themselves are in gbh_kernels.hip,
// compiled into this translation unit so the library is one code object.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gbh.h"
#include "gbh_kernels.hip"

struct gbh_ctx {
    int device = 0;
    hipStream_t stream[2] = {nullptr, nullptr};
    gbh_device_info_t info{};
};
struct gbh_event {
    hipEvent_t ev = nullptr;
};

namespace {

thread_local std::string g_err;

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

hipStream_t S(gbh_ctx* c, int s) { return c->stream[s & 1]; }

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
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&gbh_dev::k_window),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) (void)hipGetLastError();   // optional: windows up to the default limit still work
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

int gbh_launch_setup(gbh_ctx* c, int s, const gbh_setup_args* a)
{
    if (!a->nprimes) return GBH_OK;
    hipLaunchKernelGGL(gbh_dev::k_setup, dim3(grid_of(a->nprimes, 256, 1u << 30)), dim3(256), 0, S(c, s),
                       *a->geom, a->primes, a->nprimes, a->i0);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_window(gbh_ctx* c, int s, const gbh_window_args* a)
{
    const gbh_geom& g = *a->geom;
    if (g.wbits < (1u << 12) || g.wbits > (1u << 20) || (g.wbits & (g.wbits - 1)))
        return fail(GBH_E_ARG, "window bits must be a power of two in [2^12, 2^20]");
    if (uint64_t(g.nwin) * (g.wbits / 64) > g.cw) return fail(GBH_E_ARG, "windows exceed the class stride");
    for (int k = 0; k < GBH_CLASSES; ++k)
        if (g.nc[k] > uint64_t(g.nwin) * g.wbits) return fail(GBH_E_ARG, "class array longer than its windows");
    if (a->ncoop > a->nprimes) return fail(GBH_E_ARG, "ncoop > nprimes");
    const size_t lds = g.wbits / 8;
    hipLaunchKernelGGL(gbh_dev::k_window, dim3(g.nwin, GBH_CLASSES), dim3(256), lds, S(c, s), g, a->primes,
                       a->nprimes, a->ncoop, a->i0, a->map);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_large(gbh_ctx* c, int s, const gbh_large_args* a)
{
    if (!a->nprimes) return GBH_OK;
    hipLaunchKernelGGL(gbh_dev::k_large, dim3(grid_of(a->nprimes, 256, 65536)), dim3(256), 0, S(c, s), *a->geom,
                       a->primes, a->nprimes, reinterpret_cast<unsigned long long*>(a->map));
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

namespace {
uint32_t max_class_evens(const gbh_geom& g, uint64_t n_gpu)
{
    uint32_t m = 0;
    for (uint32_t cr = 0; cr < GBH_EVEN_CLASSES; ++cr) {
        const uint64_t n0 = n_gpu + (2 * cr + 30 - n_gpu % 30) % 30;
        const uint64_t cnt = n0 <= g.seg_end ? (g.seg_end - n0) / 30 + 1 : 0;
        if (cnt > m) m = uint32_t(cnt);
    }
    return m;
}
}  // namespace

int gbh_launch_check_a(gbh_ctx* c, int s, const gbh_check_args* a)
{
    const gbh_geom& g = *a->geom;
    if (a->ka > gbh_dev::kKaMax) return fail(GBH_E_ARG, "ka above 128");
    if (a->n_gpu < g.q_first + 9780 || a->n_gpu < g.seg_start || (a->n_gpu & 1))
        return fail(GBH_E_ARG, "n_gpu below the first checkable even");
    hipLaunchKernelGGL(gbh_dev::k_coff, dim3((GBH_NCAND + 255) / 256, GBH_EVEN_CLASSES), dim3(256), 0, S(c, s), g,
                       a->n_gpu, a->coff);
    GBH_HIP(hipGetLastError());
    if (a->n_gpu > g.seg_end) return GBH_OK;
    const uint32_t words = (max_class_evens(g, a->n_gpu) + 63) / 64;
    // every tested bit (k + offset < class bits) lies inside the class stride;
    // the staging loads past it read zeros
    if (uint64_t(words) > g.cw) return fail(GBH_E_ARG, "class stride too small for the checker");
    hipLaunchKernelGGL(gbh_dev::k_check_a, dim3(grid_of(words, GBH_CHECK_WG, 1u << 30)), dim3(GBH_CHECK_WG), 0,
                       S(c, s), g, a->map, a->coff, a->n_gpu, a->ka, a->hist_on, a->pending, a->pending_cap, a->out);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_check_b(gbh_ctx* c, int s, const gbh_checkb_args* a)
{
    const gbh_geom& g = *a->geom;
    if (a->n_gpu > g.seg_end) return GBH_OK;
    hipLaunchKernelGGL(gbh_dev::k_check_b, dim3(grid_of(a->pending_max, 256, 2048)), dim3(256), 0, S(c, s), g,
                       a->map, a->coff, a->n_gpu, a->ka, a->tail_t, a->pending, a->pending_max, a->out);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

int gbh_launch_count(gbh_ctx* c, int s, const gbh_count_args* a)
{
    const gbh_geom& g = *a->geom;
    if (a->scratch_words < 2ull * a->nwords) return fail(GBH_E_ARG, "count scratch too small");
    if (uint64_t(a->nwords) + 1 > g.cw) return fail(GBH_E_ARG, "count words exceed the class stride");
    if (!a->nwords) return GBH_OK;
    gbh_dev::CountArgs k{};
    for (int i = 0; i < 8; ++i) { k.cls[i] = a->cls[i]; k.pos0[i] = a->pos0[i]; k.lo[i] = a->lo[i]; k.hi[i] = a->hi[i]; }
    for (int j = 0; j < 3; ++j) { k.tk[j] = a->tk[j]; k.tk2[j] = a->tk2[j]; k.td[j] = a->td[j]; k.tlo[j] = a->tlo[j]; k.thi[j] = a->thi[j]; }
    k.nwords = a->nwords;
    const unsigned grid = grid_of(a->nwords, 256, 1u << 30);
    hipLaunchKernelGGL(gbh_dev::k_count_a, dim3(grid), dim3(256), 0, S(c, s), g, a->map, k, a->scratch, a->out);
    GBH_HIP(hipGetLastError());
    hipLaunchKernelGGL(gbh_dev::k_count_b, dim3(grid), dim3(256), 0, S(c, s), k, a->scratch, a->out);
    GBH_HIP(hipGetLastError());
    return GBH_OK;
}

}  // extern "C"
