// Multiplier bitmap build and cache (contracts in mbits.hpp). The build runs
// the map pipeline itself over m: consecutive ranges of 2^22 bits per class
// (2^22 x 30 multipliers), all eight class arrays aligned on one byte, are
// sieved by the window stage (masks 7..61, LDS marks of 67..sqrt(m_end)) and
// interleaved into bitmap words by k_mtrans.
#include "mbits.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "env.hpp"
#include "gbh.h"
#include "geometry.hpp"
#include "primes.hpp"

namespace gb {
namespace {

std::mutex g_mu;
// process lifetime, never destroyed: no device call runs at process exit
std::shared_ptr<MultBitmap>* g_cur = new std::shared_ptr<MultBitmap>();
gbh_ctx* g_ctx = nullptr;

void chk(int rc, const char* what)
{
    if (rc) throw std::runtime_error(std::string("multiplier bitmap: ") + what + ": " + gbh_last_error());
}

struct DevBuf {
    void* p = nullptr;
    int kind = 0;
    DevBuf(size_t n, int k) : kind(k) { chk(gbh_alloc(g_ctx, std::max<size_t>(n, 256), k, &p), "alloc"); }
    ~DevBuf() { if (p) (void)gbh_free(g_ctx, p, kind); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

uint64_t isqrt64(uint64_t x)
{
    uint64_t r = uint64_t(std::sqrt(double(x)));
    while (r * r > x) --r;
    while ((r + 1) * (r + 1) <= x) ++r;
    return r;
}

std::shared_ptr<MultBitmap> build(uint64_t m_lo, uint64_t m_hi, const std::vector<uint32_t>& base, bool verbose)
{
    const auto t0 = std::chrono::steady_clock::now();
    constexpr uint32_t kNc = 1u << 22, kWbits = 1u << 19, kNwin = kNc / kWbits, kBatch = 32;
    const uint64_t y0 = (m_lo / 30) & ~uint64_t(7);
    const uint64_t ny = m_hi / 30 + 1 - y0;
    const uint64_t nps = (ny + kNc - 1) / kNc;
    auto mb = std::make_shared<MultBitmap>();
    mb->words = nps * (kNc / 8);
    mb->m_base = 30 * y0;
    mb->m_span = 240 * mb->words;
    const uint64_t m_end = mb->m_base + mb->m_span;
    const uint64_t root = isqrt64(m_end - 1);
    if (base.empty() || base.back() < root)
        throw std::runtime_error("multiplier bitmap: base primes stop below sqrt(" + std::to_string(m_end) + ")");
    const size_t i_lo = size_t(std::lower_bound(base.begin(), base.end(), 67u) - base.begin());
    const size_t i_hi = size_t(std::upper_bound(base.begin(), base.end(), uint32_t(root)) - base.begin());
    const uint32_t np = uint32_t(i_hi > i_lo ? i_hi - i_lo : 0);
    uint32_t ncoop = 0;
    while (ncoop < np && base[i_lo + ncoop] < kWbits / 128) ++ncoop;
    const uint32_t cw = kNc / 64 + 16;

    void* bits = nullptr;
    chk(gbh_alloc(g_ctx, mb->words * 8, 0, &bits), "alloc bitmap");
    mb->bits = static_cast<const uint64_t*>(bits);
    mb->owner = g_ctx;
    DevBuf d_primes(size_t(np) * 4, 0), d_i0(size_t(kBatch) * 8 * std::max(np, 1u) * 4, 0),
        d_map(size_t(kBatch) * 8 * cw * 8, 0), d_geoms(sizeof(gbh_geom) * kBatch, 0),
        h_geoms(sizeof(gbh_geom) * kBatch, 1);
    if (np) chk(gbh_copy_h2d(g_ctx, 0, d_primes.p, base.data() + i_lo, size_t(np) * 4), "upload primes");
    chk(gbh_stream_sync(g_ctx, 0), "upload sync");
    auto* hg = static_cast<gbh_geom*>(h_geoms.p);
    for (uint64_t k0 = 0; k0 < nps; k0 += kBatch) {
        const uint32_t ks = uint32_t(std::min<uint64_t>(kBatch, nps - k0));
        for (uint32_t s = 0; s < ks; ++s) {
            const uint64_t yk = y0 + (k0 + s) * kNc;
            SegmentGeometry sg;
            sg.q_first = 30 * yk + 1;
            sg.q_last = 30 * (yk + kNc) - 1;
            sg.q_odds = 15ull * kNc;
            sg.seg_start = sg.q_first;
            sg.seg_end = sg.q_last;
            const ClassGeo cg = class_geometry(sg.q_first, sg.q_odds);
            for (int c = 0; c < 8; ++c)
                if (cg.n[c] != kNc || cg.Q[c] != 30 * yk + kRes[c])
                    throw std::runtime_error("multiplier bitmap: class arrays not aligned");
            hg[s] = make_geom(sg, cg, kWbits);
        }
        chk(gbh_copy_h2d(g_ctx, 0, d_geoms.p, hg, sizeof(gbh_geom) * ks), "upload geometry");
        const gbh_batch b{static_cast<const gbh_geom*>(d_geoms.p), ks, cw, kWbits, kNwin,
                          static_cast<uint64_t*>(d_map.p)};
        if (np) {
            const gbh_setup_args a{b, static_cast<const uint32_t*>(d_primes.p), np, 0, static_cast<uint32_t*>(d_i0.p)};
            chk(gbh_launch_setup(g_ctx, 0, &a), "setup");
        }
        const gbh_window_args w{b, static_cast<const uint32_t*>(d_primes.p), np, ncoop,
                                static_cast<const uint32_t*>(d_i0.p), nullptr, nullptr, 0, 0};
        chk(gbh_launch_window(g_ctx, 0, &w), "window sieve");
        const gbh_mtrans_args t{b, kNc, 0, static_cast<uint64_t*>(bits), k0 * (kNc / 8), mb->words};
        chk(gbh_launch_mtrans(g_ctx, 0, &t), "interleave");
        chk(gbh_stream_sync(g_ctx, 0), "batch sync");   // the pinned geometry is rewritten next
    }
    mb->build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (verbose)
        std::fprintf(stderr, "multiplier bitmap: m in [%llu, %llu), %.2f GiB, %u sieving primes, %.3f s\n",
                     (unsigned long long)mb->m_base, (unsigned long long)m_end, double(mb->words * 8) / (1u << 30), np,
                     mb->build_s);
    return mb;
}

}  // namespace

MultBitmap::~MultBitmap()
{
    if (bits) (void)gbh_free(static_cast<gbh_ctx*>(owner), const_cast<uint64_t*>(bits), 0);
}

uint64_t mult_bitmap_budget() { return env::u64("GB_MT_GB", 80, 0, 1024) << 30; }

std::shared_ptr<const MultBitmap> mult_bitmap(uint64_t m_lo, uint64_t m_hi, const std::vector<uint32_t>& base,
                                              bool verbose)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (m_lo <= 61 || m_hi < m_lo) throw std::runtime_error("multiplier bitmap: range");
    std::shared_ptr<MultBitmap>& cur = *g_cur;
    if (cur && cur->covers(m_lo, m_hi)) return cur;
    const uint64_t budget = mult_bitmap_budget();
    const uint64_t root = base.empty() ? 0 : base.back();
    const auto fits = [&](uint64_t lo, uint64_t hi) {
        return (hi - lo) / 30 + (uint64_t(1) << 23) <= budget && root * root > hi + (uint64_t(30) << 22);
    };
    if (!fits(m_lo, m_hi))
        throw std::runtime_error("multiplier bitmap: m range [" + std::to_string(m_lo) + ", " + std::to_string(m_hi) +
                                 "] exceeds the GB_MT_GB budget or the base primes");
    if (!g_ctx) chk(gbh_open(0, &g_ctx), "gbh_open");
    // a small range next to a large cached one: a bitmap of its own, not cached
    if (cur && (m_hi - m_lo) / 30 < (uint64_t(1) << 30) && cur->m_span / 30 > (uint64_t(4) << 30))
        return build(m_lo, m_hi, base, verbose);
    uint64_t lo = m_lo, hi = m_hi;
    if (cur) {
        const uint64_t lo2 = std::min(lo, cur->m_base), hi2 = std::max(hi, cur->m_base + cur->m_span - 1);
        if (fits(lo2, hi2)) { lo = lo2; hi = hi2; }
    }
    cur.reset();   // its memory goes before the new one is allocated (unless a run still holds it)
    cur = build(lo, hi, base, verbose);
    return cur;
}

std::shared_ptr<const MultBitmap> mult_bitmap_peek()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return *g_cur;
}

void mult_bitmap_release()
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_cur->reset();
}

}  // namespace gb
