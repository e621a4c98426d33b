/* gbh.h — device-level C ABI of the MI355X (gfx950) segment pipeline.
 *
 * This layer replaces the reference's GPU runtime boundary and kernel ABI:
 * the Level Zero wrapper include/l0_context.hpp:14-215 (Context, Module,
 * Kernel::set_arg/launch, alloc_device/alloc_shared, host_sync, timestamp
 * events) together with the positional kernel argument lists that
 * src/verifier.cpp sets in enqueue_sieve / enqueue_large_primes /
 * enqueue_check (verifier.cpp:1808-2330). Every stage launch takes one POD
 * argument struct instead of set_arg calls; all pointers are device
 * pointers obtained from gbh_alloc, sizes are in bytes unless stated.
 *
 * Work is issued per BATCH of consecutive segments (up to a few dozen):
 * every stage covers all segments of the batch in one launch, the batch's
 * class maps being resident in HBM side by side. (The reference launches
 * every stage once per segment; its hit-list harvest alone spans a batch.)
 *
 * Error convention: every function returns 0 on success and a non-zero
 * GBH_E_* code on failure; gbh_last_error() then holds the message (the
 * ZE_CHECK text of l0_context.hpp:14-20 becomes this string). Callers that
 * want the reference's fail-closed behaviour turn a non-zero status into an
 * exception (the C++ driver does).
 *
 * Threading: one driver thread per gbh_ctx (as the reference's single
 * immediate command list). Streams are in-order HIP streams at the default
 * priority; a context holds GBH_STREAMS of them, named 0 .. GBH_STREAMS - 1
 * in every `stream` argument (the C++ driver runs a batch's main chain on
 * stream k and its window sieve on stream 2 + k, k the batch's slot).
 */
#ifndef GBH_H
#define GBH_H

#include <stddef.h>
#include <stdint.h>

#define GBH_STREAMS 4

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GBH_OK = 0,
    GBH_E_HIP = 1,        /* a HIP runtime call failed */
    GBH_E_ARG = 2,        /* an argument violates the stage's contract */
    GBH_E_NODEV = 3,      /* no usable gfx950 device */
    GBH_E_NOMEM = 4
};

/* Layout constants shared by host and kernels. */
#define GBH_CLASSES 8            /* residues 1 7 11 13 17 19 23 29 mod 30 */
#define GBH_EVEN_CLASSES 15      /* n mod 30 = 0, 2, ..., 28 */
#define GBH_NCAND 1205           /* odd primes 3..9781 (MAX_GB_PRIME, config.hpp:86) */
#define GBH_NCAND_PAD 1208       /* candidate rows padded to a multiple of 4 */
#define GBH_NSMALL 15            /* mask-stage primes 7..61 (MASK_MAX_P = 64, config.hpp:141) */
#define GBH_GAP_BINS 1024        /* gap histogram bins g / 2 (GAPS30_HBINS, gstats.hpp:75) */
#define GBH_GAP_TAIL 700         /* tail gaps (GAPS30_TAIL, gstats.hpp:74) */
#define GBH_GAP_TCAP 4096        /* tail entries per segment (GAPS30_TCAP, gstats.hpp:114) */
#define GBH_LEFT_CAP 65536       /* leftovers per segment copied back */
#define GBH_CHECK_WG 256         /* check_a work-group: 256 lanes x 64 evens */
#define GBH_KA_MAX 128           /* phase-A candidates per even class, at most */
#define GBH_RIDGE_CAP 1024       /* ridge records per segment kept in the output block */
#define GBH_RIDGE_MORE 64512     /* further ridge records per segment in gbh_checkb_args.ridge_more: 65536 in
                                    all, the reference's RIDGE_CAP (config.hpp:103) */
#define GBH_MAX_BATCH 64         /* segments per batch, at most */

typedef struct gbh_ctx gbh_ctx;
typedef struct gbh_event gbh_event;

/* Per-segment geometry of the mod-30 class map (host_math.hpp:53-107
 * SegmentGeometry / ClassGeo): bit t of class array c <-> q = Q[c] + 30 t. */
typedef struct {
    uint64_t seg_start, seg_end;   /* first / last even of the segment */
    uint64_t q_first, q_last;      /* odd q range sieved */
    uint64_t Q[GBH_CLASSES];       /* first q of each class array */
    uint64_t n_gpu;                /* first even the checker takes (smaller ones: CPU) */
    uint32_t nc[GBH_CLASSES];      /* bits of each class array */
    uint32_t nwin;                 /* windows of this segment's class arrays */
    uint32_t hist_on;              /* record p(n) of every even (a sampled segment) */
    uint32_t np_win;               /* window primes [0, np_win) can divide a q of the segment (the rest are
                                      above sqrt(q_last): skipped) */
    uint32_t ridge_t;              /* phase B records (n, p(n)) for p(n) >= ridge_t (~0u: none) */
    uint64_t restore[GBH_CLASSES]; /* bits of word 0 set again after the mask stage (q = p <= 61) */
    uint32_t s0[GBH_CLASSES][16];  /* mask primes: class bit of every multiple, residue mod p */
} gbh_geom;

/* Counter parameters of one segment (count30_limits, gstats.hpp:262-296):
 * the eight class arrays in ascending phase (position order inside a block
 * of 30), each with its class index, the position of bit 0 (q - q_first)
 * and the counted bits [lo, hi); the three twin pairs (11,13) (17,19)
 * (29,31) with the lower / partner class, the partner's bit offset (0 or
 * 1) and the counted bits of the lower class. */
typedef struct {
    uint32_t cls[GBH_CLASSES], pos0[GBH_CLASSES], lo[GBH_CLASSES], hi[GBH_CLASSES];
    uint32_t tk[3], tk2[3], td[3], tlo[3], thi[3];
    uint32_t nwords;               /* words walked per class array */
    uint32_t on;                   /* 0: no counting for this segment */
} gbh_countseg;

/* A batch: nseg segments, their class maps side by side (map + s * 8 * cw +
 * c * cw), one LDS window size. The segments are consecutive segments of one
 * range, or (batched ranges) runs of consecutive segments of several ranges
 * (gbh_hl_run). */
typedef struct {
    const gbh_geom* geoms;         /* device array [nseg] */
    uint32_t nseg;
    uint32_t cw;                   /* class array stride, 64-bit words (>= nwin * wbits / 64 + 16) */
    uint32_t wbits;                /* LDS window bits (power of two, 2^12..2^20) */
    uint32_t nwin;                 /* windows per class array, the batch's maximum */
    uint64_t* map;
} gbh_batch;

/* Per-segment outputs of the checker and the counter (reference: the slot's
 * shared USM counters, verifier.cpp:42-87). The leftover list comes last:
 * a caller copies back the first GBH_SEGOUT_HEAD bytes of each block and
 * the list only when left_n > 0 (it is almost always empty). */
typedef struct {
    uint32_t acct[2];              /* evens resolved by phase A / phase B */
    uint32_t pending_n;            /* evens phase A left */
    uint32_t left_n;               /* evens no candidate resolved (CPU fallback) */
    uint32_t err;                  /* error flags (1: hit-list overflow) */
    uint32_t cnt[GBH_CLASSES];     /* primes per class array in the counting range */
    uint32_t twins[3];             /* twin pairs (11,13) (17,19) (29,31) */
    uint32_t pfirst, plast;        /* first / last counted prime: position q - q_first (~0u / 0: none) */
    uint32_t gtail_n;              /* tail gaps (g >= 700) inside the segment */
    uint32_t vdiff;                /* map words that differ from the verification map (gbh_launch_mapcmp) */
    uint64_t gmax;                 /* largest inner gap: g << 32 | ~(upper position) (0: none) */
    uint32_t ghist[GBH_GAP_BINS + 1];   /* inner gaps by g / 2; [GBH_GAP_BINS]: overflow */
    uint32_t pad1;
    uint32_t gtail[2 * GBH_GAP_TCAP];   /* (upper position, g) */
    uint32_t hist[GBH_EVEN_CLASSES * GBH_NCAND];   /* p(n) by even class and candidate slot */
    uint32_t first[GBH_NCAND];     /* smallest even index with p(n) = candidate (by prime index; ~0u none) */
    uint32_t ridge_n;              /* phase-B evens with p(n) >= the ridge threshold */
    uint32_t pad2;
    uint64_t ridge[2 * GBH_RIDGE_CAP];   /* (n, p(n)) of the first GBH_RIDGE_CAP of them */
    uint32_t left[GBH_LEFT_CAP];   /* even indices of the leftovers */
} gbh_segout;
#define GBH_SEGOUT_HEAD offsetof(gbh_segout, left)

/* Stage 1: first class bit of every window prime in every segment
 * (reference mid_setup, goldbach.cl:3528-3563, and the host's
 * first_bits_class_range, host_math.cpp:80-113). */
typedef struct {
    gbh_batch b;
    const uint32_t* primes;        /* window primes (device), ascending, all >= 67 */
    uint32_t nprimes;
    uint32_t pad0;
    uint32_t* i0;                  /* [nseg][8][nprimes] first bit, ~0u: none in the segment */
} gbh_setup_args;

/* Stage 2: hit lists of the large primes for the whole batch (reference
 * hl_harvest_mt, goldbach.cl:3251-3426, with sieve_large_packed's wheel,
 * 480-748, for the primes below t_hl), filed by (segment, class, window)
 * into the lists the window sieve applies. A prime p >= t_hl lists the
 * multiples p m with m >= p prime, read from the multiplier bitmap; a prime
 * below t_hl lists every multiple p m with m >= p coprime to 2 3 5 7 11 13.
 * Exact when every q of the batch is below t_hl^3 (device guard: err bit 2)
 * and the bitmap covers every multiplier asked for (err bit 3). Work-group g
 * takes the primes [chunks[g], chunks[g + 1]). The batch's q range must stay
 * below 2^33, and nseg * 8 * nwin <= GBH_HL_BUCKETS. */
#define GBH_HL_BUCKETS 12288
/* The large primes as a packed stream, one byte per prime (the reference's
 * pack_large_primes, src/host_math.cpp:186-209, decoded on the device like
 * its decode_large, kernels/goldbach.cl:1678-1716; here in groups of one
 * wave): prime j = base[j / 64] + 2 * (gap[64 (j / 64) + 1] + ... + gap[j]),
 * gap[j] = half the gap below prime j for j not a multiple of 64 and 0 at
 * every group's first prime. Both arrays are padded past the last prime by
 * GBH_PS_PAD entries (gaps 0, bases the last prime), so a wave may decode a
 * whole group of indices beyond n; base[g] for every group g holding a prime
 * is at least every prime of group g - 1. The harvest reads only this stream
 * and takes every chunk of primes from a multiple of 64 (err bit 0 else). */
#define GBH_PS_GROUP 64
#define GBH_PS_PAD 4096
typedef struct {
    const uint32_t* base;          /* [ceil((n + GBH_PS_PAD) / 64)] */
    const uint8_t* gap;            /* [64 * ceil((n + GBH_PS_PAD) / 64)] */
    uint32_t n;                    /* primes in the stream */
    uint32_t pad0;
} gbh_pstream;
#define GBH_HL_CHUNK 65536     /* primes per harvest work-group, at most */
#define GBH_HL_STAGE 4096      /* staged entries per harvest work-group, at most */
/* One run of a batch: segments [seg0, seg0 + nseg), consecutive segments of
 * one range (their q ranges overlap or touch), served by the large primes
 * [0, nprimes) of the list (those up to the range's sqrt bound). A batch of
 * several ranges has one run per range; every work-group walks its chunk's
 * primes once per run. nseg * 8 * nwin <= GBH_HL_BUCKETS and the run's q
 * range below 2^33. */
typedef struct {
    uint32_t seg0, nseg;
    uint32_t nprimes;
    uint32_t pad0;
} gbh_hl_run;
typedef struct {
    gbh_batch b;
    gbh_pstream primes;            /* large primes (device), ascending, packed */
    const uint32_t* chunks;        /* [nchunks + 1] prime index boundaries, each chunk's first a multiple of 64 */
    const gbh_hl_run* runs;        /* device array [nruns]; nullptr: one run of the whole batch, every prime */
    uint32_t nruns;
    uint32_t run_nseg_max;         /* the largest run's nseg (sizes the work-group's counters) */
    uint32_t nchunks;
    uint32_t cap;                  /* entries per window list */
    uint32_t t_hl;                 /* first prime served by the bitmap */
    uint32_t t_hi, t_hi_end;       /* primes in [t_hi, t_hi_end) walk the wheel too (UINT32_MAX: none);
                                      chunks never straddle t_hl, t_hi or t_hi_end */
    uint32_t stage;                /* LDS entries a work-group may stage (0: always two walks), <= GBH_HL_STAGE */
    const uint64_t* mbits;         /* multiplier bitmap (gbh_launch_mtrans layout) */
    uint64_t m_base;               /* a multiple of 240 */
    uint64_t m_span;               /* the bitmap covers m in [m_base, m_base + m_span) */
    uint32_t* hits;                /* [nseg][8][nwin][cap] window offsets */
    uint32_t* hit_n;               /* [nseg][8][nwin] fill counts (zeroed before the launch) */
    /* Direct (one-pass) form, when rdesc is not null: work-group k appends
     * the class bit of every hit to its own region of each (segment, class)
     * list, region + rdesc[2k] + (s * 8 + c) * rdesc[2k + 1] (rdesc[2k + 1]
     * entries each), and writes the counts to fills[k * nseg * 8 + s * 8 + c]
     * (every entry written, 0 for a run without primes in the chunk: no
     * zeroing needed); gbh_launch_partition then builds the window lists
     * (each count taken at most rdesc[2k + 1]). hits / hit_n / cap / stage
     * are unused then (stage 0). */
    const uint32_t* rdesc;
    uint32_t* region;
    uint32_t* fills;
    uint32_t* err;                 /* bit 0: a list or region overflow or a run beyond the contract;
                                      bit 2: t_hl^3 <= q; bit 3: m outside the bitmap */
    /* Persistent form, when queue is not null: a device word (cleared by the
     * launch, in stream order, unless queue_zeroed) from which as many
     * work-groups as the GPU holds at once take chunk indices until all are
     * done; null: one work-group per chunk. Either way chunk k files into
     * region k and fills row k. */
    uint32_t* queue;
    uint32_t queue_wg;             /* persistent form: work-groups per CU at most (0: as many as fit) */
    /* Nonzero: the batch has chunks of bitmap primes with at most four
     * multipliers in their run's q span; the launch then reserves the LDS
     * queues of the compacted path for them (zero: such chunks take the
     * bitmap walk). */
    uint32_t few;
    /* Nonzero: *queue is already zero in stream order (the caller cleared it
     * after the queue's previous launch), so the launch does not clear it. */
    uint32_t queue_zeroed;
    /* Nonzero (direct form with few): the bitmap walk of primes with several
     * multipliers queues each wave's hits in LDS and files them with every
     * lane at once (10 KB more LDS per work-group). */
    uint32_t hit_queue;
} gbh_harvest_args;

/* Second step of the direct harvest: the regions of every chunk split into
 * the window lists hits / hit_n of gbh_harvest_args (hit_n zeroed before).
 * The chunks go in groups [groups[g], groups[g + 1]) of at most
 * GBH_PART_CHUNKS chunks whose regions (rdesc[2k + 1] entries per list)
 * hold at most part_cap entries together: GBH_PART_CAP_SMALL (32 KB of LDS
 * per work-group, more of them per CU) or GBH_PART_CAP. */
#define GBH_PART_CHUNKS 64
#define GBH_PART_CAP_SMALL 8192
#define GBH_PART_CAP 16384
typedef struct {
    gbh_batch b;
    const uint32_t* rdesc;
    const uint32_t* region;
    const uint32_t* fills;
    const uint32_t* groups;        /* [ngroups + 1] chunk boundaries */
    uint32_t ngroups;
    uint32_t cap;                  /* entries per window list */
    uint32_t* hits;
    uint32_t* hit_n;
    uint32_t* err;                 /* bit 0: a window list overflow */
    uint32_t part_cap;             /* GBH_PART_CAP_SMALL or GBH_PART_CAP */
    uint32_t pad0;
} gbh_partition_args;

/* Map self-check (reference GB_BUCKET_VERIFY compare kernel,
 * src/verifier.cpp:2059-2085): counts the words of one segment's class map
 * (bits below nc[c] of each class) that differ from a second map of the
 * same segment built another way, into *diff. */
typedef struct {
    const uint64_t* map;           /* the production map of the segment */
    const uint64_t* ref;           /* the verification map */
    const gbh_geom* geom;          /* the segment's geometry (device) */
    uint32_t cw;                   /* class stride (words) of both maps */
    uint32_t pad0;
    uint32_t* diff;
} gbh_mapcmp_args;

/* Multiplier bitmap build (reference mt_build, goldbach.cl:3181-3242): the
 * class maps of a batch sieved over m (segment s: class arrays of nc bits
 * starting at the common byte y0 + s nc, Q[c] = 30 (y0 + s nc) + R[c]) are
 * interleaved into bitmap words: word w, bit 8 j + i <-> m = m_base + 240 w +
 * 30 j + R[i], R the unit residues mod 30 ascending; bit set iff m has no
 * prime factor among the sieving primes. Writes words [word0, word0 +
 * nseg nc / 8) below mwords. */
typedef struct {
    gbh_batch b;
    uint32_t nc;                   /* bits per class array, a multiple of 64 */
    uint32_t pad0;
    uint64_t* mbits;
    uint64_t word0;
    uint64_t mwords;
} gbh_mtrans_args;

/* Stage 3: LDS window sieve of every class array of every segment
 * (reference make_prime_map + mid_window + hl_apply, goldbach.cl:1092-1608,
 * 3434-3457, 3600-3701): masks for 7..61, marks of the window primes, the
 * window's hit list, one coalesced store. phases: bit 0 the masks and window
 * primes (the window starts from the masks), bit 1 the hit lists (without
 * bit 0 the window starts from the map as stored); 0 means both. The sieve
 * phase does not read the hit lists, so it can run beside the harvest. */
typedef struct {
    gbh_batch b;
    const uint32_t* primes;        /* window primes (device) */
    uint32_t nprimes;
    uint32_t ncoop;                /* primes [0, ncoop) are marked by the whole work-group */
    const uint32_t* i0;            /* stage-1 table */
    const uint32_t* hits;          /* stage-2 lists (nullptr: no large primes) */
    const uint32_t* hit_n;
    uint32_t cap;
    uint32_t phases;               /* 1: sieve, 2: hit lists, 0 or 3: both */
} gbh_window_args;

/* Stage 4: checker phase A (reference check30_a, goldbach.cl:2028-2189):
 * exact ascending candidates, 64 evens per lane, the first ka candidates of
 * every even class; the p(n) histogram on segments with hist_on. */
typedef struct {
    gbh_batch b;
    uint32_t ka;                   /* phase-A candidates per even class (a multiple of 16, <= GBH_KA_MAX) */
    uint32_t pcap;                 /* pending entries per segment */
    uint32_t* coff;                /* [nseg][15][GBH_NCAND] per-segment candidate table (written here) */
    uint32_t* coffa;               /* [nseg][15][GBH_NCAND_PAD] the same as staged-word offsets (written here) */
    uint32_t* pending;             /* [nseg][pcap] even indices left for phase B */
    gbh_segout* out;               /* [nseg] */
    uint32_t kinds;                /* bit 0: the batch has segments without histogram, bit 1: with;
                                      bit 2: the candidate tables coff / coffa are already built
                                      for this batch (an earlier call), skip k_coff */
    uint32_t class_groups;         /* 0 or 1: a work-group checks all 15 even classes of its words;
                                      3 or 5: the classes split over that many work-groups (more
                                      waves for short batches) */
} gbh_check_args;

/* Stage 5: phase B (reference check30_b, goldbach.cl:2210-2306): the
 * remaining candidates of every pending even, ascending; the first hit is
 * p(n). Evens with p(n) >= the segment's ridge_t are recorded (Goldbach
 * ridges). */
typedef struct {
    gbh_batch b;
    uint32_t ka;                   /* first candidate slot to probe */
    uint32_t tail_t;               /* first occurrences are tracked for p >= tail_t */
    const uint32_t* coff;
    const uint32_t* pending;
    uint32_t pcap;
    uint32_t kend;                 /* probe candidate slots [ka, kend) only (0: every slot); an even whose p(n)
                                      lies beyond is a leftover for the CPU (a test of the fallback) */
    gbh_segout* out;
    uint64_t* ridge_more;          /* [nseg][GBH_RIDGE_MORE][2] (n, p(n)) of the ridge records beyond the first
                                      GBH_RIDGE_CAP (nullptr: not kept; ridge_n still counts them) */
} gbh_checkb_args;

/* Stage 6: prime / twin counts and the gap walk of the counting range
 * (reference count30 + gap_tail_exact, goldbach.cl:2499-2825). The walk of
 * each map word position runs as GBH_COUNT_SUB pieces of 64 / GBH_COUNT_SUB
 * blocks of 30 (one lane each; the gaps across pieces come from their first
 * and last prime positions, kept in the scratch). */
#define GBH_COUNT_SUB 2
typedef struct {
    gbh_batch b;
    const gbh_countseg* cs;        /* device array [nseg] */
    uint32_t nwords;               /* the batch's largest nwords */
    uint32_t swords;               /* scratch u32 per segment (>= 2 * GBH_COUNT_SUB * nwords) */
    uint32_t* scratch;             /* [nseg][swords] per-piece first / last prime positions */
    gbh_segout* out;
} gbh_count_args;

typedef struct {
    char name[128];
    char arch[64];
    int cus;
    int xcds_hint;
    size_t lds_per_block;
    size_t hbm_bytes;
    int clock_khz;
    int mem_clock_khz;
} gbh_device_info_t;

/* ---- context, memory, streams, events ---- */
int gbh_open(int device, gbh_ctx** out);
void gbh_close(gbh_ctx* ctx);
const char* gbh_last_error(void);
int gbh_device_info(gbh_ctx* ctx, gbh_device_info_t* info);
/* kind: 0 device, 1 pinned host (device-visible) */
int gbh_alloc(gbh_ctx* ctx, size_t bytes, int kind, void** out);
int gbh_free(gbh_ctx* ctx, void* ptr, int kind);
int gbh_copy_h2d(gbh_ctx* ctx, int stream, void* dst, const void* src, size_t bytes);
int gbh_copy_d2h(gbh_ctx* ctx, int stream, void* dst, const void* src, size_t bytes);
/* rows pieces of width bytes, at src + k * spitch, to dst + k * dpitch */
int gbh_copy_d2h_rows(gbh_ctx* ctx, int stream, void* dst, size_t dpitch, const void* src, size_t spitch,
                      size_t width, size_t rows);
int gbh_zero(gbh_ctx* ctx, int stream, void* dst, size_t bytes);  /* fill kernel (graph-safe) */
int gbh_fill32(gbh_ctx* ctx, int stream, void* dst, size_t count, uint32_t value);
/* out[s] of every segment of a batch reset for a new run: zeros, with
 * first[] and pfirst set to ~0u */
int gbh_reset_outputs(gbh_ctx* ctx, int stream, gbh_segout* out, uint32_t nseg);
int gbh_stream_sync(gbh_ctx* ctx, int stream);
int gbh_event_create(gbh_ctx* ctx, gbh_event** ev);
int gbh_event_destroy(gbh_ctx* ctx, gbh_event* ev);
int gbh_event_record(gbh_ctx* ctx, int stream, gbh_event* ev);
int gbh_event_sync(gbh_ctx* ctx, gbh_event* ev);
int gbh_event_elapsed_ms(gbh_ctx* ctx, gbh_event* a, gbh_event* b, float* ms);
/* Work queued on `stream` after this call waits for `ev` (recorded on the
 * other stream). */
int gbh_stream_wait(gbh_ctx* ctx, int stream, gbh_event* ev);

/* ---- a batch's launches as one HIP graph ----
 * No counterpart in the reference (its host appends a segment's kernels to
 * Level Zero command lists one by one, src/verifier.cpp:2487-2517): a
 * batch's whole pipeline captured once and launched as one graph, for
 * callers that replay batches of one shape. gbh_graph_begin makes
 * `stream` capture instead of run: every copy, launch, event record and
 * wait queued on it, and on every stream that waits on an event recorded in
 * the capture, becomes a node. gbh_graph_end ends the capture (every such
 * stream must have been joined back by a wait on `stream`), makes the
 * graph's executable form (in place, when the graph `g` ran before with
 * the same nodes and edges; else anew) and launches it on `stream`. While
 * capturing, gbh_event_record records a dependency only (time a captured
 * batch by events recorded before gbh_graph_begin and after gbh_graph_end).
 * The library's launchers never queue a memset (clears are fill kernels),
 * so a captured graph holds kernels and copies only. A graph's executable
 * form is relaunched only after its previous launch completed (the caller
 * synchronises on an event recorded after gbh_graph_end). */
typedef struct gbh_graph gbh_graph;
int gbh_graph_create(gbh_ctx* ctx, gbh_graph** g);
int gbh_graph_destroy(gbh_ctx* ctx, gbh_graph* g);
int gbh_graph_begin(gbh_ctx* ctx, int stream);
int gbh_graph_end(gbh_ctx* ctx, int stream, gbh_graph* g);
/* Instantiations and in-place updates made by gbh_graph_end on `g`. */
int gbh_graph_stats(gbh_ctx* ctx, const gbh_graph* g, uint64_t* instantiated, uint64_t* updated);

/* ---- stage launches (all asynchronous on `stream`) ---- */
int gbh_launch_setup(gbh_ctx* ctx, int stream, const gbh_setup_args* a);
int gbh_launch_harvest(gbh_ctx* ctx, int stream, const gbh_harvest_args* a);
int gbh_launch_partition(gbh_ctx* ctx, int stream, const gbh_partition_args* a);
int gbh_launch_mtrans(gbh_ctx* ctx, int stream, const gbh_mtrans_args* a);
int gbh_launch_window(gbh_ctx* ctx, int stream, const gbh_window_args* a);
int gbh_launch_check_a(gbh_ctx* ctx, int stream, const gbh_check_args* a);
int gbh_launch_check_b(gbh_ctx* ctx, int stream, const gbh_checkb_args* a);
int gbh_launch_count(gbh_ctx* ctx, int stream, const gbh_count_args* a);
int gbh_launch_mapcmp(gbh_ctx* ctx, int stream, const gbh_mapcmp_args* a);

/* ---- allocation-free form ----
 * gbh_launch_count without a workspace: the counter's scratch (the first
 * and last prime position of every piece of every map word, read back by
 * the gap walk across pieces) is the library's own static device array of
 * GBH_NOALLOC_SCRATCH_WORDS u32 (16 MiB per device, held until the process
 * ends; gbh_open looks up its address on the context's device). No
 * counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The caller passes inputs and outputs as for gbh_launch_count; a->scratch
 * must be NULL and nseg * swords <= GBH_NOALLOC_SCRATCH_WORDS, else
 * GBH_E_ARG. The call allocates and frees nothing. Every call on one device
 * shares that device's array, which keeps the last call's contents (each
 * call writes every entry it reads, so nothing needs clearing): calls that
 * overlap on one device (two streams, contexts or threads) race. */
#define GBH_NOALLOC_SCRATCH_WORDS (4u << 20)
int gbh_launch_count_noalloc(gbh_ctx* ctx, int stream, const gbh_count_args* a);

#ifdef __cplusplus
}
#endif
#endif
