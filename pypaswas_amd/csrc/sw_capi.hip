// C ABI of the gfx950 Smith-Waterman library (declared in
// include/pypaswas_hip.h).  One execution of a prepared plan is a single
// sequence of device work on one stream pair, with no host round trip in the
// middle:
//   1. clear the control block (work counters, bump allocators);
//   2. scoring pass over all pairs (plain: padded shapes, static schedule,
//      one launch per batch; batched: exact shapes, lane-per-pair jobs and
//      wave-per-pair tiles pulled by persistent workers from device counters);
//   3. the traceback gate of default_main_cpu.cl:408-410, which writes one
//      task per passing pair and takes its direction and candidate areas from
//      device bump allocators;
//   4. direction pass, start selection + path marking, walk, each over the
//      device-side task count;
//   5. read back the control block and the pair maxima, then the tasks and
//      the compacted starting points, hits and op characters.
// Tasks that did not fit the work areas are finished by host-driven tail
// rounds with exact sizes.  Because nothing in 1-4 depends on the host, the
// sequence can be captured once in a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cctype>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/pypaswas_hip.h"
#include "sw_device.hip"

using namespace pw;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Scoring {
    std::vector<float> subst;
    int dim = 0, char_offset = 0;
    float gap = 0, ext = 0, ll = 1, minimum = 0;
    bool affine = false;
    bool integral = false;  // every substitution and gap score is a whole number
    float max_abs = 0;      // largest magnitude of those scores
};

// Two streams and the events of one execution.  ev: 0 score start, 1 score
// end, 2 gate end, 3 directions end, 4 traceback end, 5 walk end; fork/join
// order the tile-kernel stream against the main one.
struct Lane {
    hipStream_t s = nullptr, s2 = nullptr;
    hipEvent_t ev[6] = {};
    hipEvent_t fork = nullptr, join = nullptr;
    hipEvent_t fork2 = nullptr, join2 = nullptr;  // maxima copy on s2 beside the traceback stages
    // two more streams for the lane-per-pair kernels of a two-stream lane
    hipStream_t sx[2] = {};
    hipEvent_t joinx[2] = {};
};

// Work-area sizing rules (pw_set_limits).
struct Limits {
    int64_t dir_init = (int64_t)256 << 20;  // first direction area, bytes
    int64_t dir_max = 0;                    // largest direction area the pools grow to (0: a quarter of HBM)
    int32_t cand0 = 32;                     // candidate slots per gated pair
    int64_t ops_init = (int64_t)16 << 20;   // first ops area, bytes
    bool grow = true;                       // resize the areas after an execution that overflowed them
};

// Device work areas of one execution, plus the pinned host staging of the
// read-back.  want_* are the sizes the next execution allocates.
struct Workspace {
    DevBuf ctl, maxima, gated, tasks, touts, dirs, cands, spslot, keys, out_sps, out_hits, ops, gscratch, tscratch,
        dscratch, units, dlinks, dprog, rq_slots, rq_needs;
    // the gate's block counts; the result layout's block sums, per-slot
    // offsets and the laid-out records, op characters and alignment rows
    DevBuf gsums, lsums, lay_off, f_sps, f_hits, f_ops, f_aln;
    int64_t want_dir = 0, want_cand = 0, want_ops = 0;
    int32_t cand0 = 0;
    int64_t dir_cap = 0, cand_cap = 0, ops_cap = 0;  // usable sizes of the current buffers
    int64_t out_cap = 0;                             // records the read-back areas (out_sps, out_hits) hold
    int64_t units_cap = 0, dlinks_cap = 0, dprog_cap = 0;
    PinBuf h_ctl, h_tasks, h_touts, h_sps, h_hits, h_ops;
    // direction area lent by a caller (a range of a serving arena) instead of dirs
    uint8_t *dir_ext = nullptr;
    int64_t dir_ext_cap = 0;
};

uint8_t *dir_ptr(const Workspace &W) { return W.dir_ext ? W.dir_ext : (uint8_t *)W.dirs.p; }

// Candidate slots (cands, spslot, keys: one round's) and record slots
// (out_sps, out_hits: every round's) of the current buffers.
void set_slot_caps(Workspace &W) {
    W.cand_cap = std::min<int64_t>({(int64_t)(W.cands.cap / sizeof(Cand)), (int64_t)(W.spslot.cap / sizeof(SpRec)),
                                    (int64_t)(W.keys.cap / sizeof(uint64_t)), (int64_t)INT32_MAX});
    W.out_cap = std::min<int64_t>({(int64_t)(W.out_sps.cap / sizeof(SpRec)),
                                   (int64_t)(W.out_hits.cap / sizeof(HitRec)), (int64_t)INT32_MAX});
}

}  // namespace

struct PlanScratch;  // host buffers of the planner (defined with the planner)

namespace {

// Measured shape table (pw_autotune), one per gap model.  Buckets of the
// lane-side length a (<= 64, 128, 256, 512, 1024, 2048, more) and of the
// swept length b (<= 32, 64, 128, 256, 512, 1024, more); rate = cells per
// microsecond of the tile kernel with 1, 2, 4 or 8 residues per lane on a
// device full of such pairs.
constexpr int TUNE_A = 7, TUNE_B = 7, TUNE_R = 4;
constexpr int TUNE_RS[TUNE_R] = {1, 2, 4, 8};
constexpr int TUNE_REP_A[TUNE_A] = {48, 96, 192, 384, 768, 1536, 3072};
constexpr int TUNE_REP_B[TUNE_B] = {24, 48, 96, 192, 384, 768, 1536};
// Chunk width of multi-chunk lane-per-pair jobs (column sequence longer than 96
// residues).  80 rather than 64: fewer chunk edges through the scratch column
// per cell outweigh the wider chunk's padding and its register use (two waves
// per SIMD instead of three); measured on the 2,000 x 20,000 protein set
// (columns the queries, ~350 residues): score pass 1071 ms at 64, 993 at 80,
// 1008 at 96, 1433 at 48.
constexpr int GROUP_NB_LONG = 80;
struct Tuning {
    bool valid = false;
    float rate[TUNE_A][TUNE_B][TUNE_R] = {};
    int group_nb_long = GROUP_NB_LONG;  // chunk width of lane-per-pair jobs whose column sequence exceeds 96 residues
    float group_rate[5] = {};  // measured for widths 32, 48, 64, 80, 96
};

// Planning choices besides the inputs: device facts, the tuned shapes, and
// the overrides the autotuner uses to time one shape.
struct PlanOpts {
    size_t total_mem = 0;
    int num_cus = 256;
    const Tuning *tune = nullptr;
    int force_R = 0;        // > 0: every tile item uses these residues per lane, untransposed
    bool force_tile = false;  // every pair to the tile kernel
    bool force_group = false; // every pair to lane-per-pair jobs (no fill or length rule)
    bool no_group_split = false;  // no lane-per-pair job split into chunk items
    int force_nb = 0;       // > 0: chunk width of multi-chunk lane-per-pair jobs
    double group_max_share = 0.25;  // longest lane-per-pair job, as a share of the cells per device lane
    int orient = -1;  // column sequence of lane-per-pair jobs: 0 target, 1 query, -1 the cheaper one
    int dir_rmax = 0;  // > 0: largest residues per lane of the direction pass (2, 4 or 8)
    double grid_share = 1.0;  // persistent grids cover at most this share of the device's resident waves
    bool checkpoints = true;  // long linear pairs save checkpoints for a block-wise direction pass (CkPair)
    PlanScratch *scratch = nullptr;  // reusable host buffers of the planner (the context's; may be null)
};

// Bucket k holds lengths up to 4/3 of its representative length rep[k]
// (64, 128, ... for the lane side; 32, 64, ... for the swept side).
int tune_bucket(int v, const int *rep, int n) {
    for (int k = 0; k < n - 1; ++k)
        if (v <= rep[k] * 4 / 3) return k;
    return n - 1;
}

}  // namespace

// A result as the library owns it until pw_free_result; pub is its public
// view (first member, so a pw_result pointer converts back).
// Host array of a result that only grows (a recycled result fills it without
// touching or zeroing new memory): pinned when the device copies into it,
// else ordinary memory.  A pinned block replaced while frees are deferred goes
// to *dead (hipHostFree waits for the whole device).
template <class T>
struct ResultArray {
    T *p = nullptr;
    size_t cap = 0;  // elements
    bool pinned = false;
    ResultArray() = default;
    ResultArray(const ResultArray &) = delete;
    ResultArray &operator=(const ResultArray &) = delete;
    ~ResultArray() { drop(nullptr); }
    void drop(std::vector<void *> *dead) {
        if (p) {
            if (!pinned) std::free(p);
            else if (dead) dead->push_back(p);
            else (void)hipHostFree(p);
        }
        p = nullptr;
        cap = 0;
        pinned = false;
    }
    // room for n elements (pinned if pin); false when no memory
    bool reserve(size_t n, bool pin, std::vector<void *> *dead) {
        if (n <= cap && (pinned || !pin)) return true;
        drop(dead);
        const size_t want = std::max<size_t>(n + n / 8, 16);
        void *q = nullptr;
        if (pin && hipHostMalloc(&q, want * sizeof(T), hipHostMallocDefault) == hipSuccess) {
            pinned = true;
        } else {
            if (pin) (void)hipGetLastError();
            q = std::malloc(want * sizeof(T));
            if (!q) return false;
        }
        p = static_cast<T *>(q);
        cap = want;
        return true;
    }
    T *data() { return p; }
    const T *data() const { return p; }
    T &operator[](size_t i) { return p[i]; }
    const T &operator[](size_t i) const { return p[i]; }
};

struct HostResult {
    pw_result pub;
    float *maxima = nullptr;  // pinned host memory of maxima_cap floats
    size_t maxima_cap = 0;
    std::vector<float> maxima_vec;  // maxima of a served job (when maxima is NULL)
    // records; the vectors only grow (a recycled result fills them without
    // touching or zeroing new memory), n_sps and n_ops are the sizes in use
    ResultArray<pw_starting_point> sps;
    ResultArray<pw_hit> hits;
    ResultArray<char> ops;
    ResultArray<char> aligned;  // 3 chars per op: the three alignment rows of each shown hit
    int64_t n_sps = 0, n_ops = 0;
};

struct pw_ctx {
    int device = 0;
    int num_cus = 0;
    size_t total_mem = 0;
    Lane lane;
    std::string err;
    bool scoring_set = false;
    Scoring sc;
    Limits lim;
    Workspace ws;  // shared by the one-shot calls (pw_align_pairs, pw_score_pairs)
    // pinned host buffers for result maxima, recycled by pw_free_result
    std::vector<std::pair<float *, size_t>> pinned_free;
    int64_t reallocs = 0;  // device or pinned buffers grown (statistics)
    // while a server has batches in flight, buffers replaced by larger ones
    // are kept here instead of freed (hipFree and hipHostFree wait for the
    // whole device) and released once the device is idle
    bool defer_free = false;
    std::vector<std::pair<void *, size_t>> dead_dev;
    std::vector<void *> dead_pinned;
    // device buffers given back, by size, kept for reuse (dev_alloc, dev_release)
    std::multimap<size_t, void *> dev_cache;
    size_t dev_cache_bytes = 0;
    std::map<std::pair<const void *, size_t>, int> occupancy;
    PinBuf h_stage;  // upload staging of the plans of pw_prepare and the one-shot calls
    Tuning tune[2];  // linear, affine
    int route_orient = -1;        // pw_set_routing: orientation of lane-per-pair jobs (-1: the cheaper)
    double route_share = 0.25;    // pw_set_routing: longest lane-per-pair job (share of cells per lane)
    bool layout_on_device = true;  // pw_set_assembly: lay results out on the device when every pair finished there
    int route_dir_rmax = 0;       // pw_set_routing: residues per lane of the direction pass (0: by plan)
    std::vector<HostResult *> result_pool;  // freed results kept for reuse
    // this device's copy of pw_static_scratch (hipGetSymbolAddress at pw_init),
    // the only device memory pw_score_pairs_noalloc uses
    void *static_scratch = nullptr;
    std::string static_err;
    unsigned wait_polls = WAIT_POLLS;  // pw_set_wait_limit: polls before a hand-over wait gives up
    PlanScratch *plan_scratch = nullptr;  // planner buffers reused across this context's plans
    pw_plan *oneshot = nullptr;           // the plan of the one-shot calls, reused (its buffers keep their size)
};

static thread_local std::string g_init_err;

namespace {

struct Encoded {
    int K = 0, KP = 8, lkp = 5;
    std::vector<uint8_t> chars;     // code -> the residue character it stands for
    std::vector<uint8_t> codes;     // all residues, queries then targets
    std::vector<SeqRef> refs;       // nq + nt
    std::vector<float> tables;      // 2 * KP * KP
    int64_t max_len_q = 0, max_len_t = 0;
};

// chunk widths of the lane-per-pair kernel
constexpr int NB_WIDTHS[NB_KINDS] = {16, 24, 32, 40, 48, 56, 64, 72, 80, 88, 96};

// Launch geometry of one execution, computed before any launch (and before
// a graph capture, during which no occupancy query is made).
struct Geometry {
    int group_grid[NB_KINDS] = {};
    int64_t gscr_stride = 0;  // float2 per workgroup
    int tile_grid = 0;
    int64_t tscr_stride = 0;  // floats per workgroup
    int gate_grid = 0;
    int build_grid = 0;  // workgroups of 256 threads of the task builder
    int dir_grid = 0;
    int64_t dscr_stride = 0;
    int task_grid = 0;  // workgroups of 64 threads of the start-selection and walk kernels (a pair per thread)
    int wave_grid = 0;  // one-wave workgroups of their long-path forms (a pair per wave); 0: no long pairs
};

}  // namespace

// A prepared job: sequences, substitution tables, the scoring work list and
// the pair list resident on the device, plus the host copies the read-back
// needs and (once executed through a graph) the instantiated graph.
struct pw_plan {
    Encoded E;
    Scoring sc;
    int32_t nq = 0, nt = 0;
    int64_t n_pairs = 0;
    int32_t mode = 1;
    bool cross = false;  // the pairs are every query with every target, query-major (pair = q * nt + t)
    std::vector<int32_t> pq, pt, qlen, tlen;
    std::vector<float> mss;
    std::vector<uint8_t> same;
    int64_t max_b = 0;
    int64_t max_t = 0;          // longest target of any pair
    int64_t dir_worst = 0;      // direction bytes if every pair passed the gate
    int64_t dir_largest = 0;    // largest single pair's direction area
    double cells = 0;
    int64_t n_tile = 0;        // items of the wave-per-pair tile kernel (one per stripe of a split pair)
    int64_t score_links = 0;   // floats of stripe-boundary rows of the split pairs of the scoring pass
    int64_t score_prog = 0;    // their progress counters
    int64_t units_worst = 0;   // direction-pass work units if every pair passed the gate
    int64_t dlinks_worst = 0;  // boundary-row floats of the direction pass in that case
    int64_t dprog_worst = 0;   // and its progress counters
    int32_t dir_rmax = 8;      // largest residues per lane of the direction pass
    int64_t n_groups = 0;      // lane-per-pair items (a split job counts one per chunk)
    int64_t group_split_items = 0;  // of which items of split jobs
    int32_t orient = 0;        // column sequence of the lane-per-pair jobs: 0 target, 1 query
    double group_cells = 0;    // cells scored lane-per-pair
    int64_t group_rows = 0;    // longest row sequence of a multi-chunk job (scratch sizing)
    int32_t max_nb = 16;
    int64_t nb_begin[NB_KINDS] = {}, nb_end[NB_KINDS] = {};  // job ranges per chunk width
    std::vector<std::pair<int64_t, int64_t>> batches;         // plain launch: item range of each batch
    std::vector<ScoreItem> items;
    std::vector<uint8_t> trans;
    std::vector<CkPair> ckpairs;  // checkpointed pairs, by pair index
    int64_t ck_floats = 0;        // their checkpoints and block maxima
    int64_t ck_units = 0;         // their direction-pass blocks (beyond units_worst)
    int64_t ck_dir_extra = 0;     // their direction bytes at CK_R beyond those at the plan's cap (in dir_worst)
    std::vector<GroupItem> gitems;
    std::vector<GroupEntry> gents;
    // per-plan arrays: views into one device allocation (blob) filled by one
    // copy from pinned staging
    DevBuf res, seqs, tables, d_items, d_trans, d_gitems, d_gents, d_mss, d_pairs, d_same;
    DevBuf d_ckpairs;
    DevBuf d_chars;  // 512 bytes: the residue character of each code, then its lower case
    std::vector<uint8_t> chartab;
    DevBuf blob, d_links, d_prog, d_ck;
    PinBuf h_blob;
    Workspace own;
    PlanOpts opts;  // routing choices the work list was built with
    // graph of the execution and the launch parameters it was captured with
    hipGraphExec_t graph = nullptr;
    std::vector<uint64_t> graph_sig;
};

#define HIPCHK(ctx, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                 \
            return PW_EDEVICE;                                                              \
        }                                                                                   \
    } while (0)

// Device memory is not given back to the runtime while the context lives:
// after one hipFree of a large allocation (256 MiB and up, measured on
// MI355X with ROCm 7) every later device-to-host copy of the process runs
// about three times slower (a 5 MiB copy into pinned memory, 0.10 -> 0.35 ms),
// and hipFree waits for the whole device besides.  Released buffers go to a
// cache by size instead; dev_alloc takes a cached buffer that is large enough
// and not much larger, else allocates.  The cache is emptied only when an
// allocation fails, or by pw_destroy.  A buffer enters the cache only once no
// queued work uses it (the callers release after the stream synchronized, or
// through the deferred list while a server runs).
static void dev_cache_flush(pw_ctx *ctx) {
    for (auto &kv : ctx->dev_cache) (void)hipFree(kv.second);
    ctx->dev_cache.clear();
    ctx->dev_cache_bytes = 0;
}

static void dev_release(pw_ctx *ctx, void *p, size_t cap) {
    if (!p) return;
    ctx->dev_cache.emplace(cap, p);
    ctx->dev_cache_bytes += cap;
}

// A buffer of at least need bytes; want bytes when a new one is allocated.
static int dev_alloc(pw_ctx *ctx, size_t need, size_t want, void **p, size_t *cap) {
    auto it = ctx->dev_cache.lower_bound(need);
    if (it != ctx->dev_cache.end() && it->first <= std::max(4 * need, need + ((size_t)16 << 20))) {
        *p = it->second;
        *cap = it->first;
        ctx->dev_cache_bytes -= it->first;
        ctx->dev_cache.erase(it);
        return PW_OK;
    }
    hipError_t e = hipMalloc(p, want);
    if (e != hipSuccess && !ctx->dev_cache.empty()) {
        (void)hipGetLastError();
        dev_cache_flush(ctx);
        e = hipMalloc(p, want);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        *cap = 0;
        ctx->err = "hipMalloc of " + std::to_string(want) + " bytes: " + hipGetErrorString(e);
        return PW_ENOMEM;
    }
    *cap = want;
    return PW_OK;
}

// Buffers grow to twice the request, so that a sequence of growing requests
// reallocates rarely.
static int ensure(pw_ctx *ctx, DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return PW_OK;
    ctx->reallocs += 1;
    if (b.p) {
        if (ctx->defer_free) ctx->dead_dev.emplace_back(b.p, b.cap);
        else dev_release(ctx, b.p, b.cap);
        b.p = nullptr;
        b.cap = 0;
    }
    const size_t want = bytes < ((size_t)1 << 30) ? 2 * bytes : bytes + bytes / 4;
    return dev_alloc(ctx, bytes, want, &b.p, &b.cap);
}

static int ensure_pinned(pw_ctx *ctx, PinBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 64;
    if (b.cap >= bytes) return PW_OK;
    ctx->reallocs += 1;
    if (b.p) {
        if (ctx->defer_free) ctx->dead_pinned.push_back(b.p);
        else (void)hipHostFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = 2 * bytes;
    hipError_t e = hipHostMalloc(&b.p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        b.p = nullptr;
        ctx->err = "hipHostMalloc of " + std::to_string(want) + " bytes: " + hipGetErrorString(e);
        return PW_ENOMEM;
    }
    b.cap = want;
    return PW_OK;
}

#define ENSURE(ctx, buf, bytes)                      \
    do {                                             \
        int rc_ = ensure((ctx), (buf), (bytes));     \
        if (rc_ != PW_OK) return rc_;                \
    } while (0)

#define ENSURE_PINNED(ctx, buf, bytes)                  \
    do {                                                \
        int rc_ = ensure_pinned((ctx), (buf), (bytes)); \
        if (rc_ != PW_OK) return rc_;                   \
    } while (0)

// Releases the buffers kept while frees were deferred.
static void release_dead(pw_ctx *ctx) {
    for (auto &d : ctx->dead_dev) dev_release(ctx, d.first, d.second);
    for (void *p : ctx->dead_pinned) (void)hipHostFree(p);
    ctx->dead_dev.clear();
    ctx->dead_pinned.clear();
}

namespace {

void free_buf(pw_ctx *ctx, DevBuf &b) {
    dev_release(ctx, b.p, b.cap);
    b.p = nullptr;
    b.cap = 0;
}

void free_pinned(PinBuf &b) {
    if (b.p) (void)hipHostFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

void free_workspace(pw_ctx *ctx, Workspace &W) {
    DevBuf *bufs[] = {&W.ctl, &W.maxima, &W.gated, &W.tasks, &W.touts, &W.dirs, &W.cands, &W.spslot, &W.keys,
                      &W.out_sps, &W.out_hits, &W.ops, &W.gscratch, &W.tscratch, &W.dscratch, &W.units, &W.dlinks,
                      &W.dprog, &W.rq_slots, &W.rq_needs, &W.gsums, &W.lsums, &W.lay_off, &W.f_sps, &W.f_hits,
                      &W.f_ops, &W.f_aln};
    for (DevBuf *b : bufs) free_buf(ctx, *b);
    PinBuf *pins[] = {&W.h_ctl, &W.h_tasks, &W.h_touts, &W.h_sps, &W.h_hits, &W.h_ops};
    for (PinBuf *b : pins) free_pinned(*b);
}

// single: one stream for everything (s2 == s), so that a lane occupies one
// hardware queue; the tile kernels of a batch then follow its lane-per-pair
// kernels instead of overlapping them.
int create_lane(Lane &L, std::string &err, bool single = false) {
    hipError_t e = hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking);
    if (single) L.s2 = L.s;
    else if (e == hipSuccess) e = hipStreamCreateWithFlags(&L.s2, hipStreamNonBlocking);
    for (int i = 0; i < 2 && !single; ++i) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&L.sx[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&L.joinx[i], hipEventDisableTiming);
    }
    for (auto &ev : L.ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.fork2, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.join2, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.join, hipEventDisableTiming);
    if (e != hipSuccess) {
        err = std::string("stream/event creation: ") + hipGetErrorString(e);
        return PW_EDEVICE;
    }
    return PW_OK;
}

void destroy_lane(Lane &L) {
    if (L.s) (void)hipStreamSynchronize(L.s);
    if (L.s2 && L.s2 != L.s) (void)hipStreamSynchronize(L.s2);
    for (int i = 0; i < 2; ++i) {
        if (L.sx[i]) {
            (void)hipStreamSynchronize(L.sx[i]);
            (void)hipStreamDestroy(L.sx[i]);
        }
        if (L.joinx[i]) (void)hipEventDestroy(L.joinx[i]);
    }
    for (auto &ev : L.ev)
        if (ev) (void)hipEventDestroy(ev);
    if (L.fork) (void)hipEventDestroy(L.fork);
    if (L.fork2) (void)hipEventDestroy(L.fork2);
    if (L.join2) (void)hipEventDestroy(L.join2);
    if (L.join) (void)hipEventDestroy(L.join);
    if (L.s2 && L.s2 != L.s) (void)hipStreamDestroy(L.s2);
    if (L.s) (void)hipStreamDestroy(L.s);
    L = Lane();
}

}  // namespace

// Device memory of pw_score_pairs_noalloc: a file-scope array that each GPU
// the library runs on holds from module load until the process ends.  Host
// code never passes the array's name to a kernel or a copy (on the host it is
// only a stand-in for the symbol); pw_init looks up the device address of this
// GPU's copy and keeps it in the context.
static __device__ __align__(256) unsigned char pw_static_scratch[PW_NOALLOC_SCRATCH_BYTES];

extern "C" int pw_init(int device, pw_ctx **out) {
    if (!out) {
        g_init_err = "pw_init: out is NULL";
        return PW_EINVAL;
    }
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        g_init_err = std::string("no HIP device: ") + (e != hipSuccess ? hipGetErrorString(e) : "count is 0");
        return PW_EDEVICE;
    }
    if (device < 0 || device >= n) {
        g_init_err = "device " + std::to_string(device) + " out of range (" + std::to_string(n) + " devices)";
        return PW_EINVAL;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        g_init_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return PW_EDEVICE;
    }
    pw_ctx *ctx = new pw_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->num_cus = prop.multiProcessorCount;
        ctx->total_mem = prop.totalGlobalMem;
    }
    if (ctx->num_cus <= 0) ctx->num_cus = 256;
    if (ctx->total_mem == 0) {
        size_t f = 0, t = 0;
        if (hipMemGetInfo(&f, &t) == hipSuccess) ctx->total_mem = t;
    }
    if (create_lane(ctx->lane, g_init_err) != PW_OK) {
        destroy_lane(ctx->lane);
        delete ctx;
        return PW_EDEVICE;
    }
    // with this device current: the address of its copy of the static scratch
    e = hipGetSymbolAddress(&ctx->static_scratch, HIP_SYMBOL(pw_static_scratch));
    if (e != hipSuccess) {
        ctx->static_scratch = nullptr;
        ctx->static_err = std::string("hipGetSymbolAddress(pw_static_scratch): ") + hipGetErrorString(e);
        (void)hipGetLastError();
    }
    *out = ctx;
    return PW_OK;
}

namespace {
void free_context_plans(pw_ctx *ctx);
}

extern "C" int pw_destroy(pw_ctx *ctx) {
    if (!ctx) return PW_EINVAL;
    (void)hipSetDevice(ctx->device);
    destroy_lane(ctx->lane);
    free_context_plans(ctx);
    free_workspace(ctx, ctx->ws);
    free_pinned(ctx->h_stage);
    release_dead(ctx);
    dev_cache_flush(ctx);
    for (HostResult *R : ctx->result_pool) delete R;
    for (auto &pb : ctx->pinned_free) (void)hipHostFree(pb.first);
    delete ctx;
    return PW_OK;
}

extern "C" const char *pw_last_error(pw_ctx *ctx) { return ctx ? ctx->err.c_str() : g_init_err.c_str(); }

extern "C" int pw_device_mem_bytes(pw_ctx *ctx, uint64_t *bytes) {
    if (!ctx || !bytes) return PW_EINVAL;
    size_t free_b = 0, total_b = 0;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemGetInfo(&free_b, &total_b));
    *bytes = total_b;
    return PW_OK;
}

namespace {

int make_scoring(std::string &err, const float *subst, int dim, int char_offset, float gap, float gap_ext,
                 float lower_limit, float minimum_score, Scoring &sc) {
    if (!subst || dim <= 0 || dim > 255) {
        err = "scoring: substitution matrix must be dim x dim with 0 < dim <= 255";
        return PW_EINVAL;
    }
    if (!(gap <= 0.0f) || !(gap_ext <= 0.0f)) {
        err = "scoring: gap and gap extension scores must be <= 0";
        return PW_EINVAL;
    }
    if (!(gap > GAP_INIT) || !(gap_ext > GAP_INIT) || !std::isfinite(lower_limit) || !std::isfinite(minimum_score)) {
        err = "scoring: scoring constants out of range";
        return PW_EINVAL;
    }
    for (int i = 0; i < dim * dim; ++i)
        if (!std::isfinite(subst[i]) || std::fabs(subst[i]) >= 1e9f) {
            err = "scoring: substitution scores must be finite and below 1e9 in magnitude";
            return PW_EINVAL;
        }
    sc.subst.assign(subst, subst + (size_t)dim * dim);
    sc.integral = std::floor(gap) == gap && std::floor(gap_ext) == gap_ext;
    sc.max_abs = std::max(std::fabs(gap), std::fabs(gap_ext));
    for (int i = 0; i < dim * dim; ++i) {
        sc.integral = sc.integral && std::floor(subst[i]) == subst[i];
        sc.max_abs = std::max(sc.max_abs, std::fabs(subst[i]));
    }
    sc.dim = dim;
    sc.char_offset = char_offset;
    sc.gap = gap;
    sc.ext = gap_ext;
    sc.affine = gap_ext != 0.0f;
    sc.ll = lower_limit;
    sc.minimum = minimum_score;
    return PW_OK;
}

}  // namespace

extern "C" int pw_set_scoring(pw_ctx *ctx, const float *subst, int dim, int char_offset, float gap, float gap_ext,
                              float lower_limit, float minimum_score) {
    if (!ctx) return PW_EINVAL;
    Scoring sc;
    int rc = make_scoring(ctx->err, subst, dim, char_offset, gap, gap_ext, lower_limit, minimum_score, sc);
    if (rc) return rc;
    ctx->sc = std::move(sc);
    ctx->scoring_set = true;
    return PW_OK;
}

extern "C" int pw_set_limits(pw_ctx *ctx, int64_t max_direction_bytes, int32_t candidates_per_pair,
                             int64_t ops_bytes, int32_t grow) {
    if (!ctx) return PW_EINVAL;
    if (max_direction_bytes < 0 || candidates_per_pair < 0 || ops_bytes < 0) {
        ctx->err = "pw_set_limits: negative size";
        return PW_EINVAL;
    }
    if (max_direction_bytes > 0) {
        ctx->lim.dir_max = max_direction_bytes;
        ctx->lim.dir_init = std::min<int64_t>(ctx->lim.dir_init, max_direction_bytes);
    }
    if (candidates_per_pair > 0) ctx->lim.cand0 = std::min<int32_t>(candidates_per_pair, 1 << 16);
    if (ops_bytes > 0) ctx->lim.ops_init = ops_bytes;
    ctx->lim.grow = grow != 0;
    // the shared work areas follow the new rules from the next call on
    ctx->ws.want_dir = ctx->ws.want_cand = ctx->ws.want_ops = 0;
    ctx->ws.cand0 = 0;
    return PW_OK;
}

extern "C" int pw_set_wait_limit(pw_ctx *ctx, int64_t polls) {
    if (!ctx) return PW_EINVAL;
    if (polls > (int64_t)UINT32_MAX) {
        ctx->err = "pw_set_wait_limit: at most 2^32 - 1 polls";
        return PW_EINVAL;
    }
    ctx->wait_polls = polls < 0 ? WAIT_POLLS : (unsigned)polls;
    return PW_OK;
}

extern "C" int pw_set_routing(pw_ctx *ctx, int32_t orient, double group_max_share, int32_t dir_residues) {
    if (!ctx) return PW_EINVAL;
    if (orient < -1 || orient > 1) {
        ctx->err = "pw_set_routing: orient must be -1, 0 or 1";
        return PW_EINVAL;
    }
    if (dir_residues != 0 && dir_residues != 2 && dir_residues != 4 && dir_residues != 8) {
        ctx->err = "pw_set_routing: dir_residues must be 0, 2, 4 or 8";
        return PW_EINVAL;
    }
    ctx->route_orient = orient;
    ctx->route_share = group_max_share > 0 ? group_max_share : 0.25;
    ctx->route_dir_rmax = dir_residues;
    return PW_OK;
}

extern "C" int pw_set_assembly(pw_ctx *ctx, int32_t on_device) {
    if (!ctx) return PW_EINVAL;
    ctx->layout_on_device = on_device != 0;
    return PW_OK;
}

namespace {

// A pool of host threads for the planner's parallel loops, created on first
// use and kept for the life of the process (its threads wait on a condition
// variable between runs; thread creation would cost tens of microseconds per
// thread and loop).  run(n, f) calls f(0) on the caller and f(1..n-1) on
// pool threads, and returns when all have returned; runs from different
// callers take turns.  f must not start a run itself.
class HostPool {
  public:
    static HostPool &get() {
        static HostPool *pool = new HostPool();  // never destroyed: its threads outlive static destruction
        return *pool;
    }
    int size() const { return nthreads_; }
    void run(int n, const std::function<void(int)> &f) {
        n = std::max(1, std::min(n, nthreads_));
        std::lock_guard<std::mutex> turn(run_mutex_);
        if (n > 1) {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &f;
            njobs_ = n;
            remaining_ = n - 1;
            ++gen_;
            cv_.notify_all();
        }
        f(0);
        if (n > 1) {
            std::unique_lock<std::mutex> lk(m_);
            done_cv_.wait(lk, [this] { return remaining_ == 0; });
            job_ = nullptr;
        }
    }

  private:
    HostPool() {
        nthreads_ = (int)std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency()));
        for (int k = 1; k < nthreads_; ++k) std::thread([this, k] { loop(k); }).detach();
    }
    void loop(int k) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)> *fn = nullptr;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (k >= njobs_) continue;
                fn = job_;
            }
            (*fn)(k);
            std::lock_guard<std::mutex> lk(m_);
            if (--remaining_ == 0) done_cv_.notify_one();
        }
    }
    int nthreads_ = 1;
    std::mutex run_mutex_, m_;
    std::condition_variable cv_, done_cv_;
    uint64_t gen_ = 0;
    int njobs_ = 0, remaining_ = 0;
    const std::function<void(int)> *job_ = nullptr;
};

// Host threads for n elements: at most the pool's, and one per min_per_thread elements.
int host_threads(int64_t n, int64_t min_per_thread) {
    const int64_t hw = HostPool::get().size();
    return (int)std::max<int64_t>(1, std::min<int64_t>(hw, n / std::max<int64_t>(1, min_per_thread)));
}

// Runs f(lo, hi, k) over [0, n) in contiguous ranges, thread k of the
// host_threads(n, min_per_thread) threads taking the k-th range.
template <class F>
void parallel_ranges(int64_t n, int64_t min_per_thread, F f) {
    const int nth = host_threads(n, min_per_thread);
    if (nth <= 1) {
        f((int64_t)0, n, 0);
        return;
    }
    HostPool::get().run(nth, [&](int k) { f(n * k / nth, n * (k + 1) / nth, k); });
}

// Residues into the compact alphabet: the characters present get codes in
// the order of their first appearance (queries, then targets).  Host threads
// find each character's first position in their share of the sequences, then
// encode their share.
int encode(std::string &err, const Scoring &sc, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen,
           int32_t nq, const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt, Encoded &E) {
    const int64_t ns = (int64_t)nq + nt;
    int64_t total = 0;
    for (int64_t i = 0; i < ns; ++i) {
        const int32_t len = i < nq ? qlen[i] : tlen[i - nq];
        const int64_t off = i < nq ? qoff[i] : toff[i - nq];
        if (len < 0 || off < 0) {
            err = "negative sequence length or offset";
            return PW_EINVAL;
        }
        total += len;
    }
    auto seq = [&](int64_t i, const uint8_t *&p, int32_t &len) {
        if (i < nq) {
            p = qres + qoff[i];
            len = qlen[i];
        } else {
            p = tres + toff[i - nq];
            len = tlen[i - nq];
        }
    };
    // first appearance of every byte value, as (sequence << 32) + position
    constexpr int64_t NONE = INT64_MAX;
    const int64_t per = std::max<int64_t>(1, ns * 4096 / std::max<int64_t>(1, total));  // ~4096 residues per element
    std::vector<std::array<int64_t, 256>> first(16);
    for (auto &f : first) f.fill(NONE);
    parallel_ranges(ns, per * 16, [&](int64_t lo, int64_t hi, int k) {
        std::array<int64_t, 256> &f = first[k];
        for (int64_t i = lo; i < hi; ++i) {
            const uint8_t *p;
            int32_t len;
            seq(i, p, len);
            for (int32_t j = 0; j < len; ++j)
                if (f[p[j]] == NONE) f[p[j]] = (i << 32) + j;
        }
    });
    std::array<int64_t, 256> pos;
    pos.fill(NONE);
    for (const auto &f : first)
        for (int c = 0; c < 256; ++c) pos[c] = std::min(pos[c], f[c]);
    std::vector<int> used;
    for (int c = 0; c < 256; ++c)
        if (pos[c] != NONE) used.push_back(c);
    std::sort(used.begin(), used.end(), [&](int a, int b) { return pos[a] < pos[b]; });
    for (int c : used) {
        const int idx = c - sc.char_offset;
        if (idx < 0 || idx >= sc.dim) {
            char buf[160];
            snprintf(buf, sizeof buf, "character %d ('%c') is outside the scoring alphabet (offset %d, dimension %d)", c,
                     (c >= 32 && c < 127) ? c : '?', sc.char_offset, sc.dim);
            err = buf;
            return PW_EINVAL;
        }
    }
    int map[256];
    for (int &m : map) m = -1;
    for (size_t k = 0; k < used.size(); ++k) map[used[k]] = (int)k;
    E.chars.assign(used.begin(), used.end());
    E.K = (int)used.size();
    if (E.K > KPMAX - 1) {
        err = "more than 63 distinct residues in one call";
        return PW_EINVAL;
    }
    E.KP = 8;
    while (E.KP < E.K + 1) E.KP *= 2;
    E.lkp = 2;
    for (int v = E.KP; v > 1; v >>= 1) ++E.lkp;  // log2(KP) + 2: byte shift of a table row
    const int KP = E.KP;
    E.tables.assign(2 * KP * KP, FILL_SCORE);
    for (int a = 0; a < E.K; ++a)
        for (int b = 0; b < E.K; ++b) {
            const int ia = used[a] - sc.char_offset, ib = used[b] - sc.char_offset;
            // table 0: A = query  -> entry [b][a] = S(query a, target b)
            E.tables[b * KP + a] = sc.subst[(size_t)ia * sc.dim + ib];
            // table 1: A = target -> entry [b][a] = S(query b, target a)
            E.tables[KP * KP + b * KP + a] = sc.subst[(size_t)ib * sc.dim + ia];
        }
    E.codes.resize((size_t)total + 64);
    E.refs.resize((size_t)ns);
    E.max_len_q = E.max_len_t = 0;
    int64_t run = 0;
    for (int64_t i = 0; i < ns; ++i) {
        const int32_t len = i < nq ? qlen[i] : tlen[i - nq];
        E.refs[i] = SeqRef{run, len, 0};
        run += len;
        if (i < nq) E.max_len_q = std::max<int64_t>(E.max_len_q, len);
        else E.max_len_t = std::max<int64_t>(E.max_len_t, len);
    }
    parallel_ranges(ns, per * 16, [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i) {
            const uint8_t *p;
            int32_t len;
            seq(i, p, len);
            uint8_t *dst = E.codes.data() + E.refs[i].off;
            for (int32_t j = 0; j < len; ++j) dst[j] = (uint8_t)map[p[j]];
        }
    });
    return PW_OK;
}

inline int pad8(int n) { return (n + 7) / 8 * 8; }

// Tile shape for the batched pass: orientation and residues per lane with the
// highest measured rate of the pair's length buckets when the device has been
// tuned, else those that minimise the modelled cost
// stripes * (columns + 63) * (R + overhead).
void choose_shape_model(int qlen, int tlen, int &trans, int &R);

int tune_rk(int R) { return R == 1 ? 0 : R == 2 ? 1 : R == 4 ? 2 : 3; }

// Measured rate of a shape (0 when the table has none).
double tuned_rate(const Tuning *tu, int qlen, int tlen, int o, int R) {
    const int a = o ? tlen : qlen, b = o ? qlen : tlen;
    return tu->rate[tune_bucket(a, TUNE_REP_A, TUNE_A)][tune_bucket(b, TUNE_REP_B, TUNE_B)][tune_rk(R)];
}

// Measured rates decide only when they beat the cost model's shape clearly
// (by TUNE_MARGIN): a shape timed alone on uniform pairs can lose a few per
// cent inside a mixed batch that shares the device with other kernels.
constexpr double TUNE_MARGIN = 1.08;

void choose_shape(int qlen, int tlen, int &trans, int &R, const Tuning *tu) {
    choose_shape_model(qlen, tlen, trans, R);
    if (!tu || !tu->valid) return;
    const double base = tuned_rate(tu, qlen, tlen, trans, R);
    double rbest = base;
    int ob = trans, Rb = R;
    for (int o = 0; o < 2; ++o)
        for (int k = 0; k < TUNE_R; ++k) {
            const double r = tuned_rate(tu, qlen, tlen, o, TUNE_RS[k]);
            if (r > rbest) {
                rbest = r;
                ob = o;
                Rb = TUNE_RS[k];
            }
        }
    if (rbest > TUNE_MARGIN * base) {
        trans = ob;
        R = Rb;
    }
}

void choose_shape_model(int qlen, int tlen, int &trans, int &R) {
    double best = 1e300;
    trans = 0;
    R = 4;
    for (int o = 0; o < 2; ++o) {
        const int a = o ? tlen : qlen, b = o ? qlen : tlen;
        for (int r : {1, 2, 4, 8}) {
            const int rows = 64 * r;
            const int stripes = std::max(1, (a + rows - 1) / rows);
            const double cost = (double)stripes * (b + 63) * (r + 1.5) + (stripes > 1 ? stripes * (b + 63) * 0.5 : 0);
            if (cost < best) {
                best = cost;
                trans = o;
                R = r;
            }
        }
    }
}

// a 64-lane job is used when its lanes are busy for at least this fraction
// of its rows; the longest row sequence such a job takes
constexpr double GROUP_MIN_FILL = 0.6;
// cost of a cell in the wave-per-pair tile kernel relative to a lane-cell of a
// lane-per-pair job (measured on MI355X: ~2.5x)
constexpr double TILE_COST_FACTOR = 2.5;
// a multi-chunk job longer than the length rule allows runs as one item per
// chunk in a row pipeline if its waves would wait at most this multiple of the
// job's work at the start (GROUP_PUB rows per chunk of the pipeline); measured
// on the protein sample, 0.2 / 0.5 / 2: score pass 21.3 / 20.0 / 19.2 ms
constexpr double GROUP_SPLIT_MAX_WAIT = 2.0;
constexpr int GROUP_MAX_ROWS = 1 << 15;

// columns per chunk of a lane-per-pair job: one chunk up to 96 columns,
// chunks of GROUP_NB_LONG (or the tuned width) beyond
int group_nb(int clen, const Tuning *tu) {
    if (clen <= 96) return std::max(16, (clen + 7) / 8 * 8);
    return (tu && tu->valid) ? tu->group_nb_long : GROUP_NB_LONG;
}


// two transposed substitution tables, the swept-residue window, the staging
// of the stripe boundary rows (up to 64 columns in, 64 out) and, for the direction
// pass, the candidate buffer
// LDS of the wave-per-pair kernels: the scoring pass holds both orientations
// of the substitution table, the direction pass (query along the lanes) one.
size_t lds_bytes(int KP, bool dirs = false) {
    return (size_t)(dirs ? 1 : 2) * KP * KP * sizeof(float) + BWIN + WAVE + 2 * WAVE * sizeof(float4) +
           (dirs ? CAND_BUF * sizeof(Cand) : 0);
}

// Checks that every pair indexes the sequence lists; *cross (when given) tells
// whether the pairs are every query with every target in query-major order
// (pair q * nt + t), the all-vs-all batch of the engine, which the planner
// handles without per-pair lists.
int check_pairs(std::string &err, int32_t nq, int32_t nt, const int32_t *pq, const int32_t *pt, int64_t n_pairs,
                bool *cross = nullptr) {
    if (n_pairs < 0 || n_pairs > (int64_t)INT32_MAX) {
        err = "n_pairs out of range";
        return PW_EINVAL;
    }
    // the first pair out of range, and whether each range follows the
    // query-major grid, over host threads
    std::vector<int64_t> bad(16, INT64_MAX);
    std::vector<uint8_t> grid(16, 1);
    const bool may_cross = nt > 0 && n_pairs > 0 && n_pairs == (int64_t)nq * nt;
    parallel_ranges(n_pairs, 1 << 18, [&](int64_t lo, int64_t hi, int k) {
        int32_t gq = may_cross ? (int32_t)(lo / nt) : 0, gt = may_cross ? (int32_t)(lo % nt) : 0;
        bool g = may_cross;
        for (int64_t i = lo; i < hi; ++i) {
            if (pq[i] < 0 || pq[i] >= nq || pt[i] < 0 || pt[i] >= nt) {
                bad[k] = i;
                break;
            }
            g = g && pq[i] == gq && pt[i] == gt;
            if (++gt == nt) {
                gt = 0;
                ++gq;
            }
        }
        grid[k] = g ? 1 : 0;
    });
    const int64_t first = *std::min_element(bad.begin(), bad.end());
    if (first != INT64_MAX) {
        err = "pair " + std::to_string(first) + " indexes outside the sequence lists";
        return PW_EINVAL;
    }
    if (cross) *cross = may_cross && std::all_of(grid.begin(), grid.end(), [](uint8_t v) { return v != 0; });
    return PW_OK;
}

void free_plan_buffers(pw_ctx *ctx, pw_plan *P) {
    DevBuf *bufs[] = {&P->blob, &P->d_links, &P->d_prog, &P->d_ck};
    for (DevBuf *b : bufs) free_buf(ctx, *b);
    free_pinned(P->h_blob);
    free_workspace(ctx, P->own);
    if (P->graph) (void)hipGraphExecDestroy(P->graph);
    P->graph = nullptr;
    P->graph_sig.clear();
}

// The plan's buffers are reused at once: work that an error left queued on the
// context's lane must end first (a server's lanes are destroyed before their
// plans are freed).
void free_plan(pw_ctx *ctx, pw_plan *P) {
    if (!P) return;
    (void)hipStreamSynchronize(ctx->lane.s);
    if (ctx->lane.s2 != ctx->lane.s) (void)hipStreamSynchronize(ctx->lane.s2);
    free_plan_buffers(ctx, P);
    delete P;
}

// Host half of pw_prepare: encodes the residues and builds the scoring work
// list.  No device call.
int plan_worklist(const PlanOpts &opt, pw_plan *P);

int plan_host(std::string &err, const Scoring &sc, const PlanOpts &opt, pw_plan *P, const uint8_t *qres,
              const int64_t *qoff, const int32_t *qlen, int32_t nq, const uint8_t *tres, const int64_t *toff,
              const int32_t *tlen, int32_t nt, const int32_t *pq, const int32_t *pt, int64_t n_pairs,
              const float *mss, const uint8_t *same_id, int32_t mode) {
    if (mode != 0 && mode != 1) {
        err = "mode must be 0 (plain per-batch launch) or 1 (batched persistent)";
        return PW_EINVAL;
    }
    bool cross = false;
    int rc = check_pairs(err, nq, nt, pq, pt, n_pairs, &cross);
    if (rc) return rc;
    P->cross = cross;
    P->E = Encoded();
    rc = encode(err, sc, qres, qoff, qlen, nq, tres, toff, tlen, nt, P->E);
    if (rc) return rc;
    P->sc = sc;
    P->nq = nq;
    P->nt = nt;
    P->n_pairs = n_pairs;
    P->mode = mode;
    // the caller's per-pair arrays, copied on host threads (a reused plan keeps their capacity)
    auto copy_in = [n_pairs](auto &v, const auto *src) {
        using T = typename std::decay<decltype(v)>::type::value_type;
        v.resize((size_t)n_pairs);
        parallel_ranges(n_pairs, 1 << 18, [&](int64_t lo, int64_t hi, int) {
            if (src) std::memcpy(v.data() + lo, src + lo, (size_t)(hi - lo) * sizeof(T));
            else std::memset((void *)(v.data() + lo), 0, (size_t)(hi - lo) * sizeof(T));
        });
    };
    copy_in(P->pq, pq);
    copy_in(P->pt, pt);
    P->qlen.assign(qlen, qlen + nq);
    P->tlen.assign(tlen, tlen + nt);
    copy_in(P->mss, mss);
    copy_in(P->same, same_id);
    // per-pair sizes of the traceback stages for each cap of the direction
    // pass's residues per lane (2, 4, 8), in one pass over the pairs on
    // several host threads; the cap is then the largest of 8, 4, 2 for which
    // the pairs, all gated, would give at least 4 work units per SIMD (more,
    // shorter stripes for few long pairs: their waves run latency-bound)
    struct Acc {
        int64_t cells = 0, max_t = 0;
        int64_t units[3] = {}, dir_worst[3] = {}, dir_largest[3] = {}, dlinks[3] = {}, dprog[3] = {};
    };
    const int caps[3] = {2, 4, 8};
    // per query and cap: residues per lane and stripes (the divisions of
    // dir_residues, dir_area and split_stripes, once per query instead of per pair)
    struct QShape {
        int32_t R[3], S[3];
    };
    std::vector<QShape> qs((size_t)nq);
    parallel_ranges(nq, 1 << 12, [&](int64_t lo, int64_t hi, int) {
        for (int64_t q = lo; q < hi; ++q)
            for (int c = 0; c < 3; ++c) {
                const int R = dir_residues(qlen[q], caps[c]), rows = WAVE * R;
                qs[q].R[c] = R;
                qs[q].S[c] = qlen[q] > 0 ? (qlen[q] + rows - 1) / rows : 1;
            }
    });
    Acc tot;
    if (cross) {
        // every query with every target: each sum over the pairs is a product
        // of a sum over the queries and one over the targets (a pair's
        // direction area S * (tlen + 63) * 32R needs no rounding)
        int64_t sum_tl = 0, sum_steps = 0, max_steps = 0, nlong = 0, links_long = 0;
        for (int32_t t = 0; t < nt; ++t) {
            const int64_t tl = tlen[t];
            sum_tl += tl;
            sum_steps += tl + WAVE - 1;
            max_steps = std::max<int64_t>(max_steps, tl + WAVE - 1);
            tot.max_t = std::max<int64_t>(tot.max_t, tl);
            if (tl >= SPLIT_MIN_COLS) {
                ++nlong;
                links_long += link_row_floats((int)tl);
            }
        }
        int64_t sum_ql = 0;
        for (int32_t q = 0; q < nq; ++q) {
            sum_ql += qlen[q];
            for (int c = 0; c < 3; ++c) {
                const int64_t S = qs[q].S[c], per = S * (WAVE / 2) * qs[q].R[c];
                tot.dir_worst[c] += per * sum_steps;
                tot.dir_largest[c] = std::max(tot.dir_largest[c], per * max_steps);
                if (S >= 2) {
                    tot.units[c] += S * nlong + (nt - nlong);
                    tot.dlinks[c] += (S - 1) * links_long;
                    tot.dprog[c] += S * nlong;
                } else {
                    tot.units[c] += nt;
                }
            }
        }
        tot.cells = sum_ql * sum_tl;
    } else {
        std::vector<Acc> acc(16);  // per thread (parallel_ranges runs at most 16)
        parallel_ranges(n_pairs, 1 << 16, [&](int64_t lo, int64_t hi, int k) {
            Acc a;
            for (int64_t i = lo; i < hi; ++i) {
                const QShape &qh = qs[pq[i]];
                const int ql = qlen[pq[i]], tl = tlen[pt[i]];
                a.cells += (int64_t)ql * tl;
                a.max_t = std::max<int64_t>(a.max_t, tl);
                const int64_t steps = (int64_t)tl + WAVE - 1;
                const bool long_cols = tl >= SPLIT_MIN_COLS;
                for (int c = 0; c < 3; ++c) {
                    const int64_t S = qh.S[c];
                    // dir_area(ql, tl, R) and split_stripes(ql, tl, R)
                    const int64_t area = (S * steps * (WAVE / 2) * qh.R[c] + 15) & ~(int64_t)15;
                    a.dir_worst[c] += area;
                    a.dir_largest[c] = std::max(a.dir_largest[c], area);
                    const bool split = S >= 2 && long_cols;
                    a.units[c] += split ? S : 1;
                    if (split) {
                        a.dlinks[c] += (S - 1) * link_row_floats(tl);
                        a.dprog[c] += S;
                    }
                }
            }
            acc[k] = a;
        });
        for (const Acc &a : acc) {
            tot.cells += a.cells;
            tot.max_t = std::max(tot.max_t, a.max_t);
            for (int c = 0; c < 3; ++c) {
                tot.units[c] += a.units[c];
                tot.dir_worst[c] += a.dir_worst[c];
                tot.dir_largest[c] = std::max(tot.dir_largest[c], a.dir_largest[c]);
                tot.dlinks[c] += a.dlinks[c];
                tot.dprog[c] += a.dprog[c];
            }
        }
    }
    int ci = 0;
    if (opt.dir_rmax > 0) {
        ci = opt.dir_rmax >= 8 ? 2 : (opt.dir_rmax >= 4 ? 1 : 0);
    } else {
        for (int c : {2, 1})
            if (tot.units[c] >= (int64_t)std::max(1, opt.num_cus) * 4 * 4) {
                ci = c;
                break;
            }
    }
    P->dir_rmax = opt.dir_rmax > 0 ? opt.dir_rmax : caps[ci];
    P->cells = (double)tot.cells;  // an exact integer below 2^53
    P->max_t = tot.max_t;
    P->dir_worst = tot.dir_worst[ci];
    P->dir_largest = tot.dir_largest[ci];
    P->units_worst = tot.units[ci];
    P->dlinks_worst = tot.dlinks[ci];
    P->dprog_worst = tot.dprog[ci];
    P->opts = opt;
    return plan_worklist(opt, P);
}

// Lane-per-pair jobs of one orientation of a batched plan: for every column
// sequence (the target for o = 0, the query for o = 1) its pairs in pair
// order, the candidates among them stably sorted by row length, longest
// first, cut into 64-lane jobs that are kept (or kept split into chunk
// items) or sent to the tile kernel, and the cost estimate of the choice.
// Per column c the arrays hold, from off[c]: the tile-kernel pairs (ntile[c]:
// the non-candidates in pair order, then the pairs of rejected jobs in job
// order), the kept pairs (ngrp[c], 64 per job) and, at the position of each
// kept job's first pair, whether it is split.  Columns are independent and
// run on several host threads.
struct PairRef {
    int32_t pair;  // pair index (n_pairs <= INT32_MAX)
    int32_t row;   // its row sequence: the query (o = 0) or the target (o = 1)
};
struct OrientPlan {
    double cost = 0;
    std::vector<int64_t> off, ntile, ngrp;
    std::vector<PairRef> by_col, tile, grp;
    std::vector<uint8_t> split;
    std::vector<int> nb;
    // the all-vs-all batch: per column and 64-lane group of the shared row
    // order, 0 (tile kernel), 1 (one wave) or 2 (one item per chunk), in place
    // of the tile, grp and split lists
    std::vector<uint8_t> choice;
};

}  // namespace

// Host buffers of the planner, kept by a context across plans so that their
// pages are not faulted in again for every call.
struct PlanScratch {
    OrientPlan op[2];
    // work-list temporaries of plan_worklist (kept so that their pages are
    // not faulted in and zeroed again for every plan)
    std::vector<int64_t> tile_pairs, gorder;
    std::vector<GroupItem> gitems, gjobs;
    std::vector<double> gcost;
};

namespace {

// The one-shot plan and the planner buffers of a context being destroyed.
void free_context_plans(pw_ctx *ctx) {
    if (ctx->oneshot) free_plan(ctx, ctx->oneshot);
    ctx->oneshot = nullptr;
    delete ctx->plan_scratch;
    ctx->plan_scratch = nullptr;
}

void orient_plan(const PlanOpts &opt, const pw_plan *P, int o, double per_lane, OrientPlan &R) {
    const int32_t nq = P->nq, nt = P->nt;
    const int64_t n = P->n_pairs;
    const int32_t *pq = P->pq.data(), *pt = P->pt.data(), *qlen = P->qlen.data(), *tlen = P->tlen.data();
    const int32_t ncol = o ? nq : nt;
    auto col_of = [&](int64_t i) { return o ? pq[i] : pt[i]; };
    const int32_t *rlen = o ? tlen : qlen;  // row length of a row sequence
    // pairs grouped by column, in pair order, each with its row sequence (read
    // here, in pair order, instead of gathered later): per-thread column
    // counts, then each thread fills its own slots
    const int nth = host_threads(n, 1 << 16);  // the thread count parallel_ranges(n, 1 << 16, ...) uses
    std::vector<int64_t> cnt((size_t)nth * ncol, 0);
    parallel_ranges(n, 1 << 16, [&](int64_t lo, int64_t hi, int k) {
        int64_t *c = cnt.data() + (size_t)k * ncol;
        for (int64_t i = lo; i < hi; ++i) ++c[col_of(i)];
    });
    R.off.assign((size_t)ncol + 1, 0);
    std::vector<int64_t> pos((size_t)nth * ncol);
    int64_t run = 0;
    for (int32_t c = 0; c < ncol; ++c) {
        R.off[c] = run;
        for (int k = 0; k < nth; ++k) {
            pos[(size_t)k * ncol + c] = run;
            run += cnt[(size_t)k * ncol + c];
        }
    }
    R.off[ncol] = run;
    R.by_col.resize((size_t)n);
    parallel_ranges(n, 1 << 16, [&](int64_t lo, int64_t hi, int k) {
        int64_t *ps = pos.data() + (size_t)k * ncol;
        for (int64_t i = lo; i < hi; ++i) R.by_col[ps[col_of(i)]++] = PairRef{(int32_t)i, o ? pt[i] : pq[i]};
    });
    R.tile.resize((size_t)n);
    R.grp.resize((size_t)n);
    R.split.resize((size_t)n);
    R.ntile.assign((size_t)ncol, 0);
    R.ngrp.assign((size_t)ncol, 0);
    R.nb.assign((size_t)ncol, 16);
    std::vector<double> ccost((size_t)ncol, 0.0);
    // columns in blocks, dynamically over the threads
    std::atomic<int32_t> next_col{0};
    const int cth = host_threads(n, 1 << 15);
    const int32_t step = std::max<int32_t>(1, std::min<int32_t>(16, ncol / (8 * cth)));
    auto work = [&]() {
        std::vector<PairRef> cand, sorted;
        std::vector<int64_t> bucket;
        for (;;) {
            const int32_t c0 = next_col.fetch_add(step);
            if (c0 >= ncol) break;
            for (int32_t cs = c0; cs < std::min(ncol, c0 + step); ++cs) {
                const int64_t base = R.off[cs], m = R.off[cs + 1] - base;
                if (m == 0) continue;
                const PairRef *L = R.by_col.data() + base;
                const int cl = o ? qlen[cs] : tlen[cs];
                double cost = 0;
                int64_t nti = 0, ngr = 0;
                cand.clear();
                int lmin = INT32_MAX, lmax = 0;
                for (int64_t u = 0; u < m; ++u) {
                    const PairRef pr = L[u];
                    const int rl = rlen[pr.row];
                    if (!opt.force_tile && rl <= GROUP_MAX_ROWS) {
                        cand.push_back(pr);
                        lmin = std::min(lmin, rl);
                        lmax = std::max(lmax, rl);
                    } else {
                        cost += TILE_COST_FACTOR * (double)rl * cl;
                        R.tile[base + nti++] = pr;
                    }
                }
                // stable sort by row length, longest first: counting sort over
                // the column's length range when that is small
                const int64_t nc = (int64_t)cand.size();
                if (nc > 1) {
                    const int64_t range = (int64_t)lmax - lmin + 1;
                    if (range <= 4 * nc + 1024) {
                        bucket.assign((size_t)range + 1, 0);
                        for (const PairRef &pr : cand) ++bucket[(size_t)(lmax - rlen[pr.row]) + 1];
                        for (int64_t b = 1; b <= range; ++b) bucket[b] += bucket[b - 1];
                        sorted.resize((size_t)nc);
                        for (const PairRef &pr : cand) sorted[bucket[(size_t)(lmax - rlen[pr.row])]++] = pr;
                        cand.swap(sorted);
                    } else {
                        std::stable_sort(cand.begin(), cand.end(),
                                         [&](const PairRef &x, const PairRef &y) { return rlen[x.row] > rlen[y.row]; });
                    }
                }
                // keep a 64-lane job only if its lanes are busy for most of its
                // rows (sum of row lengths vs 64 x the longest) and its cells per
                // lane stay below a share of the plan's cells per lane of the
                // device: the job runs as one wave, which alone on a SIMD needs as
                // long per cell as the whole SIMD needs per lane-cell, so a longer
                // job would set the makespan (split tiles finish it sooner);
                // otherwise its pairs go to the tile kernel
                const int nb = (opt.force_nb > 0 && cl > 96) ? opt.force_nb : group_nb(cl, opt.tune);
                const int nch = cl > nb ? (cl + nb - 1) / nb : 1;
                R.nb[cs] = nb;
                for (int64_t g0 = 0; g0 < nc; g0 += 64) {
                    const int64_t g1 = std::min(nc, g0 + 64);
                    double sum = 0;
                    for (int64_t u = g0; u < g1; ++u) sum += rlen[cand[u].row];
                    const int top = std::max(1, rlen[cand[g0].row]);
                    const double fill = sum / (64.0 * top);
                    const double job_steps = (double)top * cl;
                    bool keep =
                        opt.force_group || (fill >= GROUP_MIN_FILL && job_steps <= opt.group_max_share * per_lane);
                    // too long for one wave: one item per chunk, pipelined over rows
                    const bool split = !keep && !opt.no_group_split && fill >= GROUP_MIN_FILL && nch > 1 &&
                                       (double)top * nb <= opt.group_max_share * per_lane &&
                                       (double)GROUP_PUB * (nch - 1) <= 2.0 * GROUP_SPLIT_MAX_WAIT * top;
                    keep = keep || split;
                    const double wait = split ? (double)GROUP_PUB * (nch - 1) / (2.0 * top) : 0.0;
                    cost += keep ? 64.0 * job_steps * (1.0 + wait) : TILE_COST_FACTOR * sum * cl;
                    if (keep) {
                        R.split[base + ngr] = split ? 1 : 0;
                        for (int64_t u = g0; u < g1; ++u) R.grp[base + ngr++] = cand[u];
                    } else {
                        for (int64_t u = g0; u < g1; ++u) R.tile[base + nti++] = cand[u];
                    }
                }
                R.ntile[cs] = nti;
                R.ngrp[cs] = ngr;
                ccost[cs] = cost;
            }
        }
    };
    HostPool::get().run(cth, [&](int) { work(); });
    // lane slots of the jobs (64 x longest row x columns) plus the tile
    // kernel's cells at its lower rate per cell, column by column
    R.cost = 0;
    for (int32_t c = 0; c < ncol; ++c) R.cost += ccost[c];
}

// The all-vs-all batch (P->cross): every column holds every row sequence, in
// row order, so the candidates' order (by row length, longest first, stable),
// their 64-lane groups, length sums and longest rows are the same for every
// column; only the choices that depend on the column's length differ.  The
// functions below make exactly the choices and cost sums of orient_plan
// without its per-pair lists.
struct CrossRows {
    std::vector<int32_t> cand;   // candidate rows, sorted
    std::vector<int32_t> other;  // rows too long for a lane-per-pair job, in row order
    std::vector<double> gsum;    // per 64-lane group: sum of its row lengths
    std::vector<int32_t> gtop;   // per group: its longest row length (at least 1)
};

void cross_rows(const PlanOpts &opt, const pw_plan *P, int o, CrossRows &C) {
    const int32_t nrows = o ? P->nt : P->nq;
    const int32_t *rlen = o ? P->tlen.data() : P->qlen.data();
    C.cand.clear();
    C.other.clear();
    for (int32_t r = 0; r < nrows; ++r) {
        if (!opt.force_tile && rlen[r] <= GROUP_MAX_ROWS) C.cand.push_back(r);
        else C.other.push_back(r);
    }
    std::stable_sort(C.cand.begin(), C.cand.end(), [&](int32_t x, int32_t y) { return rlen[x] > rlen[y]; });
    const int64_t nc = (int64_t)C.cand.size();
    C.gsum.clear();
    C.gtop.clear();
    for (int64_t g0 = 0; g0 < nc; g0 += 64) {
        const int64_t g1 = std::min<int64_t>(nc, g0 + 64);
        double sum = 0;
        for (int64_t u = g0; u < g1; ++u) sum += rlen[C.cand[u]];
        C.gsum.push_back(sum);
        C.gtop.push_back(std::max(1, rlen[C.cand[g0]]));
    }
}

// One column of length cl: its cost (summed in orient_plan's order), its
// chunk width, and per group 0 (to the tile kernel), 1 (one wave) or 2 (one
// item per chunk).
double cross_column(const PlanOpts &opt, const CrossRows &C, const int32_t *rlen, int cl, double per_lane, int *nb_out,
                    uint8_t *choice) {
    double cost = 0;
    for (int32_t r : C.other) cost += TILE_COST_FACTOR * (double)rlen[r] * cl;
    const int nb = (opt.force_nb > 0 && cl > 96) ? opt.force_nb : group_nb(cl, opt.tune);
    const int nch = cl > nb ? (cl + nb - 1) / nb : 1;
    for (size_t g = 0; g < C.gsum.size(); ++g) {
        const double sum = C.gsum[g];
        const int top = C.gtop[g];
        const double fill = sum / (64.0 * top);
        const double job_steps = (double)top * cl;
        bool keep = opt.force_group || (fill >= GROUP_MIN_FILL && job_steps <= opt.group_max_share * per_lane);
        const bool split = !keep && !opt.no_group_split && fill >= GROUP_MIN_FILL && nch > 1 &&
                           (double)top * nb <= opt.group_max_share * per_lane &&
                           (double)GROUP_PUB * (nch - 1) <= 2.0 * GROUP_SPLIT_MAX_WAIT * top;
        keep = keep || split;
        const double wait = split ? (double)GROUP_PUB * (nch - 1) / (2.0 * top) : 0.0;
        cost += keep ? 64.0 * job_steps * (1.0 + wait) : TILE_COST_FACTOR * sum * cl;
        if (choice) choice[g] = keep ? (split ? 2 : 1) : 0;
    }
    if (nb_out) *nb_out = nb;
    return cost;
}

// The orientation's cost: per-column costs on host threads, summed in column order.
double cross_cost(const PlanOpts &opt, const pw_plan *P, int o, double per_lane, const CrossRows &C) {
    const int32_t ncol = o ? P->nq : P->nt;
    const int32_t *rlen = o ? P->tlen.data() : P->qlen.data();
    const int32_t *clen = o ? P->qlen.data() : P->tlen.data();
    std::vector<double> cc((size_t)ncol);
    parallel_ranges(ncol, std::max<int64_t>(1, (int64_t)(1 << 14) / std::max<int64_t>(1, (int64_t)C.gsum.size())),
                    [&](int64_t lo, int64_t hi, int) {
                        for (int64_t c = lo; c < hi; ++c)
                            cc[c] = cross_column(opt, C, rlen, clen[c], per_lane, nullptr, nullptr);
                    });
    double cost = 0;
    for (double v : cc) cost += v;
    return cost;
}

// orient_plan's choices for the all-vs-all batch: per column its chunk width,
// its counts of tile-kernel and lane-per-pair pairs and the choice of every
// group (the jobs are built from these and the shared row order).
void cross_plan(const PlanOpts &opt, const pw_plan *P, int o, double per_lane, const CrossRows &C, OrientPlan &R) {
    const int32_t ncol = o ? P->nq : P->nt;
    const int32_t *rlen = o ? P->tlen.data() : P->qlen.data();
    const int32_t *clen = o ? P->qlen.data() : P->tlen.data();
    const int64_t ng = (int64_t)C.gsum.size(), nc = (int64_t)C.cand.size();
    R.ntile.assign((size_t)ncol, 0);
    R.ngrp.assign((size_t)ncol, 0);
    R.nb.assign((size_t)ncol, 16);
    R.choice.resize((size_t)ncol * ng);
    std::vector<double> cc((size_t)ncol);
    parallel_ranges(ncol, std::max<int64_t>(1, (int64_t)(1 << 14) / std::max<int64_t>(1, ng)),
                    [&](int64_t lo, int64_t hi, int) {
        for (int64_t cs = lo; cs < hi; ++cs) {
            int nb = 16;
            uint8_t *ch = R.choice.data() + cs * ng;
            cc[cs] = cross_column(opt, C, rlen, clen[cs], per_lane, &nb, ch);
            R.nb[cs] = nb;
            int64_t kept = 0;
            for (int64_t g = 0; g < ng; ++g)
                if (ch[g]) kept += std::min<int64_t>(nc, g * 64 + 64) - g * 64;
            R.ngrp[cs] = kept;
            R.ntile[cs] = (int64_t)C.other.size() + nc - kept;
        }
    });
    R.cost = 0;
    for (double v : cc) R.cost += v;
}

// Builds the scoring work list of a plan whose inputs are in place (plain
// batches, or lane-per-pair jobs and tile items under the routing choices of opt).
int plan_worklist(const PlanOpts &opt, pw_plan *P) {
    const size_t total_mem = opt.total_mem;
    const int num_cus = opt.num_cus;
    const Scoring &sc = P->sc;
    const int32_t nq = P->nq, nt = P->nt;
    const int64_t n_pairs = P->n_pairs;
    const int32_t mode = P->mode;
    const int32_t *pq = P->pq.data(), *pt = P->pt.data(), *qlen = P->qlen.data(), *tlen = P->tlen.data();
    P->batches.clear();
    P->score_links = P->score_prog = 0;
    P->ckpairs.clear();
    P->ck_floats = 0;
    P->ck_units = 0;
    P->ck_dir_extra = 0;
    P->gitems.clear();
    P->gents.clear();
    P->n_groups = 0;
    P->group_split_items = 0;
    P->orient = 0;
    P->group_cells = 0;
    P->group_rows = 0;
    P->max_nb = 16;
    for (int w = 0; w < NB_KINDS; ++w) P->nb_begin[w] = P->nb_end[w] = 0;

    std::vector<ScoreItem> &items = P->items;
    std::vector<uint8_t> &trans = P->trans;
    int64_t max_b = 0;
    if (mode == 0) {
        items.assign((size_t)n_pairs, ScoreItem{});
        trans.assign((size_t)n_pairs, 0);
        // The reference's batching: consecutive pairs form a batch while the
        // batch's full matrices fit in 80% of device memory (SmithWaterman.py:
        // 202-221, 420-432: 1 direction byte + 4 (linear) or 12 (affine) bytes
        // of FP32 matrix per cell), every pair of a batch padded to the
        // batch's longest query and target rounded up to 8 (:459-461, :325).
        const double budget = 0.8 * (double)total_mem;
        const double bpc = sc.affine ? 13.0 : 5.0;
        int64_t begin = 0;
        int bx = 0, by = 0;
        auto close = [&](int64_t end) {
            for (int64_t k = begin; k < end; ++k)
                items[k] = ScoreItem{pq[k], nq + pt[k], (int32_t)k, 4, bx, by, 0, 1 << 30, -1, 0, -1, 0};
            P->batches.push_back({begin, end});
            max_b = std::max<int64_t>(max_b, by);
        };
        for (int64_t i = 0; i < n_pairs; ++i) {
            const int x = std::max(bx, pad8(qlen[pq[i]])), y = std::max(by, pad8(tlen[pt[i]]));
            if (i > begin && (double)(i - begin + 1) * x * y * bpc > budget) {
                close(i);
                begin = i;
                bx = pad8(qlen[pq[i]]);
                by = pad8(tlen[pt[i]]);
            } else {
                bx = x;
                by = y;
            }
        }
        if (n_pairs > begin) close(n_pairs);
    } else {
        // lane-per-pair jobs for pairs whose column sequence is shared by
        // enough pairs; the rest (long or unshared pairs) go to the
        // wave-per-pair tile kernel.  The column sequence of a job is the
        // target (orient 0, rows = queries) or the query (orient 1, rows =
        // targets); the orientation with the lower estimated cost is used.
        const double per_lane = P->cells / ((double)std::max(1, num_cus) * 4.0 * WAVE);
        std::unique_ptr<PlanScratch> own_scratch;
        PlanScratch *scr = opt.scratch;
        if (!scr) {
            own_scratch.reset(new PlanScratch());
            scr = own_scratch.get();
        }
        std::vector<int64_t> &tile_pairs = scr->tile_pairs;
        std::vector<GroupItem> &gitems = scr->gitems;
        std::vector<GroupEntry> &gents = P->gents;
        std::vector<double> &gcost = scr->gcost;
        OrientPlan *op = scr->op;
        int orient = opt.orient;
        CrossRows cr[2];
        if (P->cross) {
            // the all-vs-all batch: both orientations costed from one sort of
            // each side's lengths, then the chosen one's choices made
            if (orient < 0) {
                cross_rows(opt, P, 1, cr[1]);
                cross_rows(opt, P, 0, cr[0]);
                orient = cross_cost(opt, P, 1, per_lane, cr[1]) < cross_cost(opt, P, 0, per_lane, cr[0]) ? 1 : 0;
            } else {
                cross_rows(opt, P, orient, cr[orient]);
            }
            cross_plan(opt, P, orient, per_lane, cr[orient], op[orient]);
        } else if (orient < 0) {
            orient_plan(opt, P, 1, per_lane, op[1]);
            orient_plan(opt, P, 0, per_lane, op[0]);
            orient = op[1].cost < op[0].cost ? 1 : 0;
        } else {
            orient_plan(opt, P, orient, per_lane, op[orient]);
        }
        const OrientPlan &R = op[orient];
        {
            // jobs of every column in column order, filled on several host
            // threads at offsets from prefix sums over the columns
            const int o = orient;
            const int32_t ncol = o ? nq : nt;
            const int32_t *rlen = o ? tlen : qlen;
            std::vector<int64_t> job0((size_t)ncol + 1, 0), ent0((size_t)ncol + 1, 0), tile0((size_t)ncol + 1, 0);
            for (int32_t cs = 0; cs < ncol; ++cs) {
                job0[cs + 1] = job0[cs] + (R.ngrp[cs] + 63) / 64;
                ent0[cs + 1] = ent0[cs] + R.ngrp[cs];
                tile0[cs + 1] = tile0[cs] + R.ntile[cs];
            }
            gitems.resize((size_t)job0[ncol]);
            gcost.resize((size_t)job0[ncol]);
            gents.resize((size_t)ent0[ncol]);
            tile_pairs.resize((size_t)tile0[ncol]);
            // per-thread results (parallel_ranges runs at most 16 threads)
            std::vector<int64_t> cells_t(16, 0), rows_t(16, 0);
            std::vector<int> nb_t(16, 16);
            parallel_ranges(ncol, std::max<int64_t>(1, (int64_t)ncol * (1 << 16) / std::max<int64_t>(1, n_pairs)),
                            [&](int64_t c_lo, int64_t c_hi, int k) {
                int64_t cells = 0, rows = 0;
                int max_nb = 16;
                for (int64_t cs = c_lo; cs < c_hi; ++cs) {
                    const int cl = o ? qlen[cs] : tlen[cs];
                    const int nb = R.nb[cs];
                    const int nch = cl > nb ? (cl + nb - 1) / nb : 1;
                    int64_t j = job0[cs], e = ent0[cs], tp = tile0[cs];
                    // one 64-lane job of cnt pairs, ref(u) the u-th
                    auto job = [&](int cnt, int split, auto ref) {
                        GroupItem gi;
                        gi.c_seq = o ? (int32_t)cs : nq + (int32_t)cs;
                        gi.entry_off = (int32_t)e;
                        gi.count = cnt;
                        gi.rows = rlen[ref(0).row];
                        gi.nb = nb;
                        gi.trans = o;
                        gi.k0 = 0;
                        gi.k1 = nch;
                        gi.link_in = gi.link_out = -1;
                        gi.prog_in = gi.prog_out = -1;
                        gi.split = split;
                        gi.pad = 0;
                        for (int u = 0; u < cnt; ++u) {
                            const PairRef pr = ref(u);
                            gents[e++] = GroupEntry{o ? nq + pr.row : pr.row, pr.pair};
                            cells += (int64_t)rlen[pr.row] * cl;
                        }
                        gitems[j] = gi;
                        // critical path: the whole job in one wave, or the row pipeline of a split one
                        gcost[j] = gi.split ? (double)(gi.rows + GROUP_PUB * (nch - 1)) * nb : (double)gi.rows * nch * nb;
                        if (nch > 1 && !gi.split) rows = std::max<int64_t>(rows, gi.rows);
                        max_nb = std::max(max_nb, nb);
                        ++j;
                    };
                    if (P->cross) {
                        // the shared row order and this column's group choices
                        const CrossRows &C = cr[o];
                        const int64_t ng = (int64_t)C.gsum.size(), nc = (int64_t)C.cand.size();
                        const uint8_t *ch = R.choice.data() + cs * ng;
                        auto pref = [&](int32_t row) {
                            return PairRef{(int32_t)(o ? cs * nt + row : (int64_t)row * nt + cs), row};
                        };
                        for (int32_t r : C.other) tile_pairs[tp++] = pref(r).pair;
                        for (int64_t g = 0; g < ng; ++g) {
                            const int64_t g0 = g * 64, cnt = std::min<int64_t>(nc, g0 + 64) - g0;
                            const int32_t *rows_g = C.cand.data() + g0;
                            if (!ch[g]) {
                                for (int64_t u = 0; u < cnt; ++u) tile_pairs[tp++] = pref(rows_g[u]).pair;
                                continue;
                            }
                            job((int)cnt, ch[g] == 2 ? 1 : 0, [&](int u) { return pref(rows_g[u]); });
                        }
                    } else {
                        const int64_t base = R.off[cs];
                        for (int64_t u = 0; u < R.ntile[cs]; ++u) tile_pairs[tp++] = R.tile[base + u].pair;
                        const int64_t ng = R.ngrp[cs];
                        for (int64_t g0 = 0; g0 < ng; g0 += 64) {
                            const PairRef *grp = R.grp.data() + base + g0;
                            job((int)std::min<int64_t>(64, ng - g0), R.split[base + g0],
                                [&](int u) { return grp[u]; });
                        }
                    }
                }
                cells_t[k] = cells;
                rows_t[k] = rows;
                nb_t[k] = max_nb;
            });
            int64_t cells = 0;
            for (int k = 0; k < (int)cells_t.size(); ++k) {
                cells += cells_t[k];
                P->group_rows = std::max<int64_t>(P->group_rows, rows_t[k]);
                P->max_nb = std::max(P->max_nb, nb_t[k]);
            }
            P->group_cells = (double)cells;  // an exact integer below 2^53
        }
        P->orient = orient;
        // longest jobs first for the persistent workers
        std::vector<int64_t> &gorder = scr->gorder;
        gorder.resize(gitems.size());
        std::iota(gorder.begin(), gorder.end(), 0);
        std::stable_sort(gorder.begin(), gorder.end(), [&](int64_t x, int64_t y) {
            if (gitems[x].nb != gitems[y].nb) return gitems[x].nb < gitems[y].nb;
            return gcost[x] > gcost[y];
        });
        std::vector<GroupItem> &gjobs = scr->gjobs;
        gjobs.resize(gitems.size());
        for (size_t i = 0; i < gorder.size(); ++i) gjobs[i] = gitems[gorder[i]];
        // split jobs are expanded into their chunk items (below) after the
        // tile items have taken their link rows

        // pairs with several stripes and long rows are split into one item
        // per stripe (consecutive, in stripe order), run by as many waves in a
        // pipeline; the others are one item each
        const int64_t ntile = (int64_t)tile_pairs.size();
        std::vector<ScoreItem> titems((size_t)ntile);
        std::vector<uint8_t> ttrans((size_t)ntile, 0);
        std::vector<int> tsplit((size_t)ntile, 0);
        std::vector<uint8_t> tck((size_t)ntile, 0);
        std::vector<std::pair<double, int64_t>> order((size_t)ntile);
        // long linear pairs are checkpointed: query along the lanes, CK_R
        // residues per lane (the direction pass's shape), split into stripes
        const bool ck_on = opt.checkpoints && !sc.affine && opt.force_R <= 0;
        for (int64_t k = 0; k < ntile; ++k) {
            const int64_t i = tile_pairs[k];
            const int ql = qlen[pq[i]], tl = tlen[pt[i]];
            int tr, R;
            if (ck_on && ql >= CK_MIN_LEN && tl >= CK_MIN_LEN) {
                tr = 0;
                R = CK_R;
                tck[k] = 1;
            } else if (opt.force_R > 0) {
                tr = 0;
                R = opt.force_R;
            } else {
                choose_shape(ql, tl, tr, R, opt.tune);
            }
            const int a = tr ? tl : ql, b = tr ? ql : tl;
            const int aseq = tr ? nq + pt[i] : pq[i], bseq = tr ? pq[i] : nq + pt[i];
            titems[k] = ScoreItem{aseq, bseq, (int32_t)i, R, a, b, 0, 1 << 30, -1, 0, -1, 0};
            ttrans[k] = (uint8_t)tr;
            const int rows = 64 * R;
            const int ns = (a + rows - 1) / rows;
            tsplit[k] = split_stripes(a, b, R);
            // critical path: all stripes in a row, or the pipeline of a split pair
            const double path = tsplit[k] ? (double)(b + 63 + (ns - 1) * 128) : (double)ns * (b + 63);
            order[k] = {-path * (R + 1.5), k};
            max_b = std::max<int64_t>(max_b, b);
        }
        std::stable_sort(order.begin(), order.end(),
                         [](const std::pair<double, int64_t> &x, const std::pair<double, int64_t> &y) {
                             return x.first < y.first;
                         });
        items.clear();
        trans.clear();
        P->score_links = P->score_prog = 0;
        for (int64_t k = 0; k < ntile; ++k) {
            const int64_t o = order[k].second;
            const int ns = tsplit[o];
            if (ns == 0) {
                items.push_back(titems[o]);
                trans.push_back(ttrans[o]);
                continue;
            }
            ScoreItem it = titems[o];
            it.rows_off = P->score_links;
            it.prog_off = P->score_prog;
            P->score_links += (int64_t)(ns - 1) * link_row_floats(it.blen);
            P->score_prog += ns;
            int64_t ck0 = -1, bm0 = 0;
            int nb = 0;
            if (tck[o]) {
                // checkpoints [stripe][block][slot][lane], then block maxima [stripe][block]
                nb = ck_blocks(it.blen);
                ck0 = P->ck_floats;
                bm0 = ck0 + (int64_t)ns * nb * CK_SLOTS * WAVE;
                P->ck_floats = bm0 + (int64_t)ns * nb;
                P->ckpairs.push_back(CkPair{it.out, nb, ck0, bm0, it.rows_off});
                P->ck_units += (int64_t)ns * nb;
                const int ql = qlen[pq[it.out]], tl = tlen[pt[it.out]];
                P->ck_dir_extra += std::max<int64_t>(
                    0, dir_area(ql, tl, CK_R) - dir_area(ql, tl, dir_residues(ql, P->dir_rmax)));
            }
            for (int st = 0; st < ns; ++st) {
                it.s0 = st;
                it.s1 = st + 1;
                if (ck0 >= 0) {
                    it.ck_off = ck0 + (int64_t)st * nb * CK_SLOTS * WAVE;
                    it.bmax_off = bm0 + (int64_t)st * nb;
                }
                items.push_back(it);
                trans.push_back(ttrans[o]);
            }
        }
        std::sort(P->ckpairs.begin(), P->ckpairs.end(),
                  [](const CkPair &x, const CkPair &y) { return x.pair < y.pair; });

        // lane-per-pair jobs in order; a split job becomes one item per chunk,
        // consecutive in chunk order (the persistent workers dequeue in order,
        // so the item a wave waits for is held by a running wave), with a link
        // column of rows x 64 (T, I) pairs and a progress counter per boundary
        P->score_links = (P->score_links + 1) / 2 * 2;  // float2 alignment
        P->gitems.clear();
        for (const GroupItem &j : gjobs) {
            if (!j.split) {
                P->gitems.push_back(j);
                continue;
            }
            const int64_t col = (int64_t)j.rows * WAVE;  // float2 per link column
            const int64_t link0 = P->score_links / 2;
            const int32_t prog0 = (int32_t)P->score_prog;
            for (int k = j.k0; k < j.k1; ++k) {
                GroupItem it = j;
                it.k0 = k;
                it.k1 = k + 1;
                it.link_in = k > j.k0 ? link0 + (int64_t)(k - 1 - j.k0) * col : -1;
                it.prog_in = k > j.k0 ? prog0 + (k - 1 - j.k0) : -1;
                it.link_out = k + 1 < j.k1 ? link0 + (int64_t)(k - j.k0) * col : -1;
                it.prog_out = k + 1 < j.k1 ? prog0 + (k - j.k0) : -1;
                P->gitems.push_back(it);
            }
            P->score_links += 2 * (int64_t)(j.k1 - j.k0 - 1) * col;
            P->score_prog += j.k1 - j.k0 - 1;
            P->group_split_items += j.k1 - j.k0;
        }
        for (int w = 0; w < NB_KINDS; ++w) {
            for (size_t i = 0; i < P->gitems.size(); ++i)
                if (P->gitems[i].nb == NB_WIDTHS[w]) {
                    if (P->nb_end[w] == 0) P->nb_begin[w] = (int64_t)i;
                    P->nb_end[w] = (int64_t)i + 1;
                }
        }
        P->n_groups = (int64_t)P->gitems.size();
    }
    P->max_b = max_b;
    P->n_tile = (int64_t)items.size();
    return PW_OK;
}

// Device half of pw_prepare: uploads everything plan_host built, on stream s,
// with one copy from the pinned staging buffer stage (which must not be in use
// by an earlier copy); sync waits for the copy, so that stage can be reused.
int plan_upload(pw_ctx *ctx, pw_plan *P, hipStream_t s, PinBuf &stage, bool sync) {
    const Encoded &E = P->E;
    const int64_t n = P->n_pairs;
    // residue characters of the alignment rows, by code ('?' beyond the
    // alphabet), and their lower-case forms
    P->chartab.assign(512, 0);
    for (int c = 0; c < 256; ++c) {
        const uint8_t ch = c < (int)E.chars.size() ? E.chars[c] : (uint8_t)'?';
        P->chartab[c] = ch;
        P->chartab[256 + c] = (uint8_t)std::tolower(ch);
    }
    struct Part {
        DevBuf *d;
        const void *h;  // nullptr: the pair list, packed from pq and pt
        size_t bytes;
        size_t off;
    };
    Part parts[] = {{&P->res, E.codes.data(), E.codes.size(), 0},
                    {&P->seqs, E.refs.data(), E.refs.size() * sizeof(SeqRef), 0},
                    {&P->tables, E.tables.data(), E.tables.size() * sizeof(float), 0},
                    {&P->d_items, P->items.data(), P->items.size() * sizeof(ScoreItem), 0},
                    {&P->d_trans, P->trans.data(), P->trans.size(), 0},
                    {&P->d_gitems, P->gitems.data(), P->gitems.size() * sizeof(GroupItem), 0},
                    {&P->d_gents, P->gents.data(), P->gents.size() * sizeof(GroupEntry), 0},
                    {&P->d_mss, P->mss.data(), (size_t)n * sizeof(float), 0},
                    {&P->d_pairs, nullptr, P->cross ? 0 : (size_t)n * sizeof(int2), 0},
                    {&P->d_same, P->same.data(), (size_t)n, 0},
                    {&P->d_ckpairs, P->ckpairs.data(), P->ckpairs.size() * sizeof(CkPair), 0},
                    {&P->d_chars, P->chartab.data(), P->chartab.size(), 0}};
    size_t total = 0;
    for (Part &q : parts) {
        q.off = total;
        total += (q.bytes + 255) / 256 * 256;
    }
    ENSURE(ctx, P->blob, total);
    ENSURE_PINNED(ctx, stage, total);
    for (Part &q : parts) {
        q.d->p = (char *)P->blob.p + q.off;
        q.d->cap = q.bytes;
    }
    // staging on several host threads: each copies its share of every part
    // (and packs its share of the pair list)
    char *st = (char *)stage.p;
    const int32_t *pq = P->pq.data(), *pt = P->pt.data();
    parallel_ranges((int64_t)total, 1 << 20, [&](int64_t lo, int64_t hi, int) {
        for (const Part &q : parts) {
            const int64_t a = std::max<int64_t>(lo, (int64_t)q.off), b = std::min<int64_t>(hi, (int64_t)(q.off + q.bytes));
            if (a >= b) continue;
            if (q.h) {
                std::memcpy(st + a, (const char *)q.h + (a - (int64_t)q.off), (size_t)(b - a));
            } else {
                // whole int2 records: a thread writes every record that starts in its range
                const int64_t r0 = (a - (int64_t)q.off + (int64_t)sizeof(int2) - 1) / (int64_t)sizeof(int2);
                const int64_t r1 = std::min<int64_t>(n, (b - (int64_t)q.off + (int64_t)sizeof(int2) - 1) / (int64_t)sizeof(int2));
                int2 *dst = (int2 *)(st + q.off);
                for (int64_t i = r0; i < r1; ++i) dst[i] = make_int2(pq[i], pt[i]);
            }
        }
    });
    HIPCHK(ctx, hipMemcpyAsync(P->blob.p, stage.p, total, hipMemcpyHostToDevice, s));
    ENSURE(ctx, P->d_links, (size_t)P->score_links * sizeof(float));
    ENSURE(ctx, P->d_prog, (size_t)P->score_prog * sizeof(unsigned));
    ENSURE(ctx, P->d_ck, (size_t)P->ck_floats * sizeof(float));
    if (sync) HIPCHK(ctx, hipStreamSynchronize(s));
    return PW_OK;
}

// Whether the lane-per-pair kernels of P may run the row-shifted affine
// recurrence: it needs whole-number scores and every shifted sum exact in
// FP32.  Shifted values stay within (rows + columns + 2) x (largest score
// magnitude) of zero, so that product is held below 2^22 (FILL padding
// sums are -1e10-sized and never win a maximum in either form).
bool group_shift_exact(const pw_plan *P) {
    if (!P->sc.affine || !P->sc.integral) return false;
    const double span = (double)(P->E.max_len_q + P->E.max_len_t + 2) * (3.0 * (double)P->sc.max_abs + 1.0);
    return span < 4194304.0;
}

template <bool AFF, bool SH>
const void *group_kernel_width(int nb) {
    switch (nb) {
    case 16: return (const void *)group_kernel<AFF, 16, SH>;
    case 24: return (const void *)group_kernel<AFF, 24, SH>;
    case 32: return (const void *)group_kernel<AFF, 32, SH>;
    case 40: return (const void *)group_kernel<AFF, 40, SH>;
    case 48: return (const void *)group_kernel<AFF, 48, SH>;
    case 56: return (const void *)group_kernel<AFF, 56, SH>;
    case 64: return (const void *)group_kernel<AFF, 64, SH>;
    case 72: return (const void *)group_kernel<AFF, 72, SH>;
    case 80: return (const void *)group_kernel<AFF, 80, SH>;
    case 88: return (const void *)group_kernel<AFF, 88, SH>;
    default: return (const void *)group_kernel<AFF, 96, SH>;
    }
}

const void *group_kernel_for(const pw_plan *P, int nb) {
    if (!P->sc.affine) return group_kernel_width<false, false>(nb);
    return group_shift_exact(P) ? group_kernel_width<true, true>(nb) : group_kernel_width<true, false>(nb);
}

// Workgroups of 64 threads resident per CU (cached per kernel and LDS size).
int blocks_per_cu(pw_ctx *ctx, const void *kernel, size_t lds) {
    const auto key = std::make_pair(kernel, lds);
    auto it = ctx->occupancy.find(key);
    if (it != ctx->occupancy.end()) return it->second;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, lds) != hipSuccess || per_cu <= 0)
        per_cu = 8;
    ctx->occupancy[key] = per_cu;
    return per_cu;
}

int persistent_grid(pw_ctx *ctx, const void *kernel, size_t lds, int64_t work, double share = 1.0) {
    const int64_t full = (int64_t)ctx->num_cus * blocks_per_cu(ctx, kernel, lds);
    const int64_t g = std::max<int64_t>(1, (int64_t)(share * (double)full));
    return (int)std::max<int64_t>(1, std::min<int64_t>(work, g));
}

// Per-lane score profile of the checkpointed pairs' scoring pass (small
// alphabets only: nres + 1 slots of 64 lanes x CK_R floats after the
// wave-per-pair kernel's LDS, 10-12 KiB for DNA, so that a CU still holds
// ten or more of its waves); 0: the pass reads the substitution table.
constexpr int PROF_MAX_SLOTS = 6;
int prof_nres(const pw_plan *P) {
    return (!P->ckpairs.empty() && P->E.K + 1 <= PROF_MAX_SLOTS) ? P->E.K : 0;
}
size_t score_lds(const pw_plan *P) {
    const int n = prof_nres(P);
    return lds_bytes(P->E.KP) + (n > 0 ? (size_t)(n + 1) * WAVE * CK_R * sizeof(float) : 0);
}

size_t group_lds(const pw_plan *P, int w) { return (size_t)P->E.KP * (NB_WIDTHS[w] + 4) * sizeof(float); }

Geometry geometry(pw_ctx *ctx, const pw_plan *P) {
    Geometry g;
    const bool aff = P->sc.affine;
    if (P->n_groups > 0) {
        g.gscr_stride = std::max<int64_t>(1, P->group_rows) * WAVE;
        for (int w = 0; w < NB_KINDS; ++w) {
            const int64_t n = P->nb_end[w] - P->nb_begin[w];
            if (n > 0)
                g.group_grid[w] =
                    persistent_grid(ctx, group_kernel_for(P, NB_WIDTHS[w]), group_lds(P, w), n, P->opts.grid_share);
        }
    }
    if (P->n_tile > 0) {
        g.tile_grid = persistent_grid(ctx, (const void *)(aff ? score_kernel<true> : score_kernel<false>),
                                      score_lds(P), P->n_tile, P->opts.grid_share);
        g.tscr_stride = 3 * (P->max_b + 64);
    }
    const int64_t n = P->n_pairs;
    // gate: contiguous ranges of at least 2048 pairs per workgroup (one atomic each)
    g.gate_grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + 2047) / 2048, (int64_t)ctx->num_cus * 4));
    g.build_grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 4));
    // persistent direction-pass workers: as many as the units of every pair
    // passing the gate could use (one per stripe of a split pair), up to the
    // resident limit; split stripes need their producer stripes running
    g.dir_grid = persistent_grid(ctx, (const void *)(aff ? dir_kernel<true> : dir_kernel<false>),
                                 lds_bytes(P->E.KP, true), std::max<int64_t>(1, P->units_worst + P->ck_units),
                                 P->opts.grid_share);
    g.dscr_stride = 3 * (P->max_t + 64);
    g.task_grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + 63) / 64, (int64_t)ctx->num_cus * 4));
    g.wave_grid = long_path_lengths(P->E.max_len_q, P->E.max_len_t)
                      ? (int)std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)ctx->num_cus * 8))
                      : 0;
    return g;
}

int64_t dir_limit(pw_ctx *ctx) {
    return ctx->lim.dir_max > 0 ? ctx->lim.dir_max : (int64_t)(ctx->total_mem / 4);
}

// Candidate slots an execution may hold: their five per-slot records
// (candidate, starting point twice, sort key, hit) within a quarter of the
// direction limit.
int64_t cand_slot_limit(pw_ctx *ctx) {
    const int64_t slot_bytes = (int64_t)(sizeof(Cand) + 2 * sizeof(SpRec) + sizeof(uint64_t) + sizeof(HitRec));
    return std::max<int64_t>(1 << 20, dir_limit(ctx) / 4 / slot_bytes);
}

// Sizes the work areas of W for an execution of P; every allocation happens
// here, before the first launch.
int size_workspace(pw_ctx *ctx, pw_plan *P, Workspace &W, const Geometry &g) {
    const int64_t n = P->n_pairs;
    if (W.cand0 <= 0) W.cand0 = ctx->lim.cand0;
    if (W.want_dir <= 0) W.want_dir = ctx->lim.dir_init;
    if (W.want_ops <= 0) W.want_ops = ctx->lim.ops_init;
    if (W.want_cand <= 0) W.want_cand = (int64_t)W.cand0 * 16384;
    ENSURE(ctx, W.ctl, sizeof(Ctl));
    ENSURE(ctx, W.maxima, (size_t)n * sizeof(float));
    ENSURE(ctx, W.tasks, (size_t)n * sizeof(PairTask));
    ENSURE(ctx, W.touts, (size_t)n * sizeof(TaskOut));
    ENSURE(ctx, W.gated, (size_t)n * sizeof(int));
    ENSURE(ctx, W.rq_slots, (size_t)n * sizeof(int));
    ENSURE(ctx, W.rq_needs, (size_t)n * sizeof(int));
    ENSURE(ctx, W.gsums, (size_t)std::max(1, g.gate_grid) * sizeof(int));
    if (W.dir_ext) {
        W.dir_cap = W.dir_ext_cap;
    } else {
        // the whole worst case when it is at most half the limit, else the
        // first size (grown after an execution that needed more)
        const int64_t worst = P->dir_worst + P->ck_dir_extra;
        const int64_t want = worst <= dir_limit(ctx) / 2 ? std::max(W.want_dir, worst) : W.want_dir;
        const int64_t dcap = std::max<int64_t>(16, std::min(worst, std::min(want, dir_limit(ctx))));
        ENSURE(ctx, W.dirs, (size_t)dcap);
        W.dir_cap = (int64_t)W.dirs.cap;
    }
    const int64_t ccap = std::max<int64_t>(16, std::min<int64_t>(n * (int64_t)W.cand0, W.want_cand));
    ENSURE(ctx, W.cands, (size_t)ccap * sizeof(Cand));
    ENSURE(ctx, W.spslot, (size_t)ccap * sizeof(SpRec));
    ENSURE(ctx, W.keys, (size_t)ccap * sizeof(uint64_t));
    ENSURE(ctx, W.out_sps, (size_t)ccap * sizeof(SpRec));
    ENSURE(ctx, W.out_hits, (size_t)ccap * sizeof(HitRec));
    set_slot_caps(W);
    ENSURE(ctx, W.ops, (size_t)std::max<int64_t>(16, W.want_ops));
    W.ops_cap = (int64_t)W.ops.cap;
    // direction-pass work units: sized for every pair passing the gate; the
    // split pairs' boundary rows up to a quarter of the direction limit (pairs
    // beyond it run unsplit)
    ENSURE(ctx, W.units, (size_t)std::max<int64_t>(1, P->units_worst + P->ck_units) * sizeof(DirUnit));
    W.units_cap = (int64_t)(W.units.cap / sizeof(DirUnit));
    const int64_t lk = std::min<int64_t>(P->dlinks_worst, dir_limit(ctx) / 4 / (int64_t)sizeof(float));
    ENSURE(ctx, W.dlinks, (size_t)std::max<int64_t>(1, lk) * sizeof(float));
    W.dlinks_cap = lk > 0 ? (int64_t)(W.dlinks.cap / sizeof(float)) : 0;
    ENSURE(ctx, W.dprog, (size_t)std::max<int64_t>(1, P->dprog_worst) * sizeof(unsigned));
    W.dprog_cap = (int64_t)(W.dprog.cap / sizeof(unsigned));
    int gmax = 0;
    for (int w = 0; w < NB_KINDS; ++w) gmax = std::max(gmax, g.group_grid[w]);
    if (gmax > 0) ENSURE(ctx, W.gscratch, (size_t)gmax * g.gscr_stride * sizeof(float2));
    if (g.tile_grid > 0) ENSURE(ctx, W.tscratch, (size_t)g.tile_grid * g.tscr_stride * sizeof(float));
    ENSURE(ctx, W.dscratch, (size_t)g.dir_grid * g.dscr_stride * sizeof(float));
    ENSURE_PINNED(ctx, W.h_ctl, sizeof(Ctl));
    return PW_OK;
}

// Scoring pass: queue counters live in the (already cleared) control block.
int launch_score(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, const Geometry &g, bool events,
                 int64_t *launches) {
    const bool aff = P->sc.affine;
    Ctl *ctl = (Ctl *)W.ctl.p;
    if (events) HIPCHK(ctx, hipEventRecord(L.ev[0], L.s));
    const uint8_t *res = (const uint8_t *)P->res.p;
    const SeqRef *seqs = (const SeqRef *)P->seqs.p;
    const float *tabs = (const float *)P->tables.p;
    float *mx = (float *)W.maxima.p;
    // the lane-per-pair kernels of the chunk widths present are spread round
    // robin over the lane's streams (four on a two-stream lane), so that one
    // kernel's tail overlaps the others' work (only the multi-chunk width
    // uses the per-block scratch)
    hipStream_t gstreams[4] = {L.s, L.s2, L.sx[0], L.sx[1]};
    const int ngs = (L.s2 == L.s) ? 1 : (L.sx[0] && L.sx[1] ? 4 : 2);
    bool used[4] = {true, false, false, false};
    bool forked = false;
    auto fork = [&]() -> int {
        if (forked) return PW_OK;
        HIPCHK(ctx, hipEventRecord(L.fork, L.s));
        forked = true;
        return PW_OK;
    };
    auto use = [&](int i) -> int {
        if (used[i]) return PW_OK;
        if (fork()) return PW_EDEVICE;
        HIPCHK(ctx, hipStreamWaitEvent(gstreams[i], L.fork, 0));
        used[i] = true;
        return PW_OK;
    };
    // the fork point is the end of the work issued so far (the clears), taken
    // before the first score kernel goes onto the main stream: the other
    // streams must not wait for that kernel
    if (ngs > 1 && fork()) return PW_EDEVICE;
    if (P->n_groups > 0) {
        // launch order: the two narrowest widths (short jobs, done early) and
        // the two widest (the longest jobs, the critical path) first, then the
        // middle widths widest first, round robin over the streams
        int pres[NB_KINDS], npres = 0;
        for (int w = 0; w < NB_KINDS; ++w)
            if (P->nb_end[w] > P->nb_begin[w]) pres[npres++] = w;
        int order[NB_KINDS], no = 0;
        if (npres >= 4) {
            order[no++] = pres[0];
            order[no++] = pres[1];
            order[no++] = pres[npres - 2];
            order[no++] = pres[npres - 1];
            for (int i = npres - 3; i >= 2; --i) order[no++] = pres[i];
        } else {
            for (int i = 0; i < npres; ++i) order[no++] = pres[i];
        }
        int nk = 0;
        for (int oi = 0; oi < no; ++oi) {
            const int w = order[oi];
            const int64_t n = P->nb_end[w] - P->nb_begin[w];
            const int si = P->mode == 1 ? (nk++ % ngs) : 0;
            if (use(si)) return PW_EDEVICE;
            hipStream_t gs = gstreams[si];
            const void *kern = group_kernel_for(P, NB_WIDTHS[w]);
            const GroupItem *items = (const GroupItem *)P->d_gitems.p + P->nb_begin[w];
            const GroupEntry *ents = (const GroupEntry *)P->d_gents.p;
            int kp = P->E.KP, ni = (int)n;
            float gp = P->sc.gap, e = P->sc.ext;
            float2 *scr = (float2 *)W.gscratch.p;
            int64_t ss = g.gscr_stride;
            unsigned *q = ctl->queue + 16 * (1 + w);
            float2 *lk = (float2 *)P->d_links.p;
            unsigned *pg = (unsigned *)P->d_prog.p;
            unsigned *ab = &ctl->abort_wait;
            unsigned wl = ctx->wait_polls;
            void *args[] = {&res, &seqs, &items, &ni, &ents, &tabs, &kp, &gp, &e, &scr, &ss, &q, &mx, &lk, &pg, &ab, &wl};
            HIPCHK(ctx, hipLaunchKernel(kern, dim3(g.group_grid[w]), dim3(64), args, group_lds(P, w), gs));
            if (launches) *launches += 1;
        }
    }
    if (P->n_tile > 0) {
        auto kern = aff ? score_kernel<true> : score_kernel<false>;
        const size_t lds = score_lds(P);
        const int nres = prof_nres(P);
        const ScoreItem *items = (const ScoreItem *)P->d_items.p;
        const uint8_t *trans = (const uint8_t *)P->d_trans.p;
        if (P->mode == 0) {
            // plain per-batch launches: one launch per batch, static schedule
            for (const auto &b : P->batches) {
                const int64_t n = b.second - b.first;
                const int grid = (int)std::min<int64_t>(n, g.tile_grid);
                hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, L.s, res, seqs, items + b.first, (int)n,
                                   trans + b.first, tabs, P->E.KP, P->E.lkp, P->sc.gap, P->sc.ext,
                                   (float *)W.tscratch.p, g.tscr_stride, (unsigned *)nullptr, mx, (float *)nullptr,
                                   (unsigned *)nullptr, &ctl->abort_wait, ctx->wait_polls, (float *)P->d_ck.p, nres);
                HIPCHK(ctx, hipGetLastError());
                if (launches) *launches += 1;
            }
        } else {
            // persistent workers on the second stream, overlapping the lane-per-pair kernels
            const int si = (P->n_groups > 0 && ngs > 1) ? 1 : 0;
            if (use(si)) return PW_EDEVICE;
            hipStream_t st = gstreams[si];
            hipLaunchKernelGGL(kern, dim3(g.tile_grid), dim3(64), lds, st, res, seqs, items, (int)P->n_tile, trans,
                               tabs, P->E.KP, P->E.lkp, P->sc.gap, P->sc.ext, (float *)W.tscratch.p, g.tscr_stride,
                               ctl->queue, mx, (float *)P->d_links.p, (unsigned *)P->d_prog.p, &ctl->abort_wait,
                               ctx->wait_polls, (float *)P->d_ck.p, nres);
            HIPCHK(ctx, hipGetLastError());
            if (launches) *launches += 1;
        }
    }
    if (forked) {
        hipEvent_t joins[4] = {nullptr, L.join, L.joinx[0], L.joinx[1]};
        for (int i = 1; i < ngs; ++i)
            if (used[i]) {
                HIPCHK(ctx, hipEventRecord(joins[i], gstreams[i]));
                HIPCHK(ctx, hipStreamWaitEvent(L.s, joins[i], 0));
            }
    }
    if (events) HIPCHK(ctx, hipEventRecord(L.ev[1], L.s));
    return PW_OK;
}

// Start selection over the device-side task list: a thread per pair of short
// paths, then (when the plan has pairs whose paths may be long) a wave per
// long pair.
int launch_start_selection(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, const Geometry &g, Ctl *ctl) {
    const bool aff = P->sc.affine;
    const SeqRef *seqs = (const SeqRef *)P->seqs.p;
    PairTask *tasks = (PairTask *)W.tasks.p;
    auto tk = aff ? tb_kernel<true> : tb_kernel<false>;
    hipLaunchKernelGGL(tk, dim3(g.task_grid), dim3(64), 0, L.s, seqs, tasks, (const Ctl *)ctl, dir_ptr(W),
                       (Cand *)W.cands.p, (SpRec *)W.spslot.p);
    HIPCHK(ctx, hipGetLastError());
    if (g.wave_grid > 0) {
        auto tw = aff ? tb_wave_kernel<true> : tb_wave_kernel<false>;
        hipLaunchKernelGGL(tw, dim3(g.wave_grid), dim3(64), WALK_LDS, L.s, seqs, tasks, (const Ctl *)ctl,
                           dir_ptr(W), (Cand *)W.cands.p, (SpRec *)W.spslot.p);
        HIPCHK(ctx, hipGetLastError());
    }
    return PW_OK;
}

// The walk over the device-side task list, split the same way.
int launch_walk(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, const Geometry &g, Ctl *ctl, double minimum) {
    const bool aff = P->sc.affine;
    const uint8_t *res = (const uint8_t *)P->res.p;
    const SeqRef *seqs = (const SeqRef *)P->seqs.p;
    PairTask *tasks = (PairTask *)W.tasks.p;
    auto wk = aff ? walk_kernel<true> : walk_kernel<false>;
    hipLaunchKernelGGL(wk, dim3(g.task_grid), dim3(64), 0, L.s, res, seqs, tasks, ctl, dir_ptr(W),
                       (const SpRec *)W.spslot.p, (uint64_t *)W.keys.p, (SpRec *)W.out_sps.p, (HitRec *)W.out_hits.p,
                       (char *)W.ops.p, W.ops_cap, minimum, (TaskOut *)W.touts.p, W.out_cap);
    HIPCHK(ctx, hipGetLastError());
    if (g.wave_grid > 0) {
        auto ww = aff ? walk_wave_kernel<true> : walk_wave_kernel<false>;
        hipLaunchKernelGGL(ww, dim3(g.wave_grid), dim3(64), WALK_LDS, L.s, res, seqs, tasks, ctl, dir_ptr(W),
                           (const SpRec *)W.spslot.p, (uint64_t *)W.keys.p, (SpRec *)W.out_sps.p,
                           (HitRec *)W.out_hits.p, (char *)W.ops.p, W.ops_cap, minimum, (TaskOut *)W.touts.p,
                           W.out_cap);
        HIPCHK(ctx, hipGetLastError());
    }
    return PW_OK;
}

// Direction pass, start selection and walk over the device-side task list.
int launch_traceback(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, const Geometry &g, double minimum,
                     bool events) {
    const bool aff = P->sc.affine;
    Ctl *ctl = (Ctl *)W.ctl.p;
    PairTask *tasks = (PairTask *)W.tasks.p;
    const uint8_t *res = (const uint8_t *)P->res.p;
    const SeqRef *seqs = (const SeqRef *)P->seqs.p;
    if (!dir_ptr(W) || !W.cands.p || !W.tasks.p || !W.ops.p) {
        ctx->err = "traceback work areas are not allocated";
        return PW_EDEVICE;
    }
    auto dk = aff ? dir_kernel<true> : dir_kernel<false>;
    hipLaunchKernelGGL(dk, dim3(g.dir_grid), dim3(64), lds_bytes(P->E.KP, true), L.s, res, seqs, tasks,
                       (const DirUnit *)W.units.p, ctl, (const float *)P->tables.p, P->E.KP, P->E.lkp, P->sc.gap,
                       P->sc.ext, (float *)W.dscratch.p, g.dscr_stride, dir_ptr(W), (Cand *)W.cands.p,
                       (float *)W.dlinks.p, (unsigned *)W.dprog.p, W.units_cap, ctx->wait_polls,
                       (const CkPair *)P->d_ckpairs.p, (float *)P->d_ck.p, (float *)P->d_links.p);
    HIPCHK(ctx, hipGetLastError());
    if (events) HIPCHK(ctx, hipEventRecord(L.ev[3], L.s));
    int rc = launch_start_selection(ctx, P, W, L, g, ctl);
    if (rc) return rc;
    if (events) HIPCHK(ctx, hipEventRecord(L.ev[4], L.s));
    rc = launch_walk(ctx, P, W, L, g, ctl, minimum);
    if (rc) return rc;
    if (events) HIPCHK(ctx, hipEventRecord(L.ev[5], L.s));
    return PW_OK;
}

int launch_clear(pw_ctx *ctx, void *p, int64_t words, hipStream_t s) {
    if (words <= 0) return PW_OK;
    const int grid = (int)std::min<int64_t>((words + 255) / 256, 1024);
    hipLaunchKernelGGL(clear_kernel, dim3(grid), dim3(256), 0, s, (uint32_t *)p, (int)words);
    HIPCHK(ctx, hipGetLastError());
    return PW_OK;
}

// Zeroes the control block, the split pairs' progress counters and (when
// split pairs combine their stripes' maxima) the pair maxima, with kernels.
int launch_clears(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L) {
    int rc = launch_clear(ctx, W.ctl.p, (int64_t)(sizeof(Ctl) / 4), L.s);
    if (!rc && P->score_prog > 0) rc = launch_clear(ctx, P->d_prog.p, P->score_prog, L.s);
    if (!rc && P->score_prog > 0) rc = launch_clear(ctx, W.maxima.p, P->n_pairs, L.s);
    return rc;
}

// Task kernel over the gated list (first round) or over the list of requeued
// tasks with their read-back slots and candidate needs (requeued).
int launch_tasks(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, const Geometry &g, bool requeued) {
    hipLaunchKernelGGL(task_kernel, dim3(g.build_grid), dim3(256), 0, L.s, (const float *)W.maxima.p,
                       (const float *)P->d_mss.p, (const int2 *)P->d_pairs.p, (const uint8_t *)P->d_same.p,
                       (const SeqRef *)P->seqs.p, P->nq, P->sc.ll, (Ctl *)W.ctl.p, (const int *)W.gated.p,
                       (PairTask *)W.tasks.p, W.dir_cap, W.cand_cap, W.cand0, (DirUnit *)W.units.p, W.units_cap,
                       W.dlinks_cap, (unsigned *)W.dprog.p, W.dprog_cap, P->dir_rmax,
                       (const CkPair *)P->d_ckpairs.p, (int)P->ckpairs.size(), (const float *)P->d_ck.p,
                       P->cross ? (int)P->nt : 0, requeued ? (const int *)W.rq_slots.p : (const int *)nullptr,
                       requeued ? (const int *)W.rq_needs.p : (const int *)nullptr);
    HIPCHK(ctx, hipGetLastError());
    return PW_OK;
}

// The whole device sequence of one execution (capturable: kernels, event
// fork/join and one device-to-host copy into pinned memory).
int launch_pipeline(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, const Geometry &g, double minimum, bool events,
                    int64_t *launches, float *maxima_dst = nullptr) {
    Ctl *ctl = (Ctl *)W.ctl.p;
    int rc = launch_clears(ctx, P, W, L);
    if (rc) return rc;
    rc = launch_score(ctx, P, W, L, g, events, launches);
    if (rc) return rc;
    // the pair maxima go to the host on the second stream while the traceback
    // stages run (a lane with one stream copies them in order)
    const bool copy_beside = maxima_dst && P->n_pairs > 0 && L.s2 != L.s;
    if (copy_beside) {
        HIPCHK(ctx, hipEventRecord(L.fork2, L.s));
        HIPCHK(ctx, hipStreamWaitEvent(L.s2, L.fork2, 0));
        HIPCHK(ctx, hipMemcpyAsync(maxima_dst, W.maxima.p, (size_t)P->n_pairs * sizeof(float),
                                   hipMemcpyDeviceToHost, L.s2));
        HIPCHK(ctx, hipEventRecord(L.join2, L.s2));
    } else if (maxima_dst && P->n_pairs > 0) {
        HIPCHK(ctx, hipMemcpyAsync(maxima_dst, W.maxima.p, (size_t)P->n_pairs * sizeof(float),
                                   hipMemcpyDeviceToHost, L.s));
    }
    if (P->n_pairs > 0) {
        hipLaunchKernelGGL(gate_count_kernel, dim3(g.gate_grid), dim3(256), 0, L.s, (const float *)W.maxima.p,
                           (const float *)P->d_mss.p, P->n_pairs, P->sc.minimum, ctl, (int *)W.gsums.p);
        HIPCHK(ctx, hipGetLastError());
        hipLaunchKernelGGL(gate_scatter_kernel, dim3(g.gate_grid), dim3(256), 0, L.s, (const float *)W.maxima.p,
                           (const float *)P->d_mss.p, P->n_pairs, P->sc.minimum, (const int *)W.gsums.p,
                           (int *)W.gated.p);
        HIPCHK(ctx, hipGetLastError());
        int rc2 = launch_tasks(ctx, P, W, L, g, false);
        if (rc2) return rc2;
    }
    if (events) HIPCHK(ctx, hipEventRecord(L.ev[2], L.s));
    rc = launch_traceback(ctx, P, W, L, g, minimum, events);
    if (rc) return rc;
    if (copy_beside) HIPCHK(ctx, hipStreamWaitEvent(L.s, L.join2, 0));
    HIPCHK(ctx, hipMemcpyAsync(W.h_ctl.p, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, L.s));
    return PW_OK;
}


// Finished traceback work of one gated pair: where its records sit in the
// read-back arrays.
struct PairOut {
    int32_t pair;
    int32_t nsp;
    const SpRec *sps;
    const HitRec *hits;
    const char *ops;
    int64_t nops;
};

// Host copies of the records of the tail rounds (kept alive until assembly).
struct TailStore {
    std::vector<std::vector<SpRec>> sps;
    std::vector<std::vector<HitRec>> hits;
    std::vector<std::vector<char>> ops;
    // copies the records of outs[from, end) out of the staging buffers, one
    // block per kind, and points those outs at the copies
    void keep(std::vector<PairOut> &outs, size_t from) {
        size_t nsp = 0, nops = 0;
        for (size_t i = from; i < outs.size(); ++i) {
            nsp += (size_t)outs[i].nsp;
            nops += (size_t)outs[i].nops;
        }
        sps.emplace_back(nsp);
        hits.emplace_back(nsp);
        ops.emplace_back(nops);
        SpRec *ps = sps.back().data();
        HitRec *ph = hits.back().data();
        char *po = ops.back().data();
        for (size_t i = from; i < outs.size(); ++i) {
            PairOut &o = outs[i];
            if (o.nsp > 0) {
                std::memcpy(ps, o.sps, (size_t)o.nsp * sizeof(SpRec));
                std::memcpy(ph, o.hits, (size_t)o.nsp * sizeof(HitRec));
            }
            if (o.nops > 0) std::memcpy(po, o.ops, (size_t)o.nops);
            o.sps = ps;
            o.hits = ph;
            o.ops = po;
            ps += o.nsp;
            ph += o.nsp;
            po += o.nops;
        }
    }
};

int take_maxima_buffer(pw_ctx *ctx, int64_t n, HostResult *R) {
    if (n <= 0) return PW_OK;
    for (size_t i = 0; i < ctx->pinned_free.size(); ++i)
        if (ctx->pinned_free[i].second >= (size_t)n) {
            R->maxima = ctx->pinned_free[i].first;
            R->maxima_cap = ctx->pinned_free[i].second;
            ctx->pinned_free.erase(ctx->pinned_free.begin() + i);
            return PW_OK;
        }
    HIPCHK(ctx, hipHostMalloc((void **)&R->maxima, (size_t)n * sizeof(float), hipHostMallocDefault));
    R->maxima_cap = (size_t)n;
    return PW_OK;
}

// Reads the tasks and compacted records of a finished device sequence whose
// control block is in W.h_ctl; finished pairs go to outs, the others to pending.
// After device rounds (n_last >= 0) the summaries cover every slot of the
// first round, and the tasks still unfinished are those of the last round
// (tasks[0, n_last)) that did not finish; otherwise they are tasks[i] for the
// summaries i that did not finish.
int read_tasks(pw_ctx *ctx, Workspace &W, Lane &L, int ntasks, std::vector<PairOut> &outs,
               std::vector<PairTask> &pending, int n_last = -1) {
    const Ctl &c = *(const Ctl *)W.h_ctl.p;
    if (ntasks <= 0) return PW_OK;
    const int64_t nout = std::min<int64_t>((int64_t)c.out_used, W.out_cap);
    const int64_t nops = std::min<int64_t>((int64_t)c.ops_used, W.ops_cap);
    // per-task summaries (32 bytes; the whole 120-byte tasks only when some
    // did not finish and go to a tail round)
    ENSURE_PINNED(ctx, W.h_touts, (size_t)ntasks * sizeof(TaskOut));
    ENSURE_PINNED(ctx, W.h_sps, (size_t)nout * sizeof(SpRec));
    ENSURE_PINNED(ctx, W.h_hits, (size_t)nout * sizeof(HitRec));
    ENSURE_PINNED(ctx, W.h_ops, (size_t)nops);
    HIPCHK(ctx, hipMemcpyAsync(W.h_touts.p, W.touts.p, (size_t)ntasks * sizeof(TaskOut), hipMemcpyDeviceToHost,
                               L.s));
    if (nout > 0) {
        HIPCHK(ctx, hipMemcpyAsync(W.h_sps.p, W.out_sps.p, (size_t)nout * sizeof(SpRec), hipMemcpyDeviceToHost, L.s));
        HIPCHK(ctx, hipMemcpyAsync(W.h_hits.p, W.out_hits.p, (size_t)nout * sizeof(HitRec), hipMemcpyDeviceToHost,
                                   L.s));
    }
    if (nops > 0) HIPCHK(ctx, hipMemcpyAsync(W.h_ops.p, W.ops.p, (size_t)nops, hipMemcpyDeviceToHost, L.s));
    HIPCHK(ctx, hipStreamSynchronize(L.s));
    const TaskOut *T = (const TaskOut *)W.h_touts.p;
    bool unfinished = false;
    for (int i = 0; i < ntasks && !unfinished; ++i) unfinished = T[i].status != TASK_OK;
    const PairTask *full = nullptr;
    if (unfinished) {
        const int nfull = n_last >= 0 ? n_last : ntasks;
        ENSURE_PINNED(ctx, W.h_tasks, (size_t)std::max(1, nfull) * sizeof(PairTask));
        HIPCHK(ctx, hipMemcpyAsync(W.h_tasks.p, W.tasks.p, (size_t)nfull * sizeof(PairTask), hipMemcpyDeviceToHost,
                                   L.s));
        HIPCHK(ctx, hipStreamSynchronize(L.s));
        full = (const PairTask *)W.h_tasks.p;
        if (n_last >= 0)
            for (int i = 0; i < n_last; ++i)
                if (full[i].status != TASK_OK) pending.push_back(full[i]);
    }
    for (int i = 0; i < ntasks; ++i) {
        const TaskOut &t = T[i];
        if (t.status != TASK_OK) {
            if (n_last < 0) pending.push_back(full[i]);
            continue;
        }
        if ((int64_t)t.out_off + t.nsp > nout || t.ops_off + t.nops > nops) {
            ctx->err = "traceback records outside the read-back range";
            return PW_EDEVICE;
        }
        outs.push_back(PairOut{t.pair, t.nsp, (const SpRec *)W.h_sps.p + t.out_off,
                               (const HitRec *)W.h_hits.p + t.out_off, (const char *)W.h_ops.p + t.ops_off, t.nops});
    }
    return PW_OK;
}

double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms;
}

// Tail rounds: pairs that did not fit the work areas (or overflowed their
// candidate or ops ranges) run again, host-driven, with exact sizes.
int tail_rounds(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, const Geometry &g, double minimum,
                std::vector<PairTask> &pending, std::vector<PairOut> &outs, TailStore &store, double *ms_dir,
                double *ms_tb, double *ms_walk) {
    // Pairs are taken in order from `pending` (a cursor, no copies of the rest),
    // after the pairs of the previous round that overflowed (`retry`).  Every
    // round takes at least its first pair (the areas are sized for it); a pair
    // goes round again only after an overflow, then first and with its exact
    // candidate count, so the round after completes it.  Two rounds in a row
    // that complete no pair therefore mean a sizing mistake and end the call.
    int idle_rounds = 0;
    const int64_t cand_limit = cand_slot_limit(ctx);
    size_t cursor = 0;
    std::vector<PairTask> retry;
    auto area_of = [&](const PairTask &t) { return dir_area(P->qlen[P->pq[t.pair]], P->tlen[P->pt[t.pair]], t.R); };
    // exact candidate capacity where the count is known (an overflowed pair)
    auto exact_cap = [](PairTask &t) {
        if (t.status == TASK_CAND_OVERFLOW || t.status == TASK_OPS_OVERFLOW) t.cand_cap = std::max(t.ncand, 1);
        else t.cand_cap = std::max(t.cand_cap, 1);
    };
    for (PairTask &t : pending) exact_cap(t);
    while (cursor < pending.size() || !retry.empty()) {
        if (idle_rounds >= 2) {
            ctx->err = "traceback tail rounds did not converge";
            return PW_EDEVICE;
        }
        for (PairTask &t : retry) exact_cap(t);
        const size_t nret = retry.size();
        auto at = [&](size_t k) -> PairTask & { return k < nret ? retry[k] : pending[cursor + (k - nret)]; };
        const size_t avail = nret + (pending.size() - cursor);
        // areas large enough for the first pair, and grown toward what the next
        // pairs need (up to the limits), so that a large deferred set runs in
        // few rounds; then take pairs in order while they fit
        const PairTask &f = at(0);
        const int64_t farea = area_of(f);
        if (W.dir_ext && farea > W.dir_cap) {
            // a lent area too small for this pair: continue in an own one
            W.dir_ext = nullptr;
            W.dir_ext_cap = 0;
            W.dir_cap = 0;
        }
        int64_t dneed = 0, cneed = 0;
        for (size_t k = 0; k < avail && (dneed < dir_limit(ctx) || cneed < cand_limit); ++k) {
            dneed += area_of(at(k));
            cneed += at(k).cand_cap;
        }
        if (!W.dir_ext) {
            const int64_t want = std::max<int64_t>({farea, W.dir_cap, std::min(dneed, dir_limit(ctx))});
            if (want > W.dir_cap) {
                ENSURE(ctx, W.dirs, (size_t)want);
                W.dir_cap = (int64_t)W.dirs.cap;
            }
        }
        // a round's records are at most its candidates: both areas take cwant
        const int64_t cwant = std::max<int64_t>((int64_t)f.cand_cap, std::min(cneed, cand_limit));
        if (cwant > std::min(W.cand_cap, W.out_cap)) {
            const size_t c = (size_t)cwant;
            ENSURE(ctx, W.cands, c * sizeof(Cand));
            ENSURE(ctx, W.spslot, c * sizeof(SpRec));
            ENSURE(ctx, W.keys, c * sizeof(uint64_t));
            ENSURE(ctx, W.out_sps, c * sizeof(SpRec));
            ENSURE(ctx, W.out_hits, c * sizeof(HitRec));
            set_slot_caps(W);
        }
        const int64_t slot_cap = std::min(W.cand_cap, W.out_cap);
        std::vector<PairTask> chunk;
        int64_t dused = 0, cused = 0;
        const int64_t max_take = std::min<int64_t>(P->n_pairs, W.units_cap);
        size_t k = 0;
        for (; k < avail && (int64_t)chunk.size() < max_take; ++k) {
            PairTask t = at(k);
            const int64_t a = area_of(t);
            if (dused + a > W.dir_cap || cused + t.cand_cap > slot_cap) break;
            t.dir_off = dused;
            t.cand_off = cused;
            t.status = TASK_OK;
            t.ncand = 0;
            t.nsp = 0;
            t.ops_off = 0;
            t.nops = 0;
            t.out_off = 0;
            dused += a;
            cused += t.cand_cap;
            chunk.push_back(t);
        }
        // the taken pairs leave retry and pending
        if (k >= nret) {
            cursor += k - nret;
            retry.clear();
        } else {
            retry.erase(retry.begin(), retry.begin() + (std::ptrdiff_t)k);
        }
        const int nc = (int)chunk.size();
        std::vector<DirUnit> units((size_t)nc);
        for (int u = 0; u < nc; ++u) {
            chunk[u].split = 0;  // tail rounds run each pair's stripes in one unit
            chunk[u].gslot = u;  // read back by this round's summaries
            units[u] = DirUnit{u, -1, -1};
        }
        Ctl c;
        std::memset(&c, 0, sizeof c);
        c.ngated = nc;
        c.units_used = nc;
        HIPCHK(ctx, hipMemcpyAsync(W.tasks.p, chunk.data(), (size_t)nc * sizeof(PairTask), hipMemcpyHostToDevice,
                                   L.s));
        HIPCHK(ctx, hipMemcpyAsync(W.units.p, units.data(), (size_t)nc * sizeof(DirUnit), hipMemcpyHostToDevice,
                                   L.s));
        HIPCHK(ctx, hipMemcpyAsync(W.ctl.p, &c, sizeof c, hipMemcpyHostToDevice, L.s));
        const bool aff = P->sc.affine;
        Ctl *dctl = (Ctl *)W.ctl.p;
        PairTask *tasks = (PairTask *)W.tasks.p;
        const uint8_t *res = (const uint8_t *)P->res.p;
        const SeqRef *seqs = (const SeqRef *)P->seqs.p;
        if (!dir_ptr(W)) {
            ctx->err = "direction area is not allocated";
            return PW_EDEVICE;
        }
        HIPCHK(ctx, hipEventRecord(L.ev[2], L.s));
        auto dk = aff ? dir_kernel<true> : dir_kernel<false>;
        hipLaunchKernelGGL(dk, dim3(g.dir_grid), dim3(64), lds_bytes(P->E.KP, true), L.s, res, seqs, tasks,
                           (const DirUnit *)W.units.p, dctl, (const float *)P->tables.p, P->E.KP, P->E.lkp, P->sc.gap,
                           P->sc.ext, (float *)W.dscratch.p, g.dscr_stride, dir_ptr(W), (Cand *)W.cands.p,
                           (float *)W.dlinks.p, (unsigned *)W.dprog.p, W.units_cap, ctx->wait_polls,
                           (const CkPair *)P->d_ckpairs.p, (float *)P->d_ck.p, (float *)P->d_links.p);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(L.ev[3], L.s));
        int rc0 = launch_start_selection(ctx, P, W, L, g, dctl);
        if (rc0) return rc0;
        HIPCHK(ctx, hipEventRecord(L.ev[4], L.s));
        // exact ops capacity from the starting-point counts
        std::vector<PairTask> after((size_t)nc);
        HIPCHK(ctx, hipMemcpyAsync(after.data(), W.tasks.p, (size_t)nc * sizeof(PairTask), hipMemcpyDeviceToHost,
                                   L.s));
        HIPCHK(ctx, hipStreamSynchronize(L.s));
        *ms_dir += elapsed(L.ev[2], L.ev[3]);
        *ms_tb += elapsed(L.ev[3], L.ev[4]);
        int64_t ops_need = 0;
        for (const PairTask &t : after)
            if (t.status == TASK_OK && t.nsp > 0) {
                const int64_t ql = P->qlen[P->pq[t.pair]], tl = P->tlen[P->pt[t.pair]];
                ops_need += std::min<int64_t>((int64_t)t.nsp * (ql + tl), ql * tl + t.nsp);
            }
        if (ops_need > W.ops_cap) {
            ENSURE(ctx, W.ops, (size_t)ops_need);
            W.ops_cap = (int64_t)W.ops.cap;
        }
        HIPCHK(ctx, hipEventRecord(L.ev[4], L.s));
        rc0 = launch_walk(ctx, P, W, L, g, dctl, minimum);
        if (rc0) return rc0;
        HIPCHK(ctx, hipEventRecord(L.ev[5], L.s));
        HIPCHK(ctx, hipMemcpyAsync(W.h_ctl.p, W.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, L.s));
        HIPCHK(ctx, hipStreamSynchronize(L.s));
        *ms_walk += elapsed(L.ev[4], L.ev[5]);
        std::vector<PairOut> got;
        std::vector<PairTask> again;
        int rc = read_tasks(ctx, W, L, nc, got, again);
        if (rc) return rc;
        idle_rounds = got.empty() ? idle_rounds + 1 : 0;
        // copy this round's records out of the staging buffers the next round reuses
        const size_t first = outs.size();
        outs.insert(outs.end(), got.begin(), got.end());
        store.keep(outs, first);
        // overflowed pairs first: the next round takes them with exact sizes
        again.insert(again.end(), retry.begin(), retry.end());
        retry.swap(again);
    }
    return PW_OK;
}

// Sizes for the next execution after one that overflowed its work areas.
// c: the first round's control block (what the direction and candidate
// areas were asked for); last: that of the last device round (ops asked for
// over every round).
void adapt_workspace(pw_ctx *ctx, Workspace &W, const Ctl &c, const Ctl &last,
                     const std::vector<PairTask> &pending) {
    if (!ctx->lim.grow) return;
    W.want_dir = std::min<int64_t>(std::max<int64_t>(W.want_dir, (int64_t)c.dir_used), dir_limit(ctx));
    int32_t maxc = W.cand0;
    for (const PairTask &t : pending)
        if (t.status == TASK_CAND_OVERFLOW) maxc = std::max(maxc, t.ncand);
    W.cand0 = std::min<int32_t>(maxc, 1024);
    W.want_cand = std::min(std::max<int64_t>({W.want_cand, (int64_t)c.cand_used, (int64_t)c.ngated * W.cand0}),
                           cand_slot_limit(ctx));
    W.want_ops = std::max<int64_t>({W.want_ops, (int64_t)c.ops_used, (int64_t)last.ops_used});
}

// Per-pair records in pair order into the host result.
// Sorts finished pairs by pair index: a stable LSD radix sort on
// (pair << 32 | position) keys, one byte of the pair index per pass.
void sort_by_pair(std::vector<PairOut> &outs) {
    const size_t n = outs.size();
    bool sorted = true;
    for (size_t i = 1; i < n && sorted; ++i) sorted = outs[i - 1].pair < outs[i].pair;
    if (sorted) return;
    std::vector<uint64_t> keys(n), tmp(n);
    uint32_t maxp = 0;
    for (size_t i = 0; i < n; ++i) {
        keys[i] = ((uint64_t)(uint32_t)outs[i].pair << 32) | (uint64_t)i;
        maxp = std::max(maxp, (uint32_t)outs[i].pair);
    }
    for (int shift = 32; shift < 64 && (maxp >> (shift - 32)) != 0; shift += 8) {
        size_t cnt[257] = {};
        for (uint64_t k : keys) ++cnt[((k >> shift) & 255u) + 1];
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint64_t k : keys) tmp[cnt[(k >> shift) & 255u]++] = k;
        keys.swap(tmp);
    }
    std::vector<PairOut> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = outs[keys[i] & 0xffffffffu];
    outs.swap(out);
}

// Reports how many starting points the reference's per-pair cap would drop:
// it keeps at most sp_cap starting points per pair of a batch in total
// (SmithWaterman.py:683-685).
int64_t sp_over_cap(int64_t nsp, int32_t sp_cap, int64_t n_pairs) {
    if (sp_cap <= 0) return 0;
    const int64_t cap_total = (int64_t)sp_cap * n_pairs;
    return nsp >= cap_total ? nsp - cap_total : 0;
}

int assemble(pw_ctx *ctx, std::vector<PairOut> &outs, int32_t sp_cap, int64_t n_pairs, const pw_plan *P,
             HostResult *R) {
    sort_by_pair(outs);
    int64_t nsp = 0, nops = 0;
    for (const PairOut &o : outs) {
        nsp += o.nsp;
        nops += o.nops;
    }
    std::vector<void *> *dead = ctx->defer_free ? &ctx->dead_pinned : nullptr;
    if (!R->sps.reserve((size_t)nsp, false, dead) || !R->hits.reserve((size_t)nsp, false, dead) ||
        !R->ops.reserve((size_t)nops, false, dead) || !R->aligned.reserve((size_t)(3 * nops), false, dead)) {
        ctx->err = "no host memory for the result records";
        return PW_ENOMEM;
    }
    R->n_sps = nsp;
    R->n_ops = nops;
    static_assert(sizeof(SpRec) == sizeof(pw_starting_point), "starting point layout");
    const Encoded &E = P->E;
    const int32_t nq = P->nq;
    // residue characters by code, and their lower-case forms for the match test
    // of the alignment row (SmithWaterman.py:794-799 compares lower-cased)
    char chr[256], low[256];
    for (int c = 0; c < 256; ++c) {
        chr[c] = c < (int)E.chars.size() ? (char)E.chars[c] : '?';
        low[c] = (char)std::tolower((unsigned char)chr[c]);
    }
    // records of pairs [b, e) of outs, starting at sp0 / op0 in the result
    auto fill = [R, &outs, &E, nq, &chr, &low](size_t b, size_t e, int64_t sp0, int64_t op0) {
        for (size_t i = b; i < e; ++i) {
            const PairOut &o = outs[i];
            if (o.nsp > 0) std::memcpy(&R->sps[sp0], o.sps, (size_t)o.nsp * sizeof(SpRec));
            if (o.nops > 0) std::memcpy(&R->ops[op0], o.ops, (size_t)o.nops);
            for (int k = 0; k < o.nsp; ++k) {
                const HitRec &hr = o.hits[k];
                pw_hit &h = R->hits[sp0 + k];
                h.query = hr.query;
                h.target = hr.target;
                h.shown = hr.shown;
                h.s_start = hr.s_start;
                h.s_end = hr.s_end;
                h.t_start = hr.t_start;
                h.t_end = hr.t_end;
                h.matches = hr.matches;
                h.mismatches = hr.mismatches;
                h.length = hr.length;
                h.score = hr.score;
                h.sp_index = (int32_t)(sp0 + hr.sp_index);
                h.ops_offset = op0 + hr.ops_offset;
                // the three rows of SmithWaterman.py:777-830 (query, alignment,
                // target), start to end: the op codes run from the end back
                if (hr.shown && hr.length > 0) {
                    const SeqRef qr = E.refs[hr.query], tr = E.refs[nq + hr.target];
                    const uint8_t *qc = E.codes.data() + qr.off, *tc = E.codes.data() + tr.off;
                    const char *op = o.ops + hr.ops_offset + hr.length - 1;
                    char *qrow = R->aligned.data() + 3 * h.ops_offset, *arow = qrow + hr.length, *trow = arow + hr.length;
                    int x = hr.s_start, y = hr.t_start;
                    for (int c = 0; c < hr.length; ++c) {
                        const char v = op[-c];
                        if (v == 'D') {
                            const uint8_t a = qc[x++], t = tc[y++];
                            qrow[c] = chr[a];
                            trow[c] = chr[t];
                            arow[c] = low[a] == low[t] ? '|' : '.';
                        } else if (v == 'L') {
                            qrow[c] = chr[qc[x++]];
                            trow[c] = '-';
                            arow[c] = ' ';
                        } else {
                            qrow[c] = '-';
                            trow[c] = chr[tc[y++]];
                            arow[c] = ' ';
                        }
                    }
                }
            }
            sp0 += o.nsp;
            op0 += o.nops;
        }
    };
    // large results are filled by the host pool, each thread a contiguous range
    // of pairs of about equal cost at its precomputed offsets (the same records
    // as one thread); the fill costs ~30 ns per pair plus ~0.1 ns per record
    // byte and ~1 ns per alignment column
    const size_t n = outs.size();
    std::vector<double> cost(n + 1, 0.0);
    std::vector<int64_t> sp_at(n + 1, 0), op_at(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        const PairOut &o = outs[i];
        cost[i + 1] = cost[i] + 300.0 + 11.0 * (double)o.nops + (double)o.nsp * (sizeof(SpRec) + sizeof(pw_hit));
        sp_at[i + 1] = sp_at[i] + o.nsp;
        op_at[i + 1] = op_at[i] + o.nops;
    }
    const double total = cost[n];  // units of 0.1 ns
    const int nth = (int)std::max<int64_t>(1, std::min<int64_t>(HostPool::get().size(), (int64_t)(total / 5e5)));
    if (nth == 1) {
        fill(0, n, 0, 0);
    } else {
        auto cut = [&](int k) {
            return (size_t)(std::lower_bound(cost.begin(), cost.end(), total * k / nth) - cost.begin());
        };
        HostPool::get().run(nth, [&](int k) {
            const size_t b = k == 0 ? 0 : cut(k), e = k == nth - 1 ? n : cut(k + 1);
            if (e > b) fill(b, e, sp_at[b], op_at[b]);
        });
    }
    R->pub.sp_truncated = sp_over_cap(nsp, sp_cap, n_pairs);
    return PW_OK;
}

// Launch parameters a captured graph depends on.
std::vector<uint64_t> graph_signature(const pw_plan *P, const Workspace &W, const Lane &L, const Geometry &g,
                                      double minimum, unsigned wait_polls) {
    std::vector<uint64_t> s;
    auto add = [&](uint64_t v) { s.push_back(v); };
    add(wait_polls);
    const DevBuf *bufs[] = {&W.ctl,      &W.maxima,   &W.gated,    &W.tasks,    &W.touts,    &W.dirs,
                            &W.cands,    &W.spslot,   &W.keys,     &W.out_sps,  &W.out_hits, &W.ops,
                            &W.gscratch, &W.tscratch, &W.dscratch, &W.units,    &W.dlinks,   &W.dprog,
                            &P->d_links, &P->d_prog,   &P->d_ck,    &P->d_ckpairs, &W.gsums};
    for (const DevBuf *b : bufs) {
        add((uint64_t)(uintptr_t)b->p);
        add(b->cap);
    }
    add((uint64_t)(uintptr_t)W.h_ctl.p);
    add((uint64_t)(uintptr_t)dir_ptr(W));
    add((uint64_t)W.dir_cap);
    add((uint64_t)W.cand_cap);
    add((uint64_t)W.out_cap);
    add((uint64_t)W.ops_cap);
    add((uint64_t)W.cand0);
    add((uint64_t)W.units_cap);
    add((uint64_t)W.dlinks_cap);
    add((uint64_t)W.dprog_cap);
    add((uint64_t)(uintptr_t)L.s);
    add((uint64_t)(uintptr_t)L.s2);
    uint64_t mb;
    std::memcpy(&mb, &minimum, sizeof mb);
    add(mb);
    for (int w = 0; w < NB_KINDS; ++w) add((uint64_t)g.group_grid[w]);
    add((uint64_t)g.tile_grid);
    add((uint64_t)g.dir_grid);
    add((uint64_t)g.task_grid);
    add((uint64_t)g.wave_grid);
    (void)P;
    return s;
}

// State of one execution between its launch and its read-back.
struct Exec {
    Geometry g;
    int64_t launches = 0;
    bool graph = false;
    double minimum = 0;
    std::chrono::steady_clock::time_point wall0;
};

// Results of a finished execution before they are split into host records.
struct Finished {
    std::vector<PairOut> outs;  // sorted by pair
    TailStore store;
    int ng = 0;
    double ms_score = 0, ms_dir = 0, ms_tb = 0, ms_walk = 0;
    double ms_readback = 0, ms_tail = 0, ms_assemble = 0;  // host phases
    double ms_device_rounds = 0;                           // further rounds on the device (wall)
    bool on_device = false;  // every pair finished on the device: the result is laid out there (outs empty)
};

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Sizes the work areas and issues the device work of one execution on L, the
// pair maxima going to maxima_dst (pinned host memory, n_pairs floats).  Nothing
// waits for the device.
int execute_launch(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, double minimum, bool use_graph,
                   float *maxima_dst, Exec &X) {
    X.wall0 = std::chrono::steady_clock::now();
    X.graph = use_graph;
    X.minimum = minimum;
    X.launches = 0;
    X.g = geometry(ctx, P);
    const Geometry &g = X.g;
    int rc = size_workspace(ctx, P, W, g);
    if (rc) return rc;
    if (use_graph) {
        std::vector<uint64_t> sig = graph_signature(P, W, L, g, minimum, ctx->wait_polls);
        if (!P->graph || sig != P->graph_sig) {
            if (P->graph) {
                (void)hipGraphExecDestroy(P->graph);
                P->graph = nullptr;
            }
            hipGraph_t graph = nullptr;
            HIPCHK(ctx, hipStreamBeginCapture(L.s, hipStreamCaptureModeThreadLocal));
            rc = launch_pipeline(ctx, P, W, L, g, minimum, false, &X.launches);
            hipError_t e = hipStreamEndCapture(L.s, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) {
                ctx->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
                return PW_EDEVICE;
            }
            e = hipGraphInstantiate(&P->graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                P->graph = nullptr;
                ctx->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
                return PW_EDEVICE;
            }
            P->graph_sig = sig;
        } else {
            for (int w = 0; w < NB_KINDS; ++w) X.launches += P->nb_end[w] > P->nb_begin[w] ? 1 : 0;
            X.launches += P->n_tile > 0 ? (P->mode == 0 ? (int64_t)P->batches.size() : 1) : 0;
        }
        HIPCHK(ctx, hipEventRecord(L.ev[0], L.s));
        HIPCHK(ctx, hipGraphLaunch(P->graph, L.s));
        HIPCHK(ctx, hipEventRecord(L.ev[5], L.s));
        // the destination differs per call, so the maxima copy stays out of the graph
        if (P->n_pairs > 0)
            HIPCHK(ctx, hipMemcpyAsync(maxima_dst, W.maxima.p, (size_t)P->n_pairs * sizeof(float),
                                       hipMemcpyDeviceToHost, L.s));
    } else {
        rc = launch_pipeline(ctx, P, W, L, g, minimum, true, &X.launches, maxima_dst);
        if (rc) return rc;
    }
    return PW_OK;
}

// Waits for the launched execution, reads its tasks and records back and runs
// the tail rounds of the pairs that did not fit.
// Rounds after the first on the device: requeue_kernel lists the unfinished
// tasks, round_reset_kernel restarts the allocators, then the task kernel
// (over that list), direction pass, start selection and walk run as in the
// first round.  Stops when every task finished, when a round finished none
// (n_last: that round's task count, its unfinished tasks left for the host),
// or after MAX_DEVICE_ROUNDS.
constexpr int MAX_DEVICE_ROUNDS = 64;

// Direction and candidate areas for the next device round: what the last
// round asked for (its allocators count every request, met or not), within
// the limits.
// (the capacities always describe the buffers, also after a failed growth)
int grow_round_areas(pw_ctx *ctx, Workspace &W, const Ctl &c, bool &grew) {
    if (!W.dir_ext && (int64_t)c.dir_used > W.dir_cap && W.dir_cap < dir_limit(ctx)) {
        const int64_t before = W.dir_cap;
        const int rc = ensure(ctx, W.dirs, (size_t)std::min<int64_t>((int64_t)c.dir_used, dir_limit(ctx)));
        W.dir_cap = (int64_t)W.dirs.cap;
        if (rc) return rc;
        grew = grew || W.dir_cap > before;
    }
    const int64_t cl = cand_slot_limit(ctx);
    if ((int64_t)c.cand_used > W.cand_cap && W.cand_cap < cl) {
        const int64_t before = W.cand_cap;
        const size_t n = (size_t)std::min<int64_t>((int64_t)c.cand_used, cl);
        int rc = ensure(ctx, W.cands, n * sizeof(Cand));
        if (!rc) rc = ensure(ctx, W.spslot, n * sizeof(SpRec));
        if (!rc) rc = ensure(ctx, W.keys, n * sizeof(uint64_t));
        set_slot_caps(W);
        if (rc) return rc;
        grew = grew || W.cand_cap > before;
    }
    return PW_OK;
}

// Replaces b by a buffer of at least need bytes whose first keep bytes are
// b's (a device-to-device copy); the old buffer goes back to the cache.
int regrow(pw_ctx *ctx, Lane &L, DevBuf &b, size_t keep, size_t need) {
    DevBuf old = b;
    b = DevBuf{};
    const int rc = ensure(ctx, b, need);
    if (rc) {
        b = old;
        return rc;
    }
    if (keep > 0) HIPCHK(ctx, hipMemcpyAsync(b.p, old.p, keep, hipMemcpyDeviceToDevice, L.s));
    HIPCHK(ctx, hipStreamSynchronize(L.s));
    if (ctx->defer_free) ctx->dead_dev.emplace_back(old.p, old.cap);
    else dev_release(ctx, old.p, old.cap);
    return PW_OK;
}

// The record areas filled up during a round: every allocation that succeeded
// ends at or below their capacity, the allocator having counted the failed
// ones beyond it.  Each area is replaced by one sized for what was asked
// (`used`) plus as much again beyond the old capacity, the old contents
// copied over; *kept = the old capacity, where the allocator restarts.
int grow_ops(pw_ctx *ctx, Workspace &W, Lane &L, int64_t used, long long *kept) {
    const int64_t keep = W.ops_cap;
    const int64_t need = std::min<int64_t>(used + (used - keep), std::max<int64_t>(used, dir_limit(ctx)));
    const int rc = regrow(ctx, L, W.ops, (size_t)keep, (size_t)need);
    if (rc) return rc;
    W.ops_cap = (int64_t)W.ops.cap;
    W.want_ops = std::max<int64_t>(W.want_ops, need);
    *kept = keep;
    return PW_OK;
}

int grow_records(pw_ctx *ctx, Workspace &W, Lane &L, int64_t used, long long *kept) {
    const int64_t keep = W.out_cap;
    const int64_t need = std::min<int64_t>(std::max<int64_t>(2 * keep, used + (used - keep)), (int64_t)INT32_MAX);
    if (need <= keep) return PW_ENOMEM;
    int rc = regrow(ctx, L, W.out_sps, (size_t)keep * sizeof(SpRec), (size_t)need * sizeof(SpRec));
    if (!rc) rc = regrow(ctx, L, W.out_hits, (size_t)keep * sizeof(HitRec), (size_t)need * sizeof(HitRec));
    set_slot_caps(W);
    if (rc) return rc;
    W.want_cand = std::max<int64_t>(W.want_cand, std::min<int64_t>(need, cand_slot_limit(ctx)));
    *kept = keep;
    return PW_OK;
}

int device_rounds(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, const Geometry &g, double minimum, Finished &F,
                  int &n_last, bool &done) {
    Ctl *dctl = (Ctl *)W.ctl.p;
    Ctl c = *(const Ctl *)W.h_ctl.p;
    int prev = c.ngated;
    n_last = prev;
    const int rq_grid = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)prev + 255) / 256, (int64_t)ctx->num_cus * 4));
    for (int round = 0; round < MAX_DEVICE_ROUNDS && c.unfinished > 0; ++round) {
        hipLaunchKernelGGL(requeue_kernel, dim3(rq_grid), dim3(256), 0, L.s, (const PairTask *)W.tasks.p, dctl,
                           (int *)W.gated.p, (int *)W.rq_slots.p, (int *)W.rq_needs.p);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemcpyAsync(W.h_ctl.p, dctl, sizeof(Ctl), hipMemcpyDeviceToHost, L.s));
        HIPCHK(ctx, hipStreamSynchronize(L.s));
        c = *(const Ctl *)W.h_ctl.p;
        const int next = c.next_ngated;
        if (next <= 0) break;
        // where growth is on: the direction and candidate areas, which hold one
        // round's work, toward what the round just walked asked for (nothing to
        // keep); the record and ops areas, full, with their contents copied.
        // Without growth a full record or ops area ends the rounds.
        bool grew = false;
        long long ops_keep = -1, out_keep = -1;
        const bool full = (int64_t)c.out_used > W.out_cap || (int64_t)c.ops_used > W.ops_cap;
        if (ctx->lim.grow) {
            int rc = grow_round_areas(ctx, W, c, grew);
            if (!rc && (int64_t)c.out_used > W.out_cap) rc = grow_records(ctx, W, L, (int64_t)c.out_used, &out_keep);
            if (!rc && (int64_t)c.ops_used > W.ops_cap) rc = grow_ops(ctx, W, L, (int64_t)c.ops_used, &ops_keep);
            if (rc == PW_ENOMEM) {
                // no room to grow: what is left goes to the host's tail rounds
                (void)hipGetLastError();
                ctx->err.clear();
                break;
            }
            if (rc) return rc;
            grew = grew || out_keep >= 0 || ops_keep >= 0;
        } else if (full) {
            break;
        }
        // the round just walked finished nothing and no area grew
        if (next >= prev && !grew) break;
        hipLaunchKernelGGL(round_reset_kernel, dim3(1), dim3(64), 0, L.s, dctl, next, ops_keep, out_keep);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(L.ev[2], L.s));
        int rc = launch_tasks(ctx, P, W, L, g, true);
        if (!rc) rc = launch_traceback(ctx, P, W, L, g, minimum, true);
        if (rc) return rc;
        HIPCHK(ctx, hipMemcpyAsync(W.h_ctl.p, dctl, sizeof(Ctl), hipMemcpyDeviceToHost, L.s));
        HIPCHK(ctx, hipStreamSynchronize(L.s));
        F.ms_dir += elapsed(L.ev[2], L.ev[3]);
        F.ms_tb += elapsed(L.ev[3], L.ev[4]);
        F.ms_walk += elapsed(L.ev[4], L.ev[5]);
        c = *(const Ctl *)W.h_ctl.p;
        if (c.abort_wait) {
            ctx->err = "a stripe or chunk hand-over wait on the device gave up after " +
                       std::to_string(ctx->wait_polls) + " polls; the execution's results are discarded";
            return PW_EDEVICE;
        }
        prev = next;
        n_last = next;
    }
    done = c.unfinished == 0;
    return PW_OK;
}

// layout_on_device: when every pair finishes on the device, leave the records
// there for device_assemble (F.on_device) instead of reading them back.
int execute_finish(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, const Exec &X, Finished &F,
                   bool layout_on_device = false) {
    HIPCHK(ctx, hipStreamSynchronize(L.s));
    if (X.graph) {
        F.ms_score = elapsed(L.ev[0], L.ev[5]);
    } else {
        F.ms_score = elapsed(L.ev[0], L.ev[1]);
        F.ms_dir = elapsed(L.ev[2], L.ev[3]);
        F.ms_tb = elapsed(L.ev[3], L.ev[4]);
        F.ms_walk = elapsed(L.ev[4], L.ev[5]);
    }
    const Ctl c = *(const Ctl *)W.h_ctl.p;
    if (c.abort_wait) {
        ctx->err = "a stripe or chunk hand-over wait on the device gave up after " + std::to_string(ctx->wait_polls) +
                   " polls; the execution's results are discarded";
        return PW_EDEVICE;
    }
    F.ng = c.ngated;
    if (F.ng < 0 || F.ng > P->n_pairs) {
        ctx->err = "gate count out of range";
        return PW_EDEVICE;
    }
    // pairs that did not fit (directions, candidates, records) run further
    // rounds on the device, the records collecting in the read-back areas;
    // what a round cannot finish goes to the host's tail rounds
    int n_last = -1;
    bool done = c.unfinished == 0;
    auto t0 = std::chrono::steady_clock::now();
    if (!done) {
        int rc = device_rounds(ctx, P, W, L, X.g, X.minimum, F, n_last, done);
        if (rc) return rc;
    }
    F.ms_device_rounds = ms_since(t0);
    if (done && layout_on_device) {
        if (c.unfinished > 0) adapt_workspace(ctx, W, c, *(const Ctl *)W.h_ctl.p, {});
        F.ms_tail = F.ms_device_rounds;
        F.on_device = true;
        return PW_OK;
    }
    std::vector<PairTask> pending;
    t0 = std::chrono::steady_clock::now();
    int rc = read_tasks(ctx, W, L, F.ng, F.outs, pending, n_last);
    if (rc) return rc;
    F.ms_readback = ms_since(t0);
    t0 = std::chrono::steady_clock::now();
    // areas sized for what this execution asked for, for the next one
    if (c.unfinished > 0) adapt_workspace(ctx, W, c, *(const Ctl *)W.h_ctl.p, pending);
    if (!pending.empty()) {
        // the tail rounds reuse the staging buffers: keep copies of what they hold
        F.store.keep(F.outs, 0);
        rc = tail_rounds(ctx, P, W, L, X.g, X.minimum, pending, F.outs, F.store, &F.ms_dir, &F.ms_tb, &F.ms_walk);
        if (rc) return rc;
    }
    F.ms_tail = ms_since(t0) + F.ms_device_rounds;  // further rounds: on the device, then on the host
    t0 = std::chrono::steady_clock::now();
    sort_by_pair(F.outs);
    F.ms_assemble = ms_since(t0);
    return PW_OK;
}

static_assert(sizeof(HitRec) == sizeof(pw_hit) && offsetof(HitRec, score) == offsetof(pw_hit, score) &&
                  offsetof(HitRec, sp_index) == offsetof(pw_hit, sp_index) &&
                  offsetof(HitRec, ops_offset) == offsetof(pw_hit, ops_offset),
              "the device writes the public hit records");

// The result of an execution whose pairs all finished on the device: the
// records, op characters and alignment rows laid out in pair order there, then
// copied into the result's pinned arrays.
int device_assemble(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, int ng, int32_t sp_cap, HostResult *R,
                    Finished &F) {
    auto t0 = std::chrono::steady_clock::now();
    int64_t nsp = 0, nops = 0;
    Ctl *dctl = (Ctl *)W.ctl.p;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)ng + 255) / 256, (int64_t)ctx->num_cus * 4));
    if (ng > 0) {
        ENSURE(ctx, W.lsums, (size_t)grid * 2 * sizeof(long long));
        ENSURE(ctx, W.lay_off, (size_t)ng * 2 * sizeof(long long));
        hipLaunchKernelGGL(layout_sums_kernel, dim3(grid), dim3(256), 0, L.s, (const TaskOut *)W.touts.p, ng, dctl,
                           (long long *)W.lsums.p);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemcpyAsync(W.h_ctl.p, dctl, sizeof(Ctl), hipMemcpyDeviceToHost, L.s));
        HIPCHK(ctx, hipStreamSynchronize(L.s));
        const Ctl &c = *(const Ctl *)W.h_ctl.p;
        nsp = c.lay_sps;
        nops = c.lay_ops;
        if (nsp < 0 || nops < 0 || nsp > (int64_t)INT32_MAX) {
            ctx->err = "result sizes out of range";
            return PW_EDEVICE;
        }
    }
    std::vector<void *> *dead = ctx->defer_free ? &ctx->dead_pinned : nullptr;
    if (!R->sps.reserve((size_t)nsp, true, dead) || !R->hits.reserve((size_t)nsp, true, dead) ||
        !R->ops.reserve((size_t)nops, true, dead) || !R->aligned.reserve((size_t)(3 * nops), true, dead)) {
        ctx->err = "no host memory for the result records";
        return PW_ENOMEM;
    }
    if (nsp > 0) {
        if (!P->d_chars.p || P->d_chars.cap < 512) {
            ctx->err = "the plan's residue characters are not on the device";
            return PW_EDEVICE;
        }
        ENSURE(ctx, W.f_sps, (size_t)nsp * sizeof(SpRec));
        ENSURE(ctx, W.f_hits, (size_t)nsp * sizeof(HitRec));
        ENSURE(ctx, W.f_ops, (size_t)std::max<int64_t>(1, nops));
        ENSURE(ctx, W.f_aln, (size_t)std::max<int64_t>(1, 3 * nops));
        long long *off = (long long *)W.lay_off.p;
        hipLaunchKernelGGL(layout_offsets_kernel, dim3(grid), dim3(256), 0, L.s, (const TaskOut *)W.touts.p, ng,
                           (const long long *)W.lsums.p, off, off + ng);
        HIPCHK(ctx, hipGetLastError());
        const int ggrid = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)ng + 3) / 4, (int64_t)ctx->num_cus * 8));
        hipLaunchKernelGGL(layout_gather_kernel, dim3(ggrid), dim3(256), 0, L.s, (const uint8_t *)P->res.p,
                           (const SeqRef *)P->seqs.p, P->nq, (const TaskOut *)W.touts.p, ng, (const long long *)off,
                           (const long long *)(off + ng), (const SpRec *)W.out_sps.p, (const HitRec *)W.out_hits.p,
                           (const char *)W.ops.p, (const uint8_t *)P->d_chars.p, (SpRec *)W.f_sps.p,
                           (HitRec *)W.f_hits.p, (char *)W.f_ops.p, (char *)W.f_aln.p);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipStreamSynchronize(L.s));
        F.ms_assemble = ms_since(t0);
        t0 = std::chrono::steady_clock::now();
        HIPCHK(ctx, hipMemcpyAsync(R->sps.data(), W.f_sps.p, (size_t)nsp * sizeof(SpRec), hipMemcpyDeviceToHost, L.s));
        HIPCHK(ctx, hipMemcpyAsync(R->hits.data(), W.f_hits.p, (size_t)nsp * sizeof(HitRec), hipMemcpyDeviceToHost,
                                   L.s));
        if (nops > 0) {
            HIPCHK(ctx, hipMemcpyAsync(R->ops.data(), W.f_ops.p, (size_t)nops, hipMemcpyDeviceToHost, L.s));
            HIPCHK(ctx, hipMemcpyAsync(R->aligned.data(), W.f_aln.p, (size_t)(3 * nops), hipMemcpyDeviceToHost, L.s));
        }
        HIPCHK(ctx, hipStreamSynchronize(L.s));
        F.ms_readback = ms_since(t0);
    } else {
        F.ms_assemble = ms_since(t0);
    }
    R->n_sps = nsp;
    R->n_ops = nops;
    R->pub.sp_truncated = sp_over_cap(nsp, sp_cap, P->n_pairs);
    return PW_OK;
}

void fill_public(HostResult *R, const float *maxima, int64_t n_pairs, double cells, int64_t ng, const Finished &F,
                 const Exec &X) {
    pw_result &O = R->pub;
    O.n_pairs = n_pairs;
    O.maxima = maxima;
    O.n_starting_points = R->n_sps;
    O.starting_points = R->sps.data();
    O.n_hits = R->n_sps;
    O.hits = R->hits.data();
    O.n_ops = R->n_ops;
    O.ops = R->ops.data();
    O.aligned = R->aligned.data();
    O.n_traceback_pairs = ng;
    O.cells = cells;
    O.ms_score = F.ms_score;
    O.ms_directions = F.ms_dir;
    O.ms_traceback = F.ms_tb;
    O.ms_walk = F.ms_walk;
    O.score_launches = X.launches;
    O.ms_readback = F.ms_readback;
    O.ms_tail = F.ms_tail;
    O.ms_assemble = F.ms_assemble;
    O.ms_plan = 0;  // set by the one-shot call that planned the execution
    O.ms_upload = 0;
    O.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - X.wall0).count();
}

int execute(pw_ctx *ctx, pw_plan *P, Workspace &W, Lane &L, double minimum, int32_t sp_cap, bool use_graph,
            HostResult *R) {
    int rc = take_maxima_buffer(ctx, P->n_pairs, R);
    if (rc) return rc;
    Exec X;
    rc = execute_launch(ctx, P, W, L, minimum, use_graph, R->maxima, X);
    if (rc) return rc;
    Finished F;
    rc = execute_finish(ctx, P, W, L, X, F, ctx->layout_on_device);
    if (rc) return rc;
    if (F.on_device) {
        rc = device_assemble(ctx, P, W, L, F.ng, sp_cap, R, F);
        if (rc) return rc;
    } else {
        const auto t0 = std::chrono::steady_clock::now();
        rc = assemble(ctx, F.outs, sp_cap, P->n_pairs, P, R);
        if (rc) return rc;
        F.ms_assemble += ms_since(t0);
    }
    fill_public(R, R->maxima, P->n_pairs, P->cells, F.ng, F, X);
    return PW_OK;
}

PlanOpts plan_opts(pw_ctx *ctx, const Scoring &sc) {
    PlanOpts o;
    o.total_mem = ctx->total_mem;
    o.num_cus = ctx->num_cus;
    o.tune = &ctx->tune[sc.affine ? 1 : 0];
    o.orient = ctx->route_orient;
    o.group_max_share = ctx->route_share;
    o.dir_rmax = ctx->route_dir_rmax;
    if (!ctx->plan_scratch) ctx->plan_scratch = new PlanScratch();
    o.scratch = ctx->plan_scratch;
    return o;
}

// Calls that only score: no checkpoints for a direction pass.
PlanOpts score_only_opts(pw_ctx *ctx, const Scoring &sc) {
    PlanOpts o = plan_opts(ctx, sc);
    o.checkpoints = false;
    return o;
}

int ready(pw_ctx *ctx) {
    if (!ctx) return PW_EINVAL;
    if (!ctx->scoring_set) {
        ctx->err = "pw_set_scoring has not been called";
        return PW_EINVAL;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return PW_OK;
}

// A result record, recycled from the context's pool when one is there (its
// vectors keep their capacity, so filling them touches no new pages).
int new_result(pw_ctx *ctx, HostResult **out) {
    HostResult *R = nullptr;
    if (ctx && !ctx->result_pool.empty()) {
        R = ctx->result_pool.back();
        ctx->result_pool.pop_back();
    } else {
        R = new HostResult();
    }
    std::memset(&R->pub, 0, sizeof(R->pub));
    R->n_sps = R->n_ops = 0;
    R->maxima_vec.clear();
    *out = R;
    return PW_OK;
}

}  // namespace

extern "C" int pw_free_result(pw_ctx *ctx, pw_result *res);

namespace {

// Times the scoring pass of a synthetic job: nq queries of length A and nt
// targets of length B over the first residues of the scoring alphabet, pairs
// from pairs_of; returns the best of two device times in milliseconds.
int time_synthetic(pw_ctx *ctx, const PlanOpts &opt, int nq, int A, int nt, int B,
                   const std::vector<int32_t> &pq, const std::vector<int32_t> &pt, uint64_t &seed, double &ms) {
    const Scoring &sc = ctx->sc;
    const int K = std::max(1, std::min(sc.dim, 20));
    std::vector<uint8_t> qres((size_t)nq * A + 1), tres((size_t)nt * B + 1);
    auto rnd = [&]() {
        seed = seed * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(seed >> 33);
    };
    for (auto &c : qres) c = (uint8_t)(sc.char_offset + rnd() % K);
    for (auto &c : tres) c = (uint8_t)(sc.char_offset + rnd() % K);
    std::vector<int64_t> qoff(nq), toff(nt);
    std::vector<int32_t> qlen(nq, A), tlen(nt, B);
    for (int i = 0; i < nq; ++i) qoff[i] = (int64_t)i * A;
    for (int i = 0; i < nt; ++i) toff[i] = (int64_t)i * B;
    pw_plan P;
    PlanOpts o = opt;
    o.checkpoints = false;  // the tuned shapes are those of the scoring pass alone
    int rc = plan_host(ctx->err, sc, o, &P, qres.data(), qoff.data(), qlen.data(), nq, tres.data(), toff.data(),
                       tlen.data(), nt, pq.data(), pt.data(), (int64_t)pq.size(), nullptr, nullptr, 1);
    if (!rc) rc = plan_upload(ctx, &P, ctx->lane.s, ctx->h_stage, true);
    Workspace &W = P.own;
    const Geometry g = geometry(ctx, &P);
    if (!rc) rc = size_workspace(ctx, &P, W, g);
    ms = 1e300;
    for (int rep = 0; rep < 3 && !rc; ++rep) {
        rc = launch_clears(ctx, &P, W, ctx->lane);
        if (!rc) rc = launch_score(ctx, &P, W, ctx->lane, g, true, nullptr);
        if (!rc) {
            hipError_t e = hipStreamSynchronize(ctx->lane.s);
            if (e != hipSuccess) {
                ctx->err = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
                rc = PW_EDEVICE;
            }
        }
        if (!rc && rep > 0) ms = std::min(ms, elapsed(ctx->lane.ev[0], ctx->lane.ev[1]));
    }
    free_plan_buffers(ctx, &P);
    return rc;
}

// Measures the tile kernel for every length bucket and residues per lane, and
// the chunk widths of multi-chunk lane-per-pair jobs, for the current scoring
// scheme's gap model.
int autotune(pw_ctx *ctx, double *seconds) {
    const auto t0 = std::chrono::steady_clock::now();
    Tuning tu;
    uint64_t seed = 12345;
    PlanOpts opt = plan_opts(ctx, ctx->sc);
    opt.tune = nullptr;
    for (int ia = 0; ia < TUNE_A; ++ia)
        for (int ib = 0; ib < TUNE_B; ++ib) {
            const int A = TUNE_REP_A[ia], B = TUNE_REP_B[ib];
            // enough pairs to fill the device several times over, at most 8192
            const int n = (int)std::max<int64_t>(1024, std::min<int64_t>(8192, (int64_t)3e8 / ((int64_t)A * B)));
            std::vector<int32_t> pq(n), pt(n);
            for (int i = 0; i < n; ++i) pq[i] = pt[i] = i;
            for (int k = 0; k < TUNE_R; ++k) {
                opt.force_tile = true;
                opt.force_R = TUNE_RS[k];
                double ms = 0;
                int rc = time_synthetic(ctx, opt, n, A, n, B, pq, pt, seed, ms);
                if (rc) return rc;
                tu.rate[ia][ib][k] = (float)((double)n * A * B / (ms * 1e3));
            }
        }
    // multi-chunk lane-per-pair jobs: 512 targets of 1024 residues, each with 64 queries of 384
    {
        const int NT = 512, A = 384, B = 1024;
        std::vector<int32_t> pq, pt;
        for (int t = 0; t < NT; ++t)
            for (int q = 0; q < 64; ++q) {
                pq.push_back(t * 64 + q);
                pt.push_back(t);
            }
        const int widths[5] = {32, 48, 64, 80, 96};
        opt = plan_opts(ctx, ctx->sc);
        opt.tune = nullptr;
        opt.force_group = true;
        for (int w = 0; w < 5; ++w) {
            opt.force_nb = widths[w];
            double ms = 0;
            int rc = time_synthetic(ctx, opt, NT * 64, A, NT, B, pq, pt, seed, ms);
            if (rc) return rc;
            tu.group_rate[w] = (float)((double)NT * 64 * A * B / (ms * 1e3));
        }
        // the built-in width unless another is clearly faster
        const int wdef = (int)(std::find(widths, widths + 5, GROUP_NB_LONG) - widths);
        float best = tu.group_rate[wdef];
        tu.group_nb_long = GROUP_NB_LONG;
        for (int w = 0; w < 5; ++w)
            if (tu.group_rate[w] > TUNE_MARGIN * tu.group_rate[wdef] && tu.group_rate[w] > best) {
                best = tu.group_rate[w];
                tu.group_nb_long = widths[w];
            }
    }
    tu.valid = true;
    ctx->tune[ctx->sc.affine ? 1 : 0] = tu;
    // the synthetic jobs' upload staging is large; give it back
    free_pinned(ctx->h_stage);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PW_OK;
}

// Plan-level tuning: the plan's own scoring pass timed under a few routing
// choices (the longest lane-per-pair job as a share of the per-lane work, the
// chunk width of multi-chunk lane-per-pair jobs, tile shapes from the measured
// table or the cost model); the fastest work list is kept.
int tune_plan(pw_ctx *ctx, pw_plan *P, double *seconds, double *best_ms, int32_t *choice) {
    const auto t0 = std::chrono::steady_clock::now();
    if (P->mode != 1) {
        if (seconds) *seconds = 0.0;
        if (best_ms) *best_ms = 0.0;
        return PW_OK;
    }
    struct Cand {
        double share;
        int nb;
        bool table;
    };
    const Tuning *tu = &ctx->tune[P->sc.affine ? 1 : 0];
    std::vector<Cand> cands;
    for (double share : {0.125, 0.25, 0.5})
        for (int nb : {GROUP_NB_LONG, 64, 96})
            for (int table = 0; table < (tu->valid ? 2 : 1); ++table) cands.push_back(Cand{share, nb, table != 0});
    auto opts_of = [&](const Cand &c) {
        PlanOpts o = plan_opts(ctx, P->sc);
        o.group_max_share = c.share;
        o.force_nb = c.nb;
        o.tune = c.table ? tu : nullptr;
        return o;
    };
    // the default choices are timed first and kept unless another candidate
    // is clearly faster (differences within a few per cent are noise)
    std::stable_sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) {
        const bool da = a.share == 0.25 && a.nb == GROUP_NB_LONG && !a.table,
                   db = b.share == 0.25 && b.nb == GROUP_NB_LONG && !b.table;
        return da && !db;
    });
    constexpr double PLAN_TUNE_MARGIN = 0.97;
    double best = 1e300, base = 1e300;
    size_t bi = 0;
    Lane &L = ctx->lane;
    for (size_t i = 0; i < cands.size(); ++i) {
        int rc = plan_worklist(opts_of(cands[i]), P);
        if (!rc) rc = plan_upload(ctx, P, L.s, ctx->h_stage, true);
        const Geometry g = geometry(ctx, P);
        if (!rc) rc = size_workspace(ctx, P, P->own, g);
        double ms = 1e300;
        for (int rep = 0; rep < 4 && !rc; ++rep) {
            rc = launch_clears(ctx, P, P->own, L);
            if (!rc) rc = launch_score(ctx, P, P->own, L, g, true, nullptr);
            if (!rc) {
                hipError_t e = hipStreamSynchronize(L.s);
                if (e != hipSuccess) {
                    ctx->err = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
                    rc = PW_EDEVICE;
                }
            }
            if (!rc && rep > 0) ms = std::min(ms, elapsed(L.ev[0], L.ev[1]));
        }
        if (rc) return rc;
        if (i == 0) {
            base = best = ms;
        } else if (ms < best && ms < PLAN_TUNE_MARGIN * base) {
            best = ms;
            bi = i;
        }
    }
    int rc = plan_worklist(opts_of(cands[bi]), P);
    if (!rc) rc = plan_upload(ctx, P, L.s, ctx->h_stage, true);
    if (rc) return rc;
    if (P->graph) {
        (void)hipGraphExecDestroy(P->graph);
        P->graph = nullptr;
        P->graph_sig.clear();
    }
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (best_ms) *best_ms = best;
    if (choice) {
        choice[0] = (int32_t)(cands[bi].share * 1000.0 + 0.5);
        choice[1] = cands[bi].nb;
        choice[2] = cands[bi].table ? 1 : 0;
    }
    return PW_OK;
}

}  // namespace

extern "C" int pw_tune_plan(pw_ctx *ctx, pw_plan *plan, double *seconds, double *score_ms, int32_t *choice) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (!plan) {
        ctx->err = "pw_tune_plan: missing plan";
        return PW_EINVAL;
    }
    return tune_plan(ctx, plan, seconds, score_ms, choice);
}

extern "C" int pw_autotune(pw_ctx *ctx, double *seconds) {
    int rc = ready(ctx);
    if (rc) return rc;
    return autotune(ctx, seconds);
}

extern "C" int pw_get_tuning(pw_ctx *ctx, int32_t affine, float *rates, float *group_rates, int32_t *group_nb_long,
                             int32_t *valid) {
    if (!ctx || !rates || !group_rates || !group_nb_long || !valid) return PW_EINVAL;
    const Tuning &tu = ctx->tune[affine ? 1 : 0];
    std::memcpy(rates, tu.rate, sizeof tu.rate);
    std::memcpy(group_rates, tu.group_rate, sizeof tu.group_rate);
    *group_nb_long = tu.group_nb_long;
    *valid = tu.valid ? 1 : 0;
    return PW_OK;
}

extern "C" int pw_set_tuning(pw_ctx *ctx, int32_t affine, const float *rates, int32_t group_nb_long) {
    if (!ctx) return PW_EINVAL;
    if (!rates) {
        ctx->tune[affine ? 1 : 0] = Tuning();  // back to the cost model
        return PW_OK;
    }
    if (group_nb_long != 32 && group_nb_long != 48 && group_nb_long != 64 && group_nb_long != 80 &&
        group_nb_long != 96) {
        ctx->err = "pw_set_tuning: chunk width must be 32, 48, 64, 80 or 96";
        return PW_EINVAL;
    }
    Tuning tu;
    std::memcpy(tu.rate, rates, sizeof tu.rate);
    for (int a = 0; a < TUNE_A; ++a)
        for (int b = 0; b < TUNE_B; ++b)
            for (int k = 0; k < TUNE_R; ++k)
                if (!std::isfinite(tu.rate[a][b][k]) || tu.rate[a][b][k] < 0.0f) {
                    ctx->err = "pw_set_tuning: rates must be finite and non-negative";
                    return PW_EINVAL;
                }
    tu.group_nb_long = group_nb_long;
    tu.valid = true;
    ctx->tune[affine ? 1 : 0] = tu;
    return PW_OK;
}

extern "C" int pw_prepare(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
                          const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt,
                          const int32_t *pq, const int32_t *pt, int64_t n_pairs, const float *min_start_score,
                          const uint8_t *same_id, int32_t mode, pw_plan **out) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (!out || nq < 0 || nt < 0 || (n_pairs > 0 && (!pq || !pt))) {
        ctx->err = "pw_prepare: missing argument";
        return PW_EINVAL;
    }
    *out = nullptr;
    pw_plan *P = new pw_plan();
    rc = plan_host(ctx->err, ctx->sc, plan_opts(ctx, ctx->sc), P, qres, qoff, qlen, nq, tres, toff, tlen, nt, pq, pt, n_pairs,
                   min_start_score, same_id, mode);
    if (!rc) rc = plan_upload(ctx, P, ctx->lane.s, ctx->h_stage, true);
    if (rc) {
        free_plan(ctx, P);
        return rc;
    }
    *out = P;
    return PW_OK;
}

static int execute_entry(pw_ctx *ctx, pw_plan *plan, double minimum, int32_t sp_cap, bool graph, pw_result **out,
                         const char *what) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (!plan || !out) {
        ctx->err = std::string(what) + ": missing argument";
        return PW_EINVAL;
    }
    *out = nullptr;
    HostResult *R = nullptr;
    new_result(ctx, &R);
    rc = execute(ctx, plan, plan->own, ctx->lane, minimum, sp_cap, graph, R);
    if (rc) {
        pw_free_result(ctx, &R->pub);
        return rc;
    }
    *out = &R->pub;
    return PW_OK;
}

extern "C" int pw_execute(pw_ctx *ctx, pw_plan *plan, double minimum, int32_t sp_cap, pw_result **out) {
    return execute_entry(ctx, plan, minimum, sp_cap, false, out, "pw_execute");
}

extern "C" int pw_execute_graph(pw_ctx *ctx, pw_plan *plan, double minimum, int32_t sp_cap, pw_result **out) {
    return execute_entry(ctx, plan, minimum, sp_cap, true, out, "pw_execute_graph");
}

extern "C" int pw_plan_info(const pw_plan *plan, int64_t *info, int32_t n) {
    if (!plan || !info || n < 0) return PW_EINVAL;
    const int64_t v[10] = {plan->n_pairs,
                           plan->n_groups,
                           plan->n_tile,
                           plan->orient,
                           (int64_t)plan->group_cells,
                           (int64_t)(plan->cells - plan->group_cells),
                           plan->mode == 1 ? plan->max_nb : 0,
                           plan->dir_rmax,
                           plan->group_split_items,
                           (int64_t)plan->ckpairs.size()};
    for (int32_t i = 0; i < n && i < 10; ++i) info[i] = v[i];
    return PW_OK;
}

extern "C" int pw_release_plan(pw_ctx *ctx, pw_plan *plan) {
    if (!ctx || !plan) return PW_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->lane.s);
    (void)hipStreamSynchronize(ctx->lane.s2);
    free_plan(ctx, plan);
    return PW_OK;
}

namespace {

// After a score-only execution on L.s has finished: PW_EDEVICE when one of its
// waves gave up on a hand-over wait (the maxima are then discarded).
int check_wait_abort(pw_ctx *ctx, const Workspace &W, Lane &L) {
    unsigned ab = 0;
    HIPCHK(ctx, hipMemcpyAsync(&ab, &((const Ctl *)W.ctl.p)->abort_wait, sizeof ab, hipMemcpyDeviceToHost, L.s));
    HIPCHK(ctx, hipStreamSynchronize(L.s));
    if (ab) {
        ctx->err = "a stripe or chunk hand-over wait on the device gave up after " + std::to_string(ctx->wait_polls) +
                   " polls; the maxima are discarded";
        return PW_EDEVICE;
    }
    return PW_OK;
}

}  // namespace

extern "C" int pw_score_pairs(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
                              const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt,
                              const int32_t *pq, const int32_t *pt, int64_t n_pairs, int32_t mode, float *maxima_out,
                              double *ms_out) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (nq < 0 || nt < 0 || (n_pairs > 0 && (!pq || !pt || !maxima_out))) {
        ctx->err = "pw_score_pairs: missing argument";
        return PW_EINVAL;
    }
    if (!ctx->oneshot) ctx->oneshot = new pw_plan();
    pw_plan *P = ctx->oneshot;
    rc = plan_host(ctx->err, ctx->sc, score_only_opts(ctx, ctx->sc), P, qres, qoff, qlen, nq, tres, toff, tlen, nt, pq,
                   pt, n_pairs, nullptr, nullptr, mode);
    if (!rc) rc = plan_upload(ctx, P, ctx->lane.s, ctx->h_stage, true);
    Workspace &W = ctx->ws;
    Lane &L = ctx->lane;
    if (!rc) {
        const Geometry g = geometry(ctx, P);
        rc = size_workspace(ctx, P, W, g);
        if (!rc) {
            rc = launch_clears(ctx, P, W, L);
            if (!rc) rc = launch_score(ctx, P, W, L, g, true, nullptr);
        }
    }
    if (!rc && n_pairs > 0) {
        hipError_t e = hipMemcpyAsync(maxima_out, W.maxima.p, (size_t)n_pairs * sizeof(float), hipMemcpyDeviceToHost,
                                      L.s);
        if (e != hipSuccess) {
            ctx->err = std::string("hipMemcpyAsync maxima: ") + hipGetErrorString(e);
            rc = PW_EDEVICE;
        }
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(L.s);
        if (e != hipSuccess) {
            ctx->err = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
            rc = PW_EDEVICE;
        }
    }
    if (!rc) rc = check_wait_abort(ctx, W, L);
    if (!rc && ms_out) *ms_out = elapsed(L.ev[0], L.ev[1]);
    return rc;
}

namespace {

// Places the device arrays of a score-only execution of P in the static
// scratch at base (cap bytes, 256-byte aligned parts): the plan's inputs, the
// control block, the maxima, the split pairs' boundary rows and progress
// counters, and the per-workgroup scratch of the scoring kernels.  Persistent
// grids are lowered until the per-workgroup scratch fits (workers pull their
// items from device counters, and a split pair's stripes or a split job's
// chunks are dequeued in order, so any grid of at least one workgroup gives
// the same maxima).  Returns PW_EINVAL when the inputs alone, or one
// workgroup's scratch per kernel, do not fit.
int noalloc_layout(pw_ctx *ctx, pw_plan *P, Workspace &W, Geometry &g, char *base, size_t cap) {
    const Encoded &E = P->E;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    struct Part {
        DevBuf *d;
        size_t bytes;
    };
    const Part parts[] = {{&P->res, E.codes.size()},
                          {&P->seqs, E.refs.size() * sizeof(SeqRef)},
                          {&P->tables, E.tables.size() * sizeof(float)},
                          {&P->d_items, P->items.size() * sizeof(ScoreItem)},
                          {&P->d_trans, P->trans.size()},
                          {&P->d_gitems, P->gitems.size() * sizeof(GroupItem)},
                          {&P->d_gents, P->gents.size() * sizeof(GroupEntry)},
                          {&P->d_links, (size_t)P->score_links * sizeof(float)},
                          {&P->d_prog, (size_t)P->score_prog * sizeof(unsigned)},
                          {&W.ctl, sizeof(Ctl)},
                          {&W.maxima, (size_t)P->n_pairs * sizeof(float)}};
    size_t used = 0;
    for (const Part &q : parts) {
        q.d->p = base + used;
        q.d->cap = q.bytes;
        used += up(q.bytes);
    }
    int gmax = 0;
    for (int w = 0; w < NB_KINDS; ++w) gmax = std::max(gmax, g.group_grid[w]);
    const size_t gper = up((size_t)g.gscr_stride * sizeof(float2)), tper = up((size_t)g.tscr_stride * sizeof(float));
    const size_t fixed = used + (gmax > 0 ? gper : 0) + (g.tile_grid > 0 ? tper : 0);
    if (fixed > cap) {
        ctx->err = "pw_score_pairs_noalloc: the job needs at least " + std::to_string(fixed) +
                   " bytes of device memory, more than the static scratch of " + std::to_string(cap) + " bytes";
        return PW_EINVAL;
    }
    // what is left is shared by the two scratch arrays in proportion to the
    // grids the plan asked for
    const size_t rest = cap - used;
    const double want_g = gmax > 0 ? (double)gmax * gper : 0.0, want_t = g.tile_grid > 0 ? (double)g.tile_grid * tper : 0.0;
    int ng = 0, nt = 0;
    if (gmax > 0) {
        const size_t share = want_t > 0 ? (size_t)((double)rest * want_g / (want_g + want_t)) : rest;
        ng = (int)std::max<size_t>(1, std::min<size_t>((size_t)gmax, share / gper));
    }
    if (g.tile_grid > 0) {
        const size_t left = rest - (size_t)ng * (gmax > 0 ? gper : 0);
        nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)g.tile_grid, left / tper));
    }
    for (int w = 0; w < NB_KINDS; ++w) g.group_grid[w] = std::min(g.group_grid[w], ng);
    g.tile_grid = std::min(g.tile_grid, nt);
    W.gscratch.p = ng > 0 ? base + used : nullptr;
    W.gscratch.cap = (size_t)ng * (ng > 0 ? gper : 0);
    used += W.gscratch.cap;
    W.tscratch.p = nt > 0 ? base + used : nullptr;
    W.tscratch.cap = (size_t)nt * (nt > 0 ? tper : 0);
    used += W.tscratch.cap;
    return used <= cap ? PW_OK : PW_EINVAL;
}

}  // namespace

extern "C" int pw_score_pairs_noalloc(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen,
                                      int32_t nq, const uint8_t *tres, const int64_t *toff, const int32_t *tlen,
                                      int32_t nt, const int32_t *pq, const int32_t *pt, int64_t n_pairs, int32_t mode,
                                      float *maxima_out, double *ms_out) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (nq < 0 || nt < 0 || (n_pairs > 0 && (!pq || !pt || !maxima_out))) {
        ctx->err = "pw_score_pairs_noalloc: missing argument";
        return PW_EINVAL;
    }
    if (!ctx->static_scratch) {
        ctx->err = "pw_score_pairs_noalloc: no static scratch on this device (" + ctx->static_err + ")";
        return PW_EDEVICE;
    }
    // the plan is built in host memory only; its device arrays are views of
    // the static scratch, never released to the context's buffer cache
    std::unique_ptr<pw_plan> P(new pw_plan());
    rc = plan_host(ctx->err, ctx->sc, score_only_opts(ctx, ctx->sc), P.get(), qres, qoff, qlen, nq, tres, toff, tlen,
                   nt, pq, pt, n_pairs, nullptr, nullptr, mode);
    if (rc) return rc;
    Workspace W;
    Geometry g = geometry(ctx, P.get());
    rc = noalloc_layout(ctx, P.get(), W, g, (char *)ctx->static_scratch, PW_NOALLOC_SCRATCH_BYTES);
    if (rc) return rc;
    Lane &L = ctx->lane;
    const Encoded &E = P->E;
    const std::pair<DevBuf *, const void *> inputs[] = {
        {&P->res, E.codes.data()},        {&P->seqs, E.refs.data()},      {&P->tables, E.tables.data()},
        {&P->d_items, P->items.data()},   {&P->d_trans, P->trans.data()}, {&P->d_gitems, P->gitems.data()},
        {&P->d_gents, P->gents.data()}};
    for (const auto &in : inputs) {
        if (!in.first->cap) continue;
        const hipError_t ce = hipMemcpyAsync(in.first->p, in.second, in.first->cap, hipMemcpyHostToDevice, L.s);
        if (ce != hipSuccess) {
            ctx->err = std::string("hipMemcpyAsync inputs: ") + hipGetErrorString(ce);
            rc = PW_EDEVICE;
            break;
        }
    }
    if (!rc) rc = launch_clears(ctx, P.get(), W, L);
    if (!rc) rc = launch_score(ctx, P.get(), W, L, g, true, nullptr);
    if (!rc && n_pairs > 0) {
        hipError_t e = hipMemcpyAsync(maxima_out, W.maxima.p, (size_t)n_pairs * sizeof(float), hipMemcpyDeviceToHost,
                                      L.s);
        if (e != hipSuccess) {
            ctx->err = std::string("hipMemcpyAsync maxima: ") + hipGetErrorString(e);
            rc = PW_EDEVICE;
        }
    }
    // the host copies of the inputs (in P) and the scratch must outlive the work
    hipError_t e = hipStreamSynchronize(L.s);
    if (L.s2 != L.s) (void)hipStreamSynchronize(L.s2);
    if (!rc && e != hipSuccess) {
        ctx->err = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
        rc = PW_EDEVICE;
    }
    if (!rc) rc = check_wait_abort(ctx, W, L);
    if (!rc && ms_out) *ms_out = elapsed(L.ev[0], L.ev[1]);
    return rc;
}

extern "C" int pw_align_pairs(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
                              const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt,
                              const int32_t *pq, const int32_t *pt, int64_t n_pairs, const float *min_start_score,
                              const uint8_t *same_id, double minimum, int32_t sp_cap, int32_t mode, pw_result **out) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (!out || nq < 0 || nt < 0 || (n_pairs > 0 && (!pq || !pt || !min_start_score))) {
        ctx->err = "pw_align_pairs: missing argument";
        return PW_EINVAL;
    }
    *out = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    // the context's one-shot plan: planned anew, its host vectors and device
    // arrays reused at their current sizes
    if (!ctx->oneshot) ctx->oneshot = new pw_plan();
    pw_plan *P = ctx->oneshot;
    rc = plan_host(ctx->err, ctx->sc, plan_opts(ctx, ctx->sc), P, qres, qoff, qlen, nq, tres, toff, tlen, nt, pq, pt, n_pairs,
                   min_start_score, same_id, mode);
    const double ms_plan = ms_since(t0);
    const auto t1 = std::chrono::steady_clock::now();
    if (!rc) rc = plan_upload(ctx, P, ctx->lane.s, ctx->h_stage, true);
    const double ms_upload = ms_since(t1);
    if (!rc) {
        HostResult *R = nullptr;
        new_result(ctx, &R);
        rc = execute(ctx, P, ctx->ws, ctx->lane, minimum, sp_cap, false, R);
        if (rc) {
            pw_free_result(ctx, &R->pub);
        } else {
            R->pub.ms_plan = ms_plan;
            R->pub.ms_upload = ms_upload;
            R->pub.ms_total = ms_since(t0);
            *out = &R->pub;
        }
    }
    return rc;
}

extern "C" int pw_free_result(pw_ctx *ctx, pw_result *res) {
    if (!res) return PW_EINVAL;
    // pub is the first member of HostResult
    HostResult *R = reinterpret_cast<HostResult *>(res);
    if (R->maxima) {
        if (ctx && ctx->pinned_free.size() < 4) ctx->pinned_free.push_back({R->maxima, R->maxima_cap});
        else (void)hipHostFree(R->maxima);
        R->maxima = nullptr;
        R->maxima_cap = 0;
    }
    if (ctx && ctx->result_pool.size() < 4) ctx->result_pool.push_back(R);
    else delete R;
    return PW_OK;
}

#include "sw_server.hip"
