// C ABI of the gfx950 Smith-Waterman library (declared in
// include/pypaswas_hip.h).  Host orchestration of one call:
//   1. encode residues into a compact alphabet, upload sequences and the two
//      transposed substitution tables;
//   2. scoring pass over all pairs (plain: padded shapes, static schedule;
//      batched: exact shapes, per-pair tile shape, persistent workers pulling
//      from a device counter, longest pairs first);
//   3. the traceback gate of default_main_cpu.cl:408-410 on the pair maxima;
//   4. direction pass over gated pairs, start selection + marking, walk;
//   5. gather results into one host record.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/pypaswas_hip.h"
#include "sw_device.hip"

using namespace pw;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

}  // namespace

struct pw_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // tile-kernel work that overlaps the lane-per-pair kernels
    hipEvent_t ev[10] = {};
    std::string err;
    bool scoring_set = false;
    std::vector<float> subst;
    int dim = 0, char_offset = 0;
    float gap = 0, ext = 0, ll = 1, minimum = 0;
    bool affine = false;
    DevBuf res, seqs, items, trans, tables, scratch, scratch2, queue, maxima, ditems, dirs, ccount, cands, tbitems,
        sps, spcount, keys, hits, ops, opsused, gated;
    // pinned host buffers for result maxima, recycled by pw_free_result
    std::vector<std::pair<float *, size_t>> pinned_free;
};

static thread_local std::string g_init_err;

namespace {

struct Encoded {
    int K = 0, KP = 8, lkp = 5;
    std::vector<uint8_t> codes;     // all residues, queries then targets
    std::vector<SeqRef> refs;       // nq + nt
    std::vector<float> tables;      // 2 * KP * KP
    int64_t max_len_q = 0, max_len_t = 0;
};

}  // namespace

// chunk widths of the lane-per-pair kernel
constexpr int NB_KINDS = 6;
constexpr int NB_WIDTHS[NB_KINDS] = {16, 32, 48, 64, 80, 96};

// A prepared job: sequences, substitution tables and the scoring work list
// resident on the device, plus the host copies the later stages need.
struct pw_plan {
    Encoded E;
    int32_t nq = 0, nt = 0;
    int64_t n_pairs = 0;
    int32_t mode = 1;
    std::vector<int32_t> pq, pt, qlen, tlen;
    std::vector<float> mss;
    std::vector<uint8_t> same;
    int64_t max_b = 0;
    double cells = 0;
    int64_t n_tile = 0;        // pairs scored by the wave-per-pair tile kernel
    int64_t n_groups = 0;      // lane-per-pair jobs
    int64_t group_rows = 0;    // longest row sequence of a multi-chunk job (scratch sizing)
    int32_t max_nb = 16;
    int64_t nb_begin[NB_KINDS] = {}, nb_end[NB_KINDS] = {};  // job ranges per chunk width
    std::vector<std::pair<int64_t, int64_t>> batches;         // plain launch: item range of each batch
    DevBuf res, seqs, tables, items, trans, gitems, gents, d_mss;
};

#define HIPCHK(ctx, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                 \
            return PW_EDEVICE;                                                              \
        }                                                                                   \
    } while (0)

static int ensure(pw_ctx *ctx, DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return PW_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) {
            ctx->err = std::string("hipFree: ") + hipGetErrorString(e);
            return PW_EDEVICE;
        }
    }
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        ctx->err = "hipMalloc of " + std::to_string(want) + " bytes: " + hipGetErrorString(e);
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
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    if (ctx->num_cus <= 0) ctx->num_cus = 256;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_init_err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        delete ctx;
        return PW_EDEVICE;
    }
    for (auto &ev : ctx->ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) {
            g_init_err = std::string("hipEventCreate: ") + hipGetErrorString(e);
            delete ctx;
            return PW_EDEVICE;
        }
    }
    *out = ctx;
    return PW_OK;
}

extern "C" int pw_destroy(pw_ctx *ctx) {
    if (!ctx) return PW_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DevBuf *bufs[] = {&ctx->res, &ctx->seqs, &ctx->items, &ctx->trans, &ctx->tables, &ctx->scratch, &ctx->scratch2,
                      &ctx->queue,
                      &ctx->maxima, &ctx->ditems, &ctx->dirs, &ctx->ccount, &ctx->cands, &ctx->tbitems, &ctx->sps,
                      &ctx->spcount, &ctx->keys, &ctx->hits, &ctx->ops, &ctx->opsused, &ctx->gated};
    for (DevBuf *b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto &pb : ctx->pinned_free) (void)hipHostFree(pb.first);
    for (auto &ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
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

extern "C" int pw_set_scoring(pw_ctx *ctx, const float *subst, int dim, int char_offset, float gap, float gap_ext,
                              float lower_limit, float minimum_score) {
    if (!ctx) return PW_EINVAL;
    if (!subst || dim <= 0 || dim > 255) {
        ctx->err = "pw_set_scoring: substitution matrix must be dim x dim with 0 < dim <= 255";
        return PW_EINVAL;
    }
    if (!(gap <= 0.0f) || !(gap_ext <= 0.0f)) {
        ctx->err = "pw_set_scoring: gap and gap extension scores must be <= 0";
        return PW_EINVAL;
    }
    if (!(gap > GAP_INIT) || !std::isfinite(lower_limit) || !std::isfinite(minimum_score)) {
        ctx->err = "pw_set_scoring: scoring constants out of range";
        return PW_EINVAL;
    }
    for (int i = 0; i < dim * dim; ++i)
        if (!std::isfinite(subst[i]) || std::fabs(subst[i]) >= 1e9f) {
            ctx->err = "pw_set_scoring: substitution scores must be finite and below 1e9 in magnitude";
            return PW_EINVAL;
        }
    ctx->subst.assign(subst, subst + (size_t)dim * dim);
    ctx->dim = dim;
    ctx->char_offset = char_offset;
    ctx->gap = gap;
    ctx->ext = gap_ext;
    ctx->affine = gap_ext != 0.0f;
    ctx->ll = lower_limit;
    ctx->minimum = minimum_score;
    ctx->scoring_set = true;
    return PW_OK;
}

namespace {

int encode(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
           const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt, Encoded &E) {
    int map[256];
    for (int &m : map) m = -1;
    std::vector<int> used;
    auto scan = [&](const uint8_t *res, const int64_t *off, const int32_t *len, int32_t n) -> int {
        for (int32_t i = 0; i < n; ++i) {
            if (len[i] < 0 || off[i] < 0) {
                ctx->err = "negative sequence length or offset";
                return PW_EINVAL;
            }
            const uint8_t *p = res + off[i];
            for (int32_t k = 0; k < len[i]; ++k) {
                const int c = p[k];
                if (map[c] < 0) {
                    const int idx = c - ctx->char_offset;
                    if (idx < 0 || idx >= ctx->dim) {
                        char buf[160];
                        snprintf(buf, sizeof buf,
                                 "character %d ('%c') is outside the scoring alphabet (offset %d, dimension %d)", c,
                                 (c >= 32 && c < 127) ? c : '?', ctx->char_offset, ctx->dim);
                        ctx->err = buf;
                        return PW_EINVAL;
                    }
                    map[c] = (int)used.size();
                    used.push_back(c);
                }
            }
        }
        return PW_OK;
    };
    int rc = scan(qres, qoff, qlen, nq);
    if (rc) return rc;
    rc = scan(tres, toff, tlen, nt);
    if (rc) return rc;
    E.K = (int)used.size();
    if (E.K > KPMAX - 1) {
        ctx->err = "more than 63 distinct residues in one call";
        return PW_EINVAL;
    }
    E.KP = 8;
    while (E.KP < E.K + 1) E.KP *= 2;
    E.lkp = 2;
    for (int v = E.KP; v > 1; v >>= 1) ++E.lkp;  // log2(KP) + 2: byte shift of a table row
    const int KP = E.KP, FILL = KP - 1;
    E.tables.assign(2 * KP * KP, FILL_SCORE);
    for (int a = 0; a < E.K; ++a)
        for (int b = 0; b < E.K; ++b) {
            const int ia = used[a] - ctx->char_offset, ib = used[b] - ctx->char_offset;
            // table 0: A = query  -> entry [b][a] = S(query a, target b)
            E.tables[b * KP + a] = ctx->subst[(size_t)ia * ctx->dim + ib];
            // table 1: A = target -> entry [b][a] = S(query b, target a)
            E.tables[KP * KP + b * KP + a] = ctx->subst[(size_t)ib * ctx->dim + ia];
        }
    (void)FILL;
    int64_t total = 0;
    for (int32_t i = 0; i < nq; ++i) total += qlen[i];
    for (int32_t i = 0; i < nt; ++i) total += tlen[i];
    E.codes.resize((size_t)total + 64);
    E.refs.resize((size_t)nq + nt);
    int64_t pos = 0;
    for (int32_t i = 0; i < nq; ++i) {
        E.refs[i] = SeqRef{pos, qlen[i], 0};
        for (int32_t k = 0; k < qlen[i]; ++k) E.codes[pos++] = (uint8_t)map[qres[qoff[i] + k]];
        E.max_len_q = std::max<int64_t>(E.max_len_q, qlen[i]);
    }
    for (int32_t i = 0; i < nt; ++i) {
        E.refs[nq + i] = SeqRef{pos, tlen[i], 0};
        for (int32_t k = 0; k < tlen[i]; ++k) E.codes[pos++] = (uint8_t)map[tres[toff[i] + k]];
        E.max_len_t = std::max<int64_t>(E.max_len_t, tlen[i]);
    }
    return PW_OK;
}

inline int pad8(int n) { return (n + 7) / 8 * 8; }

// Tile shape for the batched pass: orientation and residues per lane that
// minimise the modelled cost stripes * (columns + 63) * (R + overhead).
void choose_shape(int qlen, int tlen, int &trans, int &R) {
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
constexpr int GROUP_MAX_ROWS = 1 << 15;

// columns per chunk of a lane-per-pair job: one chunk up to 96 columns,
// chunks of 64 beyond
int group_nb(int clen) {
    if (clen <= 96) return std::max(16, (clen + 15) / 16 * 16);
    return 64;
}

int dir_R(int qlen) {
    int r = 1;
    while (r < 8 && 64 * r < qlen) r *= 2;
    return r;
}

size_t lds_bytes(int KP) { return (size_t)2 * KP * KP * sizeof(float) + BWIN + WAVE; }

int check_pairs(pw_ctx *ctx, int32_t nq, int32_t nt, const int32_t *pq, const int32_t *pt, int64_t n_pairs) {
    if (n_pairs < 0 || n_pairs > (int64_t)INT32_MAX) {
        ctx->err = "n_pairs out of range";
        return PW_EINVAL;
    }
    for (int64_t i = 0; i < n_pairs; ++i)
        if (pq[i] < 0 || pq[i] >= nq || pt[i] < 0 || pt[i] >= nt) {
            ctx->err = "pair " + std::to_string(i) + " indexes outside the sequence lists";
            return PW_EINVAL;
        }
    return PW_OK;
}

void free_buf(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

void free_plan(pw_plan *P) {
    if (!P) return;
    free_buf(P->res);
    free_buf(P->seqs);
    free_buf(P->tables);
    free_buf(P->items);
    free_buf(P->trans);
    free_buf(P->gitems);
    free_buf(P->gents);
    free_buf(P->d_mss);
    delete P;
}

// Encodes and uploads the sequences and builds the scoring work list.
int prepare(pw_ctx *ctx, pw_plan *P, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
            const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt, const int32_t *pq,
            const int32_t *pt, int64_t n_pairs, const float *mss, const uint8_t *same_id, int32_t mode) {
    if (mode != 0 && mode != 1) {
        ctx->err = "mode must be 0 (plain per-batch launch) or 1 (batched persistent)";
        return PW_EINVAL;
    }
    int rc = check_pairs(ctx, nq, nt, pq, pt, n_pairs);
    if (rc) return rc;
    rc = encode(ctx, qres, qoff, qlen, nq, tres, toff, tlen, nt, P->E);
    if (rc) return rc;
    Encoded &E = P->E;
    P->nq = nq;
    P->nt = nt;
    P->n_pairs = n_pairs;
    P->mode = mode;
    P->pq.assign(pq, pq + n_pairs);
    P->pt.assign(pt, pt + n_pairs);
    P->qlen.assign(qlen, qlen + nq);
    P->tlen.assign(tlen, tlen + nt);
    if (mss) P->mss.assign(mss, mss + n_pairs);
    else P->mss.assign((size_t)n_pairs, 0.0f);
    if (same_id) P->same.assign(same_id, same_id + n_pairs);
    else P->same.assign((size_t)n_pairs, 0);
    P->cells = 0;
    for (int64_t i = 0; i < n_pairs; ++i) P->cells += (double)qlen[pq[i]] * (double)tlen[pt[i]];

    std::vector<ScoreItem> items((size_t)n_pairs);
    std::vector<uint8_t> trans((size_t)n_pairs, 0);
    int64_t max_b = 0;
    if (mode == 0) {
        // The reference's batching: consecutive pairs form a batch while the
        // batch's full matrices fit in 80% of device memory (SmithWaterman.py:
        // 202-221, 420-432: 1 direction byte + 4 (linear) or 12 (affine) bytes
        // of FP32 matrix per cell), every pair of a batch padded to the
        // batch's longest query and target rounded up to 8 (:459-461, :325).
        size_t free_b = 0, total_b = 0;
        HIPCHK(ctx, hipMemGetInfo(&free_b, &total_b));
        const double budget = 0.8 * (double)total_b;
        const double bpc = ctx->affine ? 13.0 : 5.0;
        int64_t begin = 0;
        int bx = 0, by = 0;
        auto close = [&](int64_t end) {
            for (int64_t k = begin; k < end; ++k) items[k] = ScoreItem{pq[k], nq + pt[k], (int32_t)k, 4, bx, by};
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
        // lane-per-pair jobs for pairs whose target is shared by enough pairs;
        // the rest (long or unshared pairs) go to the wave-per-pair tile kernel
        std::vector<std::vector<int64_t>> by_target((size_t)nt);
        for (int64_t i = 0; i < n_pairs; ++i) by_target[pt[i]].push_back(i);
        std::vector<int64_t> tile_pairs;
        std::vector<GroupItem> gitems;
        std::vector<GroupEntry> gents;
        std::vector<double> gcost;
        for (int32_t t = 0; t < nt; ++t) {
            std::vector<int64_t> &L = by_target[t];
            if (L.empty()) continue;
            std::vector<int64_t> cand;
            for (int64_t i : L)
                if (qlen[pq[i]] <= GROUP_MAX_ROWS) cand.push_back(i);
                else tile_pairs.push_back(i);
            std::stable_sort(cand.begin(), cand.end(),
                             [&](int64_t x, int64_t y) { return qlen[pq[x]] > qlen[pq[y]]; });
            // keep a 64-lane job only if its lanes are busy for most of its
            // rows (sum of row lengths vs 64 x the longest); otherwise its
            // pairs go to the tile kernel
            std::vector<int64_t> grp;
            for (size_t g0 = 0; g0 < cand.size(); g0 += 64) {
                const size_t g1 = std::min(cand.size(), g0 + 64);
                double sum = 0;
                for (size_t u = g0; u < g1; ++u) sum += qlen[pq[cand[u]]];
                const double fill = sum / (64.0 * std::max(1, qlen[pq[cand[g0]]]));
                for (size_t u = g0; u < g1; ++u) (fill >= GROUP_MIN_FILL ? grp : tile_pairs).push_back(cand[u]);
            }
            const int tl = tlen[t];
            const int nb = group_nb(tl);
            const int nch = tl > nb ? (tl + nb - 1) / nb : 1;
            for (size_t g0 = 0; g0 < grp.size(); g0 += 64) {
                const int cnt = (int)std::min<size_t>(64, grp.size() - g0);
                GroupItem gi;
                gi.c_seq = nq + t;
                gi.entry_off = (int32_t)gents.size();
                gi.count = cnt;
                gi.rows = qlen[pq[grp[g0]]];
                gi.nb = nb;
                gi.trans = 0;
                for (int u = 0; u < cnt; ++u) gents.push_back(GroupEntry{pq[grp[g0 + u]], (int32_t)grp[g0 + u]});
                gitems.push_back(gi);
                gcost.push_back((double)gi.rows * nch * nb);
                if (nch > 1) P->group_rows = std::max<int64_t>(P->group_rows, gi.rows);
                P->max_nb = std::max(P->max_nb, nb);
            }
        }
        // longest jobs first for the persistent workers
        std::vector<int64_t> gorder(gitems.size());
        std::iota(gorder.begin(), gorder.end(), 0);
        std::stable_sort(gorder.begin(), gorder.end(), [&](int64_t x, int64_t y) {
            if (gitems[x].nb != gitems[y].nb) return gitems[x].nb < gitems[y].nb;
            return gcost[x] > gcost[y];
        });
        std::vector<GroupItem> gsorted(gitems.size());
        for (size_t i = 0; i < gorder.size(); ++i) gsorted[i] = gitems[gorder[i]];
        for (int w = 0; w < NB_KINDS; ++w) {
            P->nb_begin[w] = P->nb_end[w] = 0;
            for (size_t i = 0; i < gsorted.size(); ++i)
                if (gsorted[i].nb == NB_WIDTHS[w]) {
                    if (P->nb_end[w] == 0) P->nb_begin[w] = (int64_t)i;
                    P->nb_end[w] = (int64_t)i + 1;
                }
        }
        P->n_groups = (int64_t)gsorted.size();
        ENSURE(ctx, P->gitems, gsorted.size() * sizeof(GroupItem));
        ENSURE(ctx, P->gents, gents.size() * sizeof(GroupEntry));
        if (!gsorted.empty()) {
            HIPCHK(ctx, hipMemcpyAsync(P->gitems.p, gsorted.data(), gsorted.size() * sizeof(GroupItem),
                                       hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(ctx, hipMemcpyAsync(P->gents.p, gents.data(), gents.size() * sizeof(GroupEntry),
                                       hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        }

        const int64_t ntile = (int64_t)tile_pairs.size();
        items.resize((size_t)ntile);
        trans.assign((size_t)ntile, 0);
        std::vector<std::pair<double, int64_t>> order((size_t)ntile);
        for (int64_t k = 0; k < ntile; ++k) {
            const int64_t i = tile_pairs[k];
            const int ql = qlen[pq[i]], tl = tlen[pt[i]];
            int tr, R;
            choose_shape(ql, tl, tr, R);
            const int a = tr ? tl : ql, b = tr ? ql : tl;
            const int aseq = tr ? nq + pt[i] : pq[i], bseq = tr ? pq[i] : nq + pt[i];
            items[k] = ScoreItem{aseq, bseq, (int32_t)i, R, a, b};
            trans[k] = (uint8_t)tr;
            const int rows = 64 * R;
            order[k] = {-(double)((a + rows - 1) / rows) * (b + 63) * (R + 1.5), k};
            max_b = std::max<int64_t>(max_b, b);
        }
        std::stable_sort(order.begin(), order.end(),
                         [](const std::pair<double, int64_t> &x, const std::pair<double, int64_t> &y) {
                             return x.first < y.first;
                         });
        std::vector<ScoreItem> sorted((size_t)ntile);
        std::vector<uint8_t> strans((size_t)ntile);
        for (int64_t k = 0; k < ntile; ++k) {
            sorted[k] = items[order[k].second];
            strans[k] = trans[order[k].second];
        }
        items.swap(sorted);
        trans.swap(strans);
    }
    P->max_b = max_b;
    P->n_tile = (int64_t)items.size();
    ENSURE(ctx, P->res, E.codes.size());
    ENSURE(ctx, P->seqs, E.refs.size() * sizeof(SeqRef));
    ENSURE(ctx, P->tables, E.tables.size() * sizeof(float));
    ENSURE(ctx, P->items, items.size() * sizeof(ScoreItem));
    ENSURE(ctx, P->trans, trans.size());
    HIPCHK(ctx, hipMemcpyAsync(P->res.p, E.codes.data(), E.codes.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(P->seqs.p, E.refs.data(), E.refs.size() * sizeof(SeqRef), hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(P->tables.p, E.tables.data(), E.tables.size() * sizeof(float),
                               hipMemcpyHostToDevice, ctx->stream));
    ENSURE(ctx, P->d_mss, (size_t)n_pairs * sizeof(float));
    if (n_pairs > 0)
        HIPCHK(ctx, hipMemcpyAsync(P->d_mss.p, P->mss.data(), (size_t)n_pairs * sizeof(float),
                                   hipMemcpyHostToDevice, ctx->stream));
    if (!items.empty()) {
        HIPCHK(ctx, hipMemcpyAsync(P->items.p, items.data(), items.size() * sizeof(ScoreItem),
                                   hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(P->trans.p, trans.data(), trans.size(), hipMemcpyHostToDevice, ctx->stream));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return PW_OK;
}

const void *group_kernel_for(bool affine, int nb) {
    switch (nb) {
    case 16: return affine ? (const void *)group_kernel<true, 16> : (const void *)group_kernel<false, 16>;
    case 32: return affine ? (const void *)group_kernel<true, 32> : (const void *)group_kernel<false, 32>;
    case 48: return affine ? (const void *)group_kernel<true, 48> : (const void *)group_kernel<false, 48>;
    case 64: return affine ? (const void *)group_kernel<true, 64> : (const void *)group_kernel<false, 64>;
    case 80: return affine ? (const void *)group_kernel<true, 80> : (const void *)group_kernel<false, 80>;
    default: return affine ? (const void *)group_kernel<true, 96> : (const void *)group_kernel<false, 96>;
    }
}

int persistent_grid(pw_ctx *ctx, const void *kernel, size_t lds, int64_t work) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, lds) != hipSuccess || per_cu <= 0)
        per_cu = 8;
    const int64_t g = (int64_t)ctx->num_cus * per_cu;
    return (int)std::max<int64_t>(1, std::min<int64_t>(work, g));
}

// Scoring pass over all pairs of the plan; maxima land in ctx->maxima.
int run_score(pw_ctx *ctx, pw_plan *P, double *ms_score, int64_t *launches) {
    const int64_t n_pairs = P->n_pairs;
    ENSURE(ctx, ctx->maxima, (size_t)n_pairs * sizeof(float));
    if (P->n_tile > 0) {  // sized before any launch, for the same reason as the group scratch
        const size_t lds = lds_bytes(P->E.KP);
        const int g = persistent_grid(ctx, (const void *)(ctx->affine ? score_kernel<true> : score_kernel<false>), lds,
                                      P->n_tile);
        ENSURE(ctx, ctx->scratch2, (size_t)g * 3 * (P->max_b + 64) * sizeof(float));
    }
    // work counters on separate 64-byte lines: [0] tile jobs, [16*(1+w)] jobs of chunk width w
    ENSURE(ctx, ctx->queue, 64 * (1 + NB_KINDS));
    HIPCHK(ctx, hipMemsetAsync(ctx->queue.p, 0, 64 * (1 + NB_KINDS), ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    if (P->n_groups > 0) {
        const int64_t scr_stride = std::max<int64_t>(1, P->group_rows) * WAVE;
        // size the scratch once for the largest grid: resizing frees memory a
        // kernel launched just before may still be using
        int grid_all = 1;
        for (int w = 0; w < NB_KINDS; ++w) {
            const int64_t n = P->nb_end[w] - P->nb_begin[w];
            if (n <= 0) continue;
            const size_t lds = (size_t)P->E.KP * (NB_WIDTHS[w] + 4) * sizeof(float);
            grid_all = std::max(grid_all, persistent_grid(ctx, group_kernel_for(ctx->affine, NB_WIDTHS[w]), lds, n));
        }
        ENSURE(ctx, ctx->scratch, (size_t)grid_all * scr_stride * sizeof(float2));
        for (int w = 0; w < NB_KINDS; ++w) {
            const int64_t n = P->nb_end[w] - P->nb_begin[w];
            if (n <= 0) continue;
            const void *kern = group_kernel_for(ctx->affine, NB_WIDTHS[w]);
            const size_t lds = (size_t)P->E.KP * (NB_WIDTHS[w] + 4) * sizeof(float);
            const int grid = persistent_grid(ctx, kern, lds, n);
            const GroupItem *items = (const GroupItem *)P->gitems.p + P->nb_begin[w];
            const uint8_t *res = (const uint8_t *)P->res.p;
            const SeqRef *seqs = (const SeqRef *)P->seqs.p;
            const GroupEntry *ents = (const GroupEntry *)P->gents.p;
            const float *tabs = (const float *)P->tables.p;
            int kp = P->E.KP, ni = (int)n;
            float g = ctx->gap, e = ctx->ext;
            float2 *scr = (float2 *)ctx->scratch.p;
            int64_t ss = scr_stride;
            unsigned *q = (unsigned *)ctx->queue.p + 16 * (1 + w);
            float *mx = (float *)ctx->maxima.p;
            void *args[] = {&res, &seqs, &items, &ni, &ents, &tabs, &kp, &g, &e, &scr, &ss, &q, &mx};
            HIPCHK(ctx, hipLaunchKernel(kern, dim3(grid), dim3(64), args, lds, ctx->stream));
            if (launches) *launches += 1;
        }
    }
    if (P->n_tile > 0) {
        auto kern = ctx->affine ? score_kernel<true> : score_kernel<false>;
        const size_t lds = lds_bytes(P->E.KP);
        const int64_t scr_stride = 3 * (P->max_b + 64);
        const int grid_max = persistent_grid(ctx, (const void *)kern, lds, P->n_tile);
        const ScoreItem *items = (const ScoreItem *)P->items.p;
        const uint8_t *trans = (const uint8_t *)P->trans.p;
        if (P->mode == 0) {
            // plain per-batch launches: one launch per batch, static schedule
            for (const auto &b : P->batches) {
                const int64_t n = b.second - b.first;
                const int grid = (int)std::min<int64_t>(n, grid_max);
                hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, ctx->stream, (const uint8_t *)P->res.p,
                                   (const SeqRef *)P->seqs.p, items + b.first, (int)n, trans + b.first,
                                   (const float *)P->tables.p, P->E.KP, P->E.lkp, ctx->gap, ctx->ext,
                                   (float *)ctx->scratch2.p, scr_stride, (unsigned *)nullptr,
                                   (float *)ctx->maxima.p);
                HIPCHK(ctx, hipGetLastError());
                if (launches) *launches += 1;
            }
        } else {
            // persistent workers on the second stream, overlapping the lane-per-pair kernels
            hipStream_t st = P->n_groups > 0 ? ctx->stream2 : ctx->stream;
            if (st != ctx->stream) HIPCHK(ctx, hipStreamWaitEvent(st, ctx->ev[0], 0));
            hipLaunchKernelGGL(kern, dim3(grid_max), dim3(64), lds, st, (const uint8_t *)P->res.p,
                               (const SeqRef *)P->seqs.p, items, (int)P->n_tile, trans, (const float *)P->tables.p,
                               P->E.KP, P->E.lkp, ctx->gap, ctx->ext, (float *)ctx->scratch2.p, scr_stride,
                               (unsigned *)ctx->queue.p, (float *)ctx->maxima.p);
            HIPCHK(ctx, hipGetLastError());
            if (launches) *launches += 1;
            if (st != ctx->stream) {
                HIPCHK(ctx, hipEventRecord(ctx->ev[8], st));
                HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev[8], 0));
            }
        }
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    (void)ms_score;  // read from events ev[0]..ev[1] once the stream has been synchronised
    return PW_OK;
}

int score_ms(pw_ctx *ctx, double *ms_score) {
    float ms = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    if (ms_score) *ms_score += ms;
    return PW_OK;
}

struct HostResult {
    pw_result pub;
    float *maxima = nullptr;  // pinned host memory of maxima_cap floats
    size_t maxima_cap = 0;
    std::vector<pw_starting_point> sps;
    std::vector<pw_hit> hits;
    std::vector<char> ops;
};

int execute(pw_ctx *ctx, pw_plan *P, double minimum, int32_t sp_cap, HostResult *R) {
    const auto wall0 = std::chrono::steady_clock::now();
    const int64_t n_pairs = P->n_pairs;
    const int32_t nq = P->nq;
    const int32_t *pq = P->pq.data(), *pt = P->pt.data(), *qlen = P->qlen.data(), *tlen = P->tlen.data();
    const float *min_start_score = P->mss.data();
    Encoded &E = P->E;
    double ms_score = 0;
    int64_t launches = 0;
    int rc = run_score(ctx, P, &ms_score, &launches);
    if (rc) return rc;
    // traceback gate on the device: maximum >= MINIMUM_SCORE (default_main_cpu.cl:410)
    // and >= the pair's minimum start score, without which no cell can start (:427)
    ENSURE(ctx, ctx->gated, (size_t)(n_pairs + 16) * sizeof(int));
    HIPCHK(ctx, hipMemsetAsync(ctx->gated.p, 0, sizeof(int), ctx->stream));
    if (n_pairs > 0) {
        const int grid = (int)std::min<int64_t>((n_pairs + 255) / 256, (int64_t)ctx->num_cus * 16);
        hipLaunchKernelGGL(gate_kernel, dim3(grid), dim3(256), 0, ctx->stream, (const float *)ctx->maxima.p,
                           (const float *)P->d_mss.p, n_pairs, ctx->minimum, (int *)ctx->gated.p,
                           (int *)ctx->gated.p + 16);
        HIPCHK(ctx, hipGetLastError());
    }
    if (n_pairs > 0) {
        // reuse a pinned buffer from an earlier, freed result when one is large enough
        for (size_t i = 0; i < ctx->pinned_free.size(); ++i)
            if (ctx->pinned_free[i].second >= (size_t)n_pairs) {
                R->maxima = ctx->pinned_free[i].first;
                R->maxima_cap = ctx->pinned_free[i].second;
                ctx->pinned_free.erase(ctx->pinned_free.begin() + i);
                break;
            }
        if (!R->maxima) {
            HIPCHK(ctx, hipHostMalloc((void **)&R->maxima, (size_t)n_pairs * sizeof(float), hipHostMallocDefault));
            R->maxima_cap = (size_t)n_pairs;
        }
        HIPCHK(ctx, hipMemcpyAsync(R->maxima, ctx->maxima.p, (size_t)n_pairs * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream));
    }
    int ngated = 0;
    HIPCHK(ctx, hipMemcpyAsync(&ngated, ctx->gated.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    rc = score_ms(ctx, &ms_score);
    if (rc) return rc;
    std::vector<int64_t> gated((size_t)ngated);
    if (ngated > 0) {
        std::vector<int> g32((size_t)ngated);
        HIPCHK(ctx, hipMemcpy(g32.data(), (int *)ctx->gated.p + 16, (size_t)ngated * sizeof(int),
                              hipMemcpyDeviceToHost));
        std::sort(g32.begin(), g32.end());
        for (int i = 0; i < ngated; ++i) gated[i] = g32[i];
    }
    const int ng = (int)gated.size();
    double ms_dir = 0, ms_tb = 0, ms_walk = 0;
    if (ng > 0) {
        std::vector<TbItem> tb(ng);
        std::vector<int> spc(ng, 0);
        std::vector<int64_t> opsused(ng, 0);
        // --- direction pass ---
        std::vector<DirItem> di(ng);
        int64_t dir_total = 0, cand_total = 0;
        const int CAND0 = 16;
        int64_t max_t = 0;
        for (int g = 0; g < ng; ++g) {
            const int64_t p = gated[g];
            const int ql = qlen[pq[p]], tl = tlen[pt[p]];
            const int Rr = dir_R(ql);
            const int rows = 64 * Rr, stripes = std::max(1, (ql + rows - 1) / rows);
            DirItem d;
            d.q_seq = pq[p];
            d.t_seq = nq + pt[p];
            d.pair = (int32_t)p;
            d.R = Rr;
            d.dir_off = dir_total;
            d.t1 = ctx->ll * R->maxima[p];  // LOWER_LIMIT_SCORE * maximum, f32
            d.mp = min_start_score[p];
            d.cand_off = cand_total;
            d.cand_cap = CAND0;
            d.pad = 0;
            di[g] = d;
            dir_total += (int64_t)stripes * (tl + 63) * 64 * Rr;
            dir_total = (dir_total + 15) & ~(int64_t)15;
            cand_total += CAND0;
            max_t = std::max<int64_t>(max_t, tl);
        }
        const size_t lds = lds_bytes(E.KP);
        const int64_t scr_stride = 3 * (max_t + 64);
        auto run_dir = [&](const std::vector<DirItem> &list, double &ms) -> int {
            const int n = (int)list.size();
            const int grid = std::max(1, std::min(n, ctx->num_cus * 8));
            ENSURE(ctx, ctx->ditems, list.size() * sizeof(DirItem));
            ENSURE(ctx, ctx->scratch, (size_t)grid * scr_stride * sizeof(float));
            HIPCHK(ctx, hipMemcpyAsync(ctx->ditems.p, list.data(), list.size() * sizeof(DirItem),
                                       hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(ctx, hipMemsetAsync(ctx->ccount.p, 0, (size_t)n * sizeof(int), ctx->stream));
            HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
            auto kern = ctx->affine ? dir_kernel<true> : dir_kernel<false>;
            hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, ctx->stream, (const uint8_t *)P->res.p,
                               (const SeqRef *)P->seqs.p, (const DirItem *)ctx->ditems.p, n,
                               (const float *)P->tables.p, E.KP, E.lkp, ctx->gap, ctx->ext, (float *)ctx->scratch.p,
                               scr_stride, (uint8_t *)ctx->dirs.p, (int *)ctx->ccount.p, (Cand *)ctx->cands.p);
            HIPCHK(ctx, hipGetLastError());
            HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
            HIPCHK(ctx, hipEventSynchronize(ctx->ev[3]));
            float t = 0;
            HIPCHK(ctx, hipEventElapsedTime(&t, ctx->ev[2], ctx->ev[3]));
            ms += t;
            return PW_OK;
        };
        ENSURE(ctx, ctx->dirs, (size_t)dir_total);
        ENSURE(ctx, ctx->ccount, (size_t)ng * sizeof(int));
        ENSURE(ctx, ctx->cands, (size_t)cand_total * sizeof(Cand));
        rc = run_dir(di, ms_dir);
        if (rc) return rc;
        std::vector<int> cc(ng);
        HIPCHK(ctx, hipMemcpy(cc.data(), ctx->ccount.p, (size_t)ng * sizeof(int), hipMemcpyDeviceToHost));
        // pairs whose candidates overflowed the first guess run again with the
        // exact capacity; the recomputed directions are identical
        bool over = false;
        for (int g = 0; g < ng; ++g)
            if (cc[g] > di[g].cand_cap) over = true;
        if (over) {
            int64_t ctot = 0;
            for (int g = 0; g < ng; ++g) {
                const int cap = std::max(cc[g], 1);
                di[g].cand_off = ctot;
                di[g].cand_cap = cap;
                ctot += cap;
            }
            ENSURE(ctx, ctx->cands, (size_t)ctot * sizeof(Cand));
            rc = run_dir(di, ms_dir);
            if (rc) return rc;
            HIPCHK(ctx, hipMemcpy(cc.data(), ctx->ccount.p, (size_t)ng * sizeof(int), hipMemcpyDeviceToHost));
        }
        // --- start selection + marking ---
        int64_t sp_total = 0;
        for (int g = 0; g < ng; ++g) {
            const int64_t p = gated[g];
            TbItem t;
            t.pair = (int32_t)p;
            t.q_seq = pq[p];
            t.t_seq = nq + pt[p];
            t.R = di[g].R;
            t.dir_off = di[g].dir_off;
            t.cand_off = di[g].cand_off;
            t.sp_off = sp_total;
            t.ops_off = 0;
            t.hit_off = sp_total;
            t.key_off = sp_total;
            t.maxv = R->maxima[p];
            t.mp = min_start_score[p];
            t.qidx = pq[p];
            t.tidx = pt[p];
            t.same_id = P->same[p];
            t.ncand = std::min(cc[g], di[g].cand_cap);
            tb[g] = t;
            sp_total += t.ncand;
        }
        ENSURE(ctx, ctx->tbitems, (size_t)ng * sizeof(TbItem));
        ENSURE(ctx, ctx->sps, (size_t)sp_total * sizeof(SpRec));
        ENSURE(ctx, ctx->spcount, (size_t)ng * sizeof(int));
        HIPCHK(ctx, hipMemcpyAsync(ctx->tbitems.p, tb.data(), tb.size() * sizeof(TbItem), hipMemcpyHostToDevice,
                                   ctx->stream));
        HIPCHK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
        const int tb_grid = (ng + 63) / 64;
        auto tbk = ctx->affine ? tb_kernel<true> : tb_kernel<false>;
        hipLaunchKernelGGL(tbk, dim3(tb_grid), dim3(64), 0, ctx->stream, (const SeqRef *)P->seqs.p,
                           (const TbItem *)ctx->tbitems.p, ng, (uint8_t *)ctx->dirs.p, (Cand *)ctx->cands.p,
                           (SpRec *)ctx->sps.p, (int *)ctx->spcount.p);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(spc.data(), ctx->spcount.p, (size_t)ng * sizeof(int), hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        {
            float t = 0;
            HIPCHK(ctx, hipEventElapsedTime(&t, ctx->ev[4], ctx->ev[5]));
            ms_tb += t;
        }
        // --- walk ---
        int64_t ops_total = 0;
        for (int g = 0; g < ng; ++g) {
            const int64_t p = gated[g];
            const int64_t ql = qlen[pq[p]], tl = tlen[pt[p]];
            const int64_t cap = std::min<int64_t>((int64_t)spc[g] * (ql + tl), ql * tl + spc[g]);
            tb[g].ops_off = ops_total;
            ops_total += cap;
        }
        ENSURE(ctx, ctx->keys, (size_t)sp_total * sizeof(uint64_t));
        ENSURE(ctx, ctx->hits, (size_t)sp_total * sizeof(HitRec));
        ENSURE(ctx, ctx->ops, (size_t)ops_total + 1);
        ENSURE(ctx, ctx->opsused, (size_t)ng * sizeof(int64_t));
        HIPCHK(ctx, hipMemcpyAsync(ctx->tbitems.p, tb.data(), tb.size() * sizeof(TbItem), hipMemcpyHostToDevice,
                                   ctx->stream));
        HIPCHK(ctx, hipEventRecord(ctx->ev[6], ctx->stream));
        auto wk = ctx->affine ? walk_kernel<true> : walk_kernel<false>;
        hipLaunchKernelGGL(wk, dim3(tb_grid), dim3(64), 0, ctx->stream, (const uint8_t *)P->res.p,
                           (const SeqRef *)P->seqs.p, (const TbItem *)ctx->tbitems.p, ng, (uint8_t *)ctx->dirs.p,
                           (const SpRec *)ctx->sps.p, (const int *)ctx->spcount.p, (uint64_t *)ctx->keys.p,
                           (HitRec *)ctx->hits.p, (char *)ctx->ops.p, (int64_t *)ctx->opsused.p, minimum);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(ctx->ev[7], ctx->stream));
        std::vector<SpRec> hsp((size_t)sp_total);
        std::vector<HitRec> hh((size_t)sp_total);
        std::vector<char> hops((size_t)ops_total);
        if (sp_total > 0) {
            HIPCHK(ctx, hipMemcpyAsync(hsp.data(), ctx->sps.p, (size_t)sp_total * sizeof(SpRec),
                                       hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(ctx, hipMemcpyAsync(hh.data(), ctx->hits.p, (size_t)sp_total * sizeof(HitRec),
                                       hipMemcpyDeviceToHost, ctx->stream));
        }
        if (ops_total > 0)
            HIPCHK(ctx, hipMemcpyAsync(hops.data(), ctx->ops.p, (size_t)ops_total, hipMemcpyDeviceToHost,
                                       ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(opsused.data(), ctx->opsused.p, (size_t)ng * sizeof(int64_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        {
            float t = 0;
            HIPCHK(ctx, hipEventElapsedTime(&t, ctx->ev[6], ctx->ev[7]));
            ms_walk += t;
        }
        // per gated pair: starting points (emission order), hits (walk order), op characters
        static_assert(sizeof(SpRec) == sizeof(pw_starting_point), "starting point layout");
        for (int g = 0; g < ng; ++g) {
            const int64_t sp0 = (int64_t)R->sps.size();
            const int n = spc[g];
            for (int k = 0; k < n; ++k) {
                pw_starting_point sp;
                std::memcpy(&sp, &hsp[tb[g].sp_off + k], sizeof sp);
                R->sps.push_back(sp);
            }
            const int64_t op0 = (int64_t)R->ops.size();
            R->ops.insert(R->ops.end(), hops.begin() + tb[g].ops_off, hops.begin() + tb[g].ops_off + opsused[g]);
            for (int k = 0; k < n; ++k) {
                const HitRec &hr = hh[tb[g].hit_off + k];
                pw_hit h;
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
                R->hits.push_back(h);
            }
        }
    }
    // the reference keeps at most sp_cap starting points per pair of a batch in
    // total (SmithWaterman.py:683-685); report how many a cap would drop
    int64_t truncated = 0;
    if (sp_cap > 0) {
        const int64_t cap_total = (int64_t)sp_cap * n_pairs;
        if ((int64_t)R->sps.size() >= cap_total) truncated = (int64_t)R->sps.size() - cap_total;
    }
    pw_result &O = R->pub;
    O.n_pairs = n_pairs;
    O.maxima = R->maxima;
    O.n_starting_points = (int64_t)R->sps.size();
    O.starting_points = R->sps.data();
    O.n_hits = (int64_t)R->hits.size();
    O.hits = R->hits.data();
    O.n_ops = (int64_t)R->ops.size();
    O.ops = R->ops.data();
    O.n_traceback_pairs = ng;
    O.sp_truncated = truncated;
    O.cells = P->cells;
    O.ms_score = ms_score;
    O.ms_directions = ms_dir;
    O.ms_traceback = ms_tb;
    O.ms_walk = ms_walk;
    O.score_launches = launches;
    O.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    return PW_OK;
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

}  // namespace

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
    rc = prepare(ctx, P, qres, qoff, qlen, nq, tres, toff, tlen, nt, pq, pt, n_pairs, min_start_score, same_id,
                 mode);
    if (rc) {
        free_plan(P);
        return rc;
    }
    *out = P;
    return PW_OK;
}

extern "C" int pw_execute(pw_ctx *ctx, pw_plan *plan, double minimum, int32_t sp_cap, pw_result **out) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (!plan || !out) {
        ctx->err = "pw_execute: missing argument";
        return PW_EINVAL;
    }
    *out = nullptr;
    HostResult *R = new HostResult();
    std::memset(&R->pub, 0, sizeof(R->pub));
    rc = execute(ctx, plan, minimum, sp_cap, R);
    if (rc) {
        pw_free_result(ctx, &R->pub);
        return rc;
    }
    *out = &R->pub;
    return PW_OK;
}

extern "C" int pw_release_plan(pw_ctx *ctx, pw_plan *plan) {
    if (!ctx || !plan) return PW_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_plan(plan);
    return PW_OK;
}

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
    pw_plan *P = new pw_plan();
    rc = prepare(ctx, P, qres, qoff, qlen, nq, tres, toff, tlen, nt, pq, pt, n_pairs, nullptr, nullptr, mode);
    double ms = 0;
    if (!rc) rc = run_score(ctx, P, &ms, nullptr);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            ctx->err = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
            rc = PW_EDEVICE;
        }
    }
    if (!rc) rc = score_ms(ctx, &ms);
    if (!rc && n_pairs > 0) {
        hipError_t e = hipMemcpy(maxima_out, ctx->maxima.p, (size_t)n_pairs * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            ctx->err = std::string("hipMemcpy maxima: ") + hipGetErrorString(e);
            rc = PW_EDEVICE;
        }
    }
    free_plan(P);
    if (!rc && ms_out) *ms_out = ms;
    return rc;
}

extern "C" int pw_align_pairs(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
                              const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt,
                              const int32_t *pq, const int32_t *pt, int64_t n_pairs, const float *min_start_score,
                              const uint8_t *same_id, double minimum, int32_t sp_cap, int32_t mode, pw_result **out) {
    int rc = ready(ctx);
    if (rc) return rc;
    if (!out || (n_pairs > 0 && !min_start_score)) {
        ctx->err = "pw_align_pairs: missing argument";
        return PW_EINVAL;
    }
    pw_plan *P = nullptr;
    rc = pw_prepare(ctx, qres, qoff, qlen, nq, tres, toff, tlen, nt, pq, pt, n_pairs, min_start_score, same_id, mode,
                    &P);
    if (rc) return rc;
    rc = pw_execute(ctx, P, minimum, sp_cap, out);
    free_plan(P);
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
    }
    delete R;
    return PW_OK;
}
