// Serving loop of the gfx950 Smith-Waterman library: continuous batching of
// alignment jobs over several stream lanes.  Included at the end of
// sw_capi.hip (one translation unit).
//
// A job is what one SmithWaterman.align_sequences pass handles
// (pyPaSWAS/Core/SmithWaterman.py:441-560): query and target sequences, the
// (query, target) pairs to align, their minimum start scores and one scoring
// scheme.  pw_server_poll admits queued jobs whenever a lane is idle and the
// direction arena has room: the queue head and every later queued job with
// the same scheme and minimum score are merged into one batch (one device
// execution), up to a pair limit and to the largest free range of the arena.
// Each job reserves the direction bytes its pairs would need if all of them
// passed the traceback gate, so admission never over-commits the arena.
// Batches on different lanes run concurrently on the device; each lane is one
// stream (one hardware queue), and jobs above LARGE_JOB_CELLS run on lane 0
// only, with three quarters of the arena, so that long alignments never queue
// in front of short ones nor take their direction memory.  Pairs are
// independent, so a job's records are the same whichever batch it runs in;
// they are split back per job with job-local sequence indices.
#include <deque>
#include <map>
#include <unordered_map>

namespace {

constexpr int SERVE_SCHEMES = 8;
constexpr double LARGE_JOB_CELLS = 5e7;
constexpr double LARGE_LANE_GRID_SHARE = 0.5;
constexpr int64_t ARENA_ALIGN = 256;

// First-fit range allocator over [base, base + cap), with coalescing on release.
struct RangeAlloc {
    std::map<int64_t, int64_t> free_;  // offset -> length
    int64_t cap = 0;

    void reset(int64_t base, int64_t c) {
        cap = c;
        free_.clear();
        if (c > 0) free_[base] = c;
    }
    int64_t largest() const {
        int64_t m = 0;
        for (const auto &f : free_) m = std::max(m, f.second);
        return m;
    }
    bool all_free() const { return free_.size() == 1 && free_.begin()->second == cap; }
    int64_t take(int64_t len) {
        len = (len + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= len) {
                const int64_t off = it->first, rest = it->second - len;
                free_.erase(it);
                if (rest > 0) free_[off + len] = rest;
                return off;
            }
        return -1;
    }
    void give(int64_t off, int64_t len) {
        len = (len + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
        auto it = free_.emplace(off, len).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct ServeJob {
    int64_t id = 0;
    int32_t scheme = 0;
    double minimum = 0;
    int32_t sp_cap = 0;
    std::vector<uint8_t> qres, tres;
    std::vector<int64_t> qoff, toff;
    std::vector<int32_t> qlen, tlen, pq, pt;
    std::vector<float> mss;
    std::vector<uint8_t> same;
    int64_t dir_need = 0;  // direction bytes if every pair passed the gate
    double cells = 0;
    // placement in its batch
    int32_t q0 = 0, t0 = 0;
    int64_t p0 = 0;
    HostResult *result = nullptr;
    bool done = false;
};

struct ServeLane {
    Lane L;
    pw_plan *P = nullptr;
    std::vector<ServeJob *> jobs;
    int64_t dir_off = -1, dir_len = 0;
    hipEvent_t done_ev = nullptr;
    bool busy = false;
    Exec X;
    PinBuf h_max;  // pair maxima of the batch
    uint64_t seq = 0;  // launch order
};

bool same_key(const ServeJob *a, const ServeJob *b) { return a->scheme == b->scheme && a->minimum == b->minimum; }

bool is_large(const ServeJob *j) { return j->cells > LARGE_JOB_CELLS; }

// Lane 0 takes any job, the others small jobs only.
size_t large_lanes(size_t nlanes) { return nlanes < 1 ? 0 : 1; }
bool lane_takes(size_t lane, size_t nlanes, const ServeJob *j) { return lane < large_lanes(nlanes) || !is_large(j); }
// arena allocator of a lane: large-capable lanes share ra[0]
int lane_pool(size_t lane, size_t nlanes) { return lane < large_lanes(nlanes) ? 0 : 1; }

}  // namespace

struct pw_server {
    pw_ctx *ctx = nullptr;
    Scoring schemes[SERVE_SCHEMES];
    bool scheme_set[SERVE_SCHEMES] = {};
    std::vector<ServeLane> lanes;
    DevBuf arena;
    // direction arena: lane 0 (the only lane for large jobs) allocates from
    // ra[0], the other lanes from ra[1], so that large jobs cannot take the
    // room short ones need
    RangeAlloc ra[2];
    int64_t arena_cap = 0;
    std::deque<ServeJob *> queue;
    std::unordered_map<int64_t, ServeJob *> jobs;
    std::vector<int64_t> completed;
    int64_t next_id = 1;
    int64_t max_batch_pairs = 1 << 20;
    uint64_t launch_seq = 0;
    // statistics: batches launched, jobs admitted, jobs completed, arena bytes reserved (peak, now)
    int64_t n_batches = 0, n_admitted = 0, n_completed = 0, peak_reserved = 0, reserved = 0;
    // host time spent forming and launching batches, and reading them back; device time of the batches
    double us_launch = 0, us_finish = 0, us_device = 0;
};

double now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {

int serve_fail(pw_server *S, const std::string &msg, int rc) {
    S->ctx->err = msg;
    return rc;
}

// Merges the jobs into the lane's plan, uploads it and launches the batch.
int launch_batch(pw_server *S, ServeLane &ln, std::vector<ServeJob *> &batch, int64_t dir_off, int64_t dir_len) {
    pw_ctx *ctx = S->ctx;
    const double t0 = now_us();
    const Scoring &sc = S->schemes[batch.front()->scheme];
    std::vector<uint8_t> qres, tres;
    std::vector<int64_t> qoff, toff;
    std::vector<int32_t> qlen, tlen, pq, pt;
    std::vector<float> mss;
    std::vector<uint8_t> same;
    for (ServeJob *j : batch) {
        j->q0 = (int32_t)qlen.size();
        j->t0 = (int32_t)tlen.size();
        j->p0 = (int64_t)pq.size();
        const int64_t qb = (int64_t)qres.size(), tb = (int64_t)tres.size();
        qres.insert(qres.end(), j->qres.begin(), j->qres.end());
        tres.insert(tres.end(), j->tres.begin(), j->tres.end());
        for (size_t i = 0; i < j->qlen.size(); ++i) {
            qoff.push_back(qb + j->qoff[i]);
            qlen.push_back(j->qlen[i]);
        }
        for (size_t i = 0; i < j->tlen.size(); ++i) {
            toff.push_back(tb + j->toff[i]);
            tlen.push_back(j->tlen[i]);
        }
        for (size_t i = 0; i < j->pq.size(); ++i) {
            pq.push_back(j->q0 + j->pq[i]);
            pt.push_back(j->t0 + j->pt[i]);
        }
        mss.insert(mss.end(), j->mss.begin(), j->mss.end());
        same.insert(same.end(), j->same.begin(), j->same.end());
    }
    if (qres.empty()) qres.push_back(0);
    if (tres.empty()) tres.push_back(0);
    pw_plan *P = ln.P;
    PlanOpts opt = plan_opts(ctx, sc);
    // the large-job lane's persistent kernels take at most half of the
    // device's resident waves, so that the small-job lanes always find room
    if (&ln == &S->lanes[0] && large_lanes(S->lanes.size()) > 0) opt.grid_share = LARGE_LANE_GRID_SHARE;
    int rc = plan_host(ctx->err, sc, opt, P, qres.data(), qoff.data(), qlen.data(), (int32_t)qlen.size(),
                       tres.data(), toff.data(), tlen.data(), (int32_t)tlen.size(), pq.data(), pt.data(),
                       (int64_t)pq.size(), mss.data(), same.data(), 1);
    if (rc) return rc;
    // the lane's previous batch (and with it the copy out of this staging) has finished
    rc = plan_upload(ctx, P, ln.L.s, P->h_blob, false);
    if (rc) return rc;
    Workspace &W = P->own;
    W.dir_ext = (uint8_t *)S->arena.p + dir_off;
    W.dir_ext_cap = dir_len;
    ENSURE_PINNED(ctx, ln.h_max, (size_t)std::max<int64_t>(1, P->n_pairs) * sizeof(float));
    rc = execute_launch(ctx, P, W, ln.L, batch.front()->minimum, false, (float *)ln.h_max.p, ln.X);
    if (rc) return rc;
    HIPCHK(ctx, hipEventRecord(ln.done_ev, ln.L.s));
    ln.jobs = batch;
    ln.dir_off = dir_off;
    ln.dir_len = dir_len;
    ln.busy = true;
    ln.seq = ++S->launch_seq;
    S->n_batches += 1;
    S->n_admitted += (int64_t)batch.size();
    S->us_launch += now_us() - t0;
    return PW_OK;
}

// One job's records out of its batch's (pair-sorted) records.
int split_job(pw_ctx *ctx, const Finished &F, const ServeLane &ln, ServeJob *j, size_t &cursor) {
    HostResult *R = nullptr;
    new_result(ctx, &R);
    const int64_t p1 = j->p0 + (int64_t)j->pq.size();
    std::vector<PairOut> mine;
    while (cursor < F.outs.size() && F.outs[cursor].pair < p1) {
        if (F.outs[cursor].pair >= j->p0) mine.push_back(F.outs[cursor]);
        ++cursor;
    }
    for (PairOut &o : mine) o.pair -= (int32_t)j->p0;
    const int rc = assemble(ctx, mine, j->sp_cap, (int64_t)j->pq.size(), ln.P, R);
    if (rc) {
        (void)pw_free_result(ctx, &R->pub);
        return rc;
    }
    for (int64_t k = 0; k < R->n_sps; ++k) {
        pw_starting_point &sp = R->sps[(size_t)k];
        sp.sequence -= (uint32_t)j->q0;
        sp.target -= (uint32_t)j->t0;
        pw_hit &h = R->hits[(size_t)k];
        h.query -= j->q0;
        h.target -= j->t0;
    }
    R->maxima_vec.assign((const float *)ln.h_max.p + j->p0, (const float *)ln.h_max.p + p1);
    Finished stats;
    stats.ms_score = F.ms_score;
    stats.ms_dir = F.ms_dir;
    stats.ms_tb = F.ms_tb;
    stats.ms_walk = F.ms_walk;
    stats.ms_readback = F.ms_readback;
    stats.ms_tail = F.ms_tail;
    stats.ms_assemble = F.ms_assemble;
    fill_public(R, R->maxima_vec.data(), (int64_t)j->pq.size(), j->cells, (int64_t)mine.size(), stats, ln.X);
    j->result = R;
    j->done = true;
    return PW_OK;
}

int finish_batch(pw_server *S, ServeLane &ln) {
    pw_ctx *ctx = S->ctx;
    const double t0 = now_us();
    Finished F;
    int rc = execute_finish(ctx, ln.P, ln.P->own, ln.L, ln.X, F);
    if (rc) return rc;
    S->us_device += 1e3 * elapsed(ln.L.ev[0], ln.L.ev[5]);
    size_t cursor = 0;
    for (ServeJob *j : ln.jobs) {
        rc = split_job(ctx, F, ln, j, cursor);
        if (rc) return rc;
        S->completed.push_back(j->id);
        S->n_completed += 1;
    }
    // the lent range goes back to the arena; the plan keeps its own buffers for the next batch
    ln.P->own.dir_ext = nullptr;
    ln.P->own.dir_ext_cap = 0;
    S->ra[lane_pool((size_t)(&ln - &S->lanes[0]), S->lanes.size())].give(ln.dir_off, ln.dir_len);
    S->reserved -= ln.dir_len;
    ln.jobs.clear();
    ln.busy = false;
    S->us_finish += now_us() - t0;
    return PW_OK;
}

// Admits queued jobs onto idle lanes while the arena has room.
int admit(pw_server *S) {
    // small-job lanes first, so that the large-capable lanes stay free for
    // large jobs when the others suffice
    const size_t nl = S->lanes.size();
    const size_t nlarge = large_lanes(nl);
    for (size_t li = 0; li < nl; ++li) {
        const size_t lane = (li + nlarge) % nl;
        ServeLane &ln = S->lanes[lane];
        if (ln.busy) continue;
        if (S->queue.empty()) break;
        ServeJob *head = nullptr;
        for (ServeJob *j : S->queue)
            if (lane_takes(lane, nl, j)) {
                head = j;
                break;
            }
        if (!head) continue;
        RangeAlloc &ra = S->ra[lane_pool(lane, nl)];
        const int64_t largest = ra.largest();
        const int64_t cap = ra.cap;
        // a job larger than the whole arena runs alone once the arena is free;
        // its pairs that do not fit finish in tail rounds
        if (head->dir_need > largest && !(head->dir_need > cap && ra.all_free())) continue;
        std::vector<ServeJob *> batch;
        std::deque<ServeJob *> rest;
        int64_t need = 0, pairs = 0;
        for (ServeJob *j : S->queue) {
            const bool take = j == head || (!batch.empty() && same_key(j, head) && lane_takes(lane, nl, j) &&
                                            need + j->dir_need <= largest &&
                                            pairs + (int64_t)j->pq.size() <= S->max_batch_pairs);
            if (take) {
                batch.push_back(j);
                need += j->dir_need;
                pairs += (int64_t)j->pq.size();
            } else {
                rest.push_back(j);
            }
        }
        const int64_t len = std::max<int64_t>(ARENA_ALIGN, std::min(need, cap));
        const int64_t off = ra.take(len);
        if (off < 0) return serve_fail(S, "serving arena: no free range for an admitted batch", PW_EDEVICE);
        S->queue.swap(rest);
        S->reserved += (len + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
        S->peak_reserved = std::max(S->peak_reserved, S->reserved);
        int rc = launch_batch(S, ln, batch, off, (len + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN);
        if (rc) return rc;
    }
    return PW_OK;
}

int check_alphabet(pw_ctx *ctx, const Scoring &sc, const uint8_t *res, const int64_t *off, const int32_t *len,
                   int32_t n) {
    for (int32_t i = 0; i < n; ++i) {
        if (len[i] < 0 || off[i] < 0) {
            ctx->err = "negative sequence length or offset";
            return PW_EINVAL;
        }
        for (int32_t k = 0; k < len[i]; ++k) {
            const int idx = (int)res[off[i] + k] - sc.char_offset;
            if (idx < 0 || idx >= sc.dim) {
                ctx->err = "character outside the scoring alphabet of the job's scheme";
                return PW_EINVAL;
            }
        }
    }
    return PW_OK;
}

}  // namespace

extern "C" int pw_server_create(pw_ctx *ctx, int32_t lanes, int64_t direction_bytes, pw_server **out) {
    if (!ctx || !out) return PW_EINVAL;
    *out = nullptr;
    if (lanes < 1 || lanes > 16 || direction_bytes < (int64_t)ARENA_ALIGN) {
        ctx->err = "pw_server_create: lanes must be 1..16 and the direction arena at least 256 bytes";
        return PW_EINVAL;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    pw_server *S = new pw_server();
    S->ctx = ctx;
    const int64_t cap = direction_bytes / ARENA_ALIGN * ARENA_ALIGN;
    int rc = ensure(ctx, S->arena, (size_t)cap);
    if (rc) {
        delete S;
        return rc;
    }
    S->arena_cap = cap;
    if (lanes == 1) {
        S->ra[0].reset(0, cap);
    } else {
        const int64_t small = std::max<int64_t>(ARENA_ALIGN, cap / 4 / ARENA_ALIGN * ARENA_ALIGN);
        S->ra[0].reset(0, cap - small);
        S->ra[1].reset(cap - small, small);
    }
    S->lanes.resize((size_t)lanes);
    ctx->defer_free = true;
    for (ServeLane &ln : S->lanes) {
        rc = create_lane(ln.L, ctx->err, true);
        hipError_t e = rc ? hipSuccess : hipEventCreateWithFlags(&ln.done_ev, hipEventDisableTiming);
        if (rc || e != hipSuccess) {
            if (!rc) ctx->err = std::string("hipEventCreate: ") + hipGetErrorString(e);
            for (ServeLane &l2 : S->lanes) {
                destroy_lane(l2.L);
                if (l2.done_ev) (void)hipEventDestroy(l2.done_ev);
                if (l2.P) free_plan(ctx, l2.P);
            }
            free_buf(ctx, S->arena);
            ctx->defer_free = false;
            release_dead(ctx);
            delete S;
            return PW_EDEVICE;
        }
        ln.P = new pw_plan();
        // each lane keeps its work areas between batches: start them small
        ln.P->own.want_dir = 0;
    }
    *out = S;
    return PW_OK;
}

extern "C" int pw_server_set_scoring(pw_server *S, int32_t scheme, const float *subst, int dim, int char_offset,
                                     float gap, float gap_ext, float lower_limit, float minimum_score) {
    if (!S) return PW_EINVAL;
    if (scheme < 0 || scheme >= SERVE_SCHEMES) return serve_fail(S, "scheme index must be 0..7", PW_EINVAL);
    for (const ServeLane &ln : S->lanes)
        if (ln.busy) return serve_fail(S, "pw_server_set_scoring: jobs are running", PW_EINVAL);
    for (const ServeJob *j : S->queue)
        if (j->scheme == scheme) return serve_fail(S, "pw_server_set_scoring: queued jobs use this scheme", PW_EINVAL);
    Scoring sc;
    int rc = make_scoring(S->ctx->err, subst, dim, char_offset, gap, gap_ext, lower_limit, minimum_score, sc);
    if (rc) return rc;
    S->schemes[scheme] = std::move(sc);
    S->scheme_set[scheme] = true;
    return PW_OK;
}

extern "C" int pw_server_submit(pw_server *S, int32_t scheme, const uint8_t *qres, const int64_t *qoff,
                                const int32_t *qlen, int32_t nq, const uint8_t *tres, const int64_t *toff,
                                const int32_t *tlen, int32_t nt, const int32_t *pq, const int32_t *pt,
                                int64_t n_pairs, const float *min_start_score, const uint8_t *same_id,
                                double minimum, int32_t sp_cap, int64_t *job_id) {
    if (!S) return PW_EINVAL;
    pw_ctx *ctx = S->ctx;
    if (!job_id || nq < 0 || nt < 0 || (nq > 0 && (!qres || !qoff || !qlen)) ||
        (nt > 0 && (!tres || !toff || !tlen)) || (n_pairs > 0 && (!pq || !pt || !min_start_score)))
        return serve_fail(S, "pw_server_submit: missing argument", PW_EINVAL);
    if (scheme < 0 || scheme >= SERVE_SCHEMES || !S->scheme_set[scheme])
        return serve_fail(S, "pw_server_submit: scheme not set", PW_EINVAL);
    const Scoring &sc = S->schemes[scheme];
    int rc = check_pairs(ctx->err, nq, nt, pq, pt, n_pairs);
    if (rc) return rc;
    rc = check_alphabet(ctx, sc, qres, qoff, qlen, nq);
    if (!rc) rc = check_alphabet(ctx, sc, tres, toff, tlen, nt);
    if (rc) return rc;
    ServeJob *j = new ServeJob();
    j->id = S->next_id++;
    j->scheme = scheme;
    j->minimum = minimum;
    j->sp_cap = sp_cap;
    // compact copies of the sequences
    j->qlen.assign(qlen, qlen + nq);
    j->tlen.assign(tlen, tlen + nt);
    for (int32_t i = 0; i < nq; ++i) {
        j->qoff.push_back((int64_t)j->qres.size());
        j->qres.insert(j->qres.end(), qres + qoff[i], qres + qoff[i] + qlen[i]);
    }
    for (int32_t i = 0; i < nt; ++i) {
        j->toff.push_back((int64_t)j->tres.size());
        j->tres.insert(j->tres.end(), tres + toff[i], tres + toff[i] + tlen[i]);
    }
    j->pq.assign(pq, pq + n_pairs);
    j->pt.assign(pt, pt + n_pairs);
    j->mss.assign(min_start_score, min_start_score + n_pairs);
    if (same_id) j->same.assign(same_id, same_id + n_pairs);
    else j->same.assign((size_t)n_pairs, 0);
    for (int64_t i = 0; i < n_pairs; ++i) {
        const int ql = qlen[pq[i]], tl = tlen[pt[i]];
        j->dir_need += dir_area(ql, tl, dir_residues_bound(ql));
        j->cells += (double)ql * tl;
    }
    S->jobs[j->id] = j;
    S->queue.push_back(j);
    *job_id = j->id;
    return PW_OK;
}

extern "C" int pw_server_poll(pw_server *S, int32_t block, int64_t *done_ids, int32_t max_ids, int32_t *n_done) {
    if (!S || !n_done || (max_ids > 0 && !done_ids)) return PW_EINVAL;
    pw_ctx *ctx = S->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    *n_done = 0;
    for (;;) {
        // finish batches whose device work has ended, then refill the lanes
        bool any_busy = false;
        for (ServeLane &ln : S->lanes) {
            if (!ln.busy) continue;
            hipError_t e = hipEventQuery(ln.done_ev);
            if (e == hipErrorNotReady) {
                any_busy = true;
                continue;
            }
            if (e != hipSuccess) return serve_fail(S, std::string("hipEventQuery: ") + hipGetErrorString(e), PW_EDEVICE);
            int rc = finish_batch(S, ln);
            if (rc) return rc;
        }
        // buffers outgrown during the batches are freed while nothing runs
        if (!any_busy && (!ctx->dead_dev.empty() || !ctx->dead_pinned.empty())) release_dead(ctx);
        int rc = admit(S);
        if (rc) return rc;
        if (!block || !S->completed.empty()) break;
        // nothing finished yet: wait for the oldest running batch
        ServeLane *oldest = nullptr;
        for (ServeLane &ln : S->lanes)
            if (ln.busy && (!oldest || ln.seq < oldest->seq)) oldest = &ln;
        if (!oldest) break;
        HIPCHK(ctx, hipEventSynchronize(oldest->done_ev));
    }
    const int32_t n = (int32_t)std::min<size_t>(S->completed.size(), (size_t)std::max(0, max_ids));
    for (int32_t i = 0; i < n; ++i) done_ids[i] = S->completed[i];
    S->completed.erase(S->completed.begin(), S->completed.begin() + n);
    *n_done = n;
    return PW_OK;
}

extern "C" int pw_server_take(pw_server *S, int64_t job_id, pw_result **out) {
    if (!S || !out) return PW_EINVAL;
    *out = nullptr;
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end() || !it->second->done)
        return serve_fail(S, "pw_server_take: job " + std::to_string(job_id) + " is not completed", PW_EINVAL);
    *out = &it->second->result->pub;
    delete it->second;
    S->jobs.erase(it);
    return PW_OK;
}

extern "C" int pw_server_stats(pw_server *S, int64_t *stats, int32_t n) {
    if (!S || (n > 0 && !stats)) return PW_EINVAL;
    int64_t running = 0;
    for (const ServeLane &ln : S->lanes) running += ln.busy ? (int64_t)ln.jobs.size() : 0;
    const int64_t v[11] = {S->n_batches, S->n_admitted, S->n_completed, S->peak_reserved,
                           (int64_t)S->queue.size(), running, S->arena_cap, (int64_t)S->us_launch,
                           (int64_t)S->us_finish, (int64_t)S->us_device, S->ctx->reallocs};
    for (int32_t i = 0; i < n && i < 11; ++i) stats[i] = v[i];
    return PW_OK;
}

extern "C" int pw_server_destroy(pw_server *S) {
    if (!S) return PW_EINVAL;
    pw_ctx *ctx = S->ctx;
    (void)hipSetDevice(ctx->device);
    for (ServeLane &ln : S->lanes) {
        destroy_lane(ln.L);
        if (ln.done_ev) (void)hipEventDestroy(ln.done_ev);
        if (ln.P) free_plan(ctx, ln.P);
        free_pinned(ln.h_max);
    }
    for (auto &kv : S->jobs) {
        if (kv.second->result) pw_free_result(ctx, &kv.second->result->pub);
        delete kv.second;
    }
    // queued jobs are in the map too
    free_buf(ctx, S->arena);
    ctx->defer_free = false;
    release_dead(ctx);
    delete S;
    return PW_OK;
}
