/*
 * ORACLE — test infrastructure only. Nothing in the product links, loads or
 * calls this file; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py do, as the checker.
 *
 * CPU restatement of pyPaSWAS's OpenCL-CPU Smith-Waterman path:
 *   - pyPaSWAS/Core/ocl/default_main_cpu.cl  (calculateScore :108-222,
 *     calculateScoreAffineGap :224-387, traceback :389-496,
 *     tracebackAffineGap :498-604)
 *   - the host launch loops SmithWaterman._calculate_score / _traceback_host
 *     (pyPaSWAS/Core/SmithWaterman.py:592-653) and the launch geometry of
 *     SmithWatermanCPU (pyPaSWAS/Core/SmithWatermanOcl.py:457-518)
 *   - the host alignment walk SmithWaterman._print_alignments
 *     (pyPaSWAS/Core/SmithWaterman.py:665-841)
 *
 * Two independent restatements live here:
 *   oracle_sched_batch  keeps the reference's data layout and schedule:
 *                       (X+1)x(Y+1) FP32 matrices per pair, 8x8 blocks, 2x2
 *                       work-items of 4x4 cells, one sweep per block
 *                       anti-diagonal, the 2-D prefix max into globalMaxima,
 *                       and the reverse sweep for the traceback.  OpenMP runs
 *                       the independent work-groups of one sweep in parallel,
 *                       like the OpenCL CPU runtime does.
 *   oracle_serial_pair  one pair at its real lengths, row-major cell loop and
 *                       a reverse row-major traceback sweep; no blocks, no
 *                       padding.
 * Per-pair results do not depend on the schedule (every schedule processes a
 * cell before all componentwise-smaller cells), so the two must agree on
 * maxima, post-traceback directions and the starting-point set.  Starting
 * points are reported in one canonical order (see sp_order_key).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* direction codes, pyPaSWAS/Core/__init__.py:7-13 */
#define DIR_UL 1
#define DIR_UP 2
#define DIR_LEFT 3
#define DIR_NONE 32
#define DIR_STOP 64
#define DIR_IN_ALIGNMENT 96
/* affine source bits, default_direction.cl:34-43 (PaSWAS.py:46-53 fills them) */
#define MAIN_MATRIX (DIR_UL << 2)
#define I_MATRIX (DIR_UP << 2)
#define J_MATRIX (DIR_LEFT << 2)
#define DIRECTION_MASK 3
#define MATRIX_MASK 12

#define SHARED_X 8
#define SHARED_Y 8
#define WORKLOAD_X 4
#define WORKLOAD_Y 4
#define WORKGROUP_X 2
#define WORKGROUP_Y 2

/* default_main_cpu.cl:21 — a double literal assigned to float; exact in f32 */
#define FILL_SCORE (-1E10f)
/* default_main_cpu.cl:24 */
#define AFFINE_GAP_INIT (-1E10f)

typedef struct {
    uint32_t sequence, target, blockX, blockY, valueX, valueY;
    float score, maxScore, posScore;
} sp_record; /* default_main_cpu.cl:66-76, 36 bytes */

typedef struct {
    const float *subst; /* dim x dim, row = query residue */
    int dim;
    int char_offset;
    float gap;       /* gapScore, default_score.cl:19 */
    float gap_ext;   /* gapExtension, default_score.cl:20 */
    int affine;      /* SmithWaterman.py:149: gap_extension != None */
    float lower_limit;   /* LOWER_LIMIT_SCORE */
    float minimum;       /* MINIMUM_SCORE */
} sw_params;

static inline float as_neg(float v) {
    /* as_float(SIGN_BIT_MASK | as_int(v)), default_main_cpu.cl:446 */
    uint32_t u;
    memcpy(&u, &v, 4);
    u |= 0x80000000u;
    memcpy(&v, &u, 4);
    return v;
}

static inline float subst_score(const sw_params *p, char a, char b) {
    /* default_main_cpu.cl:161 */
    if (a == 0 || b == 0) return FILL_SCORE;
    return p->subst[(a - p->char_offset) * p->dim + (b - p->char_offset)];
}

/* Canonical emission order of starting points within one pair: the order of
 * a serial reading of the reverse sweep — block anti-diagonal (bx+by)
 * descending (SmithWaterman.py:626-653), blockx descending within it (groups
 * are launched with blockx = idx - g), work-item anti-diagonal (tx+ty)
 * descending (default_main_cpu.cl:414), tx descending among concurrent
 * work-items, then the j and k loops descending (:416-420).  Larger key =
 * emitted earlier. */
static inline uint64_t sp_order_key(uint32_t x, uint32_t y) {
    uint64_t bx = x >> 3, by = y >> 3, tx = (x >> 2) & 1, ty = (y >> 2) & 1;
    uint64_t j = x & 3, k = y & 3;
    return ((bx + by) << 45) | (bx << 26) | ((tx + ty) << 5) | (tx << 4) | (j << 2) | k;
}

typedef struct {
    uint64_t pair;  /* q * T + t */
    uint64_t key;
    sp_record rec;
} sp_tmp;

static int cmp_sp_tmp(const void *a, const void *b) {
    const sp_tmp *x = (const sp_tmp *)a, *y = (const sp_tmp *)b;
    if (x->pair != y->pair) return x->pair < y->pair ? -1 : 1;
    if (x->key != y->key) return x->key > y->key ? -1 : 1; /* descending */
    return 0;
}

/* ------------------------------------------------------------------ */
/* Schedule-faithful restatement                                       */
/* ------------------------------------------------------------------ */

typedef struct {
    int N, T, X, Y, XB, YB;
    size_t mstride; /* (X+1)*(Y+1) */
    float *H, *I, *J;
    float *gmax; /* [N][T][XB][YB][2][2] */
    uint8_t *dir; /* [N][T][X][Y] */
    const char *seqs, *tgts;
    const float *minsc; /* [T][N] */
    const sw_params *p;
} sched_state;

#define MAT(S, M, q, t, ax, ay) ((M)[((size_t)(q) * (S)->T + (t)) * (S)->mstride + (size_t)(ax) * ((S)->Y + 1) + (ay)])
#define DIRC(S, q, t, x, y) ((S)->dir[(((size_t)(q) * (S)->T + (t)) * (S)->X + (x)) * (size_t)(S)->Y + (y)])
#define GMAX(S, q, t, bx, by, wx, wy) ((S)->gmax[((((size_t)(q) * (S)->T + (t)) * (S)->XB + (bx)) * (S)->YB + (by)) * 4 + (wx) * 2 + (wy)])

/* one cell of calculateScore, default_main_cpu.cl:154-181 */
static inline float cell_linear(sched_state *S, int q, int t, unsigned aXM1, unsigned aYM1) {
    const sw_params *p = S->p;
    unsigned aIDx = aXM1 + 1, aIDy = aYM1 + 1;
    char c1 = S->seqs[(size_t)q * S->X + aXM1];
    char c2 = S->tgts[(size_t)t * S->Y + aYM1];
    float inner = subst_score(p, c1, c2);
    float ulS = MAT(S, S->H, q, t, aXM1, aYM1) + inner;
    float lS = MAT(S, S->H, q, t, aXM1, aIDy) + p->gap;
    float uS = MAT(S, S->H, q, t, aIDx, aYM1) + p->gap;
    float cur = 0.0f;
    uint8_t d = DIR_NONE;
    if (cur < lS) { cur = lS; d = DIR_LEFT; }
    if (cur < uS) { cur = uS; d = DIR_UP; }
    if (cur < ulS) { cur = ulS; d = DIR_UL; }
    MAT(S, S->H, q, t, aIDx, aIDy) = cur;
    DIRC(S, q, t, aXM1, aYM1) = d;
    return cur;
}

/* one cell of calculateScoreAffineGap, default_main_cpu.cl:275-346 */
static inline float cell_affine(sched_state *S, int q, int t, unsigned aXM1, unsigned aYM1) {
    const sw_params *p = S->p;
    unsigned aIDx = aXM1 + 1, aIDy = aYM1 + 1;
    char c1 = S->seqs[(size_t)q * S->X + aXM1];
    char c2 = S->tgts[(size_t)t * S->Y + aYM1];
    float inner = subst_score(p, c1, c2);
    uint8_t d = DIR_NONE, di = DIR_NONE, dj = DIR_NONE;
    float cur = 0.0f, ci, cj, mM, mI, mJ;
    mM = MAT(S, S->H, q, t, aXM1, aYM1) + inner;
    mI = MAT(S, S->I, q, t, aXM1, aYM1) + inner;
    mJ = MAT(S, S->J, q, t, aXM1, aYM1) + inner;
    if (cur < mI) { cur = mI; d = DIR_UL | MAIN_MATRIX; }
    if (cur < mJ) { cur = mJ; d = DIR_UL | MAIN_MATRIX; }
    if (cur < mM) { cur = mM; d = DIR_UL | MAIN_MATRIX; }
    MAT(S, S->H, q, t, aIDx, aIDy) = inner == FILL_SCORE ? 0.0f : cur;

    ci = AFFINE_GAP_INIT;
    mM = p->gap + MAT(S, S->H, q, t, aIDx, aYM1);
    mI = p->gap_ext + MAT(S, S->I, q, t, aIDx, aYM1);
    if (ci < mI) { ci = mI; di = DIR_UP | I_MATRIX; }
    if (ci < mM) { ci = mM; di = DIR_UP | I_MATRIX; }
    MAT(S, S->I, q, t, aIDx, aIDy) = ci < 0 ? AFFINE_GAP_INIT : ci;

    cj = AFFINE_GAP_INIT;
    mM = p->gap + MAT(S, S->H, q, t, aXM1, aIDy);
    mJ = p->gap_ext + MAT(S, S->J, q, t, aXM1, aIDy);
    if (cj < mJ) { cj = mJ; dj = DIR_LEFT | J_MATRIX; }
    if (cj < mM) { cj = mM; dj = DIR_LEFT | J_MATRIX; }
    MAT(S, S->J, q, t, aIDx, aIDy) = cj < 0 ? AFFINE_GAP_INIT : cj;

    cur = fmaxf(cur, fmaxf(ci, cj));
    if (cur > 0) {
        if (cur == MAT(S, S->H, q, t, aIDx, aIDy)) d = d;
        else if (cur == MAT(S, S->I, q, t, aIDx, aIDy)) d = di;
        else if (cur == MAT(S, S->J, q, t, aIDx, aIDy)) d = dj;
    }
    DIRC(S, q, t, aXM1, aYM1) = d;
    return cur;
}

/* one work-group of calculateScore(AffineGap): default_main_cpu.cl:119-221 */
static void score_group(sched_state *S, int q, int t, unsigned blockx, unsigned blocky) {
    float thread_max[WORKGROUP_X][WORKGROUP_Y] = {{0.0f}};
    for (int i = 0; i < WORKGROUP_X + WORKGROUP_Y; ++i) {
        for (int tx = 0; tx < WORKGROUP_X; ++tx)
            for (int ty = 0; ty < WORKGROUP_Y; ++ty) {
                if (i != tx + ty) continue;
                for (int j = 0; j < WORKLOAD_X; j++) {
                    unsigned aXM1 = tx * WORKLOAD_X + j + blockx * SHARED_X;
                    for (int k = 0; k < WORKLOAD_Y; k++) {
                        unsigned aYM1 = ty * WORKLOAD_Y + k + blocky * SHARED_Y;
                        float cur = S->p->affine ? cell_affine(S, q, t, aXM1, aYM1)
                                                 : cell_linear(S, q, t, aXM1, aYM1);
                        thread_max[tx][ty] = fmaxf(cur, thread_max[tx][ty]);
                    }
                }
            }
        /* the 2-D prefix max of thread maxima, default_main_cpu.cl:188-210 */
        for (int tx = 0; tx < WORKGROUP_X; ++tx)
            for (int ty = 0; ty < WORKGROUP_Y; ++ty) {
                if (i - 1 != tx + ty) continue;
                if (i == 1) {
                    float maxPrev = 0.0f;
                    if (blockx && blocky)
                        maxPrev = fmaxf(fmaxf(GMAX(S, q, t, blockx - 1, blocky - 1, 1, 1),
                                              GMAX(S, q, t, blockx - 1, blocky, 1, 1)),
                                        GMAX(S, q, t, blockx, blocky - 1, 1, 1));
                    else if (blockx)
                        maxPrev = GMAX(S, q, t, blockx - 1, blocky, 1, 1);
                    else if (blocky)
                        maxPrev = GMAX(S, q, t, blockx, blocky - 1, 1, 1);
                    GMAX(S, q, t, blockx, blocky, 0, 0) = fmaxf(maxPrev, thread_max[tx][ty]);
                } else if (!tx && ty) {
                    GMAX(S, q, t, blockx, blocky, 0, ty) =
                        fmaxf(GMAX(S, q, t, blockx, blocky, 0, ty - 1), thread_max[tx][ty]);
                } else if (tx && !ty) {
                    GMAX(S, q, t, blockx, blocky, tx, 0) =
                        fmaxf(GMAX(S, q, t, blockx, blocky, tx - 1, 0), thread_max[tx][ty]);
                } else if (tx && ty) {
                    GMAX(S, q, t, blockx, blocky, tx, ty) =
                        fmaxf(GMAX(S, q, t, blockx, blocky, tx - 1, ty),
                              fmaxf(thread_max[tx][ty], GMAX(S, q, t, blockx, blocky, tx, ty - 1)));
                }
            }
    }
    GMAX(S, q, t, blockx, blocky, 1, 1) =
        fmaxf(thread_max[1][1], fmaxf(GMAX(S, q, t, blockx, blocky, 0, 1), GMAX(S, q, t, blockx, blocky, 1, 0)));
}

/* one cell of traceback, default_main_cpu.cl:424-485 */
static inline int tb_cell_linear(sched_state *S, int q, int t, unsigned blockx, unsigned blocky,
                                 unsigned vx, unsigned vy, float maximum, float mp, sp_record *out) {
    const sw_params *p = S->p;
    unsigned aXM1 = blockx * SHARED_X + vx, aYM1 = blocky * SHARED_Y + vy;
    unsigned aIDx = aXM1 + 1, aIDy = aYM1 + 1;
    int emitted = 0;
    float score = MAT(S, S->H, q, t, aIDx, aIDy);
    uint8_t direction = DIRC(S, q, t, aXM1, aYM1);
    float ll_max = p->lower_limit * maximum;
    if (direction == DIR_UL && score >= ll_max && score >= mp) {
        out->sequence = q; out->target = t; out->blockX = blockx; out->blockY = blocky;
        out->valueX = vx; out->valueY = vy;
        out->score = MAT(S, S->H, q, t, aIDx, aIDy);
        out->maxScore = maximum; out->posScore = mp;
        emitted = 1;
        MAT(S, S->H, q, t, aIDx, aIDy) = as_neg(score);
    }
    score = MAT(S, S->H, q, t, aIDx, aIDy);
    if (score < 0 && direction == DIR_UL) {
        score = MAT(S, S->H, q, t, aIDx - 1, aIDy - 1);
        if (score == 0.0f) { direction = DIR_STOP; DIRC(S, q, t, aXM1, aYM1) = direction; }
        else MAT(S, S->H, q, t, aIDx - 1, aIDy - 1) = as_neg(score);
    }
    if (score < 0 && direction == DIR_LEFT) {
        score = MAT(S, S->H, q, t, aIDx - 1, aIDy);
        if (score == 0.0f) { direction = DIR_STOP; DIRC(S, q, t, aXM1, aYM1) = direction; }
        else MAT(S, S->H, q, t, aIDx - 1, aIDy) = as_neg(score);
    }
    if (score < 0 && direction == DIR_UP) {
        score = MAT(S, S->H, q, t, aIDx, aIDy - 1);
        if (score == 0.0f) { direction = DIR_STOP; DIRC(S, q, t, aXM1, aYM1) = direction; }
        else MAT(S, S->H, q, t, aIDx, aIDy - 1) = as_neg(score);
    }
    return emitted;
}

/* one cell of tracebackAffineGap, default_main_cpu.cl:535-596 */
static inline int tb_cell_affine(sched_state *S, int q, int t, unsigned blockx, unsigned blocky,
                                 unsigned vx, unsigned vy, float maximum, float mp, sp_record *out) {
    const sw_params *p = S->p;
    unsigned aXM1 = blockx * SHARED_X + vx, aYM1 = blocky * SHARED_Y + vy;
    unsigned aIDx = aXM1 + 1, aIDy = aYM1 + 1;
    int emitted = 0;
    float score = MAT(S, S->H, q, t, aIDx, aIDy);
    uint8_t raw = DIRC(S, q, t, aXM1, aYM1);
    uint8_t direction = DIRECTION_MASK & raw;
    uint8_t source = MATRIX_MASK & raw;
    float ll_max = p->lower_limit * maximum;
    if (source == MAIN_MATRIX && score >= ll_max && score >= mp) {
        out->sequence = q; out->target = t; out->blockX = blockx; out->blockY = blocky;
        out->valueX = vx; out->valueY = vy;
        out->score = MAT(S, S->H, q, t, aIDx, aIDy);
        out->maxScore = maximum; out->posScore = mp;
        emitted = 1;
        MAT(S, S->H, q, t, aIDx, aIDy) = as_neg(score);
    }
    float iv = MAT(S, S->I, q, t, aIDx, aIDy), jv = MAT(S, S->J, q, t, aIDx, aIDy);
    if ((MAT(S, S->H, q, t, aIDx, aIDy) < 0 && source == MAIN_MATRIX) ||
        (iv < 0 && iv > AFFINE_GAP_INIT && source == I_MATRIX) ||
        (jv < 0 && jv > AFFINE_GAP_INIT && source == J_MATRIX)) {
        switch (source) {
        case MAIN_MATRIX:
            score = MAT(S, S->H, q, t, aIDx - 1, aIDy - 1);
            if (score == 0.0f) direction = DIR_STOP;
            else { MAT(S, S->H, q, t, aIDx - 1, aIDy - 1) = as_neg(score); direction = DIR_UL; }
            break;
        case I_MATRIX:
            score = MAT(S, S->I, q, t, aIDx, aIDy - 1);
            MAT(S, S->I, q, t, aIDx, aIDy - 1) = as_neg(score);
            direction = DIR_UP;
            break;
        case J_MATRIX:
            score = MAT(S, S->J, q, t, aIDx - 1, aIDy);
            MAT(S, S->J, q, t, aIDx - 1, aIDy) = as_neg(score);
            direction = DIR_LEFT;
            break;
        }
    }
    DIRC(S, q, t, aXM1, aYM1) = direction;
    return emitted;
}

/* one work-group of traceback(AffineGap): default_main_cpu.cl:401-494 */
static int tb_group(sched_state *S, int q, int t, unsigned blockx, unsigned blocky,
                    sp_tmp *buf) {
    float maximum = GMAX(S, q, t, S->XB - 1, S->YB - 1, 1, 1);
    int n = 0;
    if (!(maximum >= S->p->minimum)) return 0;
    float mp = S->minsc[(size_t)t * S->N + q];
    for (int i = WORKGROUP_X + WORKGROUP_Y - 1; i >= 0; --i)
        for (int tx = WORKGROUP_X - 1; tx >= 0; --tx)
            for (int ty = WORKGROUP_Y - 1; ty >= 0; --ty) {
                if (i != tx + ty) continue;
                for (int j = WORKLOAD_X - 1; j >= 0; j--)
                    for (int k = WORKLOAD_Y - 1; k >= 0; k--) {
                        unsigned vx = tx * WORKLOAD_X + j, vy = ty * WORKLOAD_Y + k;
                        sp_record rec;
                        int e = S->p->affine ? tb_cell_affine(S, q, t, blockx, blocky, vx, vy, maximum, mp, &rec)
                                             : tb_cell_linear(S, q, t, blockx, blocky, vx, vy, maximum, mp, &rec);
                        if (e) {
                            buf[n].pair = (uint64_t)q * S->T + t;
                            buf[n].key = sp_order_key(blockx * SHARED_X + vx, blocky * SHARED_Y + vy);
                            buf[n].rec = rec;
                            n++;
                        }
                    }
            }
    return n;
}

/*
 * Score + traceback of one batch exactly as SmithWatermanCPU runs it.
 *   seqs  N*X chars ('\0'-padded, SWSeq.extentToFillGPU), tgts T*Y chars
 *   minsc per-pair minimum start score, index t*N + q (default_main_cpu.cl:412)
 * Outputs: pair maxima [N][T], post-traceback directions [N][T][X][Y] and the
 * starting points ordered by pair then canonical emission order.  Returns the
 * number of starting points found (which may exceed sp_cap; only sp_cap are
 * written), or -1 on allocation failure.
 */
long oracle_sched_batch(const char *seqs, int N, int X, const char *tgts, int T, int Y,
                        const float *minsc, const sw_params *p, int nthreads,
                        float *max_out, uint8_t *dir_out, sp_record *sp_out, long sp_cap) {
    sched_state S;
    memset(&S, 0, sizeof S);
    S.N = N; S.T = T; S.X = X; S.Y = Y; S.XB = X / SHARED_X; S.YB = Y / SHARED_Y;
    S.mstride = (size_t)(X + 1) * (Y + 1);
    S.seqs = seqs; S.tgts = tgts; S.minsc = minsc; S.p = p;
    S.dir = dir_out;
    size_t npair = (size_t)N * T;
    S.H = (float *)malloc(npair * S.mstride * sizeof(float));
    S.gmax = (float *)calloc(npair * S.XB * S.YB * 4, sizeof(float));
    if (p->affine) {
        S.I = (float *)malloc(npair * S.mstride * sizeof(float));
        S.J = (float *)malloc(npair * S.mstride * sizeof(float));
    }
    if (!S.H || !S.gmax || (p->affine && (!S.I || !S.J))) {
        free(S.H); free(S.gmax); free(S.I); free(S.J);
        return -1;
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    /* SmithWatermanOcl.py:371-383: matrix filled with 0, I/J with -1E10 (in parallel,
     * as the runtime's buffer fill is) */
    const long total_cells = (long)(npair * S.mstride);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < total_cells; i++) {
        S.H[i] = 0.0f;
        if (p->affine) { S.I[i] = -1E10f; S.J[i] = -1E10f; }
    }

    /* SmithWaterman._calculate_score, SmithWaterman.py:592-618 */
    int xb = S.XB, yb = S.YB;
    int max_blocks = xb < yb ? xb : yb;
    int start_decrease_at = xb + yb - max_blocks;
    int idx = 0, idy = 0;
    for (int i = 1; i < xb + yb; i++) {
        int nb = (i <= max_blocks) ? i : (i >= start_decrease_at ? xb + yb - i : max_blocks);
        long groups = (long)N * T * nb;
#pragma omp parallel for schedule(dynamic, 16)
        for (long gid = 0; gid < groups; gid++) {
            int q = (int)(gid / ((long)T * nb));
            long rem = gid % ((long)T * nb);
            int g = (int)(rem / T), t = (int)(rem % T);
            score_group(&S, q, t, (unsigned)(idx - g), (unsigned)(idy + g));
        }
        if (idx == xb - 1) idy += 1;
        if (idx < xb - 1) idx += 1;
    }
    for (size_t pi = 0; pi < npair; pi++) {
        int q = (int)(pi / T), t = (int)(pi % T);
        max_out[pi] = GMAX(&S, q, t, xb - 1, yb - 1, 1, 1);
    }

    /* SmithWaterman._traceback_host, SmithWaterman.py:626-653 */
    long total = 0, cap_tmp = 1024;
    sp_tmp *all = (sp_tmp *)malloc(cap_tmp * sizeof(sp_tmp));
    idx = xb - 1; idy = yb - 1;
    for (int i = 1; i < xb + yb; i++) {
        int nb = (i <= max_blocks) ? i : (i >= start_decrease_at ? xb + yb - i : max_blocks);
        long groups = (long)N * T * nb;
#pragma omp parallel
        {
            sp_tmp local[64];
            long lcap = 0, ln = 0;
            sp_tmp *lbuf = NULL;
#pragma omp for schedule(dynamic, 16)
            for (long gid = 0; gid < groups; gid++) {
                int q = (int)(gid / ((long)T * nb));
                long rem = gid % ((long)T * nb);
                int g = (int)(rem / T), t = (int)(rem % T);
                int n = tb_group(&S, q, t, (unsigned)(idx - g), (unsigned)(idy + g), local);
                if (n) {
                    if (ln + n > lcap) {
                        lcap = (ln + n) * 2 + 64;
                        lbuf = (sp_tmp *)realloc(lbuf, lcap * sizeof(sp_tmp));
                    }
                    memcpy(lbuf + ln, local, n * sizeof(sp_tmp));
                    ln += n;
                }
            }
            if (ln) {
#pragma omp critical
                {
                    if (total + ln > cap_tmp) {
                        cap_tmp = (total + ln) * 2;
                        all = (sp_tmp *)realloc(all, cap_tmp * sizeof(sp_tmp));
                    }
                    memcpy(all + total, lbuf, ln * sizeof(sp_tmp));
                    total += ln;
                }
            }
            free(lbuf);
        }
        if (idy == 0) idx -= 1;
        if (idy > 0) idy -= 1;
    }
    qsort(all, total, sizeof(sp_tmp), cmp_sp_tmp);
    for (long i = 0; i < total && i < sp_cap; i++) sp_out[i] = all[i].rec;
    free(all);
    free(S.H); free(S.gmax); free(S.I); free(S.J);
    return total;
}

/* ------------------------------------------------------------------ */
/* Independent serial restatement (real lengths, row-major)            */
/* ------------------------------------------------------------------ */

/*
 * One pair, no padding.  The recurrences are the same cell formulas
 * (default_main_cpu.cl:154-181 and :275-346); the traceback applies the
 * per-cell rule of default_main_cpu.cl:424-485 / :535-596 in reverse
 * row-major order, which is another valid reverse topological order.
 * dir_out is m*n (row = query position).  Starting points are returned in
 * canonical order (sp_order_key) with blockX/valueX etc. filled.
 */
long oracle_serial_pair(const char *q, int m, const char *t, int n, float mp,
                        const sw_params *p, float *max_out, uint8_t *dir_out,
                        sp_record *sp_out, long sp_cap) {
    size_t W = (size_t)n + 1;
    float *H = (float *)malloc((size_t)(m + 1) * W * sizeof(float));
    float *I = NULL, *J = NULL;
    if (!H) return -1;
    for (size_t i = 0; i < (size_t)(m + 1) * W; i++) H[i] = 0.0f;
    if (p->affine) {
        I = (float *)malloc((size_t)(m + 1) * W * sizeof(float));
        J = (float *)malloc((size_t)(m + 1) * W * sizeof(float));
        if (!I || !J) { free(H); free(I); free(J); return -1; }
        for (size_t i = 0; i < (size_t)(m + 1) * W; i++) { I[i] = -1E10f; J[i] = -1E10f; }
    }
    float best = 0.0f;
    for (int x = 1; x <= m; x++)
        for (int y = 1; y <= n; y++) {
            float s = p->subst[(q[x - 1] - p->char_offset) * p->dim + (t[y - 1] - p->char_offset)];
            uint8_t d = DIR_NONE;
            float cur = 0.0f;
            if (!p->affine) {
                float ul = H[(x - 1) * W + y - 1] + s, l = H[(x - 1) * W + y] + p->gap,
                      u = H[x * W + y - 1] + p->gap;
                if (cur < l) { cur = l; d = DIR_LEFT; }
                if (cur < u) { cur = u; d = DIR_UP; }
                if (cur < ul) { cur = ul; d = DIR_UL; }
                H[x * W + y] = cur;
            } else {
                float a = I[(x - 1) * W + y - 1] + s, b = J[(x - 1) * W + y - 1] + s,
                      c = H[(x - 1) * W + y - 1] + s;
                if (cur < a) { cur = a; d = DIR_UL | MAIN_MATRIX; }
                if (cur < b) { cur = b; d = DIR_UL | MAIN_MATRIX; }
                if (cur < c) { cur = c; d = DIR_UL | MAIN_MATRIX; }
                H[x * W + y] = cur;
                uint8_t di = DIR_NONE, dj = DIR_NONE;
                float ci = -1E10f, cj = -1E10f;
                float e1 = p->gap_ext + I[x * W + y - 1], e2 = p->gap + H[x * W + y - 1];
                if (ci < e1) { ci = e1; di = DIR_UP | I_MATRIX; }
                if (ci < e2) { ci = e2; di = DIR_UP | I_MATRIX; }
                I[x * W + y] = ci < 0 ? -1E10f : ci;
                float f1 = p->gap_ext + J[(x - 1) * W + y], f2 = p->gap + H[(x - 1) * W + y];
                if (cj < f1) { cj = f1; dj = DIR_LEFT | J_MATRIX; }
                if (cj < f2) { cj = f2; dj = DIR_LEFT | J_MATRIX; }
                J[x * W + y] = cj < 0 ? -1E10f : cj;
                cur = fmaxf(cur, fmaxf(ci, cj));
                if (cur > 0) {
                    if (cur == H[x * W + y]) d = d;
                    else if (cur == I[x * W + y]) d = di;
                    else if (cur == J[x * W + y]) d = dj;
                }
            }
            dir_out[(size_t)(x - 1) * n + (y - 1)] = d;
            best = fmaxf(best, cur);
        }
    *max_out = best;
    long total = 0, cap = 64;
    sp_tmp *buf = (sp_tmp *)malloc(cap * sizeof(sp_tmp));
    if (best >= p->minimum) {
        float ll_max = p->lower_limit * best;
        for (int x = m; x >= 1; x--)
            for (int y = n; y >= 1; y--) {
                uint8_t *dp = &dir_out[(size_t)(x - 1) * n + (y - 1)];
                float score = H[x * W + y];
                if (!p->affine) {
                    uint8_t direction = *dp;
                    if (direction == DIR_UL && score >= ll_max && score >= mp) {
                        if (total == cap) { cap *= 2; buf = (sp_tmp *)realloc(buf, cap * sizeof(sp_tmp)); }
                        sp_tmp *o = &buf[total++];
                        o->pair = 0; o->key = sp_order_key(x - 1, y - 1);
                        o->rec.sequence = 0; o->rec.target = 0;
                        o->rec.blockX = (x - 1) >> 3; o->rec.blockY = (y - 1) >> 3;
                        o->rec.valueX = (x - 1) & 7; o->rec.valueY = (y - 1) & 7;
                        o->rec.score = score; o->rec.maxScore = best; o->rec.posScore = mp;
                        H[x * W + y] = as_neg(score);
                    }
                    score = H[x * W + y];
                    if (score < 0 && (direction == DIR_UL || direction == DIR_LEFT || direction == DIR_UP)) {
                        int px = x - (direction != DIR_UP), py = y - (direction != DIR_LEFT);
                        float pv = H[px * W + py];
                        if (pv == 0.0f) *dp = DIR_STOP;
                        else H[px * W + py] = as_neg(pv);
                    }
                } else {
                    uint8_t direction = *dp & DIRECTION_MASK, source = *dp & MATRIX_MASK;
                    if (source == MAIN_MATRIX && score >= ll_max && score >= mp) {
                        if (total == cap) { cap *= 2; buf = (sp_tmp *)realloc(buf, cap * sizeof(sp_tmp)); }
                        sp_tmp *o = &buf[total++];
                        o->pair = 0; o->key = sp_order_key(x - 1, y - 1);
                        o->rec.sequence = 0; o->rec.target = 0;
                        o->rec.blockX = (x - 1) >> 3; o->rec.blockY = (y - 1) >> 3;
                        o->rec.valueX = (x - 1) & 7; o->rec.valueY = (y - 1) & 7;
                        o->rec.score = score; o->rec.maxScore = best; o->rec.posScore = mp;
                        H[x * W + y] = as_neg(score);
                    }
                    float iv = I[x * W + y], jv = J[x * W + y];
                    if (H[x * W + y] < 0 && source == MAIN_MATRIX) {
                        float pv = H[(x - 1) * W + y - 1];
                        if (pv == 0.0f) direction = DIR_STOP;
                        else { H[(x - 1) * W + y - 1] = as_neg(pv); direction = DIR_UL; }
                    } else if (iv < 0 && iv > -1E10f && source == I_MATRIX) {
                        I[x * W + y - 1] = as_neg(I[x * W + y - 1]);
                        direction = DIR_UP;
                    } else if (jv < 0 && jv > -1E10f && source == J_MATRIX) {
                        J[(x - 1) * W + y] = as_neg(J[(x - 1) * W + y]);
                        direction = DIR_LEFT;
                    }
                    *dp = direction;
                }
            }
    }
    qsort(buf, total, sizeof(sp_tmp), cmp_sp_tmp);
    for (long i = 0; i < total && i < sp_cap; i++) sp_out[i] = buf[i].rec;
    free(buf);
    free(H); free(I); free(J);
    return total;
}

/* ------------------------------------------------------------------ */
/* Host alignment walk                                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t shown;       /* 1: a Hit is built from this starting point */
    int32_t s_start, s_end, t_start, t_end;
    int32_t matches, mismatches;
    int32_t length;      /* characters in the alignment strings */
    int64_t ops_offset;  /* into the ops buffer: 'D' pair, 'L' query+gap, 'U' gap+target */
} walk_result;

static inline char ascii_lower(char c) { return (c >= 'A' && c <= 'Z') ? (char)(c + 32) : c; }

static int cmp_sp_score(const void *a, const void *b) {
    const sp_tmp *x = (const sp_tmp *)a, *y = (const sp_tmp *)b;
    if (x->rec.score != y->rec.score) return x->rec.score < y->rec.score ? -1 : 1;
    return x->key < y->key ? -1 : (x->key > y->key ? 1 : 0); /* key = input position: stable */
}

/*
 * SmithWaterman._print_alignments, SmithWaterman.py:689-841, on the direction
 * array dir [npairs][X][Y] (modified in place with IN_ALIGNMENT marks).
 * sps are sorted ascending by score with a stable sort (:699); order[] receives
 * the index of each processed starting point in that order.  same_id[pair]
 * is 1 when the query and target ids are equal (:752-755).  qs/ts hold the
 * real (uppercase) sequences of each starting point's pair, qlen/tlen their
 * lengths.  Characters compare case-insensitively (:771).
 * Returns the total number of op characters written (or needed, if above ops_cap).
 */
long oracle_walk(uint8_t *dir, int X, int Y, const sp_record *sps, long nsp,
                 const int32_t *pair_of_sp, const uint8_t *same_id, double minimum,
                 const char *const *qs, const char *const *ts,
                 walk_result *res, int64_t *order, char *ops, long ops_cap) {
    sp_tmp *tmp = (sp_tmp *)malloc((nsp ? nsp : 1) * sizeof(sp_tmp));
    for (long i = 0; i < nsp; i++) { tmp[i].pair = (uint64_t)pair_of_sp[i]; tmp[i].key = (uint64_t)i; tmp[i].rec = sps[i]; }
    qsort(tmp, nsp, sizeof(sp_tmp), cmp_sp_score);
    long nops = 0;
    for (long s = 0; s < nsp; s++) {
        long idx = (long)tmp[s].key;
        const sp_record *sp = &tmp[s].rec;
        int pr = (int)tmp[s].pair;
        uint8_t *D = dir + (size_t)pr * X * Y;
        order[s] = idx;
        walk_result *r = &res[s];
        int bx = (int)sp->blockX, by = (int)sp->blockY, vx = (int)sp->valueX, vy = (int)sp->valueY;
        int s_end = bx * 8 + vx, t_end = by * 8 + vy;
        int s_start = s_end + 1, t_start = t_end + 1;
        int matches = 0, mismatches = 0, len = 0;
        r->ops_offset = nops;
        int direction = D[(size_t)(bx * 8 + vx) * Y + by * 8 + vy];
        int show = 1;
        if (same_id[pr]) { direction = DIR_STOP; show = 0; }
        if ((double)sp->score < minimum) show = 0;
        const char *qq = qs[idx], *tt = ts[idx];
        while (show && bx >= 0 && by >= 0 && vx >= 0 && vy >= 0 && direction != DIR_STOP && direction != DIR_NONE) {
            size_t c = (size_t)(bx * 8 + vx) * Y + by * 8 + vy;
            direction = D[c];
            D[c] = DIR_IN_ALIGNMENT;
            if (direction == DIR_IN_ALIGNMENT) {
                show = 0;
            } else if (direction == DIR_UL || direction == DIR_STOP) {
                char a = qq[bx * 8 + vx], b = tt[by * 8 + vy];
                if (ascii_lower(a) == ascii_lower(b)) matches++; else mismatches++;
                if (nops < ops_cap) ops[nops] = 'D';
                nops++; len++;
                s_start--; t_start--;
                if (direction == DIR_UL) {
                    if (vx == 0) { bx--; vx = 7; } else vx--;
                    if (vy == 0) { by--; vy = 7; } else vy--;
                }
            } else if (direction == DIR_LEFT) {
                if (nops < ops_cap) ops[nops] = 'L';
                nops++; len++;
                s_start--;
                if (vx == 0) { bx--; vx = 7; } else vx--;
            } else if (direction == DIR_UP) {
                if (nops < ops_cap) ops[nops] = 'U';
                nops++; len++;
                t_start--;
                if (vy == 0) { by--; vy = 7; } else vy--;
            } else {
                bx = -1;
            }
        }
        r->shown = show;
        r->s_start = s_start; r->s_end = s_end; r->t_start = t_start; r->t_end = t_end;
        r->matches = matches; r->mismatches = mismatches; r->length = len;
        if (!show) nops = r->ops_offset; /* no Hit: drop its characters */
    }
    free(tmp);
    return nops;
}
