// Device-side data structures shared by the kernels and the host code of
// the gfx950 Smith-Waterman library.
#pragma once
#include <stdint.h>

namespace pw {

// Compact residue alphabet: the host maps every character present in a call
// to a code 0..K-1; FILL pads tiles.  The LDS substitution table is KP x KP
// (KP a power of two >= K+1) with the FILL row and column at -1e10, the
// FILL_SCORE of pyPaSWAS/Core/ocl/default_main_cpu.cl:21.
constexpr int KPMAX = 64;
constexpr float FILL_SCORE = -1e10f;
constexpr float GAP_INIT = -1e10f;  // AFFINE_GAP_INIT, default_main_cpu.cl:24

constexpr int WAVE = 64;
constexpr int BWIN = 1024;  // swept residues staged in LDS per refill
constexpr int CAND_BUF = 128;  // candidate starts gathered in LDS per wave before an append

struct SeqRef {
    int64_t off;
    int32_t len;
    int32_t pad;
};

// One tile job of the scoring pass: sequence a_seq along the 64 lanes
// (R residues per lane, stripes of 64*R), b_seq swept column by column.
struct ScoreItem {
    int32_t a_seq, b_seq;
    int32_t out;     // pair index
    int32_t R;       // residues per lane: 1, 2, 4 or 8
    int32_t alen;    // rows to iterate (>= real length; FILL beyond it)
    int32_t blen;    // columns to iterate
    int32_t s0, s1;  // stripes of this item: all of the pair's, or one stripe of a split pair
    int64_t rows_off;  // split pair: its stripe-boundary rows in the link area (floats); -1 otherwise
    int64_t prog_off;  // split pair: its progress counters in the counter area
    int64_t ck_off;    // checkpointed pair: this stripe's checkpoints in the checkpoint area (floats); -1 otherwise
    int64_t bmax_off;  // checkpointed pair: this stripe's block maxima in the checkpoint area (floats)
};

// Checkpointed pairs (long linear-gap pairs of a batched plan).  Their
// scoring pass runs with the query along the lanes, CK_R residues per lane and
// one item per stripe, and every CKB steps of a stripe it saves the wave's
// state (per lane: the R values of its rows at its previous column, then the
// row-above inputs of the next step, upT and dgT) and the largest value of the
// CKB steps before.  The direction pass of such a pair then recomputes only
// the blocks of CKB steps that a path from a candidate start can reach (cells
// up and left of the candidates), every block from its checkpoint and the
// boundary row the scoring pass left above its stripe, with no hand-over
// between waves.
constexpr int CKB = 512;                 // steps per block (divides BWIN, a multiple of HANDOFF and of 64)
constexpr int CK_R = 8;                  // residues per lane of a checkpointed pair (4 or 8)
constexpr int CK_SLOTS = CK_R + 2;       // checkpoint values per lane (linear gaps)
constexpr int CK_MIN_LEN = 1024;         // both sequences at least this long (two stripes, two blocks)
struct CkPair {
    int32_t pair;
    int32_t nblocks;   // blocks per stripe: ceil((tlen + 63) / CKB)
    int64_t ck_off;    // floats: checkpoints [stripe][block][slot][64 lanes]
    int64_t bmax_off;  // floats: block maxima [stripe][block]
    int64_t rows_off;  // floats: the scoring pass's stripe-boundary rows of the pair (link area)
};
__host__ __device__ inline int ck_blocks(int tlen) { return (tlen + 63 + CKB - 1) / CKB; }

// Stripe-boundary rows of a split pair: boundary k (between stripes k and
// k+1) holds LINK_VALS floats per column for blen + 64 columns; counter k
// holds the number of columns of boundary k already written.
constexpr int LINK_VALS = 3;
__host__ __device__ inline int64_t link_row_floats(int blen) { return (int64_t)LINK_VALS * (blen + 64); }

// One work unit of the direction pass: a task, and the stripe of it this
// unit runs (-1: all its stripes, one after the other); for a checkpointed
// pair, the block of CKB steps of that stripe (-1: the whole stripe).
struct DirUnit {
    int32_t task;
    int32_t stripe;
    int32_t block;
};

// One job of the lane-per-pair scoring kernel: a column sequence shared by up
// to 64 pairs (one per lane), each lane sweeping its own row sequence.
// A job too long for one wave is split into one item per chunk, consecutive in
// chunk order; item k reads the left edge of its chunk (T and the I candidate
// of every row) from the link column item k-1 writes, row by row, and waits on
// that column's progress counter (rows published).
struct GroupItem {
    int32_t c_seq;      // column sequence (held as a substitution profile in LDS)
    int32_t entry_off;  // first GroupEntry of this job
    int32_t count;      // lanes in use (<= 64)
    int32_t rows;       // longest row sequence of the job
    int32_t nb;         // columns per chunk: 16 to 96 in steps of 8
    int32_t trans;      // 1 when the column sequence is the query
    int32_t k0, k1;     // chunks [k0, k1) of the column sequence
    int64_t link_in;    // float2 offset of the link column read by chunk k0 (-1: none)
    int64_t link_out;   // float2 offset of the link column written by chunk k1-1 (-1: none)
    int32_t prog_in;    // progress counter of link_in (-1: none)
    int32_t prog_out;   // progress counter of link_out (-1: none)
    int32_t split;      // 1: the job spans several items (maxima combined with an atomic max)
    int32_t pad;
};
constexpr int GROUP_PUB = 16;  // rows of a link column published at a time

struct GroupEntry {
    int32_t row_seq;
    int32_t pair;
};

// One pair of the traceback stages (direction pass, start selection, walk),
// written by the gate kernel for every pair that passes the traceback gate
// (or by the host for a tail round).  Candidates, the pair's starting points
// and the walk's sort keys share the index range [cand_off, cand_off + cand_cap).
struct PairTask {
    int32_t pair;        // pair index
    int32_t q_seq, t_seq;
    int32_t R;           // query residues per lane of the direction pass
    int64_t dir_off;     // byte offset of the pair's tile-major direction area
    int64_t cand_off;
    int64_t ops_off;     // into the ops area (set by the walk)
    int64_t nops;        // op characters written by the walk
    float t1;            // LOWER_LIMIT_SCORE * maximum (f32 product)
    float mp;            // minimum start score of the pair
    float maxv;          // pair maximum
    int32_t cand_cap;
    int32_t ncand;       // candidate cells found (may exceed cand_cap: overflow)
    int32_t nsp;         // starting points emitted
    int32_t same_id;
    int32_t status;      // TASK_*
    int32_t qidx, tidx;  // indices reported in the records
    int32_t out_off;     // first compacted starting point / hit of the pair
    int32_t split;       // 1: the direction pass runs each stripe as its own unit
    int64_t rows_off;    // split: stripe-boundary rows in the link area (floats)
    int64_t prog_off;    // split: progress counters in the counter area
    int32_t ck;          // checkpointed pair: its CkPair index (-1: none)
    int32_t gslot;       // its slot in the read-back summaries (its index in the first round's gated list)
};

// What the host reads back per task: where the pair's records sit (written
// by the walk; the status of a task that did not finish, so that only then the
// whole PairTask has to be read).
struct TaskOut {
    int32_t pair;
    int32_t status;
    int32_t nsp;
    int32_t out_off;
    int64_t ops_off;
    int64_t nops;
};

// status of a PairTask
constexpr int32_t TASK_OK = 0;
constexpr int32_t TASK_DEFERRED = 1;       // no room in the direction or candidate area
constexpr int32_t TASK_CAND_OVERFLOW = 2;  // more candidates than cand_cap
constexpr int32_t TASK_OPS_OVERFLOW = 3;   // no room in the ops area

// Device control block of one execution: work counters of the scoring pass
// (one 64-byte line each) and the bump allocators of the traceback stages.
constexpr int NB_KINDS = 11;
struct Ctl {
    unsigned queue[16 * (1 + NB_KINDS)];
    int ngated;
    unsigned dir_queue;  // next direction-pass unit
    unsigned long long dir_used, cand_used, ops_used, out_used;
    unsigned long long link_used, prog_used;  // split-pair areas of the direction pass
    long long units_used;                     // work units of the direction pass
    unsigned abort_wait;  // set by a wave whose stripe or chunk hand-over wait gave up (see WAIT_POLLS)
    unsigned unfinished;  // tasks of the round left unfinished (no room, or a candidate or record overflow)
    int next_ngated;      // tasks requeued for the next round (requeue_kernel)
    unsigned gate_done;   // blocks of gate_count_kernel finished
    unsigned lay_done;    // blocks of layout_sums_kernel finished
    long long lay_sps, lay_ops;  // records and op characters of the execution (layout_sums_kernel)
};

// Polls a wave makes on a hand-over counter before it gives up, sets
// Ctl::abort_wait and leaves its work item (the host then reports a device
// error and discards the execution).  A poll is one device-scope load plus a
// short sleep, about a microsecond, so the default bound is seconds, far
// beyond any hand-over of a working schedule (a few microseconds per chunk).
constexpr unsigned WAIT_POLLS = 1u << 22;

// A candidate start: x carries two flags in its top bits, whether the cell
// qualifies by its value (unmarked) and by its negated value (once marked)
struct Cand {
    int32_t x, y;
    float v;
};
constexpr int32_t CAND_PLAIN = 1 << 29, CAND_MARKED = 1 << 28;

struct SpRec {
    uint32_t sequence, target, blockX, blockY, valueX, valueY;
    float score, maxScore, posScore;
};

struct HitRec {
    int32_t query, target;
    int32_t shown;
    int32_t s_start, s_end, t_start, t_end;
    int32_t matches, mismatches;
    int32_t length;
    float score;
    int32_t sp_index;
    int64_t ops_offset;
};

// Direction nibble of the direction pass, two cells per byte (the cell of even
// index in the low nibble; the candidate flags travel in the candidate record):
//   bits 1:0  code: 0 none, 1 upper-left / main matrix, 2 upper / I matrix, 3 left / J matrix
//   bit 2     the predecessor along the code has value 0 (a mark here becomes STOP)
//   bit 3     marked by the traceback
// A marked cell always has a code (the marking never reaches a cell of value
// 0), so "marked with code 0" is free and means consumed by the walk
// (IN_ALIGNMENT): the walk reads a cell's code before it clears it.
constexpr uint8_t DB_MAIN = 1, DB_STOP = 4, DB_MARK = 8;
__host__ __device__ inline uint32_t nib_walked(uint32_t b) { return (b & ~3u) | DB_MARK; }
__host__ __device__ inline bool is_walked(uint32_t b) { return (b & DB_MARK) && (b & 3u) == 0; }

}  // namespace pw
