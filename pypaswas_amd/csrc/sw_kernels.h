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
};

// One job of the lane-per-pair scoring kernel: a column sequence shared by up
// to 64 pairs (one per lane), each lane sweeping its own row sequence.
struct GroupItem {
    int32_t c_seq;      // column sequence (held as a substitution profile in LDS)
    int32_t entry_off;  // first GroupEntry of this job
    int32_t count;      // lanes in use (<= 64)
    int32_t rows;       // longest row sequence of the job
    int32_t nb;         // columns per chunk: 16, 32, 48, 64, 80 or 96
    int32_t trans;      // 1 when the column sequence is the query
};

struct GroupEntry {
    int32_t row_seq;
    int32_t pair;
};

// One pair of the direction pass (query along lanes, target swept).
struct DirItem {
    int32_t q_seq, t_seq;
    int32_t pair;    // pair index
    int32_t R;
    int64_t dir_off; // byte offset of the pair's tile-major direction area
    float t1;        // LOWER_LIMIT_SCORE * maximum (f32 product)
    float mp;        // minimum start score of the pair
    int64_t cand_off;
    int32_t cand_cap;
    int32_t pad;
};

struct Cand {
    int32_t x, y;
    float v;
};

// Per-pair traceback / walk bookkeeping.
struct TbItem {
    int32_t pair, q_seq, t_seq, R;
    int64_t dir_off;
    int64_t cand_off;
    int64_t sp_off;      // into the starting-point area (capacity = candidate count)
    int64_t ops_off;     // into the ops area
    int64_t hit_off;     // into the hit area (capacity = starting points)
    int64_t key_off;     // scratch keys for the walk order
    float maxv, mp;
    int32_t qidx, tidx;  // indices reported in the records
    int32_t same_id;
    int32_t ncand;
};

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

// direction byte of the direction pass
//   bits 1:0  code: 0 none, 1 upper-left / main matrix, 2 upper / I matrix, 3 left / J matrix
//   bit 2     the predecessor along the code has value 0 (a mark here becomes STOP)
//   bit 3     candidate start with the unmarked value
//   bit 4     candidate start with the sign-marked value
//   bit 5     marked by the traceback
//   bit 6     consumed by the walk (IN_ALIGNMENT)
constexpr uint8_t DB_STOP = 4, DB_CAND = 8, DB_CANDM = 16, DB_MARK = 32, DB_WALK = 64;

}  // namespace pw
