/*
 * pypaswas_hip.h — C ABI of the MI355X (gfx950) Smith-Waterman scoring and
 * traceback library.  Plain pointers and sizes only; every call returns an
 * int status (PW_OK or a negative PW_E* code) and leaves a message for
 * pw_last_error().  A context is used by one host thread at a time.
 *
 * Each entry point replaces the part of pyPaSWAS's device backend named in
 * its comment (paths relative to the pyPaSWAS source tree).
 */
#ifndef PYPASWAS_HIP_H
#define PYPASWAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PW_OK 0
#define PW_EINVAL (-1)   /* bad argument or option  -> InvalidOptionException */
#define PW_ENOMEM (-2)   /* device or host allocation failed */
#define PW_EDEVICE (-3)  /* HIP runtime / kernel error -> HardwareException */

typedef struct pw_ctx pw_ctx;
typedef struct pw_plan pw_plan;

/* Starting point record, 36 bytes, layout of the StartingPoint struct of
 * pyPaSWAS/Core/ocl/default_main_cpu.cl:66-76 (parsed by
 * pyPaSWAS/Core/StartingPoint.py:54-93).  sequence/target are indices into
 * the call's query/target lists; blockX*8+valueX is the 0-based query
 * position, blockY*8+valueY the 0-based target position. */
typedef struct {
    uint32_t sequence, target, blockX, blockY, valueX, valueY;
    float score, maxScore, posScore;
} pw_starting_point;

/* One processed starting point of the host alignment walk
 * (pyPaSWAS/Core/SmithWaterman.py:703-838), in walk order. */
typedef struct {
    int32_t query, target;   /* indices into the call's lists */
    int32_t shown;           /* 1: the walk produced a hit (a Hit object is built) */
    int32_t s_start, s_end, t_start, t_end;
    int32_t matches, mismatches;
    int32_t length;          /* alignment columns = entries in ops */
    float score;             /* starting-point score (f32) */
    int32_t sp_index;        /* index into starting_points */
    int64_t ops_offset;      /* into ops: 'D' pair column, 'L' query char over '-', 'U' '-' over target char */
} pw_hit;

typedef struct {
    int64_t n_pairs;
    const float *maxima;                 /* [n_pairs] pair maximum (globalMaxima of the last block), in pair order */
    int64_t n_starting_points;
    const pw_starting_point *starting_points; /* per pair, emission order */
    int64_t n_hits;
    const pw_hit *hits;                  /* per pair, walk order (score ascending, stable) */
    int64_t n_ops;
    const char *ops;
    int64_t n_traceback_pairs;           /* pairs whose maximum passed the traceback gate */
    int64_t sp_truncated;                /* starting points dropped by the per-pair cap */
    double cells;                        /* sum over pairs of |q|*|t| (real lengths) */
    double ms_score;                     /* device time of the scoring kernel(s), HIP events */
    double ms_directions;                /* direction recompute of gated pairs */
    double ms_traceback;                 /* start selection + marking */
    double ms_walk;                      /* walk */
    double ms_total;                     /* wall time of the call, host clock */
    int64_t score_launches;
    double ms_readback;                  /* copy of the records to the host (or, host assembly: reading tasks and records back) */
    double ms_tail;                      /* further rounds of pairs that did not fit the work areas (on the device, then on the host) */
    double ms_assemble;                  /* records into pair order (the device layout kernels, or the host assembly) */
    double ms_plan;                      /* host: encoding and work list of a one-shot call (pw_align_pairs) */
    double ms_upload;                    /* host: staging and copy of the plan to the device (one-shot call) */
    const char *aligned;                 /* [3 * n_ops] for each shown hit at 3 * ops_offset: its query row,
                                            alignment row ('|', '.', ' ') and target row, `length` chars each
                                            (SmithWaterman.py:777-830) */
} pw_result;

/* SmithWatermanOcl._initialize_device (pyPaSWAS/Core/SmithWatermanOcl.py:116-129)
 * and SmithWatermanCuda._initialize_device: select a device, create the
 * stream and the persistent device state. */
int pw_init(int device, pw_ctx **out);

/* Release everything the context owns (SmithWatermanOcl._clear_memory, :137-147). */
int pw_destroy(pw_ctx *ctx);

/* SmithWatermanOcl._device_global_mem_size (:131-134). */
int pw_device_mem_bytes(pw_ctx *ctx, uint64_t *bytes);

/* Scoring constants, replacing the kernel-template substitution of
 * pyPaSWAS/Core/PaSWAS.py:55-91 (default_score.cl).  subst is dim x dim,
 * row = query residue index (char - char_offset), FP32.  gap_ext == 0 selects
 * the linear kernels (Scores.py:58, SmithWaterman.py:149). */
int pw_set_scoring(pw_ctx *ctx, const float *subst, int dim, int char_offset, float gap, float gap_ext,
                   float lower_limit, float minimum_score);

/* Align explicit (query, target) pairs; replaces one pass of
 * SmithWaterman.align_sequences (pyPaSWAS/Core/SmithWaterman.py:441-560):
 * copy_sequences, _calculate_score, _traceback_host and the walk of
 * _print_alignments, with the per-pair minimum start score of
 * SmithWatermanOcl._fill_max_possible_score (:302-307).
 *   qres/qoff/qlen  nq query sequences (uppercase residues), tres/... nt targets
 *   pq/pt           n_pairs pair indices into the two lists
 *   min_start_score [n_pairs] f32 (maxPossibleScore of default_main_cpu.cl:412)
 *   same_id         [n_pairs] 1 when the two record ids are equal (SmithWaterman.py:752), may be NULL
 *   minimum         settings.minimum_score as a double (SmithWaterman.py:758)
 *   sp_cap          starting points kept per pair (SmithWaterman.py:144, 683-685)
 *   mode            0 = plain per-batch launch, 1 = batched persistent kernel
 * The result is owned by the library until pw_free_result. */
int pw_align_pairs(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
                   const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt,
                   const int32_t *pq, const int32_t *pt, int64_t n_pairs, const float *min_start_score,
                   const uint8_t *same_id, double minimum, int32_t sp_cap, int32_t mode, pw_result **out);

/* The same work as pw_align_pairs split in two, so that sequences, the
 * substitution tables and the scoring work list stay resident in device
 * memory across executions (the role of copy_sequences + _init_memory of
 * SmithWatermanOcl.py:223-293, done once).  pw_prepare encodes and uploads;
 * pw_execute runs scoring, gate, traceback and walk on the resident plan;
 * pw_release_plan frees it. */
int pw_prepare(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
               const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt,
               const int32_t *pq, const int32_t *pt, int64_t n_pairs, const float *min_start_score,
               const uint8_t *same_id, int32_t mode, pw_plan **out);
int pw_execute(pw_ctx *ctx, pw_plan *plan, double minimum, int32_t sp_cap, pw_result **out);
int pw_release_plan(pw_ctx *ctx, pw_plan *plan);

/* Routing of a prepared plan, for diagnostics and tests (no counterpart in
 * pyPaSWAS, whose kernels have one launch shape): info[0..n) receives, in this
 * order, pairs, lane-per-pair items, wave-per-pair items, the orientation of
 * the lane-per-pair jobs (0: the target is the shared column sequence, 1: the
 * query), cells scored lane-per-pair, cells scored wave-per-pair, the widest
 * chunk of a lane-per-pair job, the direction pass's largest residues per lane,
 * the lane-per-pair items that belong to jobs split into one item per chunk
 * and the checkpointed pairs (long linear-gap pairs whose scoring pass saves
 * checkpoints, so that their direction pass recomputes only the blocks a path
 * can reach) (at most 10 values). */
int pw_plan_info(const pw_plan *plan, int64_t *info, int32_t n);

/* pw_execute with the device work of the execution (scoring, gate, direction
 * pass, start selection, walk and the control-block read-back) captured once
 * in a HIP graph and replayed on later calls; the graph is captured again
 * when a work area was resized or `minimum` changed.  Results are identical
 * to pw_execute.  Replaces the per-block launch loops of
 * SmithWaterman._calculate_score / _traceback_host
 * (pyPaSWAS/Core/SmithWaterman.py:592-653) for a job run repeatedly.  In the
 * result, ms_score holds the device time of the whole graph and the other
 * stage times only cover tail rounds. */
int pw_execute_graph(pw_ctx *ctx, pw_plan *plan, double minimum, int32_t sp_cap, pw_result **out);

/* Sizing of the device work areas of the traceback stages, the role of the
 * memory model of SmithWaterman._get_max_number_sequences and
 * maximum_memory_usage (pyPaSWAS/Core/SmithWaterman.py:202-221).  Values <= 0
 * keep the current setting.
 *   max_direction_bytes  upper bound of the direction area (default: a quarter
 *                        of device memory; the first area is at most 256 MiB)
 *   candidates_per_pair  candidate-start slots a gated pair gets (default 32)
 *   ops_bytes            first size of the op-character area (default 16 MiB)
 *   grow                 1: after an execution whose pairs did not fit, the
 *                        areas of the next one are sized to what it needed
 * Pairs that do not fit finish in host-driven tail rounds with exact sizes,
 * so the limits change speed and memory, never results.  Applies to plans
 * prepared afterwards and to the one-shot calls. */
int pw_set_limits(pw_ctx *ctx, int64_t max_direction_bytes, int32_t candidates_per_pair, int64_t ops_bytes,
                  int32_t grow);

/* Routing rules of the batched scoring pass for plans prepared afterwards (no
 * counterpart in pyPaSWAS, whose kernels have one launch shape; results never
 * depend on them).  orient: shared column sequence of lane-per-pair jobs, 0 the
 * target, 1 the query, -1 whichever the cost estimate prefers (default).
 * group_max_share: longest lane-per-pair job as a share of the plan's cells
 * per device lane (default 0.25; <= 0 restores it; a large value keeps every
 * job whose lanes are busy enough, e.g. for small plans in tests).
 * dir_residues: largest query residues per lane of the direction pass, 2, 4
 * or 8 (0: chosen per plan so that the pairs give enough work units). */
int pw_set_routing(pw_ctx *ctx, int32_t orient, double group_max_share, int32_t dir_residues);

/* Where the records of executions afterwards are put into pair order (no
 * counterpart in pyPaSWAS, which assembles its hits on the host; results never
 * depend on it).  on_device != 0 (default): when every pair of an execution
 * finished on the device, its records, op characters and alignment rows are laid
 * out in pair order there and copied once into the result's pinned arrays;
 * 0: they are read back per pair and assembled on the host. */
int pw_set_assembly(pw_ctx *ctx, int32_t on_device);

/* Bound of the device-side waits of the stripe and chunk hand-over of split
 * pairs and split lane-per-pair jobs (no counterpart in pyPaSWAS, whose
 * traceback spins on semaphores without a bound, default_main_cpu.cl:457-483):
 * a wave that has polled a hand-over counter `polls` times gives up, the
 * execution's other waves leave their items, and the call returns PW_EDEVICE
 * with its results discarded, instead of the GPU hanging on a scheduling
 * mistake.  polls < 0 restores the default (2^22 polls, seconds: far beyond
 * any working hand-over); 0 gives up on the first wait that is not already
 * satisfied (for tests of the error path).  Applies to later calls. */
int pw_set_wait_limit(pw_ctx *ctx, int64_t polls);

/* Score pass only: pair maxima, the quantity the traceback gate of
 * default_main_cpu.cl:408-410 reads.  maxima_out is [n_pairs]. */
int pw_score_pairs(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
                   const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt,
                   const int32_t *pq, const int32_t *pt, int64_t n_pairs, int32_t mode, float *maxima_out,
                   double *ms_out);

/* Device memory pw_score_pairs_noalloc may use, per GPU (64 MiB). */
#define PW_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* pw_score_pairs without any allocation: no counterpart in pyPaSWAS.  For
 * callers that cannot allocate device memory (code that must not call
 * hipMalloc, such as tight serving loops or harnesses that forbid it).  Same
 * arguments, results and errors as pw_score_pairs; the plan is built in host
 * memory and every device array of the call (the encoded sequences, tables and
 * work list, control block, maxima, the split pairs' boundary rows and the
 * kernels' per-workgroup scratch) lives in one static device array of
 * PW_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends; no
 * device or pinned memory is allocated or freed.  A job that needs more than
 * that (inputs plus one workgroup's scratch per kernel) is rejected with
 * PW_EINVAL; persistent grids are lowered to fit, which changes speed, never
 * the maxima.  Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (from two contexts or threads) race and must not be made. */
int pw_score_pairs_noalloc(pw_ctx *ctx, const uint8_t *qres, const int64_t *qoff, const int32_t *qlen, int32_t nq,
                           const uint8_t *tres, const int64_t *toff, const int32_t *tlen, int32_t nt,
                           const int32_t *pq, const int32_t *pt, int64_t n_pairs, int32_t mode, float *maxima_out,
                           double *ms_out);

/* Autotuner.  Times the scoring kernels on synthetic pairs for the gap model
 * of the current scoring scheme: the wave-per-pair tile kernel with 1, 2, 4
 * and 8 residues per lane for every bucket of lane-side length (<= 64, 128,
 * 256, 512, 1024, 2048, more) and swept length (<= 32, 64, 128, 256, 512,
 * 1024, more), and the chunk widths 32..96 of lane-per-pair jobs whose target
 * exceeds 96 residues.  Plans prepared afterwards take, per pair, the
 * orientation and residues per lane with the highest measured rate of its
 * buckets (instead of the built-in cost model), and the fastest chunk width.
 * Replaces nothing in pyPaSWAS, whose kernels have one fixed shape
 * (8x8 blocks of 2x2 work-items x 4x4 cells, default_main_cpu.cl:108-222);
 * results never depend on the shape.  *seconds: time the tuning took. */
int pw_autotune(pw_ctx *ctx, double *seconds);

/* The tuning table of a gap model (affine != 0: affine): rates[7][7][4] in
 * cells per microsecond (lane-side bucket, swept bucket, 1/2/4/8 residues per
 * lane), group_rates[5] for chunk widths 32, 48, 64, 80, 96, the chosen width
 * and whether the table is in use. */
int pw_get_tuning(pw_ctx *ctx, int32_t affine, float *rates, float *group_rates, int32_t *group_nb_long,
                  int32_t *valid);

/* Installs a saved table (rates as pw_get_tuning returns them); rates == NULL
 * drops the table and restores the cost model. */
int pw_set_tuning(pw_ctx *ctx, int32_t affine, const float *rates, int32_t group_nb_long);

/* Plan-level tuning of a batched plan (pw_prepare, mode 1): its scoring pass
 * is timed under each combination of three routing choices -- the longest
 * lane-per-pair job allowed (1/8, 1/4 or 1/2 of the plan's cells per device
 * lane), the chunk width of lane-per-pair jobs whose column sequence exceeds
 * 96 residues (80, 64 or 96), and tile shapes from the pw_autotune table or
 * from the cost model (when a table exists) -- and the fastest work list is
 * kept for the plan's executions; the defaults (1/4, 80, cost model) stay unless
 * another combination is at least 3% faster.  Results never depend on these
 * choices.
 * *seconds: tuning time; *score_ms: the kept scoring pass; choice[3]: share x
 * 1000, chunk width, table used.  Plain plans (mode 0) are left as they are. */
int pw_tune_plan(pw_ctx *ctx, pw_plan *plan, double *seconds, double *score_ms, int32_t *choice);

/* ------------------------------------------------------------------------
 * Serving: a request loop with continuous batching.  pyPaSWAS has no request
 * loop; a job here is the work of one SmithWaterman.align_sequences pass
 * (pyPaSWAS/Core/SmithWaterman.py:441-560) with the arguments of
 * pw_align_pairs.  The caller's thread drives the loop through pw_server_poll;
 * the server runs batches on `lanes` stream pairs of its own, concurrently on
 * the device.  Whenever a lane is idle, the queue head and the later queued
 * jobs with the same scheme and minimum score are merged into one batch while
 * the direction bytes they could need (every pair passing the traceback gate)
 * fit in the largest free range of the direction arena of `direction_bytes`.
 * A job's records equal those of the same job run alone through
 * pw_align_pairs (pairs are independent), with the job's own sequence indices.
 * ------------------------------------------------------------------------ */
typedef struct pw_server pw_server;

int pw_server_create(pw_ctx *ctx, int32_t lanes, int64_t direction_bytes, pw_server **out);

/* Scoring scheme `scheme` (0..7) of later jobs; same arguments as pw_set_scoring. */
int pw_server_set_scoring(pw_server *srv, int32_t scheme, const float *subst, int dim, int char_offset, float gap,
                          float gap_ext, float lower_limit, float minimum_score);

/* Queues a job (inputs are copied); *job_id identifies it in poll and take. */
int pw_server_submit(pw_server *srv, int32_t scheme, const uint8_t *qres, const int64_t *qoff,
                     const int32_t *qlen, int32_t nq, const uint8_t *tres, const int64_t *toff,
                     const int32_t *tlen, int32_t nt, const int32_t *pq, const int32_t *pt, int64_t n_pairs,
                     const float *min_start_score, const uint8_t *same_id, double minimum, int32_t sp_cap,
                     int64_t *job_id);

/* Finishes batches whose device work ended, admits queued jobs onto idle
 * lanes and returns up to max_ids ids of completed jobs.  block != 0 waits
 * until at least one job has completed (or nothing is queued or running). */
int pw_server_poll(pw_server *srv, int32_t block, int64_t *done_ids, int32_t max_ids, int32_t *n_done);

/* The result of a completed job, owned by the caller until pw_free_result. */
int pw_server_take(pw_server *srv, int64_t job_id, pw_result **out);

/* stats[0..10]: batches launched, jobs admitted, jobs completed, peak arena
 * bytes reserved, jobs queued, jobs running, arena bytes, host microseconds
 * forming and launching batches, host microseconds reading them back, device
 * microseconds of the batches, device or pinned buffers grown so far. */
int pw_server_stats(pw_server *srv, int64_t *stats, int32_t n);

int pw_server_destroy(pw_server *srv);

/* Free a result of pw_align_pairs. */
int pw_free_result(pw_ctx *ctx, pw_result *res);

/* Last error message of the context (or of pw_init when ctx is NULL). */
const char *pw_last_error(pw_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* PYPASWAS_HIP_H */
