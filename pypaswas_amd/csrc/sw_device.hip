// gfx950 Smith-Waterman kernels: scoring pass, direction pass, start
// selection + path marking, and the alignment walk.
//
// Semantics follow pyPaSWAS's OpenCL-CPU kernels
// (pyPaSWAS/Core/ocl/default_main_cpu.cl) and its host walk
// (pyPaSWAS/Core/SmithWaterman.py:665-841); the layout and schedule are this
// library's own:
//
//  * wave-per-pair tiles: one 64-lane wavefront per tile; lane k owns R
//    consecutive residues of the tile sequence A (a stripe of 64*R rows) and
//    sweeps the other sequence B column by column, lane k running one column
//    behind lane k-1 (anti-diagonal wavefront inside the wave);
//  * the values a lane needs from the row above (H, or T = max(M, I, J) and
//    the J candidate for affine gaps) arrive from lane k-1 through a DPP
//    wave_shr:1 move; lane 0 receives the previous stripe's bottom row and
//    lane 63 hands its own down, HANDOFF (32) columns at a time through LDS
//    and a scratch row in device memory;
//  * the substitution matrix (transposed, KP x KP, FILL row/col -1e10) and a
//    window of B's residue codes live in LDS;
//  * lane-per-pair jobs (the scoring pass only): 64 pairs sharing one column
//    sequence, one pair per lane, see group_tile;
//  * all score arithmetic is IEEE FP32 adds and max, in the order that yields
//    exactly the reference's values: max(a+s, b+s) == max(a, b)+s because FP32
//    rounding is monotone, and the affine recurrence is rewritten only in
//    ways that keep T, the direction bytes and the maxima bit for bit (gap
//    candidates from the unclamped diagonal sum with J clamped at 0, see
//    group_tile; a row shift for whole-number scores in the lane-per-pair
//    kernels);
//  * start selection and the walk follow paths serially, one thread per pair,
//    or for long pairs one wave per pair through LDS windows of the direction
//    area (PathWindow).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sw_kernels.h"

namespace pw {

__device__ __forceinline__ float shr1(float old_lane0, float v) {
    // DPP wave_shr:1 — lane k receives lane k-1's v; lane 0 keeps old_lane0.
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old_lane0), __float_as_int(v), 0x138, 0xf, 0xf, false));
}

__device__ __forceinline__ int shr1i(int old_lane0, int v) {
    return __builtin_amdgcn_update_dpp(old_lane0, v, 0x138, 0xf, 0xf, false);
}

// DPP wave_shl:1 — lane k receives lane k+1's v; lane 63 keeps old_lane63.
__device__ __forceinline__ float shl1(float old_lane63, float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old_lane63), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ int shl1i(int old_lane63, int v) {
    return __builtin_amdgcn_update_dpp(old_lane63, v, 0x130, 0xf, 0xf, false);
}

__device__ __forceinline__ float fmax3(float a, float b, float c) { return __builtin_fmaxf(__builtin_fmaxf(a, b), c); }

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = __builtin_fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Scratch rows handed from one stripe or chunk to the next (by the same wave,
// or by another wave for a split pair) are written and read only with
// device-scope relaxed atomics: coherent per location, they need no fence and
// no cache maintenance.
__device__ __forceinline__ float load_scr(const float *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_scr(float *p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Stripe hand-over of a split pair.  Boundary rows and counters are only
// accessed with device-scope (relaxed) atomics, which are coherent across the
// XCDs' L2 caches without whole-cache write-back or invalidation (the cost of
// an agent-scope release or acquire fence on this chip).  The producer wave
// waits until its row stores are performed (s_waitcnt), then one lane stores
// the new column count; the consumer wave spins (with s_sleep) on the count
// and only then issues its row loads.  Waits end: the producer stripe was
// dequeued earlier by a running wave, and it publishes every chunk up to the
// last column.  Every wait is bounded all the same (wait_rows), so that a
// planning or routing mistake ends in a device error reported by the host
// instead of a hung GPU: once one wave gave up, every later wait of the
// execution gives up at its first unsatisfied poll and every kernel drains.
__device__ __forceinline__ void publish_progress(unsigned *p, unsigned v) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned load_counter(const unsigned *p) {
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Waits until *p >= need and returns the count it saw.  After `limit` polls,
// or at once when another wave already gave up, the wave sets *abort and
// returns `need` as if the rows were there: the item then runs to its end on
// stale boundary values (addresses never depend on values, so every access
// stays in bounds and the work stays bounded), and the host discards the
// execution when it reads the flag.
__device__ __forceinline__ unsigned wait_rows(const unsigned *p, unsigned need, unsigned *abort, unsigned limit) {
    unsigned v = load_counter(p);
    for (unsigned k = 0; v < need; ++k) {
        if (k >= limit || load_counter(abort) != 0u) {
            if ((threadIdx.x & 63) == 0) __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return need;
        }
        __builtin_amdgcn_s_sleep(2);
        v = load_counter(p);
    }
    __builtin_amdgcn_s_waitcnt(0);
    return v;
}

__device__ __forceinline__ void wait_progress(const unsigned *p, unsigned need, unsigned *abort, unsigned limit) {
    (void)wait_rows(p, need, abort, limit);
}

// Columns of a stripe boundary row handed over at a time (through LDS, then
// device memory, then a progress count for a split pair): the stripe below
// can start a chunk once lane 63 of the stripe above has finished it, so the
// pipeline lag per stripe is HANDOFF + 63 steps.
constexpr int HANDOFF = 32;
static_assert((HANDOFF & (HANDOFF - 1)) == 0 && HANDOFF <= WAVE, "HANDOFF: a power of two up to the wave size");
static_assert(CKB % WAVE == 0 && CKB % HANDOFF == 0 && BWIN % CKB == 0, "a block starts on a hand-over chunk, "
              "a residue-offset reload and inside one residue window");

// Block-wise work of a checkpointed pair (CkPair).  Scoring pass (CK): ck is
// this stripe's checkpoint array, written at every block start after the
// first, and bmax its block maxima.  Direction pass (t1 >= 0): the unit runs
// steps [t0, t1) of one stripe, starting from the checkpoint at ck (nullptr
// for the first block), with the row above read from the scoring pass's
// boundary row and no hand-over to the stripe below.
struct CkArgs {
    float *ck = nullptr;
    float *bmax = nullptr;
    int t0 = 0, t1 = -1;
    // scoring pass with a per-lane profile (PROF): this wave's profile in LDS
    // and the residue codes it covers (slot c < nres: code c, slot nres: FILL)
    float *prof = nullptr;
    int nres = 0;
};
// The block arguments for the tile instance of R residues per lane: only the
// CK_R instance runs checkpoint blocks.
template <int R>
__device__ __forceinline__ CkArgs ck_for(const CkArgs &c) {
    if constexpr (R == CK_R) return c;
    else return CkArgs();
}

// ---------------------------------------------------------------------------
// One tile job: A along lanes (stripes of 64*R rows), B swept.
//   DIRS = false: returns the pair maximum (max over all cells of H, or of
//                 max(M, I, J) for affine gaps), the globalMaxima value of
//                 default_main_cpu.cl:188-220.
//   DIRS = true : A must be the query; writes one direction byte per cell
//                 (tile-major, see cell_offset) and appends candidate starts.
// Stripes [s_first, s_last) are processed.  links == nullptr: one wave runs
// them in order, handing each stripe's bottom row to the next through scr in
// place.  Otherwise (a split pair, one stripe per wave) the row above comes
// from boundary s-1 of links and the bottom row goes to boundary s, chunk by
// chunk, synchronised through the progress counters prog.
// ---------------------------------------------------------------------------
template <bool AFF, int R, bool DIRS, bool CK = false, bool PROF = false>
__device__ float tile(const uint8_t *__restrict__ res, SeqRef A, SeqRef B, int alen_it, int blen_it,
                      const float *subT, int lkp, uint8_t *bwin, float4 *stg, float *__restrict__ scr, float G,
                      float E, int fill, uint8_t *__restrict__ dir, float t1, float mp, int *__restrict__ cand_count,
                      Cand *__restrict__ cands, int cand_cap, int s_first, int s_last, float *__restrict__ links,
                      unsigned *__restrict__ prog, unsigned *abort, unsigned wlimit, CkArgs ckv = CkArgs()) {
    static_assert(!CK || (!DIRS && !AFF && R == CK_R), "checkpoints: the linear scoring pass at CK_R residues per lane");
    static_assert(!PROF || (CK && R % 4 == 0), "per-lane profile: the checkpointing scoring pass");
    const int lane = threadIdx.x & 63;
    const int rows = WAVE * R;
    const int nstripes = (alen_it + rows - 1) / rows;
    const int nsteps = blen_it + WAVE - 1;
    const char *subB = reinterpret_cast<const char *>(subT);
    // The gap scores as vector registers: a VOP2 add whose operands are all
    // VGPRs dual-issues (measured on MI355X: ~1.4 wave instructions per SIMD
    // quad-cycle with 4 waves), one that reads an SGPR does not (~0.9)
    asm volatile("" : "+v"(G), "+v"(E));
    float rmax = 0.0f;
    // J candidate entering row 1: max(-1e10 + E, 0 + G), which both passes
    // keep clamped at 0 (see the recurrences below)
    const float bJ = 0.0f;
    const int64_t row_floats = link_row_floats(blen_it);
    if (s_last > nstripes) s_last = nstripes;
    // a candidate start needs value >= t1 and >= mp (or -value for a marked cell)
    const float thr = __builtin_fmaxf(t1, mp);
    // direction pass over one block of a checkpointed pair: steps [tb0, te)
    const bool blk = DIRS && ckv.t1 >= 0;
    const int tb0 = blk ? ckv.t0 : 0;
    const int te = blk ? min(ckv.t1, nsteps) : nsteps;
    float pm = 0.0f;  // checkpointing scoring pass: maximum of the blocks done
    // LDS staging of the row above, HANDOFF columns at a time (lane 0 takes one
    // column per step, as a broadcast read)
    float4 *stg_in = stg;

    // window of the swept residues for steps [t0, t0 + BWIN): every load is
    // issued before the first LDS store, so the wave waits for memory once
    auto load_window = [&](int t0) {
        constexpr int PER_LANE = (BWIN + WAVE - 1 + WAVE - 1) / WAVE;
        uint8_t v[PER_LANE];
#pragma unroll
        for (int k = 0; k < PER_LANE; ++k) {
            const int i = lane + k * WAVE;
            const int c = t0 - (WAVE - 1) + i;
            v[k] = (i < BWIN + WAVE - 1 && c >= 0 && c < B.len && c < blen_it) ? res[B.off + c] : (uint8_t)fill;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER_LANE; ++k) {
            const int i = lane + k * WAVE;
            if (i < BWIN + WAVE - 1) bwin[i] = v[k];
        }
        __syncthreads();
    };
    // Candidate starts are gathered in LDS and appended to the pair's list
    // in blocks, one atomic per block instead of one per candidate cell
    Cand *cbuf = reinterpret_cast<Cand *>(stg + 2 * WAVE);
    int nbuf = 0;  // wave-uniform
    auto flush_cands = [&]() {
        if (nbuf == 0) return;
        __syncthreads();
        unsigned base = 0;
        if (lane == 0) base = (unsigned)atomicAdd(cand_count, nbuf);
        base = __builtin_amdgcn_readfirstlane(base);
        for (int k = lane; k < nbuf; k += WAVE) {
            const int idx = (int)base + k;
            if (idx < cand_cap) cands[idx] = cbuf[k];
        }
        __syncthreads();
        nbuf = 0;
    };

    // Score reads.  Table form: the entry of the swept residue's row (byte
    // offset boff = code << lkp) at this lane's residue of row r (a4[r]), one
    // address add and one LDS read per cell.  Profile form (PROF, small
    // alphabets): the lane's R scores against every residue code, as float4s
    // of rows 4g..4g+3 laid out [slot][g][lane] (the wave's 16-byte reads of
    // one g are contiguous, free of bank conflicts), and boff is the slot's
    // byte offset; one add and R/4 wide reads per step.
    constexpr int PSHIFT = PROF ? 6 + (R == 8 ? 3 : 2) + 2 : 0;  // bytes per slot: 64 lanes x R floats
    const char *profl = PROF ? reinterpret_cast<const char *>(ckv.prof) + lane * 16 : nullptr;
    auto off_of = [&](int code) {
        if constexpr (PROF) return min(code, ckv.nres) << PSHIFT;
        else return code << lkp;
    };
    for (int s = s_first; s < s_last; ++s) {
        int a4[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = s * rows + lane * R + r;
            a4[r] = ((i < A.len) ? (int)res[A.off + i] : fill) * 4;
        }
        if constexpr (PROF) {
            // this stripe's profile; each lane writes and reads only its own
            // entries, so LDS order within the lane is all it needs
            float *pw = ckv.prof + lane * 4;
            for (int c = 0; c <= ckv.nres; ++c) {
                const int b = c < ckv.nres ? c : fill;
#pragma unroll
                for (int r4 = 0; r4 < R; r4 += 4) {
                    float4 v;
                    v.x = *reinterpret_cast<const float *>(subB + (b << lkp) + a4[r4]);
                    v.y = *reinterpret_cast<const float *>(subB + (b << lkp) + a4[r4 + 1]);
                    v.z = *reinterpret_cast<const float *>(subB + (b << lkp) + a4[r4 + 2]);
                    v.w = *reinterpret_cast<const float *>(subB + (b << lkp) + a4[r4 + 3]);
                    *reinterpret_cast<float4 *>(pw + (c << (PSHIFT - 2)) + r4 * WAVE) = v;
                }
            }
        }
        // state of this lane's rows at the previous column: T, I and, for
        // affine gaps, d + G with d = T[i-1][j-1] + s (and, for the direction
        // pass, d itself)
        float Tp[R], Mp[R], Ip[R], Dp[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Tp[r] = 0.0f;
            Mp[r] = 0.0f + G;  // M + G of the virtual column left of the matrix
            Ip[r] = GAP_INIT;
            Dp[r] = 0.0f;      // M = 0 there
        }
        const bool above = s > 0;
        const bool below = !blk && s + 1 < nstripes;
        const float *sin = links ? links + (int64_t)(s - 1) * row_floats : scr;  // row above (when above)
        float *sout = links ? links + (int64_t)s * row_floats : scr;             // bottom row (when below)
        // columns [c0, c0 + HANDOFF) of the row above into LDS (one column per lane)
        auto stage_in = [&](int c0) {
            float4 v = make_float4(0.0f, bJ, 0.0f, 0.0f);
            const int c = c0 + lane;
            if (lane < HANDOFF && c < blen_it) {
                v.x = load_scr(&sin[3 * c]);
                if (AFF) v.y = load_scr(&sin[3 * c + 1]);
                if (DIRS && AFF) v.z = load_scr(&sin[3 * c + 2]);
            }
            __syncthreads();
            stg_in[lane] = v;
            __syncthreads();
        };
        if (above) {
            if (prog && blen_it > 0) wait_progress(prog + (s - 1), (unsigned)min(HANDOFF, blen_it), abort, wlimit);
            stage_in(tb0);
        }
        // inputs of step 0: lane 0 reads column 0 of the row above; the other
        // lanes start in the virtual columns left of the matrix
        float upT = 0.0f, upJ = bJ, upM = 0.0f;
        if (above && lane == 0) {
            const float4 v = stg_in[0];
            upT = v.x;
            upJ = v.y;
            upM = v.z;
        }
        float dgT = 0.0f, dgM = 0.0f;
        if (blk && tb0 > 0) {
            // the state the scoring pass saved before step tb0 (linear gaps:
            // every lane's R values at its previous column, and the row-above
            // inputs of step tb0)
            const float *c = ckv.ck + lane;
#pragma unroll
            for (int r = 0; r < R; ++r) Tp[r] = c[r * WAVE];
            upT = c[R * WAVE];
            dgT = c[(R + 1) * WAVE];
        }

        // Table offsets of the swept residues: lane k sweeps column t - k at
        // step t, so its column at step t+1 is lane k-1's at step t (a DPP
        // shift), and only lane 0 takes a new one, column t+1, from fB.  fB
        // holds the offsets of 64 consecutive columns (lane l: column base + l,
        // read from the LDS window every 64 steps) and moves down one lane per
        // step.  Substitution scores are read one step ahead, so that the LDS
        // latency overlaps a step of arithmetic.
        // (an opaque window base: letting the compiler fold the block's first
        // step into the window's addressing raised the kernel's register use
        // from 112 to 166 VGPRs, one wave per SIMD fewer)
        int wbase = tb0 & ~(BWIN - 1);
        asm volatile("" : "+s"(wbase));
        load_window(wbase);
        const int w0 = tb0 & (BWIN - 1);
        int fB = off_of((int)bwin[w0 + (WAVE - 1) + lane]);  // columns tb0 .. tb0 + 63
        // column tb0 - lane (< 0 for lanes 1..63 at step 0)
        int boff = tb0 > 0 ? off_of((int)bwin[w0 + (WAVE - 1) - lane]) : (lane == 0 ? fB : off_of(fill));
        auto read_scores = [&](float (&o)[R]) {
            if constexpr (PROF) {
#pragma unroll
                for (int r4 = 0; r4 < R; r4 += 4) {
                    const float4 v = *reinterpret_cast<const float4 *>(profl + boff + r4 * (WAVE * 4));
                    o[r4] = v.x;
                    o[r4 + 1] = v.y;
                    o[r4 + 2] = v.z;
                    o[r4 + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) o[r] = *reinterpret_cast<const float *>(subB + boff + a4[r]);
            }
        };
        float scn[R];
        read_scores(scn);
        // lane 63's bottom-row values of the latest steps, the newest in lane 63
        // (lane L holds the one of the step 63 - L steps back)
        float oT = 0.0f, oJ = 0.0f, oM = 0.0f;
        // the column after which lane 63's next hand-over group is complete: the
        // last of each HANDOFF columns, and the row's last column (none for an
        // empty row)
        int next_ho = blen_it > 0 ? min(HANDOFF - 1, blen_it - 1) : INT32_MIN;

        // one step, scores in sc, the next step's read into nx; two steps per
        // loop iteration alternate the two register sets
        auto step = [&](const int t, float (&sc)[R], float (&nx)[R]) {
            const int j = t - lane;  // this lane's column
            const int tn = t + 1;
            // input of step t+1 for lane 0: column tn of the row above (read
            // before the score reads below, so that the wait for it at the end
            // of the step leaves those in flight)
            float4 inj = make_float4(0.0f, bJ, 0.0f, 0.0f);
            if (above) {
                if ((tn & (HANDOFF - 1)) == 0) {
                    if (prog && tn < blen_it)
                        wait_progress(prog + (s - 1), (unsigned)min(tn + HANDOFF, blen_it), abort, wlimit);
                    stage_in(tn);
                }
                inj = stg_in[tn & (HANDOFF - 1)];
            }
            // scores of step t+1 (also at t+1 == nsteps: offsets always lie in the table)
            if ((tn & (WAVE - 1)) == 0) {
                if ((tn & (BWIN - 1)) == 0) load_window(tn);
                fB = off_of((int)bwin[(tn & (BWIN - 1)) + (WAVE - 1) + lane]);
            } else {
                fB = shl1i(fB, fB);
            }
            boff = shr1i(fB, boff);
            read_scores(nx);

            float tup = upT, jin = upJ;
            uint32_t bits[R];
            if (!DIRS) {
                // diagonal sums T[i-1][j-1] + s of all of this step's cells
                // first, so that each row's T is replaced in place
                float dg[R];
                dg[0] = dgT + sc[0];
#pragma unroll
                for (int r = 1; r < R; ++r) dg[r] = Tp[r - 1] + sc[r];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (AFF) {
                        // With G, E <= 0 a negative gap value never reaches
                        // T and M + G <= 0 whenever M is clamped, so the
                        // gap candidates take the unclamped d + G and J is
                        // clamped at 0 instead: I' = max(I' + E, d + G) <= I
                        // and J' = max(J' + E, d + G, 0) = max(J, 0) keep the
                        // positive parts of I and J (by induction), and
                        // T = max(d, I', J') is the reference's T exactly.
                        const float MG = dg[r] + G;
                        const float I = __builtin_fmaxf(Ip[r] + E, Mp[r]);  // Mp holds d + G
                        const float T = fmax3(dg[r], I, jin);
                        rmax = __builtin_fmaxf(rmax, T);
                        jin = fmax3(jin + E, MG, 0.0f);
                        Tp[r] = T;
                        Mp[r] = MG;
                        Ip[r] = I;
                    } else {
                        const float h = fmax3(dg[r], __builtin_fmaxf(tup, Tp[r]) + G, 0.0f);
                        rmax = __builtin_fmaxf(rmax, h);
                        Tp[r] = h;
                        tup = h;
                    }
                }
            } else {
                float td = dgT, md = dgM;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (AFF) {
                        // the scoring pass's clamp-free form: d unclamped in
                        // the gap candidates, J clamped at 0, T exact; with
                        // T > 0, T == M exactly when T == d, T == I exactly
                        // when T == I', and M = 0 exactly when d <= 0, so the
                        // direction byte and the candidate test are the
                        // reference's
                        const float d = td + sc[r];
                        const float dG = d + G;
                        const float I = __builtin_fmaxf(Ip[r] + E, Mp[r]);  // Mp holds d + G
                        const float T = fmax3(d, I, jin);
                        // default_main_cpu.cl:283-345: source = M, I, J in that
                        // priority among equal positive values, none when 0; a
                        // main-matrix cell stops on a zero diagonal predecessor
                        // (M = 0 there, i.e. d <= 0) and is a candidate start
                        // when M (or -M, for a marked cell) reaches both
                        // thresholds (the two candidate flags go with the
                        // candidate record, not the byte).  Every value is
                        // formed, then selected: no divergent branch per cell.
                        uint32_t main = DB_MAIN | (md <= 0.0f ? DB_STOP : 0u);
                        uint32_t other = T == I ? 2u : 3u;
                        asm volatile("" : "+v"(main), "+v"(other));
                        const bool isM = T == d, pos = T > 0.0f;
                        bits[r] = pos ? (isM ? main : other) : 0u;
                        const float jout = fmax3(jin + E, dG, 0.0f);
                        td = Tp[r];
                        md = Dp[r];
                        Tp[r] = T;
                        Dp[r] = d;
                        Mp[r] = dG;
                        Ip[r] = I;
                        jin = jout;
                    } else {
                        const float hd = td;
                        const float d = hd + sc[r];
                        // default_main_cpu.cl:162-177: LEFT (x-1), UP (y-1), UL,
                        // first strict improvement over 0 wins: the value is
                        // max(0, l, u, d) and the code that of the first of
                        // l, u, d equal to it (when it is positive)
                        const float l = tup + G, u = Tp[r] + G;
                        const float cur = fmax3(__builtin_fmaxf(l, u), d, 0.0f);
                        // STOP when the predecessor along the code is 0; only
                        // diagonal cells are candidate starts.  With G <= 0 a
                        // positive gap value cur = v + G has v >= cur > 0, so
                        // only the diagonal predecessor hd can be 0 where the
                        // byte is kept (cur > 0)
                        const bool isl = cur == l, isu = cur == u;
                        uint32_t dgc = hd == 0.0f ? (uint32_t)(DB_MAIN | DB_STOP) : (uint32_t)DB_MAIN;
                        asm volatile("" : "+v"(dgc));
                        const uint32_t b = isl ? 3u : (isu ? 2u : dgc);
                        const bool pos = cur > 0.0f;
                        bits[r] = pos ? b : 0u;
                        td = Tp[r];
                        Tp[r] = cur;
                        tup = cur;
                    }
                }
            }
            if (DIRS) {
                // R nibbles per lane and step, tile-major: nibble
                // ((s * nsteps + t) * 64 + lane) * R + r, two per byte (nibble r
                // at bit 4r of the lane's word)
                // (byte address: 32R bytes per step, lane * R / 2 into it; for
                // R = 1 the even lane's byte)
                uint8_t *const dbyte = dir + (int64_t)(s * (int64_t)nsteps + t) * (WAVE / 2 * R) + ((lane * R) >> 1);
                uint32_t word = 0;
#pragma unroll
                for (int r = 0; r < R; ++r) word |= bits[r] << (4 * r);
                if (R == 1) {
                    // lanes 2k and 2k+1 share a byte: the even lane stores both
                    const uint32_t other = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)word, 0xB1, 0xf, 0xf, false);
                    if ((lane & 1) == 0) *dbyte = (uint8_t)(word | (other << 4));
                } else if (R == 2) {
                    *dbyte = (uint8_t)word;
                } else if (R == 6) {
                    dbyte[0] = (uint8_t)word;
                    dbyte[1] = (uint8_t)(word >> 8);
                    dbyte[2] = (uint8_t)(word >> 16);
                } else if (R == 4) {
                    *reinterpret_cast<uint16_t *>(dbyte) = (uint16_t)word;
                } else {
                    *reinterpret_cast<uint32_t *>(dbyte) = word;
                }
                // A candidate start is a main-matrix (diagonal) cell whose M
                // reaches the threshold (M >= 0, so "M >= thr or M <= -thr" is
                // M >= thr; for the linear form M is the cell value); then the
                // step's largest T reaches it too, which is the cheap test of
                // every step; the cells are tested only when it passes.
                float smax = Tp[0];
#pragma unroll
                for (int r = 1; r < R; ++r) smax = __builtin_fmaxf(smax, Tp[r]);
                rmax = __builtin_fmaxf(rmax, smax);
                if (__builtin_amdgcn_ballot_w64(smax >= thr && j >= 0 && j < B.len) != 0) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int x = s * rows + lane * R + r;
                        const float m = AFF ? Dp[r] : Tp[r];
                        const bool c = (bits[r] & 3u) == DB_MAIN && m >= thr && x < A.len && j >= 0 && j < B.len;
                        const uint64_t mask = __builtin_amdgcn_ballot_w64(c);
                        if (mask == 0) continue;
                        if (nbuf > CAND_BUF - WAVE) flush_cands();
                        if (c) {
                            const int pos = nbuf + (int)__builtin_amdgcn_mbcnt_hi(
                                                       (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                            Cand cd;
                            const float v = Tp[r];
                            cd.x = x | (v >= thr ? CAND_PLAIN : 0) | (v <= -thr ? CAND_MARKED : 0);
                            cd.y = j;
                            cd.v = v;
                            cbuf[pos] = cd;
                        }
                        nbuf += __builtin_popcountll(mask);
                    }
                }
            }

            // bottom row of this lane for column j
            const float botT = Tp[R - 1];
            const float botJ = AFF ? jin : 0.0f;
            const float botM = DIRS ? Dp[R - 1] : 0.0f;  // d of the bottom row: the next stripe's STOP test
            if (below) {
                // lane 63 at step t finishes column t - 63 of the bottom row; the
                // values move down one lane per step, so at the end of a
                // hand-over group lane L holds column c63 - 63 + L
                oT = shl1(botT, oT);
                if (AFF) oJ = shl1(botJ, oJ);
                if (DIRS && AFF) oM = shl1(botM, oM);
                const int c63 = t - (WAVE - 1);
                if (c63 == next_ho) {
                    const int c = c63 - (WAVE - 1) + lane;
                    if (c >= (c63 & ~(HANDOFF - 1))) {
                        store_scr(&sout[3 * c], oT);
                        if (AFF) store_scr(&sout[3 * c + 1], oJ);
                        if (DIRS && AFF) store_scr(&sout[3 * c + 2], oM);
                    }
                    if (links) publish_progress(prog + s, (unsigned)(c63 + 1));
                    next_ho = min(next_ho + HANDOFF, blen_it - 1);
                }
            }
            // inputs of step t+1
            dgT = upT;
            dgM = upM;
            upT = shr1(inj.x, botT);
            if (AFF) upJ = shr1(inj.y, botJ);
            if (DIRS && AFF) upM = shr1(inj.z, botM);
        };
        float scm[R];
        if constexpr (DIRS) {
            // (two steps per iteration here raise the direction pass's register
            // use past 200 VGPRs, so its steps copy the next scores instead)
            for (int t = tb0; t < te; ++t) {
                step(t, scn, scm);
#pragma unroll
                for (int r = 0; r < R; ++r) scn[r] = scm[r];
            }
            flush_cands();
        } else {
            // blocks of CKB steps when checkpointing (an even count, so that the
            // two score register sets keep their roles), else one block
            for (int b0 = 0; b0 < nsteps; b0 += (CK ? CKB : nsteps)) {
                const int b1 = CK ? min(b0 + CKB, nsteps) : nsteps;
                if (CK && b0 > 0) {
                    float *c = ckv.ck + (int64_t)(b0 / CKB) * (CK_SLOTS * WAVE) + lane;
#pragma unroll
                    for (int r = 0; r < R; ++r) c[r * WAVE] = Tp[r];
                    c[R * WAVE] = upT;
                    c[(R + 1) * WAVE] = dgT;
                }
                int t = b0;
                for (; t + 1 < b1; t += 2) {
                    step(t, scn, scm);
                    step(t + 1, scm, scn);
                }
                if (t < b1) step(t, scn, scm);
                if (CK) {
                    const float bm = wave_max(rmax);
                    if (lane == 0) ckv.bmax[b0 / CKB] = bm;
                    pm = __builtin_fmaxf(pm, rmax);
                    rmax = 0.0f;
                }
            }
        }
    }
    return CK ? __builtin_fmaxf(rmax, pm) : rmax;
}

// Loads the two transposed substitution tables (A = query, A = target) into LDS.
__device__ __forceinline__ void load_tables(float *lds, const float *__restrict__ g, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = g[i];
    __syncthreads();
}

template <bool AFF, int R, bool CK = false, bool PROF = false>
__device__ __forceinline__ float score_item(const uint8_t *res, const SeqRef *seqs, const ScoreItem &it,
                                            const float *lds_tab, int kp2, int lkp, uint8_t *bwin, float4 *stg,
                                            float *scr,
                                            float G, float E, int fill, int trans, float *links, unsigned *prog,
                                            unsigned *abort, unsigned wlimit, float *ckarea = nullptr,
                                            float *prof = nullptr, int nres = 0) {
    const float *tab = lds_tab + (trans ? kp2 : 0);
    float *lk = it.rows_off >= 0 ? links + it.rows_off : nullptr;
    unsigned *pg = it.rows_off >= 0 ? prog + it.prog_off : nullptr;
    CkArgs ckv;
    if (CK) {
        ckv.ck = ckarea + it.ck_off;
        ckv.bmax = ckarea + it.bmax_off;
        ckv.prof = prof;
        ckv.nres = nres;
    }
    return tile<AFF, R, false, CK, PROF>(res, seqs[it.a_seq], seqs[it.b_seq], it.alen, it.blen, tab, lkp, bwin, stg, scr,
                                   G, E, fill, nullptr, 0.0f, 0.0f, nullptr, nullptr, 0, it.s0, it.s1, lk, pg, abort,
                                   wlimit, ckv);
}

// Scoring pass.  queue != nullptr: persistent workers pull items from a
// device counter (longest first, as the host orders them; the stripes of a
// split pair are consecutive items in stripe order, so the wave a stripe waits
// for has always been dequeued before it); otherwise each workgroup takes
// items blockIdx.x, blockIdx.x + gridDim.x, ... (no split pairs then).  A
// split pair's maximum is the maximum over its stripes, combined with an
// integer atomic max on the (non-negative) float bits into a zeroed slot.
// The stripes of a checkpointed pair also write their checkpoints and block
// maxima into ckarea.  nres > 0: they read their scores from a per-lane
// profile of nres + 1 slots in LDS after the hand-over staging (the host
// passes nres, the number of residue codes, only when the launch's LDS holds
// that profile).
template <bool AFF>
__global__ void __launch_bounds__(64) score_kernel(const uint8_t *__restrict__ res, const SeqRef *__restrict__ seqs,
                                                   const ScoreItem *__restrict__ items, int nitems,
                                                   const uint8_t *__restrict__ trans_flags,
                                                   const float *__restrict__ tables, int kp, int lkp, float G,
                                                   float E, float *__restrict__ scratch, int64_t scr_stride,
                                                   unsigned *__restrict__ queue, float *__restrict__ maxima,
                                                   float *__restrict__ links, unsigned *__restrict__ prog,
                                                   unsigned *abort, unsigned wlimit, float *__restrict__ ckarea,
                                                   int nres) {
    extern __shared__ float lds[];
    const int kp2 = kp * kp;
    load_tables(lds, tables, 2 * kp2);
    uint8_t *bwin = reinterpret_cast<uint8_t *>(lds + 2 * kp2);
    float4 *stg = reinterpret_cast<float4 *>(bwin + BWIN + WAVE);
    float *prof = reinterpret_cast<float *>(stg + 2 * WAVE);
    float *scr = scratch + (int64_t)blockIdx.x * scr_stride;
    const int fill = kp - 1;
    int k = 0;
    for (;;) {
        int item;
        if (queue) {
            unsigned v = 0;
            if (threadIdx.x == 0) v = atomicAdd(queue, 1u);
            item = (int)__builtin_amdgcn_readfirstlane(v);
        } else {
            item = (int)blockIdx.x + k * (int)gridDim.x;
            ++k;
        }
        if (item >= nitems) break;
        const ScoreItem it = items[item];
        const int tr = trans_flags ? trans_flags[item] : 0;
        float m = 0.0f;
        bool done = false;
        if constexpr (!AFF) {
            if (it.ck_off >= 0) {
                if (nres > 0)
                    m = score_item<AFF, CK_R, true, true>(res, seqs, it, lds, kp2, lkp, bwin, stg, scr, G, E, fill, tr,
                                                          links, prog, abort, wlimit, ckarea, prof, nres);
                else
                    m = score_item<AFF, CK_R, true>(res, seqs, it, lds, kp2, lkp, bwin, stg, scr, G, E, fill, tr,
                                                    links, prog, abort, wlimit, ckarea);
                done = true;
            }
        }
        if (!done) {
            switch (it.R) {
            case 1: m = score_item<AFF, 1>(res, seqs, it, lds, kp2, lkp, bwin, stg, scr, G, E, fill, tr, links, prog, abort, wlimit); break;
            case 2: m = score_item<AFF, 2>(res, seqs, it, lds, kp2, lkp, bwin, stg, scr, G, E, fill, tr, links, prog, abort, wlimit); break;
            case 4: m = score_item<AFF, 4>(res, seqs, it, lds, kp2, lkp, bwin, stg, scr, G, E, fill, tr, links, prog, abort, wlimit); break;
            default: m = score_item<AFF, 8>(res, seqs, it, lds, kp2, lkp, bwin, stg, scr, G, E, fill, tr, links, prog, abort, wlimit); break;
            }
        }
        m = wave_max(m);
        if (threadIdx.x == 0) {
            if (it.rows_off >= 0) atomicMax(reinterpret_cast<int *>(maxima) + it.out, __float_as_int(m));
            else maxima[it.out] = m;
        }
    }
}

// ---------------------------------------------------------------------------
// Lane-per-pair scoring.  One job = one column sequence C, shared by up to 64
// pairs, and one row sequence per lane.  The wave builds a substitution
// profile of C in LDS (one row of NB columns per residue code); each lane
// sweeps its own row sequence top to bottom with the NB column states of the
// previous row in registers, so there is no skew and no lane exchange.
// C longer than NB is processed in chunks of NB columns; the right edge of a
// chunk (T and the I candidate of every row) goes through a per-wave scratch
// column, row-major with the 64 lanes contiguous.
// Recurrences per cell (r, c), row = lane sequence, column = C:
//   M = max(T[r-1][c-1] + s, 0)           T = max(M, I, J)
//   I = max(I[r][c-1] + E, M[r][c-1] + G) J = max(J[r-1][c] + E, M[r-1][c] + G)
// (linear: H = max(H[r-1][c-1] + s, max(H[r-1][c], H[r][c-1]) + G, 0)).
// With G, E <= 0 (pw_set_scoring rejects positive gap scores) a negative gap
// value never reaches T, and M[r][c] + G <= 0 whenever M is clamped.  So the
// kernel feeds the unclamped d = T[r-1][c-1] + s into the gap candidates and
// clamps J at 0 instead: I' = max(I' + E, d + G) <= I and
// J' = max(J' + E, d + G, 0) = max(J, 0) keep the positive parts of I and J
// (by induction over the row and column), and T = max(d, I', J') is the
// reference's T exactly, with one instruction fewer per cell.
// Row-shifted form (SH, integral scores only, see group_shift_exact on the
// host): with sigma_r = -r * E >= 0 the kernel carries T^ = T + sigma_r,
// J^ = J' + sigma_r and I^ = I' + sigma_r - E, and its profile holds s - E.
// Then, with d^ = T^[r-1][c-1] + (s - E) = d + sigma_r and
// m = d^ + (G - E),
//   P = I^ + E,  T^ = max(d^, P, J^),  I^ <- max(P, m),
//   J^ <- max(J^, m, sigma_{r+1})
// which is the recurrence above shifted by sigma_r, one add fewer per cell
// (E leaves the J update and the I add is shared with T).  Integral scores
// keep every sum exact, so the row maximum of T^ minus sigma_r is the row
// maximum of T bit for bit.  The boundary row above holds T^ = E (T = 0,
// sigma_-1 = E); the left boundary of row r holds T^ = sigma_r and
// I^ = G - E + sigma_r.
// ---------------------------------------------------------------------------
template <bool AFF, int NB, bool SH>
__device__ float group_tile(const uint8_t *__restrict__ res, const SeqRef *__restrict__ seqs, const GroupItem &it,
                            const GroupEntry *__restrict__ ents, const float *__restrict__ gtab, int kp,
                            float *prof, float2 *__restrict__ scr, float2 *__restrict__ links,
                            unsigned *__restrict__ prog, float G, float E, int fill, unsigned *abort,
                            unsigned wlimit) {
    constexpr int NBP = NB + 4;
    const int lane = threadIdx.x & 63;
    const SeqRef C = seqs[it.c_seq];
    const bool active = lane < it.count;
    SeqRef Rw;
    Rw.off = 0;
    Rw.len = 0;
    if (active) Rw = seqs[ents[it.entry_off + lane].row_seq];
    const int nchunks = C.len > NB ? (C.len + NB - 1) / NB : 1;
    const float *tab = gtab + (it.trans ? kp * kp : 0);
    static_assert(AFF || !SH, "the row shift applies to affine gaps");
    const float bI = AFF ? G : 0.0f;
    // wave-uniform values kept in scalar registers (readfirstlane), so that
    // the shifted form costs no vector registers
    auto uniform = [](float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); };
    const float GmE = uniform(G - E);
    const float t0 = SH ? E : 0.0f;  // T (T^) of the boundary row above the first row
    float rmax = 0.0f;
    const int k_end = min(it.k1, nchunks);
    for (int k = it.k0; k < k_end; ++k) {
        const int c0 = k * NB;
        // the left edge comes from the previous item's link column (published
        // row by row) for the first chunk of a split item, else from the
        // wave's own scratch column; likewise the right edge goes out
        const bool from_link = k > 0 && k == it.k0 && it.link_in >= 0;
        const bool to_link = k + 1 < nchunks && k == k_end - 1 && it.link_out >= 0;
        const float2 *src = from_link ? links + it.link_in : scr;
        float2 *dst = to_link ? links + it.link_out : scr;
        unsigned avail = from_link ? 0u : 0x7fffffffu;  // rows of src known to be published
        __syncthreads();
        // profile row of residue a, shifted by one column: entry p < NB holds
        // the score of column c0+p+1, entry NB the score of column c0; so the
        // 4-column read of block b supplies columns b+1..b+4, i.e. three of
        // the block's own diagonal sums and the first one of the next block
        for (int idx = lane; idx < kp * (NB + 1); idx += WAVE) {
            const int a = idx / (NB + 1), p = idx - a * (NB + 1);
            const int col = c0 + (p < NB ? p + 1 : 0);
            const int b = col < C.len ? (int)res[C.off + col] : fill;
            prof[a * NBP + p] = SH ? tab[b * kp + a] - E : tab[b * kp + a];
        }
        __syncthreads();
        const bool left_in = k > 0, right_out = k + 1 < nchunks;
        float Tp[NB], Jc[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            Tp[c] = t0;
            Jc[c] = 0.0f;
        }
        float dl = t0;  // T[r-1][c0-1]
        int a_next = (active && 0 < Rw.len) ? (int)res[Rw.off] : fill;
        float2 lb_next = make_float2(0.0f, SH ? GmE : bI);
        if (left_in && it.rows > 0) {
            if (from_link) avail = wait_rows(prog + it.prog_in, 1u, abort, wlimit);
            lb_next.x = load_scr(&src[lane].x);
            lb_next.y = load_scr(&src[lane].y);
        }
        // profile values of the first row: columns 1..4 and column 0
        const float *prow = prof + a_next * NBP;
        float4 s4n = *reinterpret_cast<const float4 *>(prow);
        float s0n = prow[NB];
        float sig1 = uniform(-E);  // sigma_{r+1} (SH)
#pragma unroll 1
        for (int r = 0; r < it.rows; ++r) {
            const float2 lb = lb_next;
            a_next = (active && r + 1 < Rw.len) ? (int)res[Rw.off + r + 1] : fill;
            if (left_in && r + 1 < it.rows) {
                if ((unsigned)(r + 2) > avail) avail = wait_rows(prog + it.prog_in, (unsigned)(r + 2), abort, wlimit);
                lb_next.x = load_scr(&src[(r + 1) * WAVE + lane].x);
                lb_next.y = load_scr(&src[(r + 1) * WAVE + lane].y);
            } else if (SH) {
                lb_next = make_float2(sig1, GmE + sig1);  // left boundary of row r+1
            }
            float ic = lb.y;  // affine: I candidate entering the column; linear: H[r][c-1]
            // diagonal sum of the block's first column, T[r-1][c-1] + s; each
            // block forms the next block's from the profile values it already
            // holds, before its last column overwrites T[r-1][c-1]
            float dgc = dl + s0n;
            // SH: the running maximum in the row's shifted terms, rmax + sigma_r
            if (SH) rmax = rmax + (sig1 + E);
#pragma unroll
            for (int b = 0; b < NB; b += 4) {
                // the next block's profile values (the next row's first block
                // after the last block) are read one block ahead, so that the
                // LDS latency overlaps this block's arithmetic
                const float4 s4 = s4n;
                if (b + 4 < NB) {
                    s4n = *reinterpret_cast<const float4 *>(prow + b + 4);
                } else {
                    // the next row's first block and column 0
                    prow = prof + a_next * NBP;
                    s4n = *reinterpret_cast<const float4 *>(prow);
                    s0n = prow[NB];
                }
                const float dg0 = dgc;
                const float dg1 = Tp[b] + s4.x;
                const float dg2 = Tp[b + 1] + s4.y;
                const float dg3 = Tp[b + 2] + s4.z;
                if (b + 4 < NB) dgc = Tp[b + 3] + s4.w;
                const float dg[4] = {dg0, dg1, dg2, dg3};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int c = b + u;
                    if (AFF && SH) {
                        const float m = dg[u] + GmE;
                        const float P = ic + E;
                        const float T = fmax3(dg[u], P, Jc[c]);
                        Jc[c] = fmax3(Jc[c], m, sig1);
                        ic = __builtin_fmaxf(P, m);
                        Tp[c] = T;
                    } else if (AFF) {
                        const float MG = dg[u] + G;
                        const float T = fmax3(dg[u], ic, Jc[c]);
                        Jc[c] = fmax3(Jc[c] + E, MG, 0.0f);
                        ic = __builtin_fmaxf(ic + E, MG);
                        Tp[c] = T;
                    } else {
                        const float h = fmax3(dg[u], __builtin_fmaxf(Tp[c], ic) + G, 0.0f);
                        Tp[c] = h;
                        ic = h;
                    }
                }
                // Materialise this block's column state here: without it the
                // compiler sinks the T and J updates to the end of the row and
                // keeps ~4 registers per column alive.  The empty asm emits no
                // instruction; it only pins the values in registers.
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (AFF) asm volatile("" : "+v"(Tp[b + u]), "+v"(Jc[b + u]));
                    else asm volatile("" : "+v"(Tp[b + u]));
                }
                rmax = fmax3(rmax, fmax3(Tp[b], Tp[b + 1], Tp[b + 2]), Tp[b + 3]);
                asm volatile("" : "+v"(rmax), "+v"(ic));
            }
            if (SH) rmax = rmax - (sig1 + E);
            if (right_out) {
                store_scr(&dst[r * WAVE + lane].x, Tp[NB - 1]);
                store_scr(&dst[r * WAVE + lane].y, ic);
                if (to_link && (((r + 1) % GROUP_PUB) == 0 || r + 1 == it.rows))
                    publish_progress(prog + it.prog_out, (unsigned)(r + 1));
            }
            dl = lb.x;
            if (SH) sig1 = uniform(sig1 - E);
        }
    }
    return rmax;
}

// One kernel per chunk width, so that each width gets its own register
// budget; the host launches one per width present, over that width's jobs.
// Width 72 is held to 168 VGPRs (3 waves per SIMD instead of 2; no spill).
template <bool AFF, int NB, bool SH>
__global__ void __launch_bounds__(64, NB == 72 ? 3 : 1) group_kernel(const uint8_t *__restrict__ res, const SeqRef *__restrict__ seqs,
                                                   const GroupItem *__restrict__ items, int nitems,
                                                   const GroupEntry *__restrict__ ents,
                                                   const float *__restrict__ tables, int kp, float G, float E,
                                                   float2 *__restrict__ scratch, int64_t scr_stride,
                                                   unsigned *__restrict__ queue, float *__restrict__ maxima,
                                                   float2 *__restrict__ links, unsigned *__restrict__ prog,
                                                   unsigned *abort, unsigned wlimit) {
    extern __shared__ float lds[];
    float2 *scr = scratch + (int64_t)blockIdx.x * scr_stride;
    const int fill = kp - 1;
    int k = 0;
    for (;;) {
        int item;
        if (queue) {
            unsigned v = 0;
            if (threadIdx.x == 0) v = atomicAdd(queue, 1u);
            item = (int)__builtin_amdgcn_readfirstlane(v);
        } else {
            item = (int)blockIdx.x + k * (int)gridDim.x;
            ++k;
        }
        if (item >= nitems) break;
        const GroupItem it = items[item];
        const float m =
            group_tile<AFF, NB, SH>(res, seqs, it, ents, tables, kp, lds, scr, links, prog, G, E, fill, abort, wlimit);
        if ((int)threadIdx.x < it.count) {
            const int pair = ents[it.entry_off + threadIdx.x].pair;
            // a split job's maximum: the maximum over its items, combined with
            // an integer atomic max on the (non-negative) float bits into a
            // slot zeroed at the start of the execution
            if (it.split) atomicMax(reinterpret_cast<int *>(maxima) + pair, __float_as_int(m));
            else maxima[pair] = m;
        }
    }
}

// Number of traceback tasks of the current execution (written by the gate
// kernel or by the host before the kernel that reads it).
__device__ __forceinline__ int task_count(const Ctl *ctl) {
    return __hip_atomic_load(&ctl->ngated, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Direction pass over the work units of the traceback tasks (one unit per
// task, or one per stripe of a split task, consecutive in stripe order, or
// one per needed block of a checkpointed task, in any order); persistent
// workgroups pull units from a device counter up to the unit count.
// Held to four waves per SIMD (128 VGPRs, a few spilled outside the step
// loop): the affine form would take 140 and run three; its step is bound by
// instruction issue, so the fourth wave is ~12% more throughput (measured on
// the 2,000 x 20,000 protein set: direction pass 402 -> 370 ms).
template <bool AFF>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) dir_kernel(const uint8_t *__restrict__ res, const SeqRef *__restrict__ seqs,
                                                 PairTask *__restrict__ tasks, const DirUnit *__restrict__ units,
                                                 Ctl *__restrict__ ctl, const float *__restrict__ tables, int kp,
                                                 int lkp, float G, float E, float *__restrict__ scratch,
                                                 int64_t scr_stride, uint8_t *__restrict__ dirs,
                                                 Cand *__restrict__ cands, float *__restrict__ links,
                                                 unsigned *__restrict__ prog, int64_t units_cap, unsigned wlimit,
                                                 const CkPair *__restrict__ ckpairs, float *__restrict__ ckarea,
                                                 float *__restrict__ score_links) {
    extern __shared__ float lds[];
    const int kp2 = kp * kp;
    // only the table with the query along the lanes (the first): 4 KB less LDS
    // per wave for protein alphabets, four waves per SIMD instead of three
    load_tables(lds, tables, kp2);
    uint8_t *bwin = reinterpret_cast<uint8_t *>(lds + kp2);
    float4 *stg = reinterpret_cast<float4 *>(bwin + BWIN + WAVE);
    float *scr = scratch + (int64_t)blockIdx.x * scr_stride;
    const int fill = kp - 1;
    const int n = (int)min((int64_t)__hip_atomic_load(&ctl->units_used, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                           units_cap);
    unsigned *abort = &ctl->abort_wait;
    for (;;) {
        unsigned v = 0;
        if (threadIdx.x == 0) v = atomicAdd(&ctl->dir_queue, 1u);
        const int u = (int)__builtin_amdgcn_readfirstlane(v);
        if (u >= n) break;
        const DirUnit du = units[u];
        const PairTask it = tasks[du.task];
        if (it.status != TASK_OK) continue;
        const SeqRef q = seqs[it.q_seq], t = seqs[it.t_seq];
        uint8_t *d = dirs + it.dir_off;
        int *cc = &tasks[du.task].ncand;
        Cand *cl = cands + it.cand_off;
        const int s0 = du.stripe < 0 ? 0 : du.stripe;
        const int s1 = du.stripe < 0 ? (1 << 30) : du.stripe + 1;
        float *lk = du.stripe < 0 ? nullptr : links + it.rows_off;
        unsigned *pg = du.stripe < 0 ? nullptr : prog + it.prog_off;
        CkArgs ckv;
        if (!AFF && du.block >= 0 && it.ck >= 0 && it.R == CK_R) {
            // one block of a checkpointed pair: the scoring pass's checkpoint
            // and boundary row above, no waits (the same instance of the tile
            // as the pair's whole-stripe form, so that the kernel keeps one
            // register budget)
            const CkPair cp = ckpairs[it.ck];
            ckv.t0 = du.block * CKB;
            ckv.t1 = ckv.t0 + CKB;
            ckv.ck = du.block > 0
                         ? ckarea + cp.ck_off + ((int64_t)du.stripe * cp.nblocks + du.block) * (CK_SLOTS * WAVE)
                         : nullptr;
            lk = score_links + cp.rows_off;
            pg = nullptr;
        }
        switch (it.R) {
        case 1: tile<AFF, 1, true>(res, q, t, q.len, t.len, lds, lkp, bwin, stg, scr, G, E, fill, d, it.t1, it.mp, cc, cl, it.cand_cap, s0, s1, lk, pg, abort, wlimit); break;
        case 2: tile<AFF, 2, true>(res, q, t, q.len, t.len, lds, lkp, bwin, stg, scr, G, E, fill, d, it.t1, it.mp, cc, cl, it.cand_cap, s0, s1, lk, pg, abort, wlimit); break;
        case 4: tile<AFF, 4, true>(res, q, t, q.len, t.len, lds, lkp, bwin, stg, scr, G, E, fill, d, it.t1, it.mp, cc, cl, it.cand_cap, s0, s1, lk, pg, abort, wlimit, ck_for<4>(ckv)); break;
        case 6: tile<AFF, 6, true>(res, q, t, q.len, t.len, lds, lkp, bwin, stg, scr, G, E, fill, d, it.t1, it.mp, cc, cl, it.cand_cap, s0, s1, lk, pg, abort, wlimit); break;
        default: tile<AFF, 8, true>(res, q, t, q.len, t.len, lds, lkp, bwin, stg, scr, G, E, fill, d, it.t1, it.mp, cc, cl, it.cand_cap, s0, s1, lk, pg, abort, wlimit, ck_for<8>(ckv)); break;
        }
    }
}

// Clears n 32-bit words (control blocks), in place of a memset node so that
// the execution can be captured in a graph.
__global__ void __launch_bounds__(256) clear_kernel(uint32_t *__restrict__ p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Query residues per lane of the direction pass: the smallest power of two
// whose 64-lane stripe covers the query, at most rmax (a plan lowers rmax when
// its pairs would otherwise give too few stripe units to fill the device),
// except that a query of 257-384 residues takes 6 (one stripe of 384 rows
// instead of 512: a quarter fewer padded rows).
__device__ __host__ __forceinline__ int dir_residues(int qlen, int rmax = 8) {
    int r = 1;
    while (r < rmax && WAVE * r < qlen) r *= 2;
    if (r == 8 && qlen <= WAVE * 6) r = 6;
    return r;
}

// Residues per lane whose direction area bounds that of every rmax above (an
// arena reservation made before the plan's rmax is known).
__device__ __host__ __forceinline__ int dir_residues_bound(int qlen) {
    int r = 1;
    while (r < 8 && WAVE * r < qlen) r *= 2;
    return r;
}

// Bytes of a pair's tile-major direction area (16-byte aligned).
__device__ __host__ __forceinline__ int64_t dir_area(int qlen, int tlen, int R) {
    const int rows = WAVE * R;
    const int64_t stripes = qlen > 0 ? (qlen + rows - 1) / rows : 1;
    // one nibble per cell (64R, an even count, per step), 16-byte multiples
    return (stripes * (int64_t)(tlen + WAVE - 1) * WAVE * R / 2 + 15) & ~(int64_t)15;
}

// A pair is split into one work unit per stripe when it has several stripes
// and enough columns for the stripes to overlap.
constexpr int SPLIT_MIN_COLS = 256;
__device__ __host__ __forceinline__ int split_stripes(int alen, int blen, int R) {
    const int rows = WAVE * R;
    const int ns = alen > 0 ? (alen + rows - 1) / rows : 1;
    return (ns >= 2 && blen >= SPLIT_MIN_COLS) ? ns : 0;
}

// Exclusive prefix sum over a 256-thread block of v, and the block total (lds:
// four entries; every thread of the block calls it).
template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T &total, T *lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    T before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const T t = lds[w];
        if (w < wave) before += t;
        total += t;
    }
    __syncthreads();
    return before + x - v;
}

// Called by every thread of each block of a grid whose blocks stored their
// sums (device-scope atomic stores) before: true in the block that finished
// last, which then sees every block's sum.
__device__ __forceinline__ bool last_block(unsigned *done) {
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

// Exclusive prefix sums, in place, of the n sums at a[0], a[stride], ... (one
// 256-thread block, device-scope atomic accesses); returns the total.
template <class T>
__device__ T scan_sums(T *a, int n, int stride, T *lds) {
    T run = 0;
    for (int b0 = 0; b0 < n; b0 += 256) {
        const int b = b0 + threadIdx.x;
        const T v = b < n ? __hip_atomic_load(&a[(int64_t)b * stride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        T tot;
        const T ex = block_excl_scan<T>(v, tot, lds);
        if (b < n) __hip_atomic_store(&a[(int64_t)b * stride], run + ex, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        run += tot;
    }
    return run;
}

// Traceback gate of default_main_cpu.cl:408-410 (maximum >= MINIMUM_SCORE)
// plus the pair's minimum start score, without which no cell can start
// (:427): the indices of the passing pairs are compacted into `gated` in
// ascending order, so that the read-back slots of an execution are in pair
// order.  Block b of both passes takes the contiguous range [b*per, (b+1)*per).
__device__ __forceinline__ bool gate_pass(const float *__restrict__ maxima, const float *__restrict__ mss, int64_t i,
                                          float minimum) {
    const float m = maxima[i];
    return m >= minimum && m >= mss[i];
}

// Pass 1: each block counts its passing pairs into bsum[b]; the block that
// finishes last turns the counts into exclusive offsets and sets ctl->ngated.
__global__ void __launch_bounds__(256) gate_count_kernel(const float *__restrict__ maxima,
                                                         const float *__restrict__ mss, int64_t n, float minimum,
                                                         Ctl *__restrict__ ctl, int *__restrict__ bsum) {
    __shared__ int lds[4];
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per, hi = min(n, lo + per);
    int cnt = 0;
    for (int64_t i0 = lo; i0 < hi; i0 += blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        cnt += (i < hi && gate_pass(maxima, mss, i, minimum)) ? 1 : 0;
    }
    int tot;
    (void)block_excl_scan<int>(cnt, tot, lds);
    if (threadIdx.x == 0) __hip_atomic_store(&bsum[blockIdx.x], tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!last_block(&ctl->gate_done)) return;
    const int all = scan_sums<int>(bsum, gridDim.x, 1, lds);
    if (threadIdx.x == 0) __hip_atomic_store(&ctl->ngated, all, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pass 2: each block writes its passing pairs in order from its offset.
__global__ void __launch_bounds__(256) gate_scatter_kernel(const float *__restrict__ maxima,
                                                           const float *__restrict__ mss, int64_t n, float minimum,
                                                           const int *__restrict__ boff, int *__restrict__ gated) {
    __shared__ int lds[4];
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per, hi = min(n, lo + per);
    int run = __hip_atomic_load(&boff[blockIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int64_t i0 = lo; i0 < hi; i0 += blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        const bool pass = i < hi && gate_pass(maxima, mss, i, minimum);
        int tot;
        const int k = block_excl_scan<int>(pass ? 1 : 0, tot, lds);
        if (pass) gated[run + k] = (int)i;
        run += tot;
    }
}

// ---------------------------------------------------------------------------
// Result layout on the device: every pair's records, op characters and the
// three alignment rows of its shown hits, in pair order (the read-back slots
// are in pair order), at the offsets of exclusive sums over the slots.
// ---------------------------------------------------------------------------

// Pass 1: per-block sums of the slots' record and op counts (bsum[2b],
// bsum[2b + 1]); the block finishing last scans them and stores the totals in
// ctl->lay_sps / ctl->lay_ops.
__global__ void __launch_bounds__(256) layout_sums_kernel(const TaskOut *__restrict__ touts, int ng,
                                                          Ctl *__restrict__ ctl, long long *__restrict__ bsum) {
    __shared__ long long lds[4];
    const int per = (ng + gridDim.x - 1) / gridDim.x;
    const int lo = blockIdx.x * per, hi = min(ng, lo + per);
    long long ns = 0, no = 0;
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        ns += touts[i].nsp;
        no += touts[i].nops;
    }
    long long ts, to;
    (void)block_excl_scan<long long>(ns, ts, lds);
    (void)block_excl_scan<long long>(no, to, lds);
    if (threadIdx.x == 0) {
        __hip_atomic_store(&bsum[2 * blockIdx.x], ts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&bsum[2 * blockIdx.x + 1], to, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!last_block(&ctl->lay_done)) return;
    const long long as = scan_sums<long long>(bsum, gridDim.x, 2, lds);
    const long long ao = scan_sums<long long>(bsum + 1, gridDim.x, 2, lds);
    if (threadIdx.x == 0) {
        __hip_atomic_store(&ctl->lay_sps, as, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->lay_ops, ao, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Pass 2: each slot's first record and first op character in the result.
__global__ void __launch_bounds__(256) layout_offsets_kernel(const TaskOut *__restrict__ touts, int ng,
                                                             const long long *__restrict__ bsum,
                                                             long long *__restrict__ sp_off,
                                                             long long *__restrict__ op_off) {
    __shared__ long long lds[4];
    const int per = (ng + gridDim.x - 1) / gridDim.x;
    const int lo = blockIdx.x * per, hi = min(ng, lo + per);
    long long rs = __hip_atomic_load(&bsum[2 * blockIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    long long ro = __hip_atomic_load(&bsum[2 * blockIdx.x + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int i0 = lo; i0 < hi; i0 += blockDim.x) {
        const int i = i0 + threadIdx.x;
        const bool on = i < hi;
        long long ts, to;
        const long long es = block_excl_scan<long long>(on ? (long long)touts[i].nsp : 0ll, ts, lds);
        const long long eo = block_excl_scan<long long>(on ? (long long)touts[i].nops : 0ll, to, lds);
        if (on) {
            sp_off[i] = rs + es;
            op_off[i] = ro + eo;
        }
        rs += ts;
        ro += to;
    }
}

// Pass 3, one wave per slot: starting points and op characters copied, hits
// with their indices moved to the result's, and each shown hit's query,
// alignment and target rows (SmithWaterman.py:777-830) written start to end
// at three times its op offset.  The ops of a hit run from its end back; a
// column's query (target) position is the hit's start plus the columns
// before it that consume a query (target) residue: a ballot prefix count
// per 64 columns.  ctab: the residue character of each code, then its lower
// case (the match test compares lower-cased, :794-799).
__global__ void __launch_bounds__(256) layout_gather_kernel(
    const uint8_t *__restrict__ res, const SeqRef *__restrict__ seqs, int nq, const TaskOut *__restrict__ touts,
    int ng, const long long *__restrict__ sp_off, const long long *__restrict__ op_off,
    const SpRec *__restrict__ out_sps, const HitRec *__restrict__ out_hits, const char *__restrict__ ops,
    const uint8_t *__restrict__ ctab, SpRec *__restrict__ dsp, HitRec *__restrict__ dhit, char *__restrict__ dops,
    char *__restrict__ daln) {
    __shared__ uint8_t tab[512];
    for (int i = threadIdx.x; i < 512; i += blockDim.x) tab[i] = ctab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < ng; s += waves) {
        const TaskOut t = touts[s];
        const long long sp0 = sp_off[s], op0 = op_off[s];
        const uint32_t *ssrc = reinterpret_cast<const uint32_t *>(out_sps + t.out_off);
        uint32_t *sdst = reinterpret_cast<uint32_t *>(dsp + sp0);
        constexpr int SPW = (int)(sizeof(SpRec) / 4);
        for (int w = lane; w < t.nsp * SPW; w += WAVE) sdst[w] = ssrc[w];
        for (int k = lane; k < t.nsp; k += WAVE) {
            HitRec h = out_hits[t.out_off + k];
            h.sp_index = (int32_t)(sp0 + h.sp_index);
            h.ops_offset = op0 + h.ops_offset;
            dhit[sp0 + k] = h;
        }
        const char *osrc = ops + t.ops_off;
        char *odst = dops + op0;
        for (int64_t i = lane; i < t.nops; i += WAVE) odst[i] = osrc[i];
        for (int k = 0; k < t.nsp; ++k) {
            const HitRec h = out_hits[t.out_off + k];
            if (!h.shown || h.length <= 0) continue;
            const uint8_t *qc = res + seqs[h.query].off, *tc = res + seqs[nq + h.target].off;
            const char *op = osrc + h.ops_offset + h.length - 1;
            char *qrow = daln + 3 * (op0 + h.ops_offset), *arow = qrow + h.length, *trow = arow + h.length;
            int x = h.s_start, y = h.t_start;
            for (int c0 = 0; c0 < h.length; c0 += WAVE) {
                const int c = c0 + lane;
                const bool on = c < h.length;
                const char v = on ? op[-c] : 'D';
                const uint64_t bq = __ballot(on && v != 'U'), bt = __ballot(on && v != 'L');
                const int xi = x + __popcll(bq & below), yi = y + __popcll(bt & below);
                if (on) {
                    uint8_t qch = '-', tch = '-', ach = ' ';
                    if (v == 'D') {
                        const uint8_t a = qc[xi], b = tc[yi];
                        qch = tab[a];
                        tch = tab[b];
                        ach = tab[256 + a] == tab[256 + b] ? '|' : '.';
                    } else if (v == 'L') {
                        qch = tab[qc[xi]];
                    } else {
                        tch = tab[tc[yi]];
                    }
                    qrow[c] = (char)qch;
                    arow[c] = (char)ach;
                    trow[c] = (char)tch;
                }
                x += __popcll(bq);
                y += __popcll(bt);
            }
        }
    }
}

// Exclusive prefix sum over the wave of a 64-bit amount, and the wave total.
__device__ __forceinline__ uint64_t wave_excl_scan(uint64_t v, uint64_t &total) {
    const int lane = threadIdx.x & 63;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// One amount per lane from a shared counter: one atomic per wave.
__device__ __forceinline__ unsigned long long wave_alloc(unsigned long long *counter, uint64_t amount) {
    uint64_t total = 0;
    const uint64_t off = wave_excl_scan(amount, total);
    unsigned long long b = 0;
    if ((threadIdx.x & 63) == 0 && total) b = atomicAdd(counter, (unsigned long long)total);
    b = __shfl(b, 0, 64);
    return b + off;
}

// Direction-pass blocks a checkpointed pair needs: a path starts at a candidate
// (a cell whose value reaches thr, so its block's maximum does too) and moves
// only up and left, so stripe s needs the columns up to ycol(s), the last
// column of any block of stripe s or below whose maximum reaches thr, i.e.
// its blocks up to the one that holds step ycol(s) + 63 (lane 63's cell of
// that column).  Counts the units, and writes them to out when given.
__device__ int ck_units(const CkPair &cp, const float *__restrict__ ckarea, int ns, float thr,
                        DirUnit *__restrict__ out, int task) {
    const float *bm = ckarea + cp.bmax_off;
    const int nb = cp.nblocks;
    int ycol = -1, n = 0;
    for (int s = ns - 1; s >= 0; --s) {
        for (int b = nb - 1; b >= 0 && (b + 1) * CKB - 1 > ycol; --b)
            if (bm[(int64_t)s * nb + b] >= thr) {
                ycol = (b + 1) * CKB - 1;
                break;
            }
        if (ycol < 0) continue;
        const int k = min(nb, (ycol + WAVE - 1) / CKB + 1);
        if (out)
            for (int b = 0; b < k; ++b) out[n + b] = DirUnit{task, s, b};
        n += k;
    }
    return n;
}

// Index of pair i among the checkpointed pairs (sorted by pair), or -1.
__device__ __forceinline__ int ck_find(const CkPair *__restrict__ ckp, int nck, int i) {
    int lo = 0, hi = nck;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ckp[mid].pair < i) lo = mid + 1;
        else hi = mid;
    }
    return (lo < nck && ckp[lo].pair == i) ? lo : -1;
}

// Builds the task of every gated pair (slot = its position in `gated`), with
// its direction and candidate areas taken from the bump allocators in ctl,
// and its direction-pass work units (one per stripe for a split pair, whose
// boundary rows and zeroed progress counters are allocated here too; without
// room it runs unsplit; for a checkpointed pair, one per block it needs).  A
// pair that does not fit is marked TASK_DEFERRED for a later round.  With
// cross_nt > 0 the pairs are every query with every target, query-major, and
// pairs is not read.  A requeued round (requeue_kernel) passes each pair's
// read-back slot and candidate need in slots and needs; the first round takes
// slot g and cand0 candidate slots per pair.
// Allocations are summed over the wave: one atomic per wave and counter.
__global__ void __launch_bounds__(256) task_kernel(const float *__restrict__ maxima, const float *__restrict__ mss,
                                                   const int2 *__restrict__ pairs,
                                                   const uint8_t *__restrict__ same_id,
                                                   const SeqRef *__restrict__ seqs, int nq, float ll,
                                                   Ctl *__restrict__ ctl, const int *__restrict__ gated,
                                                   PairTask *__restrict__ tasks, int64_t dir_cap,
                                                   int64_t cand_cap_total, int cand0, DirUnit *__restrict__ units,
                                                   int64_t units_cap, int64_t link_cap,
                                                   unsigned *__restrict__ prog, int64_t prog_cap, int rmax,
                                                   const CkPair *__restrict__ ckpairs, int nck,
                                                   const float *__restrict__ ckarea, int cross_nt,
                                                   const int *__restrict__ slots, const int *__restrict__ needs) {
    const int ng = __hip_atomic_load(&ctl->ngated, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int stride = gridDim.x * blockDim.x;
    // whole waves iterate together (the scans need every lane)
    for (int g0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63); g0 < ng; g0 += stride) {
        const int g = g0 + (threadIdx.x & 63);
        const bool on = g < ng;
        PairTask t;
        int ql = 0, tl = 0, ns = 0, ck = -1, ckn = 0;
        int64_t area = 0;
        float m = 0.0f;
        if (on) {
            const int i = gated[g];
            // the all-vs-all batch has no pair list: pair i is query i / nt, target i % nt
            const int2 pr = cross_nt > 0 ? make_int2(i / cross_nt, i % cross_nt) : pairs[i];
            ql = seqs[pr.x].len;
            tl = seqs[nq + pr.y].len;
            m = maxima[i];
            t.pair = i;
            t.q_seq = pr.x;
            t.t_seq = nq + pr.y;
            if (nck > 0 && ql >= CK_MIN_LEN && tl >= CK_MIN_LEN) ck = ck_find(ckpairs, nck, i);
            t.R = ck >= 0 ? CK_R : dir_residues(ql, rmax);
            area = dir_area(ql, tl, t.R);
            ns = ck >= 0 ? 0 : split_stripes(ql, tl, t.R);
            t.ops_off = 0;
            t.nops = 0;
            t.t1 = ll * m;
            t.mp = mss[i];
            t.maxv = m;
            t.cand_cap = needs ? max(cand0, needs[g]) : cand0;
            t.ncand = 0;
            t.nsp = 0;
            t.same_id = same_id[i];
            t.qidx = pr.x;
            t.tidx = pr.y;
            t.out_off = 0;
            t.split = 0;
            t.rows_off = 0;
            t.prog_off = 0;
            t.ck = ck;
            t.gslot = slots ? slots[g] : g;
            if (ck >= 0)
                ckn = ck_units(ckpairs[ck], ckarea, (ql + WAVE * CK_R - 1) / (WAVE * CK_R),
                               __builtin_fmaxf(t.t1, t.mp), nullptr, g);
        }
        const unsigned long long doff = wave_alloc(&ctl->dir_used, on ? (uint64_t)area : 0ull);
        const int cneed = on ? t.cand_cap : 0;
        const unsigned long long coff = wave_alloc(&ctl->cand_used, (uint64_t)cneed);
        const bool ok = on && (int64_t)doff + area <= dir_cap && (int64_t)coff + cneed <= cand_cap_total;
        const int64_t rows = (ok && ns > 0) ? (int64_t)(ns - 1) * link_row_floats(tl) : 0;
        const unsigned long long lo = wave_alloc(&ctl->link_used, (uint64_t)rows);
        const unsigned long long po = wave_alloc(&ctl->prog_used, (ok && ns > 0) ? (uint64_t)ns : 0ull);
        const bool split = ok && ns > 0 && (int64_t)lo + rows <= link_cap && (int64_t)po + ns <= prog_cap;
        const int cnt = ok ? (ck >= 0 ? ckn : (split ? ns : 1)) : 0;
        const unsigned long long k0 = wave_alloc(reinterpret_cast<unsigned long long *>(&ctl->units_used),
                                                 (uint64_t)cnt);
        if (!on) continue;
        t.dir_off = (int64_t)doff;
        t.cand_off = (int64_t)coff;
        t.status = ok ? TASK_OK : TASK_DEFERRED;
        if (split) {
            t.split = 1;
            t.rows_off = (int64_t)lo;
            t.prog_off = (int64_t)po;
            for (int k = 0; k < ns; ++k) prog[po + k] = 0u;
        }
        if (ok) {
            if ((int64_t)k0 + cnt <= units_cap) {
                if (ck >= 0) {
                    (void)ck_units(ckpairs[ck], ckarea, (ql + WAVE * CK_R - 1) / (WAVE * CK_R),
                                   __builtin_fmaxf(t.t1, t.mp), units + k0, g);
                } else {
                    for (int k = 0; k < cnt; ++k) units[k0 + k] = DirUnit{g, split ? k : -1, -1};
                }
            } else {
                // no unit slots (the host sizes them for every pair): finish the
                // pair in a tail round; its in-range slots become no-ops
                t.status = TASK_DEFERRED;
                for (int k = 0; k < cnt && (int64_t)k0 + k < units_cap; ++k) units[k0 + k] = DirUnit{g, -1, -1};
            }
        }
        tasks[g] = t;
    }
}

// Next round of an execution: the tasks of the round just walked that did not
// finish (no room for their directions or candidates, a candidate count
// beyond their slots, no room for their records) are listed for task_kernel,
// each with its read-back slot and the candidate slots it needs (its exact
// count after a candidate overflow).  The list replaces the gated list; the
// count goes to ctl->next_ngated (one atomic per wave).
__global__ void __launch_bounds__(256) requeue_kernel(const PairTask *__restrict__ tasks, Ctl *__restrict__ ctl,
                                                      int *__restrict__ gated, int *__restrict__ slots,
                                                      int *__restrict__ needs) {
    const int n = task_count(ctl);
    const int lane = threadIdx.x & 63;
    const int stride = gridDim.x * blockDim.x;
    for (int g0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63); g0 < n; g0 += stride) {
        const int g = g0 + lane;
        const bool redo = g < n && tasks[g].status != TASK_OK;
        const uint64_t mask = __ballot(redo);
        if (!mask) continue;
        int base = 0;
        if (lane == 0) base = atomicAdd(&ctl->next_ngated, __popcll(mask));
        base = __shfl(base, 0, 64);
        if (redo) {
            const PairTask &t = tasks[g];
            const int k = base + __popcll(mask & ((1ull << lane) - 1ull));
            gated[k] = t.pair;
            slots[k] = t.gslot;
            needs[k] = t.status == TASK_CAND_OVERFLOW ? max(t.ncand, t.cand_cap) : t.cand_cap;
        }
    }
}

// Starts the next round (one thread): the ntasks requeued tasks (the count the
// host read back) become the round's tasks and the traceback areas' allocators
// restart; the record areas (out, ops) keep what earlier rounds wrote (after
// an area grew, ops_keep / out_keep >= 0 restarts its allocator there: every
// allocation that succeeded ends at or below the old capacity).  The
// control block is written with device-scope atomic stores, as every kernel
// reads it with device-scope atomics: a plain access can be served by this
// XCD's L2 from a line cached before the requeue's atomics (seen on MI355X:
// a plain load of next_ngated returned 0 after the host had read the count).
__global__ void round_reset_kernel(Ctl *__restrict__ ctl, int ntasks, long long ops_keep, long long out_keep) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (ops_keep >= 0)
            __hip_atomic_store(&ctl->ops_used, (unsigned long long)ops_keep, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (out_keep >= 0)
            __hip_atomic_store(&ctl->out_used, (unsigned long long)out_keep, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->ngated, ntasks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->next_ngated, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->unfinished, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->dir_queue, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->dir_used, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->cand_used, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->link_used, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->prog_used, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->units_used, 0ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// Start selection, marking and the walk: one thread per pair, or one wave per
// pair for pairs whose paths may be long.
// ---------------------------------------------------------------------------

// Nibble index (in the whole direction area, two per byte) of cell (x, y) of a
// task, and reading / or-ing / replacing it.
__device__ __forceinline__ int64_t cell_offset(const PairTask &it, int tlen, int x, int y) {
    const int R = it.R, rows = WAVE * R;
    const int s = x / rows, rem = x - s * rows, k = rem / R, r = rem - k * R;
    const int64_t nsteps = (int64_t)tlen + WAVE - 1;
    return 2 * it.dir_off + (((int64_t)s * nsteps + (y + k)) * WAVE + k) * R + r;
}

__device__ __forceinline__ uint32_t nib_get(const uint8_t *p, int64_t n) {
    return ((uint32_t)p[n >> 1] >> ((uint32_t)(n & 1) << 2)) & 15u;
}

__device__ __forceinline__ void nib_put(uint8_t *p, int64_t n, uint32_t v) {
    const uint32_t sh = (uint32_t)(n & 1) << 2;
    p[n >> 1] = (uint8_t)(((uint32_t)p[n >> 1] & ~(15u << sh)) | ((v & 15u) << sh));
}

// Canonical emission order (larger key = emitted first): the serial reading
// of the reverse sweep of SmithWaterman._traceback_host (SmithWaterman.py:
// 626-653) and default_main_cpu.cl:414-420 — block anti-diagonal, block x,
// work-item anti-diagonal, work-item x, then the j and k loops, all descending.
__device__ __forceinline__ uint64_t order_key(uint32_t x, uint32_t y) {
    const uint64_t bx = x >> 3, by = y >> 3, tx = (x >> 2) & 1, ty = (y >> 2) & 1, j = x & 3, k = y & 3;
    return ((bx + by) << 45) | (bx << 26) | ((tx + ty) << 5) | (tx << 4) | (j << 2) | k;
}

template <class T, class Less>
__device__ __forceinline__ void heap_sift(T *a, int root, int end, Less less) {
    while (2 * root + 1 < end) {
        int child = 2 * root + 1;
        if (child + 1 < end && less(a[child], a[child + 1])) ++child;
        const T vr = a[root], vc = a[child];
        if (!less(vr, vc)) return;
        a[root] = vc;
        a[child] = vr;
        root = child;
    }
}

template <class T, class Less>
__device__ void heap_sort(T *a, int n, Less less) {
    for (int i = n / 2 - 1; i >= 0; --i) heap_sift(a, i, n, less);
    for (int end = n - 1; end > 0; --end) {
        const T v0 = a[0], ve = a[end];
        a[0] = ve;
        a[end] = v0;
        heap_sift(a, 0, end, less);
    }
}

// LDS window over one pair's direction area for the serial path walks of the
// start selection and the walk: stripe s, steps [t0, t0 + nt) of the
// tile-major layout (nt x 64R contiguous bytes), loaded by the whole wave with
// 16-byte loads when a cell outside it is needed.  A path moves to smaller
// steps (1 or 2 per cell), so a window ends at the step that missed.  With
// RES the window also holds the query residues of the stripe's rows and the
// target residues of the columns its steps cover (the walk counts matches).
// One wave runs one pair: every lane follows the same path (wave-uniform
// values, read from LDS by broadcast), lane 0 alone stores to memory.
constexpr int WALK_WIN = 16384;

// Pairs whose paths may be long (|q| + |t| >= LONG_PATH_MIN) take the
// wave-per-pair kernels below; the others one thread each (tb_kernel,
// walk_kernel), where a short path costs less than loading windows.  The two
// kernels run one after the other, so the threshold keeps mid-length pairs
// (proteins) out of the wave form: its time adds to the thread form's.
constexpr int LONG_PATH_MIN = 8192;
__device__ __host__ __forceinline__ bool long_path_lengths(int64_t qlen, int64_t tlen) {
    return qlen + tlen >= LONG_PATH_MIN;
}
__device__ __forceinline__ bool long_path_pair(const SeqRef *__restrict__ seqs, const PairTask &it) {
    return long_path_lengths(seqs[it.q_seq].len, seqs[it.t_seq].len);
}
constexpr int WALK_QWIN = WAVE * 8;                      // 64R rows, R <= 8
constexpr int WALK_TWIN = 2 * WALK_WIN / WAVE + WAVE;    // steps of an R = 1 window (2 cells per byte) plus the skew
constexpr int WALK_LDS = WALK_WIN + WALK_QWIN + WALK_TWIN;

struct PathWindow {
    uint8_t *blk, *qw, *tw;  // LDS
    int64_t dir_off, nsteps;
    int R, lgR, s, t0, nt;  // lgR: log2 R, or -1 for R = 6
    SeqRef q, t;
};

// v / R for v >= 0 and R in {1, 2, 4, 8} (a shift) or 6 (a multiply by the
// reciprocal: 0xAAAAAAAB / 2^34 divides every 32-bit unsigned value by 6 exactly)
__device__ __forceinline__ int div_r(int v, int lgR) {
    return lgR >= 0 ? (v >> lgR) : (int)(((uint64_t)(uint32_t)v * 0xAAAAAAABull) >> 34);
}

__device__ __forceinline__ void window_init(PathWindow &w, uint8_t *lds, const PairTask &it, SeqRef q, SeqRef t) {
    w.blk = lds;
    w.qw = lds + WALK_WIN;
    w.tw = lds + WALK_WIN + WALK_QWIN;
    w.dir_off = it.dir_off;
    w.nsteps = (int64_t)t.len + WAVE - 1;
    w.R = it.R;
    w.lgR = it.R == 1 ? 0 : it.R == 2 ? 1 : it.R == 4 ? 2 : it.R == 8 ? 3 : -1;
    w.s = -1;
    w.t0 = 0;
    w.nt = 0;
    w.q = q;
    w.t = t;
}

// LDS index of cell (x, y) (query row x, target column y), loading the window
// around it first when it lies outside.
template <bool RES>
__device__ __forceinline__ int window_cell(PathWindow &w, const uint8_t *__restrict__ dirs,
                                           const uint8_t *__restrict__ res, int x, int y) {
    const int R = w.R, rows = WAVE * R;
    const int s = div_r(x >> 6, w.lgR), rem = x - s * rows, k = div_r(rem, w.lgR), r = rem - k * R;
    const int t = y + k;
    if (s != w.s || t < w.t0 || t >= w.t0 + w.nt) {
        const int lane = threadIdx.x & 63;
        const int B = 2 * WALK_WIN / rows;  // steps of rows nibbles
        const int t0 = max(0, t - B + 1);
        const int nt = (int)min((int64_t)B, w.nsteps - t0);
        // lane 0's earlier stores (marks) reach memory before the wave reads
        // it back, and every lane's reads of the old window are done
        __syncthreads();
        // (a step holds 32R bytes and the area starts 16-byte aligned, so the
        // window does too; loads go out eight at a time before their stores,
        // one memory round trip per 8 KB instead of one per kilobyte)
        const uint4 *s4 = reinterpret_cast<const uint4 *>(dirs + w.dir_off + ((int64_t)s * w.nsteps + t0) * rows / 2);
        uint4 *d4 = reinterpret_cast<uint4 *>(w.blk);
        const int n16 = nt * rows / 32;
        constexpr int BATCH = 8;
        for (int i0 = 0; i0 < n16; i0 += BATCH * WAVE) {
            // unconditional loads (indices clamped into the window), so that
            // the batch stays in registers
            uint4 v[BATCH];
#pragma unroll
            for (int b = 0; b < BATCH; ++b) v[b] = s4[min(i0 + b * WAVE + lane, n16 - 1)];
            // (every value taken here: otherwise the compiler sinks each load
            // into its store's branch, one round trip per load again)
#pragma unroll
            for (int b = 0; b < BATCH; ++b) asm volatile("" : "+v"(v[b].x), "+v"(v[b].y), "+v"(v[b].z), "+v"(v[b].w));
#pragma unroll
            for (int b = 0; b < BATCH; ++b) {
                const int i = i0 + b * WAVE + lane;
                if (i < n16) d4[i] = v[b];
            }
        }
        if (RES) {
            // the stripe's query residues (64R <= WALK_QWIN) and the window's
            // target columns (nt + 63 <= WALK_TWIN), every load before any store
            constexpr int QB = WALK_QWIN / WAVE, TB = WALK_TWIN / WAVE + 1;
            int qv[QB], tv[TB];
            // (clamped indices: every load is made, the value kept only in range)
            const int qn = (int)w.q.len, tn = (int)w.t.len;
#pragma unroll
            for (int b = 0; b < QB; ++b) {
                const int i = b * WAVE + lane, xx = s * rows + i;
                const uint8_t c = res[w.q.off + min(max(xx, 0), max(qn - 1, 0))];
                qv[b] = (i < rows && xx < qn) ? (int)c : 0;
            }
#pragma unroll
            for (int b = 0; b < TB; ++b) {
                const int i = b * WAVE + lane, yy = t0 - (WAVE - 1) + i;
                const uint8_t c = res[w.t.off + min(max(yy, 0), max(tn - 1, 0))];
                tv[b] = (i < nt + WAVE - 1 && yy >= 0 && yy < tn) ? (int)c : 0;
            }
#pragma unroll
            for (int b = 0; b < QB; ++b) asm volatile("" : "+v"(qv[b]));
#pragma unroll
            for (int b = 0; b < TB; ++b) asm volatile("" : "+v"(tv[b]));
#pragma unroll
            for (int b = 0; b < QB; ++b)
                if (b * WAVE + lane < rows) w.qw[b * WAVE + lane] = (uint8_t)qv[b];
#pragma unroll
            for (int b = 0; b < TB; ++b)
                if (b * WAVE + lane < nt + WAVE - 1) w.tw[b * WAVE + lane] = (uint8_t)tv[b];
        }
        __syncthreads();
        w.s = s;
        w.t0 = t0;
        w.nt = nt;
    }
    return ((t - w.t0) * WAVE + k) * R + r;
}

// Window index of cell (x, y) (x, y >= 0) with no window load, per lane; ok
// says whether the current window holds the cell.
__device__ __forceinline__ int window_index(const PathWindow &w, int x, int y, bool &ok) {
    const int R = w.R, rows = WAVE * R;
    const int s = div_r(x >> 6, w.lgR), rem = x - s * rows, k = div_r(rem, w.lgR), r = rem - k * R;
    const int t = y + k;
    ok = s == w.s && t >= w.t0 && t < w.t0 + w.nt;
    return ((t - w.t0) * WAVE + k) * R + r;
}

// Writes nibble b of window index idx (old: its byte as read) into the window
// and memory, from the calling lane.
__device__ __forceinline__ void window_write_lane(PathWindow &w, uint8_t *__restrict__ dirs, int idx, uint32_t b,
                                                  uint32_t old) {
    const uint32_t sh = (uint32_t)(idx & 1) << 2;
    const uint8_t nb = (uint8_t)((old & ~(15u << sh)) | ((b & 15u) << sh));
    w.blk[idx >> 1] = nb;
    dirs[w.dir_off + ((int64_t)w.s * w.nsteps + w.t0) * (WAVE * w.R) / 2 + (idx >> 1)] = nb;
}

// Diagonal runs of a path, one cell per lane: lane i reads cell (x - i, y - i)
// when the current window holds it (no load); its window index, byte and
// nibble come back in o, byte and nib (valid when ok).  Two cells of one
// diagonal never share a byte (their steps differ), so the lanes can write
// their cells' nibbles at once.
__device__ __forceinline__ void diag_cells(const PathWindow &w, int x, int y, bool &ok, int &o, uint32_t &byte,
                                           uint32_t &nib) {
    const int lane = threadIdx.x & 63;
    const int xi = x - lane, yi = y - lane;
    ok = xi >= 0 && yi >= 0;
    o = 0;
    byte = 0;
    nib = 0;
    if (ok) o = window_index(w, xi, yi, ok);
    if (ok) {
        byte = w.blk[o >> 1];
        nib = (byte >> ((uint32_t)(o & 1) << 2)) & 15u;
    }
}

// The direction nibble of cell (x, y), wave-uniform; idx is its nibble index
// in the window.
template <bool RES>
__device__ __forceinline__ uint32_t window_read(PathWindow &w, const uint8_t *dirs, const uint8_t *res, int x, int y,
                                                int &idx, uint32_t &byte) {
    idx = window_cell<RES>(w, dirs, res, x, y);
    byte = (uint32_t)__builtin_amdgcn_readfirstlane((int)w.blk[idx >> 1]);
    return (byte >> ((uint32_t)(idx & 1) << 2)) & 15u;
}

// Replaces the nibble at window index idx with b, old being its byte as read
// (its global offset derived from the window's stripe and steps): LDS by every
// lane, memory by lane 0.
__device__ __forceinline__ void window_write(PathWindow &w, uint8_t *__restrict__ dirs, int idx, uint32_t b,
                                             uint32_t old) {
    const uint32_t sh = (uint32_t)(idx & 1) << 2;
    const uint8_t nb = (uint8_t)((old & ~(15u << sh)) | ((b & 15u) << sh));
    w.blk[idx >> 1] = nb;
    if ((threadIdx.x & 63) == 0)
        dirs[w.dir_off + ((int64_t)w.s * w.nsteps + w.t0) * (WAVE * w.R) / 2 + (idx >> 1)] = nb;
}

// default_main_cpu.cl:389-604 restated per pair: candidates in canonical
// reverse-sweep order; an unmarked candidate becomes a starting point and its
// path is marked (linear: along the direction codes until a predecessor of
// value 0; affine: along main-matrix cells only — I/J marks never arise
// because they would need an existing I/J mark); a marked candidate becomes
// a starting point only if its sign-marked value passes the tests.  One
// thread per pair of short paths (long_path_pair: tb_wave_kernel).
template <bool AFF>
__global__ void __launch_bounds__(64) tb_kernel(const SeqRef *__restrict__ seqs, PairTask *__restrict__ tasks,
                                                const Ctl *__restrict__ ctl, uint8_t *__restrict__ dirs,
                                                Cand *__restrict__ cands, SpRec *__restrict__ sps) {
    const int n = task_count(ctl);
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
        const PairTask it = tasks[g];
        if (it.status != TASK_OK || long_path_pair(seqs, it)) continue;
        if (it.ncand > it.cand_cap) {
            tasks[g].status = TASK_CAND_OVERFLOW;
            continue;
        }
        const int tlen = seqs[it.t_seq].len;
        Cand *c = cands + it.cand_off;
        const int nc = it.ncand;
        constexpr int32_t XM = ~(CAND_PLAIN | CAND_MARKED);
        heap_sort(c, nc, [](const Cand &a, const Cand &b) {
            return order_key(a.x & XM, a.y) > order_key(b.x & XM, b.y);
        });
        SpRec *out = sps + it.cand_off;
        int nsp = 0;
        for (int i = 0; i < nc; ++i) {
            const int fl = c[i].x & (CAND_PLAIN | CAND_MARKED);
            int x = c[i].x & ~(CAND_PLAIN | CAND_MARKED), y = c[i].y;
            const float v = c[i].v;
            const int64_t off = cell_offset(it, tlen, x, y);
            const uint32_t b = nib_get(dirs, off);
            float score;
            bool emit = false;
            if (!(b & DB_MARK)) {
                if (fl & CAND_PLAIN) {
                    emit = true;
                    score = v;
                    // mark the path
                    int cx = x, cy = y;
                    for (;;) {
                        const int64_t o = cell_offset(it, tlen, cx, cy);
                        const uint32_t cb = nib_get(dirs, o);
                        nib_put(dirs, o, cb | DB_MARK);
                        const uint32_t code = cb & 3u;
                        if (cb & DB_STOP) break;
                        if (code == 0u) break;
                        if (AFF && code != 1u) break;
                        const int px = cx - (code != 2u ? 1 : 0), py = cy - (code != 3u ? 1 : 0);
                        if (px < 0 || py < 0) break;
                        if (nib_get(dirs, cell_offset(it, tlen, px, py)) & DB_MARK) break;
                        cx = px;
                        cy = py;
                    }
                }
            } else if (fl & CAND_MARKED) {
                emit = true;
                score = -v;
            }
            if (emit) {
                SpRec r;
                r.sequence = (uint32_t)it.qidx;
                r.target = (uint32_t)it.tidx;
                r.blockX = (uint32_t)x >> 3;
                r.blockY = (uint32_t)y >> 3;
                r.valueX = (uint32_t)x & 7;
                r.valueY = (uint32_t)y & 7;
                r.score = score;
                r.maxScore = it.maxv;
                r.posScore = it.mp;
                out[nsp++] = r;
            }
        }
        tasks[g].nsp = nsp;
    }
}

// The same per pair of long paths, one wave per pair (persistent over the
// task list), following the marking paths through an LDS window.
template <bool AFF>
__global__ void __launch_bounds__(64) tb_wave_kernel(const SeqRef *__restrict__ seqs, PairTask *__restrict__ tasks,
                                                     const Ctl *__restrict__ ctl, uint8_t *__restrict__ dirs,
                                                     Cand *__restrict__ cands, SpRec *__restrict__ sps) {
    extern __shared__ uint8_t wlds[];
    const int n = task_count(ctl);
    const bool lead = (threadIdx.x & 63) == 0;
    for (int g = blockIdx.x; g < n; g += gridDim.x) {
        const PairTask it = tasks[g];
        if (it.status != TASK_OK || !long_path_pair(seqs, it)) continue;
        if (it.ncand > it.cand_cap) {
            if (lead) tasks[g].status = TASK_CAND_OVERFLOW;
            continue;
        }
        Cand *c = cands + it.cand_off;
        const int nc = it.ncand;
        constexpr int32_t XM = ~(CAND_PLAIN | CAND_MARKED);
        // every lane runs the same sort on the same addresses (identical
        // values, so the wave's coinciding stores agree)
        heap_sort(c, nc, [](const Cand &a, const Cand &b) {
            return order_key(a.x & XM, a.y) > order_key(b.x & XM, b.y);
        });
        PathWindow w;
        window_init(w, wlds, it, seqs[it.q_seq], seqs[it.t_seq]);
        SpRec *out = sps + it.cand_off;
        int nsp = 0;
        for (int i = 0; i < nc; ++i) {
            const Cand ci = c[i];
            const int fl = __builtin_amdgcn_readfirstlane(ci.x & (CAND_PLAIN | CAND_MARKED));
            const int x = __builtin_amdgcn_readfirstlane(ci.x & ~(CAND_PLAIN | CAND_MARKED));
            const int y = __builtin_amdgcn_readfirstlane(ci.y);
            const float v = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(ci.v)));
            int idx;
            uint32_t byte;
            const uint32_t b = window_read<false>(w, dirs, nullptr, x, y, idx, byte);
            float score = 0.0f;
            bool emit = false;
            if (!(b & DB_MARK)) {
                if (fl & CAND_PLAIN) {
                    emit = true;
                    score = v;
                    // mark the path
                    // (the predecessor's nibble, read for its mark, is the next
                    // cell's: read once; writes go to the cell just read, so the
                    // window holds it when the next read needs it)
                    int cx = x, cy = y, o = idx;
                    uint32_t cb = b, cbyte = byte;
                    for (;;) {
                        if ((cb & 3u) == DB_MAIN && !(cb & DB_STOP)) {
                            // a diagonal run, one cell per lane: cell i (i >= 1)
                            // is reached when cells 0..i-1 go on diagonally and
                            // cells 1..i are in the window and unmarked; every
                            // reached cell is marked, then the last one goes on
                            // below as the current cell (marked again, the same
                            // byte)
                            bool ok;
                            int oi;
                            uint32_t byi, ni;
                            diag_cells(w, cx, cy, ok, oi, byi, ni);
                            const int lane = threadIdx.x & 63;
                            const int dc = (ok && (ni & 3u) == DB_MAIN && !(ni & DB_STOP)) ? 1 : 0;
                            const int dprev = shr1i(0, dc);  // lane i: dc of lane i-1 (every lane takes part)
                            const bool reach = ok && (lane == 0 || (!(ni & DB_MARK) && dprev != 0));
                            const uint64_t miss = __builtin_amdgcn_ballot_w64(!reach);
                            const int m = miss ? __builtin_ctzll(miss) : WAVE;  // cells 0..m-1 reached
                            if (lane < m) window_write_lane(w, dirs, oi, ni | DB_MARK, byi);
                            if (m > 1) {
                                cx -= m - 1;
                                cy -= m - 1;
                                o = __builtin_amdgcn_readlane(oi, m - 1);
                                cb = (uint32_t)__builtin_amdgcn_readlane((int)ni, m - 1);
                                cbyte = (uint32_t)__builtin_amdgcn_readlane((int)byi, m - 1);
                            }
                        }
                        window_write(w, dirs, o, cb | DB_MARK, cbyte);
                        const uint32_t code = cb & 3u;
                        if (cb & DB_STOP) break;
                        if (code == 0u) break;
                        if (AFF && code != 1u) break;
                        const int px = cx - (code != 2u ? 1 : 0), py = cy - (code != 3u ? 1 : 0);
                        if (px < 0 || py < 0) break;
                        const uint32_t pb = window_read<false>(w, dirs, nullptr, px, py, o, cbyte);
                        if (pb & DB_MARK) break;
                        cb = pb;
                        cx = px;
                        cy = py;
                    }
                }
            } else if (fl & CAND_MARKED) {
                emit = true;
                score = -v;
            }
            if (emit) {
                if (lead) {
                    SpRec r;
                    r.sequence = (uint32_t)it.qidx;
                    r.target = (uint32_t)it.tidx;
                    r.blockX = (uint32_t)x >> 3;
                    r.blockY = (uint32_t)y >> 3;
                    r.valueX = (uint32_t)x & 7;
                    r.valueY = (uint32_t)y & 7;
                    r.score = score;
                    r.maxScore = it.maxv;
                    r.posScore = it.mp;
                    out[nsp] = r;
                }
                ++nsp;
            }
        }
        if (lead) tasks[g].nsp = nsp;
    }
}

__device__ __forceinline__ uint32_t float_order(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Direction values of SmithWaterman.py:761-821 after the traceback rewrote
// the matrix: STOP where a marked cell's predecessor is 0, otherwise the
// code (linear none = 32; affine keeps direction & 3, so none = 0).
template <bool AFF>
__device__ __forceinline__ int eff_dir(uint32_t b) {
    if (is_walked(b)) return 96;
    const int code = b & 3;
    if ((b & DB_MARK) && (b & DB_STOP) && (AFF ? code == 1 : code != 0)) return 64;
    if (code == 0) return AFF ? 0 : 32;
    return code;
}

// The walk of SmithWaterman.py:657-848 per pair: starting points in
// ascending score order (stable), each walked until STOP / NONE / an
// IN_ALIGNMENT cell.  The pair's op characters go to a range taken from the
// ops bump allocator (capacity: every walk at most |q|+|t| columns and all
// walks together at most |q||t| cells plus one column per walk); its starting
// points (emission order) and hits (walk order) are compacted into out_sps /
// out_hits at a range taken from the output allocator.
template <bool AFF>
__global__ void __launch_bounds__(64) walk_kernel(const uint8_t *__restrict__ res, const SeqRef *__restrict__ seqs,
                                                  PairTask *__restrict__ tasks, Ctl *__restrict__ ctl,
                                                  uint8_t *__restrict__ dirs, const SpRec *__restrict__ sps,
                                                  uint64_t *__restrict__ keys, SpRec *__restrict__ out_sps,
                                                  HitRec *__restrict__ out_hits, char *__restrict__ ops,
                                                  int64_t ops_cap, double minimum, TaskOut *__restrict__ touts,
                                                  int64_t out_cap) {
    const int n = task_count(ctl);
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
        const PairTask it = tasks[g];
        if (it.status != TASK_OK) {
            // (deferred or candidate overflow: counted once per wave for the next round)
            touts[it.gslot] = TaskOut{it.pair, it.status, 0, 0, 0, 0};
            const uint64_t m = __ballot(1);
            if ((int)(threadIdx.x & 63) == __ffsll((long long)m) - 1)
                atomicAdd(&ctl->unfinished, (unsigned)__popcll(m));
            continue;
        }
        if (long_path_pair(seqs, it)) continue;
        const SeqRef q = seqs[it.q_seq], t = seqs[it.t_seq];
        const int nsp = it.nsp;
        const int64_t ql = q.len, tl = t.len;
        const int64_t cap = nsp > 0 ? min((int64_t)nsp * (ql + tl), ql * tl + nsp) : 0;
        const unsigned long long ooff = cap > 0 ? atomicAdd(&ctl->ops_used, (unsigned long long)cap) : 0ull;
        const unsigned long long out0 =
            ((int64_t)ooff + cap <= ops_cap && nsp > 0) ? atomicAdd(&ctl->out_used, (unsigned long long)nsp) : 0ull;
        // no room for its op characters or records (the record areas collect
        // every round of an execution): the pair goes to a later round
        if ((int64_t)ooff + cap > ops_cap || (int64_t)out0 + nsp > out_cap) {
            tasks[g].status = TASK_OPS_OVERFLOW;
            touts[it.gslot] = TaskOut{it.pair, TASK_OPS_OVERFLOW, 0, 0, 0, 0};
            atomicAdd(&ctl->unfinished, 1u);
            continue;
        }
        const SpRec *sp = sps + it.cand_off;
        uint64_t *key = keys + it.cand_off;
        for (int i = 0; i < nsp; ++i) {
            key[i] = ((uint64_t)float_order(sp[i].score) << 32) | (uint32_t)i;
            out_sps[out0 + i] = sp[i];
        }
        // ascending score, stable (SmithWaterman.py:699)
        heap_sort(key, nsp, [](const uint64_t &a, const uint64_t &b) { return a < b; });
        char *op = ops + ooff;
        int64_t nops = 0;
        HitRec *h = out_hits + out0;
        for (int k = 0; k < nsp; ++k) {
            const int i = (int)(key[k] & 0xffffffffu);
            const SpRec r = sp[i];
            int x = (int)(r.blockX * 8 + r.valueX), y = (int)(r.blockY * 8 + r.valueY);
            const int s_end = x, t_end = y;
            int s_start = s_end + 1, t_start = t_end + 1;
            int matches = 0, mismatches = 0, len = 0;
            int direction = eff_dir<AFF>(nib_get(dirs, cell_offset(it, t.len, x, y)));
            bool show = true;
            if (it.same_id) {
                direction = 64;
                show = false;
            }
            if ((double)r.score < minimum) show = false;
            const int64_t start_ops = nops;
            bool term = false;
            while (show && !term && x >= 0 && y >= 0 && direction != 64 && direction != 32) {
                const int64_t o = cell_offset(it, t.len, x, y);
                const uint32_t b = nib_get(dirs, o);
                direction = eff_dir<AFF>(b);
                nib_put(dirs, o, nib_walked(b));
                if (direction == 96) {
                    show = false;
                } else if (direction == 1 || direction == 64) {
                    if (res[q.off + x] == res[t.off + y]) ++matches; else ++mismatches;
                    op[nops++] = 'D';
                    ++len;
                    --s_start;
                    --t_start;
                    if (direction == 1) {
                        --x;
                        --y;
                    }
                } else if (direction == 3) {
                    op[nops++] = 'L';
                    ++len;
                    --s_start;
                    --x;
                } else if (direction == 2) {
                    op[nops++] = 'U';
                    ++len;
                    --t_start;
                    --y;
                } else {
                    term = true;
                }
            }
            if (!show) nops = start_ops;
            HitRec hr;
            hr.query = it.qidx;
            hr.target = it.tidx;
            hr.shown = show ? 1 : 0;
            hr.s_start = s_start;
            hr.s_end = s_end;
            hr.t_start = t_start;
            hr.t_end = t_end;
            hr.matches = matches;
            hr.mismatches = mismatches;
            hr.length = show ? len : 0;
            hr.score = r.score;
            hr.sp_index = i;
            hr.ops_offset = start_ops;
            h[k] = hr;
        }
        tasks[g].ops_off = (int64_t)ooff;
        tasks[g].nops = nops;
        tasks[g].out_off = (int32_t)out0;
        touts[it.gslot] = TaskOut{it.pair, TASK_OK, nsp, (int32_t)out0, (int64_t)ooff, nops};
    }
}

// The same per pair of long paths, one wave per pair through an LDS window
// (match counts from the window's residues; lane 0 stores).
template <bool AFF>
__global__ void __launch_bounds__(64) walk_wave_kernel(const uint8_t *__restrict__ res, const SeqRef *__restrict__ seqs,
                                                       PairTask *__restrict__ tasks, Ctl *__restrict__ ctl,
                                                       uint8_t *__restrict__ dirs, const SpRec *__restrict__ sps,
                                                       uint64_t *__restrict__ keys, SpRec *__restrict__ out_sps,
                                                       HitRec *__restrict__ out_hits, char *__restrict__ ops,
                                                       int64_t ops_cap, double minimum, TaskOut *__restrict__ touts,
                                                       int64_t out_cap) {
    extern __shared__ uint8_t wlds[];
    const int n = task_count(ctl);
    const int lane = threadIdx.x & 63;
    const bool lead = lane == 0;
    for (int g = blockIdx.x; g < n; g += gridDim.x) {
        const PairTask it = tasks[g];
        if (it.status != TASK_OK || !long_path_pair(seqs, it)) continue;
        const SeqRef q = seqs[it.q_seq], t = seqs[it.t_seq];
        const int nsp = it.nsp;
        const int64_t ql = q.len, tl = t.len;
        const int64_t cap = nsp > 0 ? min((int64_t)nsp * (ql + tl), ql * tl + nsp) : 0;
        unsigned long long ooff = 0, out0 = 0;
        if (lead && cap > 0) ooff = atomicAdd(&ctl->ops_used, (unsigned long long)cap);
        ooff = __shfl(ooff, 0, 64);
        if (lead && (int64_t)ooff + cap <= ops_cap && nsp > 0) out0 = atomicAdd(&ctl->out_used, (unsigned long long)nsp);
        out0 = __shfl(out0, 0, 64);
        // no room for its op characters or records: a later round
        if ((int64_t)ooff + cap > ops_cap || (int64_t)out0 + nsp > out_cap) {
            if (lead) {
                tasks[g].status = TASK_OPS_OVERFLOW;
                touts[it.gslot] = TaskOut{it.pair, TASK_OPS_OVERFLOW, 0, 0, 0, 0};
                atomicAdd(&ctl->unfinished, 1u);
            }
            continue;
        }
        const SpRec *sp = sps + it.cand_off;
        uint64_t *key = keys + it.cand_off;
        for (int i = lane; i < nsp; i += WAVE) out_sps[out0 + i] = sp[i];
        // every lane builds and sorts the same keys at the same addresses
        // (identical values, so the wave's coinciding stores agree)
        for (int i = 0; i < nsp; ++i) key[i] = ((uint64_t)float_order(sp[i].score) << 32) | (uint32_t)i;
        // ascending score, stable (SmithWaterman.py:699)
        heap_sort(key, nsp, [](const uint64_t &a, const uint64_t &b) { return a < b; });
        PathWindow w;
        window_init(w, wlds, it, q, t);
        char *op = ops + ooff;
        int64_t nops = 0;
        HitRec *h = out_hits + out0;
        for (int k = 0; k < nsp; ++k) {
            const int i = __builtin_amdgcn_readfirstlane((int)(key[k] & 0xffffffffu));
            const SpRec r = sp[i];
            int x = __builtin_amdgcn_readfirstlane((int)(r.blockX * 8 + r.valueX));
            int y = __builtin_amdgcn_readfirstlane((int)(r.blockY * 8 + r.valueY));
            const float rscore = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r.score)));
            const int s_end = x, t_end = y;
            int s_start = s_end + 1, t_start = t_end + 1;
            int matches = 0, mismatches = 0, len = 0;
            int idx;
            uint32_t byte;
            int direction = eff_dir<AFF>(window_read<true>(w, dirs, res, x, y, idx, byte));
            bool show = true;
            if (it.same_id) {
                direction = 64;
                show = false;
            }
            if ((double)rscore < minimum) show = false;
            const int64_t start_ops = nops;
            bool term = false;
            while (show && !term && x >= 0 && y >= 0 && direction != 64 && direction != 32) {
                if (direction == 1) {
                    // a diagonal run, one cell per lane: the cells from (x, y)
                    // on whose direction is 1 (up-left, not a stop), in the
                    // window, taken at once (walked marks, op characters,
                    // matches); the first other cell is taken below
                    bool ok;
                    int oi;
                    uint32_t byi, ni;
                    diag_cells(w, x, y, ok, oi, byi, ni);
                    const int lane = threadIdx.x & 63;
                    const bool dg = ok && eff_dir<AFF>(ni) == 1;
                    const uint64_t miss = __builtin_amdgcn_ballot_w64(!dg);
                    const int m = miss ? __builtin_ctzll(miss) : WAVE;
                    if (m > 0) {
                        bool same = false;
                        if (lane < m) {
                            same = w.qw[x - lane - w.s * WAVE * w.R] == w.tw[y - lane - (w.t0 - (WAVE - 1))];
                            window_write_lane(w, dirs, oi, nib_walked(ni), byi);
                            op[nops + lane] = 'D';
                        }
                        const int nm = __builtin_popcountll(__builtin_amdgcn_ballot_w64(lane < m && same));
                        matches += nm;
                        mismatches += m - nm;
                        nops += m;
                        len += m;
                        s_start -= m;
                        t_start -= m;
                        x -= m;
                        y -= m;
                        continue;
                    }
                }
                int o;
                const uint32_t b = window_read<true>(w, dirs, res, x, y, o, byte);
                // the window holds the residues of the cell just read (read
                // beside its nibble, used on diagonal moves)
                const int qa = __builtin_amdgcn_readfirstlane((int)w.qw[x - w.s * WAVE * w.R]);
                const int tb = __builtin_amdgcn_readfirstlane((int)w.tw[y - (w.t0 - (WAVE - 1))]);
                direction = eff_dir<AFF>(b);
                window_write(w, dirs, o, nib_walked(b), byte);
                if (direction == 96) {
                    show = false;
                } else if (direction == 1 || direction == 64) {
                    if (qa == tb) ++matches; else ++mismatches;
                    if (lead) op[nops] = 'D';
                    ++nops;
                    ++len;
                    --s_start;
                    --t_start;
                    if (direction == 1) {
                        --x;
                        --y;
                    }
                } else if (direction == 3) {
                    if (lead) op[nops] = 'L';
                    ++nops;
                    ++len;
                    --s_start;
                    --x;
                } else if (direction == 2) {
                    if (lead) op[nops] = 'U';
                    ++nops;
                    ++len;
                    --t_start;
                    --y;
                } else {
                    term = true;
                }
            }
            if (!show) nops = start_ops;
            if (lead) {
                HitRec hr;
                hr.query = it.qidx;
                hr.target = it.tidx;
                hr.shown = show ? 1 : 0;
                hr.s_start = s_start;
                hr.s_end = s_end;
                hr.t_start = t_start;
                hr.t_end = t_end;
                hr.matches = matches;
                hr.mismatches = mismatches;
                hr.length = show ? len : 0;
                hr.score = rscore;
                hr.sp_index = i;
                hr.ops_offset = start_ops;
                h[k] = hr;
            }
        }
        if (lead) {
            tasks[g].ops_off = (int64_t)ooff;
            tasks[g].nops = nops;
            tasks[g].out_off = (int32_t)out0;
            touts[it.gslot] = TaskOut{it.pair, TASK_OK, nsp, (int32_t)out0, (int64_t)ooff, nops};
        }
    }
}

}  // namespace pw
