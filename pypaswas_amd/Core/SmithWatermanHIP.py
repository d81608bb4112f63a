"""HIP backend of the engine: scoring, traceback and walk on an MI355X.

Plays the role SmithWatermanCPU (pyPaSWAS/Core/SmithWatermanOcl.py:338-518)
and SmithWatermanCuda (pyPaSWAS/Core/SmithWatermanCuda.py) play in the
reference, with the semantics of the OpenCL-CPU kernels.  One
``align_sequences`` call sends every query x target pair of the block to the
library in as few calls as the pair budget allows; the library returns the
pair maxima, the starting points and the walked alignments, and this class
builds Hit objects, applies the filter and fills the HitList exactly as
SmithWaterman._print_alignments (SmithWaterman.py:665-841) does.
"""
import time

import numpy as np

from pypaswas_amd.Core.Hit import Hit
from pypaswas_amd.Core.HitList import HitList
from pypaswas_amd.Core.SmithWaterman import SmithWaterman
from pypaswas_amd._native import Device, PackedSeqs

# pairs per library call; bounds host-side staging, not device memory
MAX_PAIRS_PER_CALL = 1 << 23


def min_start_scores(highest, filter_factor, qlens, tlens):
    """SmithWatermanOcl._fill_max_possible_score (:302-307): highest * min(len) * filter_factor, in double, stored f32."""
    shortest = np.minimum(np.asarray(qlens, dtype=np.float64), np.asarray(tlens, dtype=np.float64))
    return ((float(highest) * shortest) * float(filter_factor)).astype(np.float32)


def gap_open(score, settings):
    """The gap-open constant the kernels take.

    ``kernel_semantics`` CPU (the default) is the OpenCL-CPU kernels' recurrence,
    which opens an affine gap with G (default_main_cpu.cl:304, 319).  GPU is
    the scoring recurrence of the GPU-OpenCL and CUDA kernels, which open it
    with G + E (default_main_gpu.cl:298, 305; default_main.cu:487, 502), the
    sum formed in FP32 before it is added, as those kernels do.  Their other
    scoring differences change no value: zeroed fill cells are padding, and a
    linear cell whose best candidate ties at 0 has value 0 (it takes a
    direction there, default_main_gpu.cl:140-150, which matters only for a
    starting point of score 0, possible only with lower_limit_score 0 and
    filter_factor 0).  Start selection, marking and the walk are the OpenCL-CPU
    ones in both settings.  Linear gaps are scored the same in both."""
    semantics = str(getattr(settings, 'kernel_semantics', 'CPU') or 'CPU').upper()
    if semantics not in ('CPU', 'GPU'):
        from pypaswas_amd.Core.Exceptions import InvalidOptionException
        raise InvalidOptionException('kernel_semantics must be CPU or GPU, not {0}'.format(semantics))
    if semantics == 'GPU' and score.gap_extension is not None:
        return float(np.float32(score.gap_score) + np.float32(score.gap_extension))
    return float(score.gap_score)


def job_arrays(queries, targets, highest, filter_factor):
    """All-vs-all pair arrays for (id, sequence) records or SWSeqRecords.

    Returns packed queries, packed targets, pair query / target indices
    (query-major), minimum start scores and same-id flags.
    """
    def ident(r):
        return r.id if hasattr(r, 'id') else r[0]

    def seq(r):
        return str(r.seq) if hasattr(r, 'seq') else r[1]

    qstr = [seq(r) for r in queries]
    tstr = [seq(r) for r in targets]
    nq, nt = len(qstr), len(tstr)
    pq = np.repeat(np.arange(nq, dtype=np.int32), nt)
    pt = np.tile(np.arange(nt, dtype=np.int32), nq)
    qlen = np.array([len(s) for s in qstr], dtype=np.float64)
    tlen = np.array([len(s) for s in tstr], dtype=np.float64)
    # query-major nq x nt grids, raveled in the order of pq / pt
    mss = min_start_scores(highest, filter_factor, qlen[:, None], tlen[None, :]).ravel()
    # ids as integer codes over both lists, so that the per-pair comparison is numeric
    codes = {}
    qcode = np.array([codes.setdefault(ident(r), len(codes)) for r in queries], dtype=np.int64)
    tcode = np.array([codes.setdefault(ident(r), len(codes)) for r in targets], dtype=np.int64)
    same = (qcode[:, None] == tcode[None, :]).ravel().view(np.uint8)
    return PackedSeqs.pack(qstr), PackedSeqs.pack(tstr), pq, pt, mss, same


class SmithWatermanHIP(SmithWaterman):

    def __init__(self, logger, score, settings, device=None, mode=1):
        SmithWaterman.__init__(self, logger, score, settings)
        self.mode = mode
        self.dev = device if device is not None else Device(self.device)
        self.last_result = None
        self.dev.set_scoring(score.to_float32(), ord(score.char_offset), gap_open(score, settings),
                             score.gap_extension if score.gap_extension is not None else 0.0,
                             score.lower_limit_score, score.minimum_score)

    def _initialize_device(self, device_number):
        self.dev = Device(device_number)

    def _device_global_mem_size(self):
        return self.dev.global_mem_size()

    def _targets_per_block(self, targets, target_index, records_seqs):
        n = len(records_seqs) if records_seqs else 1
        return max(1, MAX_PAIRS_PER_CALL // max(1, n))

    def align_sequences(self, records_seqs, targets, target_index):
        hitlist = HitList(self.logger)
        block = targets[target_index:target_index + self.number_of_targets]
        if not records_seqs or not block:
            return hitlist
        per_call = max(1, MAX_PAIRS_PER_CALL // len(block))
        for q0 in range(0, len(records_seqs), per_call):
            t0 = time.time()
            queries = records_seqs[q0:q0 + per_call]
            self._align_block(queries, block, hitlist)
            processed = sum(len(s.seq) for s in queries) * sum(len(s.seq) for s in block)
            self._log_progress(processed, time.time() - t0)
        return hitlist

    def _align_block(self, queries, targets, hitlist):
        nq, nt = len(queries), len(targets)
        qp, tp, pq, pt, mss, same = job_arrays(queries, targets, self.score.highest_score, self.filter_factor)
        res = self.dev.align_pairs(qp, tp, pq, pt, mss, same, float(self.settings.minimum_score),
                                   self.maximum_number_starting_points, self.mode)
        self.last_result = res
        self.logger.debug('Number of starting points is: {0}.'.format(len(res.starting_points)))
        cap = self.maximum_number_starting_points * nq * nt
        if len(res.starting_points) >= cap:
            # SmithWaterman.py:683-685 drops the starting points beyond this count of its
            # memory-sized batch; here every one is walked and reported
            self.logger.warning("{0} starting points, at or above maximum_number_starting_points x pairs ({1}); "
                                "all are reported. Set lower_limit_score higher in the config to find fewer."
                                .format(len(res.starting_points), cap))
        self._collect_hits(res, queries, targets, hitlist)

    def _collect_hits(self, res, queries, targets, hitlist):
        hits = res.hits
        if len(hits) == 0:
            return
        # the records as Python values in one conversion, and the library's
        # alignment rows (SmithWaterman.py:777-830, start-to-end order, 3 chars
        # per op) as one string
        recs = hits[['shown', 'query', 'target', 's_start', 's_end', 't_start', 't_end', 'matches', 'mismatches',
                     'length', 'score', 'ops_offset']].tolist()
        rows = bytes(res.aligned).decode('latin-1')
        # starting points are processed in ascending score order, stable (SmithWaterman.py:699)
        order = np.argsort(hits['score'], kind='stable').tolist()
        for k in order:
            shown, qi, ti, s0, s1, t0, t1, matches, mismatches, n, score, off = recs[k]
            if not shown:
                continue
            qrec, trec = queries[qi], targets[ti]
            hit = Hit(self.logger, qrec, trec, (s0 + qrec.start_position, s1 + qrec.start_position),
                      (t0 + trec.start_position, t1 + trec.start_position))
            a0 = 3 * off
            hit.set_sequence_match(rows[a0:a0 + n])
            hit.set_alignment(rows[a0 + n:a0 + 2 * n])
            hit.set_target_match(rows[a0 + 2 * n:a0 + 3 * n])
            hit.set_scores(score, matches, mismatches)
            if self._filter_hit(hit):
                hitlist.append(hit)
            else:
                self.logger.debug("Hit {0} -vs- {1} does not meet filter requirements".format(qrec.id, trec.id))
