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
    qlen = np.array([len(s) for s in qstr], dtype=np.int64)
    tlen = np.array([len(s) for s in tstr], dtype=np.int64)
    mss = min_start_scores(highest, filter_factor, qlen[pq], tlen[pt])
    qids = np.array([ident(r) for r in queries], dtype=object)
    tids = np.array([ident(r) for r in targets], dtype=object)
    same = (qids[pq] == tids[pt]).astype(np.uint8)
    return PackedSeqs.pack(qstr), PackedSeqs.pack(tstr), pq, pt, mss, same


def decode_alignment(ops, offset, length, query, target, s_start, t_start):
    """Rebuilds the three alignment strings from the walk's column codes.

    The walk records one code per column from the alignment end back to its
    start, so the codes are read in reverse while x and y advance from the start.
    """
    x, y = s_start, t_start
    sm, al, tm = [], [], []
    for k in range(length - 1, -1, -1):
        op = ops[offset + k]
        if op == 68:  # 'D'
            a, b = query[x], target[y]
            sm.append(a)
            tm.append(b)
            al.append(SmithWaterman.MATCH_CHAR if a.lower() == b.lower() else SmithWaterman.MISMATCH_CHAR)
            x += 1
            y += 1
        elif op == 76:  # 'L'
            sm.append(query[x])
            tm.append(SmithWaterman.GAP_CHAR_SEQ)
            al.append(SmithWaterman.GAP_CHAR_ALIGN)
            x += 1
        else:  # 'U'
            sm.append(SmithWaterman.GAP_CHAR_SEQ)
            tm.append(target[y])
            al.append(SmithWaterman.GAP_CHAR_ALIGN)
            y += 1
    return ''.join(sm), ''.join(al), ''.join(tm)


class SmithWatermanHIP(SmithWaterman):

    def __init__(self, logger, score, settings, device=None, mode=1):
        SmithWaterman.__init__(self, logger, score, settings)
        self.mode = mode
        self.dev = device if device is not None else Device(self.device)
        self.last_result = None
        self.dev.set_scoring(score.to_float32(), ord(score.char_offset), score.gap_score,
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
        qstr = [str(r.seq) for r in queries]
        tstr = [str(r.seq) for r in targets]
        nq, nt = len(qstr), len(tstr)
        qp, tp, pq, pt, mss, same = job_arrays(queries, targets, self.score.highest_score, self.filter_factor)
        res = self.dev.align_pairs(qp, tp, pq, pt, mss, same, float(self.settings.minimum_score),
                                   self.maximum_number_starting_points, self.mode)
        self.last_result = res
        self.logger.debug('Number of starting points is: {0}.'.format(len(res.starting_points)))
        if len(res.starting_points) >= self.maximum_number_starting_points * nq * nt:
            self.logger.warning("Too many hits returned. Skipping the rest. Please set lower_limit_score higher in "
                                "config.")
        self._collect_hits(res, queries, targets, qstr, tstr, hitlist)

    def _collect_hits(self, res, queries, targets, qstr, tstr, hitlist):
        hits = res.hits
        if len(hits) == 0:
            return
        # starting points are processed in ascending score order, stable (SmithWaterman.py:699)
        order = np.argsort(hits['score'], kind='stable')
        for k in order:
            h = hits[k]
            if not h['shown']:
                continue
            qi, ti = int(h['query']), int(h['target'])
            qrec, trec = queries[qi], targets[ti]
            s_start = int(h['s_start']) + qrec.start_position
            s_end = int(h['s_end']) + qrec.start_position
            t_start = int(h['t_start']) + trec.start_position
            t_end = int(h['t_end']) + trec.start_position
            hit = Hit(self.logger, qrec, trec, (s_start, s_end), (t_start, t_end))
            sm, al, tm = decode_alignment(res.ops, int(h['ops_offset']), int(h['length']), qstr[qi], tstr[ti],
                                          int(h['s_start']), int(h['t_start']))
            hit.set_sequence_match(sm)
            hit.set_alignment(al)
            hit.set_target_match(tm)
            hit.set_scores(float(h['score']), int(h['matches']), int(h['mismatches']))
            if self._filter_hit(hit):
                hitlist.append(hit)
            else:
                self.logger.debug("Hit {0} -vs- {1} does not meet filter requirements".format(qrec.id, trec.id))
