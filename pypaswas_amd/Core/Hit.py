"""One local alignment between a query and a target (pyPaSWAS/Core/Hit.py:11-406).

The record text this class produces (SAM line, CIGAR, tags, trimmed FASTA)
is part of the parity contract, so every rule below reproduces the
reference's behaviour, including its quirks:

* locations must satisfy 0 <= start < end < original_length on both axes,
  otherwise CudaException (a one-column alignment is rejected too);
* ids ending in "RC" lose their last three characters in SAM records;
* for reverse-complement hits ``get_sam_line`` reverses ``alignment`` and
  ``sequence_match`` in place, writes the reverse complement of the query,
  POS = len(target) - target_end, and always starts the CIGAR with a
  soft clip (possibly "0S");
* CIGAR runs are runs of equal alignment characters; a run of ' ' is I or D
  by the query character at its last column;
* trimmed FASTA cuts the *query* at the *target* coordinates.
"""
from pypaswas_amd.Core.Exceptions import CudaException
from pypaswas_amd.Core.SWSeq import Seq


def _strip_rc(identifier):
    return identifier[:-3] if identifier[-2:] == 'RC' else identifier


class Hit(object):

    def __init__(self, logger, sequence_info, target_info, sequence_location, target_location):
        self.logger = logger
        self.sequence_info = sequence_info
        self.target_info = target_info
        self.seq_location = ()
        self.target_location = ()
        self._set_seq_location(sequence_location)
        self._set_target_location(target_location)
        self.sequence_match = ''
        self.alignment = ''
        self.target_match = ''
        self.score = 0
        self.matches = 0
        self.mismatches = 0
        self.query_identity = 0
        self.query_coverage = 0
        self.relative_score = 0
        self.distance = None
        self.rc = False

    @staticmethod
    def _is_a_valid_location(location, sequence_length):
        if not (isinstance(location, tuple) and len(location) == 2):
            return False
        start, end = location
        return (isinstance(start, int) and isinstance(end, int) and 0 <= start < sequence_length
                and 0 <= end < sequence_length and start < end)

    def _set_seq_location(self, location):
        if not self._is_a_valid_location(location, self.sequence_info.original_length):
            raise CudaException('Invalid sequence location: {0}, length={1}, {2}'.format(
                location, len(self.sequence_info.seq), self.sequence_info.original_length))
        self.seq_location = location

    def _set_target_location(self, location):
        if not self._is_a_valid_location(location, self.target_info.original_length):
            raise CudaException('Invalid target location: {0}'.format(location))
        self.target_location = location

    def set_scores(self, score, matches, mismatches):
        if not (isinstance(score, (float, int)) and score >= 0.0):
            raise CudaException('Score should be a float or an integer, not {0}'.format(score))
        if not (isinstance(matches, int) and matches >= 0):
            raise CudaException('Matches  should be integers. Passed value is {0}'.format(matches))
        if not (isinstance(mismatches, int) and mismatches >= 0):
            raise CudaException('Mismatches should be integers. Passed value is {0}.'.format(mismatches))
        self.score = score
        self.matches = matches
        self.mismatches = mismatches
        qlen = self.sequence_info.original_length
        tlen = self.target_info.original_length
        shortest = tlen if qlen > tlen else qlen
        self.relative_score = float(self.score) / float(shortest)
        span = self.seq_location[1] - self.seq_location[0] + 1
        self.query_identity = float(matches) / float(span)
        self.query_coverage = float(span) / float(qlen)
        self.base_score = float(self.score) / float(len(self.alignment))

    def set_sequence_match(self, sequence_match):
        self.sequence_match = sequence_match

    def set_target_match(self, target_match):
        self.target_match = target_match

    def set_alignment(self, alignment):
        self.alignment = alignment

    def keys(self):
        """Start key, end key, start key (HitList dedupe, HitList.py:28-52)."""
        q, t = self.get_seq_id(), self.get_target_id()
        start = (q, t, self.seq_location[0], self.target_location[0])
        end = (q, t, self.seq_location[1], self.target_location[1])
        return [start, end, start]

    def _check_rc(self):
        return self.get_seq_id()[-2:] == "RC" and self.get_target_id()[-2:] == "RC"

    def get_seq_id(self):
        return self.sequence_info.id

    def get_target_id(self):
        return self.target_info.id

    def get_euclidian_distance(self):
        return 0.0 if self.target_info.distance is None else self.target_info.distance

    def _require_strings(self):
        if not self.sequence_match or not self.target_match or not self.alignment:
            raise CudaException('sequence_match, target_match and alignment should be set and have lengths > 0.')

    def get_sam_line(self):
        query_rc = self.get_seq_id()[-2:] == 'RC'
        target_rc = self.get_target_id()[-2:] == 'RC'
        self.rc = query_rc or target_rc
        position = str(self.target_location[0] + 1)
        sam_sequence = self.sequence_info.seq
        if self.rc:
            self.alignment = self.alignment[::-1]
            self.sequence_match = self.sequence_match[::-1]
            sam_sequence = Seq(str(self.sequence_info.seq)).reverse_complement()
            position = str(self.target_info.original_length - self.target_location[1])
        self._require_strings()
        return '\t'.join([_strip_rc(self.get_seq_id()), self._get_sam_flag(), _strip_rc(self.get_target_id()),
                          position, Hit._get_sam_mapq(), self._get_sam_cigar(), '*', '0', '0', str(sam_sequence),
                          '*', self._get_sam_alignment_score(), self._get_sam_relative_score(),
                          self._get_sam_relative_base_score(), self._get_sam_query_coverage(),
                          self._get_sam_query_identity()])

    def _cut_query(self):
        seq = self.sequence_info.seq
        return str(seq[:self.target_location[0]] + seq[self.target_location[1] + 1:])

    def get_trimmed_line(self):
        self._require_strings()
        return '>' + _strip_rc(self.get_target_id()) + ' ' + self.sequence_info.id + '\n' + self._cut_query()

    def get_fasta(self):
        self._require_strings()
        return '>' + _strip_rc(self.get_target_id()) + ' ' + self.sequence_info.id + '\n' + self._cut_query()

    def get_full_fasta(self):
        return '>{}\n{}'.format(self.get_seq_id(), str(self.sequence_info.seq))

    def get_sam_sq(self):
        return '\t'.join(['@SQ', 'SN:{0}'.format(_strip_rc(self.get_target_id())),
                          'LN:{0}'.format(self.target_info.original_length)])

    def _get_sam_alignment_score(self):
        return 'AS:i:' + str(int(self.score))

    def _get_sam_flag(self):
        return '16' if (self.get_seq_id()[-2:] == 'RC' or self.get_target_id()[-2:] == 'RC') else '0'

    @staticmethod
    def _get_sam_mapq():
        return '255'

    def _get_sam_cigar(self):
        parts = []
        if self.rc:
            parts.append(str(len(self.sequence_info.seq) - self.seq_location[1] - 1) + 'S')
        elif self.seq_location[0] > 0:
            parts.append(str(self.seq_location[0]) + 'S')
        run_char = self.alignment[0]
        run = 0
        last = 0
        for last, ch in enumerate(self.alignment):
            if ch == run_char:
                run += 1
            else:
                parts.append(str(run) + self._get_cigar_char(last - 1))
                run_char = ch
                run = 1
        parts.append(str(run) + self._get_cigar_char(last))
        tail = self.seq_location[0] if self.rc else self.sequence_info.original_length - self.seq_location[1] - 1
        if tail > 0:
            parts.append(str(tail) + 'S')
        return ''.join(parts)

    def _get_sam_relative_score(self):
        return 'RS:f:' + str(self.relative_score)

    def _get_sam_query_coverage(self):
        return 'QC:f:' + str(self.query_coverage)

    def _get_sam_query_identity(self):
        return 'QI:f:' + str(self.query_identity)

    def _get_sam_relative_base_score(self):
        return 'BS:f:' + str(self.base_score)

    def _get_cigar_char(self, position):
        ch = self.alignment[position]
        if ch == '|':
            return '='
        if ch == '.':
            return 'X'
        return 'D' if self.sequence_match[position] == '-' else 'I'
