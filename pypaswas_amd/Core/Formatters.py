"""Output records (pyPaSWAS/Core/Formatters.py:3-191).

Hits are written in (query id, target id, score) order, ties in insertion
order.  TXT: 17 comma-space fields then the query, alignment and target
strings; SAM: @HD, one @SQ per target in first-seen order, @PG, one line per
hit; trimmedFasta and FASTA: one record per hit.
"""


class DefaultFormatter(object):

    def __init__(self, logger, hitlist, outputfile):
        self.logger = logger
        self.hitlist = hitlist
        self.outputfile = outputfile
        self.name = ''
        self._set_name()

    def _set_name(self):
        self.name = 'defaultformatter'

    def _get_hits(self):
        return sorted(self.hitlist.real_hits.values(),
                      key=lambda hit: (hit.get_seq_id(), hit.get_target_id(), hit.score))

    def _format_hit(self, hit):
        n = len(hit.alignment)
        fields = [hit.get_seq_id(), hit.get_target_id(), hit.seq_location[0], hit.seq_location[1],
                  hit.target_location[0], hit.target_location[1], hit.score, hit.matches, hit.mismatches,
                  n - hit.matches - hit.mismatches, n, hit.score / n, hit.sequence_info.original_length,
                  hit.target_info.original_length, hit.score / hit.sequence_info.original_length,
                  hit.score / hit.target_info.original_length, hit.distance]
        return '\n'.join([', '.join(str(f) for f in fields), hit.sequence_match, hit.alignment, hit.target_match])

    def format_lines(self):
        return [self._format_hit(hit) for hit in self._get_hits()]

    def print_results(self):
        with open(self.outputfile, 'w') as output:
            for line in self.format_lines():
                output.write(line + '\n')


class SamFormatter(DefaultFormatter):

    def __init__(self, logger, hitlist, outputfile):
        DefaultFormatter.__init__(self, logger, hitlist, outputfile)
        self.sq_lines = {}
        self.record_lines = []

    def _set_name(self):
        self.name = 'SAM formatter'

    def _format_hit(self, hit):
        target = hit.get_target_id()
        if target not in self.sq_lines:
            key = target if target[-2:] != 'RC' else target[:-3]
            self.sq_lines[key] = hit.get_sam_sq()
        self.record_lines.append(hit.get_sam_line())

    def format_lines(self):
        for hit in self._get_hits():
            self._format_hit(hit)
        return (['@HD\tVN:1.4\tSO:unknown'] + list(self.sq_lines.values()) + ['@PG\tID:0\tPN:paswas\tVN:3.0']
                + self.record_lines)


class TrimmerFormatter(DefaultFormatter):

    def __init__(self, logger, hitlist, outputfile):
        DefaultFormatter.__init__(self, logger, hitlist, outputfile)
        self.record_lines = []

    def _set_name(self):
        self.name = 'SAM formatter'

    def _format_hit(self, hit):
        self.record_lines.append(hit.get_trimmed_line())

    def format_lines(self):
        for hit in self._get_hits():
            self._format_hit(hit)
        return self.record_lines


class FASTA(DefaultFormatter):

    def __init__(self, logger, hitlist, outputfile):
        DefaultFormatter.__init__(self, logger, hitlist, outputfile)
        self.record_lines = []

    def _set_name(self):
        self.name = 'FASTA formatter'

    def _format_hit(self, hit):
        self.record_lines.append(hit.get_full_fasta())

    def format_lines(self):
        for hit in self._get_hits():
            self._format_hit(hit)
        return self.record_lines
