"""Sequence file readers (role of pyPaSWAS/Core/Readers.py:6-118) without Biopython.

Supported file types: ``fasta`` (id = first word of the header line),
``fastq`` (four-line records), ``genbank``/``gb`` (id = VERSION
accession.version, else ACCESSION, else LOCUS name; sequence from ORIGIN).
Records are uppercased, optionally filtered by length, and sorted by length,
longest first (stable).  Nucleotide targets get their reverse complements
appended with the id suffix ``_RC``.
"""
import itertools
import os.path

from pypaswas_amd.Core.Exceptions import InvalidOptionException, ReaderException
from pypaswas_amd.Core.SWSeq import Seq
from pypaswas_amd.Core.SWSeqRecord import SWSeqRecord


def _parse_fasta(handle):
    ident, chunks = None, []
    for line in handle:
        if line.startswith('>'):
            if ident is not None:
                yield ident, ''.join(chunks)
            title = line[1:].rstrip()
            ident = title.split(None, 1)[0] if title.split(None, 1) else ''
            chunks = []
        elif ident is not None:
            chunks.append(line.strip().replace(' ', '').replace('\r', ''))
    if ident is not None:
        yield ident, ''.join(chunks)


def _parse_fastq(handle):
    while True:
        header = handle.readline()
        if not header:
            return
        if not header.strip():
            continue
        if not header.startswith('@'):
            raise InvalidOptionException('FASTQ record does not start with @: {!r}'.format(header[:40]))
        title = header[1:].rstrip()
        ident = title.split(None, 1)[0] if title.split(None, 1) else ''
        seq_lines = []
        line = handle.readline()
        while line and not line.startswith('+'):
            seq_lines.append(line.strip())
            line = handle.readline()
        seq = ''.join(seq_lines)
        qual = ''
        while len(qual) < len(seq):
            q = handle.readline()
            if not q:
                break
            qual += q.strip()
        yield ident, seq


def _parse_genbank(handle):
    ident, locus, accession, seq, in_origin = None, None, None, [], False
    for line in handle:
        if line.startswith('LOCUS'):
            parts = line.split()
            locus = parts[1] if len(parts) > 1 else ''
            ident, accession, seq, in_origin = None, None, [], False
        elif line.startswith('ACCESSION'):
            parts = line.split()
            accession = parts[1] if len(parts) > 1 else None
        elif line.startswith('VERSION'):
            parts = line.split()
            ident = parts[1] if len(parts) > 1 else None
        elif line.startswith('ORIGIN'):
            in_origin = True
        elif line.startswith('//'):
            yield (ident or accession or locus or ''), ''.join(seq)
            in_origin = False
            seq = []
        elif in_origin:
            seq.append(''.join(ch for ch in line if ch.isalpha()))


_PARSERS = {'fasta': _parse_fasta, 'fa': _parse_fasta, 'fastq': _parse_fastq, 'fastq-sanger': _parse_fastq,
            'genbank': _parse_genbank, 'gb': _parse_genbank}


class Reader(object):

    def __init__(self, logger, path, filetype, limitlength=5000):
        self.logger = logger
        self.rc_string = "_RC"
        self.path = ''
        self.filetype = filetype
        self.records = None
        self._set_path(path)
        self._set_limit_length(limitlength)

    def sort_records(self, reverse=True):
        self.records.sort(key=lambda record: len(record.seq), reverse=reverse)

    @staticmethod
    def _is_a_readable_file(path):
        return os.path.isfile(path) and os.access(path, os.R_OK)

    def _set_path(self, path):
        if not self._is_a_readable_file(path):
            raise InvalidOptionException('{0} is not a file or the program is not allowed to access it.'.format(path))
        self.path = path

    def _set_limit_length(self, limit_length):
        try:
            self.limitlength = int(limit_length)
        except ValueError:
            raise InvalidOptionException('Limitlength should be an int but is {0}'.format(limit_length))

    def get_records(self):
        return self.records

    def complement_records(self):
        self.records.extend([SWSeqRecord(Seq(str(r.seq)).reverse_complement(), identifier=str(r.id) + self.rc_string)
                             for r in self.records])

    def complement_records_only(self):
        self.records = [SWSeqRecord(Seq(str(r.seq)).reverse_complement(), identifier=str(r.id) + self.rc_string)
                        for r in self.records]

    def reverse_records(self):
        self.records.extend([SWSeqRecord(Seq(str(r.seq)[::-1]), identifier=str(r.id) + self.rc_string)
                             for r in self.records])


class BioPythonReader(Reader):
    """Keeps the reference's class name so callers need no change."""

    def read_records(self, start=0, end=None):
        parser = _PARSERS.get(str(self.filetype).lower())
        if parser is None:
            raise InvalidOptionException('Unsupported file type {}'.format(self.filetype))
        with open(self.path, 'r') as handle:
            raw = list(itertools.islice(parser(handle), start, end))
        if len(raw) == 0:
            raise ReaderException('No (more) sequence data found in input file ({}), of file type {}.'.format(
                self.path, self.filetype))
        if self.limitlength > 0:
            n = len(raw)
            self.records = [SWSeqRecord(Seq(s), identifier=i) for i, s in raw if 0 < len(s) <= self.limitlength]
            if n - len(self.records) > 0:
                self.logger.info('{} sequence(s) removed with length > limit_length ({}bp)'.format(
                    n - len(self.records), self.limitlength))
        else:
            self.records = [SWSeqRecord(Seq(s), identifier=i) for i, s in raw]
