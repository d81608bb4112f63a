"""The pyPaSWAS application: settings in, aligned hits out in the chosen format.

Behaviour kept from pyPaSWAS/pypaswasall.py:20-266 (the class name, the
option names and every error message, so that scripts and tests written
against the reference need no change):

* the output path is ``out_file``, or ``output`` when that is empty or ``.``;
  an existing file is an error unless ``override_output`` is ``T``, and a
  missing parent directory is created (:78-101);
* the substitution matrix, the output format and the program are chosen by
  name from fixed tables, an unknown name raising InvalidOptionException
  (:127-171);
* the inputs are read in windows of records: query windows of ``query_step``
  records from ``start_query``, each against the target windows of
  ``sequence_step`` records from ``start_target`` (:173-245).  The window
  rules, including the ones that end the target loop early, are those of the
  reference and are spelled out in ``first_window`` and ``Pypaswas._targets``;
* the hits of every window are gathered in one HitList and written once at
  the end (:247-262).

Output to a Neo4j graph (``GRAPH``, Core/GraphFormatter.py) has no counterpart
here and is rejected when the format is chosen, before any alignment runs.
"""
import logging
import os.path

from pypaswas_amd.cli import DEFAULTS, normalize_file_path, parse_cli, set_logger
from pypaswas_amd.Core.Exceptions import InvalidOptionException, ReaderException
from pypaswas_amd.Core.Formatters import FASTA, DefaultFormatter, SamFormatter, TrimmerFormatter
from pypaswas_amd.Core.HitList import HitList
from pypaswas_amd.Core.Programs import Aligner, Trimmer
from pypaswas_amd.Core.Readers import BioPythonReader
from pypaswas_amd.Core.Scores import (BasicScore, Blosum62Score, Blosum80Score, CustomScore, DnaRnaScore,
                                      IrysScore)

# matrix_name (upper case) -> score class
SCORES = {'DNA-RNA': DnaRnaScore, 'BASIC': BasicScore, 'BLOSUM62': Blosum62Score, 'BLOSUM80': Blosum80Score,
          'CUSTOM': CustomScore, 'IRYS': IrysScore}
# out_format (upper case) -> (the format's name, its formatter class)
FORMATS = {'TXT': ('TXT', DefaultFormatter), 'SAM': ('SAM', SamFormatter),
           'TRIMMEDFASTA': ('trimmedFasta', TrimmerFormatter), 'FASTA': ('FASTA', FASTA)}
PROGRAMS = {'aligner': Aligner, 'trimmer': Trimmer}
# what the reader does to the target records of each score type before sorting
TARGET_TRANSFORMS = {'DNA_RNA': 'complement_records', 'PALINDROME': 'complement_records_only',
                     'IRYS': 'reverse_records'}


def output_path(settings):
    """The checked output path of ``settings`` (its parent directory created if missing)."""
    path = settings.out_file
    path = normalize_file_path('output' if path in (None, '', '.') else path)
    if os.path.exists(path) and settings.override_output != "T":
        raise IOError('File {0} already exists.'.format(path))
    parent = os.path.split(path)[0]
    if not os.path.isdir(parent):
        try:
            os.makedirs(parent)
        except OSError:
            raise IOError('Cannot create direcory {0}.'.format(parent))
    elif not os.access(parent, os.W_OK):
        raise IOError('Cannot write to {0}.'.format(parent))
    return path


def first_window(start, step, end):
    """The first window [lo, hi) of records: ``step`` records from ``start``,
    cut back to ``end`` when it reaches past it, unless ``start == end`` (the
    default 0 / 0, which means no end)."""
    hi = start + step
    if hi > end and start != end:
        hi = end
    return start, hi


class Pypaswas(object):
    """Reads the settings and the two sequence files, runs the chosen program
    over them window by window and writes the hits.

    ``config`` is the defaults file (the package's defaults.cfg when None) and
    ``argv`` the command line (sys.argv when None)."""

    def __init__(self, config=None, argv=None):
        self.config_file = DEFAULTS if config is None else config
        self.argv = argv
        self.logger = logging.getLogger('pypaswas_amd')
        self.settings = None
        self.arguments = None
        self.outputfile = None
        self.score = None
        self.output_format = None
        self.program = None
        self.results = None

    # -- set-up ---------------------------------------------------------

    def _choose(self, table, key, message):
        if key not in table:
            raise InvalidOptionException(message)
        return table[key]

    def _configure(self):
        self.settings, self.arguments = parse_cli(self.config_file, self.argv)
        self.logger = set_logger(self.settings)
        s = self.settings
        self.outputfile = output_path(s)
        matrix = s.matrix_name.upper()
        self.score = self._choose(SCORES, matrix, matrix + ' is not a valid substitution matrix')(self.logger, s)
        if s.out_format.upper() == 'GRAPH':
            raise InvalidOptionException('GRAPH output (Neo4j) is not part of this framework')
        self.output_format, self._formatter = self._choose(
            FORMATS, s.out_format.upper(), 'Invalid output format {0}.'.format(s.out_format))
        self.program = self._choose(PROGRAMS, s.program, 'Invalid program selected {0}'.format(s.program))(
            self.logger, self.score, s)

    # -- input ----------------------------------------------------------

    def _read(self, path, filetype, lo, hi, transform=None):
        """Records lo..hi-1 of a sequence file, optionally transformed, sorted
        by the reader.  Raises ReaderException when the window is empty."""
        path = normalize_file_path(path)
        if not os.path.exists(path):
            raise InvalidOptionException('File {0} does not exist.'.format(path))
        reader = BioPythonReader(self.logger, path, filetype, self.settings.limit_length)
        reader.read_records(lo, hi)
        if transform:
            getattr(reader, transform)()
        reader.sort_records()
        return reader.get_records()

    def _queries(self):
        """The query windows, in order.  The loop ends at the first window the
        reader finds empty, or after the window whose end passes end_query."""
        s = self.settings
        step, end = int(s.query_step), int(s.end_query)
        lo, hi = first_window(int(s.start_query), step, end)
        while True:
            try:
                records = self._read(self.arguments[0], s.filetype1, lo, hi)
            except ReaderException:
                return
            yield records
            if 0 < end < hi:
                return
            lo, hi = lo + step, hi + step

    def _targets(self):
        """The target windows for one query window.  The loop ends at a window
        the reader finds empty, after a window that holds no more records
        (reverse complements included) than its end index, or when the next
        window's end passes end_target."""
        s = self.settings
        step, end = int(s.sequence_step), int(s.end_target)
        lo, hi = first_window(int(s.start_target), step, end)
        transform = TARGET_TRANSFORMS.get(self.score.score_type)
        while True:
            try:
                records = self._read(self.arguments[1], s.filetype2, lo, hi, transform)
            except ReaderException:
                return
            if not records:
                return
            yield records
            if len(records) <= hi:
                return
            lo, hi = lo + step, hi + step
            if 0 < end < hi:
                return

    # -- run ------------------------------------------------------------

    def run(self):
        """Runs the program; returns the HitList that was written."""
        self._configure()
        results = HitList(self.logger)
        for queries in self._queries():
            for targets in self._targets():
                if not queries:
                    break
                self.logger.info('Processing {0}- vs {1}-sequences'.format(len(queries), len(targets)))
                results.extend(self.program.process(queries, targets, self))
        self.results = results
        self.logger.info('Processing OK ({} hits found).'.format(len(results.hits)))
        if results.hits:
            self._formatter(self.logger, results, self.outputfile).print_results()
        else:
            self.logger.warning('No suitable hits produced, exiting...')
        return results
