"""Substitution models (interface of pyPaSWAS/Core/Scores.py:5-402).

Each Score exposes the fields the engine reads: ``matrix`` (dimensions x
dimensions, values as given on the command line or as published), ``dimensions``,
``char_offset``, ``gap_score``, ``gap_extension`` (None selects the linear
model), ``lower_limit_score``, ``minimum_score``, ``highest_score`` and
``score_type``.  ``to_float32()`` gives the FP32 table the kernels use: every
entry goes through ``float(str(entry))``, the value a C ``float`` initialised
from the decimal literal ``str(entry)`` holds (Scores.py:87-98 writes the
matrix into the kernel source that way).
"""
import numpy as np

from pypaswas_amd.Core.Exceptions import InvalidOptionException

# BLOSUM62 over 'A'..'Z' (B, J, O, U, X, Z rows as used by pyPaSWAS, Scores.py:216-241)
_BLOSUM62 = """
 4 -2  0 -2 -1 -2  0 -2 -1 -1 -1 -1 -1 -2 -4 -1 -1 -1  1  0 -4  0 -3 -1 -2 -1
-2  4 -3  4  1 -3 -1  0 -3 -3  0 -4 -3  4 -4 -2  0 -1  0 -1 -4 -3 -4 -1 -3  0
 0 -3  9 -3 -4 -2 -3 -3 -1 -1 -3 -1 -1 -3 -4 -3 -3 -3 -1 -1 -4 -1 -2 -1 -2 -3
-2  4 -3  6  2 -3 -1 -1 -3 -3 -1 -4 -3  1 -4 -1  0 -2  0 -1 -4 -3 -4 -1 -3  1
-1  1 -4  2  5 -3 -2  0 -3 -3  1 -3 -2  0 -4 -1  2  0  0 -1 -4 -2 -3 -1 -2  4
-2 -3 -2 -3 -3  6 -3 -1  0  0 -3  0  0 -3 -4 -4 -3 -3 -2 -2 -4 -1  1 -1  3 -3
 0 -1 -3 -1 -2 -3  6 -2 -4 -4 -2 -4 -3  0 -4 -2 -2 -2  0 -2 -4 -3 -2 -1 -3 -2
-2  0 -3 -1  0 -1 -2  8 -3 -3 -1 -3 -2  1 -4 -2  0  0 -1 -2 -4 -3 -2 -1  2  0
-1 -3 -1 -3 -3  0 -4 -3  4  3 -3  2  1 -3 -4 -3 -3 -3 -2 -1 -4  3 -3 -1 -1 -3
-1 -3 -1 -3 -3  0 -4 -3  3  3 -3  3  2 -3 -4 -3 -2 -2 -2 -1 -4  2 -2 -1 -1 -3
-1  0 -3 -1  1 -3 -2 -1 -3 -3  5 -2 -1  0 -4 -1  1  2  0 -1 -4 -2 -3 -1 -2  1
-1 -4 -1 -4 -3  0 -4 -3  2  3 -2  4  2 -3 -4 -3 -2 -2 -2 -1 -4  1 -2 -1 -1 -3
-1 -3 -1 -3 -2  0 -3 -2  1  2 -1  2  5 -2 -4 -2  0 -1 -1 -1 -4  1 -1 -1 -1 -1
-2  4 -3  1  0 -3  0  1 -3 -3  0 -3 -2  6 -4 -2  0  0  1  0 -4 -3 -4 -1 -2  0
-4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4
-1 -2 -3 -1 -1 -4 -2 -2 -3 -3 -1 -3 -2 -2 -4  7 -1 -2 -1 -1 -4 -2 -4 -1 -3 -1
-1  0 -3  0  2 -3 -2  0 -3 -2  1 -2  0  0 -4 -1  5  1  0 -1 -4 -2 -2 -1 -1  4
-1 -1 -3 -2  0 -3 -2  0 -3 -2  2 -2 -1  0 -4 -2  1  5 -1 -1 -4 -3 -3 -1 -2  0
 1  0 -1  0  0 -2  0 -1 -2 -2  0 -2 -1  1 -4 -1  0 -1  4  1 -4 -2 -3 -1 -2  0
 0 -1 -1 -1 -1 -2 -2 -2 -1 -1 -1 -1 -1  0 -4 -1 -1 -1  1  5 -4  0 -2 -1 -2 -1
-4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4
 0 -3 -1 -3 -2 -1 -3 -3  3  2 -2  1  1 -3 -4 -2 -2 -3 -2  0 -4  4 -3 -1 -1 -2
-3 -4 -2 -4 -3  1 -2 -2 -3 -2 -3 -2 -1 -4 -4 -4 -2 -3 -3 -2 -4 -3 11 -1  2 -2
-1 -1 -1 -1 -1 -1 -1 -1 -1 -1 -1 -1 -1 -1 -4 -1 -1 -1 -1 -1 -4 -1 -1 -1 -1 -1
-2 -3 -2 -3 -2  3 -3  2 -1 -1 -2 -1 -1 -2 -4 -3 -1 -2 -2 -2 -4 -1  2 -1  7 -2
-1  0 -3  1  4 -3 -2  0 -3 -3  1 -3 -1  0 -4 -1  4  0  0 -1 -4 -2 -2 -1 -2  4
"""

# BLOSUM80 over 'A'..'Z' (Scores.py:164-189)
_BLOSUM80 = """
 7 -3 -1 -3 -2 -4  0 -3 -3 -8 -1 -3 -2 -3 -8 -1 -2 -3  2  0 -8 -1 -5 -1 -4 -2
-3  6 -6  6  1 -6 -2 -1 -6 -8 -1 -7 -5  5 -8 -4 -1 -2  0 -1 -8 -6 -8 -3 -5  0
-1 -6 13 -7 -7 -4 -6 -7 -2 -8 -6 -3 -3 -5 -8 -6 -5 -6 -2 -2 -8 -2 -5 -4 -5 -7
-3  6 -7 10  2 -6 -3 -2 -7 -8 -2 -7 -6  2 -8 -3 -1 -3 -1 -2 -8 -6 -8 -3 -6  1
-2  1 -7  2  8 -6 -4  0 -6 -8  1 -6 -4 -1 -8 -2  3 -1 -1 -2 -8 -4 -6 -2 -5  6
-4 -6 -4 -6 -6 10 -6 -2 -1 -8 -5  0  0 -6 -8 -6 -5 -5 -4 -4 -8 -2  0 -3  4 -6
 0 -2 -6 -3 -4 -6  9 -4 -7 -8 -3 -7 -5 -1 -8 -5 -4 -4 -1 -3 -8 -6 -6 -3 -6 -4
-3 -1 -7 -2  0 -2 -4 12 -6 -8 -1 -5 -4  1 -8 -4  1  0 -2 -3 -8 -5 -4 -2  3  0
-3 -6 -2 -7 -6 -1 -7 -6  7 -8 -5  2  2 -6 -8 -5 -5 -5 -4 -2 -8  4 -5 -2 -3 -6
-8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8
-1 -1 -6 -2  1 -5 -3 -1 -5 -8  8 -4 -3  0 -8 -2  2  3 -1 -1 -8 -4 -6 -2 -4  1
-3 -7 -3 -7 -6  0 -7 -5  2 -8 -4  6  3 -6 -8 -5 -4 -4 -4 -3 -8  1 -4 -2 -2 -5
-2 -5 -3 -6 -4  0 -5 -4  2 -8 -3  3  9 -4 -8 -4 -1 -3 -3 -1 -8  1 -3 -2 -3 -3
-3  5 -5  2 -1 -6 -1  1 -6 -8  0 -6 -4  9 -8 -4  0 -1  1  0 -8 -5 -7 -2 -4 -1
-8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8
-1 -4 -6 -3 -2 -6 -5 -4 -5 -8 -2 -5 -4 -4 -8 12 -3 -3 -2 -3 -8 -4 -7 -3 -6 -2
-2 -1 -5 -1  3 -5 -4  1 -5 -8  2 -4 -1  0 -8 -3  9  1 -1 -1 -8 -4 -4 -2 -3  5
-3 -2 -6 -3 -1 -5 -4  0 -5 -8  3 -4 -3 -1 -8 -3  1  9 -2 -2 -8 -4 -5 -2 -4  0
 2  0 -2 -1 -1 -4 -1 -2 -4 -8 -1 -4 -3  1 -8 -2 -1 -2  7  2 -8 -3 -6 -1 -3 -1
 0 -1 -2 -2 -2 -4 -3 -3 -2 -8 -1 -3 -1  0 -8 -3 -1 -2  2  8 -8  0 -5 -1 -3 -2
-8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8 -8
-1 -6 -2 -6 -4 -2 -6 -5  4 -8 -4  1  1 -5 -8 -4 -4 -4 -3  0 -8  7 -5 -2 -3 -4
-5 -8 -5 -8 -6  0 -6 -4 -5 -8 -6 -4 -3 -7 -8 -7 -4 -5 -6 -5 -8 -5 16 -5  3 -5
-1 -3 -4 -3 -2 -3 -3 -2 -2 -8 -2 -2 -2 -2 -8 -3 -2 -2 -1 -1 -8 -2 -5 -2 -3 -1
-4 -5 -5 -6 -5  4 -6  3 -3 -8 -4 -2 -3 -4 -8 -6 -3 -4 -3 -3 -8 -3  3 -3 11 -4
-2  0 -7  1  6 -6 -4  0 -6 -8  1 -5 -3 -1 -8 -2  5  0 -1 -2 -8 -4 -5 -1 -4  6
"""


def _parse_table(text):
    return [[int(v) for v in line.split()] for line in text.strip().splitlines()]


class Score(object):
    """Common scoring settings (Scores.py:5-98)."""

    def __init__(self, logger, settings):
        self.logger = logger
        self.matrix = None
        self.score_type = 'Unknown'
        self.allowed_errors = 0
        self.dimensions = 255
        self.highest_score = 0
        self.code_string = None
        self.char_offset = '\1'
        self._set_score_type()
        self.set_gap_score(settings.gap_score)
        self.set_gap_extension(settings.gap_extension)
        self.set_lower_limit_score(settings.lower_limit_score)
        self.set_lower_limit_max_score(settings.lower_limit_score)
        self.set_minimum_score(settings.minimum_score)

    def _set_score_type(self):
        pass

    def set_dimensions(self, dimension):
        self.dimensions = dimension

    def set_gap_score(self, gap_score):
        self.gap_score = float(gap_score) if gap_score else None

    def set_gap_extension(self, gap_extension):
        # a zero extension selects the linear model (Scores.py:58)
        self.gap_extension = float(gap_extension) if gap_extension and float(gap_extension) != 0 else None

    def set_lower_limit_max_score(self, value):
        self.lower_limit_max_score = float(value)

    def set_lower_limit_score(self, value):
        llscore = float(value)
        if llscore < 1.0:
            self.logger.warning('Lower limit score is set < 1.0 ({}), this can result in too many hits to '
                                'properly handle.'.format(llscore))
        self.lower_limit_score = llscore

    def set_minimum_score(self, value):
        self.minimum_score = float(value)

    def to_float32(self):
        """The FP32 substitution table, row = query residue index."""
        return np.array([[float(str(v)) for v in row] for row in self.matrix], dtype=np.float32)

    def __str__(self):
        if self.code_string is None:
            rows = ['{' + ','.join(str(v) for v in row) + '}' for row in self.matrix]
            self.code_string = '{' + ','.join(rows) + '}'
        return self.code_string


class BasicScore(Score):
    """match on the diagonal, mismatch elsewhere, 255 x 255 from '\\1' (Scores.py:101-138)."""

    def __init__(self, logger, settings):
        Score.__init__(self, logger, settings)
        self.match_score = settings.match_score
        self.mismatch_score = settings.mismatch_score
        self.matrix = [[self.match_score if r == c else self.mismatch_score for c in range(self.dimensions)]
                       for r in range(self.dimensions)]
        self.highest_score = self.match_score

    def _set_score_type(self):
        self.score_type = 'BASIC'


class Blosum80Score(Score):
    def __init__(self, logger, settings):
        Score.__init__(self, logger, settings)
        self.set_dimensions(26)
        self.char_offset = 'A'
        self.matrix = _parse_table(_BLOSUM80)
        self.highest_score = max(max(self.matrix))

    def _set_score_type(self):
        self.score_type = 'BLOSUM80'


class Blosum62Score(Score):
    def __init__(self, logger, settings):
        Score.__init__(self, logger, settings)
        self.set_dimensions(26)
        self.char_offset = 'A'
        self.matrix = _parse_table(_BLOSUM62)
        # max over the lexicographically largest row, as Scores.py:243 computes it
        self.highest_score = max(max(self.matrix))

    def _set_score_type(self):
        self.score_type = 'BLOSUM62'


class CustomScore(Score):
    """A user matrix of 26 rows of 26 entries over 'A'..'Z' (Scores.py:247-285)."""

    def __init__(self, logger, settings):
        Score.__init__(self, logger, settings)
        matrix = settings.custom_matrix
        if matrix is None:
            raise InvalidOptionException('Required argument {0} is missing'.format('matrix'))
        if isinstance(matrix, str):
            import ast
            matrix = ast.literal_eval(matrix)
        self.char_offset = 'A'
        self.set_dimensions(26)
        if len(matrix) != self.dimensions or (len(matrix) > 0 and len(matrix[0]) != self.dimensions):
            raise InvalidOptionException('A matrix should consist of a list containing 26 lists of 26 integers each.')
        self.matrix = matrix
        self.highest_score = max(max(self.matrix))

    def _set_score_type(self):
        self.score_type = 'CUSTOM'


class DnaRnaScore(Score):
    """IUPAC-aware nucleotide scores over 'A'..'Z' (Scores.py:288-360).

    match on identity and on the pairs K:{G,T}, M:{A,C}, R:{A,G}, Y:{C,T},
    S:{C,G}, W:{A,T}; else ``any`` when either residue is N; else mismatch
    when either is one of A, T, C, G, U; else ``other``.
    """

    AMBIGUOUS = {'K': 'GT', 'M': 'AC', 'R': 'AG', 'Y': 'CT', 'S': 'CG', 'W': 'AT'}

    def __init__(self, logger, settings):
        Score.__init__(self, logger, settings)
        self.match_score = settings.match_score
        self.mismatch_score = settings.mismatch_score
        self.other_score = settings.other_score
        self.any_score = settings.any_score
        self.char_offset = 'A'
        self.set_dimensions(26)
        self._create_matrix()

    def _set_score_type(self):
        self.score_type = 'DNA_RNA'

    def _create_matrix(self):
        letters = [chr(ord('A') + i) for i in range(self.dimensions)]
        pairs = set()
        for code, bases in self.AMBIGUOUS.items():
            for b in bases:
                pairs.add((code, b))
                pairs.add((b, code))
        nucleotides = set('ATCGU')
        self.matrix = []
        for r in letters:
            row = []
            for c in letters:
                if r == c or (r, c) in pairs:
                    row.append(self.match_score)
                elif r == 'N' or c == 'N':
                    row.append(self.any_score)
                elif r in nucleotides or c in nucleotides:
                    row.append(self.mismatch_score)
                else:
                    row.append(self.other_score)
            self.matrix.append(row)
        self.highest_score = self.match_score


class IrysScore(Score):
    """Optical-map label scores: match on the diagonal, mismatch * |row - col| elsewhere (Scores.py:363-402)."""

    def __init__(self, logger, settings):
        Score.__init__(self, logger, settings)
        self.match_score = settings.match_score
        self.mismatch_score = settings.mismatch_score
        self.other_score = settings.other_score
        self.any_score = settings.any_score
        self.char_offset = 'A'
        self.set_dimensions(ord('~') - ord(self.char_offset) + 1)
        n = self.dimensions
        self.matrix = [[self.match_score if r == c else str(float(self.mismatch_score) * abs(r - c))
                        for c in range(n)] for r in range(n)]
        self.highest_score = self.match_score

    def _set_score_type(self):
        self.score_type = 'IRYS'
