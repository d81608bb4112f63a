"""Sequence value type without Biopython (role of pyPaSWAS/Core/SWSeq.py:1-30).

``Seq`` is a minimal immutable residue string with the operations the
engine and the record formatters use: slicing, ``upper()``, ``str()`` and
``reverse_complement()`` for nucleotides.
"""

# IUPAC nucleotide complement, as Biopython's ambiguous DNA complement table
_COMPLEMENT = {
    'A': 'T', 'C': 'G', 'G': 'C', 'T': 'A', 'M': 'K', 'R': 'Y', 'W': 'W', 'S': 'S', 'Y': 'R', 'K': 'M',
    'V': 'B', 'H': 'D', 'D': 'H', 'B': 'V', 'X': 'X', 'N': 'N',
}
_COMPLEMENT.update({k.lower(): v.lower() for k, v in list(_COMPLEMENT.items())})
_TABLE = str.maketrans(_COMPLEMENT)


class Seq(object):
    __slots__ = ('_data',)

    def __init__(self, data, alphabet=None):
        self._data = str(data)

    def __str__(self):
        return self._data

    def __repr__(self):
        return 'Seq({!r})'.format(self._data)

    def __len__(self):
        return len(self._data)

    def __getitem__(self, index):
        if isinstance(index, slice):
            return Seq(self._data[index])
        return self._data[index]

    def __add__(self, other):
        return Seq(self._data + str(other))

    def __eq__(self, other):
        return str(self) == str(other)

    def __hash__(self):
        return hash(self._data)

    def upper(self):
        return Seq(self._data.upper())

    def lower(self):
        return Seq(self._data.lower())

    def complement(self):
        return Seq(self._data.translate(_TABLE))

    def reverse_complement(self):
        return Seq(self._data.translate(_TABLE)[::-1])


class SWSeq(Seq):
    SPECIAL_CHAR = '\0'

    @staticmethod
    def extentToFillGPU(seq, length=None):
        """Uppercase and pad with '\\0' to ``length`` (SWSeq.py:22-30)."""
        if length is None:
            return seq.upper()
        if length < len(seq):
            return seq.upper()[0:length]
        return seq.upper() + SWSeq.SPECIAL_CHAR * (length - len(seq))
