"""Sequence record (fields of pyPaSWAS/Core/SWSeqRecord.py:7-34)."""
from pypaswas_amd.Core.SWSeq import Seq


class SWSeqRecord(object):
    start_position = 0
    distance = 0
    original_length = 0

    def __init__(self, seq, identifier, start_position=0, original_length=None, distance=0, refID=None):
        self.seq = Seq(str(seq)).upper()
        self.id = identifier
        self.name = identifier
        self.description = ''
        self.start_position = start_position
        self.original_length = len(seq) if original_length is None else original_length
        self.distance = distance
        self.refID = refID

    def __len__(self):
        return len(self.seq)

    def __getitem__(self, index):
        return self.seq[index]

    def __repr__(self):
        return 'SWSeqRecord(id={!r}, len={})'.format(self.id, len(self.seq))
