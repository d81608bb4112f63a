"""36-byte starting-point records (pyPaSWAS/Core/StartingPoint.py:5-93).

Layout, little-endian: sequence, target, block_x, block_y, value_x, value_y
as uint32, then score, max_score, pos_score as float32.
"""
import struct

_REC = struct.Struct('<6I3f')


class StartingPoint(object):
    size = 9 * 4

    def __init__(self, logger=None):
        self.logger = logger
        self.sequence = 0
        self.target = 0
        self.block_x = 0
        self.block_y = 0
        self.value_x = 0
        self.value_y = 0
        self.score = 0.0
        self.max_score = 0.0
        self.pos_score = 0.0
        self.sequence_index = 0
        self.target_index = 0

    def parse_byte_string(self, byte_string, location):
        """Sets the fields from record number ``location`` of ``byte_string``."""
        (self.sequence, self.target, self.block_x, self.block_y, self.value_x, self.value_y, self.score,
         self.max_score, self.pos_score) = _REC.unpack_from(bytes(byte_string[location * self.size:
                                                                               (location + 1) * self.size]))
        # the score is a Python float holding the f32 value exactly
        self.score = float(self.score)

    def __str__(self):
        return ("Seq: {0}\nTar: {1}\nblockX: {2}\nblockY:{3}\nvalueX: {4}\nvalueY: {5}\nscore: {6}\nmaxScore: {7}"
                "\nposScore: {8}".format(self.sequence, self.target, self.block_x, self.block_y, self.value_x,
                                         self.value_y, self.score, self.max_score, self.pos_score))
