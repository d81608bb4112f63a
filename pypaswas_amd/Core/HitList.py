"""Hit collection with start/end-location dedupe (pyPaSWAS/Core/HitList.py:5-52).

A hit is kept when neither its start key nor its end key is known, or when it
beats the score stored under one of them; a better hit sharing the end key
removes the earlier one from ``real_hits``.
"""


class HitList(object):

    def __init__(self, logger):
        self.logger = logger
        self.hits = {}
        self.real_hits = {}

    def extend(self, hit_list):
        self.hits.update(hit_list.hits)
        self.real_hits.update(hit_list.real_hits)

    def add(self, hit):
        key = hit.keys()[0]
        self.hits[key] = hit
        self.real_hits[key] = hit

    def append(self, hit):
        keys = hit.keys()
        if keys is None:
            return
        start, end, real = keys
        beats_start = start in self.hits and self.hits[start].score < hit.score
        beats_end = end in self.hits and self.hits[end].score < hit.score
        unseen = start not in self.hits and end not in self.hits
        if beats_start or beats_end or unseen:
            if beats_end:
                previous = self.hits[end].keys()[2]
                if previous in self.real_hits:
                    del self.real_hits[previous]
            self.hits[start] = hit
            self.hits[end] = hit
            self.real_hits[real] = hit
            self.logger.info('Added hit (query ID: {}, target ID: {}) to hitlist'.format(hit.get_seq_id(),
                                                                                       hit.get_target_id()))
        else:
            self.logger.debug('Rejected hit (query ID: {}, target ID: {}) from addition to hitlist'.format(
                hit.get_seq_id(), hit.get_target_id()))
