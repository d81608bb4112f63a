"""Programs of the suite: the Aligner/Trimmer plugin contract (pyPaSWAS/Core/Programs.py:9-120).

``Aligner(logger, score, settings)`` picks the engine from
``settings.framework``; ``process(records_seqs, targets, pypaswas)`` returns
a HitList.  ``--framework=HIP`` selects SmithWatermanHIP.  The reference's
OpenCL and CUDA engines are not part of this framework: asking for them (or
for an unknown framework) logs the fact and uses the HIP engine, the only
device engine here.
"""
from pypaswas_amd.Core.HitList import HitList


def make_engine(logger, score, settings):
    framework = str(getattr(settings, 'framework', 'HIP')).upper()
    if framework != 'HIP':
        logger.info('Framework {} is not available in this build; using the HIP implementation'.format(framework))
    else:
        logger.debug('Using HIP implementation')
    from pypaswas_amd.Core.SmithWatermanHIP import SmithWatermanHIP
    return SmithWatermanHIP(logger, score, settings)


class Aligner(object):

    def __init__(self, logger, score, settings):
        self.logger = logger
        self.score = score
        self.hitlist = HitList(self.logger)
        self.settings = settings
        self.smith_waterman = make_engine(logger, score, settings)

    def process(self, records_seqs, targets, pypaswas):
        target_index = 0
        all_targets_length = sum(len(s.seq) for s in targets)
        all_sequences_length = sum(len(s.seq) for s in records_seqs)
        self.smith_waterman.set_total_work_size(all_targets_length * all_sequences_length)
        while target_index < len(targets):
            last_target_index = self.smith_waterman.set_targets(targets, target_index, records_seqs=records_seqs,
                                                                use_all_records_seqs=False)
            results = self.smith_waterman.align_sequences(records_seqs, targets, target_index)
            self.hitlist.extend(results)
            target_index = last_target_index
        return self.hitlist


class Trimmer(Aligner):

    def process(self, records_seqs, targets, pypaswas):
        target_index = 0
        max_length = len(targets[0]) if len(targets) > 0 and self.settings.recompile == "F" else None
        all_targets_length = sum(len(s.seq) for s in targets)
        all_sequences_length = sum(len(s.seq) for s in records_seqs)
        self.smith_waterman.set_total_work_size(all_targets_length * all_sequences_length)
        while target_index < len(targets):
            last_target_index = self.smith_waterman.set_targets(targets, target_index, max_length, records_seqs)
            results = self.smith_waterman.align_sequences(records_seqs, targets, target_index)
            self.hitlist.extend(results)
            target_index = last_target_index
        return self.hitlist
