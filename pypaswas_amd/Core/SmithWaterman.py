"""Engine interface (pyPaSWAS/Core/SmithWaterman.py:27-848).

The programs drive an engine through ``set_total_work_size``,
``set_targets`` and ``align_sequences``; a device backend implements the
hook methods.  This base class keeps the reference's settings handling and
its hook names; SmithWatermanHIP is the only backend of this framework.
"""
import math
import time

from pypaswas_amd.Core import NO_DIRECTION, STOP_DIRECTION
from pypaswas_amd.Core.Exceptions import InvalidOptionException
from pypaswas_amd.Core.StartingPoint import StartingPoint


class SmithWaterman(object):
    float_size = 4
    int_size = 4

    GAP_CHAR_SEQ = '-'
    GAP_CHAR_ALIGN = ' '
    MISMATCH_CHAR = '.'
    MATCH_CHAR = '|'

    def __init__(self, logger, score, settings):
        self.logger = logger
        self.score = score
        self.settings = settings
        self.number_of_targets = 1
        self.number_targets = 0
        self.number_of_sequences = 0
        self.starting_point = StartingPoint(self.logger)
        self.size_of_startingpoint = self.starting_point.size
        # starting points kept per pair of a batch (SmithWaterman.py:144, 683-685)
        self.maximum_number_starting_points = 100
        self.gap_extension = score.gap_extension is not None
        if self.gap_extension:
            self.logger.info("Gap extension penalty detected: using affine gap scoring algorithm")
        else:
            self.logger.info("No gap extension penalty detected: using original PaSWAS scoring algorithm")
        self.shared_x = 8
        self.shared_y = 8
        self.device = 0
        self._set_device(self.settings.device_number)
        self.filter_factor = 0.7
        self.internal_limit = 64000
        self.max_genome_length = 10000
        self._set_filter_factor(self.settings.filter_factor)
        self._set_max_genome_length(self.settings.max_genome_length)
        try:
            self.mem_fill_factor = float(settings.maximum_memory_usage)
        except ValueError:
            raise InvalidOptionException('maximux_memory_usage is not a float'.format(settings.maximum_memory_usage))
        if self.mem_fill_factor > 1.0 or self.mem_fill_factor <= 0.0:
            raise InvalidOptionException('maximux_memory_usage is not a float between 0.0 and 1.0'.format(
                settings.maximum_memory_usage))
        self.total_work_size = 0
        self.total_processed = 0
        self.start_time = time.time()

    # --- settings -------------------------------------------------------------------
    def _set_device(self, device):
        try:
            self.device = int(device) if device else self.device
        except ValueError:
            raise InvalidOptionException('device should be an int but is {0}'.format(device))

    def _set_max_genome_length(self, max_genome_length):
        try:
            self.max_genome_length = int(max_genome_length) if max_genome_length else self.max_genome_length
        except ValueError:
            raise InvalidOptionException('maxGenomeLength should be a int but is {0}'.format(max_genome_length))

    def _set_filter_factor(self, filterfactor):
        try:
            if filterfactor:
                self.filter_factor = float(filterfactor)
        except ValueError:
            raise InvalidOptionException('Filterfactor should be a float but is {0}'.format(filterfactor))

    # --- hooks a backend implements ---------------------------------------------------
    def _initialize_device(self, device_number):
        pass

    def _device_global_mem_size(self):
        pass

    def _clear_memory(self):
        pass

    def _init_memory(self):
        pass

    def _init_zero_copy(self):
        pass

    def _compile_code(self):
        pass

    def copy_sequences(self, h_sequences, h_targets):
        pass

    def _execute_calculate_score_kernel(self, number_of_blocks, idx, idy):
        pass

    def _execute_traceback_kernel(self, number_of_blocks, idx, idy):
        pass

    def _get_number_of_starting_points(self):
        pass

    def _set_max_possible_score(self, target_index, targets, i, index, records_seqs):
        pass

    def _get_starting_point_byte_array(self, number_of_starting_points):
        pass

    def _get_direction_byte_array(self):
        pass

    def _is_traceback_required(self):
        return True

    # --- engine surface ---------------------------------------------------------------
    def set_total_work_size(self, size):
        """Total cells of the run; resets the progress report."""
        self.total_work_size = size
        self.total_processed = 0
        self.start_time = time.time()

    def set_targets(self, targets, target_index, max_length=None, records_seqs=None, use_all_records_seqs=True):
        """Selects the block of targets the next align_sequences call processes; returns the next index."""
        self.number_of_targets = max(1, min(self._targets_per_block(targets, target_index, records_seqs),
                                            len(targets) - target_index))
        self.number_targets = self.number_of_targets
        return target_index + self.number_of_targets

    def _targets_per_block(self, targets, target_index, records_seqs):
        return len(targets) - target_index

    def align_sequences(self, records_seqs, targets, target_index):
        raise NotImplementedError

    def _is_in_alignment(self, show, block_x, block_y, value_x, value_y, direction):
        return (show and block_x >= 0 and block_y >= 0 and value_x >= 0 and value_y >= 0
                and direction != STOP_DIRECTION and direction != NO_DIRECTION)

    def _filter_hit(self, hit):
        """SmithWaterman.py:843-848."""
        return (hit.relative_score >= float(self.settings.relative_score) and
                hit.query_identity >= float(self.settings.query_identity) and
                hit.query_coverage >= float(self.settings.query_coverage) and
                hit.base_score >= float(self.settings.base_score))

    def _log_progress(self, processed, duration):
        """The INFO progress line of SmithWaterman.py:544-557."""
        if self.total_work_size <= 0:
            return
        import datetime
        self.total_processed += processed
        total_duration = time.time() - self.start_time
        performance = processed / 1e9 / max(duration, 1e-12)
        avg_performance = self.total_processed / 1e9 / max(total_duration, 1e-12)
        progress = self.total_processed / float(self.total_work_size)
        eta = total_duration / progress * (1.0 - progress) if progress > 0 else 0.0
        self.logger.info("Duration: {:5.3f} | Total: {} | Performance: {:5.2f} GCUPS | Avg: {:5.2f} GCUPS | "
                         "Progress: {:7.3%} | ETA: {}".format(duration, datetime.timedelta(seconds=round(total_duration)),
                                                             performance, avg_performance, progress,
                                                             datetime.timedelta(seconds=round(eta))))

    @staticmethod
    def _ceil_to(n, k):
        return int(math.ceil(n / float(k)) * k)
