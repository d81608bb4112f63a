"""Application driver (pyPaSWAS/pypaswasall.py:20-266): read, align, format."""
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


class Pypaswas(object):

    def __init__(self, config=None, argv=None):
        self.outputfile = None
        self.score = None
        self.output_format = None
        self.program = None
        self.config_file = DEFAULTS if config is None else config
        self.argv = argv
        self.logger = logging.getLogger('pypaswas_amd')
        self.settings = None
        self.arguments = None
        self.results = None

    def _get_formatter(self, results):
        if self.output_format == 'SAM':
            return SamFormatter(self.logger, results, self.outputfile)
        if self.output_format == 'trimmedFasta':
            return TrimmerFormatter(self.logger, results, self.outputfile)
        if self.output_format == 'FASTA':
            return FASTA(self.logger, results, self.outputfile)
        if self.output_format == 'GRAPH':
            raise InvalidOptionException('GRAPH output (Neo4j) is not part of this framework')
        return DefaultFormatter(self.logger, results, self.outputfile)

    def _set_outfile(self):
        path = self.settings.out_file
        if path in (None, '', '.'):
            path = 'output'
        path = normalize_file_path(path)
        if self.settings.override_output != "T" and os.path.exists(path):
            raise IOError('File {0} already exists.'.format(path))
        destination_dir = os.path.split(path)[0]
        if os.path.isdir(destination_dir):
            if not os.access(destination_dir, os.W_OK):
                raise IOError('Cannot write to {0}.'.format(destination_dir))
        else:
            try:
                os.makedirs(destination_dir)
            except OSError:
                raise IOError('Cannot create direcory {0}.'.format(destination_dir))
        self.outputfile = path

    def _get_query_sequences(self, queryfile, start=0, end=None):
        queryfile = normalize_file_path(queryfile)
        if not os.path.exists(queryfile):
            raise InvalidOptionException('File {0} does not exist.'.format(queryfile))
        reader = BioPythonReader(self.logger, queryfile, self.settings.filetype1, self.settings.limit_length)
        reader.read_records(start, end)
        reader.sort_records()
        return reader.get_records()

    def _get_target_sequences(self, databasefile, start=0, end=None):
        databasefile = normalize_file_path(databasefile)
        if not os.path.exists(databasefile):
            raise InvalidOptionException('File {0} does not exist.'.format(databasefile))
        reader = BioPythonReader(self.logger, databasefile, self.settings.filetype2, self.settings.limit_length)
        reader.read_records(start, end)
        if self.score.score_type == 'DNA_RNA':
            reader.complement_records()
        elif self.score.score_type == 'PALINDROME':
            reader.complement_records_only()
        elif self.score.score_type == 'IRYS':
            reader.reverse_records()
        reader.sort_records()
        return reader.get_records()

    def _set_scoring_matrix(self):
        name = self.settings.matrix_name.upper()
        classes = {'DNA-RNA': DnaRnaScore, 'BASIC': BasicScore, 'BLOSUM62': Blosum62Score,
                   'BLOSUM80': Blosum80Score, 'CUSTOM': CustomScore, 'IRYS': IrysScore}
        if name not in classes:
            raise InvalidOptionException(name + ' is not a valid substitution matrix')
        self.score = classes[name](self.logger, self.settings)

    def _set_output_format(self):
        fmt = self.settings.out_format.upper()
        table = {'SAM': 'SAM', 'TXT': 'TXT', 'TRIMMEDFASTA': 'trimmedFasta', 'FASTA': 'FASTA', 'GRAPH': 'GRAPH'}
        if fmt not in table:
            raise InvalidOptionException('Invalid output format {0}.'.format(self.settings.out_format))
        self.output_format = table[fmt]

    def _set_program(self):
        if self.settings.program == 'aligner':
            self.program = Aligner(self.logger, self.score, self.settings)
        elif self.settings.program == 'trimmer':
            self.program = Trimmer(self.logger, self.score, self.settings)
        else:
            raise InvalidOptionException('Invalid program selected {0}'.format(self.settings.program))

    def run(self):
        self.settings, self.arguments = parse_cli(self.config_file, self.argv)
        self.logger = set_logger(self.settings)
        self._set_outfile()
        self._set_scoring_matrix()
        self._set_output_format()
        self._set_program()

        s = self.settings
        queries_to_process = True
        query_start = int(s.start_query)
        query_end = int(s.start_query) + int(s.query_step)
        if query_end > int(s.end_query) and int(s.start_query) != int(s.end_query):
            query_end = int(s.end_query)
        results = HitList(self.logger)
        while queries_to_process:
            try:
                query_sequences = self._get_query_sequences(self.arguments[0], start=query_start, end=query_end)
            except ReaderException:
                queries_to_process = False
            sequences_to_process = True
            start_index = int(s.start_target)
            end_index = int(s.start_target) + int(s.sequence_step)
            if end_index > int(s.end_target) and int(s.start_target) != int(s.end_target):
                end_index = int(s.end_target)
            while queries_to_process and sequences_to_process:
                try:
                    target_sequences = self._get_target_sequences(self.arguments[1], start=start_index, end=end_index)
                except ReaderException:
                    sequences_to_process = False
                if not sequences_to_process or not queries_to_process or len(query_sequences) == 0 \
                        or len(target_sequences) == 0:
                    sequences_to_process = False
                else:
                    results.extend(self.program.process(query_sequences, target_sequences, self))
                if sequences_to_process and len(target_sequences) <= end_index:
                    sequences_to_process = False
                start_index += int(s.sequence_step)
                end_index += int(s.sequence_step)
                if 0 < int(s.end_target) < end_index:
                    sequences_to_process = False
            if 0 < int(s.end_query) < query_end:
                queries_to_process = False
            query_start += int(s.query_step)
            query_end += int(s.query_step)
        self.results = results
        if len(results.hits) > 0:
            self._get_formatter(results).print_results()
        else:
            self.logger.warning('No suitable hits produced, exiting...')
        return results
