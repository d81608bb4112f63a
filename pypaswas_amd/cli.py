"""Command line and settings (option set of pyPaSWAS/__init__.py:17-305).

Options, destinations and defaults are the reference's; defaults come from
``Core/cfg/defaults.cfg`` and ``-c FILE`` overrides them.  ``--framework``
accepts HIP.
"""
import configparser
import logging
import optparse
import os
from datetime import datetime

from pypaswas_amd.Core.Exceptions import InvalidOptionException

# (flags, dest, config section, config key, help)
_OPTIONS = [
    ('General', [
        (('-L', '--logfile'), 'logfile', None, None, 'log events to FILE'),
        (('--loglevel',), 'loglevel', 'General', 'loglevel', 'log level: DEBUG, INFO, WARNING, ERROR, CRITICAL'),
        (('-o', '--out'), 'out_file', 'General', 'out_file', 'output file, defaults to ./output'),
        (('--outputformat',), 'out_format', 'General', 'out_format', 'TXT, SAM, trimmedFasta or FASTA'),
        (('-p', '--program'), 'program', 'General', 'program', '"aligner" or "trimmer"'),
        (('-1', '--filetype1'), 'filetype1', 'General', 'filetype1', 'file type of the first file'),
        (('-2', '--filetype2'), 'filetype2', 'General', 'filetype2', 'file type of the second file'),
        (('-O', '--override_output'), 'override_output', 'General', 'override_output', 'override output (T/F)'),
        (('-c', '--configfile'), 'config_file', None, None, 'settings file'),
    ]),
    ('Input', [
        (('--start_query',), 'start_query', 'Input', 'start_query', 'start index in the query file'),
        (('--end_query',), 'end_query', 'Input', 'end_query', 'end index in the query file'),
        (('--start_target',), 'start_target', 'Input', 'start_target', 'start index in the target file'),
        (('--end_target',), 'end_target', 'Input', 'end_target', 'end index in the target file'),
        (('--sequence_step',), 'sequence_step', 'Input', 'sequence_step', 'targets read per step'),
        (('--query_step',), 'query_step', 'Input', 'query_step', 'queries read per step'),
    ]),
    ('Aligner', [
        (('--customMatrix',), 'custom_matrix', None, None, 'custom matrix (26 lists of 26 numbers)'),
        (('-G', '--gap_score'), 'gap_score', 'Aligner', 'gap_score', 'gap penalty'),
        (('-g', '--gap_extension'), 'gap_extension', 'Aligner', 'gap_extension', 'gap extension penalty, 0: linear'),
        (('-M', '--matrixname'), 'matrix_name', 'Aligner', 'matrix_name',
         'DNA-RNA, BASIC, BLOSUM62, BLOSUM80, CUSTOM or IRYS'),
        (('-q', '--mismatch_score'), 'mismatch_score', 'Aligner', 'mismatch_score', 'mismatch penalty'),
        (('-r', '--match_score'), 'match_score', 'Aligner', 'match_score', 'match reward'),
        (('--any',), 'any_score', 'Aligner', 'any_score', 'score when N is present'),
        (('--other',), 'other_score', 'Aligner', 'other_score', 'score for other characters'),
        (('--minimum',), 'minimum_score', 'Aligner', 'minimum_score', 'minimum score that starts a traceback'),
        (('--llimit',), 'lower_limit_score', 'Aligner', 'lower_limit_score', 'lower limit relative to the maximum'),
    ]),
    ('Filter', [
        (('--filter_factor',), 'filter_factor', 'Filter', 'filter_factor', 'filter factor'),
        (('--query_coverage',), 'query_coverage', 'Filter', 'query_coverage', 'minimum query coverage'),
        (('--query_identity',), 'query_identity', 'Filter', 'query_identity', 'minimum query identity'),
        (('--relative_score',), 'relative_score', 'Filter', 'relative_score', 'minimum relative score'),
        (('--base_score',), 'base_score', 'Filter', 'base_score', 'minimum base score'),
    ]),
    ('GraphDatabase', [
        (('--hostname',), 'hostname', 'GraphDatabase', 'hostname', 'Neo4j host'),
        (('--username',), 'username', 'GraphDatabase', 'username', 'Neo4j user'),
        (('--password',), 'password', 'GraphDatabase', 'password', 'Neo4j password'),
        (('--target_node',), 'target_node', 'GraphDatabase', 'target_node', 'target node name'),
        (('--sequence_node',), 'sequence_node', 'GraphDatabase', 'sequence_node', 'sequence node name'),
    ]),
    ('Device', [
        (('--device',), 'device_number', 'Device', 'device_number', 'GPU index'),
        (('--number_of_compute_units',), 'number_of_compute_units', 'Device', 'number_of_compute_units',
         'OpenCL sub-devices (ignored by HIP)'),
        (('--sub_device',), 'sub_device', 'Device', 'sub_device', 'OpenCL sub-device (ignored by HIP)'),
        (('--limit_length',), 'limit_length', 'Device', 'limit_length', 'longest sequence read from file'),
        (('--maximum_memory_usage',), 'maximum_memory_usage', 'Device', 'maximum_memory_usage',
         'fraction of device memory to use'),
        (('--njobs',), 'number_of_jobs', None, None, 'not implemented'),
        (('--process_id',), 'process_id', None, None, 'not implemented'),
        (('--max_genome_length',), 'max_genome_length', 'Device', 'max_genome_length', 'deprecated'),
        (('--recompile',), 'recompile', 'Device', 'recompile', 'T/F'),
        (('--short_sequences',), 'short_sequences', 'Device', 'short_sequences', 'T/F'),
    ]),
    ('Framework', [
        (('--framework',), 'framework', 'Framework', 'language', 'HIP'),
        (('--kernel_semantics',), 'kernel_semantics', 'Framework', 'kernel_semantics',
         'CPU: the OpenCL-CPU kernels (default); GPU: the recurrence of the GPU-OpenCL and CUDA kernels '
         '(affine gap open G + E)'),
    ]),
    ('OpenCL', [
        (('--device_type',), 'device_type', 'OpenCL', 'device_type', 'CPU, GPU or ACCELERATOR'),
        (('--platform_name',), 'platform_name', 'OpenCL', 'platform_name', 'platform name'),
    ]),
]

DEFAULTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'Core', 'cfg', 'defaults.cfg')


def parse_cli(config_file=DEFAULTS, argv=None):
    """Returns (settings, arguments); raises InvalidOptionException without two input files."""
    config = configparser.ConfigParser()
    config.read(config_file)
    parser = optparse.OptionParser(usage='%prog [options] FILE_1 FILE_2')
    parser.description = ('This program performs a Smith-Waterman alignment of all sequences in FILE_1 against '
                          'all sequences in FILE_2.')
    for group_name, options in _OPTIONS:
        group = optparse.OptionGroup(parser, group_name)
        for flags, dest, section, key, help_text in options:
            kw = {'dest': dest, 'help': help_text}
            if section is not None:
                kw['default'] = config.get(section, key)
            elif dest == 'config_file':
                kw['default'] = False
            group.add_option(*flags, **kw)
        parser.add_option_group(group)
    settings, arguments = parser.parse_args(argv)
    if settings.config_file:
        settings, arguments = _override_settings(settings.config_file, settings, arguments)
    if len(arguments) < 2:
        raise InvalidOptionException('Missing input files')
    return settings, arguments


def _override_settings(config_file, settings, arguments):
    config = configparser.ConfigParser()
    config.read(config_file)
    for section in config.sections():
        for name, value in config.items(section):
            if value:
                settings._update_careful({name: value})
    return settings, arguments


def set_logger(settings):
    numeric_level = getattr(logging, str(settings.loglevel).upper(), None)
    if not isinstance(numeric_level, int):
        raise InvalidOptionException('Invalid log level: %s' % settings.loglevel)
    logger = logging.getLogger('pypaswas_amd')
    logger.setLevel(numeric_level)
    for h in list(logger.handlers):
        logger.removeHandler(h)
    if settings.logfile is None:
        handler = logging.StreamHandler()
        handler.setFormatter(logging.Formatter('%(levelname)s - %(message)s'))
    else:
        try:
            with open(settings.logfile, 'a') as handle:
                _log_settings_to_file(handle, settings)
        except IOError:
            raise InvalidOptionException('Invalid log file or writing forbidden: %s' % settings.logfile)
        handler = logging.FileHandler(settings.logfile)
        handler.setFormatter(logging.Formatter('%(asctime)s - %(levelname)s - %(message)s'))
    handler.setLevel(numeric_level)
    logger.addHandler(handler)
    logger.propagate = False
    return logger


def _log_settings_to_file(handle, settings):
    handle.write("\n{}\n".format('-' * 74))
    handle.write(datetime.today().strftime("pyPaSWAS run started at: %Y-%m-%d %H:%M:%S using the following "
                                           "settings:\n"))
    handle.write("{}\n".format('-' * 74))
    for setting, value in vars(settings).items():
        handle.write("{0:30}:{1:>30}\n".format(setting, 'N/A' if value is None else str(value)))
    handle.write("{}\n".format('-' * 74))


def normalize_file_path(path):
    if not os.path.isabs(path):
        path = os.path.join(os.getcwd(), path)
    return os.path.normpath(path)
