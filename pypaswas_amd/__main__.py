"""``python -m pypaswas_amd [options] FILE_1 FILE_2`` — the pypaswas.py entry point.

An error ends the program with a one-line message and exit status 1; with
``--loglevel DEBUG`` the whole traceback is logged instead
(pyPaSWAS/pypaswasall.py:269-280).
"""
import logging
import sys

from pypaswas_amd.pypaswasall import Pypaswas


def main(argv=None):
    app = Pypaswas(argv=argv)
    try:
        app.run()
    except Exception as exc:  # the program's last word on any failure
        if app.settings is not None and app.logger.getEffectiveLevel() == logging.DEBUG:
            app.logger.exception(str(exc))
        else:
            print('Something unforeseen caused the program to terminate.\n'
                  'The internal error message was: ', ','.join(str(a) for a in exc.args))
        return 1
    return 0


if __name__ == '__main__':
    sys.exit(main())
