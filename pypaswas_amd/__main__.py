"""``python -m pypaswas_amd [options] FILE_1 FILE_2`` — the pypaswas.py entry point."""
import sys

from pypaswas_amd.pypaswasall import Pypaswas


def main(argv=None):
    Pypaswas(argv=argv).run()
    return 0


if __name__ == '__main__':
    sys.exit(main())
