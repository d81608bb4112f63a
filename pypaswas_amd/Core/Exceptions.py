"""Exception taxonomy of the drop-in boundary (pyPaSWAS/Core/Exceptions.py:5-24)."""


class InvalidOptionException(Exception):
    """Invalid settings or arguments."""


class HardwareException(Exception):
    """Device missing, unusable, or a device-side error."""


class CudaException(Exception):
    """Invalid hit or location; the name is kept for drop-in compatibility."""


class ReaderException(Exception):
    """No (more) sequences in an input file."""
