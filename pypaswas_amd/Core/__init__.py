"""Core of the HIP framework: the SmithWaterman engine interface and its data model.

Direction codes are the reference's (pyPaSWAS/Core/__init__.py:7-13).
"""
UPPER_LEFT_DIRECTION = 1
UPPER_DIRECTION = 2
LEFT_DIRECTION = 3

NO_DIRECTION = 2 ** 5
STOP_DIRECTION = 2 ** 6
IN_ALIGNMENT = 2 ** 5 + 2 ** 6

