"""Copies the data files of pyPaSWAS's own tests that pin this path into tests/golden/.

Inputs and expected output of test_pypaswasall.test_pypaswas_aligner_basic_genbank
(pyPaSWAS/Core/test/test_pypaswasall.py:62-83) and the FASTA inputs of the
disabled basic test.  Data only; run once where the reference tree exists.
"""
import os
import shutil

SRC = "/root/reference/pyPaSWAS/Core/test/data"
DST = os.path.dirname(os.path.abspath(__file__))
FILES = ["query_genbank.gb", "target_genbank.gb", "query_fasta.faa", "target_fasta.faa",
         "reference/aligner_basic_genbank_output.txt"]

if __name__ == "__main__":
    for f in FILES:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, os.path.basename(f)))
        print("copied", f)
