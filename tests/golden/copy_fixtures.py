"""Copies the reference's input data files used as reader fixtures into
tests/golden/ (gzip-compressed, contents unchanged):

  src/molcas.log                          Molcas RASSCF log: CAS(8,8) CI
                                          vector in CSFs, natural occupations
  src/molcas.rasscf.molden                Molcas molden file of the same run
  src/UCCSD_CO_4Bohr_Molden_631star.txt   PySCF molden file (CO, alpha and
                                          beta orbitals)

Run from the repository root in a tree that has the reference checkout:
    python tests/golden/copy_fixtures.py
"""
import gzip
import os
import shutil

SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ["molcas.log", "molcas.rasscf.molden", "UCCSD_CO_4Bohr_Molden_631star.txt"]

if __name__ == "__main__":
    for name in FILES:
        with open(os.path.join(SRC, name), "rb") as fi, \
                open(os.path.join(HERE, name + ".gz"), "wb") as raw, \
                gzip.GzipFile(filename="", mode="wb", fileobj=raw, mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        print("wrote", name + ".gz")
