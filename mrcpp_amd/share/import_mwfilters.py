"""Import the Legendre multiwavelet filter and cross-correlation tables.

The tables are the numerical data that MRCPP reads at run time
(share/mwfilters/L_H0_k, L_G0_k, L_c_left_k, L_c_right_k; format described in
share/mwfilters/README of MRCPP, readers at src/core/MWFilter.cpp:200-251 and
src/core/CrossCorrelation.cpp:99-124).  They are copied byte for byte: the
rows of G0 carry sign choices that are not fixed by the mathematics (each
wavelet is defined only up to sign), so identical coefficients require the
identical tables.  tests/test_filters.py checks every imported table against
the mathematics that defines it (two-scale relation, orthogonality, Alpert
vanishing moments, cross-correlation integrals).

Usage: python mrcpp_amd/share/import_mwfilters.py [/path/to/mrcpp/share/mwfilters]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DST = os.path.join(HERE, "mwfilters")
FILTER_ORDERS = range(0, 42)  # MaxOrder = 41 (api/constants.h:33)
CCC_ORDERS = range(1, 14)     # scaling orders supported by operator construction


def main(src):
    os.makedirs(DST, exist_ok=True)
    names = []
    for k in FILTER_ORDERS:
        names += ["L_H0_%d" % k, "L_G0_%d" % k]
    for k in CCC_ORDERS:
        names += ["L_c_left_%d" % k, "L_c_right_%d" % k]
    for name in names:
        shutil.copyfile(os.path.join(src, name), os.path.join(DST, name))
    print("imported %d tables into %s" % (len(names), DST))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/share/mwfilters")
