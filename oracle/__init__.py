"""CPU oracle for the MRCPP apply path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
package, and only as the checker.  The library is oracle/build/liboracle.so,
built from oracle/mrcpp_oracle.cpp by oracle/Makefile.
"""
from .orc import FILTER_DIR, Oracle, OracleMRA, OracleOperator, OracleTree, add, apply, build, multiply, prec_factor  # noqa: F401
