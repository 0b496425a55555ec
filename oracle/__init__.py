"""CPU oracle of the DMRG H*psi path -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the reference's algorithms for this path (ITensor v3
C++: tensor/contract.cc, itdata/qdense.cc + itdata/qutil.h, the combiners,
iterativesolvers.h davidson, decomp.h denmatDecomp, hermitian.cc
diag_hermitian, decomp.cc truncate, mps/localop.h, mps/localmpo.h and
mps/dmrg.h).  Each function cites the reference file:line it follows.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker (or the timed CPU baseline);
the product path (itensor_amd) never imports it.

How each part is pinned (checked against something other than itself) is
stated in the module docstrings and in tests/test_oracle_*.py.
"""
