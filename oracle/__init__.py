"""CPU oracle for the DG operator (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of ``bench.py``
may import this package.  The product package ``bosss_amd`` never does.
"""
from .bosss_ref import RefModel, OP_SIP, OP_ADVDIFF, OP_EULER_LF, OP_EULER_HLLC, ref_block_lu, ref_block_solve, build  # noqa: F401
