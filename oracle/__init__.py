"""CPU oracle for the slab direct-solver hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, in plain numpy/scipy FP64, the
reference algorithms of SimonDirckx/SslabLU on the hot path (each function cites the
reference file:line it follows).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or as the
timed CPU baseline -- never as part of the product path, which runs on the GPU through
libsslab_hip.so and fails loudly when that library is missing.

Parity status of each module is stated in its docstring ("pinned by ..." or
"parity unpinned").
"""
