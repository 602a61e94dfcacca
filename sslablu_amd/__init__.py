"""sslablu_amd -- MI355X-native slab direct solver (hot path of SimonDirckx/SslabLU).

Subpackages mirror the reference's module paths (``direct_solve``, ``multislab``,
``matAssembly``, ``solver``, ``geometry``, ``pdo``) so drivers written against the
reference API run unchanged; the numerics run in libsslab_hip.so (gfx950 HIP
kernels behind a C ABI, include/sslab_hip.h).
"""
__version__ = "0.1.0"
