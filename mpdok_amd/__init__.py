"""mpdok_amd — MI355X-native dense-operator Krylov solvers (MPDOK's hot path).

Modules mirror the reference's Python entry points:
    mpdok_amd.dense_krylov   DenseLinearOperator, cg, gmres, gmres_ir
    mpdok_amd.mpdok_ops      MPDOKSolver, LUIRSolver, _ManagedArray
    mpdok_amd.rbf_kernel     synthetic_coords, build_rbf_kernel, weather_front
    mpdok_amd.mpdok_ooc      MPDOKOOCSolver

All device work runs in hand-written HIP kernels for gfx950 in
libmpdok_hip.so, reached through ctypes; torch tensors are only buffers.
"""
__version__ = "0.1.0"
