"""MI355X-native Newton-Raphson uvh assembly (shallow-shelf velocity-thickness system).

A drop-in for Úa's ``uvhAssembly`` (GHilmarG/UaSource ``UaMain/uvhAssembly.m``):
the same mesh, fields, boundary conditions and control variables go in, the
residual ``R`` and the sparse Jacobian ``K`` come out, computed by HIP kernels
for gfx950 behind a C ABI (``include/uvh.h``).
"""
