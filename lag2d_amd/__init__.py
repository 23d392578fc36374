"""lag2d_amd: MI355X-native Lagrangian DG step of liuspace/Lag2D (srccurv-lag).

The compute path is the HIP library lag2d_amd/liblag2d.so behind the C ABI of
include/lag2d.h; this package holds its ctypes binding, the problem/mesh
generators that play the reference driver's role, and the build script.
"""
from .case import LagCase  # noqa: F401
