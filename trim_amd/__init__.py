"""MI355X-native (gfx950) Uni10 tensor-contraction hot path.

Native libraries (built in-tree by ``python -m trim_amd._build``):
  * libtn_hip.so    - HIP kernels behind the C ABI of include/tn_hip.h
  * libuni10_amd.so - the Uni10 C++ API (UniTensor/Bond/Qnum/Network) over it

``trim_amd.tn`` binds the C ABI; ``trim_amd.uni10`` binds the C++ API's C
facade.  Neither falls back to a CPU path: a missing library raises.
"""
__all__ = ["tn"]
