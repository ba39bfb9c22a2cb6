"""MI355X-native SOLWEIG 2022a radiation engine (Tmrt and UTCI per pixel-hour).

Python host code with the reference's drop-in interfaces (``model``, ``svf``, ``utci``,
``series``) over hand-written HIP kernels for gfx950, reached through the C ABI of
``libsolweig_hip.so`` (``runtime``). ``batched`` drives whole days on the device and
``autograd`` exposes the hand-written adjoint to PyTorch.
"""
__version__ = "0.1.0"
