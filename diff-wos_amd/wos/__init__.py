"""wos -- drop-in of the reference's ``wos`` package (wos/*.py) on PyTorch-ROCm.

Solvers, scenes and estimators keep the reference's module, class, method and
keyword names; arrays are torch tensors on the HIP device and every walk runs
in the HIP kernels of ``libwos_amd.so`` through ``wos_ext``.
"""
