"""MI355X-native implicit step engine of the dac_vtsa TVSA bed simulator.

The compute path is hand-written HIP for gfx950 in ``dac_vtsa_amd/csrc`` behind a C ABI
(``include/vtsa.h``); the Python classes in ``core/`` and ``steps/`` keep the upstream Python
API (core/residual.py, core/jacobian.py, core/solver.py, steps/_runner.py) on top of it.
Importing the package does not load the native library; the first compute call does, and it
raises if the library is missing.
"""
__version__ = "0.1.0"
