"""MI355X-native solver for the semiconductor Bloch equations of EDUS.

The drop-in surface is native: ``include/edus.h`` (GPU engine, libedus_hip.so) and
``include/edus_host.h`` (C++ host: input.txt, setup, time loop, output files,
libedus_host.so).  This package holds their sources (``csrc/``), the build recipe
(``build.py``) and ctypes bindings (``native.py``).
"""
__all__ = ["native"]
