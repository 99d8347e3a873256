"""MI355X-native statevector engine for the QSV2 MUX2 record stream.

Modules: isa (stream constants), gates (closed-form gate matrices), circuit (front-end, benchmark
families, measured workloads), compiler (circuit -> MUX2 -> passes -> QSV2), libsvx (ctypes C-ABI
of the HIP library built from csrc/), runtime (SVRuntime: the drop-in for the reference's run call).
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
__all__ = ["PKG_DIR"]
