"""MI355X-native statevector engine for Qiskit Aer's statevector method.

The engine (HIP kernels for gfx950 behind the C ABI of include/aerhip.h)
replaces the QubitVector method family that Aer's
``AerSimulator(method="statevector", device="GPU")`` drives.
"""
from .circuit import Circuit, Op
from .qubitvector import QubitVector, device_count
from .simulator import AerSimulator, Result

__all__ = ["AerSimulator", "Circuit", "Op", "QubitVector", "Result", "device_count"]
__version__ = "0.1.0"
