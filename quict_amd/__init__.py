"""MI355X-native circuit simulation (QuICT simulation-path drop-in).

``quict_amd.core``        Circuit / gate / noise front end (QuICT.core API).
``quict_amd.simulation``  StateVectorSimulator / DensityMatrixSimulator / Simulator.
``quict_amd.gradient``    adjoint differentiation and batched QNN.
``quict_amd._native``     ctypes binding of libqsim_hip.so (HIP kernels, C-ABI).
"""
__version__ = "0.1.0"
