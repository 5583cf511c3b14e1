"""qml_amd: the QML kernel-ridge-regression hot path, native to MI355X (gfx950).

Drop-ins with the signatures of the reference (qmlcode/qml 0.4):
  qml_amd.representations.generate_coulomb_matrix  (qml/representations.py:69-119)
  qml_amd.kernels.laplacian_kernel, gaussian_kernel (qml/kernels.py:36-92)
  qml_amd.distance.manhattan_distance, l2_distance  (qml/distance.py:28-92)
  qml_amd.math.cho_solve                            (qml/math.py:60-97)
plus the device-resident qml_amd.krr.KRR (fit / fused predict). All compute runs in
libqmlhip.so (hand-written HIP); there is no CPU fallback.
"""

__version__ = "0.1.0"
