"""janus_amd: MI355X (gfx950) implementation of janus's Lippmann-Schwinger
fixed-point hot path.

The module layout mirrors janus so that janus-style solver scripts switch by
changing the import root:

=====================================  ========================================
janus                                  janus_amd
=====================================  ========================================
janus.fft.serial.create_real           janus_amd.fft.serial.create_real
janus.green.truncated / filtered       janus_amd.green.truncated / filtered
janus.operators (isotropic_4,          janus_amd.operators
BlockDiagonalOperator2D/3D)
janus.material.elastic.linear          janus_amd.material.elastic.linear
.isotropic                             .isotropic
janus.mandelvoigt                      janus_amd.mandelvoigt
=====================================  ========================================

Additions: :mod:`janus_amd.solver` (the fused on-device fixed-point loop,
conjugate gradients on the strain and polarization forms, restarted GMRES,
plan autotuning), :mod:`janus_amd.batch` (many right-hand sides in one
sequence of launches), :mod:`janus_amd.autograd` (implicit-differentiation
adjoint for isotropic and matrix local operators, effective stress) and
:mod:`janus_amd.microstructure` (seeded test microstructures).
All compute runs in ``libjanus_amd.so`` (hand-written HIP for gfx950).
"""

__version__ = '0.1.0'
