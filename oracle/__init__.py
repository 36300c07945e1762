"""CPU oracle for the Lippmann-Schwinger hot path -- TEST INFRASTRUCTURE ONLY.

This package restates janus's serial CPU algorithm (FFT, discrete Green
operators, isotropic block-diagonal operators, basic fixed-point scheme) in
plain numpy so that the HIP implementation in ``janus_amd`` can be checked
against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` legs of ``bench.py`` may import it, and only to check
results or as the timed CPU baseline; the product path never routes through
it.

Every function cites the janus file:line it restates.  Pinning of each
function (fixtures, closed forms, published numbers) is listed in the module
docstring of :mod:`oracle.janus_oracle`.
"""
