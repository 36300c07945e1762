"""FFT wrappers (janus/fft/__init__.py).

The planner flags of janus.fft (janus/fft/__init__.py:64-72; the values are
FFTW's public ones, janus/fft/serial/_serial_fft.pyx:10-18) are accepted by
:func:`janus_amd.fft.serial.create_real` for compatibility.  The HIP
transforms need no planning, so the flags do not change the result.
"""

FFTW_MEASURE = 0
FFTW_DESTROY_INPUT = 1
FFTW_UNALIGNED = 1 << 1
FFTW_CONSERVE_MEMORY = 1 << 2
FFTW_EXHAUSTIVE = 1 << 3
FFTW_PRESERVE_INPUT = 1 << 4
FFTW_PATIENT = 1 << 5
FFTW_ESTIMATE = 1 << 6
FFTW_WISDOM_ONLY = 1 << 21
