"""Serial real FFTs on the GPU, with the interface of janus.fft.serial.

Replaces janus/fft/serial/_serial_fft.pyx: ``create_real`` (:20-33),
``_RealFFT2D``/``_RealFFT3D`` shapes (:57-74, :181-200), ``r2c`` (:152-163,
:296-309) and ``c2r`` (:165-177, :311-325).  The complex output is stored in
janus's interleaved layout: along the last axis even entries are real parts,
odd entries imaginary parts (sphinx/fft_tutorial.rst:43).  r2c is
unnormalized and c2r scales by 1/N.  Inputs are preserved.

Arrays may be numpy arrays (copied to and from the GPU, as janus copies to
and from its FFTW buffer) or float64 torch tensors on the GPU (used in place).
"""

import ctypes

import numpy as np

from .. import _device, _lib
from . import FFTW_MEASURE

INVALID_REAL_ARRAY_SHAPE = 'shape of real array must be {0} [was {1}]'
INVALID_COMPLEX_ARRAY_SHAPE = 'shape of complex array must be {0} [was {1}]'


def create_real(shape, flags=FFTW_MEASURE):
    """Return a new FFT object to compute real-to-complex transforms.

    Args:
        shape (tuple): the global shape of the input data.
        flags (int): "or" combination of planner flags (accepted, unused).
    """
    if len(shape) == 2:
        return _RealFFT2D(shape[0], shape[1])
    elif len(shape) == 3:
        return _RealFFT3D(shape[0], shape[1], shape[2])
    else:
        msg = 'length of shape can be 2 or 3 (was {0})'
        raise ValueError(msg.format(len(shape)))


class _RealFFT:
    def __init__(self, *n):
        n = tuple(int(ni) for ni in n)
        for i, ni in enumerate(n):
            if ni <= 0:
                raise ValueError('shape{0} must be > 0 (was {1})'.format(i, ni))
        self._n = n
        last = 2 * (n[-1] // 2 + 1)
        self._ishape = n
        self._oshape = n[:-1] + (last,)
        self._handle = None

    # read-only attributes of _serial_fft.pxd:17-18
    @property
    def ishape(self):
        return self._ishape

    @property
    def oshape(self):
        return self._oshape

    @property
    def global_ishape(self):
        return self._ishape

    @property
    def global_oshape(self):
        return self._oshape

    @property
    def isize(self):
        return int(np.prod(self._ishape))

    @property
    def osize(self):
        return int(np.prod(self._oshape))

    @property
    def offset0(self):
        return 0

    @property
    def idispl(self):
        return 0

    @property
    def odispl(self):
        return 0

    @property
    def scaling(self):
        return 1. / float(np.prod(self._n))

    def _h(self):
        if self._handle is None:
            _lib.require_gpu()
            L = _lib.lib()
            h = ctypes.c_void_p()
            _lib.check(L.jx_fft_create(len(self._n), _lib.shape_array(self._n),
                                       ctypes.byref(h)))
            self._handle = h
        return self._handle

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and _lib._lib is not None:
            _lib._lib.jx_fft_destroy(h)
            self._handle = None

    def _check(self, a, shape, template):
        if tuple(a.shape) != tuple(shape):
            raise ValueError(template.format(shape, tuple(a.shape)))

    def r2c(self, r, c=None):
        """Forward DFT of the real array ``r`` (unnormalized)."""
        self._check(r, self._ishape, INVALID_REAL_ARRAY_SHAPE)
        if c is not None:
            self._check(c, self._oshape, INVALID_COMPLEX_ARRAY_SHAPE)
        device = _device.is_device_tensor(r)
        if c is not None:
            _device.require_f64(c, 'c')
        rd = _device.to_device(r)
        if device and c is not None and _device.is_device_tensor(c) \
                and c.is_contiguous():
            cd = c
        else:
            cd = _device.empty_device(self._oshape)
        L = _lib.lib()
        _lib.check(L.jx_fft_r2c(self._h(), rd.data_ptr(), cd.data_ptr(),
                                _lib.current_stream()))
        if device:
            if c is not None and cd is not c:
                c.copy_(cd)
                return c
            return cd
        return _device.to_host_into(cd, c)

    def c2r(self, c, r=None):
        """Inverse DFT of the half spectrum ``c``, scaled by 1/N."""
        self._check(c, self._oshape, INVALID_COMPLEX_ARRAY_SHAPE)
        if r is not None:
            self._check(r, self._ishape, INVALID_REAL_ARRAY_SHAPE)
        device = _device.is_device_tensor(c)
        if r is not None:
            _device.require_f64(r, 'r')
        cd = _device.to_device(c)
        if device and r is not None and _device.is_device_tensor(r) \
                and r.is_contiguous():
            rd = r
        else:
            rd = _device.empty_device(self._ishape)
        L = _lib.lib()
        _lib.check(L.jx_fft_c2r(self._h(), cd.data_ptr(), rd.data_ptr(),
                                _lib.current_stream()))
        if device:
            if r is not None and rd is not r:
                r.copy_(rd)
                return r
            return rd
        return _device.to_host_into(rd, r)


class _RealFFT2D(_RealFFT):
    def __init__(self, n0, n1):
        super().__init__(n0, n1)

    @property
    def ishape0(self):
        return self._ishape[0]

    @property
    def ishape1(self):
        return self._ishape[1]

    @property
    def oshape0(self):
        return self._oshape[0]

    @property
    def oshape1(self):
        return self._oshape[1]


class _RealFFT3D(_RealFFT):
    def __init__(self, n0, n1, n2):
        super().__init__(n0, n1, n2)

    @property
    def ishape0(self):
        return self._ishape[0]

    @property
    def ishape1(self):
        return self._ishape[1]

    @property
    def ishape2(self):
        return self._ishape[2]

    @property
    def oshape0(self):
        return self._oshape[0]

    @property
    def oshape1(self):
        return self._oshape[1]

    @property
    def oshape2(self):
        return self._oshape[2]
