"""janus's module names for janus_amd.

janus callers import ``janus.green``, ``janus.fft.serial``,
``janus.material.elastic.linear.isotropic``, ``janus.operators`` and
``janus.mandelvoigt`` (sphinx/tutorials/square_basic/square_basic.py:7-12,
examples/square_inclusion_scipy.py:14-17).
:func:`install` registers janus_amd's modules under those names, so such a
script runs unchanged once it has been called::

    import janus_amd.janus_names
    janus_amd.janus_names.install()

    import janus.green as green          # janus_amd.green
    import janus.fft.serial as fft       # janus_amd.fft.serial

Only the modules janus_amd provides are registered.  ``janus.fft.parallel``
(janus's MPI slab path, which examples/square_inclusion.py:10 imports) and
``janus.utils.checkarray`` (Cython argument checks, which janus_amd folds
into every call) are not, so importing them still raises
``ModuleNotFoundError``.
"""

import importlib
import sys

# janus module -> janus_amd module, parents before children
MODULES = (
    ('', ''),
    ('.fft', '.fft'),
    ('.fft.serial', '.fft.serial'),
    ('.green', '.green'),
    ('.operators', '.operators'),
    ('.mandelvoigt', '.mandelvoigt'),
    ('.material', '.material'),
    ('.material.elastic', '.material.elastic'),
    ('.material.elastic.linear', '.material.elastic.linear'),
    ('.material.elastic.linear.isotropic',
     '.material.elastic.linear.isotropic'),
    ('.utils', '.utils'),
)


def install(root='janus'):
    """Register janus_amd's modules as ``root``, ``root.green``, ... in
    ``sys.modules``.  Idempotent.  Raises RuntimeError when ``root`` is
    already the name of another imported package (janus itself)."""
    present = sys.modules.get(root)
    if present is not None and present is not sys.modules.get('janus_amd'):
        raise RuntimeError('a module named {0!r} is already imported ({1}); '
                           'janus_amd will not replace it'.format(
                               root, getattr(present, '__file__', present)))
    for theirs, ours in MODULES:
        sys.modules[root + theirs] = importlib.import_module('janus_amd' +
                                                             ours)
    return sys.modules[root]


def installed(root='janus'):
    """True when ``root`` currently names janus_amd."""
    return sys.modules.get(root) is sys.modules.get('janus_amd', object())


def uninstall(root='janus'):
    """Remove the names :func:`install` registered (only those that still
    name janus_amd's modules)."""
    for theirs, ours in reversed(MODULES):
        if sys.modules.get(root + theirs) is sys.modules.get('janus_amd' +
                                                             ours):
            del sys.modules[root + theirs]
