"""FP64 CPU oracle for the Faceted HyperSARA primal-dual iteration.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``fhs`` package, the
HIP library, ``bench.py``'s timed GPU leg) may import, call or execute code
from this package.  It exists to check the GPU path: ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` are
its only callers.

The reference (basp-group/Faceted-HyperSARA) is MATLAB; its NUFFT and
wavelet arithmetic live in two un-vendored submodules (RI-measurement-operator
and SARA-dictionary, ``.gitmodules:1-6``, versions unknown).  Each module here
restates, in numpy float64, the semantics of the reference ``.m`` file it
cites, and the published algorithm (Fessler & Sutton 2003 min-max NUFFT;
Daubechies filters by spectral factorisation; MATLAB ``wavedec2`` 'zpd'
indexing) where the reference only calls into a submodule.

How each function is pinned (golden vectors do not exist in the reference)
is listed in the module headers; functions whose exact reference iterates
cannot be pinned say "parity unpinned" there.
"""
