"""The C-ABI library loads and exports every symbol include/fhs.h declares
(no compute call: runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'fhs.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(fhs_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_functions():
    names = declared_functions()
    assert 'fhs_apply_direct_operator' in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    from fhs import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f'not exported: {missing}'


def test_binding_covers_header():
    from fhs import _lib
    assert set(declared_functions()) == set(_lib.PROTOTYPES), \
        set(declared_functions()) ^ set(_lib.PROTOTYPES)
    _lib.lib()   # binds every prototype, raising on a missing symbol


def test_update_epsilon_host_matches_oracle():
    """fhs_update_epsilon is host logic (no GPU): compare with the oracle."""
    import numpy as np
    import fhs
    from oracle.prox import update_epsilon as ref
    rng = np.random.default_rng(0)
    for t in (10, 50, 200):
        eps = [list(rng.uniform(1, 2, 3)), list(rng.uniform(1, 2, 2))]
        tb = [[0, 30, 150], [45, 0]]
        nr = [list(rng.uniform(0.5, 2.5, 3)), list(rng.uniform(0.5, 2.5, 2))]
        e1, t1 = fhs.update_epsilon(eps, t, tb, nr, 0.99, 1.001, 40, 0.5)
        e2, t2 = ref(eps, t, tb, nr, 0.99, 1.001, 40, 0.5)
        assert np.allclose(np.concatenate(e1), np.concatenate(e2)) and t1 == [list(x) for x in t2]
