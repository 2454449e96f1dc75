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
