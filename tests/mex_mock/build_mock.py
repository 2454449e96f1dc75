"""Builds the MEX stand-in (libmexmock.so) and every MEX shim of
faceted-hypersara_amd/mex/ against it and libfhs.so, into
tests/mex_mock/_build/ (gcc, no GPU needed).  Test infrastructure only.

    python tests/mex_mock/build_mock.py
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
MEX = os.path.join(ROOT, 'faceted-hypersara_amd', 'mex')
LIBDIR = os.path.join(ROOT, 'faceted-hypersara_amd', 'fhs')
OUT = os.path.join(HERE, '_build')
CFLAGS = ['-std=c99', '-O2', '-fPIC', '-shared', '-Wall', '-Wextra', '-Wno-unused-parameter', '-Werror']


def shims():
    return sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(MEX, '*.c')))


def build():
    os.makedirs(OUT, exist_ok=True)
    mock = os.path.join(OUT, 'libmexmock.so')
    cmds = [['gcc'] + CFLAGS + ['-I', HERE, os.path.join(HERE, 'mexmock.c'), '-o', mock]]
    for name in shims():
        cmds.append(['gcc'] + CFLAGS + ['-I', HERE, '-I', os.path.join(ROOT, 'include'), os.path.join(MEX, name + '.c'),
                     '-o', os.path.join(OUT, name + '.so'), '-L', OUT, '-lmexmock', '-L', LIBDIR, '-lfhs',
                     '-Wl,-rpath,' + OUT, '-Wl,-rpath,' + LIBDIR])
    for c in cmds:
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(' '.join(c) + '\n' + r.stderr + r.stdout)
    return OUT


if __name__ == '__main__':
    try:
        print(build())
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
