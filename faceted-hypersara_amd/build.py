"""Build libfhs.so (gfx950) in-tree with hipcc.

    python faceted-hypersara_amd/build.py            # full rebuild
    python faceted-hypersara_amd/build.py --incremental

Objects go to faceted-hypersara_amd/_build/, the library to
faceted-hypersara_amd/fhs/libfhs.so (loaded by fhs._lib).
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
OUT = os.path.join(HERE, 'fhs', 'libfhs.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('FHS_OFFLOAD_ARCH', 'gfx950')

CFLAGS = [
    '-O3', '-fPIC', '-std=c++17', f'--offload-arch={ARCH}',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result',
]


def _headers_mtime():
    hs = glob.glob(os.path.join(CSRC, '*.hpp')) + glob.glob(os.path.join(HERE, '..', 'include', '*.h'))
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _compile(src, incremental, hdr_mtime):
    obj = os.path.join(BUILD, os.path.basename(src).replace('.hip', '.o'))
    if incremental and os.path.exists(obj):
        om = os.path.getmtime(obj)
        if om > os.path.getmtime(src) and om > hdr_mtime:
            return obj, None
    cmd = [HIPCC] + CFLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, r.stderr + r.stdout
    return obj, None


def build(incremental=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, '*.hip')))
    hdr = _headers_mtime()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs, errors = [], []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda s: _compile(s, incremental, hdr), srcs):
            objs.append(obj)
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError('hipcc failed:\n' + '\n'.join(errors))
    tmp = OUT + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', f'--offload-arch={ARCH}', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stderr + r.stdout)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('--incremental', action='store_true')
    ap.add_argument('-j', type=int, default=None)
    a = ap.parse_args()
    try:
        print(build(a.incremental, a.j))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
