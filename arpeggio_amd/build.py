"""Builds libarpeggio_hip.so in-tree from arpeggio_amd/csrc/*.hip with hipcc for gfx950.

    python -m arpeggio_amd.build            (or arpeggio_amd.build.build())
"""
import glob
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libarpeggio_hip.so')
ARCH = os.environ.get('ARP_OFFLOAD_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

# -packed-fp32-ops: no packed FP32 VALU forms in the device code. -ffp-contract=off: the FP32 annealers and the
# front end evaluate every expression in the written order, without fused multiply-add, like the CPU oracle.
FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH, '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-result']


def sources():
    return sorted(glob.glob(os.path.join(CSRC, '*.hip')))


def _includes(path, seen):
    """The project files a source includes ("..." includes, followed recursively)."""
    import re
    for m in re.finditer(r'^\s*#\s*include\s+"([^"]+)"', open(path).read(), re.M):
        p = os.path.normpath(os.path.join(os.path.dirname(path), m.group(1)))
        if p not in seen and os.path.exists(p):
            seen.append(p)
            _includes(p, seen)
    return seen


def _digest(src):
    """Hash of a source, the project headers it includes, and the flags: the key of a cached object."""
    import hashlib
    h = hashlib.sha256(' '.join(FLAGS).encode())
    for p in [src] + _includes(src, []):
        with open(p, 'rb') as f:
            h.update(f.read())
    return h.hexdigest()[:20]


def build(verbose=False, jobs=None, cache_dir=None):
    """Compiles every source to an object in parallel (the front end's unrolled eliminations take minutes on their
    own), then links the shared library; the library is replaced atomically. cache_dir: reuse an object compiled
    before from identical sources and flags (None: compile everything)."""
    srcs = sources()
    work = tempfile.mkdtemp(prefix='arp_build_')
    try:
        def compile_one(src):
            obj = os.path.join(work, os.path.basename(src) + '.o')
            if cache_dir:
                cached = os.path.join(cache_dir, '%s.%s.o' % (os.path.basename(src), _digest(src)))
                if os.path.exists(cached):
                    shutil.copy(cached, obj)
                    return obj, subprocess.CompletedProcess([], 0, '')
            cmd = [HIPCC] + FLAGS + ['-I' + CSRC, '-c', '-o', obj, src]
            if verbose:
                print(' '.join(cmd), flush=True)
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if cache_dir and r.returncode == 0:
                os.makedirs(cache_dir, exist_ok=True)
                shutil.copy(obj, cached)
            return obj, r
        with ThreadPoolExecutor(max_workers=jobs or min(8, len(srcs))) as ex:
            results = list(ex.map(compile_one, srcs))
        for (obj, r), src in zip(results, srcs):
            if r.returncode != 0:
                sys.stderr.write(r.stdout)
                raise RuntimeError('hipcc failed on %s' % src)
        tmp = LIB + '.tmp%d' % os.getpid()
        cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + [o for o, _ in results]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout)
            raise RuntimeError('hipcc failed linking %s' % LIB)
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return LIB


if __name__ == '__main__':
    cache = sys.argv[sys.argv.index('--cache') + 1] if '--cache' in sys.argv else None
    print(build(verbose=True, cache_dir=cache))
