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


def build(verbose=False, jobs=None):
    """Compiles every source to an object in parallel (the front end's unrolled eliminations take minutes on their
    own), then links the shared library; the library is replaced atomically."""
    srcs = sources()
    work = tempfile.mkdtemp(prefix='arp_build_')
    try:
        def compile_one(src):
            obj = os.path.join(work, os.path.basename(src) + '.o')
            cmd = [HIPCC] + FLAGS + ['-I' + CSRC, '-c', '-o', obj, src]
            if verbose:
                print(' '.join(cmd), flush=True)
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
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
    print(build(verbose=True))
