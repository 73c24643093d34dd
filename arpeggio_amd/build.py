"""Builds libarpeggio_hip.so in-tree from arpeggio_amd/csrc/*.hip with hipcc for gfx950.

    python -m arpeggio_amd.build            (or arpeggio_amd.build.build())
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libarpeggio_hip.so')
ARCH = os.environ.get('ARP_OFFLOAD_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

# -packed-fp32-ops: no packed FP32 VALU forms in the device code. -ffp-contract=off: the FP32 annealers and the
# front end evaluate every expression in the written order, without fused multiply-add, like the CPU oracle.
FLAGS = ['-O3', '-std=c++17', '-fPIC', '-shared', '--offload-arch=' + ARCH, '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-result']


def sources():
    return sorted(glob.glob(os.path.join(CSRC, '*.hip')))


def build(verbose=False):
    srcs = sources()
    tmp = LIB + '.tmp%d' % os.getpid()
    cmd = [HIPCC] + FLAGS + ['-I' + CSRC, '-o', tmp] + srcs
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError('hipcc failed building %s' % LIB)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
