"""Build libpynufft_hip.so in-tree with hipcc for gfx950.

    python -m pynufft_amd.build        # or __graft_entry__.build()

The library has no torch dependency (plain C ABI); it is loaded with ctypes
after torch so that both share torch's HIP runtime.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libpynufft_hip.so')
SOURCES = ['api.hip', 'fft.hip', 'interp.hip', 'vec.hip', 'solvers.hip']
ARCH = os.environ.get('PYNUFFT_HIP_ARCH', 'gfx950')


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError('hipcc not found')


def compile_flags():
    return ['--offload-arch=%s' % ARCH, '-O3', '-std=c++17', '-fPIC',
            '-ffp-contract=fast-honor-pragmas',
            # packed FP32 instructions stay disabled for all device code
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-Wno-unused-result']


def build(verbose=True, jobs=4):
    """Compile every source to an object in parallel, then link the .so."""
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace('.hip', '.o'))
        objs.append(obj)
        cmd = [hipcc()] + compile_flags() + ['-c', os.path.join(CSRC, src), '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '--offload-arch=%s' % ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    if verbose:
        print(' '.join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if out.returncode != 0:
        sys.stderr.write(out.stdout.decode())
        raise RuntimeError('link failed')
    os.replace(tmp, LIB)
    return LIB


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = out.decode()
    lines = [l for l in text.splitlines()
             if 'packed-fp32-ops' not in l and l.strip()]
    if p.returncode != 0:
        sys.stderr.write(text)
        raise RuntimeError('hipcc failed on %s' % src)
    if lines:
        sys.stderr.write('\n'.join(lines) + '\n')


if __name__ == '__main__':
    build()
