"""Builds the native library libmcr.so (HIP, gfx950) in-tree with hipcc.

Flags: -ffp-contract=off keeps every multiply and add separately rounded, as the reference's
JavaScript arithmetic is; packed FP32 instructions are disabled for the device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libmcr.so')
SOURCES = ['api.hip', 'engine.hip', 'server.hip', 'trace.hip', 'bvh_build.hip', 'wide.hip', 'compile.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wall', '-Wno-unused-function', '-Wno-unused-value', '-Wno-unused-result']


def build(verbose=False):
    objs = []
    os.makedirs(os.path.join(HERE, '_build'), exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(HERE, '_build', src.replace('.hip', '.o'))
        cmd = [HIPCC, '-c', *FLAGS, os.path.join(CSRC, src), '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        text = '\n'.join(l for l in out.decode().splitlines() if 'packed-fp32-ops' not in l)
        if p.returncode != 0:
            raise RuntimeError(f'hipcc failed: {" ".join(cmd)}\n{text}')
        if verbose and text.strip():
            print(text)
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '-shared', '--offload-arch=gfx950', '-fPIC', *objs, '-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed: {r.stdout}\n{r.stderr}')
    os.replace(tmp, LIB)
    return LIB


NAPI_SRC = os.path.join(HERE, 'napi', 'mcr_napi.c')
NAPI_LIB = os.path.join(HERE, 'napi', 'mcr.node')
NODE_INCLUDE = '/usr/include/node'


def build_napi():
    """Builds the Node-API addon napi/mcr.node against libmcr.so (gcc; Node's headers must be
    installed). Returns its path, or None where the headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, 'node_api.h')):
        return None
    tmp = NAPI_LIB + '.tmp'
    cmd = ['gcc', '-O2', '-Wall', '-shared', '-fPIC', '-I' + NODE_INCLUDE, NAPI_SRC, '-L' + HERE, '-lmcr',
           '-Wl,-rpath,$ORIGIN/..', '-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'addon build failed: {r.stdout}\n{r.stderr}')
    os.replace(tmp, NAPI_LIB)
    return NAPI_LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
    print(build_napi())
