"""Build libxcape_hip.so in-tree for gfx950 with hipcc.

Each ``.hip`` translation unit is compiled to an object with its own
floating-point contraction setting (the accurate kernels must not contract
a*b+c into FMA so that they round as the reference does), then all objects are
linked into ``xcape_amd/libxcape_hip.so``.

Every HIP compile carries ``-Xclang -target-feature -Xclang -packed-fp32-ops``
so the compiler emits no packed FP32 instructions.

Usage: ``python -m xcape_amd.build`` (or ``build()``).
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libxcape_hip.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
CLANG = os.environ.get('XCAPE_LINKER', '/opt/rocm/llvm/bin/clang++')
ARCH = 'gfx950'

COMMON = ['--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wall', '-Wno-unused-function', '-Wno-unknown-warning-option']

# translation unit -> extra flags
UNITS = {
    'capi.hip': ['-ffp-contract=off'],
    # MFMA operands and accumulators in VGPRs (not AGPRs): the surrogate's
    # evaluation shares the ascent kernel's register budget
    'cape_accurate.hip': ['-ffp-contract=off', '-mllvm', '-amdgpu-mfma-vgpr-form=1'],
    'cape_fast.hip': ['-ffp-contract=fast-honor-pragmas', '-mllvm', '-amdgpu-mfma-vgpr-form=1'],
    'ragged.hip': ['-ffp-contract=off'],
    'mlp.hip': ['-ffp-contract=fast'],
    'surrogate_log.hip': ['-ffp-contract=off'],
    'transpose.hip': ['-ffp-contract=off'],
    'math_check.hip': ['-ffp-contract=off'],
}


def _compile(unit, flags, outdir, csrc=CSRC):
    obj = os.path.join(outdir, unit.replace('.hip', '.o'))
    cmd = [HIPCC] + COMMON + flags + ['-c', os.path.join(csrc, unit), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = 'packed-fp32-ops'
    err = '\n'.join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {unit}:\n{err}')
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def build(verbose=True, csrc=CSRC, out=LIB):
    """Compile every unit of `csrc` and link the shared library `out`
    (always rebuilds)."""
    with tempfile.TemporaryDirectory(prefix='xcape_build_') as tmp:
        with ThreadPoolExecutor(max_workers=min(8, len(UNITS))) as ex:
            objs = list(ex.map(lambda kv: _compile(kv[0], kv[1], tmp, csrc), UNITS.items()))
        out_tmp = out + '.tmp'
        # Link without a DT_NEEDED on libamdhip64: the HIP symbols bind to the
        # runtime already loaded with global scope (the process's only one;
        # xcape_amd._hip arranges that, a C caller links -lamdhip64).  A hard
        # dependency would pull /opt/rocm's runtime into a process that already
        # runs torch's bundled copy.
        cmd = [CLANG, '-shared', '-fPIC', '-o', out_tmp] + objs + ['-lstdc++', '-lm']
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError('link failed:\n' + r.stderr)
        shutil.move(out_tmp, out)
    if verbose:
        print('built', out)
    return out


if __name__ == '__main__':
    build()
