"""Build the native parts of fresnel_amd in-tree.

* libfresnel_hip.so -- HIP kernels + C++ runtime + C ABI (hipcc, gfx950)
* _common.*.so      -- pybind11 POD types (g++)
* _gpu.*.so         -- pybind11 binding of the C ABI (g++, links libfresnel_hip.so)

Run ``python fresnel_amd/build.py [all|common|hip|gpu]``; it does not import the
package, so it works before anything is built.
"""

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')
ARCH = os.environ.get('FRESNEL_AMD_ARCH', 'gfx950')
EXT = sysconfig.get_config_var('EXT_SUFFIX')

# Device code: no FMA contraction (bit-exact float order), IEEE division and
# square root (the HIP default, stated explicitly), no packed FP32 ops.
HIP_FLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}', '-ffp-contract=off',
    '-fhip-fp32-correctly-rounded-divide-sqrt', '-fno-gpu-rdc', '-munsafe-fp-atomics',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wall',
    '-Wno-unused-function', '-Wno-unused-variable', '-Wno-unused-parameter'
]

HIP_SOURCES = [
    'runtime/runtime.hip',
]


def _run(cmd):
    print('+', ' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _pybind_flags():
    import pybind11
    return [
        '-O2', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off', f'-I{pybind11.get_include()}',
        f'-I{sysconfig.get_paths()["include"]}', f'-I{INCLUDE}'
    ]


def build_common():
    out = os.path.join(HERE, '_common' + EXT)
    _run(['g++'] + _pybind_flags() + [os.path.join(CSRC, 'bindings', 'common_module.cpp'), '-o', out])
    return out


def build_hip():
    out = os.path.join(HERE, 'libfresnel_hip.so')
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([os.path.join(ROCM, 'bin', 'hipcc')] + HIP_FLAGS + ['-shared', f'-I{INCLUDE}', f'-I{CSRC}'] +
         srcs + ['-o', out])
    return out


def build_gpu_module():
    out = os.path.join(HERE, '_gpu' + EXT)
    _run(['g++'] + _pybind_flags() + [
        os.path.join(CSRC, 'bindings', 'gpu_module.cpp'), '-o', out, f'-L{HERE}', '-lfresnel_hip',
        "-Wl,-rpath,$ORIGIN"
    ])
    return out


def build_all():
    build_common()
    build_hip()
    build_gpu_module()


if __name__ == '__main__':
    which = sys.argv[1:] or ['all']
    for w in which:
        {'all': build_all, 'common': build_common, 'hip': build_hip, 'gpu': build_gpu_module}[w]()
