"""C-ABI checks that need no GPU: exports, struct layout, host-side error paths."""
import ctypes
import os
import re

import numpy as np
import pytest

from isac_amd import binding

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpu_aln_pack.h")


def header_functions(path=HEADER):
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", txt)
    return [n for n in names if n not in ("sizeof",)]


def test_header_declares_reference_symbols():
    names = header_functions()
    for s in binding.REFERENCE_SYMBOLS + binding.EXTRA_SYMBOLS:
        assert s in names, s
    assert len(set(names)) == len(binding.REFERENCE_SYMBOLS) + len(binding.EXTRA_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(binding.LIB_PATH)
    inc = os.path.join(ROOT, "include")
    names = []
    for h in sorted(os.listdir(inc)):
        if h.endswith(".h"):
            names += header_functions(os.path.join(inc, h))
    assert "isac_nn_conv_fwd" in names and "isac_mref_proposal_device" in names
    for s in names:
        assert hasattr(lib, s), f"{s} not exported"


def test_struct_layout_matches_reference():
    # cuda/gpu_aln_common.h:62-85 on LP64
    assert ctypes.sizeof(binding.AlignConfig) == 40
    assert binding.AlignConfig.aln_type.offset == 32
    assert binding.AlignConfig.shift_step.offset == 20
    assert ctypes.sizeof(binding.AlignParam) == 24
    assert binding.AlignParam.mirror.offset == 20
    assert binding.AlignParam.shift_x.offset == 8


def test_ctypes_mirrors_match_header(tmp_path):
    """Compile the header with gcc and compare every struct's size and offsets."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"AlignConfig": binding.AlignConfig, "AlignParam": binding.AlignParam,
               "IsacMrefSubset": binding.IsacMrefSubset}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for name, cls in structs.items():
        assert got[(name, "size")] == ctypes.sizeof(cls), name
        for f, _ in cls._fields_:
            assert got[(name, f)] == getattr(cls, f).offset, (name, f)
    assert ctypes.sizeof(binding.IsacMrefSubset) == 72


def test_batched_rejects_bad_arguments_without_gpu():
    lib = binding.load()
    lib.gpu_clear_error()
    # no descriptors: nothing to do, but a NULL array with a count is an error
    rc = lib.isac_mref_align_batched_device(None, 0, binding.BATCH_GRAPH, None)
    assert rc in (0, 1)   # 1 when no device is present
    lib.gpu_clear_error()


def test_texture_limit_positive():
    lib = binding.load()
    assert lib.get_texture_limit(76) > 1000
    assert lib.get_texture_limit(0) == 0


def test_calls_without_init_report_errors():
    lib = binding.load()
    lib.gpu_clear_error()
    lib.pre_align_run(0, 10)
    assert lib.gpu_last_error() == 3
    with pytest.raises(binding.GpuAlignError):
        binding.check(lib)
    assert lib.gpu_last_error() == 0
    lib.reset_shifts(2.0, 1.0)
    assert lib.gpu_last_error() == 3
    lib.gpu_clear_error()
    pk = np.zeros(4, np.float32)
    d = np.zeros(1, np.float32)
    lib.multi_ref_align(None, None, pk.ctypes.data_as(binding.float_ptr),
                        d.ctypes.data_as(binding.float_ptr), 1, 0.9)
    assert lib.gpu_last_error() == 3
    lib.gpu_clear_error()
    lib.ref_free_alignment_2D()
    assert lib.gpu_last_error() == 3
    lib.gpu_clear_error()
    lib.pre_align_fetch(None, 0, b"sbj_batch")
    assert lib.gpu_last_error() == 3
    lib.gpu_clear_error()


def test_driver_post_rotation_matches_oracle():
    from oracle import isac_oracle as O
    rng = np.random.default_rng(0)
    pk = rng.uniform(-3, 3, (3, 4 * 5)).astype(np.float32)
    pk[:, 0::4] = rng.uniform(0, 360, (3, 5))
    out = binding.post_rotate(pk)
    sx, sy = O.driver_post_rotation(pk[:, 0::4].astype(np.float64), pk[:, 1::4].astype(np.float64),
                                    pk[:, 2::4].astype(np.float64))
    assert np.allclose(out[:, 1::4], sx, atol=1e-5) and np.allclose(out[:, 2::4], sy, atol=1e-5)
    assert np.array_equal(out[:, 0::4], pk[:, 0::4]) and np.array_equal(out[:, 3::4], pk[:, 3::4])


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_load_binds_the_hip_runtime_torch_carries():
    # PyTorch-ROCm wheels ship their own libamdhip64 (same soname as /opt/rocm's):
    # binding.load() imports torch first, so the process maps one HIP runtime.
    import subprocess
    import sys
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from isac_amd import binding\n"
        "binding.load()\n"
        "maps = open('/proc/self/maps').read().split('\\n')\n"
        "libs = sorted({l.split()[-1] for l in maps if 'libamdhip64' in l})\n"
        "print('\\n'.join(libs))\n" % os.path.join(ROOT, "gpu-isac_amd"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    libs = out.stdout.split()
    assert len(libs) == 1, libs
    import torch
    assert os.path.dirname(libs[0]) == os.path.join(os.path.dirname(torch.__file__), "lib"), libs
