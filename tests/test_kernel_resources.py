"""Register budgets of the hot kernels, read from the built library.

The site-tile sweeps are bound by each tile's walk, with about 2.5 waves
per SIMD at 41k patterns: they rely on holding every wave of the launch at
once (3 waves per SIMD, at most 168 registers) and on keeping every value
in registers (a spill adds a scratch round trip to the walk).  This test
reads the AMDGPU metadata of the gfx950 code object inside
netrax_amd/libnetrax_gpu.so (extracted with objcopy and
clang-offload-bundler, printed by llvm-readelf) and checks, per kernel, no
scratch, no spills and the waves per SIMD its launch shape needs.  CPU only;
skipped when the library or the tools are absent.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "netrax_amd", "libnetrax_gpu.so")
LLVM = "/opt/rocm/llvm/bin"

# (mangled-name fragment, waves per SIMD the kernel's launch shape needs)
BUDGETS = [
    ("k_adjoint_persistentILi4ELi3ELi64E", 3),  # batched reverse sweep, 3 ops per group
    ("k_adjoint_persistentILi4ELi2ELi64E", 3),  # batched reverse sweep, 2 ops per group
    ("k_forward_persistentILi4ELi3ELi64E", 3),
    ("k_forward_persistentILi4ELi2ELi64E", 3),
    ("k_adjoint_levelILi4ELi4E", 3),  # plain FP64 reverse sweep, one launch
    ("k_forward_walkILi4ELi4ELi4E", 3),  # plain FP64 forward pass, one launch
    ("k_adjoint_level_aaILi4E", 2),  # 20 states FP64, one wave per category
    ("k_forward_level_aaILi4E", 2),
    ("k_adjoint_mfmaILi4E", 2),  # 20 states FP32 matrix cores
    ("k_forward_mfmaILi4ELi4E", 2),
    ("k_sumtablesILi4ELi4EdE", 2),  # branch-derivative path walks
]


@pytest.fixture(scope="module")
def kernels(tmp_path_factory):
    objcopy = shutil.which("objcopy")
    bundler = os.path.join(LLVM, "clang-offload-bundler")
    readelf = os.path.join(LLVM, "llvm-readelf")
    if not os.path.exists(LIB):
        pytest.skip("the HIP library is not built")
    if objcopy is None or not (os.path.exists(bundler) and os.path.exists(readelf)):
        pytest.skip("objcopy or the ROCm LLVM tools are not available")
    d = tmp_path_factory.mktemp("codeobj")
    fat = str(d / "fatbin.bin")
    co = str(d / "gfx950.co")
    subprocess.run([objcopy, "--dump-section", f".hip_fatbin={fat}", LIB, str(d / "lib.so")],
                   check=True, capture_output=True)
    subprocess.run([bundler, "--unbundle", "--type=o", f"--input={fat}",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"],
                   check=True, capture_output=True)
    notes = subprocess.run([readelf, "--notes", co], check=True, capture_output=True, text=True).stdout
    out = {}
    for entry in re.split(r"\n  - \.", notes):
        m = re.search(r"\n\s+\.name:\s+(\S+)", "\n" + entry)
        if not m:
            continue
        fields = {}
        for key in ("private_segment_fixed_size", "vgpr_count", "vgpr_spill_count", "sgpr_spill_count"):
            f = re.search(r"(?:^|\n)\s*\.?" + key + r":\s+(\d+)", entry)
            fields[key] = int(f.group(1)) if f else None
        out[m.group(1)] = fields
    assert out, "no kernel metadata found in the code object"
    return out


@pytest.mark.parametrize("fragment,waves", BUDGETS)
def test_hot_kernel_fits_its_register_budget(kernels, fragment, waves):
    hits = {n: f for n, f in kernels.items() if fragment in n}
    assert len(hits) == 1, f"{fragment}: {sorted(hits)}"
    (name, f), = hits.items()
    assert f["private_segment_fixed_size"] == 0, f"{name} uses scratch"
    assert f["vgpr_spill_count"] == 0 and f["sgpr_spill_count"] == 0, f"{name} spills"
    # 512 registers per SIMD lane, allocated in blocks of 8 (AGPRs included)
    regs = -(-f["vgpr_count"] // 8) * 8
    assert 512 // regs >= waves, f"{name}: {f['vgpr_count']} registers allow {512 // regs} waves per SIMD"
