"""CPU checks of the built gfx950 code objects: every kernel of
libeigen_amd.so runs without scratch memory (register spills or
runtime-indexed private arrays would be served from it, at HBM latency),
read from the kernel descriptors' metadata of the library's own embedded
code objects."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "eigen_amd", "libeigen_amd.so")
LLVM = "/opt/rocm/llvm/bin"


def kernel_metadata(tmp_path):
    """(name, private_segment_fixed_size, vgpr_count) of every gfx950 kernel."""
    if not os.path.exists(os.path.join(LLVM, "llvm-objdump")):
        pytest.skip("llvm-objdump not found")
    lib = tmp_path / "lib.so"
    shutil.copy(LIB, lib)
    # extracts each offload bundle next to its input, as lib.so.<i>.<triple>
    subprocess.run([os.path.join(LLVM, "llvm-objdump"), "--offloading", str(lib)],
                   check=True, capture_output=True, cwd=tmp_path)
    out = []
    for f in sorted(os.listdir(tmp_path)):
        if "amdgcn-amd-amdhsa" not in f or "gfx950" not in f:
            continue
        notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes",
                                str(tmp_path / f)], check=True, capture_output=True,
                               text=True).stdout
        for block in notes.split("  - .")[1:]:
            name = re.search(r"\.name:\s+(\S+)", block)
            scratch = re.search(r"\.private_segment_fixed_size:\s+(\d+)", block)
            vgpr = re.search(r"\.vgpr_count:\s+(\d+)", block)
            if name and scratch:
                out.append((name.group(1), int(scratch.group(1)),
                            int(vgpr.group(1)) if vgpr else -1))
    return out


def test_no_kernel_uses_scratch(tmp_path):
    meta = kernel_metadata(tmp_path)
    assert len(meta) >= 40, f"only {len(meta)} kernels found"
    spilling = [(n, s) for n, s, _ in meta if s]
    assert not spilling, spilling
    # the batched kernels are all there (the check above covers them)
    names = " ".join(n for n, _, _ in meta)
    for k in ("llt_batch_persistent", "lu_ir_batch_kernel", "ldlt_batch_kernel",
              "ldlt_solve_batch_kernel", "llt_rank_update_kernel", "ldlt_rank_update_kernel"):
        assert k in names, k
