"""Every kernel of the built engine library (its gfx950 code object, read from the
library's fat binary) has a hidden-argument block after its explicit arguments.

A kernel whose code never reads the grid or workgroup size gets none, so its argument
segment ends at its explicit arguments; a tool that reads a dispatch's hidden arguments
(a profiler's dispatch record) then reads past the end of the argument buffer. The
replica kernel was such a kernel; it now checks its grid against the replica count.
Needs objcopy and the ROCm LLVM tools (present in this image); built-library code only.
"""

import os
import shutil
import subprocess

import pytest
import yaml

from oxdna import _engine

LLVM = "/opt/rocm/llvm/bin"


MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _code_objects(tmp_path):
    """The gfx950 code object of every offload bundle in the library's fat-binary
    section (one per translation unit with device code)."""
    lib = _engine.LIB_PATH
    if not os.path.exists(lib):
        pytest.fail(f"{lib} is missing: build it first")
    for tool in ("objcopy",):
        if not shutil.which(tool):
            pytest.skip(f"{tool} not available")
    if not os.path.exists(os.path.join(LLVM, "clang-offload-bundler")):
        pytest.skip("clang-offload-bundler not available")
    fatbin = tmp_path / "fatbin.bin"
    subprocess.check_call(["objcopy", "--dump-section", f".hip_fatbin={fatbin}", lib])
    data = fatbin.read_bytes()
    starts = []
    k = data.find(MAGIC)
    while k >= 0:
        starts.append(k)
        k = data.find(MAGIC, k + len(MAGIC))
    assert starts and starts[0] == 0
    out = []
    for n, (a, b) in enumerate(zip(starts, starts[1:] + [len(data)])):
        part, co = tmp_path / f"bundle{n}.bin", tmp_path / f"gfx950_{n}.co"
        part.write_bytes(data[a:b])
        subprocess.check_call([os.path.join(LLVM, "clang-offload-bundler"), "--type=o", f"--input={part}",
                               f"--output={co}", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--unbundle"])
        out.append(co)
    return out


def _kernels(co):
    notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", str(co)], capture_output=True,
                           text=True, check=True).stdout
    i = notes.index("---")
    j = notes.index("...", i)
    return yaml.safe_load(notes[i:j])["amdhsa.kernels"]


def test_every_kernel_has_hidden_arguments(tmp_path):
    kernels = [k for co in _code_objects(tmp_path) for k in _kernels(co)]
    assert len(kernels) > 20
    missing = []
    for k in kernels:
        explicit = sum(a[".size"] for a in k[".args"] if not a[".value_kind"].startswith("hidden"))
        hidden = [a for a in k[".args"] if a[".value_kind"].startswith("hidden")]
        if not hidden or k[".kernarg_segment_size"] <= explicit:
            missing.append(k[".name"])
    assert not missing, missing
    names = [k[".name"] for k in kernels]
    assert any("k_ensemble" in n for n in names) and any("k_forces" in n for n in names)
