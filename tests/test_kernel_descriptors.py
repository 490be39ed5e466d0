"""Every kernel of the built engine library (its gfx950 code object, read from the
library's fat binary) has a hidden-argument block after its explicit arguments.

A kernel whose code never reads the grid or workgroup size gets none, so its argument
segment ends at its explicit arguments; a tool that reads a dispatch's hidden arguments
(a profiler's dispatch record) then reads past the end of the argument buffer. The
replica kernel was such a kernel; it now checks its grid against the replica count.
Needs the ROCm LLVM tools (present in this image); built-library code only. The library
is only read: its fat-binary section is taken from the ELF section table here (an
objcopy run without an output file rewrites its input in place, which corrupts the
mapping of a process that has the library loaded).
"""

import os
import struct
import subprocess

import pytest
import yaml

from oxdna import _engine

LLVM = "/opt/rocm/llvm/bin"


MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _elf_section(path, name):
    """The bytes of section `name` of the 64-bit little-endian ELF file at `path`."""
    data = open(path, "rb").read()
    assert data[:4] == b"\x7fELF" and data[4] == 2 and data[5] == 1, "not a 64-bit little-endian ELF"
    shoff, = struct.unpack_from("<Q", data, 0x28)
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", data, 0x3A)

    def header(i):
        # sh_name, sh_type, sh_flags, sh_addr, sh_offset, sh_size
        return struct.unpack_from("<IIQQQQ", data, shoff + i * shentsize)
    stroff = header(shstrndx)[4]
    for i in range(shnum):
        nm, _, _, _, off, size = header(i)
        end = data.index(b"\0", stroff + nm)
        if data[stroff + nm:end].decode() == name:
            return data[off:off + size]
    raise KeyError(name)


def _code_objects(tmp_path):
    """The gfx950 code object of every offload bundle in the library's fat-binary
    section (one per translation unit with device code)."""
    lib = _engine.LIB_PATH
    if not os.path.exists(lib):
        pytest.fail(f"{lib} is missing: build it first")
    if not os.path.exists(os.path.join(LLVM, "clang-offload-bundler")):
        pytest.skip("clang-offload-bundler not available")
    data = _elf_section(lib, ".hip_fatbin")
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
