"""Kernel metadata of this project's own fused code objects (hipcc output):
the AMDGPU code-object notes give each kernel's private segment (lane
scratch) size, which the runtime's scratch guard reads through HIP."""
import os
import re
import subprocess
import tempfile

LLVM = "/opt/rocm/llvm/bin"


def private_segment_sizes(hsaco):
    """{kernel name: private_segment_fixed_size} of a gfx950 code object
    bundle written by fuser.compile_source."""
    with tempfile.TemporaryDirectory() as td:
        co = os.path.join(td, "k.co")
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o",
                        "--input=" + hsaco, "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--output=" + co],
                       check=True, capture_output=True)
        notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", co], check=True,
                               capture_output=True, text=True).stdout
    out = {}
    name = None
    for line in notes.splitlines():
        m = re.match(r"\s*\.name:\s+(\S+)", line)
        if m:
            name = m.group(1)
        m = re.match(r"\s*\.private_segment_fixed_size:\s+(\d+)", line)
        if m and name:
            out[name] = int(m.group(1))
    return out
