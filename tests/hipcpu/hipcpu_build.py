"""Builds tests/hipcpu's CPU execution of the device BVH build under AddressSanitizer and
UndefinedBehaviorSanitizer: bvh_build.hip is compiled as C++ against hip/hip_runtime.h of this folder,
with the library's own internal.h less its device-code includes (the DevBuf buffers, PH_CHECK and the
error types are the library's, unchanged)."""
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(ROOT, "photon_amd", "csrc", "hip")


def internal_header():
    text = open(os.path.join(SRC, "internal.h")).read()
    head = text.split("// ---- render.hip")[0]
    head = head.replace('#include "dpath.h"\n', "")
    if "struct DevBuf" not in head or "PH_CHECK" not in head:
        raise RuntimeError("internal.h no longer has DevBuf and PH_CHECK before its render.hip section")
    return "#define PH_DEV inline\n" + head + "\n}  // namespace phk\n"


def build(out_dir=None, sanitize=True):
    out_dir = out_dir or os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "internal.h"), "w") as f:
        f.write(internal_header())
    shutil.copyfile(os.path.join(SRC, "bvh_build.hip"), os.path.join(out_dir, "bvh_build.cpp"))
    exe = os.path.join(out_dir, "bvh_build_cpu")
    flags = ["-std=c++20", "-O1", "-g", "-ffp-contract=off", "-pthread", "-Wall", "-Wno-unused-function"]
    if sanitize:
        flags += ["-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=undefined,float-cast-overflow", "-fno-omit-frame-pointer"]
    cmd = ["g++", *flags, f"-I{out_dir}", f"-I{HERE}", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(out_dir, "bvh_build.cpp"), os.path.join(HERE, "bvh_build_main.cpp"), "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else None))
