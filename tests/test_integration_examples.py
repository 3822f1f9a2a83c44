"""CPU checks that the integration guide's examples match the library: its C++
wrapper (the binding a VkNRC maintainer would add next to src/VkNRCState.hpp)
compiles against include/nrc.h and links against libvknrc.so, and every module
attribute its Python snippets use exists.  Nothing is run on a device."""
import ast
import os
import re
import shutil
import subprocess

import pytest

import vknrc_amd as v
from vknrc_amd import autotune, serving

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GUIDE = os.path.join(ROOT, "INTEGRATION.md")
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)


def code_blocks(lang):
    text = open(GUIDE).read()
    return re.findall(r"^```" + lang + r"\n(.*?)^```", text, flags=re.S | re.M)


@pytest.mark.parametrize("compiler", ["hipcc", "g++"])
def test_cpp_wrapper_compiles_and_links(tmp_path, compiler):
    if HIPCC is None:
        pytest.skip("ROCm is not installed")
    if compiler == "g++" and shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    blocks = code_blocks("cpp")
    assert len(blocks) == 1
    (tmp_path / "HipNRC.hpp").write_text(blocks[0])
    names = re.findall(r"^\t(?:explicit\s+)?(?:void\s+)?(\w+)\(", blocks[0], flags=re.M)
    members = [n for n in names if n not in ("HipNRC", "check")]
    assert {"Frame", "FrameRecords", "ResetMLPBuffers", "SetUseEMAWeights"} <= set(members)
    # Taking each member's address makes the compiler emit it, so the link
    # resolves every nrc_* call the wrapper makes.
    takes = "".join(f"\tauto p_{n} = &HipNRC::{n};\n\t(void)p_{n};\n" for n in members)
    (tmp_path / "main.cpp").write_text(
        '#include "HipNRC.hpp"\nint main(int argc, char **) {\n' + takes +
        "\tif (argc > 1000) { HipNRC n(1); }\n\treturn 0;\n}\n")
    if compiler == "hipcc":
        cmd = [HIPCC, "-Wno-unused-command-line-argument"]
    else:  # a host compiler needs the platform macro and include path for hipStream_t
        rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
        cmd = ["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(rocm, "include")]
    cmd += ["-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), "-I" + str(tmp_path),
            str(tmp_path / "main.cpp"), "-L" + os.path.join(ROOT, "vknrc_amd"), "-lvknrc", "-o", str(tmp_path / "main")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]


def attribute_uses(src):
    """(module alias, attribute) pairs for v.X, autotune.X and serving.X."""
    out = set()
    for node in ast.walk(ast.parse(src)):
        if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name):
            if node.value.id in ("v", "autotune", "serving"):
                out.add((node.value.id, node.attr))
    return out


def test_python_snippets_name_existing_attributes():
    blocks = code_blocks("python")
    assert len(blocks) >= 3
    mods = {"v": v, "autotune": autotune, "serving": serving}
    uses = set()
    for b in blocks:
        uses |= attribute_uses(b.replace("for ...]", "for _ in ()]"))
    assert ("v", "NRCState") in uses and ("serving", "Server") in uses
    missing = [f"{m}.{a}" for m, a in sorted(uses) if not hasattr(mods[m], a)]
    assert not missing, missing
    # methods the snippets call on the objects they make
    for cls, meth in ((v.NRCState, "infer"), (v.NRCState, "train"), (v.NRCState, "get_state"),
                      (v.GroupGraph, "launch"), (v.DeviceScene, "pack_primitives"),
                      (serving.Server, "run_realtime")):
        assert hasattr(cls, meth), (cls.__name__, meth)
