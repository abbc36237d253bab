"""The Rust binding crate (rust/arp-sys) against the C ABI it binds (no Rust toolchain here, so
the crate is checked as text): every function of include/arp.h is declared in its extern
block, in the same order and with the same parameter count, and the struct sizes and field
offsets its const assertions pin are the C compiler's for include/arp.h."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_RS = os.path.join(ROOT, "rust", "arp-sys", "src", "lib.rs")


def _header_functions():
    t = open(os.path.join(ROOT, "include", "arp.h")).read()
    t = re.sub(r"/\*.*?\*/", "", t, flags=re.S)
    out = []
    for m in re.finditer(r"^(?:int|void|float|double|uint32_t|const char\*)\s+(arp_\w+)\((.*?)\);", t, flags=re.S | re.M):
        out.append((m.group(1), len([a for a in m.group(2).split(",") if a.strip()])))
    return out


def _rust_functions():
    t = open(LIB_RS).read()
    block = t[t.index('extern "C" {'):]
    block = block[: block.index("\n}\n")]
    out = []
    for m in re.finditer(r"pub fn (arp_\w+)\((.*?)\)", block, flags=re.S):
        out.append((m.group(1), len([a for a in m.group(2).split(",") if a.strip()])))
    return out


def test_extern_block_matches_header():
    h = _header_functions()
    r = _rust_functions()
    assert len(h) >= 50
    assert r == h


def test_struct_layouts_match_c(tmp_path):
    t = open(LIB_RS).read()
    sizes = dict(re.findall(r"size_of::<(\w+)>\(\) == (\d+)", t))
    offs = re.findall(r"offset_of!\((\w+), (\w+)\) == (\d+)", t)
    assert len(sizes) == 9 and len(offs) >= 10
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % os.path.join(ROOT, "include", "arp.h"),
           "int main(void) {"]
    for ty in sizes:
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (ty, ty))
    for ty, f, _ in offs:
        src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (ty, f, ty, f))
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    try:
        subprocess.run(["gcc", "-o", str(exe), str(c)], check=True, capture_output=True)
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.skip("no C compiler: %s" % e)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for ty, v in sizes.items():
        assert int(got[ty]) == int(v), ty
    for ty, f, v in offs:
        assert int(got["%s.%s" % (ty, f)]) == int(v), (ty, f)
