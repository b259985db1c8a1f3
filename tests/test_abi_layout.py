"""The ctypes mirrors of the C ABI's structs (crt/capi.py) have the C
layout: size and every field offset, as gcc lays out include/crt_hip.h."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from crt import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = [("crt_camera", capi.Camera), ("crt_light", capi.Light), ("crt_build_opts", capi.BuildOpts),
         ("crt_bvh_info", capi.BvhInfo), ("crt_timing", capi.Timing), ("crt_render_opts", capi.RenderOpts),
         ("crt_material", capi.Material), ("crt_texture", capi.Texture), ("crt_batch_opts", capi.BatchOpts),
         ("crt_flash_opts", capi.FlashOpts), ("crt_flux_opts", capi.FluxOpts), ("crt_server_opts", capi.ServerOpts),
         ("crt_request", capi.Request), ("crt_server_stats", capi.ServerStats)]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_struct_layouts_match_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "crt_hip.h"', "int main(void) {"]
    for cname, py in PAIRS:
        lines.append('printf("%s __sizeof__ %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            if f.startswith("pad"):
                continue
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            a, b, c = ln.split()
            got[(a, b)] = int(c)
    for cname, py in PAIRS:
        assert got[(cname, "__sizeof__")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            if f.startswith("pad"):
                continue
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
