"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only):
the FP64 oracle's every entry point on a small scene, and the OBJ reader on
well-formed and malformed files.  Built with gcc / g++ (no device code)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
           UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="gcc/g++ not found")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=300, **kw)
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    return r.stdout


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_driver")
    subprocess.run(["gcc", "-std=gnu11", *SAN, "-I", os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c"),
                    os.path.join(ROOT, "oracle", "crt_oracle.c"), "-lm", "-o", exe], check=True)
    out = _run([exe])
    assert "oracle driver ok" in out


def test_obj_reader_under_sanitizers(tmp_path):
    exe = str(tmp_path / "obj_driver")
    subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "ceres-raytracer_amd", "csrc"),
                    os.path.join(ROOT, "tests", "sanitize", "obj_driver.cpp"), "-o", exe], check=True)
    files = {
        "good.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\nf 1/1/1 3/2/2 4/3/3\nf -4 -2 -1\n",
        "quad_fan.obj": "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0\nf 1 2 3 5 4\n",
        "long_line.obj": "# " + "x" * 5000 + "\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",
        "index_out_of_range.obj": "v 0 0 0\nv 1 0 0\nf 1 2 7\n",
        "negative_too_far.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf -1 -2 -9\n",
        "zero_index.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",
        "truncated_vertex.obj": "v 1.5\nv\nv 0 0 0\nf 1 2 3\n",
        "garbage.obj": "\x00\x01\x02 f / / /\nf //\nvn 0 0 1\nvt 0 0\nf 1/\n",
        "empty.obj": "",
        "huge_index.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 99999999999999999999\n",
    }
    paths = []
    for name, text in files.items():
        p = tmp_path / name
        p.write_bytes(text.encode("latin-1"))
        paths.append(str(p))
    out = _run([exe, *paths]).splitlines()
    got = dict(zip(files, out))
    assert got["good.obj"] == "ok 4 3"
    assert got["quad_fan.obj"] == "ok 5 3"
    assert got["long_line.obj"] == "ok 3 1"
    assert got["empty.obj"] == "ok 0 0"
    for bad in ("index_out_of_range.obj", "negative_too_far.obj", "zero_index.obj", "huge_index.obj"):
        assert got[bad].startswith("error"), (bad, got[bad])
