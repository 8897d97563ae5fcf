"""The host-side case reader and layout builder under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only; no device code is built).

tests/host_sanitize/case_harness.cpp is linked with foam_case.cpp and
mesh_builder.cpp compiled by g++ with -fsanitize=address,undefined, then run
over valid cases (ASCII, binary, decomposed; 2D and 3D) and over copies with
one file corrupted (truncated, bytes overwritten, list size enlarged or made
negative).  Every case must be read or refused with a message: no sanitizer
report, no abort, and every valid case laid out.
"""
import os
import re
import shutil
import subprocess

import pytest

import small_cases as sc
from lfm_public_amd import foam, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRCS = [os.path.join(ROOT, "tests", "host_sanitize", "case_harness.cpp"),
        os.path.join(ROOT, "lfm_public_amd", "csrc", "foam_case.cpp"),
        os.path.join(ROOT, "lfm_public_amd", "csrc", "mesh_builder.cpp")]
FLAGS = ["-std=c++17", "-g", "-O0", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
         "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "lfm_public_amd", "csrc")]
FILES = ["constant/polyMesh/points", "constant/polyMesh/faces", "constant/polyMesh/owner",
         "constant/polyMesh/neighbour", "constant/polyMesh/boundary", "0/p", "0/U", "system/controlDict"]


def _build(tmp):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    objs, procs = [], []
    for s in SRCS:
        o = os.path.join(tmp, os.path.basename(s) + ".o")
        objs.append(o)
        procs.append(subprocess.Popen(["g++", *FLAGS, "-c", s, "-o", o], stderr=subprocess.PIPE))
    for p in procs:
        _, err = p.communicate()
        assert p.returncode == 0, err.decode()
    exe = os.path.join(tmp, "case_harness")
    subprocess.check_call(["g++", "-fsanitize=address,undefined", *objs, "-o", exe])
    return exe


def _corrupt(data: bytes, mode: int, k: int) -> bytes:
    b = bytearray(data)
    if mode == 0:  # truncated
        return bytes(b[: len(b) // 2 + k])
    if mode == 1:  # bytes overwritten past the header
        for j in range(16):
            i = len(b) // 3 + (j * 7919 + k * 104729) % max(1, len(b) - len(b) // 3)
            b[i] = (j * 37 + k * 11) % 256
        return bytes(b)
    txt = b.decode("latin1")
    m = re.search(r"\n(\d+)\s*\n\(", txt)
    if not m:
        return bytes(b)
    new = str(int(m.group(1)) * 3 + 7) if mode == 2 else "-5"
    return (txt[:m.start(1)] + new + txt[m.end(1):]).encode("latin1")


def test_case_reader_and_builder_under_sanitizers(tmp_path):
    exe = _build(str(tmp_path))
    cases_dir = os.path.join(str(tmp_path), "cases")
    valid, args = [], []
    for name, c in [("vortex", sc.vortex(6)), ("cyl", sc.cylinder(24)), ("ogrid", sc.ogrid(3, 8, 3)),
                    ("box", sc.box(3))]:
        for binary in (False, True):
            d = os.path.join(cases_dir, f"{name}_{'bin' if binary else 'asc'}")
            foam.write_case(d, c, end_steps=4, write_interval=2, binary=binary)
            valid.append(d)
            args.append(d)
    c = sc.cylinder(24)
    d = os.path.join(cases_dir, "cyl_dec")
    foam.write_case(d, c, end_steps=4, cell_rank=meshgen.slab_ranks(c.mesh, 2, axis=0))
    args += [d + ":0", d + ":1"]
    k = 0
    for base in ("cyl_asc", "cyl_bin", "ogrid_bin", "box_asc"):
        for f in FILES:
            src = os.path.join(cases_dir, base, f)
            data = open(src, "rb").read()
            for mode in range(4):
                dst = os.path.join(cases_dir, f"bad{k}")
                shutil.copytree(os.path.join(cases_dir, base), dst)
                open(os.path.join(dst, f), "wb").write(_corrupt(data, mode, k))
                args.append(dst)
                k += 1
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe, *args], capture_output=True, env=env, timeout=600)
    log = (r.stdout + r.stderr).decode("latin1")  # messages quote corrupted bytes
    assert r.returncode == 0, log[-4000:]
    assert "AddressSanitizer" not in log and "runtime error" not in log, log[-4000:]
    for d in valid + [os.path.join(cases_dir, "cyl_dec")]:
        blk = log[log.index(d + " rank"):]
        assert "layout: ok" in blk.split("\n", 3)[2], blk[:300]
    assert log.count("open rc 0") < len(args)  # the corrupted cases are refused, not all accepted
