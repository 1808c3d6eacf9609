"""Checks on the gfx950 assembly of the Legendre kernels that issue LDS moves
(global_load_lds_dwordx4) from inline asm.  CPU only: hipcc cross-compiles the
device code here.

* Every SALU write of M0 is followed by at least one wait state (s_nop) before
  an LDS move reads M0 (a gfx9 hazard; the compiler inserts the same s_nop for
  its own LDS-move builtin).
* M0 is written only inside those inline-asm blocks, so nothing the compiler
  emits depends on a value the asm overwrote.
* In each analysis kernel the wait before every s_barrier is vmcnt(0), so the
  moves a chunk issued have landed whatever number of table loads the compiler
  placed around them.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "redondito_amd", "csrc", "rd_legendre.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("asm") / "rd_legendre.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "--cuda-device-only", "-S", SRC, "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return out.read_text()


def functions(text):
    """{symbol: [instruction lines]} for every kernel body in the file."""
    funcs, cur = {}, None
    for line in text.splitlines():
        m = re.match(r"^(_Z\w+):\s*(;.*)?$", line)
        if m:
            cur = m.group(1)
            funcs[cur] = []
            continue
        if cur is not None:
            if line.startswith("\t.end_amdgpu_kernel") or re.match(r"^\.Lfunc_end", line):
                cur = None
                continue
            funcs[cur].append(line.strip())
    return funcs


def instructions(lines):
    """Instruction lines only (no labels, directives or comments), with a flag
    for lines inside an inline-asm block."""
    out, in_asm = [], False
    for s in lines:
        if s.startswith(";;#ASMSTART"):
            in_asm = True
            continue
        if s.startswith(";;#ASMEND"):
            in_asm = False
            continue
        if not s or s.startswith((";", ".")) or s.endswith(":"):
            continue
        out.append((s.split(";")[0].strip(), in_asm))
    return out


def test_m0_write_has_a_wait_state_before_each_lds_move(asm):
    n_moves = 0
    for name, body in functions(asm).items():
        ins = instructions(body)
        for i, (s, _) in enumerate(ins):
            if not (s.startswith("global_load_lds") or (s.startswith("buffer_load") and " lds" in s)):
                continue
            n_moves += 1
            prev = ins[i - 1][0]
            assert not prev.startswith("s_mov_b32 m0"), f"{name}: M0 written right before an LDS move"
    assert n_moves > 0, "the analysis kernels issue LDS moves"


def test_m0_is_written_only_by_the_lds_move_asm(asm):
    for name, body in functions(asm).items():
        for s, in_asm in instructions(body):
            if re.search(r"\bm0\b", s) and not in_asm:
                raise AssertionError(f"{name}: compiler-emitted use of M0: {s}")


def test_analysis_chunk_waits_cover_every_move(asm):
    kernels = {n: b for n, b in functions(asm).items() if "15analysis_kernel" in n}
    assert len(kernels) == 4, sorted(kernels)
    for name, body in kernels.items():
        ins = [s for s, _ in instructions(body)]
        assert any(s.startswith("global_load_lds") for s in ins), name
        for i, s in enumerate(ins):
            if s != "s_barrier":
                continue
            # the nearest vmcnt wait before the barrier, with nothing that
            # issues memory operations in between
            j = i - 1
            while j >= 0 and not ins[j].startswith("s_waitcnt"):
                assert not re.match(r"(global|buffer|flat)_", ins[j]), f"{name}: memory op between wait and barrier"
                j -= 1
            assert j >= 0 and "vmcnt(0)" in ins[j], f"{name}: barrier after '{ins[j] if j >= 0 else None}'"
