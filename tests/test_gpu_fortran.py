"""The Fortran drop-in layer on the GPU: prueba_rd (built from
redondito_amd/fortran/) calls the reference's routine names with host arrays
whose device copies are registered, and checks closed-form answers."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "redondito_amd", "_lib", "prueba_rd")


def test_prueba_rd_known_answers(tmp_path):
    """Runs in a scratch directory: box_init writes theta.txt, phi.txt and
    z.txt there, and the program writes the namelist it reads."""
    assert os.path.exists(EXE), "prueba_rd not built (run __graft_entry__.build())"
    env = dict(os.environ, RD_TABLE_DIR=os.path.join(ROOT, "redondito_amd", "tables"))
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120, env=env, cwd=str(tmp_path))
    print(r.stdout)
    print(r.stderr[-2000:])
    checks = [l for l in r.stdout.splitlines() if l.startswith("CHECK")]
    assert len(checks) >= 35
    assert (tmp_path / "theta.txt").exists() and (tmp_path / "z.txt").exists()
    assert r.returncode == 0 and all(l.rstrip().endswith("PASS") for l in checks)
