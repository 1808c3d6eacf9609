"""The C-ABI library loads and exports every entry point include/redondito.h
declares; host-only entry points (tables, error strings) behave like the
reference on the CPU.  No device compute here."""
import ctypes
import os
import re
import shutil

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "redondito.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(rd_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("rd_plan_create", "rd_sht_analysis", "rd_sht_synthesis", "rd_vsht_analysis",
                 "rd_vsht_synthesis", "rd_z_forward", "rd_z_inverse", "rd_3d_forward", "rd_3d_inverse",
                 "rd_derivr", "rd_dlr", "rd_prodvec", "rd_nonlinear_uxw", "rd_load_dirichlet_tables",
                 "rd_derivr_mat", "rd_bc_setup", "rd_invdl", "rd_v_imposebc", "rd_v_imposebc_fast",
                 "rd_nonlinear_coupled", "rd_batch_create", "rd_batch_capture", "rd_batch_tune"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from redondito_amd import sphp as S
    L = S.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_error_strings():
    from redondito_amd import sphp as S
    L = S.lib()
    assert L.rd_error_string(0) == b"ok"
    assert b"table" in L.rd_error_string(-3)


def test_load_dirichlet_tables_matches_oracle():
    from redondito_amd import sphp as S
    dr = S.load_dirichlet_tables(S.TABLE_DIR, 25, 5)
    A, Q = O.load_tables(S.TABLE_DIR, 25, 5)
    assert np.array_equal(dr, O.fc_dir(A, Q))          # same summation order
    assert np.max(np.abs(dr - A @ Q.T)) < 1e-15 * np.max(np.abs(dr)) * 5


def test_missing_and_corrupt_tables(tmp_path):
    """fcgram_mod.f90:232-254 stops on a missing table; here an error code."""
    from redondito_amd import sphp as S
    with pytest.raises(S.RdError, match="table"):
        S.load_dirichlet_tables(str(tmp_path), 25, 5)
    shutil.copy(os.path.join(S.TABLE_DIR, "Q5.dat"), tmp_path / "Q5.dat")
    (tmp_path / "A25-5.dat").write_bytes(b"\0" * 999)
    with pytest.raises(S.RdError, match="table"):
        S.load_dirichlet_tables(str(tmp_path), 25, 5)


def test_plan_rejects_bad_sizes_without_touching_the_device():
    """Argument checks run before any HIP call."""
    from redondito_amd import sphp as S
    L = S.lib()
    h = ctypes.c_void_p()
    assert L.rd_plan_create(ctypes.byref(h), 15, 23, 48, 6, 0, 0, 1.0, 0.0, None) == -1   # odd nlat
    assert L.rd_plan_create(ctypes.byref(h), 15, 24, 30, 6, 0, 0, 1.0, 0.0, None) == -1   # nphi <= 2 lmax
    assert L.rd_plan_create(ctypes.byref(h), 15, 24, 48, 30, 25, 5, 1.0, 0.0, None) == -1  # nr < 2d, no dir
    assert L.rd_plan_create(ctypes.byref(h), 15, 24, 48, 6, 0, 3, 1.0, 0.0, None) == -1   # C/d mismatch


def test_python_mirror_rejects_unknown_space():
    from redondito_amd import sphp as S
    with pytest.raises(S.RdError, match="Unknown space"):
        S.derivr(None, None, None, 1, "x")
    with pytest.raises(S.RdError, match="Unknown space"):
        S.fc_filter(None, None, "q")


@pytest.mark.parametrize("ord_", [1, 2])
def test_load_neumann_tables_matches_oracle(ord_):
    """rd_load_neumann_tables (host) against the oracle's restatement of
    load_neumann_tables (fcgram_mod.f90:267-371)."""
    from redondito_amd import sphp as S
    dx = 0.0123
    neu = S.load_neumann_tables(S.TABLE_DIR, 5, ord_, dx)
    Q, dxp, Qn = O.load_neumann_files(S.TABLE_DIR, 5, ord_)
    ref = O.neumann_table(5, ord_, dx, Q, Qn, dxp)
    assert np.array_equal(neu, ref)


def test_neumann_tables_missing_and_undefined_order(tmp_path, capfd):
    """A missing table is an error (fcgram_mod.f90:331-352 stops); an undefined
    order prints the reference's warning and leaves the table unwritten."""
    from redondito_amd import sphp as S
    with pytest.raises(S.RdError, match="table"):
        S.load_neumann_tables(str(tmp_path), 5, 1, 0.1)
    out = S.load_neumann_tables(S.TABLE_DIR, 5, 3, 0.1)
    assert np.all(np.isnan(out))
    assert "undefined order" in capfd.readouterr().err
