"""The C-ABI library loads and exports every entry point include/rp_abi.h
declares; argument errors come back as codes, not crashes (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

from romspath_amd import abi

ROOT = Path(__file__).resolve().parents[1]


def declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(rp_[a-z_0-9]+)\s*\(", txt))
    return names


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(str(abi.LIB_PATH))
    names = declared()
    assert "rp_step" in names and len(names) >= 14
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_binding_covers_header():
    assert declared() == set(abi.EXPORTS)


def test_bad_arguments_return_codes():
    L = abi.lib()
    assert L.rp_create(None, None) == -1
    assert L.rp_last_error(None) == b"null context"
    assert L.rp_destroy(None) == -1
    cfg = abi.RpConfig()
    cfg.precision = 7
    h = C.c_void_p()
    rc = L.rp_create(C.byref(h), C.byref(cfg))
    assert rc in (-1, -6)          # RP_E_ARG with a GPU, RP_E_NODEV without
    assert not h.value


def test_fortran_module_binds_every_entry_point():
    """romspath_hip.f90, the ISO_C_BINDING module a Fortran host uses, has an
    interface for every entry point the headers declare."""
    f90 = (ROOT / "romspath_amd" / "fortran" / "romspath_hip.f90").read_text()
    bound = set(re.findall(r'bind\(C,\s*name="(rp_[a-z_0-9]+)"\)', f90))
    assert declared() == bound
