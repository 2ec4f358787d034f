"""CPU: the C-ABI library loads and exports every symbol the public headers
declare (no compute call is made without a GPU)."""
import ctypes
import os
import re

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "yume-pdq_amd", "libyume_pdq.so")


def declared_symbols():
    names = set()
    for h in ("yume_pdq.h", "yume_pdq_amd.h"):
        text = open(os.path.join(REPO, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(yume_pdq_\w+)\s*\(", text))
        names |= set(re.findall(r"extern const [\w ]+?\b(YUME_PDQ_\w+)", text))
    return names


def test_headers_declare_the_reference_abi():
    names = declared_symbols()
    # include/yume_pdq.h of the reference (lines 22-25, 67-100, 122-127, 146-150)
    ref = {"yume_pdq_hash_smart_kernel", "yume_pdq_visit_dihedrals_smart_kernel",
           "YUME_PDQ_VERSION_MAJOR", "YUME_PDQ_VERSION_MINOR", "YUME_PDQ_VERSION_PATCH",
           "YUME_PDQ_VERSION_STRING"} | {"YUME_PDQ_DIHEDRAL_" + d for d in (
               "NORMAL", "FLIPPED", "FLOPPED", "ROTATED_180", "ROTATED_90", "ROTATED_270",
               "ROTATED_90_FLOPPED", "FLOPPED_ROTATED_270")}
    assert ref <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_data_symbols_values():
    lib = ctypes.CDLL(LIB)
    v = lambda n: ctypes.c_uint32.in_dll(lib, n).value  # noqa: E731
    assert (v("YUME_PDQ_VERSION_MAJOR"), v("YUME_PDQ_VERSION_MINOR"), v("YUME_PDQ_VERSION_PATCH")) == (1, 1, 0)
    assert ctypes.c_char_p.in_dll(lib, "YUME_PDQ_VERSION_STRING") is not None
    s = (ctypes.c_char * 6).in_dll(lib, "YUME_PDQ_VERSION_STRING").value
    assert s == b"1.1.0"
    # Dihedrals::from_tuples, src/lib.rs:96-133 (big-endian dx0, dx1, dy0, dy1)
    exp = {"NORMAL": 0x01000001, "FLIPPED": 0x010000FF, "FLOPPED": 0xFF000001,
           "ROTATED_180": 0xFF0000FF, "ROTATED_90": 0x0001FF00, "ROTATED_270": 0x00010100,
           "ROTATED_90_FLOPPED": 0x00FFFF00, "FLOPPED_ROTATED_270": 0x00FF0100}
    for k, val in exp.items():
        assert v("YUME_PDQ_DIHEDRAL_" + k) == val, k


def test_strerror_is_host_only():
    lib = ctypes.CDLL(LIB)
    lib.yume_pdq_amd_strerror.restype = ctypes.c_char_p
    assert lib.yume_pdq_amd_strerror(-1) == b"invalid argument"
    assert lib.yume_pdq_amd_strerror(0) == b"ok"


def test_c_program_compiles_against_both_headers(tmp_path):
    """A C caller of the reference ABI (like integration/dummy.c) compiles and links."""
    import subprocess
    src = tmp_path / "caller.c"
    src.write_text('#include "yume_pdq.h"\n#include "yume_pdq_amd.h"\n'
                   "int main(void){ float t; unsigned char o[32]; static float in[512*512], b[128*128], tm[128], p[256];\n"
                   "  (void)YUME_PDQ_DIHEDRAL_NORMAL; if (0) yume_pdq_hash_smart_kernel(in,&t,o,b,tm,p);\n"
                   "  return YUME_PDQ_VERSION_MAJOR == 1 ? 0 : 1; }\n")
    exe = tmp_path / "caller"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"), str(src),
                    "-o", str(exe), "-L", os.path.dirname(LIB), "-lyume_pdq",
                    "-Wl,-rpath," + os.path.dirname(LIB)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0
