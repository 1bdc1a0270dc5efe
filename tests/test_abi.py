"""The C-ABI library on the host (no GPU needed, no compute call).

  * libflups.so loads and exports every function include/*.h declares;
  * the topology object follows the reference's size rules (src/Topology.cpp:81-98);
  * flups_pencilDirs orders directions by the sum of their boundary codes, stable
    (src/flups.cpp:200-210, src/FFTW_plan_dim.cpp:21-27,59-65);
  * checks that fail before any device work (bad topology, bad boundary pairs)
    return NULL and leave a message in flups_get_last_error, as FLUPS_CHECK would
    have aborted (src/defines.hpp:301-304).
"""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "flups_amd", "libflups.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # drop the bodies of static inline helpers (header-only, nothing to export)
        text = re.sub(r"static inline[^{;]*\{(?:[^{}]|\{[^{}]*\})*\}", "", text, flags=re.S)
        for stmt in text.split(";"):
            s = " ".join(stmt.split())
            if s.startswith(("#", "typedef")) or "(" not in s:
                continue
            m = re.search(r"\b(flups_\w+)\s*\(", s)
            if m and not s.startswith("static"):
                names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libflups.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    L.flups_get_last_error.restype = ctypes.c_char_p
    L.flups_topo_new.restype = ctypes.c_void_p
    L.flups_topo_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                 ctypes.c_bool, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
    for f in ("flups_topo_get_nmem", "flups_topo_get_nloc", "flups_topo_get_nglob"):
        getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.flups_topo_get_memsize.restype = ctypes.c_size_t
    L.flups_topo_get_memsize.argtypes = [ctypes.c_void_p]
    L.flups_topo_get_locsize.restype = ctypes.c_size_t
    L.flups_topo_get_locsize.argtypes = [ctypes.c_void_p]
    L.flups_topo_free.argtypes = [ctypes.c_void_p]
    L.flups_malloc.restype = ctypes.c_void_p
    L.flups_malloc.argtypes = [ctypes.c_size_t]
    L.flups_free.argtypes = [ctypes.c_void_p]
    L.flups_set_abort_on_error(0)
    return L


def test_header_parse_finds_the_interface():
    names = declared_functions()
    for n in ("flups_topo_new", "flups_init", "flups_setup", "flups_solve", "flups_cleanup", "flups_get_spectralInfo"):
        assert n in names
    assert "flups_locID" not in names   # static inline in the header


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def ivec(*v):
    return (ctypes.c_int * len(v))(*v)


def topo(lib, axis, lda, n, cplx, align, nproc=(1, 1, 1)):
    return lib.flups_topo_new(axis, lda, ivec(*n), ivec(*nproc), cplx, None, align, 0)


@pytest.mark.parametrize("axis,lda,n,cplx,align,nmem", [
    (0, 1, (5, 7, 3), False, 16, (6, 7, 3)),      # 40 B % 16 = 8 -> one more double
    (0, 2, (5, 7, 3), True, 64, (8, 7, 3)),       # 80 B % 64 = 16 -> 6 doubles = 3 complex
    (1, 1, (5, 7, 3), False, 64, (5, 8, 3)),      # 56 B % 64 -> 1 double
    (2, 3, (4, 4, 16), False, 64, (4, 4, 16)),    # already aligned
])
def test_topology_sizes(lib, axis, lda, n, cplx, align, nmem):
    t = topo(lib, axis, lda, n, cplx, align)
    assert t
    nf = 2 if cplx else 1
    assert [lib.flups_topo_get_nloc(t, d) for d in range(3)] == list(n)
    assert [lib.flups_topo_get_nmem(t, d) for d in range(3)] == list(nmem)
    assert lib.flups_topo_get_locsize(t) == n[0] * n[1] * n[2] * nf
    assert lib.flups_topo_get_memsize(t) == nmem[0] * nmem[1] * nmem[2] * nf * lda
    lib.flups_topo_free(t)


def test_topology_errors(lib):
    lib.flups_clear_last_error()
    assert not topo(lib, 3, 1, (4, 4, 4), False, 16)
    assert b"axis" in lib.flups_get_last_error()
    assert not topo(lib, 0, 1, (4, 4, 4), False, 16, nproc=(2, 1, 1))
    assert b"one GPU" in lib.flups_get_last_error()
    lib.flups_clear_last_error()
    assert lib.flups_get_last_error() == b""


def test_malloc_is_aligned_and_zeroed(lib):
    p = lib.flups_malloc(1000)
    assert p and p % 64 == 0
    assert bytes((ctypes.c_ubyte * 1000).from_address(p)) == bytes(1000)
    lib.flups_free(p)


def bc_table(pairs):
    """FLUPS_BoundaryType* bc[3][2] with lda = 1"""
    cells = [(ctypes.c_int * 1)(v) for pair in pairs for v in pair]
    arr = (ctypes.c_void_p * 6)(*[ctypes.addressof(c) for c in cells])
    return arr, cells


@pytest.mark.parametrize("pairs,dirs", [
    ([(4, 4), (4, 4), (4, 4)], [0, 1, 2]),
    ([(4, 4), (3, 3), (0, 1)], [2, 1, 0]),   # types 8, 6, 1
    ([(3, 3), (4, 4), (3, 3)], [0, 2, 1]),   # stable among equal types
    ([(0, 4), (1, 1), (9, 9)], [1, 0, 2]),   # 4 (semi-unbounded), 2, 18 (empty)
])
def test_pencil_dirs(lib, pairs, dirs):
    arr, keep = bc_table(pairs)
    out = (ctypes.c_int * 3)()
    lib.flups_pencilDirs(arr, out)
    assert list(out) == dirs


def test_init_rejects_unpaired_periodic(lib):
    """a periodic side needs a periodic partner: flups_init fails before any device work"""
    t = topo(lib, 0, 1, (8, 8, 8), False, 16)
    arr, keep = bc_table([(3, 4), (4, 4), (4, 4)])
    lib.flups_init.restype = ctypes.c_void_p
    lib.flups_init.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    h = (ctypes.c_double * 3)(0.125, 0.125, 0.125)
    L = (ctypes.c_double * 3)(1, 1, 1)
    lib.flups_clear_last_error()
    s = lib.flups_init(t, arr, h, L, 0, ivec(1, 1, 1))
    assert not s
    assert b"boundary" in lib.flups_get_last_error()
    lib.flups_topo_free(t)


def test_profiler_handle_without_solver(lib, capfd):
    """profiler handles (src/flups.h:592-611) exist without a GPU: empty, displayable, freeable"""
    lib.flups_profiler_new_n.restype = ctypes.c_void_p
    lib.flups_profiler_new_n.argtypes = [ctypes.c_char_p]
    lib.flups_profiler_get.restype = ctypes.c_int
    lib.flups_profiler_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
    lib.flups_profiler_disp.argtypes = [ctypes.c_void_p]
    lib.flups_profiler_free.argtypes = [ctypes.c_void_p]
    p = lib.flups_profiler_new_n(b"cpu")
    assert p
    assert lib.flups_profiler_get(p, 0, None, 0, None, None, None) == 0
    lib.flups_profiler_disp(p)
    assert 'FLUPS profiler "cpu"' in capfd.readouterr().out
    lib.flups_profiler_free(p)
    assert lib.flups_profiler_get(None, 0, None, 0, None, None, None) == -1


def test_c_driver_is_built_against_the_header():
    """tests/c_driver/solve_c: compiled against include/flups.h and MPICH's <mpi.h>, linked to this
    libflups.so (a GPU test runs it)"""
    import subprocess
    exe = os.path.join(ROOT, "tests", "c_driver", "solve_c")
    assert os.path.exists(exe), "run __graft_entry__.build()"
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert os.path.realpath(LIB) in {os.path.realpath(l.split("=>")[1].split("(")[0].strip())
                                     for l in out.splitlines() if "libflups" in l}
    assert "libmpi" in out


def test_python_locid_matches_the_header(tmp_path):
    """flups_amd.locID equals the header's inline flups_locID (src/flups.h:109-119) for every
    source / target axis pair"""
    import subprocess
    import flups_amd as fl
    src = tmp_path / "loc.c"
    src.write_text('#include <stdio.h>\n#include "flups.h"\nint main(void){const int sz[3]={5,7,3};\n'
                   'for(int as=0;as<3;++as)for(int at=0;at<3;++at)for(int nf=1;nf<3;++nf)\n'
                   'printf("%zu\\n",flups_locID(as,1,2,2,1,at,sz,nf));return 0;}\n')
    exe = tmp_path / "loc"
    subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [fl.locID(a, 1, 2, 2, 1, b, [5, 7, 3], nf) for a in range(3) for b in range(3) for nf in (1, 2)]
    assert got == want
