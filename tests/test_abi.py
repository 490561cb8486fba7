"""CPU-side checks of the C ABI and of the host-side setup (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from bosss_amd import _lib
from bosss_amd.grid import CartesianGrid
from bosss_amd.refelement import Tables
from oracle.bosss_ref import RefModel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "bosss_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bosss_hip_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIBPATH)
    syms = declared_symbols()
    assert len(syms) >= 17
    for s in syms:
        assert hasattr(L, s), s
    # and the Python binding covers exactly the declared set
    assert sorted(_lib.SIGNATURES) == syms


def test_error_path_without_gpu_is_reported():
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.bosss_hip_ctx_create(-1, ctypes.byref(h))
    assert rc < 0
    assert L.bosss_hip_last_error().decode()


@pytest.mark.parametrize("D,n,p,order,per", [(2, (4, 4), 2, 4, False), (3, (3, 3, 3), 3, 9, True), (3, (3, 4, 2), 2, 4, False),
                                             (3, (2, 2, 2), 3, 6, True), (2, (3, 5), 2, 6, True)])
def test_host_setup_matches_oracle(D, n, p, order, per):
    g = CartesianGrid(n, periodic=[per] * D)
    t = Tables(D, p, order)
    m = RefModel(D, n, p, order, periodic=[int(per)] * D)
    rg, rt = m.grid(), m.tables()
    assert (g.cells == rg["cells"]).all() and (g.trafo == rg["faces"]).all()
    for a, b in ((g.lenscale, rg["lenscale"]), (g.normals, rg["normals"]), (g.sqrtgram, rg["sqrtgram"]), (g.detJ, rg["detJ"]),
                 (g.invJ, rg["invJ"]), (g.scaling, rg["scaling"])):
        assert np.abs(a - b).max() <= 1e-15 * max(1.0, np.abs(b).max())
    # closed-form basis/Gauss tables vs the reference's printed coefficient tables
    for a, b in ((t.phi, rt["phi"]), (t.dphi, rt["dphi"]), (t.trace, rt["trace"]), (t.dtrace, rt["dtrace"]),
                 (t.vw, rt["vw"]), (t.fw, rt["fw"]), (t.vnodes, rt["vnodes"])):
        assert np.abs(a - b).max() < 5e-14 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("D,per", [(2, False), (3, False), (3, True)])
def test_host_grid_on_graded_nodes_matches_oracle(D, per):
    """Per-cell metrics of a graded tensor grid (node arrays per direction, a different cell
    count in each): the product's host setup equals the oracle's."""
    rng = np.random.default_rng(4)
    nodes = []
    for d in range(D):
        w = rng.uniform(1.0, 3.0, 3 + d)
        nodes.append(np.concatenate([[0.0], np.cumsum(w) / w.sum()]))
    g = CartesianGrid(nodes=nodes, periodic=[per] * D)
    m = RefModel(D, None, 2, 4, periodic=[int(per)] * D, nodes=nodes)
    rg = m.grid()
    assert (g.cells == rg["cells"]).all() and (g.trafo == rg["faces"]).all()
    for a, b in ((g.lenscale, rg["lenscale"]), (g.normals, rg["normals"]), (g.sqrtgram, rg["sqrtgram"]), (g.detJ, rg["detJ"]),
                 (g.invJ, rg["invJ"]), (g.scaling, rg["scaling"])):
        assert np.abs(a - b).max() <= 1e-15 * max(1.0, np.abs(b).max())
    assert len(set(np.round(g.detJ, 14))) > g.J // 2          # the cells really differ


def test_block_lu_refuses_sizes_before_any_launch():
    """The batched LU checks its sizes on the host before touching the device: nmax above
    4096 is an argument error, an inverse of blocks above 128 DOF is unsupported; the error
    text names the limit (host memory stands in for the device arrays: nothing is read)."""
    L = _lib.lib()
    n = np.array([5000], np.int32)
    off = np.zeros(1, np.int64)
    A = np.zeros(1)
    info = np.zeros(1, np.int32)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = L.bosss_hip_block_lu_dev(None, 1, 5000, ptr(n), ptr(off), ptr(A), None, None, ptr(info), 0, None)
    assert rc == -1 and "4096" in L.bosss_hip_block_last_error().decode()
    rc = L.bosss_hip_block_lu_dev(None, 1, 200, ptr(n), ptr(off), ptr(A), None, None, ptr(info), 1, None)
    assert rc == -3 and "128" in L.bosss_hip_block_last_error().decode()
    rc = L.bosss_hip_block_lu_dev(None, 1, 20, None, ptr(off), ptr(A), None, None, ptr(info), 0, None)
    assert rc == -1


def test_cpp_host_binds_every_declared_symbol_by_name():
    """The plain C++ host (tests/cpp/abi_harness.cpp) lists exactly the entry points of
    include/bosss_hip.h and resolves each one by name from the built library (no GPU needed)."""
    import subprocess
    src = open(os.path.join(ROOT, "tests", "cpp", "abi_harness.cpp")).read()
    m = re.search(r"#define BOSSS_API\(X\)(.*?)\n\n", src, flags=re.S)
    listed = sorted("bosss_hip_" + s for s in re.findall(r"X\((\w+)\)", m.group(1)))
    assert listed == declared_symbols()
    exe = os.path.join(ROOT, "tests", "cpp", "abi_harness")
    if not os.path.exists(exe):
        pytest.skip("harness not built")
    r = subprocess.run([exe, _lib.LIBPATH, "bind"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
