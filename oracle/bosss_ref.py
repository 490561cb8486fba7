"""ctypes binding of ``libbosss_ref.so`` (the C++ restatement of the BoSSS evaluator).

Test infrastructure only; see ``bosss_ref.cpp`` for the reference files each
routine follows.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'libbosss_ref.so')
TABLES = os.path.join(os.path.dirname(HERE), 'tests', 'golden')

OP_SIP, OP_ADVDIFF, OP_EULER_LF, OP_EULER_HLLC = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)
_lib = None


def build():
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    subprocess.check_call(['make', '-s', '-C', HERE])


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.bref_create.restype = ctypes.c_void_p
        lib.bref_create.argtypes = [ctypes.c_char_p, ctypes.c_int, _ip, _dp, _dp, _ip, ctypes.c_int, ctypes.c_int]
        lib.bref_create_nodes.restype = ctypes.c_void_p
        lib.bref_create_nodes.argtypes = [ctypes.c_char_p, ctypes.c_int, _ip, _dp, _ip, ctypes.c_int, ctypes.c_int]
        lib.bref_destroy.argtypes = [ctypes.c_void_p]
        lib.bref_info.argtypes = [ctypes.c_void_p, _ip]
        lib.bref_last_error.restype = ctypes.c_char_p
        lib.bref_export_grid.argtypes = [ctypes.c_void_p] + [_dp] * 4 + [_ip, _ip, _dp, _dp]
        lib.bref_export_tables.argtypes = [ctypes.c_void_p] + [_dp] * 8
        lib.bref_basis_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        lib.bref_evaluate.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int]
        lib.bref_jvp.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.bref_vjp.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.bref_vjp_dense.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.bref_assemble.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.bref_block_lu.argtypes = [ctypes.c_int, _ip, _lp, _dp, _ip, _lp, _ip]
        lib.bref_block_solve.argtypes = [ctypes.c_int, _ip, _lp, _dp, _ip, _lp, _lp, _dp]
        lib.bref_num_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def params_vector(penalty=1.3, nu=1.0, vel=(0.0, 0.0, 0.0), gamma=1.4, mach=None):
    if mach is None:
        mach = 1.0 / np.sqrt(gamma)
    v = list(vel) + [0.0] * (3 - len(vel))
    return np.array([penalty, nu, v[0], v[1], v[2], gamma, mach], dtype=np.float64)


class RefModel:
    """Cartesian DG discretisation evaluated by the CPU restatement."""

    def __init__(self, D, ncells, p, quad_order, lo=None, hi=None, periodic=None, nodes=None):
        """Uniform grid of ncells on the box [lo, hi], or, with `nodes` (D increasing coordinate
        arrays, as Cartesian3DGrid takes them), the tensor grid over those nodes."""
        lib = _load()
        self.D, self.p, self.order = D, p, quad_order
        per = np.zeros(D, np.int32) if periodic is None else np.asarray(periodic, dtype=np.int32)
        if nodes is not None:
            nc = np.array([len(x) - 1 for x in nodes], dtype=np.int32)
            flat = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x in nodes]))
            self.h = lib.bref_create_nodes(TABLES.encode(), D, _p(nc, _ip), _p(flat), _p(per, _ip), p, quad_order)
        else:
            nc = np.array(ncells, dtype=np.int32)
            lo = np.zeros(D) if lo is None else np.asarray(lo, dtype=np.float64)
            hi = np.ones(D) if hi is None else np.asarray(hi, dtype=np.float64)
            self.h = lib.bref_create(TABLES.encode(), D, _p(nc, _ip), _p(lo), _p(hi), _p(per, _ip), p, quad_order)
        if not self.h:
            raise RuntimeError(lib.bref_last_error().decode())
        info = np.zeros(5, np.int32)
        lib.bref_info(self.h, _p(info, _ip))
        self.J, self.E, self.N, self.Kv, self.Ke = (int(x) for x in info)
        self.lib = lib

    def __del__(self):
        if getattr(self, 'h', None):
            self.lib.bref_destroy(self.h)
            self.h = None

    def nvars(self, kind):
        return self.D + 2 if kind in (OP_EULER_LF, OP_EULER_HLLC) else 1

    def grid(self):
        D, J, E = self.D, self.J, self.E
        g = dict(detJ=np.zeros(J), invJ=np.zeros((J, D, D)), scaling=np.zeros(J), lenscale=np.zeros(J),
                 cells=np.zeros((E, 2), np.int32), faces=np.zeros((E, 2), np.int32), normals=np.zeros((E, D)), sqrtgram=np.zeros(E))
        self.lib.bref_export_grid(self.h, _p(g['detJ']), _p(g['invJ']), _p(g['scaling']), _p(g['lenscale']), _p(g['cells'], _ip),
                                  _p(g['faces'], _ip), _p(g['normals']), _p(g['sqrtgram']))
        return g

    def tables(self):
        D, N, Kv, Ke = self.D, self.N, self.Kv, self.Ke
        t = dict(vnodes=np.zeros((Kv, D)), vw=np.zeros(Kv), phi=np.zeros((Kv, N)), dphi=np.zeros((Kv, N, D)),
                 fnodes=np.zeros((Ke, D - 1)), fw=np.zeros(Ke), trace=np.zeros((2 * D, Ke, N)), dtrace=np.zeros((2 * D, Ke, N, D)))
        self.lib.bref_export_tables(self.h, *[_p(t[k]) for k in ('vnodes', 'vw', 'phi', 'dphi', 'fnodes', 'fw', 'trace', 'dtrace')])
        return t

    def basis(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros((x.shape[0], self.N))
        dout = np.zeros((x.shape[0], self.N, self.D))
        self.lib.bref_basis_eval(self.h, x.shape[0], _p(x), _p(out), _p(dout))
        return out, dout

    def evaluate(self, kind, prm, U, alpha=1.0, beta=0.0, R=None, nthreads=0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.zeros_like(U) if R is None else R
        self.lib.bref_evaluate(self.h, kind, _p(prm), alpha, beta, _p(U), _p(R), nthreads)
        return R

    def jvp(self, kind, prm, U, dU, nthreads=0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        dU = np.ascontiguousarray(dU, dtype=np.float64)
        out = np.zeros_like(U)
        self.lib.bref_jvp(self.h, kind, _p(prm), _p(U), _p(dU), _p(out), nthreads)
        return out

    def vjp(self, kind, prm, U, W, nthreads=0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        W = np.ascontiguousarray(W, dtype=np.float64)
        out = np.zeros_like(U)
        self.lib.bref_vjp(self.h, kind, _p(prm), _p(U), _p(W), _p(out), nthreads)
        return out

    def vjp_dense(self, kind, prm, U, W, nthreads=0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        W = np.ascontiguousarray(W, dtype=np.float64)
        out = np.zeros_like(U)
        self.lib.bref_vjp_dense(self.h, kind, _p(prm), _p(U), _p(W), _p(out), nthreads)
        return out

    def assemble(self, kind, prm, U=None, nthreads=0):
        Nt = self.nvars(kind) * self.N
        U = np.zeros(self.J * Nt) if U is None else np.ascontiguousarray(U, dtype=np.float64)
        diag = np.zeros((self.J, Nt, Nt))
        offd = np.zeros((self.E, 2, Nt, Nt))
        self.lib.bref_assemble(self.h, kind, _p(prm), _p(U), _p(diag), _p(offd), nthreads)
        return diag, offd

    def dense_matrix(self, kind, prm, U=None):
        """Assemble the full Jacobian as a dense matrix (small grids only)."""
        diag, offd = self.assemble(kind, prm, U)
        Nt = diag.shape[1]
        A = np.zeros((self.J * Nt, self.J * Nt))
        for j in range(self.J):
            A[j * Nt:(j + 1) * Nt, j * Nt:(j + 1) * Nt] = diag[j]
        cells = self.grid()['cells']
        for e in range(self.E):
            ji, jo = cells[e]
            if jo >= 0:
                A[ji * Nt:(ji + 1) * Nt, jo * Nt:(jo + 1) * Nt] += offd[e, 0]
                A[jo * Nt:(jo + 1) * Nt, ji * Nt:(ji + 1) * Nt] += offd[e, 1]
        return A


def ref_block_lu(sizes, blocks):
    """Batched uneven LU (getf2 semantics).  blocks: list of square arrays; returns (LU list, piv list, info)."""
    lib = _load()
    n = np.array(sizes, np.int32)
    off = np.zeros(len(sizes), np.int64)
    poff = np.zeros(len(sizes), np.int64)
    off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    poff[1:] = np.cumsum(n.astype(np.int64))[:-1]
    A = np.concatenate([np.ascontiguousarray(b, dtype=np.float64).ravel() for b in blocks])
    piv = np.zeros(int(n.sum()), np.int32)
    info = np.zeros(len(sizes), np.int32)
    lib.bref_block_lu(len(sizes), _p(n, _ip), _p(off, _lp), _p(A), _p(piv, _ip), _p(poff, _lp), _p(info, _ip))
    return A, piv, info, off, poff


def ref_block_solve(sizes, LU, piv, off, poff, rhs):
    lib = _load()
    n = np.array(sizes, np.int32)
    x = np.ascontiguousarray(rhs, dtype=np.float64).copy()
    lib.bref_block_solve(len(sizes), _p(n, _ip), _p(off, _lp), _p(LU), _p(piv, _ip), _p(poff, _lp), _p(poff, _lp), _p(x))
    return x


def xdg_assemble(rules, p, mu, src, penalty=2.0, gD=None):
    """Oracle XDG Poisson matrix (block layout of ``rules.block_layout``) and affine vector."""
    lib = _load()
    D = rules.D
    f = lib.bref_xdg_assemble
    f.restype = ctypes.c_int
    f.argtypes = ([ctypes.c_void_p, ctypes.c_int, _ip, _ip, _dp, _dp, _dp, ctypes.c_double, _dp, _dp,
                   ctypes.c_int, _ip, _lp,
                   ctypes.c_int, _ip, _ip, _lp, _dp, _dp,
                   ctypes.c_int, _ip, _ip, _ip, _lp, _dp, _dp, _dp, _dp, _dp,
                   ctypes.c_int, _ip, _lp, _dp, _dp, _dp, _dp, _dp])
    model = RefModel(D, (1,) * D, p, 2 * p)
    N = model.N
    lay = rules.block_layout(N)
    g = rules.grid
    blocks = np.zeros(lay["nentries"])
    affine = np.zeros(lay["ndof"])
    c = np.ascontiguousarray
    i32 = lambda a: c(np.asarray(a, np.int32))
    i64 = lambda a: c(np.asarray(a, np.int64))
    f64 = lambda a: c(np.asarray(a, np.float64))
    gd = f64(np.zeros(len(rules.face_w)) if gD is None else gD)
    keep = [i32(rules.nspec), i32(rules.spec0), f64(g.invJ), f64(g.scaling), f64(rules.lenscale), f64(mu), f64(src),
            i32(lay["cols"]), i64(lay["boff"]),
            i32(rules.vol_cell), i32(rules.vol_spec), i64(rules.vol_off), f64(rules.vol_nodes), f64(rules.vol_w),
            i32(rules.face_cells), i32(lay["fslot"]), i32(rules.face_spec), i64(rules.face_off),
            f64(rules.face_nodes_in), f64(rules.face_nodes_out), f64(rules.face_w), f64(rules.face_normal), gd,
            i32(rules.ifc_cell), i64(rules.ifc_off), f64(rules.ifc_nodes), f64(rules.ifc_w), f64(rules.ifc_normal)]
    k = keep
    rc = f(model.h, g.J, _p(k[0], _ip), _p(k[1], _ip), _p(k[2]), _p(k[3]), _p(k[4]), float(penalty), _p(k[5]), _p(k[6]),
           lay["S"], _p(k[7], _ip), _p(k[8], _lp),
           len(rules.vol_cell), _p(k[9], _ip), _p(k[10], _ip), _p(k[11], _lp), _p(k[12]), _p(k[13]),
           len(rules.face_spec), _p(k[14], _ip), _p(k[15], _ip), _p(k[16], _ip), _p(k[17], _lp),
           _p(k[18]), _p(k[19]), _p(k[20]), _p(k[21]), _p(k[22]),
           len(rules.ifc_cell), _p(k[23], _ip), _p(k[24], _lp), _p(k[25]), _p(k[26]), _p(k[27]),
           _p(blocks), _p(affine))
    if rc != 0:
        raise RuntimeError("bref_xdg_assemble failed")
    return blocks, affine, lay


def xdg_to_csr(rules, lay, blocks):
    """Scatter the block-ELL matrix into a scipy CSR matrix (tests only)."""
    import scipy.sparse as sp
    rows, cols_, vals = [], [], []
    dof, cl, bo, nr = lay["dof"], lay["cols"], lay["boff"], lay["nrow"]
    for j in range(rules.J):
        for s in range(lay["S"]):
            k = cl[j, s]
            if k < 0:
                continue
            a, b = nr[j], nr[k]
            B = blocks[bo[j, s]: bo[j, s] + a * b].reshape(a, b)
            r, cc = np.nonzero(np.ones((a, b), bool))
            rows.append(dof[j] + r); cols_.append(dof[k] + cc); vals.append(B.reshape(-1))
    n = lay["ndof"]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols_))), shape=(n, n))
