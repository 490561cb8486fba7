"""DifferentialOperator / evaluator mirror over the HIP library.

Mirrors the reference interface of the hot path:
  DifferentialOperator(DomainVar, CodomainVar, QuadOrderFunc) + EquationComponents
      (DifferentialOperator.cs:81, 391-452)
  .GetEvaluatorEx(...)  -> IEvaluatorNonLin    (DifferentialOperator.cs:1013-1033)
  .Evaluate(alpha, beta, output)                (IEvaluator.cs:81-97):  output <- beta*output + alpha*R(U)
plus the linearisation the Newton solver needs (JVP replacing Newton.dirder,
Newton.cs:1450-1500) and the adjoint (VJP), which the reference lacks.

Only the forms with hand-written kernels are accepted; anything else raises
NotSupportedException-style errors (the C# side keeps the reference CPU evaluator
as the fallback for those).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, device_array, lib
from .grid import CartesianGrid
from .refelement import Tables

FORMS = {"sip": _lib.BOSSS_FORM_SIP, "advdiff": _lib.BOSSS_FORM_ADVDIFF,
         "euler_lf": _lib.BOSSS_FORM_EULER_LF, "euler_hllc": _lib.BOSSS_FORM_EULER_HLLC}


def default_quad_order(form, p):
    """QuadOrderFunc used by the reference applications (QuadOrderFunc.cs)."""
    if form in ("sip", "advdiff"):
        o = 2 * p                      # SumOfMaxDegrees(RoundUp): p + p, even
        return o + (o % 2)
    return 3 * p                       # NonLinearWithoutParameters(2): 2p + p (CNS Operator.cs:143-173)


class Context:
    _cache = {}

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().bosss_hip_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    @classmethod
    def get(cls, device=0):
        if device not in cls._cache:
            cls._cache[device] = cls(device)
        return cls._cache[device]

    def sync(self):
        check(lib().bosss_hip_sync(self.h))

    @property
    def stream(self):
        return lib().bosss_hip_ctx_stream(self.h)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class DifferentialOperator:
    """A DG operator made of one of the supported equation components on a Cartesian grid."""

    def __init__(self, grid, form, degree, penalty=1.3, nu=1.0, velocity=(0.0, 0.0, 0.0), gamma=1.4, mach=None,
                 quad_order=None, precision="fp64", device=0):
        if form not in FORMS:
            raise NotImplementedError("no GPU kernel for form %r" % (form,))
        self.grid, self.form, self.degree = grid, form, degree
        self.quad_order = default_quad_order(form, degree) if quad_order is None else quad_order
        self.ctx = Context.get(device)
        self.tables = Tables(grid.D, degree, self.quad_order)
        self.precision = precision
        mach = 1.0 / np.sqrt(gamma) if mach is None else mach
        self.params = dict(penalty=penalty, nu=nu, velocity=tuple(velocity), gamma=gamma, mach=mach)
        L = lib()
        # grid upload
        g = grid
        self._keep = [np.ascontiguousarray(x) for x in (g.detJ, g.invJ, g.scaling, g.lenscale, g.cells, g.trafo, g.normals, g.sqrtgram)]
        gd = _lib.GridDesc(g.D, g.J, g.E, _dp(self._keep[0]), _dp(self._keep[1]), _dp(self._keep[2]), _dp(self._keep[3]),
                           _ip(self._keep[4]), _ip(self._keep[5]), _dp(self._keep[6]), _dp(self._keep[7]))
        gh = ctypes.c_void_p()
        check(L.bosss_hip_grid_upload(self.ctx.h, ctypes.byref(gd), ctypes.byref(gh)))
        self._grid_h = gh
        t = self.tables
        tk = [np.ascontiguousarray(x) for x in (t.vw, t.phi, t.dphi, t.fw, t.trace, t.dtrace)]
        td = _lib.TablesDesc(g.D, degree, t.N, t.Kv, t.Ke, t.ntrafo, *[_dp(x) for x in tk])
        v = list(velocity) + [0.0] * (3 - len(velocity))
        fp = _lib.FormParams(penalty, nu, (ctypes.c_double * 3)(*v), gamma, mach)
        oh = ctypes.c_void_p()
        prec = _lib.BOSSS_FP64 if precision == "fp64" else _lib.BOSSS_FP32
        rc = L.bosss_hip_op_create(self.ctx.h, gh, ctypes.byref(td), FORMS[form], ctypes.byref(fp), prec, ctypes.byref(oh))
        if rc != 0:
            L.bosss_hip_grid_destroy(gh)
            check(rc)
        self.h = oh
        ndof = ctypes.c_int64(); nv = ctypes.c_int32(); nm = ctypes.c_int32()
        check(L.bosss_hip_op_info(oh, ctypes.byref(ndof), ctypes.byref(nv), ctypes.byref(nm)))
        self.ndof, self.nvars, self.N = ndof.value, nv.value, nm.value

    def uses_tensor_tables(self):
        """True if the operator runs the tensor-table kernels (the fast path), False if the
        generic full-table kernels (bosss_hip_op_kernel_family)."""
        t = ctypes.c_int32()
        check(lib().bosss_hip_op_kernel_family(self.h, ctypes.byref(t)))
        return bool(t.value)

    def close(self):
        if getattr(self, "h", None):
            lib().bosss_hip_op_destroy(self.h)
            lib().bosss_hip_grid_destroy(self._grid_h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- host API
    def GetEvaluatorEx(self, U):
        return EvaluatorNonLin(self, U)

    def evaluate(self, U, alpha=1.0, beta=0.0, R=None, time=0.0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.zeros(self.ndof) if R is None else R
        assert U.size == self.ndof and R.size == self.ndof and R.dtype == np.float64 and R.flags.c_contiguous
        check(lib().bosss_hip_evaluate(self.h, time, alpha, beta, _dp(U), _dp(R)))
        return R

    def jvp(self, U, dU):
        U = np.ascontiguousarray(U, dtype=np.float64); dU = np.ascontiguousarray(dU, dtype=np.float64)
        out = np.zeros(self.ndof)
        check(lib().bosss_hip_jvp(self.h, _dp(U), _dp(dU), _dp(out)))
        return out

    def vjp(self, U, W):
        U = np.ascontiguousarray(U, dtype=np.float64); W = np.ascontiguousarray(W, dtype=np.float64)
        out = np.zeros(self.ndof)
        check(lib().bosss_hip_vjp(self.h, _dp(U), _dp(W), _dp(out)))
        return out

    # ------------------------------------------------------- device-resident API
    @staticmethod
    def _ptr(t):
        return ctypes.c_void_p(t.data_ptr())

    def _check_dev(self, *ts):
        """Vectors of the device variants: ndof entries of the operator's precision."""
        import torch
        want = torch.float64 if self.precision == "fp64" else torch.float32
        for t in ts:
            device_array(t, want, self.ndof, self.ctx.device, "vector")

    def _check_matrix_args(self, blocks, *vecs):
        """The operator matrix and the vectors it is applied to are FP64 whatever the precision."""
        import torch
        device_array(blocks, torch.float64, self._entries(), self.ctx.device, "blocks", at_least=True)
        for t in vecs:
            device_array(t, torch.float64, self.ndof, self.ctx.device, "vector")

    def _entries(self):
        if getattr(self, "_nentries", None) is None:
            n = ctypes.c_int64()
            check(lib().bosss_hip_op_matrix_layout(self.h, None, ctypes.byref(n), None, None))
            self._nentries = n.value
        return self._nentries

    def _stream(self, stream):
        """HIP stream handle: torch's current stream by default (0 is HIP's null stream)."""
        if stream is None:
            import torch
            return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        return ctypes.c_void_p(stream)

    def evaluate_dev(self, U, R, alpha=1.0, beta=0.0, stream=None):
        self._check_dev(U, R)
        check(lib().bosss_hip_evaluate_dev(self.h, alpha, beta, self._ptr(U), self._ptr(R), self._stream(stream)))

    def jvp_dev(self, U, dU, out, stream=None):
        self._check_dev(U, dU, out)
        check(lib().bosss_hip_jvp_dev(self.h, self._ptr(U), self._ptr(dU), self._ptr(out), self._stream(stream)))

    def vjp_dev(self, U, W, out, stream=None):
        self._check_dev(U, W, out)
        check(lib().bosss_hip_vjp_dev(self.h, self._ptr(U), self._ptr(W), self._ptr(out), self._stream(stream)))

    def evaluate_vjp_dev(self, U, R, W, out, alpha=1.0, beta=0.0, stream=None):
        self._check_dev(U, R, W, out)
        check(lib().bosss_hip_evaluate_vjp_dev(self.h, alpha, beta, self._ptr(U), self._ptr(R), self._ptr(W), self._ptr(out),
                                               self._stream(stream)))


    # ------------------------------------------------ operator matrix (block-ELL)
    def matrix_layout(self):
        """(blocks per row S, total entries, number of probe colours, cols[J][S])."""
        S = ctypes.c_int32(); n = ctypes.c_int64(); nc = ctypes.c_int32()
        check(lib().bosss_hip_op_matrix_layout(self.h, ctypes.byref(S), ctypes.byref(n), ctypes.byref(nc), None))
        cols = np.zeros((self.grid.J, S.value), np.int32)
        check(lib().bosss_hip_op_matrix_layout(self.h, None, None, None, cols.ctypes.data_as(ctypes.c_void_p)))
        return S.value, n.value, nc.value, cols

    def assemble_dev(self, U, blocks, alpha=1.0, beta=0.0, stream=None):
        """blocks <- beta*blocks + alpha*dR/dU(U) (the operator matrix for linear forms)."""
        self._check_matrix_args(blocks, U)
        check(lib().bosss_hip_op_assemble_dev(self.h, self._ptr(U), alpha, beta, self._ptr(blocks), self._stream(stream)))

    def diag_blocks_dev(self, U, blocks, alpha=1.0, shift=0.0, transposed=False, stream=None):
        """blocks <- alpha * dR_j/dU_j(U) + shift * I for every cell j: J row-major Nt x Nt FP64
        blocks, or each block transposed (own-cell JVP probes in the operator's precision; U is a
        state vector of it)."""
        import torch
        self._check_dev(U)
        Nt = self.nvars * self.N
        device_array(blocks, torch.float64, self.grid.J * Nt * Nt, self.ctx.device, "blocks")
        check(lib().bosss_hip_op_diag_blocks_dev(self.h, self._ptr(U), alpha, shift, int(bool(transposed)), self._ptr(blocks),
                                                 self._stream(stream)))

    def spmv_dev(self, blocks, x, y, alpha=1.0, beta=0.0, stream=None):
        self._check_matrix_args(blocks, x, y)
        check(lib().bosss_hip_bell_spmv_dev(self.h, self._ptr(blocks), alpha, self._ptr(x), beta, self._ptr(y), self._stream(stream)))

    def set_check_finite(self, enable=True):
        """CheckForNanOrInf of every output (DifferentialOperator.cs:1602,1623,1658, there in DEBUG
        builds): with it on, each evaluate / jvp / vjp synchronises and raises BosssError with code
        -4 (BOSSS_ERR_NONFINITE) naming the first NaN or Inf entry."""
        check(lib().bosss_hip_op_set_check_finite(self.h, 1 if enable else 0))

    def set_spmv_tile(self, nl):
        """Fix the panel size of the block-ELL apply (0: tune again at the next apply)."""
        check(lib().bosss_hip_op_set_spmv_tile(self.h, int(nl)))

    def spmv_tiles(self):
        """{"chosen": NL or 0, "ms": {candidate NL: timed ms per apply or None}} of the apply's autotuner."""
        chosen, nc = ctypes.c_int32(), ctypes.c_int32()
        check(lib().bosss_hip_op_spmv_tiles(self.h, ctypes.byref(chosen), ctypes.byref(nc), None, None, 0))
        cand = (ctypes.c_int32 * max(nc.value, 1))()
        ms = (ctypes.c_float * max(nc.value, 1))()
        check(lib().bosss_hip_op_spmv_tiles(self.h, None, None, cand, ms, nc.value))
        return {"chosen": chosen.value,
                "ms": {int(cand[i]): (round(float(ms[i]), 5) if ms[i] >= 0 else None) for i in range(nc.value)}}

    def GetMatrixBuilder(self):
        return EvaluatorLinear(self)


class EvaluatorLinear:
    """IEvaluatorLinear: ComputeMatrix into a device block-ELL matrix (EvaluatorLinear, DifferentialOperator.cs:1721-1838)."""

    def __init__(self, op):
        self.op = op
        self.S, self.entries, self.ncolors, self.cols = op.matrix_layout()

    def new_matrix(self):
        import torch
        return torch.zeros(self.entries, dtype=torch.float64, device="cuda:%d" % self.op.ctx.device)

    def ComputeMatrix(self, blocks, alpha=1.0, U=None):
        """Accumulate alpha * (operator matrix) into blocks; U is the linearisation point (zero for linear forms)."""
        import torch
        if U is None:
            U = torch.zeros(self.op.ndof, dtype=torch.float64, device=blocks.device)
        self.op.assemble_dev(U, blocks, alpha=alpha, beta=1.0)
        return blocks

    def ComputeAffine(self):
        """Affine part b of R(U) = M U + b (R at U = 0)."""
        return self.op.evaluate(np.zeros(self.op.ndof))


class EvaluatorNonLin:
    """IEvaluatorNonLin: Evaluate(alpha, beta, output) with the domain state fixed at creation."""

    def __init__(self, op, U):
        self.op = op
        self.U = np.ascontiguousarray(U, dtype=np.float64)
        self.time = 0.0

    def Evaluate(self, alpha, beta, output):
        self.op.evaluate(self.U, alpha, beta, output, self.time)
        return output
