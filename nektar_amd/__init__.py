"""MI355X-native matrix-free spectral/hp element operators (gfx950 HIP).

Python mirror of the operator interface of Nektar++'s matrix-free path:

* ``GetOperatorFactory().CreateInstance(key, nummodes, numpoints, nelmt)``
  mirrors ``MatrixFree::GetOperatorFactory()``
  (library/MatrixFreeOps/Operator.cpp:40-90) with keys built like
  ``GetOpstring`` ("Helmholtz_Hex_Regular", "PhysDeriv_Tet_Deformed", ...);
* :class:`Operator` mirrors ``MatrixFree::Operator`` and its subclasses
  (Operator.hpp:70-366): ``SetJac``, ``SetDF``, ``SetLambda``,
  ``SetConstVarDiffusion`` and ``operator()``.

Everything runs through the C ABI of ``libnektar_amd.so`` (declared in
``include/nektar_mf.h``).  There is no CPU fallback: importing this package
raises if the HIP library has not been built.
"""
from __future__ import annotations

import ctypes
import os

try:  # torch first, so that one HIP runtime serves both torch and this library
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for host-only use
    torch = None

import numpy as np

__all__ = [
    "NektarError", "lib", "Operator", "GetOperatorFactory", "GetOpstring", "clear_tuning", "Batch", "Server",
    "basis_tables", "FP64", "FP32", "HelmholtzSolver", "SolveInfo", "hex_box_vertices",
]

FP64, FP32 = 0, 1
SHAPE_HEX, SHAPE_TET, SHAPE_PRISM, SHAPE_PYR = 0, 1, 2, 3
# hex Helmholtz tiles TILE_MFMA + E: matrix-core kernels, E elements per workgroup
TILE_MFMA = 1024

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnektar_amd.so")


class NektarError(RuntimeError):
    """Raised for a non-zero status from the C ABI (mirrors NekError)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[nmf error {code}] {msg}")
        self.code = code


_LIB = None

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)

class SolveInfo(ctypes.Structure):
    """nmf_solve_info: CG iterations, refinement steps, final relative
    residual sqrt(||r||^2/||b||^2), ||b||^2 and device solve time."""
    _fields_ = [("iterations", ctypes.c_int), ("outer_iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("final_error", ctypes.c_double),
                ("rhs_magnitude", ctypes.c_double), ("solve_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_L = ctypes.POINTER(ctypes.c_int64)

_SIGNATURES = {
    "nmf_last_error": (ctypes.c_char_p, []),
    "nmf_version": (ctypes.c_int, []),
    "nmf_device_malloc": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
    "nmf_device_free": (ctypes.c_int, [_P]),
    "nmf_memcpy_h2d": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "nmf_memcpy_d2h": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "nmf_device_synchronize": (ctypes.c_int, []),
    "nmf_basis_tables": (ctypes.c_int, [ctypes.c_int] * 4 + [_I, _I] + [_D] * 5),
    "nmf_operator_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.POINTER(_P)]),
    "nmf_operator_create_basis": (ctypes.c_int, [ctypes.c_char_p, _I, _I, ctypes.c_int, ctypes.c_int,
                                                 ctypes.POINTER(_P)]),
    "nmf_operator_runtime_size": (ctypes.c_int, [_P, _I]),
    "nmf_operator_destroy": (ctypes.c_int, [_P]),
    "nmf_operator_info": (ctypes.c_int, [_P, _I, _I, _I, _I, _I]),
    "nmf_operator_set_jac": (ctypes.c_int, [_P, _D]),
    "nmf_operator_set_df": (ctypes.c_int, [_P, _D]),
    "nmf_operator_set_lambda": (ctypes.c_int, [_P, ctypes.c_double]),
    "nmf_operator_set_const_var_diffusion": (ctypes.c_int, [_P, _D]),
    "nmf_operator_set_advection": (ctypes.c_int, [_P, _D]),
    "nmf_operator_set_scaling_factor": (ctypes.c_int, [_P, ctypes.c_double]),
    "nmf_operator_io_size": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "nmf_operator_apply": (ctypes.c_int, [_P, _P, _P, _P]),
    "nmf_operator_apply_physderiv": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "nmf_operator_wsp_size": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "nmf_operator_apply_physderiv_dir": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, _P]),
    "nmf_operator_apply_host": (ctypes.c_int, [_P, _D, _D]),
    "nmf_operator_apply_physderiv_host": (ctypes.c_int, [_P, _D, _D, _D, _D]),
    "nmf_operator_apply_iproduct_deriv": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "nmf_operator_apply_iproduct_deriv_host": (ctypes.c_int, [_P, _D, _D, _D, _D]),
    "nmf_operator_tile_candidates": (ctypes.c_int, [_P, _I, ctypes.c_int, _I]),
    "nmf_operator_set_tile": (ctypes.c_int, [_P, ctypes.c_int]),
    "nmf_operator_get_tile": (ctypes.c_int, [_P, _I]),
    "nmf_operator_autotune": (ctypes.c_int, [_P, _P, _I, _D, ctypes.c_int]),
    "nmf_solver_autotune": (ctypes.c_int, [_P, _I, _I]),
    "nmf_server_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)]),
    "nmf_server_destroy": (ctypes.c_int, [_P]),
    "nmf_server_run": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, _P, _P]),
    "nmf_batch_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "nmf_batch_destroy": (ctypes.c_int, [_P]),
    "nmf_batch_add": (ctypes.c_int, [_P, _P, _P, _P]),
    "nmf_batch_add_physderiv": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "nmf_batch_set_mode": (ctypes.c_int, [_P, ctypes.c_int]),
    "nmf_batch_set_streams": (ctypes.c_int, [_P, ctypes.c_int]),
    "nmf_batch_autotune": (ctypes.c_int, [_P, _P]),
    "nmf_batch_run": (ctypes.c_int, [_P, _P]),
    "nmf_batch_info": (ctypes.c_int, [_P, _I, _I, _L]),
    "nmf_clear_tuning": (ctypes.c_int, []),
    "nmf_solver_create_hex_box": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_uint64, ctypes.c_double, _P, ctypes.POINTER(_P)]),
    "nmf_solver_destroy": (ctypes.c_int, [_P]),
    "nmf_solver_sizes": (ctypes.c_int, [_P, _L, _L, _L, _L]),
    "nmf_solver_estimate_bytes": (ctypes.c_int64, [ctypes.c_int] * 5),
    "nmf_solver_set_forcing_manufactured": (ctypes.c_int, [_P, ctypes.c_int]),
    "nmf_solver_set_forcing_phys": (ctypes.c_int, [_P, _P]),
    "nmf_solver_solve": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.POINTER(SolveInfo)]),
    "nmf_solver_begin": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "nmf_solver_advance": (ctypes.c_int, [_P, ctypes.c_int, _I]),
    "nmf_solver_finish": (ctypes.c_int, [_P, ctypes.POINTER(SolveInfo)]),
    "nmf_solver_l2_error_manufactured": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "nmf_solver_get_solution": (ctypes.c_int, [_P, _D]),
    "nmf_solver_get_rhs": (ctypes.c_int, [_P, _D]),
    "nmf_solver_get_inv_diag": (ctypes.c_int, [_P, _D]),
    "nmf_solver_apply_global": (ctypes.c_int, [_P, _P, _P]),
    "nmf_hex_box_vertices": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _D]),
    "nmf_assembly_create": (ctypes.c_int, [_I, _D, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)]),
    "nmf_assembly_destroy": (ctypes.c_int, [_P]),
    "nmf_assembly_global_to_local": (ctypes.c_int, [_P, _P, _P, _P]),
    "nmf_assembly_assemble": (ctypes.c_int, [_P, _P, _P, _P]),
    "nmf_linsys_create": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_int, _P, ctypes.POINTER(_P)]),
    "nmf_linsys_create_solver": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_int, ctypes.c_char_p, _P,
                                                ctypes.POINTER(_P)]),
    "nmf_linsys_get_solution_local": (ctypes.c_int, [_P, _P]),
    "nmf_linsys_sc_diagonal": (ctypes.c_int, [_P, _D]),
    "nmf_linsys_destroy": (ctypes.c_int, [_P]),
    "nmf_linsys_set_rhs_local": (ctypes.c_int, [_P, _P]),
    "nmf_linsys_set_dirichlet_local": (ctypes.c_int, [_P, _P]),
    "nmf_linsys_set_robin_local": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P]),
    "nmf_linsys_solve": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(SolveInfo)]),
    "nmf_linsys_get_solution": (ctypes.c_int, [_P, _D]),
    "nmf_linsys_get_inv_diag": (ctypes.c_int, [_P, _D]),
    "nmf_linsys_apply_global": (ctypes.c_int, [_P, _P, _P]),
}


def lib() -> ctypes.CDLL:
    """Load libnektar_amd.so (built in-tree by ``__graft_entry__.build()``)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    if rc != 0:
        raise NektarError(rc, lib().nmf_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def GetOpstring(shape: str, deformed: bool = False) -> str:
    """``MatrixFree::GetOpstring`` (Operator.cpp:46-90) for Hex/Tet/Prism/Pyr."""
    return "_" + shape + ("_Deformed" if deformed else "_Regular")


def basis_tables(shape: int, direction: int, nummodes: int, numpoints: int) -> dict:
    """Host FP64 1-D tables of the MODIFIED expansion in one direction."""
    nq, nmodes = ctypes.c_int(), ctypes.c_int()
    check(lib().nmf_basis_tables(shape, direction, nummodes, numpoints, ctypes.byref(nq),
                                 ctypes.byref(nmodes), None, None, None, None, None))
    n, m = nq.value, nmodes.value
    out = {k: np.zeros(v) for k, v in
           dict(bdata=m * n, dbdata=m * n, z=n, w=n, D=n * n).items()}
    check(lib().nmf_basis_tables(shape, direction, nummodes, numpoints, None, None,
                                 _dptr(out["bdata"]), _dptr(out["dbdata"]), _dptr(out["z"]),
                                 _dptr(out["w"]), _dptr(out["D"])))
    out["nq"], out["nmodes"] = n, m
    return out


def _device_ptr(x, dtype_expected, count) -> int:
    """Device pointer of a torch tensor (checked) or a raw int address."""
    if isinstance(x, int):
        return x
    if torch is not None and isinstance(x, torch.Tensor):
        if not x.is_cuda:
            raise NektarError(1, "tensor is not on the GPU")
        if x.dtype != dtype_expected:
            raise NektarError(1, f"tensor dtype {x.dtype} != operator precision {dtype_expected}")
        if not x.is_contiguous():
            raise NektarError(1, "tensor must be contiguous")
        if x.numel() < count:
            raise NektarError(1, f"tensor has {x.numel()} entries, operator needs {count}")
        return x.data_ptr()
    raise NektarError(1, f"unsupported array type {type(x)}")


class Operator:
    """``MatrixFree::Operator`` instance for one element collection."""

    def __init__(self, key: str, nummodes, numpoints, nelmt: int,
                 precision: int = FP64):
        """nummodes / numpoints: one int (equal modes, the shape's default
        points from the direction-0 count) or three per-direction values
        (the reference's one BasisKey per direction, Operator.hpp:59-63)."""
        self._h = None
        h = ctypes.c_void_p()
        if isinstance(nummodes, int) and isinstance(numpoints, int):
            check(lib().nmf_operator_create(key.encode(), nummodes, numpoints, nelmt, precision,
                                            ctypes.byref(h)))
        else:
            m, q = list(nummodes), list(numpoints)
            if len(m) != 3 or len(q) != 3:
                raise NektarError(1, "per-direction modes and points need three values each")
            check(lib().nmf_operator_create_basis(key.encode(), (ctypes.c_int * 3)(*m), (ctypes.c_int * 3)(*q),
                                                  nelmt, precision, ctypes.byref(h)))
        self._h = h
        g = ctypes.c_int()
        check(lib().nmf_operator_runtime_size(h, ctypes.byref(g)))
        self.runtime_size = bool(g.value)
        self.key = key
        self.op = key.split("_")[0]
        nc, nq, ne, de, pr = (ctypes.c_int() for _ in range(5))
        check(lib().nmf_operator_info(h, ctypes.byref(nc), ctypes.byref(nq), ctypes.byref(ne),
                                      ctypes.byref(de), ctypes.byref(pr)))
        self.ncoeff, self.nqtot, self.nelmt = nc.value, nq.value, ne.value
        self.deformed, self.precision = bool(de.value), pr.value

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nmf_operator_destroy(self._h)
            self._h = None

    # --- MatrixFree::Operator set-up -------------------------------------
    def _geom_points(self):
        return self.nelmt * (self.nqtot if self.deformed else 1)

    def SetJac(self, jac) -> None:
        jac = np.ascontiguousarray(jac, dtype=np.float64)
        if jac.size != self._geom_points():
            raise NektarError(1, f"SetJac: expected {self._geom_points()} values, got {jac.size}")
        check(lib().nmf_operator_set_jac(self._h, _dptr(jac)))

    def SetDF(self, df) -> None:
        df = np.ascontiguousarray(df, dtype=np.float64)
        if df.size != 9 * self._geom_points():
            raise NektarError(1, f"SetDF: expected {9 * self._geom_points()} values, got {df.size}")
        check(lib().nmf_operator_set_df(self._h, _dptr(df)))

    def SetLambda(self, lam: float) -> None:
        check(lib().nmf_operator_set_lambda(self._h, float(lam)))

    def SetConstVarDiffusion(self, d6) -> None:
        if d6 is None:
            check(lib().nmf_operator_set_const_var_diffusion(self._h, None))
            return
        d = np.ascontiguousarray(d6, dtype=np.float64)
        check(lib().nmf_operator_set_const_var_diffusion(self._h, _dptr(d)))

    def SetAdvectionVelocities(self, vel) -> None:
        """LinearAdvectionDiffusionReaction: vel[(e*nqtot + q)*3 + j], the
        velocity component j at point q of element e."""
        v = np.ascontiguousarray(vel, dtype=np.float64)
        if v.size != 3 * self.nelmt * self.nqtot:
            raise NektarError(1, f"SetAdvectionVelocities: expected {3 * self.nelmt * self.nqtot} values, "
                                 f"got {v.size}")
        check(lib().nmf_operator_set_advection(self._h, _dptr(v)))

    # --- sizes (Collections::Operator::GetInputSize/GetOutputSize) --------
    def _io_size(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(lib().nmf_operator_io_size(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def input_size(self) -> int:
        return self._io_size()[0]

    def output_size(self) -> int:
        return self._io_size()[1]

    def SetScalingFactor(self, factor: float) -> None:
        """PhysInterp1DScaled: points per direction scaled by factor."""
        check(lib().nmf_operator_set_scaling_factor(self._h, float(factor)))

    def torch_dtype(self):
        return torch.float32 if self.precision == FP32 else torch.float64

    # --- operator() -------------------------------------------------------
    def __call__(self, inp, out, out1=None, out2=None, stream=None):
        """Apply on device arrays (torch tensors or raw device addresses).
        IProductWRTDerivBase takes a sequence of three inputs."""
        dt = self.torch_dtype() if torch is not None else None
        s = stream
        if s is None and torch is not None:
            s = torch.cuda.current_stream().cuda_stream
        if self.op == "IProductWRTDerivBase":
            if len(inp) != 3:
                raise NektarError(1, "IProductWRTDerivBase takes three input arrays")
            ptrs = [_device_ptr(a, dt, self.input_size()) for a in inp]
            check(lib().nmf_operator_apply_iproduct_deriv(self._h, *ptrs, _device_ptr(out, dt, self.output_size()),
                                                          s))
            return
        pin = _device_ptr(inp, dt, self.input_size())
        if self.op == "PhysDeriv":
            ptrs = [_device_ptr(o, dt, self.output_size()) for o in (out, out1, out2)]
            check(lib().nmf_operator_apply_physderiv(self._h, pin, *ptrs, s))
        else:
            check(lib().nmf_operator_apply(self._h, pin, _device_ptr(out, dt, self.output_size()), s))

    def GetWspSize(self) -> int:
        """Collections::Operator::GetWspSize: workspace values of the
        single-direction PhysDeriv form (0 for the other operators)."""
        w = ctypes.c_int64()
        check(lib().nmf_operator_wsp_size(self._h, ctypes.byref(w)))
        return w.value

    def apply_dir(self, direction: int, inp, out, wsp=None, stream=None):
        """PhysDeriv ``operator()(dir, input, output, wsp)``
        (Collections/PhysDeriv.cpp:312-317): the derivative in one
        direction; wsp (GetWspSize() values) or None for a per-call one."""
        if self.op != "PhysDeriv":
            raise NektarError(1, "apply_dir: not a PhysDeriv operator")
        dt = self.torch_dtype() if torch is not None else None
        s = stream
        if s is None and torch is not None:
            s = torch.cuda.current_stream().cuda_stream
        w = None if wsp is None else _device_ptr(wsp, dt, self.GetWspSize())
        check(lib().nmf_operator_apply_physderiv_dir(self._h, int(direction), _device_ptr(inp, dt, self.input_size()),
                                                     _device_ptr(out, dt, self.output_size()), w, s))

    # --- workgroup tiles ---------------------------------------------------
    def tile_candidates(self) -> list:
        buf = (ctypes.c_int * 16)()
        n = ctypes.c_int()
        check(lib().nmf_operator_tile_candidates(self._h, buf, 16, ctypes.byref(n)))
        return list(buf[:min(n.value, 16)])

    @property
    def tile(self) -> int:
        t = ctypes.c_int()
        check(lib().nmf_operator_get_tile(self._h, ctypes.byref(t)))
        return t.value

    @tile.setter
    def tile(self, t: int) -> None:
        check(lib().nmf_operator_set_tile(self._h, int(t)))

    def autotune(self, stream=None):
        """Time every tile on this operator's geometry and keep the fastest;
        returns (tile, {tile: mean microseconds})."""
        cand = self.tile_candidates()
        times = (ctypes.c_double * max(1, len(cand)))()
        t = ctypes.c_int()
        s = stream
        if s is None and torch is not None:
            s = torch.cuda.current_stream().cuda_stream
        check(lib().nmf_operator_autotune(self._h, s, ctypes.byref(t), times, len(cand)))
        return t.value, {c: times[i] for i, c in enumerate(cand)}

    def apply_host(self, inp):
        """Host FP64 in / out (synchronous); returns ndarray(s).
        IProductWRTDerivBase takes a sequence of three arrays."""
        if self.op == "IProductWRTDerivBase":
            xs = [np.ascontiguousarray(a, dtype=np.float64) for a in inp]
            if len(xs) != 3 or any(a.size != self.input_size() for a in xs):
                raise NektarError(1, f"expected three inputs of {self.input_size()} values")
            y = np.zeros(self.output_size())
            check(lib().nmf_operator_apply_iproduct_deriv_host(self._h, *map(_dptr, xs), _dptr(y)))
            return y
        x = np.ascontiguousarray(inp, dtype=np.float64)
        if x.size != self.input_size():
            raise NektarError(1, f"input has {x.size} values, expected {self.input_size()}")
        if self.op == "PhysDeriv":
            outs = [np.zeros(self.output_size()) for _ in range(3)]
            check(lib().nmf_operator_apply_physderiv_host(self._h, _dptr(x), *map(_dptr, outs)))
            return outs
        y = np.zeros(self.output_size())
        check(lib().nmf_operator_apply_host(self._h, _dptr(x), _dptr(y)))
        return y


class _Factory:
    """``MatrixFree::GetOperatorFactory()`` (Operator.cpp:40-44)."""

    def CreateInstance(self, key: str, nummodes, numpoints, nelmt: int,
                       precision: int = FP64) -> Operator:
        return Operator(key, nummodes, numpoints, nelmt, precision)


_FACTORY = _Factory()


def GetOperatorFactory() -> _Factory:
    return _FACTORY


class Batch:
    """Many operators of uneven kinds and sizes applied as one submission
    (grouped launches per kind, spread over streams, captured once into a
    hipGraph).  Keeps references to its operators and arrays."""

    GRAPH, GROUPED, PER_ITEM = 0, 1, 2

    def __init__(self, mode: int = 0, streams: int = 4):
        self._h = None
        h = ctypes.c_void_p()
        check(lib().nmf_batch_create(ctypes.byref(h)))
        self._h = h
        self._keep = []
        self.set_mode(mode)
        self.set_streams(streams)

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nmf_batch_destroy(self._h)
            self._h = None

    def add(self, op: "Operator", inp, out) -> None:
        dt = op.torch_dtype() if torch is not None else None
        check(lib().nmf_batch_add(self._h, op._h, _device_ptr(inp, dt, op.input_size()),
                                  _device_ptr(out, dt, op.output_size())))
        self._keep.append((op, inp, out))

    def add_physderiv(self, op: "Operator", inp, out0, out1, out2) -> None:
        dt = op.torch_dtype() if torch is not None else None
        ptrs = [_device_ptr(o, dt, op.output_size()) for o in (out0, out1, out2)]
        check(lib().nmf_batch_add_physderiv(self._h, op._h, _device_ptr(inp, dt, op.input_size()), *ptrs))
        self._keep.append((op, inp, out0, out1, out2))

    def set_mode(self, mode: int) -> None:
        check(lib().nmf_batch_set_mode(self._h, int(mode)))

    def set_streams(self, n: int) -> None:
        check(lib().nmf_batch_set_streams(self._h, int(n)))

    def _stream(self, stream):
        if stream is None and torch is not None:
            return torch.cuda.current_stream().cuda_stream
        return stream

    def autotune(self, stream=None) -> None:
        check(lib().nmf_batch_autotune(self._h, self._stream(stream)))

    def run(self, stream=None) -> None:
        check(lib().nmf_batch_run(self._h, self._stream(stream)))

    def info(self) -> dict:
        n, g, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        check(lib().nmf_batch_info(self._h, ctypes.byref(n), ctypes.byref(g), ctypes.byref(d)))
        return {"items": n.value, "groups": g.value, "dofs": d.value}


class Request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nummodes", ctypes.c_int), ("numpoints", ctypes.c_int),
                ("k", ctypes.c_int), ("mixed", ctypes.c_int), ("maxit", ctypes.c_int),
                ("lam", ctypes.c_double), ("tol", ctypes.c_double), ("arrival_s", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [("unknowns", ctypes.c_int64), ("iterations", ctypes.c_int), ("outer_iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("reused_solver", ctypes.c_int), ("stream", ctypes.c_int),
                ("final_error", ctypes.c_double), ("l2_error", ctypes.c_double), ("solution_norm", ctypes.c_double),
                ("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("finish_s", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double), ("device_ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("peak_bytes", ctypes.c_int64), ("peak_active", ctypes.c_int), ("solvers_created", ctypes.c_int),
                ("solvers_reused", ctypes.c_int), ("evictions", ctypes.c_int), ("makespan_s", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Server:
    """Serving loop for Helmholtz solves (admission by HBM budget, pooled
    resident solvers, solves advanced chunk by chunk on several streams)."""

    def __init__(self, hbm_budget_bytes: int, streams: int = 4, max_active: int = 16):
        self._h = None
        h = ctypes.c_void_p()
        check(lib().nmf_server_create(int(hbm_budget_bytes), int(streams), int(max_active), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nmf_server_destroy(self._h)
            self._h = None

    def run(self, requests, validate: bool = False):
        """requests: dicts with n, nummodes, numpoints, k, mixed, maxit, lam,
        tol, arrival_s.  Returns (list of result dicts, stats dict)."""
        count = len(requests)
        reqs = (Request * max(1, count))()
        for i, r in enumerate(requests):
            reqs[i] = Request(r["n"], r["nummodes"], r["numpoints"], r.get("k", 1), int(r.get("mixed", 0)),
                              r.get("maxit", 5000), r.get("lam", 1.0), r.get("tol", 1e-10), r.get("arrival_s", 0.0))
        res = (Result * max(1, count))()
        st = ServerStats()
        check(lib().nmf_server_run(self._h, reqs, count, int(validate), res, ctypes.byref(st)))
        return [res[i].as_dict() for i in range(count)], st.as_dict()


def clear_tuning() -> None:
    """Forget every tile the autotuner recorded in this process."""
    check(lib().nmf_clear_tuning())


def hex_box_vertices(n: int, perturb: float = 0.0, seed: int = 0) -> np.ndarray:
    """Vertices of the structured n^3 mesh used by :class:`HelmholtzSolver`."""
    out = np.zeros((n + 1) ** 3 * 3)
    check(lib().nmf_hex_box_vertices(n, perturb, seed, _dptr(out)))
    return out


class HelmholtzSolver:
    """C0 solve of -lap(u) + lambda u = f, u = 0 on the boundary, on a
    structured n^3 hex mesh of [0,1]^3 (the ContField::HelmSolve path with
    the iterative-full CG and Diagonal preconditioner)."""

    def __init__(self, n: int, nummodes: int, numpoints: int, lam: float = 1.0,
                 perturb: float = 0.0, seed: int = 0, stream=None):
        self._h = None
        h = ctypes.c_void_p()
        check(lib().nmf_solver_create_hex_box(n, nummodes, numpoints, perturb, seed, lam,
                                              stream, ctypes.byref(h)))
        self._h = h
        self.n, self.nummodes, self.numpoints, self.lam = n, nummodes, numpoints, lam
        ng, nl, nd, nb = (ctypes.c_int64() for _ in range(4))
        check(lib().nmf_solver_sizes(h, ctypes.byref(ng), ctypes.byref(nl), ctypes.byref(nd),
                                     ctypes.byref(nb)))
        self.nglobal, self.nlocal, self.ndirichlet = ng.value, nl.value, nd.value
        self.device_bytes = nb.value

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nmf_solver_destroy(self._h)
            self._h = None

    @staticmethod
    def estimate_bytes(n, nummodes, numpoints, deformed=False, mixed=False) -> int:
        return lib().nmf_solver_estimate_bytes(n, nummodes, numpoints, int(deformed), int(mixed))

    def set_forcing_manufactured(self, k: int = 1) -> None:
        check(lib().nmf_solver_set_forcing_manufactured(self._h, k))

    def set_forcing_phys(self, f) -> None:
        count = self.n ** 3 * self.numpoints ** 3
        ptr = _device_ptr(f, torch.float64, count)
        if torch is not None:
            torch.cuda.current_stream().synchronize()  # the solver runs on its own stream
        check(lib().nmf_solver_set_forcing_phys(self._h, ptr))

    def solve(self, mixed: bool = False, tol: float = 1e-10, maxit: int = 5000) -> SolveInfo:
        info = SolveInfo()
        check(lib().nmf_solver_solve(self._h, int(mixed), tol, maxit, ctypes.byref(info)))
        return info

    # --- non-blocking solve ------------------------------------------------
    def begin(self, mixed: bool = False, tol: float = 1e-10, maxit: int = 5000) -> None:
        check(lib().nmf_solver_begin(self._h, int(mixed), tol, maxit))

    def advance(self, wait: bool = False) -> bool:
        d = ctypes.c_int()
        check(lib().nmf_solver_advance(self._h, int(wait), ctypes.byref(d)))
        return bool(d.value)

    def finish(self) -> SolveInfo:
        info = SolveInfo()
        check(lib().nmf_solver_finish(self._h, ctypes.byref(info)))
        return info

    def autotune(self):
        """Choose the gather-apply tile for FP64 and FP32 iterations;
        returns (tile64, tile32)."""
        t64, t32 = ctypes.c_int(), ctypes.c_int()
        check(lib().nmf_solver_autotune(self._h, ctypes.byref(t64), ctypes.byref(t32)))
        return t64.value, t32.value

    def l2_error_manufactured(self, k: int = 1) -> float:
        v = ctypes.c_double()
        check(lib().nmf_solver_l2_error_manufactured(self._h, k, ctypes.byref(v)))
        return v.value

    def _global(self, fn):
        out = np.zeros(self.nglobal)
        check(fn(self._h, _dptr(out)))
        return out

    def solution(self) -> np.ndarray:
        return self._global(lib().nmf_solver_get_solution)

    def rhs(self) -> np.ndarray:
        return self._global(lib().nmf_solver_get_rhs)

    def inv_diag(self) -> np.ndarray:
        return self._global(lib().nmf_solver_get_inv_diag)

    def apply_global(self, x, y) -> None:
        px = _device_ptr(x, torch.float64, self.nglobal)
        py = _device_ptr(y, torch.float64, self.nglobal)
        if torch is not None:
            torch.cuda.current_stream().synchronize()  # the solver runs on its own stream
        check(lib().nmf_solver_apply_global(self._h, px, py))


class AssemblyMap:
    """``AssemblyMapCG`` on the device for any conforming mesh: map[i] and
    sign[i] (+1/-1, or None for all +1) of every entry of the concatenated
    element-local coefficient vector; the ndir Dirichlet DOFs first."""

    def __init__(self, map_, sign, nglobal: int, ndir: int):
        m = np.ascontiguousarray(map_, dtype=np.int32)
        sg = None if sign is None else np.ascontiguousarray(sign, dtype=np.float64)
        if sg is not None and sg.size != m.size:
            raise NektarError(1, "map and sign differ in length")
        self.nloc, self.nglobal, self.ndir = int(m.size), int(nglobal), int(ndir)
        self._h = ctypes.c_void_p()
        check(lib().nmf_assembly_create(m.ctypes.data_as(_I), None if sg is None else _dptr(sg), self.nloc,
                                        self.nglobal, self.ndir, ctypes.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nmf_assembly_destroy(self._h)
            self._h = None

    def global_to_local(self, glob, loc, stream=None) -> None:
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().nmf_assembly_global_to_local(self._h, _device_ptr(glob, torch.float64, self.nglobal),
                                                 _device_ptr(loc, torch.float64, self.nloc), s))

    def assemble(self, loc, glob, stream=None) -> None:
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().nmf_assembly_assemble(self._h, _device_ptr(loc, torch.float64, self.nloc),
                                          _device_ptr(glob, torch.float64, self.nglobal), s))


class LinearSystem:
    """``GlobalLinSysIterativeFull`` with the Diagonal preconditioner over
    FP64 Helmholtz operators sharing an AssemblyMap; the operators' local
    vectors, concatenated in the order given, are the map's local vector.
    ``solver`` is the reference's LinSysIterSolver
    (GlobalLinSysIterative.cpp:86-88): "ConjugateGradient"
    (NekLinSysIterCG, assembled vectors) or "ConjugateGradientLoc"
    (NekLinSysIterCGLoc, element-local vectors); or its GlobalSysSoln
    "IterativeStaticCond" (single-level static condensation, CG on the
    Schur complement)."""

    def __init__(self, amap: AssemblyMap, ops, stream=None, solver: str = "ConjugateGradient"):
        self._amap, self._ops = amap, list(ops)  # keep both alive
        arr = (ctypes.c_void_p * len(self._ops))(*[o._h.value for o in self._ops])
        self._h = ctypes.c_void_p()
        check(lib().nmf_linsys_create_solver(amap._h, arr, len(self._ops), solver.encode(), stream,
                                             ctypes.byref(self._h)))
        self.nglobal = amap.nglobal
        self.solver = solver

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nmf_linsys_destroy(self._h)
            self._h = None

    def set_rhs_local(self, bloc) -> None:
        check(lib().nmf_linsys_set_rhs_local(self._h, _device_ptr(bloc, torch.float64, self._amap.nloc)))

    def set_dirichlet_local(self, dloc) -> None:
        """Dirichlet values in the local layout (0 off the Dirichlet DOFs);
        None returns to homogeneous data (nmf_linsys_set_dirichlet_local)."""
        p = None if dloc is None else _device_ptr(dloc, torch.float64, self._amap.nloc)
        check(lib().nmf_linsys_set_dirichlet_local(self._h, p))

    def set_robin_local(self, rows, rowptr, cols, vals) -> None:
        """Robin boundary mass rows of the local vector: y[rows[k]] +=
        sum_{m in [rowptr[k], rowptr[k+1])} vals[m] x[cols[m]]
        (nmf_linsys_set_robin_local); empty rows remove it."""
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        if rows.size and (rowptr.size != rows.size + 1 or cols.size != vals.size or cols.size != rowptr[-1]):
            raise ValueError("Robin arrays of inconsistent sizes")
        self._robin = (rows, rowptr, cols, vals)  # kept alive for the call
        check(lib().nmf_linsys_set_robin_local(self._h, rows.size, rows.ctypes.data, rowptr.ctypes.data,
                                               cols.ctypes.data, vals.ctypes.data))

    def solve(self, tol: float = 1e-10, maxit: int = 5000) -> SolveInfo:
        info = SolveInfo()
        check(lib().nmf_linsys_solve(self._h, tol, maxit, ctypes.byref(info)))
        return info

    def solution(self) -> np.ndarray:
        x = np.zeros(self.nglobal)
        check(lib().nmf_linsys_get_solution(self._h, _dptr(x)))
        return x

    def solution_local(self, xloc) -> None:
        """The solution in the local layout into the device tensor xloc
        (enqueued on the solver's stream)."""
        check(lib().nmf_linsys_get_solution_local(self._h, _device_ptr(xloc, torch.float64, self._amap.nloc)))

    def inv_diag(self) -> np.ndarray:
        d = np.zeros(self.nglobal)
        check(lib().nmf_linsys_get_inv_diag(self._h, _dptr(d)))
        return d

    def sc_diagonal(self) -> np.ndarray:
        """Static-condensation diagonal (Schur complement of the interior
        modes, assembled on the boundary DOFs; Dirichlet and interior rows 0)."""
        d = np.zeros(self.nglobal)
        check(lib().nmf_linsys_sc_diagonal(self._h, _dptr(d)))
        return d

    def apply_global(self, x, y) -> None:
        check(lib().nmf_linsys_apply_global(self._h, _device_ptr(x, torch.float64, self.nglobal),
                                            _device_ptr(y, torch.float64, self.nglobal)))
