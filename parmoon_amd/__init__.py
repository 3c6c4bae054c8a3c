"""parmoon_amd -- MI355X implementation of ParMooN's CD3D/TCD3D hot path.

Python view of the C-ABI in include/pmx.h (libpmx.so, built in-tree by
``python -m parmoon_amd.build``).  The library is the product: this module
only marshals arrays.  If libpmx.so is missing, importing the binding raises
immediately -- there is no CPU fallback.
"""
import ctypes
import os
import weakref

import numpy as np

from . import prm as prm  # noqa: F401  (parameter block helpers)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmx.so")

PRM_LEN = 512
# the callback types of pmx_system_assemble_cb (CoeffFct3D, BoundCondFunct3D, BoundValueFunct3D)
COEFF_FCT = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
                             ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
BOUND_COND_FCT = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_int))
BOUND_VALUE_FCT = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_double))
LEARNED_IN = 16  # input channels of the learned smoother's network
_lib = None


class PmxError(RuntimeError):
    pass


def lib():
    """Load libpmx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PmxError(f"{LIB_PATH} not built: run `python -m parmoon_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.pmx_last_error.restype = ctypes.c_char_p
        L.pmx_ctx_create.argtypes = [ip, ctypes.POINTER(vp)]
        L.pmx_ctx_destroy.argtypes = [vp]
        L.pmx_ctx_sync.argtypes = [vp]
        L.pmx_host_space_dims.argtypes = [ip, ip, vp]
        L.pmx_host_space_arrays.argtypes = [ip, ip, vp, vp, vp, vp]
        L.pmx_host_write_vtk.argtypes = [ip, ip, ip, vp, vp, ctypes.c_char_p]
        L.pmx_system_create.argtypes = [vp, ip, ip, ip, ctypes.POINTER(vp)]
        L.pmx_system_destroy.argtypes = [vp]
        L.pmx_system_dims.argtypes = [vp, ip, vp]
        L.pmx_space_numbering.argtypes = [vp, ip, vp, vp]
        L.pmx_space_structure.argtypes = [vp, ip, vp, vp]
        L.pmx_space_colours.argtypes = [vp, ip, vp]
        L.pmx_system_assemble.argtypes = [vp, vp]
        L.pmx_matrix_entries.argtypes = [vp, ip, ip, vp]
        L.pmx_system_rhs.argtypes = [vp, ip, vp]
        L.pmx_matvec.argtypes = [vp, ip, ip, vp, vp]
        L.pmx_defect.argtypes = [vp, ip, vp, vp, vp, vp]
        L.pmx_defect_noalloc.argtypes = [vp, ip, vp, vp, vp, vp]
        L.pmx_smooth.argtypes = [vp, ip, ip, dp, ip, vp, vp]
        L.pmx_prolongate.argtypes = [vp, ip, vp, vp]
        L.pmx_restrict.argtypes = [vp, ip, vp, vp]
        L.pmx_csr_matvec.argtypes = [vp, ip, vp, vp, vp, vp, vp]
        L.pmx_system_solve.argtypes = [vp, vp, vp, vp, vp]
        L.pmx_iterate.argtypes = [vp, ip, vp, vp, vp, vp, vp, vp, vp]
        L.pmx_level_set_entries.argtypes = [vp, ip, ip, vp, vp, vp]
        L.pmx_system_assemble_cb.argtypes = [vp, vp, vp, vp, vp]
        L.pmx_system_solve_resident.argtypes = [vp, vp, vp]
        L.pmx_system_set_step_length_control.argtypes = [vp, ctypes.c_int]
        L.pmx_system_solution.argtypes = [vp, vp]
        L.pmx_system_errors.argtypes = [vp, vp, vp]
        L.pmx_time_kernel.argtypes = [vp, ip, ip, ip, ip, vp]
        L.pmx_tcd_init.argtypes = [vp, vp, vp]
        L.pmx_tcd_step.argtypes = [vp, vp, vp, vp, vp]
        L.pmx_tcd_assemble_arhs.argtypes = [vp, vp, vp]
        L.pmx_tcd_assemble_systmat.argtypes = [vp]
        L.pmx_tcd_solve.argtypes = [vp, vp, vp]
        L.pmx_tcd_restore_massmat.argtypes = [vp]
        L.pmx_learned_nparams.argtypes = [vp]
        L.pmx_learned_init.argtypes = [vp, ctypes.c_uint]
        L.pmx_learned_params.argtypes = [vp, ip, vp]
        L.pmx_learned_params_level.argtypes = [vp, ip, ip, vp]
        L.pmx_learned_forward.argtypes = [vp, ip, vp, vp]
        L.pmx_learned_gradient.argtypes = [vp, ip, vp, vp, vp, vp]
        L.pmx_learned_adam.argtypes = [vp, dp]
        L.pmx_learned_gradient_level.argtypes = [vp, ip, vp, vp, vp]
        L.pmx_learned_train.argtypes = [vp, vp, vp, vp]
        L.pmx_learned_post.argtypes = [vp, ip, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _ok(rc, what):
    if rc != 0:
        raise PmxError(f"{what} failed ({rc}): {lib().pmx_last_error().decode()}")


def host_space(nref, order):
    """(dims, BeginIndex, GlobalNumbers, RowPtr, KCol) computed on the host only."""
    L = lib()
    d = np.zeros(8, np.int64)
    _ok(L.pmx_host_space_dims(nref, order, _p(d)), "pmx_host_space_dims")
    N, _, nc, nnz, nb = (int(v) for v in d[:5])
    bi = np.zeros(nc + 1, np.int32)
    gn = np.zeros(nc * nb, np.int32)
    rp = np.zeros(N + 1, np.int32)
    kc = np.zeros(nnz, np.int32)
    _ok(L.pmx_host_space_arrays(nref, order, _p(bi), _p(gn), _p(rp), _p(kc)), "pmx_host_space_arrays")
    return d, bi, gn, rp, kc


def write_vtk(nref, order, functions, path):
    """Legacy-VTK file of scalar functions of the host space (nref, order):
    `functions` maps SCALARS names to N values in the reference numbering
    (pmx_host_write_vtk)."""
    names = list(functions)
    vals = [np.ascontiguousarray(functions[k], np.float64) for k in names]
    N = int(host_space_dims(nref, order)[0])
    for k, v in zip(names, vals):
        if v.shape != (N,):
            raise ValueError(f"{k}: {v.shape} values for {N} DOFs")
    cn = (ctypes.c_char_p * len(names))(*[k.encode() for k in names])
    cv = (ctypes.c_void_p * len(vals))(*[v.ctypes.data for v in vals])
    _ok(lib().pmx_host_write_vtk(nref, order, len(names), cn, cv, str(path).encode()), "pmx_host_write_vtk")


def host_space_dims(nref, order):
    d = np.zeros(8, np.int64)
    _ok(lib().pmx_host_space_dims(nref, order, _p(d)), "pmx_host_space_dims")
    return d


class Context:
    """One HIP device and stream (pmx_ctx)."""

    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        self._systems = weakref.WeakSet()
        _ok(lib().pmx_ctx_create(device, ctypes.byref(self.h)), "pmx_ctx_create")

    def sync(self):
        _ok(lib().pmx_ctx_sync(self.h), "pmx_ctx_sync")

    def close(self):
        """Destroys the context after the systems created on it."""
        if self.h:
            for s in list(self._systems):
                s.close()
            _ok(lib().pmx_ctx_destroy(self.h), "pmx_ctx_destroy")
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def csr_matvec(self, rowptr, kcol, entries, x):
        n = len(rowptr) - 1
        y = np.zeros(n)
        _ok(lib().pmx_csr_matvec(self.h, n, _p(np.ascontiguousarray(rowptr, np.int32)),
                                 _p(np.ascontiguousarray(kcol, np.int32)),
                                 _p(np.ascontiguousarray(entries, float)),
                                 _p(np.ascontiguousarray(x, float)), _p(y)), "pmx_csr_matvec")
        return y


class System:
    """Multilevel CD3D system (pmx_system): unit cube levels nref0..nref0+nlev-1."""

    def __init__(self, ctx, nref0, nlev, order):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        _ok(lib().pmx_system_create(ctx.h, nref0, nlev, order, ctypes.byref(self.h)), "pmx_system_create")
        ctx._systems.add(self)
        self.nlev, self.order, self.nref0 = nlev, order, nref0

    def close(self):
        if self.h:
            lib().pmx_system_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dims(self, level=-1):
        d = np.zeros(8, np.int64)
        _ok(lib().pmx_system_dims(self.h, level, _p(d)), "pmx_system_dims")
        return d

    def numbering(self, level=-1):
        d = self.dims(level)
        bi = np.zeros(int(d[2]) + 1, np.int32)
        gn = np.zeros(int(d[2]) * int(d[4]), np.int32)
        _ok(lib().pmx_space_numbering(self.h, level, _p(bi), _p(gn)), "pmx_space_numbering")
        return bi, gn

    def structure(self, level=-1):
        d = self.dims(level)
        rp = np.zeros(int(d[0]) + 1, np.int32)
        kc = np.zeros(int(d[3]), np.int32)
        _ok(lib().pmx_space_structure(self.h, level, _p(rp), _p(kc)), "pmx_space_structure")
        return rp, kc

    def colours(self, level=-1):
        c = np.zeros(int(self.dims(level)[0]), np.int32)
        _ok(lib().pmx_space_colours(self.h, level, _p(c)), "pmx_space_colours")
        return c

    def assemble(self, prm_block):
        p = np.ascontiguousarray(prm_block, float)
        assert p.size == PRM_LEN
        _ok(lib().pmx_system_assemble(self.h, _p(p)), "pmx_system_assemble")

    def set_entries(self, entries, level=-1, rowptr=None, kcol=None):
        """The caller's matrix (CSR of the level's pattern, reference
        numbering; by default the level's own rowptr / kcol) as the level's
        operator (pmx_level_set_entries)."""
        if rowptr is None:
            rowptr, kcol = self.structure(level)
        rp = np.ascontiguousarray(rowptr, np.int32)
        kc = np.ascontiguousarray(kcol, np.int32)
        e = np.ascontiguousarray(entries, np.float64)
        if e.size != kc.size or rp[-1] != kc.size:
            raise ValueError("entries, rowptr and kcol do not fit together")
        _ok(lib().pmx_level_set_entries(self.h, level, int(rp.size - 1), _p(rp), _p(kc), _p(e)),
            "pmx_level_set_entries")

    def entries(self, level=-1, which=0):
        e = np.zeros(int(self.dims(level)[3]))
        _ok(lib().pmx_matrix_entries(self.h, level, which, _p(e)), "pmx_matrix_entries")
        return e

    def rhs(self, level=-1):
        r = np.zeros(int(self.dims(level)[0]))
        _ok(lib().pmx_system_rhs(self.h, level, _p(r)), "pmx_system_rhs")
        return r

    def matvec(self, x, level=-1, precision=0):
        y = np.zeros(int(self.dims(level)[0]))
        _ok(lib().pmx_matvec(self.h, level, precision, _p(np.ascontiguousarray(x, float)), _p(y)), "pmx_matvec")
        return y

    def defect(self, x, b, level=-1):
        r = np.zeros(int(self.dims(level)[0]))
        res = np.zeros(1)
        _ok(lib().pmx_defect(self.h, level, _p(np.ascontiguousarray(x, float)),
                             _p(np.ascontiguousarray(b, float)), _p(r), _p(res)), "pmx_defect")
        return r, float(res[0])

    def defect_noalloc(self, x_ptr, b_ptr, r_ptr, res_ptr, level=-1):
        """pmx_defect_noalloc on device arrays (addresses of level-sized FP64
        lattice-ordered vectors and of one FP64 result); ordered on the
        context's stream, no synchronisation."""
        _ok(lib().pmx_defect_noalloc(self.h, level, ctypes.c_void_p(x_ptr), ctypes.c_void_p(b_ptr),
                                     ctypes.c_void_p(r_ptr), ctypes.c_void_p(res_ptr)), "pmx_defect_noalloc")

    def smooth(self, kind, omega, sweeps, sol, f, level=-1):
        s = np.array(sol, float)
        _ok(lib().pmx_smooth(self.h, level, kind, omega, sweeps, _p(s), _p(np.ascontiguousarray(f, float))),
            "pmx_smooth")
        return s

    def prolongate(self, xc, fine_level=-1):
        fl = fine_level % self.nlev
        xf = np.zeros(int(self.dims(fl)[0]))
        _ok(lib().pmx_prolongate(self.h, fl, _p(np.ascontiguousarray(xc, float)), _p(xf)), "pmx_prolongate")
        return xf

    def restrict(self, rf, fine_level=-1):
        fl = fine_level % self.nlev
        rc = np.zeros(int(self.dims(fl - 1)[0]))
        _ok(lib().pmx_restrict(self.h, fl, _p(rc), _p(np.ascontiguousarray(rf, float))), "pmx_restrict")
        return rc

    def solve(self, opt, rhs, sol, check=True):
        o = np.zeros(24)
        o[:len(opt)] = opt
        info = np.zeros(8)
        s = np.array(sol, float)
        rc = lib().pmx_system_solve(self.h, _p(o), _p(np.ascontiguousarray(rhs, float)), _p(s), _p(info))
        if check:
            _ok(rc, "pmx_system_solve")
        return s, info, rc

    def iterate(self, rowptr, kcol, entries, opt, rhs, sol, check=True):
        """pmx_iterate: the Krylov method of opt on the caller's CSR matrix of
        the finest level, preconditioned by the system's MG levels."""
        o = np.zeros(24)
        o[:len(opt)] = opt
        info = np.zeros(8)
        s = np.array(sol, float)
        rp = np.ascontiguousarray(rowptr, np.int32)
        rc = lib().pmx_iterate(self.h, len(rp) - 1, _p(rp), _p(np.ascontiguousarray(kcol, np.int32)),
                               _p(np.ascontiguousarray(entries, float)), _p(o), _p(np.ascontiguousarray(rhs, float)),
                               _p(s), _p(info))
        if check:
            _ok(rc, "pmx_iterate")
        return s, info, rc

    def assemble_cb(self, prm_block, coeffs, cond=None, value=None):
        """pmx_system_assemble_cb with Python callables:
        coeffs(X, Y, Z) -> array (n, 6) of eps, b1, b2, b3, c, f at the points;
        cond(bdcomp, x, y, z) -> int; value(bdcomp, x, y, z) -> float."""
        def cf(n, X, Y, Z, param, cptr):
            x = np.ctypeslib.as_array(X, (n,))
            y = np.ctypeslib.as_array(Y, (n,))
            z = np.ctypeslib.as_array(Z, (n,))
            c = np.asarray(coeffs(x, y, z), float)
            for i in range(n):
                row = cptr[i]
                for k in range(6):
                    row[k] = c[i, k]

        def bc(comp, x, y, z, out):
            out[0] = int(cond(comp, x, y, z))

        def bv(comp, x, y, z, out):
            out[0] = float(value(comp, x, y, z))

        f1 = COEFF_FCT(cf)
        f2 = BOUND_COND_FCT(bc) if cond else None
        f3 = BOUND_VALUE_FCT(bv) if value else None
        p = np.ascontiguousarray(prm_block, float)
        _ok(lib().pmx_system_assemble_cb(self.h, _p(p), ctypes.cast(f1, ctypes.c_void_p),
                                         ctypes.cast(f2, ctypes.c_void_p) if f2 else None,
                                         ctypes.cast(f3, ctypes.c_void_p) if f3 else None), "pmx_system_assemble_cb")

    def set_step_length_control(self, mode):
        """MG step length control (TMGLevel3D::StepLengthControl): 0 off,
        1 finest level, 2 every level above the coarsest."""
        _ok(lib().pmx_system_set_step_length_control(self.h, int(mode)), "pmx_system_set_step_length_control")

    def solve_resident(self, opt, check=True):
        o = np.zeros(24)
        o[:len(opt)] = opt
        info = np.zeros(8)
        rc = lib().pmx_system_solve_resident(self.h, _p(o), _p(info))
        if check:
            _ok(rc, "pmx_system_solve_resident")
        return info, rc

    def errors(self, prm_block):
        """(L2 error, H1 semi-norm error) of the device solution against the
        example's exact solution at prm_block[TIME]."""
        e = np.zeros(2)
        _ok(lib().pmx_system_errors(self.h, _p(np.ascontiguousarray(prm_block, float)), _p(e)), "pmx_system_errors")
        return float(e[0]), float(e[1])

    def solution(self):
        s = np.zeros(int(self.dims(-1)[0]))
        _ok(lib().pmx_system_solution(self.h, _p(s)), "pmx_system_solution")
        return s

    def time_kernel(self, kernel, level=-1, precision=0, reps=10):
        ms = np.zeros(1)
        _ok(lib().pmx_time_kernel(self.h, kernel, level, precision, reps, _p(ms)), "pmx_time_kernel")
        return float(ms[0])

    # ---- time stepping (TSystemTCD3D) --------------------------------------
    def tcd_init(self, prm_block, sol0=None):
        s0 = None if sol0 is None else np.ascontiguousarray(sol0, float)
        _ok(lib().pmx_tcd_init(self.h, _p(np.ascontiguousarray(prm_block, float)), _p(s0)), "pmx_tcd_init")

    def tcd_step(self, prm_block, td, opt, check=True):
        """One time step to prm_block[TIME]; td = (tau, theta1..theta4)."""
        o = np.zeros(24)
        o[:len(opt)] = opt
        info = np.zeros(8)
        rc = lib().pmx_tcd_step(self.h, _p(np.ascontiguousarray(prm_block, float)),
                                _p(np.ascontiguousarray(td, float)), _p(o), _p(info))
        if check:
            _ok(rc, "pmx_tcd_step")
        return info, rc

    def tcd_assemble_arhs(self, prm_block, td):
        _ok(lib().pmx_tcd_assemble_arhs(self.h, _p(np.ascontiguousarray(prm_block, float)),
                                        _p(np.ascontiguousarray(td, float))), "pmx_tcd_assemble_arhs")

    def tcd_assemble_systmat(self):
        _ok(lib().pmx_tcd_assemble_systmat(self.h), "pmx_tcd_assemble_systmat")

    def tcd_solve(self, opt, check=True):
        o = np.zeros(24)
        o[:len(opt)] = opt
        info = np.zeros(8)
        rc = lib().pmx_tcd_solve(self.h, _p(o), _p(info))
        if check:
            _ok(rc, "pmx_tcd_solve")
        return info, rc

    def tcd_restore_massmat(self):
        _ok(lib().pmx_tcd_restore_massmat(self.h), "pmx_tcd_restore_massmat")

    # ---- learned post-smoother -------------------------------------------
    @staticmethod
    def learned_nparams():
        n = np.zeros(1, np.int32)
        _ok(lib().pmx_learned_nparams(_p(n)), "pmx_learned_nparams")
        return int(n[0])

    def learned_init(self, seed=0):
        _ok(lib().pmx_learned_init(self.h, seed), "pmx_learned_init")

    def learned_params(self, params=None, level=-1):
        """Upload `params` (FP32) of a level's network if given; return them."""
        n = self.learned_nparams()
        if params is not None:
            p = np.ascontiguousarray(params, np.float32)
            if p.size != n:
                raise ValueError(f"expected {n} parameters")
            _ok(lib().pmx_learned_params_level(self.h, level, 1, _p(p)), "pmx_learned_params")
        out = np.zeros(n, np.float32)
        _ok(lib().pmx_learned_params_level(self.h, level, 0, _p(out)), "pmx_learned_params")
        return out

    def learned_forward(self, u):
        """Network output (M, M, M) for the input u (M, M, M, 16) indexed [z, y, x, channel]."""
        u = np.ascontiguousarray(u, np.float32)
        if u.ndim != 4 or u.shape[3] != LEARNED_IN:
            raise ValueError(f"input must be (M, M, M, {LEARNED_IN})")
        out = np.zeros(u.shape[:3], np.float32)
        _ok(lib().pmx_learned_forward(self.h, u.shape[0], _p(u), _p(out)), "pmx_learned_forward")
        return out

    def learned_gradient(self, u, t):
        """Squared-error gradient for the input u (M, M, M, 16) and target t (M, M, M)."""
        u = np.ascontiguousarray(u, np.float32)
        if u.ndim != 4 or u.shape[3] != LEARNED_IN:
            raise ValueError(f"input must be (M, M, M, {LEARNED_IN})")
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.learned_nparams(), np.float32)
        loss = np.zeros(1)
        _ok(lib().pmx_learned_gradient(self.h, u.shape[0], _p(u), _p(t), _p(g), _p(loss)), "pmx_learned_gradient")
        return g, float(loss[0])

    def learned_adam(self, lr):
        _ok(lib().pmx_learned_adam(self.h, lr), "pmx_learned_adam")

    def learned_gradient_level(self, u, level=-1):
        """Residual-loss gradient for the scaled defect u on a level's lattice."""
        u = np.ascontiguousarray(u, np.float32)
        g = np.zeros(self.learned_nparams(), np.float32)
        loss = np.zeros(1)
        _ok(lib().pmx_learned_gradient_level(self.h, level, _p(u), _p(g), _p(loss)), "pmx_learned_gradient_level")
        return g, float(loss[0])

    def learned_train(self, opt, samples=16, steps=400, lr=2e-3, seed=0, levels=0, loss=1, data=0, skip=0,
                      batch=1):
        o = np.zeros(24)
        o[:len(opt)] = opt
        t = np.zeros(16)
        t[:9] = [samples, steps, lr, seed, levels, loss, data, skip, batch]
        info = np.zeros(16)
        _ok(lib().pmx_learned_train(self.h, _p(o), _p(t), _p(info)), "pmx_learned_train")
        return info

    def learned_post(self, sol, f, level=-1):
        s = np.array(sol, float)
        _ok(lib().pmx_learned_post(self.h, level, _p(s), _p(np.ascontiguousarray(f, float))), "pmx_learned_post")
        return s
