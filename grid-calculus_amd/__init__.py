"""Python binding of libgridcalc_hip.so, the MI355X engine behind gridcalc's periodic-grid
operators, reductions and explicit time stepping.

This module is plumbing for the test-suite, the benchmark and scripted use: it calls the
C ABI declared in include/gridcalc_hip.h through ctypes and moves numpy arrays to and from
device buffers allocated by the library.  It has no CPU fallback: if the shared library is
missing it raises, and every computation runs in the HIP kernels.

Arrays follow gridcalc's layout: i-fastest (site (i, j, k) at i + Nx*(j + Ny*k)), i.e.
C-ordered numpy arrays of shape (Nz, Ny, Nx) for scalar fields, (Nz, Ny, Nx, 3) for Vec3
fields and (Nz, Ny, Nx, 9) for column-major Mat3 fields.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libgridcalc_hip.so")
REPO_ROOT = os.path.dirname(PKG_DIR)

# status codes (gc_status)
OK, ERR_INVALID, ERR_EXTENT, ERR_CFL, ERR_NONPOS_D, ERR_DEVICE, ERR_NOMEM, ERR_UNSUPPORTED = range(8)
F64, F32 = 0, 1
PAIRWISE, KAHAN = 0, 1
EULER, RK4 = 0, 1

_NP_DTYPE = {F64: np.float64, F32: np.float32}

from . import inputs  # noqa: E402  (seeded synthetic inputs)


class GridcalcError(RuntimeError):
    """A gc_status other than GC_OK; ``code`` holds the status, the message is the
    library's gc_last_error() text (which mirrors the reference's exception text)."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class Grid(C.Structure):
    """gc_grid: extents and cell size (core::Grid, core/Grid.hpp:28-148)."""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double)]

    @staticmethod
    def make(nx, ny, nz, h=(1.0, 1.0, 1.0)):
        if not isinstance(h, (tuple, list)):
            h = (h, h, h)
        return Grid(int(nx), int(ny), int(nz), float(h[0]), float(h[1]), float(h[2]))

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def size(self):
        return self.nx * self.ny * self.nz

    @property
    def cell_volume(self):
        return self.hx * self.hy * self.hz

    def __repr__(self):
        return f"Grid({self.nx}, {self.ny}, {self.nz}, h=({self.hx}, {self.hy}, {self.hz}))"


class CHParams(C.Structure):
    _fields_ = [("M", C.c_double), ("kappa", C.c_double)]


class Density(C.Structure):
    """gc_density: a recognised func::evaluate density and its two parameters."""
    _fields_ = [("kind", C.c_int32), ("a", C.c_double), ("b", C.c_double)]


class ExprInsn(C.Structure):
    """gc_expr_insn: one instruction of a per-site tensor-expression program."""
    _fields_ = [("op", C.c_int32), ("dst", C.c_int32), ("a", C.c_int32), ("b", C.c_int32),
                ("nc", C.c_int32), ("c", C.c_double), ("ptr", C.c_void_p)]


# gc_expr_op
(EXPR_LOAD, EXPR_IDENTITY, EXPR_SCALE, EXPR_ADD, EXPR_SUB, EXPR_MUL, EXPR_TRACE, EXPR_SYM,
 EXPR_MATMUL, EXPR_DDOT) = range(10)

# gc_density_kind
DENSITY = dict(value=0, gradient_energy=1, cahn_hilliard=2, ginzburg_landau=3, pfc=4,
               linear_elastic=5)


def build(force=False):
    """Compiles libgridcalc_hip.so in-tree for gfx950 (hipcc via make)."""
    args = ["make", "-s", "-C", PKG_DIR, "-j8"]
    if force:
        args.insert(2, "-B")
    subprocess.run(args, check=True)


_lib = None


def lib():
    """Loads libgridcalc_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    vp, gp, sz, i64p, dp = C.c_void_p, C.POINTER(Grid), C.c_size_t, C.POINTER(C.c_int64), C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    st = C.c_int
    sig = {
        "gc_last_error": ([], C.c_char_p),
        "gc_version": ([], C.c_int),
        "gc_device_count": ([C.POINTER(C.c_int)], st),
        "gc_context_create": ([C.POINTER(vp), vp], st),
        "gc_context_destroy": ([vp], st),
        "gc_context_sync": ([vp], st),
        "gc_context_stream": ([vp], vp),
        "gc_context_scratch_bytes": ([vp], sz),
        "gc_malloc": ([C.POINTER(vp), sz], st),
        "gc_free": ([vp], st),
        "gc_host_malloc": ([C.POINTER(vp), sz], st),
        "gc_host_free": ([vp], st),
        "gc_memcpy_htod": ([vp, vp, vp, sz], st),
        "gc_memcpy_dtoh": ([vp, vp, vp, sz], st),
        "gc_memcpy_dtod": ([vp, vp, vp, sz], st),
        "gc_timer_start": ([vp], st),
        "gc_timer_stop": ([vp, C.POINTER(C.c_float)], st),
        "gc_laplacian": ([vp, gp, C.c_int, C.c_int, vp, vp], st),
        "gc_gradient": ([vp, gp, C.c_int, C.c_int, vp, vp], st),
        "gc_gradient_vec": ([vp, gp, C.c_int, C.c_int, vp, vp], st),
        "gc_divergence": ([vp, gp, C.c_int, C.c_int, vp, vp], st),
        "gc_divergence_mat": ([vp, gp, C.c_int, C.c_int, vp, vp], st),
        "gc_biharmonic": ([vp, gp, C.c_int, C.c_int, vp, vp], st),
        "gc_mixed_partial": ([vp, gp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp], st),
        "gc_cell_laplacian": ([vp, gp, C.c_int, vp, vp, vp, i64p], st),
        "gc_first_nonpositive": ([vp, gp, C.c_int, vp, i64p], st),
        "gc_max_face_d": ([vp, gp, C.c_int, vp, dp], st),
        "gc_integrate": ([vp, gp, C.c_int, C.c_int, vp, dp], st),
        "gc_integrate_noalloc": ([vp, gp, C.c_int, C.c_int, vp, vp], st),
        "gc_evaluate": ([vp, gp, C.c_int, C.POINTER(Density), C.c_int, vp, dp], st),
        "gc_expr_eval": ([vp, gp, C.POINTER(ExprInsn), C.c_int, C.c_int, C.c_int, vp], st),
        "gc_snapshot_create": ([vp, sz, C.POINTER(vp)], st),
        "gc_snapshot_capture": ([vp, vp, sz], st),
        "gc_snapshot_wait": ([vp, C.POINTER(vp)], st),
        "gc_snapshot_destroy": ([vp], st),
        "gc_ch_free_energy": ([vp, gp, C.c_int, C.c_double, C.c_int, vp, dp], st),
        "gc_gradient_energy": ([vp, gp, C.c_int, C.c_int, vp, dp], st),
        "gc_axpy": ([vp, C.c_int, sz, vp, C.c_double, vp, vp], st),
        "gc_euler_update": ([vp, C.c_int, sz, vp, C.c_double, vp], st),
        "gc_rk4_combine": ([vp, C.c_int, sz, vp, C.c_double, vp, vp, vp, vp], st),
        "gc_ch_rhs": ([vp, gp, C.c_int, C.POINTER(CHParams), vp, vp], st),
        "gc_ch_integrate": ([vp, gp, C.c_int, C.POINTER(CHParams), vp, C.c_double, C.c_int, C.c_int], st),
        "gc_diffuse_const": ([vp, gp, C.c_int, vp, C.c_double, C.c_double, C.c_int, C.c_int], st),
        "gc_diffuse_field": ([vp, gp, C.c_int, vp, vp, C.c_double, C.c_int, C.c_int, i64p], st),
        "gc_ch_config_count": ([gp], C.c_int),
        "gc_ch_config_shape": ([gp, C.c_int, ip, ip, ip], st),
        "gc_ch_config_fused": ([gp, C.c_int], C.c_int),
        "gc_ch_set_config": ([gp, C.c_int, C.c_int], st),
        "gc_ch_get_config": ([gp, C.c_int, ip], st),
        "gc_ch_tune": ([vp, gp, C.c_int, ip, C.POINTER(C.c_float)], st),
        "gc_ensemble_create": ([vp, C.c_int, C.c_int, vp, C.c_double, C.POINTER(vp)], st),
        "gc_ensemble_destroy": ([vp], st),
        "gc_ensemble_set_state": ([vp, C.c_int, vp], st),
        "gc_ensemble_get_state": ([vp, C.c_int, vp], st),
        "gc_ensemble_state_ptr": ([vp, C.c_int, C.POINTER(vp)], st),
        "gc_ensemble_sites": ([vp, C.POINTER(C.c_size_t)], st),
        "gc_ensemble_step": ([vp, C.c_int, C.c_int], st),
        "gc_ensemble_config_shape": ([vp, C.c_int, ip, ip], st),
        "gc_ensemble_set_config": ([vp, C.c_int], st),
        "gc_ensemble_tune": ([vp, ip, C.POINTER(C.c_float)], st),
        "gc_ensemble_diagnostics": ([vp, C.c_int, dp, dp], st),
        "gc_cnn_param_count": ([], C.c_int),
        "gc_cnn_create": ([vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp)], st),
        "gc_cnn_destroy": ([vp], st),
        "gc_cnn_set_params": ([vp, vp, C.c_double], st),
        "gc_cnn_get_params": ([vp, vp], st),
        "gc_cnn_get_grads": ([vp, vp], st),
        "gc_cnn_forward": ([vp, C.c_int, vp, C.c_int, vp], st),
        "gc_cnn_forward_layers": ([vp, C.c_int, vp, C.c_int, vp], st),
        "gc_cnn_loss_grad": ([vp, C.c_int, vp, vp, ip, C.c_int, dp], st),
        "gc_cnn_loss_grad_chain": ([vp, C.c_int, vp, C.c_size_t, ip, C.c_int, C.c_int, dp, dp], st),
        "gc_cnn_adam": ([vp, C.c_double, C.c_double, C.c_double, C.c_double], st),
        "gc_cnn_predict": ([vp, C.c_int, vp, vp, C.c_int], st),
        "gc_cast": ([vp, C.c_int, C.c_int, sz, vp, vp], st),
        "gc_learned_integrate": ([vp, gp, C.c_int, C.POINTER(CHParams), vp, C.c_double, C.c_int,
                                  C.c_int, C.c_double, ip], st),
        "gc_learned_integrate_phases": ([C.POINTER(C.c_void_p), ip, ip, C.c_int, gp, C.c_int,
                                         C.POINTER(CHParams), vp, C.c_double, C.c_int, C.c_int,
                                         C.c_double, ip], st),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _BOUND[:] = list(sig)
    _lib = L
    return L


_BOUND = []


def bound_symbols():
    """Names of every C entry point this binding declares a signature for."""
    lib()
    return list(_BOUND)


def _check(rc):
    if rc != OK:
        msg = lib().gc_last_error().decode("utf-8", "replace")
        raise GridcalcError(rc, msg)


def device_count():
    n = C.c_int(0)
    _check(lib().gc_device_count(C.byref(n)))
    return n.value


class Context:
    """A gc_context: one HIP stream plus the scratch memory the fused kernels use."""

    def __init__(self, stream=None):
        self._h = C.c_void_p()
        self._dependents = []  # weak references to objects that use this context's stream
        _check(lib().gc_context_create(C.byref(self._h), stream))

    def _register(self, obj):
        """obj.close() runs before the context is destroyed (a gc_cnn or gc_ensemble must not
        outlive its context)."""
        self._dependents.append(weakref.ref(obj))

    @property
    def handle(self):
        return self._h

    def sync(self):
        _check(lib().gc_context_sync(self._h))

    def scratch_bytes(self):
        return lib().gc_context_scratch_bytes(self._h)

    def timer_start(self):
        _check(lib().gc_timer_start(self._h))

    def timer_stop(self):
        ms = C.c_float()
        _check(lib().gc_timer_stop(self._h, C.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            for r in self._dependents:
                obj = r()
                if obj is not None:
                    obj.close()
            self._dependents = []
            lib().gc_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeviceArray:
    """A device buffer owned by Python (gc_malloc / gc_free) with a numpy shape/dtype."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._p = C.c_void_p()
        _check(lib().gc_malloc(C.byref(self._p), self.nbytes))

    @property
    def ptr(self):
        return self._p

    @property
    def gc_dtype(self):
        return F64 if self.dtype == np.float64 else F32

    @staticmethod
    def from_numpy(ctx, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype if dtype is not None else a.dtype)
        d = DeviceArray(a.shape, a.dtype)
        _check(lib().gc_memcpy_htod(ctx.handle, d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes))
        return d

    def upload(self, ctx, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(lib().gc_memcpy_htod(ctx.handle, self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def numpy(self, ctx):
        out = np.empty(self.shape, self.dtype)
        _check(lib().gc_memcpy_dtoh(ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return out

    def copy_from(self, ctx, other):
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(lib().gc_memcpy_dtod(ctx.handle, self.ptr, other.ptr, self.nbytes))

    def free(self):
        if self._p:
            lib().gc_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _dt(a):
    return a.gc_dtype


# ---- operators ------------------------------------------------------------------------

def laplacian(ctx, g, f, order=2, out=None):
    out = out or DeviceArray(g.shape, f.dtype)
    _check(lib().gc_laplacian(ctx.handle, C.byref(g), _dt(f), order, f.ptr, out.ptr))
    return out


def gradient(ctx, g, f, order=2, out=None):
    out = out or DeviceArray(g.shape + (3,), f.dtype)
    _check(lib().gc_gradient(ctx.handle, C.byref(g), _dt(f), order, f.ptr, out.ptr))
    return out


def gradient_vec(ctx, g, v, order=2, out=None):
    out = out or DeviceArray(g.shape + (9,), v.dtype)
    _check(lib().gc_gradient_vec(ctx.handle, C.byref(g), _dt(v), order, v.ptr, out.ptr))
    return out


def divergence(ctx, g, v, order=2, out=None):
    out = out or DeviceArray(g.shape, v.dtype)
    _check(lib().gc_divergence(ctx.handle, C.byref(g), _dt(v), order, v.ptr, out.ptr))
    return out


def divergence_mat(ctx, g, m, order=2, out=None):
    out = out or DeviceArray(g.shape + (3,), m.dtype)
    _check(lib().gc_divergence_mat(ctx.handle, C.byref(g), _dt(m), order, m.ptr, out.ptr))
    return out


def biharmonic(ctx, g, f, order=2, out=None):
    out = out or DeviceArray(g.shape, f.dtype)
    _check(lib().gc_biharmonic(ctx.handle, C.byref(g), _dt(f), order, f.ptr, out.ptr))
    return out


def mixed_partial(ctx, g, f, nx, ny, nz, order=2, out=None):
    """d^(nx+ny+nz) f / dx^nx dy^ny dz^nz (1 <= nx+ny+nz <= 4)."""
    out = out or DeviceArray(g.shape, f.dtype)
    _check(lib().gc_mixed_partial(ctx.handle, C.byref(g), _dt(f), nx, ny, nz, order, f.ptr,
                                  out.ptr))
    return out


def cell_laplacian(ctx, g, psi, D, out=None):
    out = out or DeviceArray(g.shape, psi.dtype)
    bad = C.c_int64(-1)
    _check(lib().gc_cell_laplacian(ctx.handle, C.byref(g), _dt(psi), psi.ptr, D.ptr, out.ptr, C.byref(bad)))
    return out


def first_nonpositive(ctx, g, D):
    bad = C.c_int64(-1)
    _check(lib().gc_first_nonpositive(ctx.handle, C.byref(g), _dt(D), D.ptr, C.byref(bad)))
    return bad.value


def max_face_d(ctx, g, D):
    r = C.c_double()
    _check(lib().gc_max_face_d(ctx.handle, C.byref(g), _dt(D), D.ptr, C.byref(r)))
    return r.value


# ---- reductions -------------------------------------------------------------------------

def integrate(ctx, g, f, method=PAIRWISE):
    r = C.c_double()
    _check(lib().gc_integrate(ctx.handle, C.byref(g), _dt(f), method, f.ptr, C.byref(r)))
    return r.value


def integrate_noalloc(ctx, g, f, method, result_dev):
    """gc_integrate_noalloc: cellVolume * sum of f written to the device array result_dev
    (one float64), queued on the context's stream without allocating or synchronising."""
    _check(lib().gc_integrate_noalloc(ctx.handle, C.byref(g), _dt(f), method, f.ptr,
                                      result_dev.ptr))


def ch_free_energy(ctx, g, psi, kappa, method=PAIRWISE):
    r = C.c_double()
    _check(lib().gc_ch_free_energy(ctx.handle, C.byref(g), _dt(psi), kappa, method, psi.ptr, C.byref(r)))
    return r.value


def evaluate(ctx, g, kind, field, a=0.0, b=0.0, method=PAIRWISE):
    """func::evaluate of a recognised density (a name of DENSITY) on a scalar field, or on a
    Vec3 field (shape (..., 3)) for "linear_elastic"; one fused pass on the device."""
    r = C.c_double()
    D = Density(DENSITY[kind], a, b)
    _check(lib().gc_evaluate(ctx.handle, C.byref(g), _dt(field), C.byref(D), method, field.ptr,
                             C.byref(r)))
    return r.value


def expr_eval(ctx, g, prog, out_slot, out_nc):
    """Runs a list of ExprInsn at every site; returns the FP64 result field of out_nc values
    per site."""
    out = DeviceArray(g.shape + ((out_nc,) if out_nc > 1 else ()), np.float64)
    arr = (ExprInsn * len(prog))(*prog)
    _check(lib().gc_expr_eval(ctx.handle, C.byref(g), arr, len(prog), out_slot, out_nc, out.ptr))
    return out


def gradient_energy(ctx, g, psi, method=PAIRWISE):
    r = C.c_double()
    _check(lib().gc_gradient_energy(ctx.handle, C.byref(g), _dt(psi), method, psi.ptr, C.byref(r)))
    return r.value


# ---- time stepping ------------------------------------------------------------------------

def axpy(ctx, a, scale, b, out):
    n = int(np.prod(a.shape))
    _check(lib().gc_axpy(ctx.handle, _dt(a), n, a.ptr, scale, b.ptr, out.ptr))


def euler_update(ctx, psi, dt, r):
    n = int(np.prod(psi.shape))
    _check(lib().gc_euler_update(ctx.handle, _dt(psi), n, psi.ptr, dt, r.ptr))


def rk4_combine(ctx, psi, dt, k1, k2, k3, k4):
    n = int(np.prod(psi.shape))
    _check(lib().gc_rk4_combine(ctx.handle, _dt(psi), n, psi.ptr, dt, k1.ptr, k2.ptr, k3.ptr, k4.ptr))


def ch_rhs(ctx, g, psi, M=1.0, kappa=1.0, out=None):
    out = out or DeviceArray(g.shape, psi.dtype)
    P = CHParams(M, kappa)
    _check(lib().gc_ch_rhs(ctx.handle, C.byref(g), _dt(psi), C.byref(P), psi.ptr, out.ptr))
    return out


def ch_integrate(ctx, g, psi, M, kappa, dt, n_steps, integrator=RK4):
    """In place: psi advances n_steps of the fused Cahn-Hilliard integrator."""
    P = CHParams(M, kappa)
    _check(lib().gc_ch_integrate(ctx.handle, C.byref(g), _dt(psi), C.byref(P), psi.ptr, dt,
                                 n_steps, integrator))


def diffuse_const(ctx, g, psi, D, dt, n_steps, integrator=EULER):
    _check(lib().gc_diffuse_const(ctx.handle, C.byref(g), _dt(psi), psi.ptr, D, dt, n_steps, integrator))


def diffuse_field(ctx, g, psi, D, dt, n_steps, integrator=EULER):
    bad = C.c_int64(-1)
    _check(lib().gc_diffuse_field(ctx.handle, C.byref(g), _dt(psi), psi.ptr, D.ptr, dt, n_steps,
                                  integrator, C.byref(bad)))


# ---- Cahn-Hilliard stage autotuner --------------------------------------------------------

def ch_config_count(g):
    return lib().gc_ch_config_count(C.byref(g))


def ch_config_shape(g, i):
    ox, oy, zc = C.c_int(), C.c_int(), C.c_int()
    _check(lib().gc_ch_config_shape(C.byref(g), i, C.byref(ox), C.byref(oy), C.byref(zc)))
    return ox.value, oy.value, zc.value


def ch_config_kind(g, i):
    """0: one kernel per RK4 stage; 1: two-stage kernels with LDS plane rings; 2: two-stage
    kernels with register queues along z; 3: one kernel per RK4 step (2-D)."""
    r = lib().gc_ch_config_fused(C.byref(g), i)
    if r < 0:
        raise GridcalcError(1, f"gridcalc: config id {i} out of range")
    return r


def ch_config_fused(g, i):
    """True if candidate i fuses RK4 stages: two two-stage kernels (3-D) or one kernel per
    step (2-D)."""
    return ch_config_kind(g, i) > 0


def ch_set_config(g, dtype, i):
    _check(lib().gc_ch_set_config(C.byref(g), dtype, i))


def ch_get_config(g, dtype):
    i = C.c_int()
    _check(lib().gc_ch_get_config(C.byref(g), dtype, C.byref(i)))
    return i.value


def ch_tune(ctx, g, dtype):
    """Times every candidate tile shape; returns (best id, [ms per RK4 step])."""
    n = ch_config_count(g)
    ms = (C.c_float * max(n, 1))()
    best = C.c_int()
    _check(lib().gc_ch_tune(ctx.handle, C.byref(g), dtype, C.byref(best), ms))
    return best.value, [ms[i] for i in range(n)]


# ---- asynchronous snapshots ---------------------------------------------------------------

class Snapshot:
    """gc_snapshot: capture() copies a device array into the slot on the context's stream and
    queues its device-to-host copy on the slot's own stream; numpy() waits for that copy."""

    def __init__(self, ctx, shape, dtype=np.float64):
        self.ctx, self.shape, self.dtype = ctx, tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._h = C.c_void_p()
        _check(lib().gc_snapshot_create(ctx.handle, self.nbytes, C.byref(self._h)))
        ctx._register(self)

    def capture(self, arr):
        _check(lib().gc_snapshot_capture(self._h, arr.ptr, arr.nbytes))

    def numpy(self):
        p = C.c_void_p()
        _check(lib().gc_snapshot_wait(self._h, C.byref(p)))
        buf = (C.c_char * self.nbytes).from_address(p.value)
        return np.frombuffer(buf, dtype=self.dtype).reshape(self.shape).copy()

    def close(self):
        if self._h:
            lib().gc_snapshot_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- batched 2-D Cahn-Hilliard ensembles --------------------------------------------------

class EnsembleMember(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("hx", C.c_double), ("hy", C.c_double),
                ("M", C.c_double), ("kappa", C.c_double)]


class Ensemble:
    """gc_ensemble: independent 2-D Cahn-Hilliard grids advanced together by one
    temporally blocked RK4 kernel per step.  members: list of dicts with nx, ny and
    optional hx, hy, M, kappa."""

    def __init__(self, ctx, members, dt, dtype=F64):
        self.ctx = ctx
        self.dtype = dtype
        self.np_dtype = _NP_DTYPE[dtype]
        self.members = [dict(nx=m["nx"], ny=m["ny"], hx=m.get("hx", 1.0), hy=m.get("hy", 1.0),
                             M=m.get("M", 1.0), kappa=m.get("kappa", 1.0)) for m in members]
        arr = (EnsembleMember * len(members))(*[
            EnsembleMember(m["nx"], m["ny"], m["hx"], m["hy"], m["M"], m["kappa"])
            for m in self.members])
        self._h = C.c_void_p()
        _check(lib().gc_ensemble_create(ctx.handle, dtype, len(members), arr, dt, C.byref(self._h)))
        ctx._register(self)

    def set_state(self, i, a):
        m = self.members[i]
        a = np.ascontiguousarray(a, dtype=self.np_dtype).reshape(m["ny"], m["nx"])
        _check(lib().gc_ensemble_set_state(self._h, i, a.ctypes.data_as(C.c_void_p)))

    def get_state(self, i):
        m = self.members[i]
        out = np.empty((1, m["ny"], m["nx"]), self.np_dtype)
        _check(lib().gc_ensemble_get_state(self._h, i, out.ctypes.data_as(C.c_void_p)))
        return out

    def step(self, n, use_graph=True):
        _check(lib().gc_ensemble_step(self._h, n, 1 if use_graph else 0))

    def sites(self):
        s = C.c_size_t()
        _check(lib().gc_ensemble_sites(self._h, C.byref(s)))
        return s.value

    def config_shape(self, i):
        ox, oy = C.c_int(), C.c_int()
        _check(lib().gc_ensemble_config_shape(self._h, i, C.byref(ox), C.byref(oy)))
        return ox.value, oy.value

    def set_config(self, i):
        _check(lib().gc_ensemble_set_config(self._h, i))

    def tune(self):
        n = 5 if self.dtype == F32 else 4
        ms = (C.c_float * n)()
        best = C.c_int()
        _check(lib().gc_ensemble_tune(self._h, C.byref(best), ms))
        return best.value, [ms[i] for i in range(n)]

    def diagnostics(self, method=KAHAN):
        n = len(self.members)
        mass = (C.c_double * n)()
        F = (C.c_double * n)()
        _check(lib().gc_ensemble_diagnostics(self._h, method, mass, F))
        return list(mass), list(F)

    def close(self):
        if self._h:
            lib().gc_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- learned corrector (6-layer periodic CNN) ----------------------------------------------

N_CNN_PARAMS = 37601


def cast(ctx, src, dst):
    """Element-wise dtype conversion between two device arrays of equal size."""
    n = int(np.prod(src.shape))
    _check(lib().gc_cast(ctx.handle, src.gc_dtype, dst.gc_dtype, n, src.ptr, dst.ptr))


class Corrector:
    """gc_cnn: the 2-D Cahn-Hilliard corrector network on H x W images (H % 8 == 0,
    W % 32 == 0), batches up to batch_max."""

    def __init__(self, ctx, H, W, batch_max):
        self.ctx, self.H, self.W, self.batch_max = ctx, H, W, batch_max
        self._h = C.c_void_p()
        _check(lib().gc_cnn_create(ctx.handle, H, W, batch_max, C.byref(self._h)))
        self.out = DeviceArray((batch_max, H, W), np.float32)
        ctx._register(self)

    def set_params(self, p, scale):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == N_CNN_PARAMS
        _check(lib().gc_cnn_set_params(self._h, p.ctypes.data_as(C.c_void_p), scale))

    def params(self):
        p = np.empty(N_CNN_PARAMS, np.float32)
        _check(lib().gc_cnn_get_params(self._h, p.ctypes.data_as(C.c_void_p)))
        return p

    def grads(self):
        g = np.empty(N_CNN_PARAMS, np.float32)
        _check(lib().gc_cnn_get_grads(self._h, g.ctypes.data_as(C.c_void_p)))
        return g

    def forward(self, c, B, layers=False):
        """Raw network output for the first B images of device array c -> numpy (B, H, W);
        layers=True runs the per-layer kernels instead of the fused inference kernel."""
        fn = lib().gc_cnn_forward_layers if layers else lib().gc_cnn_forward
        _check(fn(self._h, c.gc_dtype, c.ptr, B, self.out.ptr))
        return self.out.numpy(self.ctx)[:B]

    def loss_grad(self, c, target, idx=None, B=None):
        if idx is not None:
            idx = np.ascontiguousarray(idx, dtype=np.int32)
            B = idx.size
            ip = idx.ctypes.data_as(C.POINTER(C.c_int))
        else:
            ip = None
        loss = C.c_double()
        _check(lib().gc_cnn_loss_grad(self._h, c.gc_dtype, c.ptr, target.ptr, ip, B, C.byref(loss)))
        return loss.value

    def chain_loss_grad(self, data, idx, K, weights=None):
        """Unrolled-chain loss and gradient (gc_cnn_loss_grad_chain): chains start at images
        idx of device array data (images along its leading axes, H x W each); targets are the
        K images after each start; weights (K values, default 1/K each) weigh the steps."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        n_images = int(np.prod(data.shape)) // (self.H * self.W)
        wp = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float64)
            assert w.size == K
            wp = w.ctypes.data_as(C.POINTER(C.c_double))
        loss = C.c_double()
        _check(lib().gc_cnn_loss_grad_chain(self._h, data.gc_dtype, data.ptr, n_images,
                                            idx.ctypes.data_as(C.POINTER(C.c_int)), idx.size, K,
                                            wp, C.byref(loss)))
        return loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(lib().gc_cnn_adam(self._h, lr, beta1, beta2, eps))

    def predict(self, c, c_new, B=1):
        _check(lib().gc_cnn_predict(self._h, c.gc_dtype, c.ptr, c_new.ptr, B))

    def integrate(self, g, psi, M, kappa, dt, substeps, n_macro, mass_tol=1e-6):
        """Learned macro steps with the plain-RK4 fallback; returns the fallback count."""
        P = CHParams(M, kappa)
        nf = C.c_int()
        _check(lib().gc_learned_integrate(self._h, C.byref(g), psi.gc_dtype, C.byref(P), psi.ptr,
                                          dt, substeps, n_macro, mass_tol, C.byref(nf)))
        return nf.value

    def close(self):
        if self._h:
            lib().gc_cnn_destroy(self._h)
            self._h = C.c_void_p()
        self.out.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
