"""Loader for the CPU oracle (libde_oracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and the benchmark's CPU-baseline leg
as the checker; the product package never imports it.
"""
import ctypes as C
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "differential-equations_amd"))

from deamd import _abi  # noqa: E402

LIB_PATH = os.path.join(HERE, "_build", "libde_oracle.so")
POW_RESTATED = 0   # de::libm_pow, the pow the GPU kernels run
POW_HOST = 1       # this host's C library pow (what the reference binary calls)

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        lib.de_oracle_ode_solve.restype = C.c_int
        lib.de_oracle_ode_solve.argtypes = [C.POINTER(_abi.OdeBatch), C.POINTER(_abi.OdeResult),
                                            C.c_int, C.c_int]
        lib.de_oracle_sde_solve.restype = C.c_int
        lib.de_oracle_sde_solve.argtypes = [C.POINTER(_abi.SdeBatch), C.POINTER(_abi.SdeResult),
                                            C.c_int]
        lib.de_oracle_pow.restype = C.c_double
        lib.de_oracle_pow.argtypes = [C.c_double, C.c_double]
        lib.de_oracle_pow_many.restype = None
        lib.de_oracle_pow_many.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        lib.de_oracle_dop853_tableau.restype = None
        lib.de_oracle_dop853_tableau.argtypes = [C.c_void_p] * 6
        lib.de_oracle_set_generic_dopri5.restype = None
        lib.de_oracle_set_generic_dopri5.argtypes = [C.c_int]
        lib.de_oracle_constrain_step_size.restype = C.c_double
        lib.de_oracle_constrain_step_size.argtypes = [C.c_double] * 3
        lib.de_oracle_validate_step_size.restype = C.c_int
        lib.de_oracle_validate_step_size.argtypes = [C.c_double] * 5
        lib.de_oracle_stiff_check.restype = None
        lib.de_oracle_stiff_check.argtypes = [C.c_void_p] * 4 + [C.c_int64]
        lib.de_oracle_host_pow_many.restype = None
        lib.de_oracle_host_pow_many.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        lib.de_oracle_noise_normal.restype = C.c_double
        lib.de_oracle_noise_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        lib.de_oracle_philox.restype = None
        lib.de_oracle_philox.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        lib.de_oracle_dopri5_tableau.restype = None
        lib.de_oracle_erk_tableau.restype = C.c_int
        lib.de_oracle_fmt_debug.restype = C.c_int
        lib.de_oracle_fmt_debug.argtypes = [C.c_double, C.c_char_p]
        lib.de_oracle_ode_csv.restype = C.c_int
        lib.de_oracle_ode_csv.argtypes = [C.POINTER(_abi.OdeBatch), C.POINTER(_abi.OdeResult),
                                          C.c_int32, C.c_void_p, C.c_void_p]
        lib.de_oracle_erk_tableau.argtypes = [C.c_int] + [C.c_void_p] * 6
        lib.de_oracle_dopri5_tableau.argtypes = [C.c_void_p] * 5
        _lib = lib
    return _lib


def dopri5_tableau():
    import numpy as np
    c, b, er, bi0 = (np.empty(7) for _ in range(4))
    a = np.empty(49)
    load().de_oracle_dopri5_tableau(c.ctypes.data, a.ctypes.data, b.ctypes.data,
                                    er.ctypes.data, bi0.ctypes.data)
    return c, a.reshape(7, 7), b, er, bi0


def dop853_tableau():
    """(c[16], a[16,16], b[12], bh[12], er[12], bi[16,16]) as the oracle uses them."""
    import numpy as np
    c, b, bh, er = np.empty(16), np.empty(12), np.empty(12), np.empty(12)
    a, bi = np.empty(256), np.empty(256)
    load().de_oracle_dop853_tableau(c.ctypes.data, a.ctypes.data, b.ctypes.data,
                                    bh.ctypes.data, er.ctypes.data, bi.ctypes.data)
    return c, a.reshape(16, 16), b, bh, er, bi.reshape(16, 16)


def erk_tableau(method):
    """The embedded family's tableau as the oracle uses it: a dict with S, I, O,
    fsal, has_bi and the arrays C[I], A[I, I], B[S], BH[S], BI[I, O]."""
    import numpy as np
    dims = np.zeros(5, dtype=np.int32)
    if load().de_oracle_erk_tableau(method, None, None, None, None, None, dims.ctypes.data) != 0:
        raise ValueError("not an embedded-family method id: %d" % method)
    S, I, O = int(dims[0]), int(dims[1]), int(dims[2])
    c, a, b, bh, bi = np.empty(I), np.empty(I * I), np.empty(S), np.empty(S), np.empty(I * O)
    load().de_oracle_erk_tableau(method, c.ctypes.data, a.ctypes.data, b.ctypes.data,
                                 bh.ctypes.data, bi.ctypes.data, dims.ctypes.data)
    return dict(S=S, I=I, O=O, fsal=bool(dims[3]), has_bi=bool(dims[4]), C=c,
                A=a.reshape(I, I), B=b, BH=bh, BI=bi.reshape(I, O))


def fmt_debug(x):
    """Rust's `{:?}` text of a double (str)."""
    buf = C.create_string_buffer(32)
    n = load().de_oracle_fmt_debug(float(x), buf)
    return buf.raw[:n].decode()


def ode_csv(ensemble, solution, source):
    """Solution::to_csv of every trajectory of a solved ensemble: (bytes, offsets)."""
    import numpy as np
    b, res, keep = ensemble.batch()
    r = solution.struct()
    offsets = np.zeros(ensemble.n + 1, dtype=np.int64)
    if load().de_oracle_ode_csv(C.byref(b), C.byref(r), source, offsets.ctypes.data, None) != 0:
        raise ValueError("bad csv source %d" % source)
    out = C.create_string_buffer(max(1, int(offsets[-1])))
    load().de_oracle_ode_csv(C.byref(b), C.byref(r), source, offsets.ctypes.data, out)
    return out.raw[:int(offsets[-1])], offsets


def solve_ode(ensemble, pow_mode=POW_RESTATED, threads=1):
    """Run an deamd.OdeEnsemble through the oracle; returns a deamd.Solution
    (with a point output, sized by a first counting solve as on the GPU)."""
    def once(ens, points_offset):
        b, res, keep = ens.batch(points_offset)
        rc = load().de_oracle_ode_solve(C.byref(b), C.byref(res.struct()), pow_mode, threads)
        if rc != 0:
            raise RuntimeError("oracle ode solve failed: %d" % rc)
        res._solved_by = (b, keep)
        return res
    return ensemble.solve(solver=once)


def solve_sde(ensemble, threads=1):
    b, res, keep = ensemble.batch()
    rc = load().de_oracle_sde_solve(C.byref(b), C.byref(res.struct()), threads)
    if rc != 0:
        raise RuntimeError("oracle sde solve failed: %d" % rc)
    return res


def pow_many(x, y, host=False):
    """de::libm_pow (host=False) or this host's C library pow (host=True)."""
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
    out = np.empty_like(x)
    fn = load().de_oracle_host_pow_many if host else load().de_oracle_pow_many
    fn(x.ctypes.data, y.ctypes.data, out.ctypes.data, x.size)
    return out


def constrain_step_size(h, h_min, h_max):
    return load().de_oracle_constrain_step_size(h, h_min, h_max)


def validate_step_size(h0, h_min, h_max, t0, tf):
    """0 if valid, else the number of the first failed check (1 .. 9)."""
    return load().de_oracle_validate_step_size(h0, h_min, h_max, t0, tf)


def set_generic_dopri5(on):
    """Run DOPRI5 through the generic Dormand-Prince restatement (test switch)."""
    load().de_oracle_set_generic_dopri5(1 if on else 0)


def philox(ctr, k0, k1):
    import numpy as np
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    out = np.empty(4, dtype=np.uint32)
    load().de_oracle_philox(c.ctypes.data, k0, k1, out.ctypes.data)
    return out
