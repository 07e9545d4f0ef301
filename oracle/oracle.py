"""CPU oracle for the LocVolCalib hot path -- TEST INFRASTRUCTURE ONLY.

Python (ctypes) access to ``oracle/_build/liblvc_oracle.so``, a plain-C
restatement of ``finpar/LocVolCalib.fut`` (see ``lvc_oracle_impl.h``).  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker; the product
library never routes through it.

Functions and what pins each one (the checks live in ``tests/test_oracle.py``):

=====================  =====================================  ==========================================
oracle function        reference                              pinned by
=====================  =====================================  ==========================================
init_grid              LocVolCalib.fut:12-23                  closed forms: x[xindex] == s0, y[NY/2] ==
                                                              log(alpha), uniform spacing, timeline ends
init_grid, grid32      :15-21 scalars in f32, arrays in f64   myXindex == the f32 restatement's index;
                                                              x[xindex] == s0; f64 prices vs the f32
                                                              restatement at the fixtures' tolerance;
                                                              closed form s0 - K where the f32 index is
                                                              integral (t=4, alpha=0.2)
init_operator          LocVolCalib.fut:26-44                  exact identities: D*1 = 0, D*x = 1,
                                                              DD*x = 0, DD*x^2 = 2, boundary literals
set_payoff             LocVolCalib.fut:46-47                  direct formula
update_params          LocVolCalib.fut:50-66                  separable closed form
                                                              x^(2 beta) exp(2y - nu^2 t)
tridag_par             LocVolCalib.fut:68-124                 dense solve (numpy.linalg.solve) for
                                                              n <= 16; agreement with tridag_seq;
                                                              ocean-sim golden vectors (tridiag32-small)
tridag_seq             misc/ocean-sim/tridiag.fut:81-95       dense solve; ocean-sim golden vectors
                                                              (bit-exact)
tridag_par4            misc/ocean-sim/tridiag.fut:4-79        dense solve for n <= 16; agreement with
                                                              tridag_seq; ocean-sim golden vectors
explicit_method        LocVolCalib.fut:126-153                dense matrix-vector product on tiny grids
value / main (f32)     LocVolCalib.fut:211-252                the reference's fixtures small/medium/large
                                                              (.in -> .out), closed form price = s0 - K
                                                              for K < x_min, monotone/convex in K
value / main (f64)     f32 -> f64 type substitution of the    closed form, monotone/convex in K, f64 vs
                       same lines                             the f32 fixtures at fixture precision
=====================  =====================================  ==========================================

Below ~1e-6 absolute the f32 bit pattern of the Futhark C backend (its fusion
order and libm calls) is not pinned by anything in the reference; the FP64
restatement is pinned only by the closed form and invariants above.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LVC_ORACLE_LIB names another build of the same sources (the tests' sanitizer
# build, oracle/_build/liblvc_oracle_asan.so)
_LIB_PATH = os.environ.get("LVC_ORACLE_LIB") or os.path.join(_HERE, "_build", "liblvc_oracle.so")
_lib = None

_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_ptr = ctypes.c_void_p


def build() -> str:
    """Compile the oracle with its Makefile (gcc) and return the .so path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        for suf, real in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
            f = getattr(_lib, "lvc_value" + suf)
            f.restype = real
            f.argtypes = [_c_i64, _c_i64, _c_i64] + [real] * 6 + [ctypes.c_int, _ptr]
            f = getattr(_lib, "lvc_main" + suf)
            f.restype = ctypes.c_int
            f.argtypes = [_c_i32] * 4 + [real] * 5 + [ctypes.c_int, ctypes.c_int, _ptr]
            f = getattr(_lib, "lvc_batch" + suf)
            f.restype = ctypes.c_int
            f.argtypes = [_c_i32] + [_ptr] * 9 + [ctypes.c_int, ctypes.c_int, _ptr]
            f = getattr(_lib, "lvc_init_grid" + suf)
            f.restype = None
            f.argtypes = [real] * 4 + [_c_i64] * 3 + [_ptr] * 5
            f = getattr(_lib, "lvc_init_operator" + suf)
            f.restype = None
            f.argtypes = [_ptr, _c_i64, _ptr, _ptr]
            f = getattr(_lib, "lvc_set_payoff" + suf)
            f.restype = None
            f.argtypes = [real, _ptr, _c_i64, _c_i64, _ptr]
            f = getattr(_lib, "lvc_update_params" + suf)
            f.restype = None
            f.argtypes = [_ptr, _ptr, _c_i64, _c_i64, real, real, real, _ptr]
            for name in ("lvc_tridag_par", "lvc_tridag_seq", "lvc_tridag_par4"):
                f = getattr(_lib, name + suf)
                f.restype = None
                f.argtypes = [_ptr] * 4 + [_c_i64, _ptr, _ptr]
            f = getattr(_lib, "lvc_explicit" + suf)
            f.restype = None
            f.argtypes = [_ptr] * 5 + [_c_i64, _c_i64, _ptr]
        _lib.lvc_oracle_set_mutation.argtypes = [ctypes.c_int]
        _lib.lvc_oracle_set_mutation.restype = None
        _lib.lvc_oracle_set_grid32.argtypes = [ctypes.c_int]
        _lib.lvc_oracle_set_grid32.restype = None
    return _lib


def _suffix(dtype) -> tuple[str, type]:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "_f32", np.float32
    if dt == np.float64:
        return "_f64", np.float64
    raise ValueError(f"oracle supports float32/float64, not {dt}")


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def set_mutation(m: int) -> None:
    """Plant mistake number ``m`` in the oracle (0 = none).  Tests only."""
    lib().lvc_oracle_set_mutation(int(m))


class _Grid32:
    """initGrid's scalars (stdX, dx, myXindex, x[0]; LocVolCalib.fut:15-21)
    in f32 for the duration of a call: with dtype float64 the GPU library's
    form of the reference's f32 entry (the f32 grid index, an FP64 grid)."""

    def __init__(self, on):
        self.on = bool(on)

    def __enter__(self):
        if self.on:
            lib().lvc_oracle_set_grid32(1)

    def __exit__(self, *exc):
        if self.on:
            lib().lvc_oracle_set_grid32(0)


def main(outer, numX, numY, numT, s0, t, alpha, nu, beta, dtype=np.float32,
         thomas=False, nthreads=1, grid32=False) -> np.ndarray:
    """LocVolCalib.fut:238-252: prices of ``outer`` options, strike 0.001*i
    (formed in ``dtype``).  ``grid32``: initGrid's scalars in f32."""
    suf, real = _suffix(dtype)
    out = np.zeros(max(int(outer), 0), dtype=real)
    with _Grid32(grid32):
        rc = getattr(lib(), "lvc_main" + suf)(outer, numX, numY, numT, s0, t, alpha,
                                              nu, beta, int(thomas), int(nthreads),
                                              _p(out) if out.size else None)
    if rc != 0:
        raise ValueError(f"oracle main failed with code {rc}")
    return out


def value(numX, numY, numT, s0, strike, t, alpha, nu, beta, dtype=np.float64,
          thomas=False, grid32=False) -> float:
    """LocVolCalib.fut:211-236 for one option."""
    suf, real = _suffix(dtype)
    err = ctypes.c_int(0)
    with _Grid32(grid32):
        v = getattr(lib(), "lvc_value" + suf)(numX, numY, numT, s0, strike, t,
                                               alpha, nu, beta, int(thomas),
                                               ctypes.addressof(err))
    if err.value:
        raise ValueError(f"oracle value failed with code {err.value}")
    return float(v)


def batch(numX, numY, numT, s0, strike, t, alpha, nu, beta, dtype=np.float64,
          thomas=False, nthreads=1, grid32=False, allow_nan=False) -> np.ndarray:
    """Per-problem sizes and parameters; same arithmetic as ``value``.  With
    ``allow_nan`` a problem whose price index lies outside its grid gives NaN
    instead of raising (the batch entry's convention)."""
    suf, real = _suffix(dtype)
    n = len(numX)
    ints = [np.ascontiguousarray(v, dtype=np.int32) for v in (numX, numY, numT)]
    reals = [np.ascontiguousarray(v, dtype=real) for v in (s0, strike, t, alpha, nu, beta)]
    out = np.zeros(n, dtype=real)
    if n == 0:
        return out
    with _Grid32(grid32):
        rc = getattr(lib(), "lvc_batch" + suf)(n, *[_p(a) for a in ints],
                                               *[_p(a) for a in reals], int(thomas),
                                               int(nthreads), _p(out))
    if rc != 0 and not (allow_nan and rc == 3):
        raise ValueError(f"oracle batch failed with code {rc}")
    return out


def init_grid(s0, alpha, nu, t, numX, numY, numT, dtype=np.float64, grid32=False):
    """LocVolCalib.fut:12-23 -> (xindex, yindex, x, y, timeline)."""
    suf, real = _suffix(dtype)
    x = np.zeros(numX, real)
    y = np.zeros(numY, real)
    tl = np.zeros(numT, real)
    xi = np.zeros(1, np.int32)
    yi = np.zeros(1, np.int32)
    with _Grid32(grid32):
        getattr(lib(), "lvc_init_grid" + suf)(s0, alpha, nu, t, numX, numY, numT,
                                              _p(xi), _p(yi), _p(x), _p(y), _p(tl))
    return int(xi[0]), int(yi[0]), x, y, tl


def init_operator(x: np.ndarray):
    """LocVolCalib.fut:26-44 -> (D [n][3], DD [n][3])."""
    suf, real = _suffix(x.dtype)
    x = np.ascontiguousarray(x)
    n = x.shape[0]
    D = np.zeros((n, 3), real)
    DD = np.zeros((n, 3), real)
    getattr(lib(), "lvc_init_operator" + suf)(_p(x), n, _p(D), _p(DD))
    return D, DD


def set_payoff(strike, x: np.ndarray, numY: int) -> np.ndarray:
    suf, real = _suffix(x.dtype)
    x = np.ascontiguousarray(x)
    out = np.zeros((numY, x.shape[0]), real)
    getattr(lib(), "lvc_set_payoff" + suf)(strike, _p(x), x.shape[0], numY, _p(out))
    return out


def update_params(x: np.ndarray, y: np.ndarray, tnow, beta, nu) -> np.ndarray:
    """VarX [numY][numX] of LocVolCalib.fut:61-65."""
    suf, real = _suffix(x.dtype)
    x = np.ascontiguousarray(x)
    y = np.ascontiguousarray(y, dtype=real)
    out = np.zeros((y.shape[0], x.shape[0]), real)
    getattr(lib(), "lvc_update_params" + suf)(_p(x), _p(y), x.shape[0], y.shape[0],
                                              tnow, beta, nu, _p(out))
    return out


def _tridag(name, a, b, c, y):
    suf, real = _suffix(np.asarray(y).dtype)
    a, b, c, y = (np.ascontiguousarray(v, dtype=real) for v in (a, b, c, y))
    n = y.shape[0]
    out = np.zeros(n, real)
    tmp = np.zeros(5 * n, real)
    getattr(lib(), name + suf)(_p(a), _p(b), _p(c), _p(y), n, _p(out), _p(tmp))
    return out


def tridag_par(a, b, c, y) -> np.ndarray:
    """LocVolCalib.fut:68-124, scan form."""
    return _tridag("lvc_tridag_par", a, b, c, y)


def tridag_seq(a, b, c, y) -> np.ndarray:
    """misc/ocean-sim/tridiag.fut:81-95, Thomas."""
    return _tridag("lvc_tridag_seq", a, b, c, y)


def tridag_par4(a, b, c, y) -> np.ndarray:
    """misc/ocean-sim/tridiag.fut:4-79, the ocean-sim program's scan form
    (un-normalised 2x2 matrices in recurrence 1)."""
    return _tridag("lvc_tridag_par4", a, b, c, y)


def explicit_method(D, DD, mu, var, res) -> np.ndarray:
    """LocVolCalib.fut:126-153 on a [rows][m] operand."""
    suf, real = _suffix(np.asarray(res).dtype)
    D, DD, mu, var, res = (np.ascontiguousarray(v, dtype=real) for v in (D, DD, mu, var, res))
    rows, m = res.shape
    out = np.zeros((rows, m), real)
    getattr(lib(), "lvc_explicit" + suf)(_p(D), _p(DD), _p(mu), _p(var), _p(res),
                                         rows, m, _p(out))
    return out


# ---- OptionPricing (finpar/OptionPricing.fut) -------------------------------
# op_oracle.c: the per-path payoffs of the reference in f32 (Sobol :12-35,
# inverse Gaussian :40-140, Brownian bridge :145-183, Black-Scholes
# :190-225, payoffs :230-293), summed over paths in f64.  Pinned by the
# reference's three datasets (OptionPricing-data/{small,medium,large}.out),
# the Sobol construction against a brute-force Gray-code recurrence, and the
# inverse Gaussian against the normal CDF (tests/test_oracle_op.py).

def op_set_mutation(m: int) -> None:
    """Plant mistake m (0 = none) in the OptionPricing oracle: 1 Gray code
    dropped, 2 sign of dp^2 in the central polynomial, 3 left weight for the
    right neighbour in the bridge, 4 the correlation summed over the whole
    row instead of the lower triangle, 5 the
    second-to-last date's prices in payoff3."""
    L = lib()
    L.op_oracle_set_mutation.argtypes = [ctypes.c_int]
    L.op_oracle_set_mutation.restype = None
    L.op_oracle_set_mutation(int(m))


def op_sobol(n: int, dir_vs: np.ndarray) -> np.ndarray:
    """sobolIndR (:29-35) of path n over dir_vs [dims][bits]."""
    dv = np.ascontiguousarray(dir_vs, dtype=np.int32)
    out = np.empty(dv.shape[0], np.float32)
    L = lib()
    L.op_sobol.argtypes = [_c_i32, _ptr, _c_i64, _c_i64, _ptr]
    L.op_sobol(n, _p(dv), dv.shape[0], dv.shape[1], _p(out))
    return out


def op_ugaussian(p: float) -> float:
    """ugaussianEl (:121-136) in f32."""
    L = lib()
    L.op_ugaussian.argtypes = [ctypes.c_float]
    L.op_ugaussian.restype = ctypes.c_float
    return float(L.op_ugaussian(ctypes.c_float(p)))


def op_bridge(bb_inds: np.ndarray, bb_data: np.ndarray, gauss: np.ndarray) -> np.ndarray:
    """brownianBridgeDates (:145-183) of one row of num_dates Gaussians."""
    bbi = np.ascontiguousarray(bb_inds, dtype=np.int32)
    bbd = np.ascontiguousarray(bb_data, dtype=np.float32)
    g = np.ascontiguousarray(gauss, dtype=np.float32)
    D = bbi.shape[1]
    if bbd.shape != (3, D) or g.shape != (D,):
        raise ValueError("bb_inds, bb_data [3][D] and gauss [D]")
    out = np.empty(D, np.float32)
    L = lib()
    L.op_bridge.argtypes = [_c_i64, _ptr, _ptr, _ptr, _ptr]
    L.op_bridge.restype = ctypes.c_int
    if L.op_bridge(D, _p(bbi), _p(bbd), _p(g), _p(out)) != 0:
        raise MemoryError("op_bridge")
    return out


def _threads(nthreads=None) -> int:
    """Worker threads for the oracle's OpenMP loops: the caller's count, else
    OMP_NUM_THREADS, else 16; never more than the host's CPUs (a shared GPU
    host's share is smaller than its core count)."""
    if nthreads:
        return int(nthreads)
    env = os.environ.get("OMP_NUM_THREADS", "")
    t = int(env) if env.isdigit() and int(env) > 0 else 16
    return max(1, min(t, os.cpu_count() or 1))


def op_main(contract, num_mc_it, dir_vs, md_cs, md_vols, md_drifts, md_sts, md_detvals,
            md_discts, bb_inds, bb_data, paths=None, nthreads=None) -> np.ndarray:
    """main (:295-322) over the first `paths` paths (default num_mc_it): the
    per-model mean payoff, the f64 sum of the f32 payoffs / paths."""
    n = int(num_mc_it if paths is None else paths)
    dv = np.ascontiguousarray(dir_vs, dtype=np.int32)
    arrs = [np.ascontiguousarray(a, dtype=np.float32)
            for a in (md_cs, md_vols, md_drifts, md_sts, md_detvals, md_discts, bb_data)]
    cs, vols, drifts, sts, detv, disc, bbd = arrs
    bbi = np.ascontiguousarray(bb_inds, dtype=np.int32)
    M, U = sts.shape
    D = bbi.shape[1]
    K = disc.shape[1]
    sums = np.zeros(M, np.float64)
    L = lib()
    L.op_sum_payoffs.argtypes = [_c_i32, _c_i64, _c_i64, _ptr, _c_i64] + [_ptr] * 6 + \
        [_c_i64, _ptr, _ptr, _c_i64, _c_i64, _c_i64, _ptr, ctypes.c_int]
    L.op_sum_payoffs.restype = ctypes.c_int
    rc = L.op_sum_payoffs(int(contract), 0, n, _p(dv), dv.shape[1], _p(cs), _p(vols),
                          _p(drifts), _p(sts), _p(detv), _p(disc), K, _p(bbi), _p(bbd),
                          D, U, M, _p(sums), _threads(nthreads))
    if rc != 0:
        raise ValueError(f"contract {contract} does not take [{D}][{U}] paths")
    return (sums / n).astype(np.float32)


def op_payoffs(contract, num_mc_it, dir_vs, md_cs, md_vols, md_drifts, md_sts, md_detvals,
               md_discts, bb_inds, bb_data, lo=0, hi=None, nthreads=None) -> np.ndarray:
    """The f32 payoffs of paths lo .. hi-1, [M][hi-lo]."""
    hi = int(num_mc_it if hi is None else hi)
    dv = np.ascontiguousarray(dir_vs, dtype=np.int32)
    cs, vols, drifts, sts, detv, disc, bbd = [
        np.ascontiguousarray(a, dtype=np.float32)
        for a in (md_cs, md_vols, md_drifts, md_sts, md_detvals, md_discts, bb_data)]
    bbi = np.ascontiguousarray(bb_inds, dtype=np.int32)
    M, U = sts.shape
    D, K = bbi.shape[1], disc.shape[1]
    out = np.zeros((M, max(hi - lo, 0)), np.float32)
    L = lib()
    L.op_payoffs.argtypes = [_c_i32, _c_i64, _c_i64, _ptr, _c_i64] + [_ptr] * 6 + \
        [_c_i64, _ptr, _ptr, _c_i64, _c_i64, _c_i64, _ptr, ctypes.c_int]
    L.op_payoffs.restype = ctypes.c_int
    if L.op_payoffs(int(contract), lo, hi, _p(dv), dv.shape[1], _p(cs), _p(vols), _p(drifts),
                    _p(sts), _p(detv), _p(disc), K, _p(bbi), _p(bbd), D, U, M, _p(out),
                    _threads(nthreads)) != 0:
        raise ValueError(f"contract {contract} does not take [{D}][{U}] paths")
    return out


# ---- InterestCalib (finpar/InterestCalib.fut), oracle/ic_oracle.c ----------
# The G2++ swaption pricer, the Cauchy log-likelihood and the genetic / DE
# MCMC calibration loop in f32 in the reference's order of operations; the
# calendar of lib/github.com/diku-dk/date.  Pinned in tests/test_ic_oracle.py.

def _ic_sig(name, res, args):
    f = getattr(lib(), name)
    f.restype = res
    f.argtypes = args
    return f


def ic_set_mutation(m: int) -> None:
    """Plant mistake m (0 = none) in the InterestCalib oracle: 1 the Gray
    code dropped from sobolInd, 2 erf's a3 coefficient negated, 3 the sign of
    the cross term t3 of bigv, 4 rhoxy without its b_fun(a+b) factor, 5 the
    last coupon without its notional, 6 add_months without the end-of-month
    clamp."""
    _ic_sig("ic_oracle_set_mutation", None, [ctypes.c_int])(int(m))


class ic_mutation:
    """Context manager planting an InterestCalib oracle mistake."""

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        ic_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        ic_set_mutation(0)


def ic_date_of_triple(y: int, m: int, d: int) -> int:
    return int(_ic_sig("ic_date_of_triple", _c_i32, [_c_i32] * 3)(y, m, d))


def ic_triple_of_date(t: int) -> tuple:
    y, m, d = _c_i32(), _c_i32(), _c_i32()
    _ic_sig("ic_triple_of_date", None, [_c_i32] + [ctypes.POINTER(_c_i32)] * 3)(
        t, ctypes.byref(y), ctypes.byref(m), ctypes.byref(d))
    return (y.value, m.value, d.value)


def ic_check_date(y: int, m: int, d: int) -> bool:
    return bool(_ic_sig("ic_check_date", ctypes.c_int, [_c_i32] * 3)(y, m, d))


def ic_add_months(t: int, months: int) -> int:
    return int(_ic_sig("ic_add_months", _c_i32, [_c_i32, _c_i32])(t, months))


def ic_diff_dates(t1: int, t2: int) -> float:
    return float(_ic_sig("ic_diff_dates", ctypes.c_double, [_c_i32, _c_i32])(t1, t2))


def ic_ugaussian_P(x: float) -> float:
    f = _ic_sig("ic_ugaussian_P", ctypes.c_float, [ctypes.c_float])
    return float(f(ctypes.c_float(x)))


def ic_ugaussian_P_exp(x: float, exp_factor: float) -> float:
    f = _ic_sig("ic_ugaussian_P_exp", ctypes.c_float, [ctypes.c_float, ctypes.c_float])
    return float(f(ctypes.c_float(x), ctypes.c_float(exp_factor)))


def ic_sobol(n: int, dir_vct: np.ndarray) -> float:
    """sobolInd (:190-202) of index n over the direction vector."""
    dv = np.ascontiguousarray(dir_vct, dtype=np.int32)
    f = _ic_sig("ic_sobol_ind", ctypes.c_float, [ctypes.c_int, _ptr, _c_i32])
    return float(f(len(dv), _p(dv), int(n)))


def ic_brent(scales, bbi, lb: float, ub: float, tol=1.0e-4, iter_max=1000) -> tuple:
    """rootFinding_Brent (:124-185) on sum(scales * exp(-bbi * y)) - 1:
    (root, error); error is +-inf where [lb, ub] does not bracket a root."""
    sc = np.ascontiguousarray(scales, np.float32)
    bb = np.ascontiguousarray(bbi, np.float32)
    err = ctypes.c_float()
    f = _ic_sig("ic_brent", ctypes.c_float, [ctypes.c_int, _ptr, _ptr, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_float, _c_i32,
                                             ctypes.POINTER(ctypes.c_float)])
    r = f(len(sc), _p(sc), _p(bb), lb, ub, tol, iter_max, ctypes.byref(err))
    return float(r), float(err.value)


def ic_eval(genome, swaptions, herm_x, herm_w) -> tuple:
    """evalGenomeOnSwap (:484-602) of one genome (a, b, rho, nu, sigma) on
    every swaption [n][4]: (quotes, prices), float32 [n] each."""
    g = np.ascontiguousarray(genome, np.float32)
    sw = np.ascontiguousarray(swaptions, np.float32).reshape(-1, 4)
    hx = np.ascontiguousarray(herm_x, np.float32)
    hw = np.ascontiguousarray(herm_w, np.float32)
    f = _ic_sig("ic_eval_genome_on_swap", None,
                [_ptr, ctypes.c_int, _ptr, _ptr, _ptr, ctypes.POINTER(ctypes.c_float),
                 ctypes.POINTER(ctypes.c_float)])
    q, p = ctypes.c_float(), ctypes.c_float()
    quotes = np.empty(len(sw), np.float32)
    prices = np.empty(len(sw), np.float32)
    for s in range(len(sw)):
        row = np.ascontiguousarray(sw[s])
        f(_p(g), len(hx), _p(hx), _p(hw), _p(row), ctypes.byref(q), ctypes.byref(p))
        quotes[s], prices[s] = q.value, p.value
    return quotes, prices


def ic_log_likelihood(y_ref: float, y: float) -> float:
    f = _ic_sig("ic_log_likelihood", ctypes.c_float, [ctypes.c_float, ctypes.c_float])
    return float(f(y_ref, y))


def ic_logliks(genomes, swaptions, herm_x, herm_w, nthreads=None) -> np.ndarray:
    """The log-likelihood (:639-645, a sequential f32 sum over the swaptions)
    of each genome [n][5]."""
    g = np.ascontiguousarray(genomes, np.float32).reshape(-1, 5)
    sw = np.ascontiguousarray(swaptions, np.float32).reshape(-1, 4)
    hx = np.ascontiguousarray(herm_x, np.float32)
    hw = np.ascontiguousarray(herm_w, np.float32)
    out = np.empty(len(g), np.float32)
    _ic_sig("ic_logliks", None, [ctypes.c_int, _ptr, ctypes.c_int, _ptr, ctypes.c_int, _ptr,
                                 _ptr, _ptr, ctypes.c_int])(
        len(g), _p(g), len(sw), _p(sw), len(hx), _p(hx), _p(hw), _p(out),
        _threads(nthreads))
    return out


def ic_propose(sob_offs: int, genomes, props, dir_vct) -> tuple:
    """One iteration's proposals (:655-724) from the Sobol offset at the top
    of the iteration: (new proposals [pop][5], fb ratios [pop], the Sobol
    offset of the acceptance draws)."""
    g = np.ascontiguousarray(genomes, np.float32).reshape(-1, 5)
    pr = np.array(props, np.float32).reshape(-1, 5)
    dv = np.ascontiguousarray(dir_vct, np.int32)
    fb = np.empty(len(g), np.float32)
    acc = _ic_sig("ic_propose", _c_i32, [_c_i32, _c_i32, _ptr, _c_i32, _ptr, _ptr, _ptr])(
        len(g), len(dv), _p(dv), int(sob_offs), _p(g), _p(pr), _p(fb))
    return pr, fb, int(acc)


def ic_accept(acc_offs: int, genomes, logliks, props, newl, fb, dir_vct) -> tuple:
    """The accept/reject step (:733-745): (genomes, logliks) after it."""
    g = np.array(genomes, np.float32).reshape(-1, 5)
    ll = np.array(logliks, np.float32)
    pr = np.ascontiguousarray(props, np.float32).reshape(-1, 5)
    nl = np.ascontiguousarray(newl, np.float32)
    f_ = np.ascontiguousarray(fb, np.float32)
    dv = np.ascontiguousarray(dir_vct, np.int32)
    _ic_sig("ic_accept", None, [_c_i32, _c_i32, _ptr, _c_i32] + [_ptr] * 5)(
        len(g), len(dv), _p(dv), int(acc_offs), _p(g), _p(ll), _p(pr), _p(nl), _p(f_))
    return g, ll


def ic_schedule_has_de(pop: int, mcmc_conv: int, dir_vct) -> bool:
    """Whether a DE move comes up in the first mcmc_conv iterations."""
    dv = np.ascontiguousarray(dir_vct, np.int32)
    return bool(_ic_sig("ic_schedule_has_de", ctypes.c_int, [_c_i32, _c_i32, _c_i32, _ptr])(
        int(pop), int(mcmc_conv), len(dv), _p(dv)))


def ic_main(pop, mcmc_conv, swaptions, herm_x, herm_w, dir_vct, nthreads=None) -> tuple:
    """main (:929-936): (six f32 scalars a, b, sigma, nu, rho, log-likelihood
    of the winner, summary [n][3], final genomes [pop][5], final
    log-likelihoods [pop]).  Raises ValueError where the reference fails."""
    sw = np.ascontiguousarray(swaptions, np.float32).reshape(-1, 4)
    hx = np.ascontiguousarray(herm_x, np.float32)
    hw = np.ascontiguousarray(herm_w, np.float32)
    dv = np.ascontiguousarray(dir_vct, np.int32)
    out6 = np.empty(6, np.float32)
    summ = np.empty((len(sw), 3), np.float32)
    fg = np.empty((max(int(pop), 0), 5), np.float32)
    fl = np.empty(max(int(pop), 0), np.float32)
    rc = _ic_sig("ic_main", ctypes.c_int,
                 [_c_i32, _c_i32, _c_i32, _ptr, _c_i32, _ptr, _ptr, _c_i32, _ptr, _ptr, _ptr,
                  _ptr, _ptr, ctypes.c_int])(
        int(pop), int(mcmc_conv), len(sw), _p(sw), len(hx), _p(hx), _p(hw), len(dv), _p(dv),
        _p(out6), _p(summ), _p(fg), _p(fl), _threads(nthreads))
    if rc != 0:
        raise ValueError("InterestCalib main: the reference fails on these arguments")
    return out6, summ, fg, fl
