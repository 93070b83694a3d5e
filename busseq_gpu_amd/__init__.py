"""MI355X-native BUSseq Gibbs sampler: Python mirror of the libbusseq C ABI.

The product is the native library (busseq_gpu_amd/lib/libbusseq.so, built for
gfx950) and the drop-in executable busseq_gpu_amd/bin/BUSseq_gpu.  This module
is a thin ctypes layer over include/busseq.h for tests, benchmarks and Python
users; it never computes anything itself and raises if the library is missing.

Reference behaviour mirrored: the BUSseq_gpu executable (BUSseq_gpu.cu:1026-2021)
with its CLI arguments -B -N -G -K -s -c -i -b -u -p -o.
"""
import ctypes as _C
import os as _os

import numpy as _np

__all__ = ["Sampler", "simulate", "run", "read_counts", "write_counts_binary", "read_dims", "kernel_names", "copy_probe",
           "autotune", "autotune_batch", "library_path", "BusseqError", "STATUS"]

_HERE = _os.path.dirname(_os.path.abspath(__file__))
_LIB_PATH = _os.path.join(_HERE, "lib", "libbusseq.so")
STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_IO", 3: "ERR_FORMAT", 4: "ERR_DATA", 5: "ERR_DEVICE", 6: "ERR_NOMEM"}


class BusseqError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


def library_path():
    return _LIB_PATH


class _Config(_C.Structure):
    _fields_ = [("B", _C.c_int), ("n_b", _C.POINTER(_C.c_int)), ("G", _C.c_int), ("K", _C.c_int),
                ("seed", _C.c_uint32), ("n_iter", _C.c_int), ("n_burnin", _C.c_int),
                ("n_unchanged", _C.c_int), ("keep_traces", _C.c_int), ("precision", _C.c_int),
                ("chains", _C.c_int), ("chain0", _C.c_int), ("tile_rows", _C.c_int), ("device", _C.c_int)]


_PI = _C.POINTER(_C.c_int32)
_PD = _C.POINTER(_C.c_double)


class _Problem(_C.Structure):
    _fields_ = [("cfg", _C.POINTER(_Config)), ("Y", _PI)]


class _State(_C.Structure):
    _fields_ = [("X", _PI), ("Z", _PI), ("W", _PI), ("alpha", _PD), ("beta", _PD), ("L", _PI),
                ("nu", _PD), ("delta", _PD), ("phi", _PD), ("gamma", _PD), ("pi", _PD), ("ptau", _PD)]


class _Result(_C.Structure):
    _fields_ = [("B", _C.c_int), ("N", _C.c_int), ("G", _C.c_int), ("K", _C.c_int), ("n_pres", _C.c_int),
                ("n_b", _C.POINTER(_C.c_int)), ("X", _PI), ("Z", _PI), ("W", _PI),
                ("gamma", _PD), ("alpha", _PD), ("L", _PD), ("beta", _PD), ("ptau", _PD),
                ("nu", _PD), ("delta", _PD), ("phi", _PD), ("pi", _PD), ("bic", _C.c_double),
                ("tr_gamma", _PD), ("tr_alpha", _PD), ("tr_L", _PI), ("tr_beta", _PD), ("tr_ptau", _PD),
                ("tr_nu", _PD), ("tr_delta", _PD), ("tr_phi", _PD), ("tr_W", _PI), ("tr_pi", _PD)]


_lib = None
BIC_KEYS = ("alpha", "beta", "nu", "delta", "phi", "gamma", "pi")


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not _os.path.exists(_LIB_PATH):
        raise BusseqError(5, f"{_LIB_PATH} is not built; run __graft_entry__.build() "
                             "(python -m busseq_gpu_amd.build)")
    lib = _C.CDLL(_LIB_PATH)
    vp = _C.c_void_p
    lib.busseq_last_error.restype = _C.c_char_p
    lib.busseq_config_default.argtypes = [_C.POINTER(_Config)]
    lib.busseq_default_tile_rows.argtypes = [_C.c_longlong]
    lib.busseq_default_tile_rows.restype = _C.c_int
    lib.busseq_create.argtypes = [_C.POINTER(_Config), _PI, _C.POINTER(vp)]
    lib.busseq_sweeps.argtypes = [vp, _C.c_int]
    lib.busseq_time_sweeps.argtypes = [vp, _C.c_int, _C.POINTER(_C.c_double)]
    lib.busseq_iteration.argtypes = [vp]
    lib.busseq_get_state.argtypes = [vp, _C.c_int, _C.POINTER(_State)]
    lib.busseq_set_state.argtypes = [vp, _C.c_int, _C.POINTER(_State)]
    lib.busseq_finish.argtypes = [vp, _C.c_int, _C.POINTER(_Result)]
    lib.busseq_bic.argtypes = [vp, _C.c_int] + [_PD] * 8
    lib.busseq_bic_noalloc.argtypes = [vp, _C.c_int] + [_PD] * 8
    lib.busseq_destroy.argtypes = [vp]
    lib.busseq_destroy.restype = None
    lib.busseq_run.argtypes = [_C.POINTER(_Config), _PI, _C.POINTER(_Result)]
    lib.busseq_write_outputs.argtypes = [_C.POINTER(_Result), _C.c_char_p, _C.c_int]
    lib.busseq_result_free.argtypes = [_C.POINTER(_Result)]
    lib.busseq_result_free.restype = None
    lib.busseq_read_dims.argtypes = [_C.c_char_p, _C.c_int, _C.POINTER(_C.c_int)]
    lib.busseq_read_counts.argtypes = [_C.c_char_p, _C.c_int, _C.c_int, _PI]
    lib.busseq_write_counts_binary.argtypes = [_C.c_char_p, _C.c_int, _C.c_int, _PI]
    lib.busseq_create_batch.argtypes = [_C.POINTER(_Problem), _C.c_int, _C.POINTER(vp)]
    lib.busseq_chain_count.argtypes = [vp]
    lib.busseq_kernel_count.restype = _C.c_int
    lib.busseq_kernel_name.argtypes = [_C.c_int]
    lib.busseq_kernel_name.restype = _C.c_char_p
    lib.busseq_kernel_times.argtypes = [vp, _C.c_int, _PD]
    lib.busseq_copy_probe.argtypes = [_C.c_size_t, _C.c_int, _C.POINTER(_C.c_double)]
    lib.busseq_rcp_check.argtypes = [_C.c_int, _C.POINTER(_C.c_uint64)]
    lib.busseq_autotune_batch.argtypes = [_C.POINTER(_Problem), _C.c_int, _C.c_int, _C.POINTER(_C.c_int),
                                          _C.POINTER(_C.c_double)]
    lib.busseq_simulate.argtypes = [_C.c_int, _C.POINTER(_C.c_int), _C.c_int, _C.c_int, _C.c_uint32,
                                    _C.c_double, _C.c_double, _PI, _PI, _PI]
    _lib = lib
    return lib


def _check(st):
    if st != 0:
        raise BusseqError(st, _load().busseq_last_error().decode())


def _ip(a):
    return a.ctypes.data_as(_PI) if a is not None else _PI()


def _dp(a):
    return a.ctypes.data_as(_PD) if a is not None else _PD()


def _config(n_b, G, K, seed, n_iter, n_burnin, n_unchanged, keep_traces=False, precision=32,
            chains=1, chain0=0, tile_rows=0, device=0):
    lib = _load()
    cfg = _Config()
    lib.busseq_config_default(_C.byref(cfg))
    nb = (_C.c_int * len(n_b))(*[int(v) for v in n_b])
    cfg.B = len(n_b)
    cfg.n_b = _C.cast(nb, _C.POINTER(_C.c_int))
    cfg.G, cfg.K, cfg.seed = int(G), int(K), int(seed) & 0xFFFFFFFF
    cfg.n_iter, cfg.n_burnin, cfg.n_unchanged = int(n_iter), int(n_burnin), int(n_unchanged)
    cfg.keep_traces, cfg.precision = int(bool(keep_traces)), int(precision)
    cfg.chains, cfg.chain0, cfg.tile_rows, cfg.device = int(chains), int(chain0), int(tile_rows), int(device)
    return cfg, nb


def default_tile_rows(cells):
    """Cells per reduction tile that a sampler with tile_rows=0 uses, for the
    number of cells over all its chains (busseq_default_tile_rows)."""
    return int(_load().busseq_default_tile_rows(int(cells)))


def simulate(n_b, G, K, data_seed, gamma0_range=(-2.0, 0.0), want_truth=False):
    """Counts from the BUSseq generative model (busseq_simulate); returns Y (N x G int32)
    and, with want_truth, (Y, W_true, X_true)."""
    lib = _load()
    N = int(sum(n_b))
    Y = _np.zeros((N, G), _np.int32)
    W = _np.zeros(N, _np.int32) if want_truth else None
    X = _np.zeros((N, G), _np.int32) if want_truth else None
    nb = (_C.c_int * len(n_b))(*[int(v) for v in n_b])
    _check(lib.busseq_simulate(len(n_b), nb, int(G), int(K), int(data_seed) & 0xFFFFFFFF,
                               float(gamma0_range[0]), float(gamma0_range[1]), _ip(Y), _ip(W), _ip(X)))
    return (Y, W, X) if want_truth else Y


def kernel_names():
    lib = _load()
    return [lib.busseq_kernel_name(j).decode() for j in range(lib.busseq_kernel_count())]


def copy_probe(nbytes, device=0):
    """One streaming device copy of nbytes; returns milliseconds."""
    lib = _load()
    ms = _C.c_double()
    _check(lib.busseq_copy_probe(int(nbytes), int(device), _C.byref(ms)))
    return ms.value


def rcp_check(device=0):
    """Number of normal floats z in [2^-126, 2^126) whose reciprocal on the
    device differs from the host's, the IEEE quotient 1/z of z held to
    [2^-126, 2^125]; single precision needs 0."""
    lib = _load()
    n = _C.c_uint64()
    _check(lib.busseq_rcp_check(int(device), _C.byref(n)))
    return n.value


def read_dims(path, B):
    lib = _load()
    nb = (_C.c_int * B)()
    _check(lib.busseq_read_dims(path.encode(), B, nb))
    return list(nb)


def read_counts(path, N, G):
    lib = _load()
    Y = _np.zeros((N, G), _np.int32)
    _check(lib.busseq_read_counts(path.encode(), N, G, _ip(Y)))
    return Y


def write_counts_binary(path, Y):
    """Binary form of a count matrix that read_counts (and the executable's -c)
    reads without parsing text."""
    lib = _load()
    Y = _np.ascontiguousarray(Y, dtype=_np.int32)
    _check(lib.busseq_write_counts_binary(path.encode(), Y.shape[0], Y.shape[1], _ip(Y)))


def _result_to_dict(r):
    B, N, G, K, P = r.B, r.N, r.G, r.K, r.n_pres

    def a(ptr, n, dt):
        return _np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt).copy() if ptr else None

    out = {
        "X": a(r.X, N * G, _np.int32).reshape(N, G), "Z": a(r.Z, N * G, _np.int32).reshape(N, G),
        "W": a(r.W, N, _np.int32), "gamma": a(r.gamma, 2 * B, _np.float64).reshape(2, B),
        "alpha": a(r.alpha, G, _np.float64), "L": a(r.L, K * G, _np.float64).reshape(K, G),
        "beta": a(r.beta, K * G, _np.float64).reshape(K, G), "ptau": a(r.ptau, 2, _np.float64),
        "nu": a(r.nu, B * G, _np.float64).reshape(B, G), "delta": a(r.delta, N, _np.float64),
        "phi": a(r.phi, B * G, _np.float64).reshape(B, G), "pi": a(r.pi, B * K, _np.float64).reshape(B, K),
        "bic": float(r.bic), "n_pres": P,
    }
    if P and r.tr_alpha:
        out["traces"] = {
            "gamma": a(r.tr_gamma, 2 * B * P, _np.float64).reshape(2 * B, P),
            "alpha": a(r.tr_alpha, G * P, _np.float64).reshape(G, P),
            "L": a(r.tr_L, (K - 1) * G * P, _np.int32).reshape((K - 1) * G, P),
            "beta": a(r.tr_beta, (K - 1) * G * P, _np.float64).reshape((K - 1) * G, P),
            "ptau": a(r.tr_ptau, 2 * P, _np.float64).reshape(2, P),
            "nu": a(r.tr_nu, (B - 1) * G * P, _np.float64).reshape((B - 1) * G, P),
            "delta": a(r.tr_delta, N * P, _np.float64).reshape(N, P),
            "phi": a(r.tr_phi, B * G * P, _np.float64).reshape(B * G, P),
            "W": a(r.tr_W, N * P, _np.int32).reshape(N, P),
            "pi": a(r.tr_pi, B * K * P, _np.float64).reshape(B * K, P),
        }
    return out


class Sampler:
    """One libbusseq sampler handle: one problem with one or more chains, or
    (Sampler.batch) several problems advanced together in the same launches."""

    def __init__(self, Y, n_b, K, seed=13579, n_iter=4000, n_burnin=2000, n_unchanged=500,
                 keep_traces=False, precision=32, chains=1, chain0=0, tile_rows=0, device=0):
        self._init([dict(Y=Y, n_b=n_b, K=K, seed=seed, n_iter=n_iter, n_burnin=n_burnin,
                         n_unchanged=n_unchanged, keep_traces=keep_traces, chains=chains, chain0=chain0)],
                   precision, tile_rows, device)

    @classmethod
    def batch(cls, problems, precision=32, tile_rows=0, device=0):
        """problems: list of dicts with Y, n_b, K and optional seed, chains, chain0,
        n_iter, n_burnin, n_unchanged, keep_traces."""
        obj = cls.__new__(cls)
        obj._init(problems, precision, tile_rows, device)
        return obj

    def _init(self, problems, precision, tile_rows, device):
        lib = _load()
        self._lib, self._h = lib, None
        probs, keep, dims = _problem_array(problems, precision, tile_rows, device)
        h = _C.c_void_p()
        _check(lib.busseq_create_batch(probs, len(problems), _C.byref(h)))
        self._h, self._keep, self._dims = h, keep, dims
        self.N, self.G, self.K, self.B, self.n_b = dims[0]
        self.chains = len(dims)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.busseq_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def iteration(self):
        return self._lib.busseq_iteration(self._h)

    def sweeps(self, n):
        _check(self._lib.busseq_sweeps(self._h, int(n)))

    def time_sweeps(self, n):
        ms = _C.c_double()
        _check(self._lib.busseq_time_sweeps(self._h, int(n), _C.byref(ms)))
        return ms.value

    def kernel_times(self, n):
        """Average milliseconds per kernel over n eagerly launched sweeps."""
        names = kernel_names()
        ms = _np.zeros(len(names))
        _check(self._lib.busseq_kernel_times(self._h, int(n), _dp(ms)))
        return dict(zip(names, ms.tolist()))

    def state(self, chain=0):
        N, G, K, B, _ = self._dims[chain]
        s = {"X": _np.zeros((N, G), _np.int32), "Z": _np.zeros((N, G), _np.int32),
             "W": _np.zeros(N, _np.int32), "alpha": _np.zeros(G), "beta": _np.zeros((K, G)),
             "L": _np.zeros((K, G), _np.int32), "nu": _np.zeros((B, G)), "delta": _np.zeros(N),
             "phi": _np.zeros((B, G)), "gamma": _np.zeros((2, B)), "pi": _np.zeros((B, K)),
             "ptau": _np.zeros(2)}
        st = _State(_ip(s["X"]), _ip(s["Z"]), _ip(s["W"]), _dp(s["alpha"]), _dp(s["beta"]), _ip(s["L"]),
                    _dp(s["nu"]), _dp(s["delta"]), _dp(s["phi"]), _dp(s["gamma"]), _dp(s["pi"]), _dp(s["ptau"]))
        _check(self._lib.busseq_get_state(self._h, int(chain), _C.byref(st)))
        return s

    def set_state(self, s, chain=0):
        def gi(k):
            return _np.ascontiguousarray(s[k], _np.int32) if k in s else None

        def gd(k):
            return _np.ascontiguousarray(s[k], _np.float64) if k in s else None

        arrs = {k: (gi(k) if k in ("X", "Z", "W", "L") else gd(k)) for k in
                ("X", "Z", "W", "alpha", "beta", "L", "nu", "delta", "phi", "gamma", "pi", "ptau")}
        st = _State(*[(_ip(v) if k in ("X", "Z", "W", "L") else _dp(v)) for k, v in arrs.items()])
        _check(self._lib.busseq_set_state(self._h, int(chain), _C.byref(st)))

    def bic(self, params, chain=0, noalloc=False):
        """BIC of a chain's counts at parameter values params (a dict with alpha,
        beta, nu, delta, phi, gamma, pi in the shapes finish() returns); with
        noalloc, through busseq_bic_noalloc (no device allocation)."""
        arrs = [_np.ascontiguousarray(params[k], _np.float64) for k in BIC_KEYS]
        out = _C.c_double()
        fn = self._lib.busseq_bic_noalloc if noalloc else self._lib.busseq_bic
        _check(fn(self._h, int(chain), *[_dp(a) for a in arrs], _C.byref(out)))
        return out.value

    def finish(self, chain=0):
        r = _Result()
        try:
            _check(self._lib.busseq_finish(self._h, int(chain), _C.byref(r)))
            return _result_to_dict(r)
        finally:
            self._lib.busseq_result_free(_C.byref(r))


def _problem_array(problems, precision, tile_rows, device):
    keep, probs, dims = [], (_Problem * len(problems))(), []
    for j, p in enumerate(problems):
        Y = _np.ascontiguousarray(p["Y"], dtype=_np.int32)
        N, G = Y.shape
        n_b = list(p["n_b"])
        if sum(n_b) != N:
            raise BusseqError(1, "sum(n_b) must equal the number of rows of Y")
        cfg, nb = _config(n_b, G, p["K"], p.get("seed", 13579), p.get("n_iter", 4000), p.get("n_burnin", 2000),
                          p.get("n_unchanged", 500), p.get("keep_traces", False), precision,
                          p.get("chains", 1), p.get("chain0", 0), tile_rows, device)
        keep += [Y, cfg, nb]
        probs[j].cfg = _C.pointer(cfg)
        probs[j].Y = _ip(Y)
        dims += [(N, G, int(p["K"]), len(n_b), n_b)] * int(p.get("chains", 1))
    return probs, keep, dims


def autotune_batch(problems, precision=32, sweeps=20, device=0):
    """Fastest tile size (cells per tile) of {16, 32, 64, 128} for these problems in
    one batched sampler (busseq_autotune_batch); returns (tile_rows, sweeps_per_s)."""
    lib = _load()
    probs, keep, _ = _problem_array(problems, precision, 0, device)
    tr, sps = _C.c_int(), _C.c_double()
    _check(lib.busseq_autotune_batch(probs, len(problems), int(sweeps), _C.byref(tr), _C.byref(sps)))
    return tr.value, sps.value


def autotune(Y, n_b, K, seed=13579, chains=1, precision=32, sweeps=20, device=0, n_burnin=2000,
             n_unchanged=500):
    """Fastest tile size for one problem (busseq_autotune); returns (tile_rows, sweeps_per_s).
    n_burnin and n_unchanged set which phase of a run the timed sweeps belong to."""
    return autotune_batch([dict(Y=Y, n_b=n_b, K=K, seed=seed, chains=chains, n_burnin=n_burnin,
                                n_unchanged=n_unchanged)], precision, sweeps, device)


def run(Y, n_b, K, seed=13579, n_iter=4000, n_burnin=2000, n_unchanged=500, keep_traces=False,
        precision=32, tile_rows=0, device=0, prefix=None):
    """Whole run (busseq_run); writes the reference's output files when prefix is given."""
    lib = _load()
    Y = _np.ascontiguousarray(Y, dtype=_np.int32)
    cfg, nb = _config(n_b, Y.shape[1], K, seed, n_iter, n_burnin, n_unchanged, keep_traces,
                      precision, 1, 0, tile_rows, device)
    r = _Result()
    try:
        _check(lib.busseq_run(_C.byref(cfg), _ip(Y), _C.byref(r)))
        if prefix is not None:
            _check(lib.busseq_write_outputs(_C.byref(r), prefix.encode(), int(bool(keep_traces))))
        return _result_to_dict(r)
    finally:
        lib.busseq_result_free(_C.byref(r))

