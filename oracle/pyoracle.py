"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (oracle/liboracle.so).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product.  See oracle/oracle.h.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_P = np.ctypeslib.ndpointer
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    vp = C.c_void_p
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [C.c_int, _P(np.int32), C.c_int, C.c_int, _P(np.int32), C.c_uint32,
                               C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_init.argtypes = [vp]
    lib.orc_sweeps.argtypes = [vp, C.c_int]
    lib.orc_iter.argtypes = [vp]
    lib.orc_set_tile_rows.argtypes = [vp, C.c_int]
    lib.orc_set_threads.argtypes = [C.c_int]
    lib.orc_get_state.argtypes = [vp, vp]
    lib.orc_set_state.argtypes = [vp, vp]
    lib.orc_get_priors.argtypes = [vp, _P(np.float64), _P(np.float64), _P(np.float64)]
    lib.orc_get_debug.argtypes = [vp, vp]
    lib.orc_posterior.argtypes = [vp, vp]
    lib.orc_bic.argtypes = [vp, vp]
    lib.orc_bic.restype = C.c_double
    lib.orc_philox.argtypes = [_P(np.uint32), _P(np.uint32), _P(np.uint32)]
    lib.orc_stream.argtypes = [C.c_uint32] * 6 + [C.c_int, C.c_double, C.c_double, C.c_int, _P(np.float64)]
    lib.orc_math.argtypes = [C.c_int, _P(np.float64), _P(np.float64), _P(np.float64), C.c_int]
    lib.orc_nb_inv.argtypes = [_P(np.float64)] * 4 + [C.c_int]
    lib.orc_terms.argtypes = [C.c_int, C.c_int, _P(np.float64), C.c_int, _P(np.float64), C.c_int]
    lib.orc_xz_chain.argtypes = [C.c_double] * 4 + [C.c_uint32, C.c_int, C.c_int, _P(np.int64)]
    lib.orc_keep_traces.argtypes = [vp, C.c_int]
    lib.orc_get_traces.argtypes = [vp, vp]
    lib.orc_step.argtypes = [vp, C.c_int, C.c_int]
    lib.orc_set_mutation.argtypes = [C.c_int]
    _lib = lib
    return lib


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data if a is not None else None for a in arrs])


STATE_KEYS = ("S", "W", "alpha", "beta", "L", "nu", "delta", "phi", "gamma", "pi", "ptau")
DEBUG_KEYS = ("alpha_prop", "alpha_sum", "beta_prop", "beta_sum", "nu_prop", "nu_sum", "delta_prop",
              "delta_sum", "phi_prop", "phi_sum", "gamma_prop", "gamma_sum", "w_logit")
POST_KEYS = ("gamma", "alpha", "L", "beta", "ptau", "nu", "delta", "phi", "pi", "W")

# names of orc_math functions
MATH = {"log": 0, "log1p": 1, "exp": 2, "expm1": 3, "lgamma": 4, "lgamma_diff64": 5, "logf": 6,
        "log1pf": 7, "lgamma_diff32": 8, "softplus": 9, "ndtri": 10, "fix": 11, "log1p_mid": 12,
        "log1pf_mid": 13}
STREAM = {"uniform": 0, "normal": 1, "gamma": 2, "poisson": 3, "nbinom": 4, "beta": 5, "u32": 6,
          "nbinom_xz": 7}


def math(fn, x, y=None):
    lib = load()
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, np.float64)
    out = np.zeros_like(x)
    lib.orc_math(MATH[fn], x, y, out, len(x))
    return out


def stream(kind, n, seed=1, chain=0, a=0, b=0, sweep=0, phase=10, p1=0.0, p2=0.0):
    lib = load()
    out = np.zeros(n)
    lib.orc_stream(seed, chain, a, b, sweep, phase, STREAM[kind], p1, p2, n, out)
    return out


def nb_inv(u, mu, phi):
    """bq_rnbinom_inv at the given uniforms"""
    lib = load()
    u, mu, phi = (np.ascontiguousarray(a, np.float64) for a in np.broadcast_arrays(u, mu, phi))
    out = np.zeros_like(u)
    lib.orc_nb_inv(u, mu, phi, out, len(u))
    return out


def philox(ctr, key):
    lib = load()
    out = np.zeros(4, np.uint32)
    lib.orc_philox(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def xz_chain(mu, phi, g0, g1, seed, n, nx):
    lib = load()
    out = np.zeros(nx * 2, np.int64)
    lib.orc_xz_chain(mu, phi, g0, g1, seed, n, nx, out)
    return out.reshape(nx, 2)


TRACE_KEYS = ("gamma", "alpha", "L", "beta", "ptau", "nu", "delta", "phi", "W", "pi")
STEP = {"pi": 0, "ptau": 1, "L": 2}


def set_mutation(m):
    """Plant mistake m in the oracle (oracle.h lists them; 0 removes it)."""
    load().orc_set_mutation(int(m))


TERMS = {"tmh": 0, "tw": 1, "tphi": 2, "ratio": 3, "lr_gamma1": 4, "lr_phi": 5, "logodds_L": 6,
         "lr_normal": 7, "bic_elem": 8, "lr_beta": 9}


def terms(fn, prec, args):
    lib = load()
    a = np.ascontiguousarray(args, np.float64)
    out = np.zeros(a.shape[0])
    lib.orc_terms(TERMS.get(fn, fn), prec, a.ravel(), a.shape[1], out, a.shape[0])
    return out


def default_tile_rows(cells):
    """The sampler's tile size for tile_rows = 0 (busseq_default_tile_rows in
    busseq_gpu_amd/csrc/engine.hip), for one chain of `cells` cells."""
    return 16 if cells < 2560 else 32 if cells < 65536 else 64


class Oracle:
    def __init__(self, Y, n_b, K, seed=13579, chain=0, precision=32, n_iter=4000, n_burnin=2000,
                 n_unchanged=500, tile_rows=None, threads=None, keep_traces=False):
        lib = load()
        if threads:
            lib.orc_set_threads(int(threads))
        Y = np.ascontiguousarray(Y, np.int32)
        self.N, self.G = Y.shape
        self.K, self.B = int(K), len(n_b)
        nb = np.asarray(n_b, np.int32)
        h = lib.orc_create(self.B, nb, self.G, self.K, Y, seed & 0xFFFFFFFF, chain, precision,
                           n_iter, n_burnin, n_unchanged)
        if not h:
            raise ValueError("oracle rejected the configuration")
        self._h, self._lib = h, lib
        lib.orc_set_tile_rows(h, int(tile_rows) if tile_rows else default_tile_rows(self.N))
        if lib.orc_init(h) != 0:
            raise ValueError("oracle init failed (zero library size)")
        self.n_pres = max(0, n_iter - n_burnin) if keep_traces else 0
        if keep_traces and lib.orc_keep_traces(h, 1) != 0:
            raise MemoryError("oracle traces")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def iteration(self):
        return self._lib.orc_iter(self._h)

    def sweeps(self, n):
        self._lib.orc_sweeps(self._h, int(n))

    def _state_arrays(self):
        N, G, K, B = self.N, self.G, self.K, self.B
        return {"S": np.zeros((N, G), np.int32), "W": np.zeros(N, np.int32), "alpha": np.zeros(G),
                "beta": np.zeros((K, G)), "L": np.zeros((K, G), np.int32), "nu": np.zeros((B, G)),
                "delta": np.zeros(N), "phi": np.zeros((B, G)), "gamma": np.zeros((2, B)),
                "pi": np.zeros((B, K)), "ptau": np.zeros(2)}

    def state(self):
        s = self._state_arrays()
        self._lib.orc_get_state(self._h, C.cast(_ptrs([s[k] for k in STATE_KEYS]), C.c_void_p))
        S = s.pop("S")
        s["X"] = np.where(S < 0, ~S, S).astype(np.int32)
        s["Z"] = (S < 0).astype(np.int32)
        s["S"] = S
        return s

    def set_state(self, st):
        s = self._state_arrays()
        cur = self.state()
        for k in STATE_KEYS:
            if k == "S":
                if "X" in st and "Z" in st:
                    S = cur["S"].copy()
                    dyn = S <= 0
                    X = np.asarray(st["X"], np.int32)
                    Z = np.asarray(st["Z"], np.int32)
                    S[dyn] = np.where(Z[dyn] != 0, ~X[dyn], 0)
                    s["S"][...] = S
                else:
                    s["S"] = None
            elif k in st:
                s[k][...] = st[k]
            else:
                s[k] = None
        self._lib.orc_set_state(self._h, C.cast(_ptrs([s[k] for k in STATE_KEYS]), C.c_void_p))

    def step(self, which, t):
        """One step ("pi", "ptau" or "L") of sweep index t on the current state."""
        self._lib.orc_step(self._h, STEP[which], int(t))

    def traces(self):
        """Recorded sweeps in the reference's layout: one row per parameter, n_pres columns."""
        N, G, K, B, P = self.N, self.G, self.K, self.B, self.n_pres
        t = {"gamma": np.zeros((2 * B, P)), "alpha": np.zeros((G, P)), "L": np.zeros(((K - 1) * G, P), np.int32),
             "beta": np.zeros(((K - 1) * G, P)), "ptau": np.zeros((2, P)), "nu": np.zeros(((B - 1) * G, P)),
             "delta": np.zeros((N, P)), "phi": np.zeros((B * G, P)), "W": np.zeros((N, P), np.int32),
             "pi": np.zeros((B * K, P))}
        self._lib.orc_get_traces(self._h, C.cast(_ptrs([t[k] for k in TRACE_KEYS]), C.c_void_p))
        return t

    def priors(self):
        ma, mn, md = np.zeros(self.G), np.zeros(self.B), np.zeros(self.N)
        self._lib.orc_get_priors(self._h, ma, mn, md)
        return {"mu_alpha": ma, "mu_nu": mn, "mu_delta": md}

    def debug(self):
        N, G, K, B = self.N, self.G, self.K, self.B
        d = {"alpha_prop": np.zeros(G), "alpha_sum": np.zeros(G), "beta_prop": np.zeros((K, G)),
             "beta_sum": np.zeros((K, G)), "nu_prop": np.zeros((B, G)), "nu_sum": np.zeros((B, G)),
             "delta_prop": np.zeros(N), "delta_sum": np.zeros(N), "phi_prop": np.zeros((B, G)),
             "phi_sum": np.zeros((B, G)), "gamma_prop": np.zeros((2, B)), "gamma_sum": np.zeros((2, B)),
             "w_logit": np.zeros((N, K))}
        self._lib.orc_get_debug(self._h, C.cast(_ptrs([d[k] for k in DEBUG_KEYS]), C.c_void_p))
        return d

    def bic(self, params):
        """orc_bic at the given parameters (alpha, beta, nu, delta, phi, gamma, pi)."""
        N, G, K, B = self.N, self.G, self.K, self.B
        shapes = {"gamma": (2, B), "alpha": (G,), "L": (K, G), "beta": (K, G), "ptau": (2,), "nu": (B, G),
                  "delta": (N,), "phi": (B, G), "pi": (B, K)}
        p = {k: (np.ascontiguousarray(params[k], np.float64).reshape(v) if k in params else np.zeros(v))
             for k, v in shapes.items()}
        p["W"] = np.zeros(N, np.int32)
        return self._lib.orc_bic(self._h, C.cast(_ptrs([p[k] for k in POST_KEYS]), C.c_void_p))

    def posterior(self):
        N, G, K, B = self.N, self.G, self.K, self.B
        p = {"gamma": np.zeros((2, B)), "alpha": np.zeros(G), "L": np.zeros((K, G)), "beta": np.zeros((K, G)),
             "ptau": np.zeros(2), "nu": np.zeros((B, G)), "delta": np.zeros(N), "phi": np.zeros((B, G)),
             "pi": np.zeros((B, K)), "W": np.zeros(N, np.int32)}
        arr = _ptrs([p[k] for k in POST_KEYS])
        self._lib.orc_posterior(self._h, C.cast(arr, C.c_void_p))
        p["bic"] = self._lib.orc_bic(self._h, C.cast(arr, C.c_void_p))
        return p
