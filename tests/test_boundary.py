"""Drop-in boundary: C ABI exports, input readers, output writers and the
BUSseq_gpu command line (BUSseq_gpu.cu:1026-1161, 1610-2018; README.md:17-49)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import busseq_gpu_amd as bq
from busseq_gpu_amd import _Result, _load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "busseq_gpu_amd", "bin", "BUSseq_gpu")
GOLD = os.path.join(ROOT, "tests", "golden")


def declared(header, prefix):
    txt = open(os.path.join(ROOT, header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", txt)))


def exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_symbol():
    syms = declared("include/busseq.h", "busseq_")
    assert len(syms) >= 20
    missing = set(syms) - exported(bq.library_path())
    assert not missing, missing
    _load()  # loads without a GPU


def test_oracle_exports_every_declared_symbol():
    syms = declared("oracle/oracle.h", "orc_")
    missing = set(syms) - exported(os.path.join(ROOT, "oracle", "liboracle.so"))
    assert not missing, missing


def test_default_tile_size_rule_of_library_and_oracle():
    # tile_rows = 0 picks the tile size from the number of cells, in the library
    # (busseq_default_tile_rows) and, for one chain, in the oracle the same way
    from oracle import pyoracle as po
    for cells in (1, 50, 1000, 2559, 2560, 20000, 65535, 65536, 150000, 10**7):
        assert bq.default_tile_rows(cells) == po.default_tile_rows(cells), cells
    assert [bq.default_tile_rows(n) for n in (2559, 2560, 65535, 65536)] == [16, 32, 32, 64]
    cfg = bq._Config()
    _load().busseq_config_default(C.byref(cfg))
    assert cfg.tile_rows == 0


def test_read_dims_reference_fixture():
    # count_data/demo_dim.txt of the reference (-N file)
    assert bq.read_dims(os.path.join(GOLD, "demo_dim.txt"), 4) == [300, 300, 200, 200]


def test_read_dims_errors(tmp_path):
    p = tmp_path / "d.txt"
    p.write_text("300 300")
    with pytest.raises(bq.BusseqError) as e:
        bq.read_dims(str(p), 4)
    assert e.value.status == 3
    p.write_text("300 x 1 2")
    with pytest.raises(bq.BusseqError) as e:
        bq.read_dims(str(p), 4)
    assert e.value.status == 3
    with pytest.raises(bq.BusseqError) as e:
        bq.read_dims(str(tmp_path / "missing.txt"), 4)
    assert e.value.status == 2


def test_read_counts_roundtrip_and_errors(tmp_path):
    rng = np.random.default_rng(0)
    Y = rng.integers(0, 50, size=(7, 5)).astype(np.int32)
    p = tmp_path / "c.txt"
    p.write_text("\n".join("\t".join(str(v) for v in row) for row in Y) + "\n")
    np.testing.assert_array_equal(bq.read_counts(str(p), 7, 5), Y)
    p.write_text(" ".join(str(v) for v in Y.ravel()))          # any whitespace layout
    np.testing.assert_array_equal(bq.read_counts(str(p), 7, 5), Y)
    p.write_text(" ".join(str(v) for v in Y.ravel()[:-1]))     # short
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), 7, 5)
    assert e.value.status == 3 and "found 34" in str(e.value)
    p.write_text("1 2 3.5 " + " ".join("1" for _ in range(40)))  # not an integer
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), 7, 5)
    assert e.value.status == 3 and "3.5" in str(e.value)
    p.write_text("-1 " + " ".join("1" for _ in range(40)))       # negative count
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), 7, 5)
    assert e.value.status == 3


def _text(Y, rng):
    # every token followed by a random run of mixed whitespace (CRLF, tabs, blanks)
    seps = np.array([" ", "  ", "\t", "\n", "\r\n", " \t "])
    pick = seps[rng.integers(0, len(seps), size=Y.size)]
    return "".join(f"{v}{s}" for v, s in zip(Y.ravel().tolist(), pick.tolist()))


def test_read_counts_parallel_text_matches_serial_semantics(tmp_path):
    # about 16 MB of text: several reader threads, token boundaries at random offsets
    rng = np.random.default_rng(4)
    N, G = 2500, 1200
    Y = rng.negative_binomial(2, 0.2, size=(N, G)).astype(np.int32)
    Y[rng.integers(0, N, 50), rng.integers(0, G, 50)] = 2147483647
    body = _text(Y, rng)
    p = tmp_path / "big.txt"
    p.write_text(body + " 7 8 junk")                            # extra tokens after N*G are ignored
    np.testing.assert_array_equal(bq.read_counts(str(p), N, G), Y)
    toks = body.split()
    late = N * G - 1000                                          # in the last thread's range
    bad = list(toks)
    bad[late] = "12x"
    bad[late - 900000] = "-4"                                    # a negative before the bad token
    p.write_text(" ".join(bad))
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), N, G)
    assert e.value.status == 3 and "'12x'" in str(e.value)      # a serial reader stops at the bad token
    bad[late] = "1"
    p.write_text(" ".join(bad))
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), N, G)
    j = late - 900000
    assert e.value.status == 3 and f"cell {j // G}, gene {j % G}" in str(e.value)
    p.write_text(" ".join(toks[:-3]))
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), N, G)
    assert f"found {N * G - 3}" in str(e.value)
    p.write_text(" ".join(toks + ["2147483648"]))               # overflow after N*G: ignored
    np.testing.assert_array_equal(bq.read_counts(str(p), N, G), Y)
    p.write_text(" ".join(toks[:-1] + ["2147483648"]))          # overflow within N*G
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), N, G)
    assert "2147483648" in str(e.value)


def test_binary_counts_roundtrip_and_errors(tmp_path):
    rng = np.random.default_rng(5)
    Y = rng.integers(0, 1000, size=(37, 11)).astype(np.int32)
    p = tmp_path / "c.bqc"
    bq.write_counts_binary(str(p), Y)
    raw = p.read_bytes()
    assert raw[:8] == b"BQCOUNT1" and len(raw) == 32 + 4 * Y.size
    assert np.frombuffer(raw[8:32], "<i8").tolist() == [37, 11, 0]
    np.testing.assert_array_equal(np.frombuffer(raw[32:], "<i4").reshape(37, 11), Y)
    np.testing.assert_array_equal(bq.read_counts(str(p), 37, 11), Y)
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), 11, 37)
    assert e.value.status == 3 and "37 x 11" in str(e.value)
    p.write_bytes(raw[:-4])
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), 37, 11)
    assert e.value.status == 3 and "size" in str(e.value)
    Y[20, 3] = -2
    bq.write_counts_binary(str(p), Y)
    with pytest.raises(bq.BusseqError) as e:
        bq.read_counts(str(p), 37, 11)
    assert "cell 20, gene 3" in str(e.value)
    with pytest.raises(bq.BusseqError) as e:
        bq.write_counts_binary(str(tmp_path / "no" / "dir.bqc"), Y)
    assert e.value.status == 2


def fmt(v):
    return "%f" % v


def make_result(B=2, N=3, G=4, K=3, P=2, seed=3):
    rng = np.random.default_rng(seed)
    keep = {}

    def arr(shape, dt):
        a = np.ascontiguousarray(rng.normal(size=shape) * 3 if dt == np.float64 else rng.integers(0, 9, size=shape),
                                 dtype=dt)
        keep[id(a)] = a
        return a

    d = {"X": arr((N, G), np.int32), "Z": arr((N, G), np.int32), "W": arr(N, np.int32),
         "gamma": arr((2, B), np.float64), "alpha": arr(G, np.float64), "L": arr((K, G), np.float64),
         "beta": arr((K, G), np.float64), "ptau": arr(2, np.float64), "nu": arr((B, G), np.float64),
         "delta": arr(N, np.float64), "phi": arr((B, G), np.float64), "pi": arr((B, K), np.float64),
         "tr_gamma": arr((2 * B, P), np.float64), "tr_alpha": arr((G, P), np.float64),
         "tr_L": arr(((K - 1) * G, P), np.int32), "tr_beta": arr(((K - 1) * G, P), np.float64),
         "tr_ptau": arr((2, P), np.float64), "tr_nu": arr(((B - 1) * G, P), np.float64),
         "tr_delta": arr((N, P), np.float64), "tr_phi": arr((B * G, P), np.float64),
         "tr_W": arr((N, P), np.int32), "tr_pi": arr((B * K, P), np.float64)}
    r = _Result()
    r.B, r.N, r.G, r.K, r.n_pres = B, N, G, K, P
    for k, a in d.items():
        ptr = a.ctypes.data_as(C.POINTER(C.c_int32 if a.dtype == np.int32 else C.c_double))
        setattr(r, k, ptr)
    r.bic = 12345.678901234
    return r, d


def mat(a, f=fmt):
    a = np.atleast_2d(a)
    return "".join("\t".join(f(v) for v in row) + "\n" for row in a)


def test_writers_match_reference_formats(tmp_path):
    lib = _load()
    r, d = make_result()
    prefix = str(tmp_path / "out")
    assert lib.busseq_write_outputs(C.byref(r), prefix.encode(), 1) == 0
    rd = lambda name: open(prefix + name).read()
    i = lambda v: "%d" % v
    B, N, G, K = 2, 3, 4, 3
    assert rd("_post_Z.txt") == mat(d["Z"], i)
    assert rd("_post_X.txt") == mat(d["X"], i)
    assert rd("_post_gamma.txt") == mat(d["gamma"])
    assert rd("_post_alpha.txt") == mat(d["alpha"][:, None])
    assert rd("_post_L.txt") == mat(d["L"][1:])
    assert rd("_post_beta.txt") == "\t".join(["0.000000"] * G) + "\n" + mat(d["beta"][1:])
    assert rd("_post_p.txt") == fmt(d["ptau"][0]) + "\t" + fmt(d["ptau"][1]) + "\t"   # no newline (cu:1725-1727)
    assert rd("_post_nu.txt") == "\t".join(["0.000000"] * G) + "\n" + mat(d["nu"][1:])
    assert rd("_post_delta.txt") == mat(d["delta"][:, None])
    assert rd("_post_phi.txt") == mat(d["phi"])
    assert rd("_post_W.txt") == mat(d["W"][:, None], i)
    assert rd("_post_pi.txt") == mat(d["pi"])
    assert rd("_BIC.txt") == "12345.678901\n"
    for name, key, f in (("gamma", "tr_gamma", fmt), ("alpha", "tr_alpha", fmt), ("L", "tr_L", i),
                         ("beta", "tr_beta", fmt), ("p", "tr_ptau", fmt), ("nu", "tr_nu", fmt),
                         ("delta", "tr_delta", fmt), ("phi", "tr_phi", fmt), ("W", "tr_W", i), ("pi", "tr_pi", fmt)):
        assert rd(f"_preserved_{name}.txt") == mat(d[key], f), name


def test_writer_reports_unwritable_prefix(tmp_path):
    lib = _load()
    r, d = make_result()
    assert lib.busseq_write_outputs(C.byref(r), str(tmp_path / "no" / "such" / "dir").encode(), 0) == 2


def cli(*args, cwd=None):
    return subprocess.run([EXE, *args], capture_output=True, text=True, cwd=cwd, timeout=120)


@pytest.mark.parametrize("args,msg", [
    (["-B", "1", "-s", "1", "-i", "10", "-b", "5"], "The batch number must be greater than one."),
    (["-B", "4", "-s", "1", "-i", "10", "-b", "20"], "Burn-in iterations must be less than total number of iterations."),
    # a burn-in as long as the run records nothing: rejected with the same message
    (["-B", "4", "-s", "1", "-i", "10", "-b", "10"], "Burn-in iterations must be less than total number of iterations."),
    (["-B", "4", "-K", "301", "-s", "1", "-i", "10", "-b", "5"],
     "The sample size in any batch must be greater than the assumed cell type number."),
])
def test_cli_validation_messages(args, msg):
    p = cli(*args)
    assert p.returncode == 255, p.stdout + p.stderr
    assert msg in p.stdout
    assert "Seed is 1" in p.stdout


def test_cli_rejects_more_batches_than_the_reference_holds():
    p = cli("-B", "129", "-s", "1", "-i", "10", "-b", "5")
    assert p.returncode == 1 and "-B must be at most 128" in p.stdout


def test_library_rejects_a_burn_in_as_long_as_the_run():
    # validation precedes any device call, so this runs without a GPU
    Y = np.ones((6, 4), np.int32)
    with pytest.raises(bq.BusseqError) as e:
        bq.Sampler(Y, [3, 3], 2, n_iter=8, n_burnin=8, n_unchanged=2)
    assert e.value.status == 1 and "Burn-in iterations must be less" in str(e.value)


def test_library_limits_on_k_and_b_times_k():
    # validation precedes any device call, so these run without a GPU
    Y = np.ones((140, 4), np.int32)
    with pytest.raises(bq.BusseqError) as e:
        bq.Sampler(Y, [70, 70], 65, n_iter=8, n_burnin=4, n_unchanged=2)
    assert e.value.status == 1 and "K must be between 2 and 64" in str(e.value)
    Y = np.ones((33 * 64, 4), np.int32)
    with pytest.raises(bq.BusseqError) as e:
        bq.Sampler(Y, [64] * 33, 64, n_iter=8, n_burnin=4, n_unchanged=2)
    assert e.value.status == 1 and "B * K must be at most 2048" in str(e.value)


def test_cli_unknown_option_prints_usage():
    p = cli("-z")
    assert p.returncode == 1
    assert "Error Usage:" in p.stdout


def test_cli_missing_count_file(tmp_path):
    p = cli("-B", "4", "-N", os.path.join(GOLD, "demo_dim.txt"), "-G", "10", "-K", "3", "-s", "3",
            "-c", str(tmp_path / "nope.txt"), "-i", "10", "-b", "5", "-o", str(tmp_path / "o"))
    assert p.returncode == 1
    assert "Start reading file." in p.stdout and "cannot open" in p.stdout
    assert not list(tmp_path.glob("o_*"))


def test_cli_derived_defaults(tmp_path):
    # without -b, burn-in = n_iter/2 (cu:1101-1103): -i 10 gives burn-in 5 < 10, valid until the count file
    p = cli("-B", "2", "-s", "4", "-i", "10", "-c", str(tmp_path / "none.txt"))
    assert "Burn-in" not in p.stdout and p.returncode == 1


def test_cli_count_cache(tmp_path):
    # the cache is written after the text is read, read back while it is current and
    # ignored once the text is newer; the run itself stops at initialization without a GPU
    rng = np.random.default_rng(6)
    n_b = [6, 4]
    Y = rng.integers(1, 30, size=(10, 8)).astype(np.int32)
    dims = tmp_path / "dim.txt"
    dims.write_text("6 4\n")
    txt = tmp_path / "c.txt"
    txt.write_text("\n".join(" ".join(map(str, r)) for r in Y) + "\n")
    cache = tmp_path / "c.bqc"
    args = ("-B", "2", "-N", str(dims), "-G", "8", "-K", "2", "-s", "1", "-c", str(txt), "-i", "4", "-b", "2",
            "-o", str(tmp_path / "o"), f"--count-cache={cache}")
    p = cli(*args)
    assert f"Counts cached in {cache}." in p.stdout
    np.testing.assert_array_equal(bq.read_counts(str(cache), 10, 8), Y)
    p = cli(*args)
    assert f"Counts read from the cache {cache}." in p.stdout
    os.utime(txt, ns=(os.stat(cache).st_mtime_ns + 10**9,) * 2)
    p = cli(*args)
    assert "read from the cache" not in p.stdout and "Counts cached in" in p.stdout
    # a binary count file is accepted as -c directly
    p = cli(*args[:11], str(cache), *args[12:-1])
    assert "Start initialization." in p.stdout


def test_f64_writer_equals_printf_rounding(tmp_path):
    # the writers' own "%lf" formatting against Python's correctly rounded "%f"
    # (round half to even on the exact binary value, as glibc's printf)
    rng = np.random.default_rng(7)
    mags = 10.0 ** rng.uniform(-12, 12.9, size=150000)
    vals = np.concatenate([
        mags * rng.choice([-1.0, 1.0], size=mags.size),
        np.arange(-4096, 4096) / 128.0, np.arange(-4096, 4096) / 1024.0,     # exact ties at the 6th decimal
        (np.arange(1, 2000) * 15625.0 + 0.5) / 1e6,                           # near ties
        [0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 0.0000005, -0.0000005, 0.0000015,
         8.999999999999e12, -8.999999999999e12, 9.0e12, 1.5e13, -3.7e19, 1e300, np.inf, -np.inf, np.nan,
         0.1, 0.7, 1.0000005, 123456.7890125, 2.5e-7, 7.5e-7]])
    N = vals.size
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int32 if dt == np.int32 else C.c_double))

    r = _Result()
    r.B, r.N, r.G, r.K, r.n_pres = 2, N, 1, 2, 0
    r.X = arr(np.zeros(N), np.int32); r.Z = arr(np.zeros(N), np.int32); r.W = arr(np.zeros(N), np.int32)
    r.delta = arr(vals, np.float64)
    for k, n in (("gamma", 4), ("alpha", 1), ("L", 2), ("beta", 2), ("ptau", 2), ("nu", 2), ("phi", 2), ("pi", 4)):
        setattr(r, k, arr(np.zeros(n), np.float64))
    prefix = str(tmp_path / "o")
    assert _load().busseq_write_outputs(C.byref(r), prefix.encode(), 0) == 0
    got = open(prefix + "_post_delta.txt").read().split("\n")[:-1]
    want = ["%f" % v for v in vals]
    bad = [(v, g, w) for v, g, w in zip(vals, got, want) if g != w]
    assert not bad, bad[:5]
