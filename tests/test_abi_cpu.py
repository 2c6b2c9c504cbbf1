"""CPU-side checks of the C-ABI library: it loads, exports every declared symbol, and its host-only parts (treedraws
text, tree parser, libm image tables, input validation) agree with the oracle. No GPU call is made here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, same_bits
from fastbart_amd import _lib, sampler
from fastbart_amd import workloads as W


def _declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            src = open(os.path.join(inc, f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", src, flags=re.M):
                if m.group(1) not in ("if", "while", "sizeof"):
                    names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared_functions()
    assert "bart_fit_cpu" in declared and "fbh_batch_create" in declared and "fbh_predict_draws" in declared
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert set(declared) == set(_lib.EXPORTS)
    assert L.fbh_version().decode().endswith("gfx950")


def test_dropin_signature_matches_reference_header():
    """bart_fit_cpu / bart_wfit / bart_set_rng keep the reference's parameter lists (src/bart_kernel.h:49-70)."""
    src = open(os.path.join(ROOT, "include", "fastbart_hip.h")).read()
    flat = " ".join(src.split())
    assert ("void bart_fit_cpu(const uint8_t* qin, const float* yin, const float* wtin, const int* numcut, "
            "const float* pgtab, float* sigma_out, float* trmean_out, uint64_t* trees_out, int n, int p, int m, "
            "int nskip, int ndpost, float tau, float lambda, float nu, float sigest, unsigned seed, int keep_trees, "
            "int* stats_out, int probit, const int* ybin, float offset);") in flat


def test_tree_text_matches_oracle(oracle, golden_cases):
    for name in ("w200_p3_extreme", "w4096_p16_short", "p300_probit"):
        g = golden_cases[name]
        trees = g["expected"]["trees"]
        p = g["problem"]["q"].shape[0]
        t1, v1 = sampler.tree_text(trees, p)
        t2, v2 = oracle.tree_text(trees, p)
        assert t1 == t2
        assert np.array_equal(v1, v2)


def test_tree_text_round_trip_reconstructs_trees(oracle, golden_cases):
    """Parse the text back (ids 2id / 2id+1) and rebuild every split and leaf value of the words."""
    g = golden_cases["w1000_p16_ones"]
    trees = g["expected"]["trees"]
    p = g["problem"]["q"].shape[0]
    text, vc = sampler.tree_text(trees, p)
    lines = text.split("\n")
    nd, m, pp = (int(x) for x in lines[0].split())
    assert (nd, m, pp) == (trees.shape[0], trees.shape[1], p)
    pos = 1
    counted = np.zeros((nd, p), np.int64)
    for dr in range(nd):
        for j in range(m):
            cnt = int(lines[pos]); pos += 1
            nodes = {}
            for _ in range(cnt):
                i, v, c, th = lines[pos].split(); pos += 1
                nodes[int(i)] = (int(v), int(c), float(th))
            T = trees[dr, j]

            def walk(s, nid):
                w = int(T[s])
                L, R = (w >> 16) & 255, (w >> 24) & 255
                v, c, th = nodes[nid]
                if L:
                    assert (v, c) == (w & 255, (w >> 8) & 255) and th == 0.0
                    counted[dr, v] += 1
                    walk(L, 2 * nid)
                    walk(R, 2 * nid + 1)
                else:
                    f = np.uint32(w >> 32).view(np.float32)
                    assert (v, c) == (0, 0)
                    assert float(np.float32(th)) == float(f)   # 10 significant digits identify the float
            walk(0, 1)
    assert np.array_equal(counted, vc)


def test_tree_text_number_format_is_ostream_precision_10(tmp_path):
    """%.10g is what an ostream with precision(10) writes for these values (src/glue.cpp:32,50)."""
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.standard_normal(3000).astype(np.float32).astype(np.float64) * 10.0 ** rng.integers(-8, 4, 3000),
                           [0.0, -0.0, 1.0, -1.5, 1e-30, 123456.789]])
    src = tmp_path / "fmt.cpp"
    src.write_text('#include <iostream>\n#include <sstream>\n#include <cstdio>\nint main(){double x; std::ostringstream o;'
                   ' o.precision(10); while (std::scanf("%la", &x) == 1) o << x << "\\n"; std::cout << o.str(); }\n')
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    r = subprocess.run([str(exe)], input="\n".join(float(v).hex() for v in vals), capture_output=True, text=True,
                       check=True)
    assert r.stdout.split("\n")[:-1] == ["%.10g" % v for v in vals]


def test_tree_text_leaf_values_are_printf_10g():
    """The writer's leaf values (std::to_chars, general format, precision 10) are printf's %.10g for every float:
    random bit patterns over all exponents, values of the samplers' range, signed zeros, subnormals, the extremes."""
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 2 ** 32, 60000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    extra = np.array([0.0, -0.0, 1.0, -1.5, 1e-45, -1e-45, 1.1754944e-38, 3.4028235e38, 123456.789, 0.1, 1e10, 1e-5,
                      9.999999e9], np.float32)
    small = (rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 3, 20000)).astype(np.float32)
    vals = np.concatenate([vals, extra, small])
    # one draw of root-only trees: the word of a leaf is its float's bits above the (zero) rule and child bytes
    words = np.zeros((1, len(vals), 64), np.uint64)
    words[0, :, 0] = vals.view(np.uint32).astype(np.uint64) << np.uint64(32)
    text, _ = sampler.tree_text(words, 1)
    lines = text.split("\n")
    assert lines[0] == f"1 {len(vals)} 1"
    got = [lines[2 + 2 * k].split(" ")[3] for k in range(len(vals))]
    assert got == ["%.10g" % float(v) for v in vals]


def test_compile_trees_parser_matches_oracle(oracle, golden_cases):
    from fastbart_amd.predict import CompiledTrees
    g = golden_cases["w200_p3_extreme"]
    pb = W.small_problem(200, 3, 10, 12, "extreme")
    text, _ = sampler.tree_text(g["expected"]["trees"], 3)
    ct = CompiledTrees(text, pb["cutpoints"])
    a = ct.arrays()
    b = oracle.compile_trees(text, pb["cutpoints"])
    for k in ("var", "child", "root"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("cut", "theta"):
        assert same_bits(a[k], b[k]), k
    assert ct.ndpost == b["ndpost"] and ct.ntree == b["ntree"]


def test_compile_trees_rank_form_eligibility():
    """The rank-form predictor needs every variable's cutpoints nondecreasing and free of NaN: x >= xi[c] is then
    monotone in c and equals c < #{k : x >= xi[k]}. Brute-force check of that equivalence on ties, infinities and NaN
    rows, and of which tables the library accepts for the rank form."""
    from fastbart_amd.predict import CompiledTrees
    text = "1 1 2\n3\n1 1 2 0\n2 0 0 -1\n3 0 0 1\n"
    xi_ok = [np.array([0.5]), np.array([0.1, 0.2, 0.2, 0.7, 0.7, 0.9])]
    assert CompiledTrees(text, xi_ok).rank_form
    assert not CompiledTrees(text, [xi_ok[0], xi_ok[1][::-1].copy()]).rank_form
    assert not CompiledTrees(text, [xi_ok[0], np.array([0.1, np.nan, 0.3])]).rank_form
    assert not CompiledTrees(text, [np.array([np.nan]), xi_ok[1]]).rank_form
    xi = xi_ok[1]
    xs = np.concatenate([xi, np.nextafter(xi, -np.inf), np.nextafter(xi, np.inf), [-np.inf, np.inf, np.nan, -0.0]])
    for x in xs:
        rank = sum(1 for k in range(len(xi)) if x >= xi[k])
        for c in range(len(xi)):
            assert (x >= xi[c]) == (c < rank), (x, c)


def test_compile_trees_rejects_malformed_text():
    from fastbart_amd.predict import CompiledTrees
    with pytest.raises(_lib.FbhError):
        CompiledTrees("1 1 1\n3\n1 0 0 0\n2 0 0 1.5\n", [np.array([0.5])])   # node 1 has a left child, no right
    with pytest.raises(_lib.FbhError):
        CompiledTrees("1 1 1\n3\n1 0 7 0\n2 0 0 1\n3 0 0 2\n", [np.array([0.5])])   # cutpoint 7 of 1
    with pytest.raises(_lib.FbhError):
        CompiledTrees("x", [])


def test_libm_image_tables_match_host_libm(oracle):
    """The device tables are built from the host libm with the reference's own expressions: compare all 2^24
    entries of each with the checker's evaluation of the same expressions."""
    L = _lib.lib()
    n = 1 << 24
    for kind, dt in ((0, np.float32), (1, np.float32), (2, np.float64)):
        got = np.zeros(n, dt)
        _lib.check(L.fbh_debug_libm_host(kind, 0, n, got.ctypes.data))
        exp = oracle.libm_image(kind)
        assert same_bits(got, exp), kind


def test_glibc_is_not_correctly_rounded_on_the_grid(oracle):
    """Why the device reads libm images instead of computing logf/cosf itself: glibc's float logf and cosf differ
    from the correctly rounded value on part of the 2^24-point grid (long double evaluation, rounded once)."""
    k = np.arange(0, 1 << 24, 97, dtype=np.int64)
    u = ((k.astype(np.float32) + np.float32(0.5)) * np.float32(1.0 / 16777216.0)).astype(np.float32)
    lf = oracle.libm_image(0)[k]
    cr = np.log(u.astype(np.longdouble)).astype(np.float32)
    assert np.count_nonzero(lf != cr) > 0


def test_batch_create_rejects_bad_inputs_before_any_device_work():
    pb = W.small_problem(40, 2, 4, 1)
    bad = dict(pb)
    bad["numcut"] = np.array([300, 2], np.int32)
    with pytest.raises(_lib.FbhError):
        sampler.BatchSampler([bad], 2, 1, 1, [1])
    bad = dict(pb)
    y = pb["y"].copy(); y[3] = np.nan
    bad["y"] = y
    with pytest.raises(_lib.FbhError):
        sampler.BatchSampler([bad], 2, 1, 1, [1])
    with pytest.raises(_lib.FbhError):
        sampler.BatchSampler([pb], 0, 1, 1, [1])        # m = 0
    with pytest.raises(_lib.FbhError):
        sampler.BatchSampler([pb], 2, 1, 1, [1], node_cap=2)
    for kw in (dict(packing=3), dict(chains_per_cu=3), dict(resident=3), dict(tile_rows=512)):
        with pytest.raises(_lib.FbhError):
            sampler.BatchSampler([pb], 2, 1, 1, [1], **kw)


def test_launch_configuration_entry_points_reject_a_null_batch():
    L = _lib.lib()
    einval = -1  # FBH_EINVAL
    assert L.fbh_batch_configure(None, 0, 0, 0, 0) == einval
    assert L.fbh_batch_config(None, *[C.pointer(C.c_int()) for _ in range(5)]) == einval
    assert L.fbh_batch_autotune(None, 1, None) == einval
    assert L.fbh_batch_autotune_cached(None, 1, b"tune.txt", None) == einval
    order = (C.c_int * 1)(0)
    assert L.fbh_batch_set_order(None, C.cast(order, C.c_void_p)) == einval
    assert b"null" in L.fbh_last_error()


def test_external_generator_is_refused():
    """With R's generator installed (FASTBART_RNG=R) the drop-in computes nothing and reports FBH_ERNG."""
    L = _lib.lib()
    cb = C.CFUNCTYPE(C.c_double)(lambda: 0.5)
    L.bart_set_rng(C.cast(cb, C.c_void_p), None, None, None)
    try:
        pb = W.small_problem(40, 2, 4, 1)
        st = np.zeros(3, np.int32)
        sig = np.zeros(3, np.float32)
        trm = np.zeros(40, np.float32)
        q, y, wt, nc, pg = (np.ascontiguousarray(pb[k]) for k in ("q", "y", "wt", "numcut", "pgtab"))
        L.bart_fit_cpu(_lib.ptr(q), _lib.ptr(y), _lib.ptr(wt), _lib.ptr(nc), _lib.ptr(pg), _lib.ptr(sig),
                       _lib.ptr(trm), None, 40, 2, 2, 2, 1, pb["tau"], pb["lambda"], pb["nu"], pb["sigest"], 1, 0,
                       _lib.ptr(st), 0, None, 0.0)
        assert L.fbh_last_status() == -5
        assert list(st) == [-1, -1, -1]
    finally:
        L.bart_set_rng(None, None, None, None)


def test_fx64_range_is_checked_before_any_device_work():
    """FX64 sums int64 images (w r at 2^32, w e^2 at 2^24); fbh_batch_create refuses a chain whose outcome scale could
    wrap them (n max(wt) (4 max|y|)^2 >= 2^38), and does not refuse the same chain under REF16 for that reason."""
    pb = W.small_problem(4000, 3, 10, 5)
    big = dict(pb)
    big["y"] = (pb["y"] * np.float32(3000.0 / np.max(np.abs(pb["y"])))).astype(np.float32)  # max|y| = 3000
    with pytest.raises(_lib.FbhError, match="FX64 range"):
        sampler.BatchSampler([big], 2, 1, 1, [1], stats_mode=1)
    for stats_mode, prob in ((0, big), (1, pb)):
        try:
            sampler.BatchSampler([prob], 2, 1, 1, [1], stats_mode=stats_mode).close()
        except _lib.FbhError as e:       # no device here: a later failure, not the range check
            assert "FX64 range" not in str(e)


def test_fx64_oracle_holds_at_a_large_outcome_scale_inside_the_range(oracle):
    """Inside the range (max|y| = 500, n = 4000: n (4 max|y|)^2 = 1.6e10 < 2^38) the FX64 twin's sigma trace stays
    within fixed-point rounding of the REF16 trace: no sum wraps."""
    pb = W.small_problem(4000, 3, 10, 5, weights="ones")
    s = np.float32(500.0 / np.max(np.abs(pb["y"])))
    prob = dict(pb)
    prob["y"] = (pb["y"] * s).astype(np.float32)
    prob["sigest"] = float(np.float32(pb["sigest"] * s))
    prob["lambda"] = float(np.float32(pb["lambda"] * s * s))
    prob["tau"] = float(np.float32(pb["tau"] * s))
    a = oracle.fit(prob, 5, 3, 2, 7, stats_mode=0)
    b = oracle.fit(prob, 5, 3, 2, 7, stats_mode=1)
    sa, sb = np.asarray(a["sigma"], np.float64), np.asarray(b["sigma"], np.float64)
    assert np.all(np.isfinite(sb)) and np.all(sb > 0)
    # the two modes take the same moves while their sums agree to rounding: the traces agree closely
    assert np.max(np.abs(sa - sb) / sa) < 1e-3
