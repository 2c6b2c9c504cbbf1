"""ctypes access to the CPU checkers -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by ``__graft_entry__.smoke()`` and by the CPU-baseline leg of ``bench.py``; never by the product.

* ``Oracle``: the project's C restatement (oracle/fastbart_oracle.c) -> liboracle.so
* ``RefCore``: the reference's own core (src/bart_kernel.cpp built by oracle/Makefile) -> _ref/libbart_ref.so.
  It keeps its state in static arrays (src/bart_kernel.cpp:117-125), so one fit at a time per process.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbart_ref.so")
REF_N8_SO = os.path.join(HERE, "_ref", "libbart_ref_n8.so")  # same core built with 8 node slots
REF_SRC = "/root/reference/src/bart_kernel.cpp"

NODES = 64
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so, and the reference core when its sources are present (ref=None: when present)."""
    targets = ["oracle"]
    if ref or (ref is None and os.path.exists(REF_SRC)):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


class OrcTrace(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("it", "j", "mode", "nx", "v", "c", "nl", "nr", "accepted", "ngood",
                                          "nnogs", "cap")] + \
               [(k, C.c_double) for k in ("pr", "lalpha", "lu", "sl0", "sl1", "sr0", "sr1")] + \
               [("S0", C.c_double * NODES), ("S1", C.c_double * NODES), ("rng_after", C.c_uint32 * 4)]


class OrcOpts(C.Structure):
    _fields_ = [("stats_mode", C.c_int), ("node_cap", C.c_int), ("trace", C.POINTER(OrcTrace)),
                ("trace_cap", C.c_long)]


class OrcCt(C.Structure):
    _fields_ = [("ndpost", C.c_int), ("ntree", C.c_int), ("p", C.c_int), ("nnodes", C.c_int),
                ("var", C.POINTER(C.c_int)), ("child", C.POINTER(C.c_int)), ("cut", C.POINTER(C.c_double)),
                ("theta", C.POINTER(C.c_double)), ("root", C.POINTER(C.c_int))]


_FIT_ARGS = [_u8p, _f32p, _f32p, _i32p, _f32p, _f32p, _f32p, C.c_void_p,
             C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
             C.c_uint, C.c_int, _i32p, C.c_int, C.c_void_p, C.c_float]


def _as_inputs(prob):
    q = np.ascontiguousarray(prob["q"], dtype=np.uint8)
    y = np.ascontiguousarray(prob["y"], dtype=np.float32)
    wt = np.ascontiguousarray(prob["wt"], dtype=np.float32)
    nc = np.ascontiguousarray(prob["numcut"], dtype=np.int32)
    pg = np.ascontiguousarray(prob["pgtab"], dtype=np.float32)
    yb = prob.get("ybin")
    yb = None if yb is None else np.ascontiguousarray(yb, dtype=np.int32)
    return q, y, wt, nc, pg, yb


def _outputs(n, m, nskip, ndpost, keep_trees):
    sig = np.zeros(nskip + ndpost, np.float32)
    trm = np.zeros(n, np.float32)
    trees = np.zeros((ndpost * m * NODES) if keep_trees else 1, np.uint64)
    st = np.zeros(8, np.int32)
    return sig, trm, trees, st


class Oracle:
    """The C restatement."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        self.L = L
        L.orc_fit.argtypes = _FIT_ARGS + [C.POINTER(OrcOpts)]
        L.orc_fit.restype = C.c_int
        L.orc_rng_seed.argtypes = [_u32p, C.c_uint32]
        L.orc_rng_next.argtypes = [_u32p]
        L.orc_rng_next.restype = C.c_uint32
        L.orc_unif.argtypes = [_u32p]
        L.orc_unif.restype = C.c_float
        L.orc_rnorm.argtypes = [_u32p]
        L.orc_rnorm.restype = C.c_float
        L.orc_rgamma.argtypes = [_u32p, C.c_double]
        L.orc_rgamma.restype = C.c_double
        L.orc_rtnorm.argtypes = [_u32p, C.c_double, C.c_double]
        L.orc_rtnorm.restype = C.c_double
        L.orc_cmp_log.argtypes = [C.c_void_p, C.c_long, C.c_long]
        L.orc_cmp_log.restype = None
        L.orc_cmp_count.argtypes = [C.POINTER(C.c_long)]
        L.orc_cmp_count.restype = C.c_long
        L.orc_chain_new.argtypes = [_u8p, _f32p, _f32p, _i32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint, C.c_int, C.c_void_p,
                                    C.c_float, C.POINTER(OrcOpts)]
        L.orc_chain_new.restype = C.c_void_p
        L.orc_chain_free.argtypes = [C.c_void_p]
        L.orc_chain_sweeps.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_chain_trmean.argtypes = [C.c_void_p, _f32p]
        L.orc_chain_stats.argtypes = [C.c_void_p, _i32p]
        L.orc_chain_iter.argtypes = [C.c_void_p]
        L.orc_chain_iter.restype = C.c_int
        L.orc_chain_rng.argtypes = [C.c_void_p, _u32p]
        L.orc_chain_sigma.argtypes = [C.c_void_p]
        L.orc_chain_sigma.restype = C.c_double
        L.orc_chain_af.argtypes = [C.c_void_p, _f64p]
        L.orc_chain_latent.argtypes = [C.c_void_p, _f64p]
        L.orc_chain_leaf.argtypes = [C.c_void_p, _u8p]
        L.orc_chain_pack_trees.argtypes = [C.c_void_p, _u64p]
        L.orc_chain_trace_len.argtypes = [C.c_void_p]
        L.orc_chain_trace_len.restype = C.c_long
        L.orc_chain_set_informed.argtypes = [C.c_void_p, C.c_double]
        L.orc_chain_set_informed.restype = None
        L.orc_chain_informed_births.argtypes = [C.c_void_p]
        L.orc_chain_informed_births.restype = C.c_long
        L.orc_split_conditional.argtypes = [C.c_int, C.c_int, _u8p, _u8p, _f64p, _f64p,
                                            np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS"), C.c_int,
                                            C.c_int, C.c_int, _f64p, C.c_double, C.c_double, _f64p,
                                            C.POINTER(C.c_double)]
        L.orc_tree_text.argtypes = [_u64p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_size_t, C.c_void_p]
        L.orc_tree_text.restype = C.c_size_t
        L.orc_yhat.argtypes = [_u64p, _u8p, C.c_int, C.c_int, C.c_int, _f64p]
        L.orc_compile_trees.argtypes = [C.c_char_p, _f64p, _i32p, C.c_int, C.POINTER(OrcCt)]
        L.orc_compile_trees.restype = C.c_int
        L.orc_ct_free.argtypes = [C.POINTER(OrcCt)]
        L.orc_predict_draw.argtypes = [C.POINTER(OrcCt), C.POINTER(C.POINTER(C.c_double)), C.c_int, C.c_int,
                                       C.c_int, C.c_double, _f64p]
        L.orc_predict_draw.restype = C.c_int
        L.orc_libm_image.argtypes = [C.c_int, C.c_long, C.c_long, C.c_void_p, C.c_void_p]
        L.orc_libm_image.restype = None
        L.orc_normal_from.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_normal_from.restype = C.c_float

    def libm_image(self, kind, k0=0, count=1 << 24):
        """Host libm on the generator grid: 0 logf, 1 cosf(2 pi u), 2 log(double), 3 sqrtf(-2 logf)."""
        out = np.zeros(count, np.float64 if kind == 2 else np.float32)
        self.L.orc_libm_image(kind, k0, count, None if kind == 2 else out.ctypes.data,
                              out.ctypes.data if kind == 2 else None)
        return out

    def normal_from(self, x1, x2):
        return np.array([self.L.orc_normal_from(int(a), int(b)) for a, b in zip(x1, x2)], np.float32)

    # -- generator
    def rng_stream(self, seed: int, count: int, kind: str = "next"):
        s = np.zeros(4, np.uint32)
        self.L.orc_rng_seed(s, seed)
        f = {"next": self.L.orc_rng_next, "unif": self.L.orc_unif, "rnorm": self.L.orc_rnorm}[kind]
        dt = {"next": np.uint32, "unif": np.float32, "rnorm": np.float32}[kind]
        return np.array([f(s) for _ in range(count)], dtype=dt), s

    # -- whole fit
    def fit(self, prob, m, nskip, ndpost, seed, keep_trees=1, probit=0, stats_mode=0, node_cap=0, trace_cap=0):
        q, y, wt, nc, pg, yb = _as_inputs(prob)
        p, n = q.shape
        sig, trm, trees, st = _outputs(n, m, nskip, ndpost, keep_trees)
        opts = OrcOpts(stats_mode, node_cap, None, 0)
        tr = None
        if trace_cap:
            tr = (OrcTrace * trace_cap)()
            opts.trace = C.cast(tr, C.POINTER(OrcTrace))
            opts.trace_cap = trace_cap
        rc = self.L.orc_fit(q, y, wt, nc, pg, sig, trm, trees.ctypes.data if keep_trees else None, n, p, m, nskip,
                            ndpost, prob["tau"], prob["lambda"], prob["nu"], prob["sigest"], seed, keep_trees, st,
                            probit, yb.ctypes.data if yb is not None else None, prob.get("offset", 0.0),
                            C.byref(opts))
        if rc != 0:
            raise RuntimeError(f"orc_fit failed: {rc}")
        out = dict(sigma=sig, trmean=trm, stats=st[:3].copy())
        if keep_trees:
            out["trees"] = trees.reshape(ndpost, m, NODES)
        if tr is not None:
            out["trace"] = tr
        return out

    def fit_logging_comparisons(self, prob, m, nskip, ndpost, seed, probit=0, cap=1_000_000, sample=200_000):
        """fit() with the libm comparisons of rtnorm and rgamma logged (orc_cmp_log): (records [k][7], comparisons
        seen, fit outputs). The first `sample` comparisons are all recorded; after them only the near ones."""
        buf = np.zeros((cap, 7), np.float64)
        self.L.orc_cmp_log(buf.ctypes.data, cap, sample)
        try:
            out = self.fit(prob, m, nskip, ndpost, seed, keep_trees=0, probit=probit)
            seen = C.c_long()
            k = self.L.orc_cmp_count(C.byref(seen))
        finally:
            self.L.orc_cmp_log(None, 0, 0)
        return buf[:k].copy(), seen.value, out

    # -- chains advanced sweep by sweep (state snapshots for parity)
    def chain(self, prob, m, nskip, ndpost, seed, probit=0, stats_mode=0, node_cap=0, trace_cap=0):
        return OracleChain(self, prob, m, nskip, ndpost, seed, probit, stats_mode, node_cap, trace_cap)

    # -- glue
    def split_conditional(self, q, member, w, r, ranges, depth, ngv, ngood, pg, sigma, tau):
        """The informed birth rule's conditional at one node (no reference counterpart): (q [p][255], entropy,
        number of rules)."""
        q = np.ascontiguousarray(q, np.uint8)
        p, n = q.shape
        out = np.zeros((p, 255), np.float64)
        h = C.c_double()
        nr = self.L.orc_split_conditional(n, p, q, np.ascontiguousarray(member, np.uint8),
                                          np.ascontiguousarray(w, np.float64), np.ascontiguousarray(r, np.float64),
                                          np.ascontiguousarray(ranges, np.uint16), int(depth), int(ngv), int(ngood),
                                          np.ascontiguousarray(pg, np.float64), float(sigma), float(tau), out,
                                          C.byref(h))
        if nr < 0:
            raise ValueError("orc_split_conditional: bad argument")
        return out, h.value, nr

    def tree_text(self, trees, p):
        trees = np.ascontiguousarray(trees, dtype=np.uint64)
        nd, m = trees.shape[0], trees.shape[1]
        flat = trees.reshape(-1)
        need = self.L.orc_tree_text(flat, nd, m, p, None, 0, None)
        buf = C.create_string_buffer(need + 1)
        vc = np.zeros(nd * p, np.int32)
        self.L.orc_tree_text(flat, nd, m, p, buf, need + 1, vc.ctypes.data)
        return buf.raw[:need].decode(), vc.reshape(nd, p)

    def yhat(self, trees, q):
        trees = np.ascontiguousarray(trees, dtype=np.uint64)
        q = np.ascontiguousarray(q, dtype=np.uint8)
        nd, m = trees.shape[0], trees.shape[1]
        out = np.zeros((nd, q.shape[1]), np.float64)
        self.L.orc_yhat(trees.reshape(-1), q, q.shape[1], nd, m, out)
        return out

    # -- predictor
    def compile_trees(self, text: str, cutpoints):
        flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64) for c in cutpoints])
                                    if len(cutpoints) else np.zeros(0), dtype=np.float64)
        off = np.zeros(len(cutpoints) + 1, np.int32)
        off[1:] = np.cumsum([len(c) for c in cutpoints])
        ct = OrcCt()
        rc = self.L.orc_compile_trees(text.encode(), flat if flat.size else np.zeros(1), off, len(cutpoints),
                                      C.byref(ct))
        if rc != 0:
            raise ValueError(f"orc_compile_trees failed: {rc}")
        k, nr = ct.nnodes, ct.ndpost * ct.ntree
        res = dict(ndpost=ct.ndpost, ntree=ct.ntree, p=ct.p,
                   var=np.ctypeslib.as_array(ct.var, (k,)).copy(), child=np.ctypeslib.as_array(ct.child, (k,)).copy(),
                   cut=np.ctypeslib.as_array(ct.cut, (k,)).copy(), theta=np.ctypeslib.as_array(ct.theta, (k,)).copy(),
                   root=np.ctypeslib.as_array(ct.root, (max(nr, 1),))[:nr].copy())
        self.L.orc_ct_free(C.byref(ct))
        return res

    def predict_draw(self, ct, cols, m_draw, mu):
        """ct: dict from compile_trees; cols: list of p float64 columns."""
        o = OrcCt(ct["ndpost"], ct["ntree"], ct["p"], len(ct["var"]))
        keep = [np.ascontiguousarray(ct[k]) for k in ("var", "child", "cut", "theta", "root")]
        o.var = keep[0].ctypes.data_as(C.POINTER(C.c_int))
        o.child = keep[1].ctypes.data_as(C.POINTER(C.c_int))
        o.cut = keep[2].ctypes.data_as(C.POINTER(C.c_double))
        o.theta = keep[3].ctypes.data_as(C.POINTER(C.c_double))
        o.root = keep[4].ctypes.data_as(C.POINTER(C.c_int))
        colsc = [np.ascontiguousarray(c, dtype=np.float64) for c in cols]
        n = len(colsc[0]) if colsc else 0
        arr = (C.POINTER(C.c_double) * max(len(colsc), 1))(*[c.ctypes.data_as(C.POINTER(C.c_double)) for c in colsc])
        out = np.zeros(n, np.float64)
        rc = self.L.orc_predict_draw(C.byref(o), arr, n, len(colsc), m_draw, mu, out)
        if rc != 0:
            raise ValueError(f"orc_predict_draw failed: {rc}")
        return out


class OracleChain:
    def __init__(self, orc: Oracle, prob, m, nskip, ndpost, seed, probit=0, stats_mode=0, node_cap=0, trace_cap=0):
        self.o = orc
        q, y, wt, nc, pg, yb = _as_inputs(prob)
        self.p, self.n = q.shape
        self.m, self.nskip, self.ndpost = m, nskip, ndpost
        self._keep = (q, y, wt, nc, pg, yb)
        opts = OrcOpts(stats_mode, node_cap, None, 0)
        self._trace = None
        if trace_cap:
            self._trace = (OrcTrace * trace_cap)()
            opts.trace = C.cast(self._trace, C.POINTER(OrcTrace))
            opts.trace_cap = trace_cap
        self.h = orc.L.orc_chain_new(q, y, wt, nc, pg, self.n, self.p, m, nskip, ndpost, prob["tau"], prob["lambda"],
                                     prob["nu"], prob["sigest"], seed, probit,
                                     yb.ctypes.data if yb is not None else None, prob.get("offset", 0.0),
                                     C.byref(opts))
        if not self.h:
            raise RuntimeError("orc_chain_new failed")
        self.sigma_out = np.zeros(nskip + ndpost, np.float32)
        self.trees_out = np.zeros(max(ndpost * m * NODES, 1), np.uint64)

    def sweeps(self, k):
        self.o.L.orc_chain_sweeps(self.h, k, self.sigma_out.ctypes.data, self.trees_out.ctypes.data, 1)

    @property
    def it(self):
        return self.o.L.orc_chain_iter(self.h)

    def state(self):
        L = self.o.L
        rng = np.zeros(4, np.uint32)
        L.orc_chain_rng(self.h, rng)
        af = np.zeros(self.n)
        L.orc_chain_af(self.h, af)
        y = np.zeros(self.n)
        L.orc_chain_latent(self.h, y)
        leaf = np.zeros(self.m * self.n, np.uint8)
        L.orc_chain_leaf(self.h, leaf)
        words = np.zeros(self.m * NODES, np.uint64)
        L.orc_chain_pack_trees(self.h, words)
        st = np.zeros(8, np.int32)
        L.orc_chain_stats(self.h, st)
        return dict(rng=rng, sigma=L.orc_chain_sigma(self.h), af=af, latent=y, leaf=leaf.reshape(self.m, self.n),
                    trees=words.reshape(self.m, NODES), stats=st[:3].copy(), it=self.it)

    def set_informed(self, alpha):
        """Proposal study (no reference counterpart): births draw their rule with probability alpha from the rule's
        conditional given the tree, sigma and the partial residuals; 0 is the reference's proposal."""
        self.o.L.orc_chain_set_informed(self.h, float(alpha))

    @property
    def informed_births(self):
        return self.o.L.orc_chain_informed_births(self.h)

    def trmean(self):
        out = np.zeros(self.n, np.float32)
        self.o.L.orc_chain_trmean(self.h, out)
        return out

    def trace(self):
        """Per-tree-update records so far (OrcTrace structures)."""
        k = self.o.L.orc_chain_trace_len(self.h) if self._trace is not None else 0
        return [self._trace[i] for i in range(k)]

    def close(self):
        if self.h:
            self.o.L.orc_chain_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefCore:
    """The reference's compiled core (bart_fit_cpu with the internal generator)."""

    def __init__(self, path: str = REF_SO, nodes: int = NODES):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.nodes = nodes
        L = C.CDLL(path)
        self.L = L
        L.bart_set_rng.argtypes = [C.c_void_p] * 4
        L.bart_fit_cpu.argtypes = _FIT_ARGS
        L.bart_fit_cpu.restype = None
        L.bart_set_rng(None, None, None, None)

    @staticmethod
    def available(path: str = REF_SO) -> bool:
        return os.path.exists(path)

    def fit(self, prob, m, nskip, ndpost, seed, keep_trees=1, probit=0):
        q, y, wt, nc, pg, yb = _as_inputs(prob)
        p, n = q.shape
        sig, trm, _, st = _outputs(n, m, nskip, ndpost, 0)
        trees = np.zeros(max(ndpost * m * self.nodes, 1), np.uint64)
        self.L.bart_set_rng(None, None, None, None)
        self.L.bart_fit_cpu(q, y, wt, nc, pg, sig, trm, trees.ctypes.data if keep_trees else None, n, p, m, nskip,
                            ndpost, prob["tau"], prob["lambda"], prob["nu"], prob["sigest"], seed, keep_trees, st,
                            probit, yb.ctypes.data if yb is not None else None, prob.get("offset", 0.0))
        out = dict(sigma=sig, trmean=trm, stats=st[:3].copy())
        if keep_trees:
            out["trees"] = trees.reshape(ndpost, m, self.nodes)
        return out
