"""Host-side interface over the C-ABI: batched chains, the single-chain drop-in, and the output formats.

The names and argument meanings follow the reference's C-ABI (``bart_fit_cpu``, src/bart_kernel.h:63-66) and its glue
(``fbart_cpp``, src/glue.cpp:10-62). Inputs are *problems* in the C-ABI form (see fastbart_amd.workloads).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import NODES, ChainIn, ChainOut, FbhError, Opts, Tune, check, lib, ptr


def _chain_in(prob, seed, probit, keep):
    q = np.ascontiguousarray(prob["q"], dtype=np.uint8)
    p, n = q.shape
    y = np.ascontiguousarray(prob["y"], dtype=np.float32)
    wt = np.ascontiguousarray(prob["wt"], dtype=np.float32)
    nc = np.ascontiguousarray(prob["numcut"], dtype=np.int32)
    yb = prob.get("ybin")
    yb = np.ascontiguousarray(yb, dtype=np.int32) if yb is not None else None
    keep.extend([q, y, wt, nc, yb])
    ci = ChainIn()
    ci.n, ci.p = n, p
    ci.q, ci.y, ci.wt, ci.numcut = ptr(q), ptr(y), ptr(wt), ptr(nc)
    ci.ybin = ptr(yb) if probit else None
    ci.tau = prob["tau"]
    ci.lambda_ = prob["lambda"]
    ci.nu = prob["nu"]
    ci.sigest = prob["sigest"]
    ci.offset = prob.get("offset", 0.0)
    ci.seed = int(seed) & 0xFFFFFFFF
    return ci


class BatchSampler:
    """Many independent chains sampled together on one GPU (fbh_batch_*)."""

    def __init__(self, problems, m, nskip, ndpost, seeds, probit=0, keep_trees=1, stats_mode=0, node_cap=0,
                 device=0, sweeps_per_launch=0, pgtab=None, tile_rows=0, row_split=0, packing=0, graph=0, chains_per_cu=0,
                 resident=0, probit_redraw=0):
        L = lib()
        self.problems = problems
        self.C = len(problems)
        self.m, self.nskip, self.ndpost, self.probit = m, nskip, ndpost, probit
        self.keep_trees = keep_trees
        self._keep = []
        arr = (ChainIn * self.C)()
        for c, pb in enumerate(problems):
            arr[c] = _chain_in(pb, seeds[c], probit, self._keep)
        pg = np.ascontiguousarray(pgtab if pgtab is not None else problems[0]["pgtab"], dtype=np.float32)
        self._keep.append(pg)
        o = Opts(m, nskip, ndpost, probit, keep_trees, ptr(pg), stats_mode, node_cap, device, sweeps_per_launch, 0,
                 tile_rows, row_split, packing, int(bool(graph)), chains_per_cu, resident, probit_redraw)
        h = C.c_void_p()
        check(L.fbh_batch_create(arr, self.C, C.byref(o), C.byref(h)), "fbh_batch_create")
        self.h = h
        self.n = [int(pb["q"].shape[1]) for pb in problems]
        self.p = [int(pb["q"].shape[0]) for pb in problems]

    @property
    def row_split(self):
        """Workgroups per chain (1 without a row split)."""
        return int(lib().fbh_batch_row_split(self.h))

    def configure(self, packing=0, tile_rows=0, chains_per_cu=0, resident=0):
        """Switch the workgroup -> chain packing, the tile rows, the running fit's place (1 streamed, 2 in LDS, 0
        automatic) and the build (one or two chains per CU; 0 keeps it) between runs. Results do not depend on them."""
        check(lib().fbh_batch_configure(self.h, int(packing), int(tile_rows), int(resident), int(chains_per_cu)),
              "fbh_batch_configure")

    def probit_redraw(self, mode=0, slots=0):
        """Probit batches: the latent redraw from the next run on -- 1 inside the sweep kernel, 2 by kernels spread over
        the device, 0 automatic; slots (a multiple of 64) shrinks mode 2's stream buffer. Results do not depend on
        either."""
        check(lib().fbh_batch_probit_redraw(self.h, int(mode), int(slots)), "fbh_batch_probit_redraw")

    @property
    def probit_info(self):
        """The redraw in effect (0 for wbart), mode 2's stream slots per chain, and the attempts it decided outside its
        precomputed bands so far (all chains)."""
        md, sl, sp = C.c_int(), C.c_int(), C.c_ulonglong()
        check(lib().fbh_batch_probit_info(self.h, C.byref(md), C.byref(sl), C.byref(sp)), "fbh_batch_probit_info")
        return dict(mode=md.value, slots=sl.value, spot=sp.value)

    def set_order(self, order):
        """Run logical chain order[w] on workgroup w (a permutation of the chains)."""
        o = np.ascontiguousarray(order, dtype=np.int32)
        check(lib().fbh_batch_set_order(self.h, ptr(o)), "fbh_batch_set_order")

    @property
    def config(self):
        """The launch configuration in effect: packing, tile_rows, resident (1 streamed, 2 in LDS), chains_per_cu,
        graph."""
        v = [C.c_int() for _ in range(5)]
        check(lib().fbh_batch_config(self.h, *[C.byref(x) for x in v]), "fbh_batch_config")
        return dict(zip(("packing", "tile_rows", "resident", "chains_per_cu", "graph"), (x.value for x in v)))

    def autotune(self, trial_sweeps=1, cache=None):
        """Time every (packing, tile rows) candidate on the next sweeps (real sweeps of every chain) and keep the
        fastest. Returns the choice and each candidate's ms per sweep. With cache (a file path) a batch of a shape
        tuned before takes the stored choice without running a sweep, and a new shape's choice is appended."""
        t = Tune()
        if cache is None:
            check(lib().fbh_batch_autotune(self.h, int(trial_sweeps), C.byref(t)), "fbh_batch_autotune")
        else:
            check(lib().fbh_batch_autotune_cached(self.h, int(trial_sweeps), os.fsencode(cache), C.byref(t)),
                  "fbh_batch_autotune_cached")
        cands = [dict(packing=t.cand_packing[k], tile_rows=t.cand_tile_rows[k], resident=t.cand_resident[k],
                      chains_per_cu=t.cand_chains_per_cu[k], ms_per_sweep=t.ms_per_sweep[k]) for k in range(t.ncand)]
        return dict(packing=t.packing, tile_rows=t.tile_rows, resident=t.resident, chains_per_cu=t.chains_per_cu,
                    sweeps=t.sweeps, candidates=cands, from_cache=bool(t.from_cache))

    def run(self, nsweeps):
        check(lib().fbh_batch_run(self.h, int(nsweeps)), "fbh_batch_run")

    def sync(self):
        check(lib().fbh_batch_sync(self.h), "fbh_batch_sync")

    @property
    def it(self):
        return lib().fbh_batch_iter(self.h)

    def fetch(self, c):
        n = self.n[c]
        sig = np.zeros(self.nskip + self.ndpost, np.float32)
        trm = np.zeros(n, np.float32)
        trees = np.zeros((self.ndpost, self.m, NODES), np.uint64) if self.keep_trees else None
        st = np.zeros(3, np.int32)
        out = ChainOut(ptr(sig), ptr(trm), ptr(trees), ptr(st))
        check(lib().fbh_batch_fetch(self.h, c, C.byref(out)), "fbh_batch_fetch")
        res = dict(sigma=sig, trmean=trm, stats=st)
        if trees is not None:
            res["trees"] = trees
        return res

    def state(self, c):
        n = self.n[c]
        rng = np.zeros(4, np.uint32)
        sig = C.c_double()
        af = np.zeros(n)
        lat = np.zeros(n)
        leaf = np.zeros((self.m, n), np.uint8)
        words = np.zeros((self.m, NODES), np.uint64)
        check(lib().fbh_batch_state(self.h, c, ptr(rng), C.byref(sig), ptr(af), ptr(lat), ptr(leaf), ptr(words)),
              "fbh_batch_state")
        return dict(rng=rng, sigma=sig.value, af=af, latent=lat, leaf=leaf, trees=words)

    def yhat(self, c):
        out = np.zeros((self.ndpost, self.n[c]), np.float64)
        check(lib().fbh_batch_yhat(self.h, c, ptr(out)), "fbh_batch_yhat")
        return out

    def timing(self):
        avg, k, by = C.c_double(), C.c_int(), C.c_double()
        check(lib().fbh_batch_timing(self.h, C.byref(avg), C.byref(k), C.byref(by)), "fbh_batch_timing")
        return avg.value, k.value, by.value

    # -- learned birth proposal
    def learned_setup(self, net=None, alpha=0.0, pair_capacity=0):
        """Mixture weight alpha of the network (fastbart_amd.mlp.MLP, 513 -> ... -> 8160) in the birth rule draw;
        pair_capacity > 0 records accepted births as training pairs."""
        self._net = net
        check(lib().fbh_batch_learned_setup(self.h, net.h if net is not None else None, float(alpha),
                                            int(pair_capacity)), "fbh_batch_learned_setup")

    def run_learned(self, nsweeps):
        check(lib().fbh_batch_run_learned(self.h, int(nsweeps)), "fbh_batch_run_learned")

    def learned_target(self, kind, fraction=1.0, seed=1):
        """0: accepted births with BART's rule; 1: proposed births (each with probability fraction), the exact split
        conditional there as a soft target (set before collecting)."""
        check(lib().fbh_batch_learned_target(self.h, int(kind), float(fraction), int(seed) & 0xFFFFFFFFFFFFFFFF),
              "fbh_batch_learned_target")

    def learned_holdout(self, count):
        """keep the last count pairs out of training"""
        check(lib().fbh_batch_learned_holdout(self.h, int(count)), "fbh_batch_learned_holdout")

    def learned_eval(self, k0, count):
        """(mean loss of the network, mean entropy, mean log allowed class count) on pairs [k0, k0 + count)"""
        l, e, a = C.c_double(), C.c_double(), C.c_double()
        check(lib().fbh_batch_learned_eval(self.h, int(k0), int(count), C.byref(l), C.byref(e), C.byref(a)),
              "fbh_batch_learned_eval")
        return l.value, e.value, a.value

    def learned_target_stats(self):
        """(mean entropy of the pairs' split conditionals, mean log of their allowed class counts), nats"""
        e, l = C.c_double(), C.c_double()
        check(lib().fbh_batch_learned_target_stats(self.h, C.byref(e), C.byref(l)), "fbh_batch_learned_target_stats")
        return e.value, l.value

    def learned_pairs_copy(self, k0=0, count=None):
        """Collected pairs [k0, k0 + count) on the host: dict of features [count][513], ranges [count][32],
        cls [count], entropy [count][2]."""
        if count is None:
            count = self.learned_pairs() - k0
        x = np.zeros((count, 513), np.float32)
        rg = np.zeros((count, 32), np.uint16)
        cls = np.zeros(count, np.int32)
        ent = np.zeros((count, 2), np.float32)
        check(lib().fbh_batch_learned_pairs_copy(self.h, int(k0), int(count), ptr(x), ptr(rg), ptr(cls), ptr(ent)),
              "fbh_batch_learned_pairs_copy")
        return dict(features=x, ranges=rg, cls=cls, entropy=ent)

    def learned_pairs(self):
        return int(lib().fbh_batch_learned_pairs(self.h))

    def learned_train(self, steps, batch=512, lr=1e-3, seed=1):
        losses = np.zeros(max(int(steps), 1), np.float32)
        check(lib().fbh_batch_learned_train(self.h, int(steps), int(batch), float(lr), int(seed) & 0xFFFFFFFF,
                                            ptr(losses)), "fbh_batch_learned_train")
        return losses[:steps]

    def learned_timing(self):
        ms, k = C.c_double(), C.c_int()
        check(lib().fbh_batch_learned_timing(self.h, C.byref(ms), C.byref(k)), "fbh_batch_learned_timing")
        return ms.value, k.value

    def enable_trace(self, cap):
        check(lib().fbh_batch_trace(self.h, int(cap), None, 0), "fbh_batch_trace")
        self._trace_cap = cap

    def trace(self, c):
        buf = (_lib.TraceRec * self._trace_cap)()
        check(lib().fbh_batch_trace(self.h, 0, C.cast(buf, C.c_void_p), c), "fbh_batch_trace")
        return buf

    def close(self):
        if getattr(self, "h", None):
            lib().fbh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def bart_fit(prob, m, nskip, ndpost, seed, keep_trees=1, probit=0):
    """The drop-in entry ``bart_fit_cpu`` (src/bart_kernel.h:63-66) on one chain; returns its output buffers."""
    L = lib()
    q = np.ascontiguousarray(prob["q"], dtype=np.uint8)
    p, n = q.shape
    y = np.ascontiguousarray(prob["y"], dtype=np.float32)
    wt = np.ascontiguousarray(prob["wt"], dtype=np.float32)
    nc = np.ascontiguousarray(prob["numcut"], dtype=np.int32)
    pg = np.ascontiguousarray(prob["pgtab"], dtype=np.float32)
    yb = prob.get("ybin")
    yb = np.ascontiguousarray(yb, dtype=np.int32) if yb is not None else None
    sig = np.zeros(nskip + ndpost, np.float32)
    trm = np.zeros(n, np.float32)
    trees = np.zeros((ndpost, m, NODES), np.uint64) if keep_trees else None
    st = np.zeros(8, np.int32)
    L.bart_set_rng(None, None, None, None)
    L.bart_fit_cpu(ptr(q), ptr(y), ptr(wt), ptr(nc), ptr(pg), ptr(sig), ptr(trm), ptr(trees), n, p, m, nskip, ndpost,
                   prob["tau"], prob["lambda"], prob["nu"], prob["sigest"], int(seed) & 0xFFFFFFFF, keep_trees,
                   ptr(st), probit, ptr(yb) if probit else None, prob.get("offset", 0.0))
    if L.fbh_last_status() != 0:
        raise FbhError("bart_fit_cpu: " + L.fbh_last_error().decode(errors="replace"))
    res = dict(sigma=sig, trmean=trm, stats=st[:3].copy())
    if keep_trees:
        res["trees"] = trees
    return res


def tree_text(trees, p):
    """Treedraws text and varcount of [ndpost][m][64] output words (src/glue.cpp:32-51)."""
    trees = np.ascontiguousarray(trees, dtype=np.uint64)
    nd, m = trees.shape[0], trees.shape[1]
    L = lib()
    vc = np.zeros((nd, p), np.int32)
    # one formatting pass in the usual case: a buffer of 128 bytes a tree (the text of a tree of the workloads is about
    # 80), a second call with the exact length only if the text is longer
    cap = nd * m * 128 + 64
    buf = C.create_string_buffer(cap)
    need = L.fbh_tree_text(ptr(trees), nd, m, p, C.cast(buf, C.c_void_p), cap, ptr(vc))
    if need + 1 > cap:
        buf = C.create_string_buffer(need + 1)
        vc[:] = 0
        L.fbh_tree_text(ptr(trees), nd, m, p, C.cast(buf, C.c_void_p), need + 1, ptr(vc))
    return buf.raw[:need].decode(), vc


def yhat(trees, q, device=0):
    """Per-draw training fits from output words (src/glue.cpp:52-56), on the device."""
    trees = np.ascontiguousarray(trees, dtype=np.uint64)
    q = np.ascontiguousarray(q, dtype=np.uint8)
    nd, m = trees.shape[0], trees.shape[1]
    p, n = q.shape
    out = np.zeros((nd, n), np.float64)
    check(lib().fbh_yhat(ptr(trees), ptr(q), n, p, nd, m, device, ptr(out)), "fbh_yhat")
    return out


def debug_math(kind, u=None, d=None, device=0):
    u = None if u is None else np.ascontiguousarray(u, dtype=np.uint32)
    d = None if d is None else np.ascontiguousarray(d, dtype=np.float64)
    count = {0: len(u) // 2 if u is not None else 0, 1: len(u) if u is not None else 0}.get(
        kind, (len(d) // 2 if kind == 5 else len(d)) if d is not None else 0)
    of = np.zeros(count, np.float32) if kind <= 1 else None
    od = np.zeros(count, np.float64) if kind >= 2 else None
    check(lib().fbh_debug_math(kind, ptr(u), ptr(d), count, ptr(of), ptr(od), device), "fbh_debug_math")
    return of if kind <= 1 else od


def debug_copy(mib, device=0):
    """One plain device copy of mib MiB (counter calibration); returns its duration in ms."""
    ms = C.c_double()
    check(lib().fbh_debug_copy(int(mib), device, C.byref(ms)), "fbh_debug_copy")
    return ms.value


def debug_split_conditional(q, member, w, r, ranges, depth, ngv, ngood, pg, sigma, tau, u=None):
    """The device's exact split conditional at one node (the target of fbh_batch_learned_target kind 1):
    (q [p][255], entropy, classes drawn at the uniforms u)."""
    q = np.ascontiguousarray(q, np.uint8)
    p, n = q.shape
    member = np.ascontiguousarray(member, np.uint8)
    w = np.ascontiguousarray(w, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    ranges = np.ascontiguousarray(ranges, np.uint16)
    pg = np.ascontiguousarray(pg, np.float64)
    u = np.ascontiguousarray(u if u is not None else [], np.float64)
    out = np.zeros((p, 255), np.float32)
    cls = np.zeros(max(len(u), 1), np.int32)
    h = C.c_double()
    check(lib().fbh_debug_split_conditional(n, p, ptr(q), ptr(member), ptr(w), ptr(r), ptr(ranges), int(depth),
                                            int(ngv), int(ngood), ptr(pg), float(sigma), float(tau), ptr(u), len(u),
                                            ptr(out), ptr(cls), C.byref(h)), "fbh_debug_split_conditional")
    return out, h.value, cls[:len(u)]
