"""Continuous-batching request loop on one MI355X.

Requests (one structure each) are admitted as they arrive, bucketed by
padded matrix dimension (identity padding keeps the padded eigenvalues at a
constant far above every band), and launched as batched solves on a small
set of HIP streams whenever a stream slot and HBM budget are free; requests
that arrive while others are on the device queue up and join the next batch
of their bucket. Per batch: one fused build launch, one batched eigensolve,
Fermi / band-edge / DOS reductions, and one asynchronous copy of the scalar
results to pinned host memory; completion is observed with events.

Each stream has its own launch thread: a batch's kernel sequence (thousands
of launches for a large matrix) is enqueued by that thread, so a stream whose
hardware queue is full blocks only its own thread while the admission loop
and the other streams keep going (ctypes releases the GIL in the library
calls).

Small buckets (ld <= graph_max_ld) replay their eigensolve from a HIP graph:
the second batch of a (ld, matrices) shape captures the solve over that
batch's H and S buffers, which become persistent; later batches of the shape
(on any stream, after the previous user's completion event) build straight
into those buffers and replay the graph; a shape whose graph is busy on
another launch thread is solved eagerly. The cached
graphs (least recently used evicted beyond graph_cache_bytes, once their last
replay has completed) hold their buffers outside the per-batch HBM budget.
For small solves the launch chain is comparable to the work, so a replay is
5-15 % shorter; from ld 512 up it gains nothing measurable and the eager path
is kept.

SK tables stay in HBM for the lifetime of the engine.
"""
from __future__ import annotations

import threading
import time
from collections import OrderedDict, deque
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib, engine
from .tables import TableSet


def bucket_dim(n: int) -> int:
    """Padded dimension: multiples of 32 to 256, of 64 to 1024, of 128 beyond."""
    if n <= 256:
        step = 32
    elif n <= 1024:
        step = 64
    else:
        step = 128
    return ((n + step - 1) // step) * step


@dataclass
class Request:
    idx: int
    inp: engine.StructureInput
    n_orb: int
    arrival: float
    ld: int
    real: bool = False   # every k-point's H and S are real symmetric


@dataclass
class Batch:
    reqs: List[Request]
    ld: int
    stream: torch.cuda.Stream
    event: torch.cuda.Event
    host: torch.Tensor
    nbytes: int
    keep: list = field(default_factory=list)
    dos: Optional[torch.Tensor] = None
    w: Optional[torch.Tensor] = None
    plan: Optional[engine.BuildPlan] = None
    t_launch: float = 0.0


@dataclass
class SolveGraph:
    """A captured eigensolve of one (ld, matrices) shape and the
    persistent buffers it reads (H, S) and writes (w, info, workspace)."""
    graph: torch.cuda.CUDAGraph
    H: torch.Tensor
    S: torch.Tensor
    w: torch.Tensor
    info: torch.Tensor
    ws: torch.Tensor
    nbytes: int
    done: Optional[torch.cuda.Event] = None
    busy: bool = True   # a batch is between its replay and the record of ``done``


class ServingEngine:
    def __init__(self, tables: TableSet, dos_grid: Optional[torch.Tensor] = None, n_streams: int = 8,
                 max_inflight: int = 8, hbm_budget_bytes: float = 64e9, max_batch_matrices: int = 512,
                 fermi_mode: int = 0, keep_outputs: bool = False, precision: str = "mixed",
                 refine_edges: bool = True, bucket=bucket_dim, graphs: bool = True, graph_max_ld: int = 256,
                 graph_cache_bytes: float = 8e9, real_symmetric: bool = True):
        self.tables = tables
        self.dev = tables.device
        self.dev_index = self.dev.index if self.dev.index is not None else torch.cuda.current_device()
        self.grid = dos_grid
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(n_streams)]
        # one launch thread per stream keeps each stream's launch order
        self.launchers = [ThreadPoolExecutor(max_workers=1) for _ in range(n_streams)]
        # graph captures run on streams of their own: a capturing stream must
        # not be one whose earlier events the admission loop still queries
        self.capture_streams = [torch.cuda.Stream(device=self.dev) for _ in range(n_streams)]
        self._pending = [0] * n_streams
        self._pending_cost = [0.0] * n_streams
        self.max_inflight = max_inflight
        self.budget = hbm_budget_bytes
        self.max_batch = max_batch_matrices
        self.fermi_mode = fermi_mode
        self.keep_outputs = keep_outputs
        self.precision = precision
        self.refine_edges = refine_edges
        self.bucket = bucket
        self.nf = len(engine.EDGE_FIELDS)
        self._rr = 0
        self._bytes_cache = {}
        self.real_symmetric = real_symmetric
        self.graphs = graphs
        self.graph_max_ld = graph_max_ld
        self.graph_cache_bytes = graph_cache_bytes
        self._graphs: "OrderedDict[tuple, SolveGraph]" = OrderedDict()
        self._graph_seen: Dict[tuple, int] = {}
        self._graph_bytes = 0
        self._graph_lock = threading.Lock()
        self.graph_replays = 0
        self.graph_captures = 0

    # ------------------------------------------------------------------
    def _orbitals(self, inp: engine.StructureInput) -> int:
        sp = self.tables.species_index(inp.z)
        return int(self.tables.norb[sp].sum())

    def batch_bytes(self, ld: int, nmat: int) -> int:
        """Device bytes a batch holds while in flight: H and S, the exact
        eigensolver workspace (slk_eigh_*_workspace_bytes), eigenvalues, the
        band-edge refinement scratch and the per-structure outputs."""
        key = (ld, nmat)
        got = self._bytes_cache.get(key)
        if got is None:
            lib = _lib.host_lib()
            ws = (lib.slk_eigh_mixed_workspace_bytes(ld, ld, nmat) if self.precision == "mixed" else
                  lib.slk_eigh_workspace_bytes(ld, ld, nmat, 0))
            ref = lib.slk_refine_scratch_bytes(ld, nmat) if self.precision == "mixed" else 0
            ne = 0 if self.grid is None else int(self.grid.numel())
            got = 2 * nmat * ld * ld * 16 + int(ws) + int(ref) + nmat * ld * 8 + 8 * nmat + \
                nmat * (ne + 3 + len(engine.EDGE_FIELDS)) * 8
            self._bytes_cache[key] = got
        return got

    @staticmethod
    def batch_cost(ld: int, nmat: int) -> float:
        """Relative time a batch holds its stream: the per-matrix chains (chase,
        panel QRs) grow as ld^2 whatever the batch size, the arithmetic as
        nmat * ld^3."""
        return float(ld) * ld + nmat * float(ld) ** 3 / 4096.0

    def _submit(self, reqs: List[Request], ld: int, real: bool = False) -> Future:
        """Queue a batch on the stream with the least pending cost, so that a
        small batch does not wait behind a large one on the same stream."""
        nmat = sum(len(r.inp.kpoints) for r in reqs)
        cost = self.batch_cost(ld, nmat)
        i = min(range(len(self.streams)), key=lambda q: (self._pending_cost[q], (q - self._rr) % len(self.streams)))
        self._rr += 1
        self._pending[i] += 1
        self._pending_cost[i] += cost
        fut = self.launchers[i].submit(self._launch, reqs, ld, self.streams[i], i, real)
        fut.stream_index = i
        fut.cost = cost
        fut.reqs = reqs
        return fut

    def close(self):
        for ex in self.launchers:
            ex.shutdown(wait=True)

    # ------------------------------------------------------------------
    def _graph_get(self, key) -> Optional[SolveGraph]:
        with self._graph_lock:
            e = self._graphs.get(key)
            if e is not None:
                if e.busy:     # another launch thread is between its replay and the record of ``done``
                    return None
                self._graphs.move_to_end(key)
                e.busy = True
                return e
            self._graph_seen[key] = self._graph_seen.get(key, 0) + 1
            return None

    def _graph_wanted(self, key, nbytes: int) -> bool:
        """Capture on the second batch of a shape, if it fits the cache and
        no other launch thread holds a graph of the shape."""
        with self._graph_lock:
            return (key not in self._graphs and self._graph_seen.get(key, 0) >= 2 and
                    nbytes <= self.graph_cache_bytes / 8)

    def _graph_insert(self, key, e: SolveGraph):
        with self._graph_lock:
            self._graphs[key] = e
            self._graph_bytes += e.nbytes
            # evict the least recently used entries whose last replay is done
            for k in list(self._graphs):
                if self._graph_bytes <= self.graph_cache_bytes:
                    break
                old = self._graphs[k]
                if k == key or old.busy or (old.done is not None and not old.done.query()):
                    continue
                del self._graphs[k]
                self._graph_bytes -= old.nbytes

    def _capture(self, H: torch.Tensor, S: torch.Tensor, si: int, real: bool) -> SolveGraph:
        """Capture the eigensolve over H, S on the launch thread's capture
        stream; the capture records without executing, the caller replays on
        its serving stream (after the build that fills H and S)."""
        g = torch.cuda.CUDAGraph()
        wsl = []
        with torch.cuda.stream(self.capture_streams[si]):
            g.capture_begin(capture_error_mode="thread_local")
            try:
                w, _, info = engine.eigh(H, S, check=False, precision=self.precision, workspace=wsl, real=real)
            finally:
                g.capture_end()
        nbytes = 2 * H.numel() * H.element_size() + wsl[0].numel() + w.numel() * 8 + info.numel() * 4
        return SolveGraph(g, H, S, w, info, wsl[0], nbytes)

    def _solve(self, plan, ld: int, nmat: int, stream, si: int, real: bool = False):
        """Build and solve one batch on the current stream: from a cached
        graph for small buckets, else eagerly. Returns H, S, w, info, workspace."""
        key = (ld, nmat, real)
        # no capture while the kernel sampler is on (its launches record events)
        graphs = self.graphs and ld <= self.graph_max_ld and not engine.SAMPLING
        e = self._graph_get(key) if graphs else None
        if e is not None:
            if e.done is not None:
                # the previous batch of this shape (maybe on another stream)
                # has finished reading the buffers before they are rebuilt
                torch.cuda.current_stream().wait_event(e.done)
            H, S = engine.run_build(self.tables, plan, out=(e.H, e.S))
        else:
            H, S = engine.run_build(self.tables, plan)
            if graphs and self._graph_wanted(key, 3 * H.numel() * 16):
                e = self._capture(H, S, si, real)
                self._graph_insert(key, e)
                self.graph_captures += 1
        if e is not None:
            e.graph.replay()
            e.done = torch.cuda.Event()
            self.graph_replays += 1
            return H, S, e.w, e.info, e.ws, e
        wsl = []
        w, _, info = engine.eigh(H, S, check=False, precision=self.precision, workspace=wsl, real=real)
        return H, S, w, info, wsl[0], None

    def _launch(self, reqs: List[Request], ld: int, stream, si: int = 0, real: bool = False) -> Batch:
        torch.cuda.set_device(self.dev_index)
        nmat = sum(len(r.inp.kpoints) for r in reqs)
        with torch.cuda.stream(stream):
            plan = engine.plan_build(self.tables, [r.inp for r in reqs], ld=ld)
            H, S, w, info, ws, ge = self._solve(plan, ld, nmat, stream, si, real)
            refine = (H, S, ws) if (self.refine_edges and self.precision == "mixed") else None
            edges, dos, stats = engine.post(plan, w, fermi_mode=self.fermi_mode, dos_grid=self.grid, refine=refine)
            if ge is not None:
                # the graph's buffers are free for the next replay once this
                # batch's readers (refinement, post) have run
                ge.done.record(stream)
                with self._graph_lock:
                    ge.busy = False
                if self.keep_outputs:
                    w = w.clone()
            parts = [edges.reshape(-1), info.to(torch.float64)]
            if stats is not None:
                parts.append(stats.reshape(-1))
            flat = torch.cat(parts)
            host = torch.empty(flat.numel(), dtype=torch.float64, pin_memory=True)
            host.copy_(flat, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        b = Batch(reqs, ld, stream, ev, host, self.batch_bytes(ld, nmat), keep=[H, S, w, edges, info, flat, ws])
        if self.keep_outputs:
            b.dos, b.w, b.plan = dos, w, plan
        return b

    def _finish(self, b: Batch, results: Dict[int, dict], done_t: Dict[int, float], now: float):
        self.batch_log.append((b.ld, sum(len(r.inp.kpoints) for r in b.reqs), len(b.reqs), b.t_launch, now))
        h = b.host.numpy()
        ns = len(b.reqs)
        edges = h[:ns * self.nf].reshape(ns, self.nf)
        nmat = sum(len(r.inp.kpoints) for r in b.reqs)
        info = h[ns * self.nf: ns * self.nf + 2 * nmat]
        stats = h[ns * self.nf + 2 * nmat:].reshape(ns, 3) if self.grid is not None else None
        m0 = 0
        for s, r in enumerate(b.reqs):
            nk = len(r.inp.kpoints)
            bad = np.nonzero(info[m0:m0 + nk])[0]
            bad_t = np.nonzero(info[nmat + m0:nmat + m0 + nk])[0]
            res = engine.edges_dict(edges[s])
            res["gap_eV"] = 0.0 if res["metal"] > 0 else res["cbm"] - res["vbm"]
            res["ok"] = len(bad) == 0 and len(bad_t) == 0
            if len(bad):
                res["error"] = f"S(k) not positive definite at k-point index {int(bad[0])}"
            elif len(bad_t):
                res["error"] = engine.tridiag_failure(int(info[nmat + m0 + bad_t[0]]), int(bad_t[0]))
            elif r.n_orb < b.ld and engine.reaches_padding(res):
                res["ok"] = False
                res["error"] = (f"the spectrum reaches the identity-padding level ({engine.PAD_H} Ha) of "
                                f"bucket {b.ld}")
            if stats is not None:
                res["dos_at_fermi"], res["dos_gap_center_eV"], res["dos_total_states"] = stats[s]
            if self.keep_outputs:
                res["eigenvalues"] = b.w[b.plan.mat0[s]:b.plan.mat0[s] + nk, :r.n_orb]
                if b.dos is not None:
                    res["dos"] = b.dos[s]
            results[r.idx] = res
            done_t[r.idx] = now
            m0 += nk

    # ------------------------------------------------------------------
    def replay(self, inputs: Sequence[engine.StructureInput], arrivals: Sequence[float]):
        """Serve ``inputs`` arriving at ``arrivals`` (seconds from start).

        Returns dict with wall time, per-request latency (arrival -> result on
        the host) and per-request results."""
        self.batch_log = []   # (ld, matrices, structures, launch_s, done_s)
        order = np.argsort(np.asarray(arrivals), kind="stable")
        pending = deque(int(i) for i in order)
        queues: Dict[int, deque] = {}
        inflight: List[list] = []      # [future, nbytes, t_submit]
        used = 0
        results: Dict[int, dict] = {}
        done_t: Dict[int, float] = {}
        t0 = time.perf_counter()
        n = len(inputs)
        while len(done_t) < n:
            now = time.perf_counter() - t0
            while pending and arrivals[pending[0]] <= now:
                i = pending.popleft()
                inp = inputs[i]
                no = self._orbitals(inp)
                ld = self.bucket(no)
                # batches are formed per (bucket, real): a real batch skips the
                # imaginary arithmetic of the solve
                real = self.real_symmetric and engine.real_kpoints(inp.kpoints)
                queues.setdefault((ld, real), deque()).append(Request(i, inp, no, float(arrivals[i]), ld, real))
            progressed = False
            for item in list(inflight):
                fut, nbytes, t_sub = item
                if not fut.done():
                    continue
                try:
                    b = fut.result()
                except Exception as exc:   # this batch failed to launch: fail its requests only
                    self._pending[fut.stream_index] -= 1
                    self._pending_cost[fut.stream_index] -= fut.cost
                    now_f = time.perf_counter() - t0
                    for r in fut.reqs:
                        results[r.idx] = {"ok": False, "error": f"{type(exc).__name__}: {exc}"}
                        done_t[r.idx] = now_f
                    inflight.remove(item)
                    used -= nbytes
                    progressed = True
                    continue
                if b.event.query():
                    self._pending[fut.stream_index] -= 1
                    self._pending_cost[fut.stream_index] -= fut.cost
                    b.t_launch = t_sub
                    self._finish(b, results, done_t, time.perf_counter() - t0)
                    inflight.remove(item)
                    used -= nbytes
                    progressed = True
            # launch: oldest bucket first while slots and budget allow
            while len(inflight) < self.max_inflight:
                live = [(q[0].arrival, key) for key, q in queues.items() if q]
                if not live:
                    break
                _, (ld, real) = min(live)
                q = queues[(ld, real)]
                take = []
                nmat = 0
                while q and len(take) < self.max_batch:
                    nk = len(q[0].inp.kpoints)
                    if take and nmat + nk > self.max_batch:
                        break
                    # the budget holds for the first request too while other
                    # batches are in flight (one alone on an idle device runs)
                    if (take or inflight) and used + self.batch_bytes(ld, nmat + nk) > self.budget:
                        break
                    take.append(q.popleft())
                    nmat += nk
                if not take:
                    break
                nbytes = self.batch_bytes(ld, nmat)
                inflight.append([self._submit(take, ld, real), nbytes, time.perf_counter() - t0])
                used += nbytes
                progressed = True
            if not progressed:
                if inflight:
                    time.sleep(5e-5)
                elif pending:
                    wait = arrivals[pending[0]] - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
        wall = time.perf_counter() - t0
        lat = np.array([done_t[i] - float(arrivals[i]) for i in range(n)])
        return {"wall_s": wall, "latency_s": lat, "results": results}
