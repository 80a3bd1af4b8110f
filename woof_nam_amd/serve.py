"""Served radiation on one MI355X: continuous batching of radiation requests.

Requests of uneven column counts (a parent grid, nests, ensemble members with their own McICA
seeds) arrive at any time.  The server

* admits a request when its device memory (inputs + outputs, FP64) fits the HBM budget;
* cuts admitted requests into column tiles and keeps ``slots`` tiles in flight at once, each slot
  on its own HIP stream with its own work space, so tiles of different requests overlap on the
  device; a tile is handed to the next free slot as soon as one frees up (requests that arrive
  while others run join the tile queue at once: continuous batching at tile granularity);
* runs each tile as one replay of a HIP graph captured once per slot (the kernels of a tile:
  column scalars, day list, 16 LW and 14 SW band kernels spread over the context's band streams,
  post; with precision "surrogate" also the layer features and the network's inference kernel,
  with the table fallback of columns outside the training envelope) over the slot's staging buffers.
  The slots are the library's (wr_slot_*): a tile is one library call that gathers it into the
  staging buffers with the request's seed and the tile's column count (one kernel), replays the
  graph, scatters the outputs back (one kernel) and records the completion event, all on the
  slot's stream (a short last tile costs only its own columns); no PyTorch operation is on the
  per-tile path;
* records per-request latency (arrival to last tile done) for p50/p99.

Every column's result depends only on its own inputs, its column id and its request's seed, so a
request's outputs are the same alone and in any mix (asserted by the tests).
"""
from __future__ import annotations

import collections
import json
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np
import torch

import ctypes as C

from . import _lib, abi
from .radiation import Radiation, _colstruct, _dptr, input_shapes

BYTES_PER_COLUMN = (6 * 60 + 61 + 4 + 2 * 60 + 4 * 62) * 8   # FP64 inputs + outputs at 60 layers


@dataclass
class Request:
    """One radiation call: device columns (float64 [k][ncol], may be shared), seed, column ids."""
    rid: int
    kind: str
    cols: dict
    seed: int = 0
    col_id: torch.Tensor | None = None
    arrival: float = 0.0
    out: dict | None = None
    admitted: float | None = None
    done: float | None = None
    tiles_left: int = 0
    nbytes: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def ncol(self) -> int:
        return int(self.cols["t_lay"].shape[1])

    @property
    def latency(self) -> float | None:
        return None if self.done is None else self.done - self.arrival


PRECISION_CODE = {"fp64": 0, "mixed": 1, "surrogate": 2}


class _Slot:
    """A library slot (wr_slot_*): the context, its stream, staging buffers, the captured graph of a
    whole-tile call and the completion event all live in the library; a tile costs one call here."""

    def __init__(self, device, tile, nlay, precision, params, graphs, filler, surrogate=None):
        self.rad = Radiation(device, max_cols=tile)
        if surrogate is not None:        # the network's inference is captured into the slot's graph
            self.rad.attach_surrogate(surrogate)
        self.tile = tile
        fill = {k: filler[k][..., :min(tile, int(filler[k].shape[-1]))].contiguous() for k in abi.INPUT_FIELDS}
        nf = int(fill["t_lay"].shape[1])
        ci = _colstruct(fill, nf, nlay, self.rad.device)
        h = C.c_void_p()
        _lib.check(_lib.load().wr_slot_create(C.byref(h), self.rad._h, PRECISION_CODE[precision], C.byref(params),
                                              int(bool(graphs)), C.byref(ci)), "wr_slot_create")
        self._h = h                      # the filler was copied and the graph captured inside the call
        self.work = None

    def launch(self, req: Request, c0: int, n: int):
        abi_args = req.meta.get("_abi")
        if abi_args is None:             # the request's pointers, built once per request
            cols = dict(req.cols)
            if req.col_id is not None:
                cols["col_id"] = req.col_id
            ci = _colstruct(cols, req.ncol, int(req.cols["t_lay"].shape[0]), self.rad.device)
            co = abi.WrOutputs(*[_dptr(req.out[k]) for k in abi.OUTPUT_FIELDS])
            abi_args = req.meta["_abi"] = (ci, co)
        ci, co = abi_args
        rc = _lib.load().wr_slot_launch(self._h, C.byref(ci), C.byref(co), int(c0), int(n),
                                        int(req.seed) & 0xFFFFFFFFFFFFFFFF)
        _lib.check(rc, "wr_slot_launch")
        self.work = req

    def ready(self) -> bool:
        if self.work is None:
            return True
        r = _lib.load().wr_slot_query(self._h)
        if r < 0:
            _lib.check(-r, "wr_slot_query")
        return r == 1

    def close(self):
        if self._h:
            _lib.load().wr_slot_destroy(self._h)
            self._h = None
        self.rad.close()


class Server:
    """Continuous-batching radiation server on one GPU (see the module docstring)."""

    def __init__(self, device: int = 0, tile: int = 65536, slots: int = 4, budget_gb: float = 200.0,
                 precision: str = "mixed", graphs: bool = True, nlay: int = 60, params=None, filler=None,
                 surrogate=None):
        """precision: "fp64", "mixed" or "surrogate"; the last needs ``surrogate``, a trained
        surrogate.MLP that every slot attaches (its tensors must outlive the server)."""
        if precision == "surrogate" and surrogate is None:
            raise ValueError("serve: precision 'surrogate' needs the trained network (surrogate=...)")
        self.device = torch.device("cuda", device)
        self.tile, self.nslots, self.budget = int(tile), int(slots), float(budget_gb) * 1e9
        self.precision, self.nlay = precision, nlay
        self.params = params or Radiation.default_params()
        if filler is None:               # a valid column set, repeated over each slot's staging tile
            from . import gen
            cols = gen.columns(min(self.tile, 8192), seed=12345)
            filler = {k: torch.from_numpy(np.ascontiguousarray(v)).to(self.device) for k, v in cols.items()}
        self.slots = [_Slot(self.device, self.tile, nlay, precision, self.params, graphs, filler,
                            surrogate if precision == "surrogate" else None)
                      for _ in range(self.nslots)]
        # the admission budget never exceeds what the device has left once the slots exist, less a
        # reserve for the runtime (queue scratch, graph and internal allocations): outputs admitted
        # beyond free memory starve the runtime, which then aborts the process
        free, total = torch.cuda.mem_get_info(self.device)
        self.reserve = max(8e9, 0.03 * total)
        self.budget = min(self.budget, max(0.0, free - self.reserve))
        self.pending = collections.deque()     # submitted, not admitted
        self.tiles = collections.deque()       # (request, c0, n) of admitted requests
        self.inflight_bytes = 0
        self.completed = []
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, req: Request) -> None:
        if req.cols["t_lay"].shape[0] != self.nlay:
            raise ValueError(f"request {req.rid}: {req.cols['t_lay'].shape[0]} layers, server built for {self.nlay}")
        req.nbytes = req.ncol * BYTES_PER_COLUMN
        if req.nbytes > self.budget:
            raise MemoryError(f"request {req.rid} needs {req.nbytes / 1e9:.1f} GB, above the HBM budget")
        self.pending.append(req)

    def _admit(self):
        while self.pending and self.inflight_bytes + self.pending[0].nbytes <= self.budget:
            req = self.pending.popleft()
            req.out = {k: torch.empty(s, dtype=torch.float64, device=self.device)
                       for k, s in abi.output_shapes(req.ncol, self.nlay).items()}
            req.admitted = self.now()
            self.inflight_bytes += req.nbytes
            n = req.ncol
            req.tiles_left = (n + self.tile - 1) // self.tile
            for c0 in range(0, n, self.tile):
                self.tiles.append((req, c0, min(self.tile, n - c0)))
            if n == 0:
                self._finish(req)

    def _finish(self, req):
        req.done = self.now()
        self.inflight_bytes -= req.nbytes
        self.completed.append(req)

    def poll(self) -> int:
        """Retire finished tiles, admit what fits, start tiles on free slots.  Returns tiles started."""
        for s in self.slots:
            if s.work is not None and s.ready():
                req = s.work
                s.work = None
                req.tiles_left -= 1
                if req.tiles_left == 0:
                    self._finish(req)
        self._admit()
        started = 0
        for s in self.slots:
            if s.work is None and self.tiles:
                req, c0, n = self.tiles.popleft()
                s.launch(req, c0, n)
                started += 1
        return started

    def idle(self) -> bool:
        return not self.pending and not self.tiles and all(s.work is None for s in self.slots)

    def drain(self, sleep: float = 2e-4) -> None:
        while not self.idle():
            if self.poll() == 0:
                time.sleep(sleep)

    def close(self):
        for s in self.slots:
            s.close()


def latency_stats(reqs) -> dict:
    lat = np.array([r.latency for r in reqs if r.latency is not None])
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None}
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}


def poisson_run(server: Server, kinds: dict, rate: float, duration: float, seed: int = 0,
                members: int = 8, keep_outputs: bool = False, progress: float = 0.0) -> dict:
    """Drive the server with Poisson arrivals for `duration` seconds (real time).

    kinds: name -> (cols dict of device tensors, col_id tensor or None).  An arrival picks a kind
    uniformly; the kind "member" stands for one ensemble member of the parent with its own seed
    (members seeds 1..members).  progress > 0: a line on stderr every `progress` seconds (arrived,
    finished, columns done).  Returns throughput and latency figures.
    """
    rng = np.random.default_rng(seed)
    names = list(kinds)
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        arrivals.append((t, names[rng.integers(len(names))], int(rng.integers(1, members + 1))))
    server.t0 = time.perf_counter()
    reqs = []
    i = 0
    t_report = progress
    while i < len(arrivals) or not server.idle():
        now = server.now()
        if progress > 0 and now >= t_report:
            done = [r for r in reqs if r.done is not None]
            print(f"# served {now:.0f} s: {i}/{len(arrivals)} arrived, {len(done)} finished, "
                  f"{sum(r.ncol for r in done) / 1e6:.1f} M columns", file=sys.stderr, flush=True)
            t_report += progress
        while i < len(arrivals) and arrivals[i][0] <= now:
            ta, kind, member = arrivals[i]
            cols, ids = kinds[kind]
            r = Request(rid=i, kind=kind, cols=cols, col_id=ids, arrival=ta,
                        seed=member if kind == "member" else 0)
            server.submit(r)
            reqs.append(r)
            i += 1
        if server.poll() == 0:
            time.sleep(2e-4)
        if not keep_outputs:              # the consumer takes finished outputs away
            for r in server.completed:
                r.out = None
    torch.cuda.synchronize(server.device)
    ncols = sum(r.ncol for r in reqs)
    t_end = max(r.done for r in reqs) if reqs else 0.0
    res = {"requests": len(reqs), "columns": ncols, "wall_s": t_end,
           "columns_per_s": ncols / t_end if t_end > 0 else 0.0,
           "offered_columns_per_s": ncols / duration if duration > 0 else 0.0, **latency_stats(reqs)}
    res["by_kind"] = {k: {"n": sum(1 for r in reqs if r.kind == k), **latency_stats([r for r in reqs if r.kind == k])}
                      for k in names}
    return res


def autotune(device: int = 0, tiles=(32768, 65536, 131072), slots=(1, 2, 4), ncol: int = 262144,
             precision: str = "mixed", cache: str | None = None, nreq: int = 4, shapes=None,
             surrogate=None) -> dict:
    """Pick the tile size and slot count with the highest columns/s on a fixed batch of requests.

    The candidates are every (tile, slots) of `tiles` x `slots`, or the explicit list `shapes` of
    (tile, slots) pairs when given (to leave out shapes whose work space would not fit beside the
    caller's).  The batch is `ncol` generated columns submitted at once as `nreq` requests of uneven
    size (a queue deep enough to keep every slot busy, as under load).  Results are cached in `cache`
    (JSON) keyed by device name, precision, batch and candidates when a path is given.
    """
    from . import gen
    name = torch.cuda.get_device_name(device)
    cand = [(int(t), int(ns)) for t, ns in shapes] if shapes is not None else [(t, ns) for t in tiles for ns in slots]
    key = f"{name}|{precision}|{ncol}|{nreq}|{cand}"
    if cache and os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
        if key in db:
            return db[key]
    cols = gen.columns(ncol, seed=7)
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to(f"cuda:{device}") for k, v in cols.items()}
    w = np.array([1 + (3 * j) % 5 for j in range(nreq)], dtype=np.float64)   # uneven request sizes
    cuts = [0] + [int(x) for x in np.round(np.cumsum(w) / w.sum() * ncol)]
    results = []
    for tile, ns in cand:
        srv = Server(device, tile=tile, slots=ns, precision=precision, surrogate=surrogate, filler=dev)
        best = None
        for rep in range(2):
            reqs = [Request(rid=j, kind="tune", cols={k: v[..., cuts[j]:cuts[j + 1]].contiguous() for k, v in dev.items()},
                            seed=j) for j in range(nreq)]
            torch.cuda.synchronize(device)
            t0 = time.perf_counter()
            for r in reqs:
                srv.submit(r)
            srv.drain()
            torch.cuda.synchronize(device)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        srv.close()
        results.append({"tile": tile, "slots": ns, "columns_per_s": ncol / best})
    top = max(results, key=lambda r: r["columns_per_s"])
    out = {"device": name, "precision": precision, "best": top, "all": results}
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
        db[key] = out
        os.makedirs(os.path.dirname(cache) or ".", exist_ok=True)
        with open(cache, "w") as f:
            json.dump(db, f, indent=1)
    return out
