"""Solve serving: diagonalisation requests at many (g, h) parameter points,
on several lattices and spinon sectors, answered by continuously batched
GPU eigensolves.

Model.  A request names a space (a lattice + spinon sector, i.e. one basis)
and a parameter point of the ring-exchange Hamiltonian with a local Zeeman
field,
    H(g, h) = sum_rings g_sl (O_r + O_r^dagger) - sum_i h_sl(i) S^z_i ,
the terms of build_hamiltonian (hamiltonian_setup.hpp:105-121) plus the
local-axis field.  The terms are built once per space with coefficient slots
(ring sublattice s -> slot s, field on sublattice s -> slot 4 + s), so every
parameter point is just 8 numbers.

Engine.  `Engine` wraps qsi_engine_t: up to `max_cols` requests on one space
run side by side, one column of a block apply each, advancing one Lanczos
step per `step()`.  Requests join free columns between steps and leave when
converged (continuous batching).  A request's result does not depend on what
else runs in the engine.

Server.  `SolveServer` keeps spaces resident on the device under an HBM
budget (basis + cached connectivity + terms + engine buffers), uploads a space
when a request for it arrives and the budget allows, evicts idle spaces
least-recently-used, and keeps every engine with work stepping on its own
stream at its own pace (a small space is not held to the step time of a large
one).  Latency is time from a request's arrival to its result.
"""
import ctypes as C
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check
from .device import GpuBasis, GpuOp, as_lohi
from .lattice import add_local_field, add_ring_terms, g_vals, get_loc_z
from .operators import SymbolicOpSum

N_SLOTS = 8


def slotted_hamiltonian(jdata):
    """Ring terms on slots 0-3 and the local field on slots 4-7."""
    H = add_ring_terms(SymbolicOpSum(), jdata, [0.0] * 4, slots=True)
    add_local_field(H, jdata, [0.0, 0.0, 0.0], slots=True)
    return H


def point_coeffs(g, h):
    """Slot coefficients of H(g, h): g_s on slot s, -h_s / 2 on slot 4 + s
    (S^z = Z / 2)."""
    g = [float(v) for v in g]
    h = [float(v) for v in h]
    if len(g) != 4 or len(h) != 4:
        raise ValueError("g and h need one value per sublattice (4 each)")
    return np.array(g + [-0.5 * v for v in h], dtype=np.float64)


def jpm_point(Jpm, B):
    """(g, h) of the effective model at transverse exchange Jpm and field B:
    g = g_vals(Jpm, B), h_s = B . z_s."""
    loc = get_loc_z()
    Bv = np.asarray(B, dtype=float)
    return g_vals(Jpm, B), [float(Bv @ loc[s]) for s in range(4)]


@dataclass
class Finished:
    id: int
    evals: np.ndarray
    nconv: int
    steps: int


class Engine:
    """One qsi_engine_t on one GpuOp (see include/qsi_gpu.h)."""

    def __init__(self, op: GpuOp, max_cols=16, nev=1, ncv=20, tol=1e-10, max_restarts=500,
                 cache_connectivity=True, autotune=True):
        lib = _lib.load()
        if cache_connectivity:
            op.cache_connectivity(True)
        if autotune:
            # block-apply configuration for this operator at max_cols columns
            self.tuned = op.autotune(max_cols, reps=2)
        h = C.c_void_p()
        check(lib.qsi_engine_create(op.handle, max_cols, nev, ncv, tol, max_restarts,
                                    C.byref(h)), "qsi_engine_create")
        self._h = h
        self.op = op
        self.nev = nev
        self.max_cols = max_cols

    @staticmethod
    def footprint(dim, max_cols, ncv):
        return int(_lib.load().qsi_engine_footprint(int(dim), int(max_cols), int(ncv)))

    def submit(self, coeffs, seed):
        c = None
        if self.op.n_slots:
            a = np.ascontiguousarray(coeffs, dtype=np.float64)
            if a.shape != (self.op.n_slots,):
                raise ValueError(f"coeffs must have {self.op.n_slots} entries")
            c = a.ctypes.data_as(C.POINTER(C.c_double))
        rid = C.c_int64()
        check(_lib.load().qsi_engine_submit(self._h, c, int(seed), C.byref(rid)),
              "qsi_engine_submit")
        return rid.value

    def step(self):
        """Keep up to two steps enqueued and finish the oldest (qsi_engine_step)."""
        n = C.c_int()
        check(_lib.load().qsi_engine_step(self._h, C.byref(n)), "qsi_engine_step")
        return n.value

    def step_begin(self):
        """Admit and enqueue one step (up to two may be in flight); False when
        no column can step."""
        n = C.c_int()
        check(_lib.load().qsi_engine_step_begin(self._h, C.byref(n)), "qsi_engine_step_begin")
        return bool(n.value)

    @property
    def inflight(self):
        """Steps enqueued and not yet finished (0, 1 or 2)."""
        return int(_lib.load().qsi_engine_inflight(self._h))

    def ready(self):
        """True when step_end would not wait."""
        rc = _lib.load().qsi_engine_step_ready(self._h)
        if rc < 0:
            check(rc, "qsi_engine_step_ready")
        return rc == 1

    def step_end(self):
        """Finish the oldest enqueued step."""
        n = C.c_int()
        check(_lib.load().qsi_engine_step_end(self._h, C.byref(n)), "qsi_engine_step_end")
        return n.value

    def poll(self):
        out = []
        lib = _lib.load()
        ev = (C.c_double * self.nev)()
        rid, nc, st = C.c_int64(), C.c_int(), C.c_int()
        while lib.qsi_engine_poll(self._h, C.byref(rid), ev, C.byref(nc), C.byref(st)):
            out.append(Finished(rid.value, np.array(list(ev)), nc.value, st.value))
        return out

    @property
    def active(self):
        return int(_lib.load().qsi_engine_active(self._h))

    @property
    def queued(self):
        return int(_lib.load().qsi_engine_queued(self._h))

    def busy(self):
        return self.inflight > 0 or self.active + self.queued > 0

    def run_all(self):
        """Step until every submitted request finished; returns them."""
        done = []
        while self.busy():
            self.step()
            done += self.poll()
        return done

    def close(self):
        if self._h:
            _lib.load().qsi_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Space:
    """A lattice + sector whose basis can be made device-resident."""
    key: str
    jdata: dict
    states: np.ndarray  # (D, 2) uint64
    n_spins: int
    basis: object = None
    op: object = None
    engine: object = None
    last_used: float = 0.0
    _padded_nnz: int = -1

    @property
    def dim(self):
        return int(self.states.shape[0])

    def padded_nnz(self):
        """Connectivity entries of the space's operator with every row padded
        to a multiple of 4 (what the cache allocates at most): each state's
        count of connecting off-diagonal terms, counted on the host once."""
        if self._padded_nnz < 0:
            H = slotted_hamiltonian(self.jdata)
            lo = np.ascontiguousarray(self.states[:, 0])
            hi = np.ascontiguousarray(self.states[:, 1])
            f = np.zeros(self.dim, dtype=np.int32)
            M = (1 << 64) - 1
            for (_, o, _) in H.off_diag_terms:
                m = o.up | o.down
                v = o.up ^ (o.X & m)
                f += ((lo & np.uint64(m & M)) == np.uint64(v & M)) & \
                     ((hi & np.uint64(m >> 64)) == np.uint64(v >> 64))
            self._padded_nnz = int(((f + 3) // 4 * 4).sum())
        return self._padded_nnz

    def footprint(self, max_cols, ncv):
        """Device bytes when resident: basis words + bucket table (2^K + 1
        entries, K at most ceil(log2 D) + 7, 30 and n_spins) + cached
        connectivity (5 bytes per entry, rows padded to a multiple of 4, + 8
        per row) + terms + engine buffers + the block autotuner's two D x
        max_cols scratch blocks (held while the engine is created)."""
        words = 1 if self.n_spins <= 64 else 2
        lg = max(0, int(self.dim - 1).bit_length())
        kbits = max(1, min(lg + 7, 30, self.n_spins))
        basis = self.dim * 8 * words + 4 * ((1 << kbits) + 1)
        conn = 8 * (self.dim + 1) + 5 * self.padded_nnz() + 64
        terms = 1 << 20  # term tables are a few KB; a MiB covers them
        tune = 2 * 8 * self.dim * max_cols
        return basis + conn + terms + tune + Engine.footprint(self.dim, max_cols, ncv)


@dataclass
class Request:
    space: str
    coeffs: np.ndarray
    seed: int
    arrival: float = 0.0
    id: int = -1
    engine_id: int = -1
    result: object = None
    finish: float = 0.0
    tag: object = None


@dataclass
class ServeStats:
    n: int = 0
    wall_s: float = 0.0
    latencies_s: list = field(default_factory=list)
    steps: int = 0
    evictions: int = 0
    uploads: int = 0

    def summary(self):
        lat = np.array(self.latencies_s) if self.latencies_s else np.zeros(1)
        return {
            "solves": self.n,
            "wall_s": self.wall_s,
            "solves_per_s": self.n / self.wall_s if self.wall_s > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "engine_steps": self.steps,
            "uploads": self.uploads,
            "evictions": self.evictions,
        }


class SolveServer:
    """Request loop over several spaces with HBM-footprint admission."""

    def __init__(self, hbm_budget_bytes, max_cols=16, nev=1, ncv=20, tol=1e-10,
                 max_restarts=500):
        self.budget = int(hbm_budget_bytes)
        self.max_cols, self.nev, self.ncv, self.tol = max_cols, nev, ncv, tol
        self.max_restarts = max_restarts
        self.spaces = OrderedDict()
        self.pending = deque()     # arrived, waiting for their space to be resident
        self.inflight = {}         # (space, engine id) -> Request
        self.stats = ServeStats()
        self._seq = 0

    def add_space(self, key, jdata, states, n_spins):
        if self.spaces.get(key) is not None:
            raise ValueError(f"space {key!r} already registered")
        sp = Space(key, jdata, as_lohi(states), int(n_spins))
        if sp.footprint(self.max_cols, self.ncv) > self.budget:
            raise ValueError(f"space {key!r} alone exceeds the HBM budget")
        self.spaces[key] = sp
        return sp

    # -- residency
    def resident_bytes(self):
        return sum(s.footprint(self.max_cols, self.ncv) for s in self.spaces.values()
                   if s.engine is not None)

    def _upload(self, sp):
        sp.basis = GpuBasis(sp.states, sp.n_spins)
        sp.op = GpuOp(sp.basis, slotted_hamiltonian(sp.jdata))
        sp.engine = Engine(sp.op, self.max_cols, self.nev, self.ncv, self.tol,
                           self.max_restarts)
        self.stats.uploads += 1

    def _evict(self, sp):
        sp.engine.close()
        sp.op.close()
        sp.basis.close()
        sp.engine = sp.op = sp.basis = None
        self.stats.evictions += 1

    def _make_resident(self, sp):
        """True when sp is (or has been made) resident within the budget."""
        if sp.engine is not None:
            return True
        need = sp.footprint(self.max_cols, self.ncv)
        idle = sorted((s for s in self.spaces.values()
                       if s.engine is not None and not s.engine.busy()),
                      key=lambda s: s.last_used)
        while self.resident_bytes() + need > self.budget and idle:
            self._evict(idle.pop(0))
        if self.resident_bytes() + need > self.budget:
            return False
        self._upload(sp)
        return True

    # -- request flow
    def submit(self, req: Request):
        if req.space not in self.spaces:
            raise KeyError(f"unknown space {req.space!r}")
        req.id = self._seq
        self._seq += 1
        self.pending.append(req)
        return req.id

    def _admit(self):
        """First come, first served: a request whose space cannot be made
        resident yet blocks the ones behind it, so the resident engines drain,
        become evictable and the head gets in (no request starves)."""
        while self.pending:
            r = self.pending[0]
            sp = self.spaces[r.space]
            if not self._make_resident(sp):
                break
            self.pending.popleft()
            r.engine_id = sp.engine.submit(r.coeffs, r.seed)
            self.inflight[(r.space, r.engine_id)] = r

    def step(self, now=None):
        """Admit pending requests and advance every engine whose previous
        step has completed (never waits); returns the requests that
        finished."""
        self._admit()
        done = []
        # each engine runs on its own stream at its own pace: finish a step
        # whose work has completed, then enqueue the next one; engines whose
        # step is still running are left alone (nothing here blocks)
        # largest basis first: its steps are the longest, so they are enqueued
        # before the small engines' ones in each pass
        for sp in sorted(self.spaces.values(), key=lambda x: -x.dim):
            eng = sp.engine
            if eng is None:
                continue
            # each engine keeps up to two steps enqueued: the next one runs
            # while the host finishes the previous
            if eng.inflight and eng.ready():
                eng.step_end()
                sp.last_used = time.perf_counter()
                self.stats.steps += 1
            while eng.inflight < 2 and eng.active + eng.queued > 0:
                if not eng.step_begin():
                    break
            for f in eng.poll():
                r = self.inflight.pop((sp.key, f.id))
                r.result = f
                r.finish = time.perf_counter() if now is None else now
                done.append(r)
        return done

    def busy(self):
        return bool(self.pending or self.inflight)

    def run(self, requests, arrivals=None):
        """Serve `requests`; arrivals[i] (seconds after the start, or None for
        all at once) is when request i arrives.  Returns the requests with
        results and fills self.stats."""
        order = np.argsort(arrivals, kind="stable") if arrivals is not None else \
            np.arange(len(requests))
        queue = deque(int(i) for i in order)
        t0 = time.perf_counter()
        out = []
        while queue or self.busy():
            now = time.perf_counter() - t0
            while queue and (arrivals is None or arrivals[queue[0]] <= now):
                i = queue.popleft()
                r = requests[i]
                r.arrival = t0 + (arrivals[i] if arrivals is not None else 0.0)
                self.submit(r)
            if not self.busy():
                # idle until the next arrival
                time.sleep(max(0.0, min(arrivals[queue[0]] - now, 0.01)))
                continue
            for r in self.step():
                self.stats.latencies_s.append(r.finish - r.arrival)
                out.append(r)
        self.stats.n += len(out)
        self.stats.wall_s += time.perf_counter() - t0
        return out

    def close(self):
        for sp in self.spaces.values():
            if sp.engine is not None:
                self._evict(sp)
