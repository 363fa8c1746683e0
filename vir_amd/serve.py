"""Serving SIRT reconstruction requests on one GPU.

A request names a geometry (2-D fan-beam or 3-D cone-beam), a size, a view
count, an iteration count and the seed of its synthetic object.  The server

  * admits requests first-come first-served while the HBM budget and the
    slot limit allow (a request larger than the whole budget is rejected;
    nothing ever waits on a request behind it that can never fit);
  * sets an admitted request up on a stream of its own: geometry on the
    device, the measured sinogram (synthesised on the device: phantom, then
    one forward projection), the normalisation sums W_sino = A 1 and
    W_img = A^T 1, and one SIRT iteration captured into a HIP graph;
  * enqueues all of its iterations as graph replays, then an asynchronous
    copy of the result into pinned host memory, and an event;
  * keeps polling: a request is complete when its event has fired; its
    memory returns to the budget and waiting requests are admitted.

Work of different requests overlaps on the device (one stream per slot);
requests that arrive while others run are admitted as soon as memory and a
slot free up (continuous batching at request granularity).  Every kernel on
the path is deterministic, so a request's result does not depend on what
else is running.

Requests of the same geometry (kind, size, views) share one projector:
the backend keeps the projector and its normalisation sums W_sino, W_img in
an LRU cache bounded in bytes, and the server (group=True) admits waiting
requests of the geometry it has just admitted right after it.  Jobs that
share a projector run on one stream (the projector's backprojection
workspace is per projector).  With batch=True the server admits up to
backend.max_batch waiting requests of one batch key (cone beam: geometry,
iterations and step) together, and the backend runs them as one batched
SIRT whose projections trace the rays once for the whole group; each
request's result is still bit-identical to its run alone.

Fan-beam requests use the distance-driven projector with an equiangular
detector (sys_mat/dd.py, the reference's fan geometry on an arc); cone-beam
requests use the on-the-fly Siddon projector with the brick-gather
backprojector (sys_mat/siddon.py) on rays generated on the device
(proj_geom.py geom_circular).  The recurrence is the reference's SIRT
(sys_mat/sirt_dd.py:37-76, tests/Test_SIRT/sirt_test.py:21-40) from x0 = 0.

The scheduling core (Server) is independent of the GPU: it drives a backend
with estimate/start/poll/release, so it runs against a simulated backend
on the CPU as well.
"""
import collections
import dataclasses
import math
import time

import numpy as np

FAN_NU_RATIO = 888.0 / 512.0


@dataclasses.dataclass
class Request:
    rid: int
    kind: str              # "fan" or "cone"
    n: int                 # image n x n (fan) or volume n^3 (cone)
    views: int
    iters: int = 20
    arrival: float = 0.0   # seconds after the start of the trace
    seed: int = 0
    gamma: float = 1.0


@dataclasses.dataclass
class Result:
    rid: int
    ok: bool
    latency: float = math.nan      # arrival -> result on the host (s)
    queued: float = math.nan       # arrival -> admission (s)
    x: object = None               # host float32 array
    error: str = ""


# --------------------------------------------------------------- geometry

def fan_geometry(n, views):
    """2-D fan beam for an n x n image of unit pixels: the 512 / 720-view /
    888-bin benchmark scaled with n (DSO = 1000 n/512, DSD = 1500 n/512),
    round(1.734 n) equiangular bins covering the image's circumscribed
    circle, views over [0, 2 pi)."""
    DSO, DSD = 1000.0 * n / 512.0, 1500.0 * n / 512.0
    nu = int(round(FAN_NU_RATIO * n))
    gmax = math.asin((n / math.sqrt(2.0)) / DSO)
    du = 2.0 * DSD * gmax * 1.02 / nu
    ang = (np.arange(views) * (2 * np.pi / views)).astype(np.float32)
    return dict(img_shape=(n, n), ang=ang, nu=nu, DSO=DSO, DSD=DSD, du=du)


def cone_geometry(n, views):
    """3-D cone beam for an n^3 volume of unit voxels: source-isocentre 4n,
    source-detector 8n (magnification 2), round(1.5 n)^2 detector of pitch 2,
    views over [0, 2 pi)."""
    m = int(round(1.5 * n))
    pitch = 2.0
    dets = (np.arange(m) - (m - 1) / 2.0) * pitch
    views_a = np.arange(views) * (2 * np.pi / views)
    return dict(img_shape=(n, n, n), views=views_a, dets_y=dets, dets_z=dets,
                src_iso=4.0 * n, det_iso=4.0 * n, m=m)


# ------------------------------------------------------------ the server

class Server:
    """Admission and completion loop over a backend.

    backend.estimate(req) -> bytes; backend.start(req, slot) -> handle;
    handle.poll() -> bool (complete); handle.result() -> host array;
    handle.release().  Slots index the backend's streams."""

    def __init__(self, backend, hbm_budget, max_active=16, clock=time.perf_counter,
                 idle_sleep=2e-4, group=False, batch=False):
        """group: after admitting a request, admit the waiting requests
        with the same backend.geometry_key next (out of arrival order).
        batch: admit waiting requests with the admitted one's
        backend.batch_key (not None) together, up to backend.max_batch, as
        one backend.start_batch call (each keeps its own slot and budget
        share)."""
        self.backend = backend
        self.group_key = getattr(backend, "geometry_key", None) if group else None
        self.batch_key = getattr(backend, "batch_key", None) if batch else None
        self.max_batch = int(getattr(backend, "max_batch", 1))
        self.grouped = 0
        self.batched = 0
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.clock = clock
        self.idle_sleep = idle_sleep
        self.in_use = 0
        self.max_in_use = 0
        self.max_concurrent = 0

    def run(self, requests, keep_results=True, progress=None):
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        waiting = collections.deque()
        active = {}   # slot -> (req, handle, est, t_admit)
        free_slots = list(range(self.max_active))[::-1]
        results = {}
        t0 = self.clock()
        while pending or waiting or active:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            progressed = False
            # completions first: they return memory and slots
            for slot in list(active):
                req, h, est, t_adm = active[slot]
                if h.poll():
                    t_done = self.clock() - t0
                    x = h.result() if keep_results else None
                    h.release()
                    del active[slot]
                    free_slots.append(slot)
                    self.in_use -= est
                    results[req.rid] = Result(req.rid, True, t_done - req.arrival,
                                              t_adm - req.arrival, x)
                    progressed = True
                    if progress:
                        progress(results[req.rid])
            # FIFO admission
            while waiting and free_slots:
                req = waiting[0]
                est = int(self.backend.estimate(req))
                if est > self.budget:
                    waiting.popleft()
                    results[req.rid] = Result(req.rid, False, error=(
                        f"request needs {est} bytes, more than the HBM budget {self.budget}"))
                    progressed = True
                    continue
                if self.in_use + est > self.budget:
                    break
                waiting.popleft()
                slot = free_slots.pop()
                t_adm = self.clock() - t0
                members = [(req, slot, est)]
                key = self.batch_key(req) if self.batch_key is not None else None
                if key is not None:
                    need = self.in_use + est
                    for w in list(waiting):
                        if len(members) >= self.max_batch or not free_slots:
                            break
                        if self.batch_key(w) != key:
                            continue
                        e_w = int(self.backend.estimate(w))
                        if need + e_w > self.budget:
                            break
                        waiting.remove(w)
                        members.append((w, free_slots.pop(), e_w))
                        need += e_w
                try:
                    if len(members) == 1:
                        handles = [self.backend.start(req, slot)]
                    else:
                        handles = self.backend.start_batch([m[0] for m in members], [m[1] for m in members])
                except Exception as e:  # one failing request does not stop the others
                    for m in members:
                        free_slots.append(m[1])
                        results[m[0].rid] = Result(m[0].rid, False, error=f"{type(e).__name__}: {e}")
                    progressed = True
                    continue
                for (m_req, m_slot, m_est), h in zip(members, handles):
                    active[m_slot] = (m_req, h, m_est, t_adm)
                    self.in_use += m_est
                self.batched += len(members) - 1
                self.max_in_use = max(self.max_in_use, self.in_use)
                self.max_concurrent = max(self.max_concurrent, len(active))
                progressed = True
                if self.group_key is not None and waiting:
                    k = self.group_key(req)
                    same = [w for w in waiting if self.group_key(w) == k]
                    for w in same:
                        waiting.remove(w)
                    waiting.extendleft(reversed(same))
                    self.grouped += len(same)
            if not progressed:
                if not active and not waiting and pending:
                    dt = pending[0].arrival - (self.clock() - t0)
                    if dt > 0:
                        time.sleep(min(dt, 0.05))
                else:
                    time.sleep(self.idle_sleep)
        self.elapsed = self.clock() - t0
        return [results[r.rid] for r in sorted(requests, key=lambda r: r.rid)]


def summarize(results, elapsed):
    ok = [r for r in results if r.ok]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(0)
    return dict(jobs=len(results), completed=len(ok), rejected=len(results) - len(ok),
                jobs_per_s=len(ok) / elapsed if elapsed > 0 else 0.0,
                p50_s=float(np.percentile(lat, 50)) if ok else math.nan,
                p99_s=float(np.percentile(lat, 99)) if ok else math.nan,
                elapsed_s=elapsed)


def poisson_trace(n_requests, rate, seed=0, iters=20, fan_frac=0.6,
                  fan_sizes=(128, 256, 512, 1024), fan_views=(180, 720),
                  cone_sizes=(64, 128, 256), cone_views=(120, 360)):
    """Requests with exponential inter-arrival times (Poisson arrivals at
    `rate` per second): a fan_frac share of 2-D fan-beam jobs (size uniform
    over fan_sizes, views uniform in fan_views), the rest 3-D cone-beam."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    out = []
    for i in range(n_requests):
        if rng.uniform() < fan_frac:
            kind, n = "fan", int(rng.choice(fan_sizes))
            v = int(rng.integers(fan_views[0], fan_views[1] + 1))
        else:
            kind, n = "cone", int(rng.choice(cone_sizes))
            v = int(rng.integers(cone_views[0], cone_views[1] + 1))
        out.append(Request(i, kind, n, v, iters, float(t[i]), seed=int(rng.integers(1 << 30))))
    return out


# ------------------------------------------------------- the GPU backend

class _Geometry:
    """A request geometry's projector and normalisation sums, shared by the
    jobs of that geometry."""

    def __init__(self, req, stream):
        import torch
        from .geometry import circular_projections, circular_rays
        from .sys_mat.dd import DDProjector
        from .sys_mat.siddon import SiddonProjector
        with torch.cuda.stream(stream):
            if req.kind == "fan":
                g = fan_geometry(req.n, req.views)
                P = DDProjector(g["img_shape"], g["ang"], g["nu"], "fan", DSO=g["DSO"],
                                DSD=g["DSD"], du=g["du"], detector="arc")
                nbytes = 8 * req.views
            elif req.kind == "cone":
                g = cone_geometry(req.n, req.views)
                src, trg = circular_rays(g["dets_y"], g["views"], "cone", g["src_iso"],
                                         g["det_iso"], g["dets_z"])
                mats = circular_projections(g["dets_y"], g["views"], "cone", g["src_iso"],
                                            g["det_iso"], g["dets_z"])
                P = SiddonProjector(src, trg, g["img_shape"], 1.0, proj=mats)
                nbytes = P.device_bytes()
            else:
                raise ValueError(f"unknown request kind {req.kind!r}")
            img_shape, sino_shape = tuple(g["img_shape"]), tuple(P.sino_shape if req.kind == "fan"
                                                                 else P.ray_shape)
            self.W_sino = P.fp(torch.ones(img_shape, dtype=torch.float32, device="cuda"))
            self.W_img = P.bp(torch.ones(sino_shape, dtype=torch.float32, device="cuda"))
        self.P, self.g = P, g
        self.nbytes = nbytes + 4 * (self.W_sino.numel() + self.W_img.numel())
        self.stream = stream
        self.users = 0


class _Job:
    def __init__(self, req, stream, graph=True, geometry=None):
        import torch
        from .sirt import SIRT
        from .workloads import phantom_device
        self.req = req
        self.stream = stream
        G = geometry if geometry is not None else _Geometry(req, stream)
        G.users += 1
        self.geometry = G
        with torch.cuda.stream(stream):
            P, g = G.P, G.g
            n_ell = 10 if req.kind == "fan" else 20
            truth = phantom_device(g["img_shape"], n_ell, req.seed, stream=stream)
            b = P.fp(truth)            # the measurement, synthesised on the device
            del truth
            S = SIRT(P, b, gamma=req.gamma, W_sino=G.W_sino, W_img=G.W_img)
            self.P, self.S = P, S
            self.host = torch.empty(S.img_shape, dtype=torch.float32, pin_memory=True)
            if graph:
                gr = torch.cuda.CUDAGraph()
                gr.capture_begin()
                try:
                    S._iteration()
                finally:
                    # close the capture even when a launch fails, so the
                    # stream stays usable for the next request
                    gr.capture_end()
                self.graph = gr
                for _ in range(req.iters):
                    gr.replay()
            else:
                self.graph = None
                for _ in range(req.iters):
                    S._iteration()
            self.host.copy_(S.x, non_blocking=True)
            self.done = torch.cuda.Event()
            self.done.record(stream)

    def poll(self):
        return self.done.query()

    def result(self):
        self.done.synchronize()
        return self.host.numpy().copy()

    def release(self):
        self.graph = None
        self.P = self.S = None
        if self.geometry is not None:
            self.geometry.users -= 1
            self.geometry = None


class _BatchJob:
    """Requests of one geometry, iteration count and step run as one batched
    SIRT (cone beam: the Siddon projector's batched fused steps)."""

    def __init__(self, reqs, stream, graph=True, geometry=None):
        import torch
        from .sirt import SIRT
        from .workloads import phantom_device
        self.reqs = list(reqs)
        self.stream = stream
        G = geometry if geometry is not None else _Geometry(self.reqs[0], stream)
        G.users += 1
        self.geometry = G
        with torch.cuda.stream(stream):
            P, g = G.P, G.g
            b = torch.empty((len(self.reqs),) + tuple(P.ray_shape), dtype=torch.float32, device="cuda")
            for j, req in enumerate(self.reqs):
                truth = phantom_device(g["img_shape"], 20, req.seed, stream=stream)
                P.fp(truth, out=b[j])   # each measurement as its job alone makes it
                del truth
            S = SIRT(P, b, gamma=self.reqs[0].gamma, W_sino=G.W_sino, W_img=G.W_img, nbatch=len(self.reqs))
            del b
            self.S = S
            self.host = torch.empty(S.x.shape, dtype=torch.float32, pin_memory=True)
            iters = self.reqs[0].iters
            if graph:
                gr = torch.cuda.CUDAGraph()
                gr.capture_begin()
                try:
                    S._iteration()
                finally:
                    gr.capture_end()
                self.graph = gr
                for _ in range(iters):
                    gr.replay()
            else:
                self.graph = None
                for _ in range(iters):
                    S._iteration()
            self.host.copy_(S.x, non_blocking=True)
            self.done = torch.cuda.Event()
            self.done.record(stream)
        self.open = len(self.reqs)

    def member(self, j):
        return _Member(self, j)

    def release_one(self):
        self.open -= 1
        if self.open == 0:
            self.graph = None
            self.S = None
            self.geometry.users -= 1
            self.geometry = None


class _Member:
    """The handle of one request of a batched job."""

    def __init__(self, job, j):
        self.job, self.j = job, j

    def poll(self):
        return self.job.done.query()

    def result(self):
        self.job.done.synchronize()
        return self.job.host[self.j].numpy().copy()

    def release(self):
        self.job.release_one()


class SirtBackend:
    """Runs requests as device-resident, graph-captured SIRT jobs.

    cache_bytes: device bytes of projectors (with W_sino, W_img) kept for
    later requests of the same geometry (least recently used evicted first,
    never while a job uses it); 0 disables the cache.  Cone-beam requests of
    one geometry, iteration count and step can run batched (start_batch, up
    to max_batch at once)."""

    max_batch = 4

    def __init__(self, n_streams=8, graph=True, cache_bytes=32 << 30):
        import torch
        from .device import require_gpu
        require_gpu()
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self.graph = graph
        self.cache_bytes = int(cache_bytes)
        self._cache = collections.OrderedDict()
        self.cache_hits = 0
        self.cache_misses = 0

    @staticmethod
    def geometry_key(req):
        return (req.kind, req.n, req.views)

    @staticmethod
    def batch_key(req):
        """Requests with one batch key run as one batched job (cone beam
        only: the Siddon projector has the batched steps)."""
        if req.kind != "cone":
            return None
        return (req.kind, req.n, req.views, req.iters, req.gamma)

    @staticmethod
    def estimate(req):
        """Upper bound of the device bytes one request holds."""
        from . import _lib
        if req.kind == "fan":
            g = fan_geometry(req.n, req.views)
            nsino = req.views * g["nu"]
            nvox = req.n * req.n
            b = 4 * (4 * nsino + 4 * nvox) + 16 * req.views
        else:
            g = cone_geometry(req.n, req.views)
            nrays = req.views * g["m"] * g["m"]
            nvox = req.n ** 3
            grid = _lib.make_grid((req.n,) * 3, (1.0,) * 3, (0.0,) * 3, 1e-5)
            ws = int(_lib.lib().vir_siddon_bp_workspace(grid, req.views))
            b = 24 * nrays + 4 * 4 * nrays + 4 * 4 * nvox + max(ws, 4 * nvox) + 4 * nvox
            b += _lib.SIDDON_RECORD_BYTES * nrays  # the backprojector's per-ray records
            b += 48 * req.views + 64
        # allocator granularity: every tensor rounded up to 2 MiB at most
        return int(b * 1.02) + 24 * (2 << 20)

    def _evict(self):
        total = sum(e.nbytes for e in self._cache.values())
        for k in list(self._cache):
            if total <= self.cache_bytes:
                break
            e = self._cache[k]
            if e.users == 0:
                del self._cache[k]
                total -= e.nbytes

    def _geometry(self, req, slot):
        key = self.geometry_key(req)
        G = self._cache.get(key)
        if G is not None:
            self._cache.move_to_end(key)
            self.cache_hits += 1
            # jobs sharing a projector share its workspace: same stream
            stream = G.stream if G.users > 0 else self.streams[slot % len(self.streams)]
            G.stream = stream
        else:
            self.cache_misses += 1
            stream = self.streams[slot % len(self.streams)]
            G = _Geometry(req, stream)
            if 0 < G.nbytes <= self.cache_bytes:
                self._cache[key] = G
        return G, stream

    def start(self, req, slot):
        G, stream = self._geometry(req, slot)
        job = _Job(req, stream, graph=self.graph, geometry=G)
        self._evict()
        return _Released(job, self)

    def start_batch(self, reqs, slots):
        """One batched job for requests of one batch_key; one handle each."""
        keys = {self.batch_key(r) for r in reqs}
        if len(keys) != 1 or None in keys or len(reqs) > self.max_batch:
            raise ValueError("start_batch takes up to max_batch requests of one batch key")
        G, stream = self._geometry(reqs[0], slots[0])
        job = _BatchJob(reqs, stream, graph=self.graph, geometry=G)
        self._evict()
        return [_Released(job.member(j), self) for j in range(len(reqs))]


class _Released:
    """A job handle whose release also trims the backend's cache."""

    def __init__(self, job, backend):
        self.job, self.backend = job, backend

    def poll(self):
        return self.job.poll()

    def result(self):
        return self.job.result()

    def release(self):
        self.job.release()
        self.backend._evict()


def run_alone(req, graph=True):
    """One request on an idle device, synchronously: the reference result
    for the alone-vs-in-the-mix check."""
    import torch
    job = _Job(req, torch.cuda.Stream(), graph=graph)
    x = job.result()
    job.release()
    return x


# ---------------------------------------------------- the per-call path

def plain_sirt(req):
    """The kernel-per-call path: the same projector kernels, but every
    projection takes a host array and returns one (upload, kernel, download,
    synchronise) and the SIRT arithmetic runs in numpy on the host, as the
    reference's loops drive their projectors (sirt_test.py:25-35)."""
    import torch
    from .geometry import circular_projections, circular_rays
    from .sys_mat.dd import DDProjector
    from .sys_mat.siddon import SiddonProjector
    from .workloads import phantom_device
    if req.kind == "fan":
        g = fan_geometry(req.n, req.views)
        P = DDProjector(g["img_shape"], g["ang"], g["nu"], "fan", DSO=g["DSO"], DSD=g["DSD"],
                        du=g["du"], detector="arc")
        n_ell = 10
    else:
        g = cone_geometry(req.n, req.views)
        src, trg = circular_rays(g["dets_y"], g["views"], "cone", g["src_iso"], g["det_iso"],
                                 g["dets_z"])
        mats = circular_projections(g["dets_y"], g["views"], "cone", g["src_iso"], g["det_iso"],
                                    g["dets_z"])
        P = SiddonProjector(src, trg, g["img_shape"], 1.0, proj=mats)
        n_ell = 20

    def fp(x):
        return P.fp(torch.from_numpy(x).cuda()).cpu().numpy()

    def bp(y):
        return P.bp(torch.from_numpy(y).cuda()).cpu().numpy()

    b = phantom_device(g["img_shape"], n_ell, req.seed)
    b = P.fp(b).cpu().numpy()
    shape = g["img_shape"]
    W_sino = fp(np.ones(shape, np.float32))
    W_img = bp(np.ones(b.shape, np.float32))
    m1, m2 = W_sino != 0, W_img != 0
    x = np.zeros(shape, np.float32)
    for _ in range(req.iters):
        r = fp(x)
        r[m1] = (b[m1] - r[m1]) / W_sino[m1]
        u = bp(r)
        u[m2] /= W_img[m2]
        x += np.float32(req.gamma) * u
    return x
