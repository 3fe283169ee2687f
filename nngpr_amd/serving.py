"""Continuous-batching prediction server over a resident NNGP model.

No reference counterpart (the reference only has a blocking `predict`,
base_nngpr.py:763-894); it serves the same only-train predictions.

Design
------
* `ResidentModel` keeps the fitted training set, its kNN-prepared coordinates
  and norms, and the kernel parameters on the device (C ABI `nngp_model_*`).
* `NngpServer` owns a few *slots*.  A slot is a HIP stream with static device
  buffers for up to `max_batch_points` queries, pinned host staging and a
  workspace.  Slots run concurrently, so batches launched on different slots
  overlap on the device.
* Admission: requests wait in a FIFO; whenever a slot is free, the scheduler packs
  queued request chunks into that slot's batch up to its point capacity (a request
  larger than a slot is split into chunks that may run on several slots).  The
  number of slots and their capacity are derived from an HBM budget, so a small
  budget makes requests queue rather than fail.  Requests that arrive while other
  batches are on the device join the next free slot: continuous batching.
* Each slot's batch is padded to a bucket size (256 * sqrt(2)^i rounded to a
  multiple of 64, up to the slot capacity: at most ~1.41x padding) and replayed
  from a HIP graph captured once per (slot, bucket): the
  kNN screen, re-rank, fallback and fused prediction kernels of the whole batch are
  one graph launch.  Graph capture can be disabled (`use_graphs=False`).
* Every query's result is independent of the other queries in its batch, so a
  request's mean/std are identical alone and in any mix.
* Latency (submit -> result available) and throughput are recorded per request.
"""
import collections
import concurrent.futures
import ctypes
import threading
import time

import numpy as np
import torch

from . import _native as nat
from . import hip_ops
from .kernel_params import get_kernel_params, get_kernel_theta_mapping


class ResidentModel:
    """A fitted model's training set and kernel, resident on one HIP device."""

    def __init__(self, x_train, y_train, kernel, alpha=1e-10, num_nn=32, nn_type="kernel-space",
                 precision=hip_ops.PRECISION_FP32_REFINED, y_mean=None, y_std=None,
                 device="cuda:0"):
        self.device = torch.device(device)
        x = torch.as_tensor(np.ascontiguousarray(x_train, dtype=np.float64)).to(self.device)
        y = np.asarray(y_train, dtype=np.float64)
        if y.ndim == 1:
            y = y[:, None]
        yt = torch.as_tensor(np.ascontiguousarray(y)).to(self.device)
        self.n, self.d = x.shape
        self.ny = y.shape[1]
        params, mapping = get_kernel_params(kernel), get_kernel_theta_mapping(kernel)
        self.kstruct = nat.kernel_struct(params, mapping, len(kernel.theta))
        ls = params.get("length_scale") if nn_type == "kernel-space" else None
        self._scale = None if ls is None else np.ascontiguousarray(
            np.asarray(ls, dtype=np.float64).reshape(-1))
        self.precision = precision
        self.y_mean = np.zeros(self.ny) if y_mean is None else np.asarray(y_mean, dtype=np.float64).reshape(-1)
        self.y_std = np.ones(self.ny) if y_std is None else np.asarray(y_std, dtype=np.float64).reshape(-1)
        handle = ctypes.c_void_p()
        stream = torch.cuda.current_stream(self.device)
        rc = nat.lib().nngp_model_create(
            ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(yt.data_ptr()), self.n, self.d, self.ny,
            ctypes.byref(self.kstruct), float(alpha), int(num_nn),
            self._scale.ctypes.data_as(ctypes.c_void_p) if self._scale is not None else None,
            0 if self._scale is None else self._scale.size, int(precision), ctypes.byref(handle),
            ctypes.c_void_p(stream.cuda_stream))
        nat.check(rc, "nngp_model_create")
        self._handle = handle

    @classmethod
    def from_estimator(cls, est, precision=None, device=None):
        """Resident copy of a fitted TorchNngpr / NumpyNngpr."""
        if precision is None:
            precision = (hip_ops.PRECISION_FP32_REFINED if est.allow_downcast_f32 == "yes"
                         else hip_ops.PRECISION_FP64)
        dev = device or est.workers[0].device
        return cls(est.X_train_, est.y_train_, est.kernel_, alpha=float(est.alpha),
                   num_nn=est.num_nn, nn_type=est.nn_type, precision=precision,
                   y_mean=est._y_train_mean, y_std=est._y_train_std, device=dev)

    @classmethod
    def from_pickle(cls, path, precision=None, device=None):
        """Resident copy of a fitted estimator saved with pickle (the estimator's own
        format, as the reference's estimators are saved: torch_nngpr.py:214-215).
        Unpickling runs code named in the file: load trusted files only."""
        import pickle
        with open(path, "rb") as f:
            est = pickle.load(f)
        return cls.from_estimator(est, precision=precision, device=device)

    def workspace_bytes(self, max_query):
        return int(nat.lib().nngp_model_predict_workspace_bytes(self._handle, int(max_query)))

    def predict_into(self, xq, mean, std, ws, stream=None):
        """Enqueue a prediction of device queries xq (nq, d) into device buffers."""
        stream = stream or torch.cuda.current_stream(self.device)
        rc = nat.lib().nngp_model_predict(
            self._handle, ctypes.c_void_p(xq.data_ptr()), xq.shape[0], ctypes.c_void_p(mean.data_ptr()),
            ctypes.c_void_p(std.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
            ctypes.c_void_p(stream.cuda_stream))
        nat.check(rc, "nngp_model_predict")

    def predict(self, x_query):
        """Blocking prediction (mean (nq, ny) and std (nq,) in the original y scale)."""
        xq = torch.as_tensor(np.ascontiguousarray(x_query, dtype=np.float64)).to(self.device)
        nq = xq.shape[0]
        mean = torch.empty((nq, self.ny), dtype=torch.float64, device=self.device)
        std = torch.empty(nq, dtype=torch.float64, device=self.device)
        ws = torch.empty(max(self.workspace_bytes(nq), 256), dtype=torch.uint8, device=self.device)
        self.predict_into(xq, mean, std, ws)
        return self.unscale(mean.cpu().numpy(), std.cpu().numpy())

    def unscale(self, mean, std):
        return mean * self.y_std[None, :] + self.y_mean[None, :], std[:, None] * self.y_std[None, :]

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            nat.lib().nngp_model_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Request:
    __slots__ = ("x", "n", "mean", "std", "future", "t_submit", "t_done", "next_row", "rows_done")

    def __init__(self, x, ny):
        self.x = x
        self.n = x.shape[0]
        self.mean = np.empty((self.n, ny))
        self.std = np.empty((self.n, ny))
        self.future = concurrent.futures.Future()
        self.t_submit = time.perf_counter()
        self.t_done = None
        self.next_row = 0
        self.rows_done = 0


class _Slot:
    def __init__(self, model, capacity, use_graphs):
        dev = model.device
        self.model = model
        self.capacity = capacity
        self.stream = torch.cuda.Stream(device=dev)
        self.q_dev = torch.zeros((capacity, model.d), dtype=torch.float64, device=dev)
        self.mean_dev = torch.zeros((capacity, model.ny), dtype=torch.float64, device=dev)
        self.std_dev = torch.zeros(capacity, dtype=torch.float64, device=dev)
        self.ws = torch.empty(max(model.workspace_bytes(capacity), 256), dtype=torch.uint8, device=dev)
        self.q_host = torch.zeros((capacity, model.d), dtype=torch.float64).pin_memory()
        self.mean_host = torch.zeros((capacity, model.ny), dtype=torch.float64).pin_memory()
        self.std_host = torch.zeros(capacity, dtype=torch.float64).pin_memory()
        self.event = torch.cuda.Event()
        self.use_graphs = use_graphs
        self.graphs = {}
        self.parts = None  # list of (request, row0, rows, offset) of the batch in flight

    def bucket(self, n):
        """Smallest captured batch size that holds n points."""
        for b in self.buckets():
            if b >= n:
                return b
        return self.capacity

    def _enqueue_predict(self, b):
        self.model.predict_into(self.q_dev[:b], self.mean_dev[:b], self.std_dev[:b], self.ws,
                                stream=torch.cuda.current_stream(self.model.device))

    def buckets(self):
        """Batch sizes with a captured graph: 256 * sqrt(2)^i rounded up to a
        multiple of 64 (one wave of queries), then the capacity, so a batch is
        padded by at most ~1.41x."""
        if getattr(self, "_buckets", None) is None:
            out, i = [], 0
            while True:
                b = -(-int(round(256 * 2.0 ** (i / 2.0))) // 64) * 64
                if b >= self.capacity:
                    break
                if not out or b > out[-1]:
                    out.append(b)
                i += 1
            out.append(self.capacity)
            self._buckets = out
        return self._buckets

    def capture_all(self):
        """Capture the prediction graph of every bucket size (before serving starts:
        capture synchronises the device)."""
        for b in self.buckets():
            if b not in self.graphs:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=self.stream):
                    self._enqueue_predict(b)
                self.graphs[b] = g

    def launch(self, parts, npts):
        self.parts = parts
        off = 0
        qh = self.q_host.numpy()
        for req, r0, rows, _ in parts:
            qh[off:off + rows] = req.x[r0:r0 + rows]
            off += rows
        b = self.bucket(npts)
        # padding rows repeat the batch's first query: rows left from earlier batches
        # or zeros could lie far outside the training set, and a far query walks
        # many grid rings (the padded batch's kNN would wait for it)
        qh[npts:b] = qh[0]
        with torch.cuda.stream(self.stream):
            self.q_dev[:b].copy_(self.q_host[:b], non_blocking=True)
            if self.use_graphs:
                self.graphs[b].replay()
            else:
                self._enqueue_predict(b)
            self.mean_host[:npts].copy_(self.mean_dev[:npts], non_blocking=True)
            self.std_host[:npts].copy_(self.std_dev[:npts], non_blocking=True)
            self.event.record(self.stream)

    def done(self):
        return self.parts is not None and self.event.query()


class NngpServer:
    """Continuous-batching predictions against a ResidentModel (see module doc).

    hbm_budget_bytes bounds the device memory of all slots together (static
    query/output buffers and kNN/solve workspaces); max_batch_points caps a
    slot's batch.  Use as a context manager or call close().
    """

    def __init__(self, model, hbm_budget_bytes=4 << 30, max_batch_points=65536, max_slots=4,
                 use_graphs=True):
        self.model = model
        cap = int(max_batch_points)
        per = model.workspace_bytes(cap) + cap * 8 * (2 * model.d + 2 * model.ny + 2)
        while cap > 256 and per > hbm_budget_bytes:
            cap //= 2
            per = model.workspace_bytes(cap) + cap * 8 * (2 * model.d + 2 * model.ny + 2)
        n_slots = max(1, min(int(max_slots), int(hbm_budget_bytes // max(per, 1))))
        self.slot_capacity = cap
        self.hbm_bytes = n_slots * per
        self.slots = [_Slot(model, cap, use_graphs) for _ in range(n_slots)]
        if use_graphs:
            for slot in self.slots:
                slot.capture_all()
        self._pending = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self.completed = []
        self._thread = threading.Thread(target=self._loop, name="nngp-server", daemon=True)
        self._thread.start()

    # ------------------------------------------------------------------ API
    def submit(self, x_query):
        """Queue a prediction request; returns a Future of (mean, std) arrays in the
        original y scale, shapes (nq, ny) each (std repeated per output)."""
        x = np.ascontiguousarray(x_query, dtype=np.float64)
        if x.ndim != 2 or x.shape[1] != self.model.d:
            raise ValueError("queries must have shape (n, %d)" % self.model.d)
        req = _Request(x, self.model.ny)
        if req.n == 0:
            req.t_done = time.perf_counter()
            req.future.set_result((req.mean, req.std))
            return req.future
        with self._cv:
            if self._stop:
                raise RuntimeError("server is closed")
            self._pending.append(req)
            self._cv.notify()
        return req.future

    def latencies_ms(self):
        return np.array([(r.t_done - r.t_submit) * 1e3 for r in self.completed])

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._thread.join()
        torch.cuda.synchronize(self.model.device)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ scheduler
    def _take_batch(self, capacity):
        """FIFO chunks of pending requests filling up to `capacity` points."""
        parts, npts = [], 0
        while self._pending and npts < capacity:
            req = self._pending[0]
            if req.future.cancelled():
                self._pending.popleft()
                continue
            if req.next_row == 0 and not req.future.set_running_or_notify_cancel():
                self._pending.popleft()
                continue
            rows = min(req.n - req.next_row, capacity - npts)
            parts.append((req, req.next_row, rows, npts))
            req.next_row += rows
            npts += rows
            if req.next_row >= req.n:
                self._pending.popleft()
        return parts, npts

    def _finish(self, slot):
        mh = slot.mean_host.numpy()
        sh = slot.std_host.numpy()
        for req, r0, rows, off in slot.parts:
            mean, std = self.model.unscale(mh[off:off + rows], sh[off:off + rows])
            req.mean[r0:r0 + rows] = mean
            req.std[r0:r0 + rows] = std
            req.rows_done += rows
            if req.rows_done == req.n:
                req.t_done = time.perf_counter()
                self.completed.append(req)
                req.future.set_result((req.mean, req.std))
        slot.parts = None

    def _loop(self):
        torch.cuda.set_device(self.model.device)
        while True:
            progressed = False
            for slot in self.slots:
                if slot.done():
                    self._finish(slot)
                    progressed = True
            with self._cv:
                for slot in self.slots:
                    if slot.parts is None and self._pending:
                        parts, npts = self._take_batch(slot.capacity)
                        if parts:
                            try:
                                slot.launch(parts, npts)
                            except Exception as e:  # surface errors on the futures
                                for req, _, _, _ in parts:
                                    if not req.future.done():
                                        req.future.set_exception(e)
                                slot.parts = None
                            progressed = True
                busy = any(s.parts is not None for s in self.slots)
                if self._stop and not busy and not self._pending:
                    return
                if not progressed:
                    if busy:
                        self._cv.wait(timeout=50e-6)
                    else:
                        self._cv.wait(timeout=0.01)
