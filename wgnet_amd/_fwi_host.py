"""Host side of the pixel-grid FWI: scipy's L-BFGS-B band loop
(python/wgnet/fwi_eval.py:37-57), the lockstep barrier that turns many
problems' objective calls into one batched request per round, and the worker
process that runs a share of a batch's problems and asks its parent for every
objective value.

This module imports only numpy and scipy, so a worker starts quickly and
never touches the GPU.  Worker protocol over stdin / stdout, each message a
little-endian u64 length and a pickle:
  parent -> worker  {"pids", "starts", "bands", "budget", "smooth_grad"}
  worker -> parent  ("eval", [(pid, band, field_f32 [70, 70]), ...])
  parent -> worker  {pid: (loss, grad_f32 [70, 70]), ...}
  worker -> parent  ("done", {pid: (final [70, 70], history)}) or ("error", text)
"""

from __future__ import annotations

import pickle
import struct
import sys
import threading
import traceback

import numpy as np
from scipy.ndimage import gaussian_filter
from scipy.optimize import minimize

N = 70
V_BOUNDS_KM = (1.4, 4.6)  # fwi_eval.py:29


def lbfgsb_bands(start_km, bands, budget, smooth_grad, band_observed, evaluate):
    """The band loop of fwi_eval.py:37-57.  band_observed(band) -> a handle of
    the band-limited observed data; evaluate(field_f32 [N, N], band, handle)
    -> (loss, grad [N, N] float32).  Returns (final [N, N] float32, history)."""
    x = np.asarray(start_km).astype(np.float64).reshape(-1)
    per_band = max(2, budget // len(bands))
    history = []
    for band in bands:
        observed = band_observed(band)
        calls = 0

        def objective(vec: np.ndarray):
            nonlocal calls
            calls += 1
            loss, grad = evaluate(vec.reshape(N, N).astype(np.float32), band, observed)
            g = gaussian_filter(grad, smooth_grad, mode="nearest") if smooth_grad > 0 else grad
            return float(loss), g.reshape(-1).astype(np.float64)

        res = minimize(objective, x, jac=True, method="L-BFGS-B", bounds=[V_BOUNDS_KM] * (N * N),
                       options={"maxfun": per_band, "maxiter": per_band, "maxcor": 10})
        x = res.x
        history.append({"band_hz": band, "misfit": float(res.fun), "solves": calls})
    return x.reshape(N, N).astype(np.float32), history


class Lockstep:
    """Collects one objective request per running problem (thread); when every
    running problem waits, `flush(pending)` answers them all at once.
    pending maps pid -> (field, band, handle); flush returns pid -> (loss, grad)."""

    def __init__(self, flush, problems: int):
        self.flush = flush
        self.cond = threading.Condition()
        self.active = problems
        self.pending: dict = {}
        self.results: dict = {}
        self.rounds = 0
        self.error: BaseException | None = None

    def _round(self) -> None:
        try:
            self.results.update(self.flush(dict(self.pending)))
        except BaseException as exc:  # every waiting problem sees the failure
            self.error = exc
        self.rounds += 1
        self.pending.clear()
        self.cond.notify_all()

    def submit(self, pid, field, band, handle):
        with self.cond:
            if self.error is not None:
                raise RuntimeError("a batched evaluation failed") from self.error
            self.pending[pid] = (field, band, handle)
            if len(self.pending) == self.active:
                self._round()
            while pid not in self.results and self.error is None:
                self.cond.wait()
            if pid not in self.results:
                raise RuntimeError("a batched evaluation failed") from self.error
            return self.results.pop(pid)

    def finish(self) -> None:
        with self.cond:
            self.active -= 1
            if self.pending and len(self.pending) == self.active:
                self._round()


def run_lockstep(pids, starts, bands, budget, smooth_grad, band_observed, flush):
    """One thread per problem, all in lockstep through `flush`.  Returns
    ({pid: (final, history)}, rounds); a problem's exception is re-raised."""
    hub = Lockstep(flush, len(pids))
    out: dict = {}
    errors: list = []

    def run(pid, start):
        try:
            out[pid] = lbfgsb_bands(start, bands, budget, smooth_grad, lambda band: band_observed(pid, band),
                                    lambda f, band, h: hub.submit(pid, f, band, h))
        except BaseException as exc:
            errors.append(exc)
        finally:
            hub.finish()

    threads = [threading.Thread(target=run, args=(p, s), daemon=True) for p, s in zip(pids, starts)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return out, hub.rounds


def send(f, obj) -> None:
    data = pickle.dumps(obj, protocol=pickle.HIGHEST_PROTOCOL)
    f.write(struct.pack("<Q", len(data)))
    f.write(data)
    f.flush()


def recv(f):
    head = f.read(8)
    if len(head) < 8:
        raise EOFError("peer closed the pipe")
    (n,) = struct.unpack("<Q", head)
    data = f.read(n)
    if len(data) < n:
        raise EOFError("peer closed the pipe")
    return pickle.loads(data)


def worker_main() -> int:
    """Runs the parent's share of problems; every round's requests go to the
    parent, which evaluates all workers' requests in one batched call."""
    inp, out = sys.stdin.buffer, sys.stdout.buffer
    sys.stdout = sys.stderr  # the pipe carries the protocol only
    try:
        job = recv(inp)
    except EOFError:
        return 0

    def flush(pending):
        send(out, ("eval", [(pid, band, field) for pid, (field, band, _) in sorted(pending.items())]))
        return recv(inp)

    try:
        res, _ = run_lockstep(job["pids"], job["starts"], job["bands"], job["budget"], job["smooth_grad"],
                              lambda pid, band: None, flush)
    except EOFError:
        return 0
    except BaseException:
        send(out, ("error", traceback.format_exc()))
        return 1
    send(out, ("done", res))
    return 0


if __name__ == "__main__":
    sys.exit(worker_main())
