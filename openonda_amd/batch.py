"""Many particle sets of uneven size advanced together on one GPU.

Each RK4 stage of every set runs in ONE persistent launch (``vpm_batch_run``:
workgroups pull (set, target block) items, largest set first), and a whole
time step -- four stages, their axpys, the per-set LES filter and the final
RK4 combination -- is captured once in a HIP graph and replayed
(``vpm_graph_*``).  Sets keep a uniform radius each.

The step integrates the coupled particle ODE with classical RK4:
dx/dt = u(x, alpha), dalpha/dt = S(x, alpha) [+ D_PSE(x, alpha; nu + nu_t)].
With ``les=True`` stage 1 also evaluates the smoothed vorticity and the
enstrophy of every set and applies the dynamic LES filter of
vpmModule.py:887-962 per set; the resulting nu_t enters the PSE term from
stage 2 on (stage 1 of a step uses the nu_t of the previous step).

The reference solver advances one set per call and has no batched path.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L
from . import physics as P


def work_list(sizes, targets_per_item):
    """(item_set, item_block, set_item_first, set_item_count): sets ordered by
    decreasing size, the target blocks of a set consecutive."""
    sizes = np.asarray(sizes, dtype=np.int64)
    order = np.argsort(-sizes, kind="stable")
    item_set, item_block = [], []
    first = np.zeros(len(sizes), dtype=np.int32)
    count = np.zeros(len(sizes), dtype=np.int32)
    for s in order:
        nb = int(math.ceil(sizes[s] / targets_per_item))
        first[s] = len(item_set)
        count[s] = nb
        item_set.extend([s] * nb)
        item_block.extend(range(nb))
    return (np.array(item_set, dtype=np.int32), np.array(item_block, dtype=np.int32), first, count)


class BatchedSystem:
    """Particle sets [(x (N_i,3), alpha (N_i,3), sigma (N_i,), nu (N_i,)), ...]."""

    def __init__(self, sets, dt, precision="f32c", tile_config=-1, pse=True, les=True,
                 device=None, grid=0, resort_every=0, autotune_cache=None):
        """``tile_config``: -1 (the library's default shape for the batch),
        an index of physics.tile_configs(), or "auto" (the batch autotuner's
        cached (shape, grid) for this output set and total pair count, else
        the default).  ``grid``: persistent workgroups (0: as many as fit on
        the device).  ``resort_every``: re-sort the sets by Morton key every
        that many steps, eager or replayed (0: only at the first step)."""
        P.require_gpu()
        if precision not in ("f32c", "f64"):
            raise ValueError("precision must be 'f32c' or 'f64'")
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self.dt = float(dt)
        self.precision = precision
        self.pse = pse
        self.les = les
        self.resort_every = int(resort_every)
        # one tile shape for every stage (the work items count its target
        # blocks), chosen for the heaviest stage's outputs
        self.stage_mask = (L.OUT_VEL | L.OUT_STRETCH | (L.OUT_PSE if pse else 0)
                           | (L.OUT_VORT | L.OUT_ENSTROPHY if les else 0))
        sizes = [len(s[0]) for s in sets]
        self.autotuned = None
        if tile_config == "auto":
            from . import autotune

            self.autotuned = autotune.best_batch_config(sizes, self.stage_mask, path=autotune_cache)
            tile_config, grid = self.autotuned if self.autotuned else (-1, grid)
        self.tile_config = (int(tile_config) if tile_config >= 0
                            else int(L.load().vpm_batch_tile_config_default(self.stage_mask)))
        for k, s in enumerate(sets):
            sig = np.asarray(s[2])
            if len(sig) and not np.all(sig == sig[0]):
                raise ValueError(f"set {k}: the batched kernels need a uniform radius per set")
        self.sizes = np.array(sizes, dtype=np.int64)
        self.nsets = len(sets)
        n = int(self.sizes.sum())
        self.n = n
        if n >= 2 ** 31:
            raise ValueError("too many particles")
        d = self.device
        cat = lambda i: np.concatenate([np.asarray(s[i], dtype=np.float64) for s in sets])
        self.x = P.to_soa(cat(0).reshape(-1, 3), d)
        self.a = P.to_soa(cat(1).reshape(-1, 3), d)
        self.sigma = torch.from_numpy(cat(2)).to(d)
        self.nu = torch.from_numpy(cat(3)).to(d)
        self.nu_t = torch.zeros_like(self.nu)
        self.nu_eff = self.nu.clone()
        off = np.zeros(self.nsets + 1, dtype=np.int32)
        off[1:] = np.cumsum(self.sizes)
        self.offsets = off
        self.ws, self.wsb = None, 0
        self.set_launch(self.tile_config, grid)
        e = lambda: torch.zeros((3, n), dtype=torch.float64, device=d)
        self.kx = [e() for _ in range(4)]
        self.ka = [e() for _ in range(4)]
        self.xs, self.as_ = e(), e()
        self.vel, self.str_, self.pse_, self.vort = e(), e(), e(), e()
        self.enstrophy = torch.zeros(max(self.nsets, 1), dtype=torch.float64, device=d)
        st = np.zeros((max(self.nsets, 1), 3))
        st[:, 0] = 0.02958                                      # vpmModule.py:92
        self.les_state = torch.from_numpy(st).to(d)
        self.stream = torch.cuda.Stream(device=d)
        self.graph = None
        self.steps_done = 0
        self._sorted = False

    def set_launch(self, tile_config, grid=0):
        """Selects the persistent launch's tile shape and grid: rebuilds the
        work list (its items are the shape's target blocks) and grows the
        workspace if needed; the sorted order in the workspace is kept.  A
        captured graph is dropped (it holds the old launch)."""
        d = self.device
        self.tile_config = int(tile_config)
        self.grid = int(grid)
        if self.precision == "f64":
            tpi = 256
        else:
            b, r = P.tile_configs()[self.tile_config][:2]
            tpi = b * r
        item_set, item_block, first, count = work_list(self.sizes, tpi)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(d)
        self._desc_t = [t(self.offsets), t(np.repeat(np.arange(self.nsets), self.sizes)),
                        t(item_set), t(item_block), t(first), t(count)]
        self.n_items = len(item_set)
        self.desc = L.BatchDesc(self.nsets, self.n_items, self.grid,
                                *[x.data_ptr() for x in self._desc_t])
        lib = L.load()
        wsb = lib.vpm_batch_workspace_bytes(self.n, self.nsets, self.n_items)
        if wsb == 0:
            raise L.VpmError("vpm_batch_workspace_bytes", -3, lib.vpm_last_error().decode())
        if self.ws is None or wsb > self.ws.numel():
            old = self.ws
            if old is not None:
                # the system's stream may still be using the old workspace
                self.stream.synchronize()
            self.ws = torch.empty(wsb, dtype=torch.uint8, device=d)
            if old is not None:  # keep the sorted order (perm, records)
                self.ws[:old.numel()].copy_(old)
                # the copy runs on the current stream, the launches on the
                # system's: finish it before any launch reads the workspace
                # (the sort's scratch moves when the layout grows)
                torch.cuda.current_stream(d).synchronize()
        self.wsb = self.ws.numel()
        if getattr(self, "graph", None) is not None:
            L.call("vpm_graph_destroy", self.graph)
            self.graph = None

    # ------------------------------------------------------------ primitives
    def _parts(self, x, a):
        return L.Particles(self.n, max(self.n, 1), x.data_ptr(), a.data_ptr(),
                           self.sigma.data_ptr(), self.nu_eff.data_ptr())

    def _mask(self, stage1):
        m = L.OUT_VEL | L.OUT_STRETCH
        if self.pse:
            m |= L.OUT_PSE
        if stage1 and self.les:
            m |= L.OUT_VORT | L.OUT_ENSTROPHY
        return m

    def evaluate(self, x, a, mask, resort, stream, vel=None, stretch=None):
        """prepare + one persistent launch on ``stream`` (the system's),
        after the work already issued on the current stream (inputs written
        there, outputs allocated there).  Outputs go to ``vel``/``stretch``
        (default self.vel/self.str_), self.pse_, self.vort, self.enstrophy."""
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        self._evaluate(x, a, mask, resort, stream, vel, stretch)

    def _evaluate(self, x, a, mask, resort, stream, vel=None, stretch=None):
        p = self._parts(x, a)
        L.call("vpm_batch_prepare", C.byref(p), C.byref(self.desc), 1 if resort else 0,
               self.ws.data_ptr(), self.wsb, stream)
        o = L.ForwardOpts(mask, P.PRECISIONS[self.precision], L.SIGMA_UNIFORM, 0, self.tile_config)
        vel = self.vel if vel is None else vel
        stretch = self.str_ if stretch is None else stretch
        fo = L.ForwardOut(max(self.n, 1), vel.data_ptr(), stretch.data_ptr(),
                          self.pse_.data_ptr(), self.vort.data_ptr(), self.enstrophy.data_ptr())
        L.call("vpm_batch_run", C.byref(p), C.byref(self.desc), C.byref(o), C.byref(fo),
               self.ws.data_ptr(), self.wsb, stream)

    def _axpy(self, x, y, c, out, stream):
        L.call("vpm_axpy3", self.n, max(self.n, 1), x.data_ptr(), y.data_ptr(), float(c),
               out.data_ptr(), stream)

    def _stage(self, k, x, a, resort, stream):
        # kx = u, ka = S (+ D)
        self._evaluate(x, a, self._mask(k == 0), resort, stream, self.kx[k], self.ka[k])
        if self.pse:
            self._axpy(self.ka[k], self.pse_, 1.0, self.ka[k], stream)
        if k == 0 and self.les:
            L.call("vpm_batch_les_update", C.byref(self.desc), max(self.n, 1),
                   self.vort.data_ptr(), self.sigma.data_ptr(), self.enstrophy.data_ptr(),
                   self.les_state.data_ptr(), self.nu.data_ptr(), self.nu_t.data_ptr(),
                   self.nu_eff.data_ptr(), self.ws.data_ptr(), self.wsb, self.n, stream)

    def _step(self, resort, stream):
        dt = self.dt
        self._stage(0, self.x, self.a, resort, stream)
        for k, h in ((1, 0.5 * dt), (2, 0.5 * dt), (3, dt)):
            self._axpy(self.x, self.kx[k - 1], h, self.xs, stream)
            self._axpy(self.a, self.ka[k - 1], h, self.as_, stream)
            self._stage(k, self.xs, self.as_, False, stream)
        for y, ks in ((self.x, self.kx), (self.a, self.ka)):
            L.call("vpm_rk4_combine", self.n, max(self.n, 1), y.data_ptr(), ks[0].data_ptr(),
                   ks[1].data_ptr(), ks[2].data_ptr(), ks[3].data_ptr(), dt, y.data_ptr(), stream)

    # ------------------------------------------------------------ public
    def _resort_due(self):
        return self.resort_every > 0 and self.steps_done % self.resort_every == 0

    def step(self, resort=None):
        """One eager step on the system's stream (sorts on the first call and
        every ``resort_every`` steps)."""
        if resort is None:
            resort = not self._sorted or self._resort_due()
        torch.cuda.current_stream(self.device).synchronize()
        self._step(resort, self.stream.cuda_stream)
        self._sorted = True
        self.steps_done += 1

    def capture(self):
        """Captures one step (without the sort) in a HIP graph."""
        if not self._sorted:
            raise RuntimeError("run one eager step (it sorts the sets) before capturing")
        s = self.stream.cuda_stream
        self.stream.synchronize()
        L.call("vpm_graph_begin", s)
        try:
            self._step(False, s)
        finally:
            g = C.c_void_p()
            rc = L.load().vpm_graph_end(s, C.byref(g))
        if rc != 0:
            raise L.VpmError("vpm_graph_end", rc, L.load().vpm_last_error().decode())
        if self.graph is not None:
            L.call("vpm_graph_destroy", self.graph)
        self.graph = g.value

    def replay(self, steps=1):
        """Replays the captured step; where a re-sort is due
        (``resort_every``) that step runs eagerly with the sort instead (the
        graph reads the sorted order from the workspace, so it stays valid)."""
        if self.graph is None:
            self.capture()
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        for _ in range(steps):
            if self._resort_due():
                self._step(True, self.stream.cuda_stream)
            else:
                L.call("vpm_graph_launch", self.graph, self.stream.cuda_stream)
            self.steps_done += 1

    def synchronize(self):
        self.stream.synchronize()

    def set_slice(self, s):
        return slice(int(self.offsets[s]), int(self.offsets[s + 1]))

    def pairs_per_evaluation(self):
        return float(np.sum(self.sizes.astype(np.float64) ** 2))

    def __del__(self):
        try:
            if self.graph is not None:
                L.load().vpm_graph_destroy(self.graph)
        except Exception:
            pass
