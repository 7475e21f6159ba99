"""Batched execution of many independent phase-field solves (libpfhip.so pf_batch_*).

An ``Instance`` is one solve: its parameters, temperature model, dtype, step count and
time grid, plus its two device state buffers.  ``BatchEngine`` owns a slot table on
the device; ``run(steps, chunk)`` advances every active slot by up to ``steps`` steps
with one persistent launch per (dtype, K, anisotropy, tile) class and step, captured
into HIP graphs of ``chunk`` steps.  Slots whose step budget is used up stop changing,
so solves of different lengths share launches.  The engine works on its own stream
(default priority).  No counterpart exists in the reference, which runs one solve per
process (applications/phase_field/*/example.py).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .allen_cahn import DTYPES
from .config import make_params, with_defaults
from .laser import Rosenthal


@dataclass
class Instance:
    """One solve.  ``cell_ori_inds`` (N,) int32 gives the one-hot initial state (ini_cond),
    or ``eta0`` an explicit (N, K) initial state."""
    pf_args: dict
    h_xyz: tuple
    grain_dirs: np.ndarray          # (3, K, 3) rotated cube axes
    laser: object                   # pfkit.laser descriptor
    dtype: str
    steps: int
    dt: float
    t0: float = 0.0
    cell_ori_inds: object = None
    eta0: object = None
    tile: object = None             # None: the autotuner's throughput choice (pfkit.autotune.lookup)
    zchunk: object = None
    tag: object = None
    toolpath: object = None         # source table (pfkit.laser.Toolpath, RosenthalRamp): step i runs
                                    # step_times()[i] -> [i+1] with its laser(i)
    # filled by the engine
    slot: int = -1
    bufs: list = field(default_factory=list)
    axes: object = None
    path: object = None             # device table of the toolpath's per-step rows
    params: object = None           # the pf_params_t the slot runs with

    @property
    def N(self):
        return int(self.pf_args["Nx"]) * int(self.pf_args["Ny"]) * int(self.pf_args["Nz"])

    @property
    def K(self):
        return int(self.pf_args["num_oris"])

    @property
    def torch_dtype(self):
        return DTYPES[self.dtype]

    def state_bytes(self):
        return 2 * self.N * self.K * torch.empty((), dtype=self.torch_dtype).element_size()


class BatchEngine:
    def __init__(self, max_instances=256, max_tiles=1 << 22, device=None):
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.stream = torch.cuda.Stream(self.device)
        self.h = self.lib.pf_batch_create(int(max_instances), int(max_tiles), self.stream.cuda_stream)
        if not self.h:
            raise _lib.PFError("pf_batch_create: " + self.lib.pf_last_error().decode())
        self.max_instances = max_instances
        self.insts = {}
        # outputs of the retirement check (pf_reduce_noalloc), allocated once
        self._red_f = torch.empty(4 + 256, dtype=torch.float64, device=self.device)
        self._red_i = torch.empty(256 + 3, dtype=torch.int64, device=self.device)
        self._red_ws = None

    # ---- instance lifecycle
    def add(self, inst: Instance) -> int:
        a = with_defaults(inst.pf_args)
        tdt = inst.torch_dtype
        self.stream.wait_stream(torch.cuda.current_stream(self.device))  # inputs made on the caller's stream
        cdt = torch.float64 if tdt == torch.float64 else torch.float32
        with torch.cuda.stream(self.stream):
            b0 = torch.empty((inst.N, inst.K), dtype=tdt, device=self.device)
            b1 = torch.empty_like(b0)
            if inst.eta0 is not None:
                b0.copy_(torch.as_tensor(inst.eta0, device=self.device).to(tdt))
            else:
                ori = inst.cell_ori_inds
                if isinstance(ori, torch.Tensor):
                    ori = ori.to(device=self.device, dtype=torch.int32)
                else:
                    ori = torch.as_tensor(np.asarray(ori), dtype=torch.int32, device=self.device)
                _lib.check(self.lib.pf_init_onehot(b0.data_ptr(), _lib.TORCH_TO_PF[tdt], ori.data_ptr(), inst.N,
                                                   inst.K, self.stream.cuda_stream), "pf_init_onehot")
            axes = torch.as_tensor(np.ascontiguousarray(inst.grain_dirs), dtype=cdt, device=self.device) \
                if float(a["anisotropy"]) > 0 else None
        bi = _lib.BatchInstT()
        bi.eta_dev[0] = b0.data_ptr()
        bi.eta_dev[1] = b1.data_ptr()
        bi.T_dev = None
        bi.axes_dev = axes.data_ptr() if axes is not None else None
        bi.dtype = _lib.TORCH_TO_PF[tdt]
        if inst.tile is None or inst.zchunk is None:
            from . import autotune
            t, z = autotune.lookup(_lib.TORCH_TO_PF[tdt], a, float(a["anisotropy"]) > 0, throughput=True)
            inst.tile = t if inst.tile is None else inst.tile
            inst.zchunk = z if inst.zchunk is None else inst.zchunk
        bi.tile = int(inst.tile)
        bi.zchunk = int(inst.zchunk)
        bi.steps = int(inst.steps)
        bi.t0 = float(inst.t0)
        bi.dt = float(inst.dt)
        laser, path = inst.laser, None
        if inst.toolpath is not None:
            # rows (t_prev, t_crt, pos0, dir, strength, vel) of the steps, as PFSolver.run_toolpath
            # passes them to pf_step (PF_PATH_ROW values; strength = Q or dT_peak)
            tp = inst.toolpath
            if getattr(tp, "apply_switch", False):
                raise ValueError("batched toolpaths run with the source always on (apply_switch=False)")
            ts = tp.step_times()
            if inst.steps > len(ts) - 1:
                raise ValueError(f"the toolpath has {len(ts) - 1} steps, {inst.steps} asked for")
            rows = np.zeros((max(int(inst.steps), 1), _lib.PF_PATH_ROW))
            kind = None
            for i in range(int(inst.steps)):
                L = tp.laser(i)
                if kind is None:
                    kind = type(L)
                elif type(L) is not kind:
                    raise ValueError("every step of a batched source table needs the same source model")
                strength = L.Q if isinstance(L, Rosenthal) else L.dT_peak
                rows[i] = (ts[i], ts[i + 1], *L.pos0, *L.dir, strength, L.vel)
            path = torch.as_tensor(rows, dtype=torch.float64, device=self.device)
            laser = tp.laser(0)
            bi.t0 = float(ts[0])
            bi.path_dev = path.data_ptr()
        else:
            bi.path_dev = None
        bi.p = make_params(a, inst.h_xyz, laser)
        inst.params = bi.p
        self.stream.wait_stream(torch.cuda.current_stream(self.device))  # the table is copied on the caller's stream
        slot = self.lib.pf_batch_add(self.h, ctypes.byref(bi))
        if slot < 0:
            _lib.check(slot, "pf_batch_add")
        inst.slot, inst.bufs, inst.axes, inst.path = slot, [b0, b1], axes, path
        self.insts[slot] = inst
        return slot

    def remove(self, slot):
        _lib.check(self.lib.pf_batch_remove(self.h, int(slot)), "pf_batch_remove")
        return self.insts.pop(slot)

    # ---- execution
    def run(self, steps, chunk=16):
        """Enqueue `steps` steps of every active slot (graph replays of `chunk` steps)."""
        _lib.check(self.lib.pf_batch_run(self.h, int(steps), int(chunk)), "pf_batch_run")

    def status(self):
        """(steps_left, t, parity) per slot (synchronizes the engine's stream)."""
        n = self.max_instances
        sl = (ctypes.c_int32 * n)()
        tt = (ctypes.c_double * n)()
        pa = (ctypes.c_int32 * n)()
        _lib.check(self.lib.pf_batch_status(self.h, sl, tt, pa, n), "pf_batch_status")
        return np.frombuffer(sl, dtype=np.int32).copy(), np.frombuffer(tt).copy(), np.frombuffer(pa, np.int32).copy()

    def nonfinite(self, slot, parity=None):
        """Number of inf/nan values in slot's current state (FP64 reduction on the engine's stream;
        pf_reduce_noalloc, which allocates nothing, when its fixed scratch holds the reduction)."""
        inst = self.insts[slot]
        eta = self.state(slot, parity)
        p = inst.params
        lib = self.lib
        st = self.stream.cuda_stream
        if inst.K > 256:
            with torch.cuda.stream(self.stream):
                return int((~torch.isfinite(eta)).sum().item())
        Tptr = None  # batch instances evaluate a source descriptor, never a temperature buffer
        need = lib.pf_reduce_workspace_bytes(inst.N, inst.K)
        if need <= _lib.REDUCE_SCRATCH_BYTES:
            rc = lib.pf_reduce_noalloc(eta.data_ptr(), _lib.TORCH_TO_PF[eta.dtype], Tptr, None, p, float(inst.t0),
                                       self._red_f.data_ptr(), self._red_i.data_ptr(), st)
        else:
            if self._red_ws is None or self._red_ws.numel() < need:
                with torch.cuda.stream(self.stream):
                    self._red_ws = torch.empty(need, dtype=torch.uint8, device=self.device)
            rc = lib.pf_reduce(eta.data_ptr(), _lib.TORCH_TO_PF[eta.dtype], Tptr, None, p, float(inst.t0),
                               self._red_f.data_ptr(), self._red_i.data_ptr(), self._red_ws.data_ptr(), need, st)
        _lib.check(rc, "nonfinite check")
        self.stream.synchronize()
        return int(self._red_i[inst.K + 2].item())

    def state(self, slot, parity=None):
        """The device tensor holding slot's current state."""
        if parity is None:
            parity = self.status()[2][slot]
        return self.insts[slot].bufs[int(parity)]

    def run_to_completion(self, chunk=16, max_chunks=1 << 20):
        for _ in range(max_chunks):
            sl = self.status()[0]
            if not any(sl[s] > 0 for s in self.insts):
                return
            self.run(chunk, chunk)
        raise RuntimeError("run_to_completion: chunk budget exhausted")

    def close(self):
        if self.h:
            self.stream.synchronize()
            self.lib.pf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
