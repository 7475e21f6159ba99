"""Drop-in for ``jax_am.phase_field.allen_cahn`` on MI355X.

Same entry points and state layout as the reference (jax_am/phase_field/allen_cahn.py):
``phase_field``, ``explicit_euler``, ``rk4``, ``get_force_eta_fn`` and ``PFSolver``
with ``stepper``, ``ini_cond``, ``clean_sols``, ``inspect_sol`` and ``write_sols``.
State is ``(eta, t)`` with eta a (N, K) device tensor, orientation innermost, cells
in C order (z, y, x).  ``ode_params = [T]`` takes the reference's (N, 1) temperature
array, or a laser descriptor (pfkit.laser) that the fused kernel evaluates in place.

Every step runs the hand-written HIP kernels of libpfhip.so through ctypes on
torch's current stream; there is no other compute path.
"""
import glob
import os
import time

import numpy as np
import torch

from . import _lib
from .config import make_params, with_defaults
from .laser import LASERS

DTYPES = {"float64": torch.float64, "float32": torch.float32, "bfloat16": torch.bfloat16,
          "f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}


def _as_dtype(d):
    if d is None:
        return torch.float64
    if isinstance(d, torch.dtype):
        return d
    return DTYPES[str(d)]


class StateRHS:
    """What ``phase_field`` returns: the closure of allen_cahn.py:10-87 as an object holding the
    packed parameters.  Calling it evaluates state_rhs on the GPU (pf_rhs)."""

    def __init__(self, polycrystal, pf_args, dtype=torch.float64, device=None):
        self.pf_args = with_defaults(pf_args)
        self.polycrystal = polycrystal
        self.dtype = _as_dtype(dtype)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.h_xyz = tuple(polycrystal.mesh_h_xyz)
        self.K = int(self.pf_args["num_oris"])
        self.N = int(self.pf_args["Nx"]) * int(self.pf_args["Ny"]) * int(self.pf_args["Nz"])
        cdt = torch.float64 if self.dtype == torch.float64 else torch.float32
        self.aniso = float(self.pf_args["anisotropy"]) > 0.0
        self.axes = torch.as_tensor(np.ascontiguousarray(polycrystal.unique_grain_directions), dtype=cdt,
                                    device=self.device) if self.aniso else None
        self.pf_dtype = _lib.TORCH_TO_PF[self.dtype]
        self._T_keep = None
        _lib.load()

    # ---- temperature argument
    def params_for(self, T):
        """pf_params_t and the T device pointer for ode_params[0] = T."""
        if isinstance(T, LASERS):
            return make_params(self.pf_args, self.h_xyz, T), None
        p = make_params(self.pf_args, self.h_xyz, None)
        Tt = torch.as_tensor(T, device=self.device)
        if Tt.dtype not in (torch.float64, torch.float32):
            Tt = Tt.to(torch.float64)
        Tt = Tt.reshape(-1).contiguous()
        if Tt.numel() != self.N:
            raise ValueError(f"T has {Tt.numel()} values, expected {self.N} (one per cell)")
        p.laser.model = _lib.PF_T_BUFFER
        p.laser.T_dtype = _lib.TORCH_TO_PF[Tt.dtype]
        self._T_keep = Tt  # alive until the stream has consumed it (stream order + caching allocator)
        return p, Tt.data_ptr()

    def check_state(self, eta):
        if not isinstance(eta, torch.Tensor):
            eta = torch.as_tensor(np.asarray(eta), device=self.device)
        if eta.device != self.device:
            eta = eta.to(self.device)
        if eta.dtype != self.dtype:
            eta = eta.to(self.dtype)
        if tuple(eta.shape) != (self.N, self.K):
            raise ValueError(f"eta has shape {tuple(eta.shape)}, expected {(self.N, self.K)}")
        return eta.contiguous()

    def __call__(self, state, t, ode_params):
        """state_rhs(state, t, ode_params) (allen_cahn.py:53-85) on the GPU."""
        eta = self.check_state(state)
        T, = ode_params
        p, Tptr = self.params_for(T)
        out = torch.empty_like(eta)
        lib = _lib.load()
        _lib.check(lib.pf_rhs(eta.data_ptr(), out.data_ptr(), self.pf_dtype, Tptr, p,
                              _lib.ptr(self.axes), float(t), _lib.stream_handle(self.device)), "pf_rhs")
        return out

    def step(self, eta_in, eta_out, T, t_prev, t_crt, tile=-1, zchunk=0):
        """Fused explicit Euler step eta_out = eta_in + (t_crt - t_prev) rhs (+ liquid reset)."""
        p, Tptr = self.params_for(T)
        lib = _lib.load()
        _lib.check(lib.pf_step(eta_in.data_ptr(), eta_out.data_ptr(), self.pf_dtype, Tptr, p,
                               _lib.ptr(self.axes), float(t_prev), float(t_crt), int(tile), int(zchunk),
                               _lib.stream_handle(self.device)), "pf_step")


def phase_field(polycrystal, pf_args, dtype=None, device=None):
    """allen_cahn.py:10-87: returns the right-hand-side object for (polycrystal, pf_args)."""
    return StateRHS(polycrystal, pf_args, dtype=_as_dtype(dtype or pf_args.get("precision")), device=device)


def explicit_euler(state_pre, t_crt, f, ode_params):
    """allen_cahn.py:106-114: ((y + (t_crt - t_prev) f(y, t_prev), t_crt), y_new), fused on the GPU."""
    y_prev, t_prev = state_pre
    y_prev = f.check_state(y_prev)
    y = torch.empty_like(y_prev)
    T, = ode_params
    f.step(y_prev, y, T, t_prev, t_crt)
    return (y, t_crt), y


def rk4(state_pre, t_crt, f, ode_params):
    """allen_cahn.py:90-103 (classical RK4 with four pf_rhs evaluations)."""
    y_prev, t_prev = state_pre
    y_prev = f.check_state(y_prev)
    h = t_crt - t_prev
    k1 = h * f(y_prev, t_prev, ode_params)
    k2 = h * f(y_prev + k1 / 2., t_prev + h / 2., ode_params)
    k3 = h * f(y_prev + k2 / 2., t_prev + h / 2., ode_params)
    k4 = h * f(y_prev + k3, t_prev + h, ode_params)
    y_crt = y_prev + 1. / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return (y_crt, t_crt), y_crt


def get_force_eta_fn(pf_args):
    """allen_cahn.py:117-128: eta <- 0 where T > T_liquidus + 1 (T: (N, 1) tensor)."""
    a = with_defaults(pf_args)

    def force_eta_zero_in_liquid(state, T):
        eta, t = state
        liquid = torch.as_tensor(T, device=eta.device).reshape(-1, 1) > a["T_liquidus"] + a["liquid_shift"]
        return (torch.where(liquid, torch.zeros((), dtype=eta.dtype, device=eta.device), eta), t)

    return force_eta_zero_in_liquid


class PFSolver:
    """allen_cahn.py:131-210 on MI355X.

    Extra keyword arguments: ``dtype`` (torch.float64 default = the reference's x64
    run; float32; bfloat16 storage with FP32 arithmetic), ``device``, ``tile`` /
    ``zchunk`` (kernel tile shape, -1/0 = default or the autotuner's choice), and
    ``clean`` (delete old VTU output like the reference constructor does).
    """

    def __init__(self, pf_args, polycrystal, dtype=None, device=None, tile=None, zchunk=None, clean=True):
        self.pf_args = pf_args
        self.polycrystal = polycrystal
        self.state_rhs = phase_field(polycrystal, pf_args, dtype=dtype, device=device)
        self.dtype = self.state_rhs.dtype
        self.device = self.state_rhs.device
        self.force_eta_fn = get_force_eta_fn(pf_args)
        self.tile, self.zchunk = tile, zchunk
        if tile is None or zchunk is None:
            from . import autotune
            t, z = autotune.lookup(self.state_rhs.pf_dtype, pf_args, self.state_rhs.aniso)
            self.tile = t if tile is None else tile
            self.zchunk = z if zchunk is None else zchunk
        self._argmax0 = None
        self._ws = None
        if clean and "data_dir" in pf_args:
            self.clean_sols()

    @property
    def N(self):
        return self.state_rhs.N

    @property
    def K(self):
        return self.state_rhs.K

    def stepper(self, state_pre, t_crt, ode_params):
        """allen_cahn.py:139-143.  Returns ((eta_new, t_crt), eta_new); eta_new is a new tensor."""
        eta_pre, t_prev = state_pre
        eta_pre = self.state_rhs.check_state(eta_pre)
        eta = torch.empty_like(eta_pre)
        T, = ode_params
        self.state_rhs.step(eta_pre, eta, T, t_prev, t_crt, self.tile, self.zchunk)
        return (eta, t_crt), eta

    def run(self, eta0, ts, T):
        """Advance through the time grid ts with a fixed temperature argument (a laser
        descriptor: T(t) is evaluated in-kernel each step), ping-ponging two buffers.
        Returns the final eta (a tensor distinct from eta0)."""
        a = self.state_rhs.check_state(eta0).clone()
        b = torch.empty_like(a)
        for i in range(len(ts) - 1):
            self.state_rhs.step(a, b, T, ts[i], ts[i + 1], self.tile, self.zchunk)
            a, b = b, a
        return a

    def run_toolpath(self, eta0, toolpath, steps=None, callback=None):
        """Advance along a source table: a ``pfkit.laser.Toolpath`` (multi_scan/example.py:88-105) or
        ``RosenthalRamp`` (simple/example_grf.py:21-67).  Step i goes from ts[i] to ts[i+1]
        (ts = ``toolpath.step_times()``) with the source of ``toolpath.laser(i)``.
        ``callback(i + 1, eta)`` runs after each step (eta is the live buffer; copy it to keep it).
        Returns the final eta."""
        ts = toolpath.step_times()
        n = len(ts) - 1 if steps is None else min(int(steps), len(ts) - 1)
        a = self.state_rhs.check_state(eta0).clone()
        b = torch.empty_like(a)
        for i in range(n):
            self.state_rhs.step(a, b, toolpath.laser(i), ts[i], ts[i + 1], self.tile, self.zchunk)
            a, b = b, a
            if callback is not None:
                callback(i + 1, a)
        return a

    def save_checkpoint(self, path, eta, t, step=0):
        """Full-state checkpoint (no counterpart in the reference): eta in its storage dtype (BF16 as
        its 16-bit pattern), t, step and the parameters, in an .npz file (no pickled objects)."""
        import json
        e = self.state_rhs.check_state(eta).detach()
        raw = e.view(torch.int16) if e.dtype == torch.bfloat16 else e
        np.savez(path, eta=raw.cpu().numpy(), dtype=str(e.dtype).replace("torch.", ""), t=float(t), step=int(step),
                 shape=np.array(e.shape), pf_args=json.dumps({k: v for k, v in self.pf_args.items()
                                                             if isinstance(v, (int, float, str, bool))}))
        return path

    def load_checkpoint(self, path):
        """(eta on this solver's device, t, step) from save_checkpoint."""
        z = np.load(path, allow_pickle=False)
        dt = str(z["dtype"])
        raw = torch.as_tensor(z["eta"])
        eta = raw.view(torch.bfloat16) if dt == "bfloat16" else raw
        if tuple(eta.shape) != (self.N, self.K):
            raise ValueError(f"checkpoint eta has shape {tuple(eta.shape)}, expected {(self.N, self.K)}")
        return self.state_rhs.check_state(eta.to(self.device)), float(z["t"]), int(z["step"])

    def ini_cond(self):
        """allen_cahn.py:145-153: one-hot eta[c, cell_ori_inds[c]] = 1."""
        eta = torch.empty((self.N, self.K), dtype=self.dtype, device=self.device)
        ori = self.polycrystal.cell_ori_inds
        if isinstance(ori, torch.Tensor):
            ori = ori.to(device=self.device, dtype=torch.int32)
        else:
            ori = torch.as_tensor(np.asarray(ori, dtype=np.int32), device=self.device)
        lib = _lib.load()
        _lib.check(lib.pf_init_onehot(eta.data_ptr(), self.state_rhs.pf_dtype, ori.data_ptr(), self.N, self.K,
                                      _lib.stream_handle(self.device)), "pf_init_onehot")
        torch.cuda.current_stream(self.device).synchronize()
        return eta

    def clean_sols(self):
        """allen_cahn.py:155-163."""
        vtk_folder = os.path.join(self.pf_args['data_dir'], "vtk/pf/sols")
        os.makedirs(vtk_folder, exist_ok=True)
        for f in glob.glob(vtk_folder + "/*"):
            os.remove(f)

    # ---- diagnostics
    def argmax(self, eta):
        out = torch.empty(self.N, dtype=torch.int32, device=self.device)
        lib = _lib.load()
        _lib.check(lib.pf_argmax(eta.data_ptr(), self.state_rhs.pf_dtype, self.N, self.K, out.data_ptr(),
                                 _lib.stream_handle(self.device)), "pf_argmax")
        return out

    def diagnostics(self, eta, T, t=0.0, eta0=None):
        """FP64 reductions (pf_reduce): free energy parts, max T, sum eta, per-orientation volumes,
        argmax counts, #argmax changes vs eta0, #cells with T >= T_liquidus, #non-finite values."""
        eta = self.state_rhs.check_state(eta)
        p, Tptr = self.state_rhs.params_for(T)
        lib = _lib.load()
        nbytes = lib.pf_reduce_workspace_bytes(self.N, self.K)
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        a0 = None
        if eta0 is not None:
            if self._argmax0 is None or self._argmax0[0] is not eta0:
                self._argmax0 = (eta0, self.argmax(self.state_rhs.check_state(eta0)))
            a0 = self._argmax0[1]
        outf = torch.empty(4 + self.K, dtype=torch.float64, device=self.device)
        outi = torch.empty(self.K + 3, dtype=torch.int64, device=self.device)
        _lib.check(lib.pf_reduce(eta.data_ptr(), self.state_rhs.pf_dtype, Tptr, _lib.ptr(a0), p, float(t),
                                 outf.data_ptr(), outi.data_ptr(), self._ws.data_ptr(), nbytes,
                                 _lib.stream_handle(self.device)), "pf_reduce")
        f = outf.cpu().numpy()
        i = outi.cpu().numpy()
        return dict(F_local=float(f[0]), F_grad=float(f[1]), F=float(f[0] + f[1]), max_T=float(f[2]),
                    eta_sum=float(f[3]), volumes=f[4:].copy(), argmax_counts=i[:self.K].copy(),
                    changed=int(i[self.K]), liquid=int(i[self.K + 1]), nonfinite=int(i[self.K + 2]))

    def inspect_sol(self, pf_sol, pf_sol0, T, ts, step, t=None):
        """allen_cahn.py:165-180: prints % orientation change, % liquid, max T; raises ValueError on
        non-finite eta.  T: the (N, 1) array or a laser descriptor (evaluated at t, default ts[step])."""
        print(f"\nstep {step} of {len(ts[1:])}, unix timestamp = {time.time()}")
        tt = float(ts[step]) if t is None else float(t)
        d = self.diagnostics(pf_sol, T, tt, eta0=pf_sol0)
        print(f"percent of change of orientations = {d['changed'] / self.N * 100}%")
        print(f"percet of T >= T_liquidus = {d['liquid'] / self.N * 100}%")
        print(f"max T = {d['max_T']}\n")
        if d["nonfinite"] > 0:
            raise ValueError("Found np.inf or np.nan in pf_sol - stop the program")
        return d

    def temperature(self, T, t=0.0):
        """T as a host (N, 1) float64 array: evaluates a laser descriptor at t on the GPU."""
        if isinstance(T, LASERS):
            p = make_params(self.state_rhs.pf_args, self.state_rhs.h_xyz, T)
            out = torch.empty(self.N, dtype=torch.float64, device=self.device)
            lib = _lib.load()
            _lib.check(lib.pf_temperature(out.data_ptr(), _lib.PF_F64, p, float(t),
                                          _lib.stream_handle(self.device)), "pf_temperature")
            return out.cpu().numpy().reshape(-1, 1)
        return torch.as_tensor(T).detach().cpu().numpy().reshape(-1, 1).astype(np.float64)

    def write_sols(self, pf_sol, T, step, t=0.0):
        """allen_cahn.py:182-210: VTU with cell data T, ipf_x/y/z (black in liquid) and ori_inds."""
        from .vtu import write_hex_vtu
        print("\nWrite phase-field sols to file...\n")
        step = step // self.pf_args['write_sol_interval']
        Th = self.temperature(T, t)
        # argmax and IPF colours on the device; liquid cells black (allen_cahn.py:186-199)
        inds_d = self.argmax(self.state_rhs.check_state(pf_sol)).long()
        liquid_d = torch.as_tensor(Th.reshape(-1) > self.pf_args['T_liquidus'], device=self.device)
        rgb_d = torch.as_tensor(np.asarray(self.polycrystal.unique_oris_rgb), dtype=torch.float64, device=self.device)
        cell_data = {"T": Th.reshape(-1).astype(np.float32), "ori_inds": inds_d.int().cpu().numpy()}
        for j, nm in enumerate(("ipf_x", "ipf_y", "ipf_z")):
            c = rgb_d[j].index_select(0, inds_d)
            c[liquid_d] = 0.
            cell_data[nm] = c.cpu().numpy()
        path = os.path.join(self.pf_args['data_dir'], f"vtk/pf/sols/u{step:03d}.vtu")
        os.makedirs(os.path.dirname(path), exist_ok=True)
        pc = self.polycrystal
        write_hex_vtu(path, pc.Nx, pc.Ny, pc.Nz, pc.mesh_h_xyz, cell_data)
        return path
