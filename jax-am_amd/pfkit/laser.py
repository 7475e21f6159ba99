"""Temperature descriptors evaluated inside the fused step kernel.

In the reference the app computes T on its own (one extra jitted launch and an
(N, 1) array per step) and passes it to ``stepper`` through ``ode_params``.  Passing
one of these descriptors instead lets the kernel evaluate T at each cell centroid
(FP64) with no temperature array in memory.  The reference's array form is still
accepted by ``PFSolver.stepper`` (PF_T_BUFFER).
"""
from dataclasses import dataclass
from typing import Tuple

from . import _lib


@dataclass
class Rosenthal:
    """Moving point source, applications/phase_field/simple/example.py:21-37
    (general direction: multi_scan/example.py:33-49).

    T = T_ambient + Q/(2 pi kappa_T) / R * exp(-vel/(2 alpha) (R + (X,Y,Z).dir)),
    source at pos0 + vel t dir.  T_cap > 0 caps T (multi_scan/example.py:47).
    """
    pos0: Tuple[float, float, float]
    vel: float
    Q: float = 25.0
    alpha: float = 5.2e-6
    kappa_T: float = 27.0
    T_ambient: float = 300.0
    dir: Tuple[float, float, float] = (1.0, 0.0, 0.0)
    T_cap: float = 0.0

    @classmethod
    def simple_app(cls, pf_args):
        """The parameters of simple/example.py:26-33: x0 = 0.2 Lx, y = Ly/2, z = Lz,
        vel = 0.6 Lx / t_OFF."""
        return cls(pos0=(0.2 * pf_args["domain_x"], 0.5 * pf_args["domain_y"], pf_args["domain_z"]),
                   vel=0.6 * pf_args["domain_x"] / pf_args["t_OFF"], T_ambient=pf_args["T_ambient"])

    def fill(self, L):
        L.model = _lib.PF_T_ROSENTHAL
        L.T_ambient = self.T_ambient
        for d in range(3):
            L.pos0[d] = self.pos0[d]
            L.dir[d] = self.dir[d]
        L.vel = self.vel
        L.Q, L.kappa_T, L.alpha, L.T_cap = self.Q, self.kappa_T, self.alpha, self.T_cap

    def as_dict(self):
        return dict(model="rosenthal", pos0=self.pos0, vel=self.vel, Q=self.Q, alpha=self.alpha,
                    kappa_T=self.kappa_T, T_ambient=self.T_ambient, dir=self.dir, T_cap=self.T_cap)


@dataclass
class Gaussian:
    """Moving Gaussian spot: T = T_ambient + dT_peak exp(-(X^2+Y^2)/(2 sigma_xy^2) - Z^2/(2 sigma_z^2)),
    centre at pos0 + vel t dir."""
    pos0: Tuple[float, float, float]
    vel: float
    dT_peak: float
    sigma_xy: float
    sigma_z: float
    T_ambient: float = 300.0
    dir: Tuple[float, float, float] = (1.0, 0.0, 0.0)

    def fill(self, L):
        L.model = _lib.PF_T_GAUSSIAN
        L.T_ambient = self.T_ambient
        for d in range(3):
            L.pos0[d] = self.pos0[d]
            L.dir[d] = self.dir[d]
        L.vel = self.vel
        L.dT_peak, L.sigma_xy, L.sigma_z = self.dT_peak, self.sigma_xy, self.sigma_z

    def as_dict(self):
        return dict(model="gaussian", pos0=self.pos0, vel=self.vel, dT_peak=self.dT_peak,
                    sigma_xy=self.sigma_xy, sigma_z=self.sigma_z, T_ambient=self.T_ambient, dir=self.dir)


@dataclass
class Quench:
    """Directional solidification, solidification/example.py:21-34: T = T_ref + grad_T z - cool_rate t."""
    T_ref: float
    grad_T: float
    cool_rate: float

    @classmethod
    def solidification_app(cls, pf_args, vel=0.2, thermal_grad=5.e5):
        return cls(T_ref=pf_args["T_liquidus"], grad_T=thermal_grad, cool_rate=thermal_grad * vel)

    def fill(self, L):
        L.model = _lib.PF_T_QUENCH
        L.T_ref, L.grad_T, L.cool_rate = self.T_ref, self.grad_T, self.cool_rate

    def as_dict(self):
        return dict(model="quench", T_ref=self.T_ref, grad_T=self.grad_T, cool_rate=self.cool_rate)


LASERS = (Rosenthal, Gaussian, Quench)


@dataclass
class Toolpath:
    """Multi-segment scan path, applications/phase_field/multi_scan/example.py:52-84 (read_path)
    driving the Rosenthal source of :33-49 (Q = 25 W, alpha = 5.2e-6, kappa = 27, T capped at 2000).

    Segment i runs from corner i to corner i+1 at speed ``vel``: its step times are
    ``numpy.arange(t_pre, t_pre + |d_i| / vel, dt)`` and its source positions are
    ``numpy.linspace(corner_i, corner_{i+1}, n_i)`` (so the last position of a segment is the
    next corner), with the segment's unit direction.  ``schedule()`` returns those arrays;
    ``laser(i)`` is the Rosenthal descriptor of step i: it places the source at the i-th position
    at time ts[i], so it works with ``PFSolver.stepper`` (T is evaluated at t_prev = ts[i]) and
    with ``PFSolver.run_toolpath``.  The reference reads a per-segment power switch but does not
    use it in T; ``apply_switch=True`` turns the source off (T = T_ambient) on switched-off segments.
    """
    x_pos: Tuple[float, ...]
    y_pos: Tuple[float, ...]
    z_pos: Tuple[float, ...]
    switch: Tuple[float, ...]
    vel: float
    dt: float
    Q: float = 25.0
    alpha: float = 5.2e-6
    kappa_T: float = 27.0
    T_ambient: float = 300.0
    T_cap: float = 2000.0
    apply_switch: bool = False

    @classmethod
    def from_args(cls, pf_args, **kw):
        lp = pf_args["laser_path"]
        return cls(tuple(lp["x_pos"]), tuple(lp["y_pos"]), tuple(lp["z_pos"]), tuple(lp["switch"]),
                   float(pf_args["vel"]), float(pf_args["dt"]), T_ambient=float(pf_args.get("T_ambient", 300.0)),
                   **kw)

    def schedule(self):
        """(ts, xs, ys, zs, ps, mov_dir) with one entry per step."""
        import numpy as np
        if getattr(self, "_sched", None) is not None:
            return self._sched
        cx, cy, cz = (np.asarray(v, dtype=np.float64) for v in (self.x_pos, self.y_pos, self.z_pos))
        ts, xs, ys, zs, ps, dirs = [], [], [], [], [], []
        t0 = 0.0
        for i in range(len(cx) - 1):
            seg = np.array([cx[i + 1] - cx[i], cy[i + 1] - cy[i], cz[i + 1] - cz[i]])
            dist = np.linalg.norm(seg)
            duration = dist / self.vel
            t_seg = np.arange(t0, t0 + duration, self.dt)
            n = len(t_seg)
            ts.append(t_seg)
            xs.append(np.linspace(cx[i], cx[i + 1], n))
            ys.append(np.linspace(cy[i], cy[i + 1], n))
            zs.append(np.linspace(cz[i], cz[i + 1], n))
            ps.append(np.full(n, float(self.switch[i])))
            dirs.append(np.repeat((seg / dist)[None, :], n, axis=0))
            t0 = t0 + duration
        self._sched = (np.concatenate(ts), np.concatenate(xs), np.concatenate(ys), np.concatenate(zs),
                       np.concatenate(ps), np.concatenate(dirs))
        return self._sched

    def __len__(self):
        return len(self.schedule()[0])

    def step_times(self):
        return self.schedule()[0]

    def laser(self, i):
        """The temperature descriptor of step i (valid at t = ts[i])."""
        ts, xs, ys, zs, ps, dirs = self.schedule()
        if self.apply_switch and ps[i] == 0.0:
            return Quench(T_ref=self.T_ambient, grad_T=0.0, cool_rate=0.0)
        d = dirs[i]
        vt = self.vel * ts[i]
        pos0 = (xs[i] - vt * d[0], ys[i] - vt * d[1], zs[i] - vt * d[2])
        return Rosenthal(pos0=tuple(float(v) for v in pos0), vel=self.vel, Q=self.Q, alpha=self.alpha,
                         kappa_T=self.kappa_T, T_ambient=self.T_ambient, dir=tuple(float(v) for v in d),
                         T_cap=self.T_cap)



@dataclass
class RosenthalRamp:
    """Rosenthal source whose power and speed ramp linearly in time, simple/example_grf.py:21-46:
    Q(t) = Q_max + (Q_min - Q_max) t / t_OFF, vel(t) = vel_min + (vel_max - vel_min) t / t_OFF, and
    the source at x0 + (vel(t) + vel_min) t / 2 (the integral of vel), y0, z0, moving along +x;
    the exponent uses vel(t).  ``laser(i)`` is the Rosenthal descriptor of step i, valid at
    t = ts[i] (pf_step evaluates T at t_prev, as the reference's loop does, example_grf.py:61-63),
    with ts = numpy.arange(0, t_OFF + 1e-10, dt) (example_grf.py:57); the batch engine takes it as a
    source table (pf_batch_inst_t.path_dev)."""
    x0: float
    y0: float
    z0: float
    t_OFF: float
    dt: float
    vel_avg: float
    Q_max: float = 35.0
    Q_min: float = 15.0
    alpha: float = 5.2e-6
    kappa_T: float = 27.0
    T_ambient: float = 300.0

    @classmethod
    def grf_app(cls, pf_args):
        """example_grf.py:26-33: x0 = 0.2 Lx, y = Ly/2, z = Lz, vel_avg = 0.6 Lx / t_OFF."""
        return cls(x0=0.2 * pf_args["domain_x"], y0=0.5 * pf_args["domain_y"], z0=pf_args["domain_z"],
                   t_OFF=pf_args["t_OFF"], dt=pf_args["dt"], vel_avg=0.6 * pf_args["domain_x"] / pf_args["t_OFF"],
                   T_ambient=pf_args.get("T_ambient", 300.0))

    def step_times(self):
        import numpy as np
        return np.arange(0., self.t_OFF + 1e-10, self.dt)

    def __len__(self):
        return len(self.step_times())

    def at(self, t):
        """The descriptor of time t: Q(t), vel(t) and pos0 such that pos0 + vel(t) t = x0 + (vel(t) + vel_min) t / 2."""
        vel_min, vel_max = 0.5 * self.vel_avg, 1.5 * self.vel_avg
        Q = t / self.t_OFF * (self.Q_min - self.Q_max) + self.Q_max
        vel = t / self.t_OFF * (vel_max - vel_min) + vel_min
        x0 = self.x0 + (vel + vel_min) / 2. * t - vel * t
        return Rosenthal(pos0=(x0, self.y0, self.z0), vel=vel, Q=Q, alpha=self.alpha, kappa_T=self.kappa_T,
                         T_ambient=self.T_ambient)

    def laser(self, i):
        return self.at(float(self.step_times()[i]))
