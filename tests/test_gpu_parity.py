"""GPU parity of the fused HIP step against the FP64 oracle (oracle/pf_oracle.py).

Tolerances (written here, checked per test):
  F64 storage/compute: relative L-inf <= 1e-12 per step (max|d| / max|eta|).
  F32 storage/compute: |d| <= 2e-6 * max|eta| + 1e-7 per step (stochastic rounding: within one
  FP32 ulp of the FP32 arithmetic's result).
  BF16 storage, F32 compute: the only extra error is the final stochastic rounding to BF16, which
  lands on one of the two BF16 neighbours, so |d| <= 2^-7 |eta_new| + 2e-6 (one bf16 ulp plus FP32
  compute error).
"""
import numpy as np
import pytest
import torch

from oracle import pf_oracle as O
from pfkit import _lib
from pfkit.allen_cahn import PFSolver
from pfkit.config import make_params
from problems import args, gaussian, poly, quench, random_eta, rosenthal

pytestmark = pytest.mark.gpu


def _solver(a, pc, dtype, cuda, **kw):
    return PFSolver(a, pc, dtype=dtype, device=cuda, clean=False, tile=kw.pop("tile", -1),
                    zchunk=kw.pop("zchunk", 0), **kw)


def _oracle_step(eta, a, pc, L, t0, t1, T=None):
    cent = O.centroids(a["Nz"], a["Ny"], a["Nx"], pc.mesh_h_xyz)
    if T is None:
        T = O.laser_T(L.as_dict(), cent, t0)
    return O.explicit_euler(eta, t0, t1, T, a, pc.mesh_h_xyz, pc.unique_grain_directions)


def _gpu_step(s, eta_np, L, t0, t1, dtype, tile=-1, zchunk=0):
    eta = torch.as_tensor(eta_np, dtype=dtype, device=s.device)
    out = torch.empty_like(eta)
    s.state_rhs.step(eta, out, L, t0, t1, tile, zchunk)
    torch.cuda.synchronize()
    return out


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


CASES = [
    # (Nx, Ny, Nz, K, aniso, laser, force_liquid)
    (16, 12, 10, 10, 0.0, "gauss", False),
    (16, 12, 10, 10, 0.15, "gauss", True),
    (33, 17, 9, 20, 0.15, "rosen", False),
    (40, 9, 13, 3, 0.3, "quench", True),
    (7, 5, 6, 5, 0.15, "rosen", True),
    (24, 20, 12, 16, 0.0, "gauss", True),
    (21, 11, 9, 17, 0.15, "gauss", True),    # odd K: 16-byte chunks + a 1-element tail, partial rows
    (35, 13, 7, 19, 0.0, "quench", False),   # odd K, 3-element tail
    (12, 10, 8, 25, 0.15, "gauss", False),   # generic-K kernel
    (9, 7, 5, 33, 0.0, "quench", False),     # generic-K kernel
    (1, 1, 12, 4, 0.15, "gauss", False),     # degenerate 1 x 1 x Nz column
    (70, 1, 1, 6, 0.15, "gauss", False),     # degenerate line
]


def _laser(kind, Nx, Ny, Nz):
    return {"gauss": gaussian(Nx, Ny, Nz), "rosen": rosenthal(Nx, Ny, Nz), "quench": quench(Nz)}[kind]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}x{c[2]}k{c[3]}e{c[4]}{c[5]}fl{int(c[6])}" for c in CASES])
def test_f64_step_matches_oracle(case, cuda):
    Nx, Ny, Nz, K, eps, lk, fl = case
    print(f"case {case}", flush=True)
    a = args(Nx, Ny, Nz, K, anisotropy=eps, force_liquid=fl)
    pc = poly(Nx, Ny, Nz, K)
    L = _laser(lk, Nx, Ny, Nz)
    eta = random_eta(Nx * Ny * Nz, K, seed=Nx + K)
    t0, t1 = 0.3, 0.3 + 0.02
    s = _solver(a, pc, torch.float64, cuda)
    got = _gpu_step(s, eta, L, t0, t1, torch.float64).cpu().numpy()
    ref = _oracle_step(eta, a, pc, L, t0, t1)
    assert _rel(got, ref) <= 1e-12, _rel(got, ref)


@pytest.mark.parametrize("case", CASES[:6], ids=[f"{c[0]}x{c[1]}x{c[2]}k{c[3]}" for c in CASES[:6]])
def test_f32_step_matches_oracle(case, cuda):
    Nx, Ny, Nz, K, eps, lk, fl = case
    print(f"case {case}", flush=True)
    a = args(Nx, Ny, Nz, K, anisotropy=eps, force_liquid=fl)
    pc = poly(Nx, Ny, Nz, K)
    L = _laser(lk, Nx, Ny, Nz)
    eta = random_eta(Nx * Ny * Nz, K, seed=Nx + K).astype(np.float32).astype(np.float64)
    s = _solver(a, pc, torch.float32, cuda)
    got = _gpu_step(s, eta, L, 0.3, 0.32, torch.float32).cpu().numpy().astype(np.float64)
    ref = _oracle_step(eta, a, pc, L, 0.3, 0.32)
    assert np.max(np.abs(got - ref)) <= 2e-6 * np.max(np.abs(ref)) + 1e-7


@pytest.mark.parametrize("case", CASES[:4], ids=[f"{c[0]}x{c[1]}x{c[2]}k{c[3]}" for c in CASES[:4]])
def test_bf16_step_matches_oracle(case, cuda):
    Nx, Ny, Nz, K, eps, lk, fl = case
    a = args(Nx, Ny, Nz, K, anisotropy=eps, force_liquid=fl)
    pc = poly(Nx, Ny, Nz, K)
    L = _laser(lk, Nx, Ny, Nz)
    eta_b = torch.as_tensor(random_eta(Nx * Ny * Nz, K, seed=Nx + K)).to(torch.bfloat16)
    eta = eta_b.to(torch.float64).numpy()
    s = _solver(a, pc, torch.bfloat16, cuda)
    got = _gpu_step(s, eta, L, 0.3, 0.32, torch.bfloat16).to(torch.float64).cpu().numpy()
    ref = _oracle_step(eta, a, pc, L, 0.3, 0.32)
    assert np.all(np.abs(got - ref) <= 2.0 ** -7 * np.abs(ref) + 2e-6)


# (37, 21, 19): partial tiles in x and y; (40, 64, 9): tile rows divisible by 8, which selects
# the XCD-interleaved tile order of the single-step launch
@pytest.mark.parametrize("K,dims", [(10, (37, 21, 19)), (17, (37, 21, 19)), (20, (40, 64, 9))])
def test_tile_shapes_and_generic_kernel_agree_bitwise(K, dims, cuda):
    Nx, Ny, Nz = dims
    a = args(Nx, Ny, Nz, K, anisotropy=0.15, force_liquid=True)
    pc = poly(Nx, Ny, Nz, K)
    L = gaussian(Nx, Ny, Nz)
    dense = random_eta(Nx * Ny * Nz, K, seed=5)
    # sparse: one-hot grains with a diffuse band and a few -0.0 entries (zero-channel skips,
    # zero-gradient anisotropy skips and cold/hot temperature paths all occur)
    sparse = O.ini_cond(pc.cell_ori_inds, K)
    band = (np.arange(Nx * Ny * Nz) % 7) == 0
    sparse[band] = 0.8 * sparse[band] + 0.2 * dense[band]
    sparse[3, 1] = -0.0
    for eta in (dense, sparse):
        for dtype in (torch.float64, torch.float32, torch.bfloat16):
            s = _solver(a, pc, dtype, cuda)
            lib = _lib.load()
            nt = lib.pf_tile_count(_lib.TORCH_TO_PF[dtype], K)
            assert nt >= 2
            base = _gpu_step(s, eta, L, 0.1, 0.12, dtype, tile=-2)  # generic kernel
            for tile in range(nt):
                # PF_ZCHUNK_TAIL: the lowest planes in half-length chunks (every plane exactly once)
                T = _lib.ZCHUNK_TAIL
                for zc in (0, 1, 3, 7, Nz, 2 | T, 3 | T, 7 | T, Nz | T):
                    got = _gpu_step(s, eta, L, 0.1, 0.12, dtype, tile=tile, zchunk=zc)
                    assert torch.equal(got, base), (dtype, tile, zc)


def test_dropped_row_stores_write_nothing_past_the_output(cuda):
    """FP64 tiles write their rows back through raw buffer stores; lanes without a piece store at an
    offset 2^31 past the plane's base, which the hardware's range check drops.  With the output at
    the start of a buffer that reaches past those addresses, every byte after the output keeps its
    fill value, and the output equals the generic kernel's bitwise (tiles cut in x and y)."""
    Nx, Ny, Nz, K = 37, 21, 5, 20
    a = args(Nx, Ny, Nz, K, anisotropy=0.15, force_liquid=True)
    pc = poly(Nx, Ny, Nz, K)
    L = gaussian(Nx, Ny, Nz)
    eta_np = random_eta(Nx * Ny * Nz, K, seed=11)
    s = _solver(a, pc, torch.float64, cuda)
    base = _gpu_step(s, eta_np, L, 0.1, 0.12, torch.float64, tile=-2)  # generic kernel
    n = Nx * Ny * Nz * K
    fill = 7.25
    buf = torch.full(((1 << 31) // 8 + (1 << 21),), fill, dtype=torch.float64, device=cuda)  # 2^31 B + 16 MiB
    out = buf[:n].view(Nx * Ny * Nz, K)
    eta = torch.as_tensor(eta_np, dtype=torch.float64, device=cuda)
    lib = _lib.load()
    for tile in range(lib.pf_tile_count(_lib.PF_F64, K)):
        for zc in (0, 2 | _lib.ZCHUNK_TAIL):
            out.fill_(fill)
            s.state_rhs.step(eta, out, L, 0.1, 0.12, tile, zc)
            torch.cuda.synchronize()
            assert torch.equal(out, base), (tile, zc)
            assert bool((buf[n:] == fill).all()), (tile, zc)
    del buf


def test_sparse_one_hot_trajectory_f32_and_f64(cuda):
    """From the one-hot state most channels are exactly zero: exercises the skip paths."""
    Nx, Ny, Nz, K = 40, 24, 12, 10
    a = args(Nx, Ny, Nz, K, anisotropy=0.15, force_liquid=True)
    pc = poly(Nx, Ny, Nz, K, grains=30)
    L = gaussian(Nx, Ny, Nz, vel=4.0)
    eta0 = O.ini_cond(pc.cell_ori_inds, K)
    ts = np.arange(0.0, 40 * 0.025 + 1e-10, 0.025)
    ref = O.run(eta0, ts, a, pc.mesh_h_xyz, pc.unique_grain_directions, laser=L.as_dict(),
                cent=O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz))
    for dtype, tol in ((torch.float64, 1e-11), (torch.float32, 2e-5)):
        s = _solver(a, pc, dtype, cuda)
        got = s.run(s.ini_cond(), ts, L).double().cpu().numpy()
        assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref)), dtype


def test_t_buffer_matches_laser_and_oracle(cuda):
    Nx, Ny, Nz, K = 20, 14, 9, 10
    a = args(Nx, Ny, Nz, K, anisotropy=0.15, force_liquid=True)
    pc = poly(Nx, Ny, Nz, K)
    L = rosenthal(Nx, Ny, Nz)
    eta = random_eta(Nx * Ny * Nz, K, seed=9)
    cent = O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz)
    T = O.laser_T(L.as_dict(), cent, 0.4)
    s = _solver(a, pc, torch.float64, cuda)
    for Tin in (T, T.astype(np.float32)):
        (eta_t, t), _ = s.stepper((torch.as_tensor(eta, device=cuda), 0.4), 0.43, [torch.as_tensor(Tin, device=cuda)])
        ref = O.explicit_euler(eta, 0.4, 0.43, Tin.astype(np.float64), a, pc.mesh_h_xyz, pc.unique_grain_directions)
        assert t == 0.43
        assert _rel(eta_t.cpu().numpy(), ref) <= 1e-12


def test_multistep_f64_trajectory(cuda):
    Nx, Ny, Nz, K = 24, 16, 10, 10
    a = args(Nx, Ny, Nz, K, anisotropy=0.15, force_liquid=True)
    pc = poly(Nx, Ny, Nz, K)
    L = gaussian(Nx, Ny, Nz, vel=3.0)
    eta0 = O.ini_cond(pc.cell_ori_inds, K)
    ts = np.arange(0.0, 30 * 0.03 + 1e-10, 0.03)
    s = _solver(a, pc, torch.float64, cuda)
    state = (s.ini_cond(), ts[0])
    for t in ts[1:]:
        state, _ = s.stepper(state, t, [L])
    ref = O.run(eta0, ts, a, pc.mesh_h_xyz, pc.unique_grain_directions, laser=L.as_dict(),
                cent=O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz))
    assert _rel(state[0].cpu().numpy(), ref) <= 1e-11


def test_rhs_matches_oracle(cuda):
    Nx, Ny, Nz, K = 13, 11, 7, 10
    for eps in (0.0, 0.15):
        a = args(Nx, Ny, Nz, K, anisotropy=eps)
        pc = poly(Nx, Ny, Nz, K)
        L = gaussian(Nx, Ny, Nz)
        eta = random_eta(Nx * Ny * Nz, K, seed=2)
        s = _solver(a, pc, torch.float64, cuda)
        got = s.state_rhs(torch.as_tensor(eta, device=cuda), 0.2, [L]).cpu().numpy()
        T = O.laser_T(L.as_dict(), O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz), 0.2)
        ref = O.state_rhs(eta, T, a, pc.mesh_h_xyz, pc.unique_grain_directions)
        assert _rel(got, ref) <= 1e-12


def test_temperature_models(cuda):
    Nx, Ny, Nz, K = 16, 12, 10, 4
    a = args(Nx, Ny, Nz, K)
    pc = poly(Nx, Ny, Nz, K)
    s = _solver(a, pc, torch.float64, cuda)
    cent = O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz)
    for L in (rosenthal(Nx, Ny, Nz), gaussian(Nx, Ny, Nz), quench(Nz)):
        for t in (0.0, 1.7):
            got = s.temperature(L, t)
            ref = O.laser_T(L.as_dict(), cent, t)
            assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-14


def test_init_argmax_and_reduce(cuda):
    Nx, Ny, Nz, K = 19, 13, 11, 10
    a = args(Nx, Ny, Nz, K, anisotropy=0.15)
    pc = poly(Nx, Ny, Nz, K)
    L = gaussian(Nx, Ny, Nz)
    s = _solver(a, pc, torch.float64, cuda)
    eta0 = s.ini_cond()
    assert np.array_equal(eta0.cpu().numpy(), O.ini_cond(pc.cell_ori_inds, K))
    eta = random_eta(Nx * Ny * Nz, K, seed=4)
    eta[3, :] = 0.5          # ties: first maximum wins
    eta[5, 2] = np.nan       # NaN counts as the maximum
    et = torch.as_tensor(eta, device=cuda)
    am = s.argmax(et).cpu().numpy()
    assert np.array_equal(am, np.argmax(eta, axis=1))
    eta[5, 2] = 0.25
    et = torch.as_tensor(eta, device=cuda)
    d = s.diagnostics(et, L, t=0.5, eta0=eta0)
    T = O.laser_T(L.as_dict(), O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz), 0.5)
    Fl, Fg = O.free_energy(eta, T, a, pc.mesh_h_xyz)
    assert abs(d["F_local"] - Fl) <= 1e-12 * abs(Fl)
    assert abs(d["F_grad"] - Fg) <= 1e-12 * abs(Fg)
    assert np.allclose(d["volumes"], O.volumes(eta, pc.mesh_h_xyz), rtol=1e-13, atol=0)
    ins = O.inspect(eta, O.ini_cond(pc.cell_ori_inds, K), T, a)
    assert d["changed"] == round(ins["change_pct"] / 100 * Nx * Ny * Nz)
    assert d["liquid"] == round(ins["liquid_pct"] / 100 * Nx * Ny * Nz)
    assert d["max_T"] == pytest.approx(ins["max_T"], rel=1e-14)
    assert np.array_equal(d["argmax_counts"], np.bincount(np.argmax(eta, axis=1), minlength=K))
    assert d["nonfinite"] == 0
    # inspect_sol raises on non-finite state, like allen_cahn.py:179-180
    eta[7, 1] = np.inf
    with pytest.raises(ValueError):
        s.inspect_sol(torch.as_tensor(eta, device=cuda), eta0, L, np.arange(3) * 0.1, 1)


def test_reduce_is_deterministic_and_dtype_consistent(cuda):
    Nx, Ny, Nz, K = 64, 48, 20, 20
    a = args(Nx, Ny, Nz, K, anisotropy=0.15)
    pc = poly(Nx, Ny, Nz, K)
    L = gaussian(Nx, Ny, Nz)
    s = _solver(a, pc, torch.float32, cuda)
    eta = torch.as_tensor(random_eta(Nx * Ny * Nz, K, seed=8), dtype=torch.float32, device=cuda)
    d1 = s.diagnostics(eta, L, 0.2)
    d2 = s.diagnostics(eta, L, 0.2)
    assert d1["F"] == d2["F"] and np.array_equal(d1["volumes"], d2["volumes"])


def test_autotune_times_every_candidate_without_touching_the_state(cuda):
    """autotune.tune times each (tile, zchunk) candidate over consecutive steps in buffers of its
    own: every candidate gets a time, the choice is one of them, the caller's state is unchanged,
    and the chosen shape steps bitwise like the generic kernel."""
    from pfkit import autotune
    Nx, Ny, Nz, K = 24, 16, 12, 8
    a = args(Nx, Ny, Nz, K, anisotropy=0.15)
    pc = poly(Nx, Ny, Nz, K)
    L = gaussian(Nx, Ny, Nz)
    s = _solver(a, pc, torch.float32, cuda)
    eta = torch.as_tensor(random_eta(Nx * Ny * Nz, K, seed=9), dtype=torch.float32, device=cuda)
    before = eta.clone()
    best, times = autotune.tune(s.state_rhs, eta, L, 0.1, 1e-9, reps=2, warmup=1, store=False)
    assert set(times) == set(autotune.candidates(s.state_rhs.pf_dtype, K, Nz))
    assert best in times and all(t > 0.0 for t in times.values())
    assert torch.equal(eta, before)
    got, ref = torch.empty_like(eta), torch.empty_like(eta)
    s.state_rhs.step(eta, got, L, 0.1, 0.1 + 1e-9, best[0], best[1])
    s.state_rhs.step(eta, ref, L, 0.1, 0.1 + 1e-9, -2, 0)
    assert torch.equal(got, ref)
