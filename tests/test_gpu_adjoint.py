"""GPU adjoint (dtm_adjoint_design_gradient, FP32) against the FP64 oracle:
J to the FP32 parity bound, dJ/d(1/eps) against FP64 central differences at
16 seeded design voxels to 1e-3 (relative, over the 16-vector)."""
import numpy as np
import pytest

from adjoint_case import fd_gradient, objective, seeded_voxels, small_case
from dtmaxwell4_amd import abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sweep", [abi.SWEEP_TWOPASS, abi.SWEEP_FUSED], ids=["twopass", "fused"])
def test_adjoint_gradient_matches_finite_differences(sweep):
    from dtmaxwell4_amd.solver import Solver
    cfg, d0, eps, mons, omega, steps = small_case()
    cfg.sweep = sweep
    s = Solver(cfg)
    s.set_design_eps(*d0, eps)
    J, grad = s.design_gradient(steps, omega, mons)
    Jo = objective(cfg, d0, eps, mons, omega, steps)
    print(f"J gpu {J:.9e} oracle {Jo:.9e}")
    assert abs(J / Jo - 1) < 1e-5
    vox = seeded_voxels(eps.shape, n=16)
    gfd = fd_gradient(cfg, d0, eps, mons, omega, steps, vox)
    ga = np.array([grad[v] for v in vox])
    err = np.linalg.norm(ga - gfd) / np.linalg.norm(gfd)
    print("adjoint vs FD rel err", err)
    assert err < 1e-3
    # the call leaves the solver at step 0 with zero fields
    assert s.time_step == 0
    assert all(np.abs(s.field(c)).max() == 0 for c in range(6))
    # and a second call reproduces the first
    J2, grad2 = s.design_gradient(steps, omega, mons)
    assert J2 == J
    np.testing.assert_array_equal(grad2, grad)
    s.close()


@pytest.mark.parametrize("sweep", [abi.SWEEP_TWOPASS, abi.SWEEP_FUSED], ids=["twopass", "fused"])
@pytest.mark.parametrize("every", [37, 100])
def test_adjoint_checkpoints_reproduce_the_history_mode(sweep, every):
    """Checkpoint-and-recompute (whole state every K steps, segments rebuilt
    on the way back; K = 37 leaves a short last segment) recomputes the same
    forward fields, so J and the gradient equal the history mode's exactly."""
    from dtmaxwell4_amd.solver import Solver
    cfg, d0, eps, mons, omega, steps = small_case()
    cfg.sweep = sweep
    s = Solver(cfg)
    s.set_design_eps(*d0, eps)
    J, grad = s.design_gradient(steps, omega, mons)
    s.set_adjoint_checkpoints(every)
    Jc, gc = s.design_gradient(steps, omega, mons)
    print("history J", J, "checkpoint J", Jc, "max |dg|", np.abs(gc - grad).max())
    assert Jc == J
    np.testing.assert_array_equal(gc, grad)
    assert s.time_step == 0
    s.close()


def test_adjoint_rejects_what_it_does_not_cover():
    from dtmaxwell4_amd.solver import Solver
    cfg, d0, eps, mons, omega, steps = small_case()
    s = Solver(cfg)
    with pytest.raises(abi.DTMError):  # no design box yet
        s.design_gradient(10, omega, mons)
    s.set_design_eps(*d0, eps)
    s.step(1)
    with pytest.raises(abi.DTMError):  # not at step 0
        s.design_gradient(10, omega, mons)
    s.close()


def test_autograd_function_matches_pytorch_fp64_unrolled_reference():
    """DesignObjective (HIP forward + hand-written adjoint behind
    torch.autograd) against autograd through the pure-PyTorch FP64 unrolled
    FDTD, over every design voxel: relative L2 <= 1e-3."""
    import torch

    from dtmaxwell4_amd.autograd import design_objective
    from dtmaxwell4_amd.solver import Solver
    from torch_fdtd import run
    cfg, d0, eps, mons, omega, steps = small_case(steps=160)
    rho0 = (eps - 1.0) / 3.0                      # eps = 1 + 3 rho
    s = Solver(cfg)
    rho = torch.tensor(rho0, requires_grad=True)
    J = design_objective(rho, s, steps, omega, mons, d0, eps_bg=1.0, eps_scale=3.0)
    J.backward()
    g_gpu = rho.grad.numpy()
    r = torch.tensor(rho0, requires_grad=True)
    J_ref = run(cfg, 1.0 + 3.0 * r, d0, mons, omega, steps)
    J_ref.backward()
    g_ref = r.grad.numpy()
    Jg, Jr = float(J.detach()), float(J_ref.detach())
    print("J", Jg, Jr)
    assert abs(Jg / Jr - 1) < 1e-5
    err = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    print("autograd vs pytorch reference rel err", err)
    assert err < 1e-3
    s.close()


def test_adjoint_downsized_inverse_design_48():
    """The inverse-design workload downsized to 48^3 (eps = 1 + 11 rho with a
    3-cell box-smoothed seeded rho, checkpoints every K = 200 steps and
    recompute): dJ/d(1/eps) against FP64 central differences at 16 seeded
    voxels (<= 1e-3), and dJ/drho through the torch.autograd.Function against
    autograd of the pure-PyTorch FP64 unrolled FDTD over every voxel (<= 1e-3)."""
    import torch

    from adjoint_case import downsized_case
    from dtmaxwell4_amd.autograd import design_objective
    from dtmaxwell4_amd.solver import Solver
    from torch_fdtd import run
    cfg, d0, eps, mons, omega, steps, rho0 = downsized_case()
    s = Solver(cfg)
    s.set_design_eps(*d0, eps)
    s.set_adjoint_checkpoints(200)
    J, grad = s.design_gradient(steps, omega, mons)
    Jo = objective(cfg, d0, eps, mons, omega, steps)
    print(f"J gpu {J:.9e} oracle {Jo:.9e}")
    assert abs(J / Jo - 1) < 1e-4
    vox = seeded_voxels(eps.shape, n=16)
    gfd = fd_gradient(cfg, d0, eps, mons, omega, steps, vox)
    ga = np.array([grad[v] for v in vox])
    err = np.linalg.norm(ga - gfd) / np.linalg.norm(gfd)
    print("48^3 adjoint (K=200) vs FD rel err", err)
    assert err < 1e-3
    # autograd path against the unrolled FP64 PyTorch reference
    rho = torch.tensor(rho0, requires_grad=True)
    Jt = design_objective(rho, s, steps, omega, mons, d0, eps_bg=1.0, eps_scale=11.0)
    Jt.backward()
    g_gpu = rho.grad.numpy()
    dev = "cuda" if torch.cuda.is_available() else None
    r = torch.tensor(rho0, requires_grad=True)
    J_ref = run(cfg, 1.0 + 11.0 * r, d0, mons, omega, steps, device=dev)
    J_ref.backward()
    g_ref = r.grad.numpy()
    err2 = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    print("48^3 autograd vs PyTorch FP64 unrolled rel err", err2, "J", float(Jt.detach()), float(J_ref.detach()))
    assert abs(float(Jt.detach()) / float(J_ref.detach()) - 1) < 1e-4
    assert err2 < 1e-3
    s.close()


@pytest.mark.parametrize("where", ["beside_cpml", "monitor_in_design"])
def test_adjoint_edge_placements(where):
    """Design box and monitors right beside the CPML slabs (the adjoint is
    exact outside the layer), and one monitor inside the design box (the
    injection and the gradient then run as separate launches): against FP64
    central differences at 8 voxels and against the history mode."""
    from dtmaxwell4_amd.solver import Solver
    cfg, d0, eps, mons, omega, steps = small_case(steps=220)
    if where == "beside_cpml":
        d0 = (cfg.npml[0], 10, 11)
        mons = [(cfg.nx - cfg.npml[1] - 1, 12 + a, 12 + b) for a in range(4) for b in range(4)]
        cfg.soft[0].i = 18
    else:
        mons = mons[:8] + [(d0[0] + 3, d0[1] + 4, d0[2] + 2)]
    s = Solver(cfg)
    s.set_design_eps(*d0, eps)
    J, grad = s.design_gradient(steps, omega, mons)
    vox = seeded_voxels(eps.shape, n=8)
    gfd = fd_gradient(cfg, d0, eps, mons, omega, steps, vox)
    ga = np.array([grad[v] for v in vox])
    err = np.linalg.norm(ga - gfd) / np.linalg.norm(gfd)
    print(where, "adjoint vs FD rel err", err)
    assert err < 1e-3
    s.set_adjoint_checkpoints(50)
    Jc, gc = s.design_gradient(steps, omega, mons)
    assert Jc == J
    np.testing.assert_array_equal(gc, grad)
    s.close()


def test_adjoint_refuses_the_absorbing_layer():
    """A design box reaching into the CPML, a monitor inside it, or the
    split-field layer: the gradient there would not be the exact transpose."""
    from dtmaxwell4_amd.solver import Solver
    cfg, d0, eps, mons, omega, steps = small_case()
    s = Solver(cfg)
    s.set_design_eps(cfg.npml[0] - 1, 10, 11, eps)
    with pytest.raises(abi.DTMError):
        s.design_gradient(10, omega, mons)
    s.close()
    s = Solver(cfg)
    s.set_design_eps(*d0, eps)
    with pytest.raises(abi.DTMError):
        s.design_gradient(10, omega, mons + [(cfg.nx - 2, 14, 14)])
    s.close()
    cfg.pml_kind = abi.PML_SPLIT
    s = Solver(cfg)
    s.set_design_eps(*d0, eps)
    with pytest.raises(abi.DTMError):
        s.design_gradient(10, omega, mons)
    s.close()
