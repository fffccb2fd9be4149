"""GPU adjoint (sparrowpy_amd.adjoint) against the FP64 reference gradient
(oracle/adjoint.py, itself pinned by finite differences) and against
central finite differences of the FP64 oracle forward, on a small room
with one BRDF per patch.  Tolerance: 1e-4 of the largest gradient
component (the GPU histograms are float32)."""
import numpy as np
import pytest

from oracle import adjoint as oadj
from test_oracle_adjoint import (FREQS, _setup, loss_and_grad, oracle_scene,
                                 run, spec_info)
from sparrowpy_amd import brdf as B
from sparrowpy_amd.coordinates import Coordinates, FrequencyData

pytestmark = pytest.mark.gpu

AIR = np.array([2e-3, 5e-3])


@pytest.fixture(scope="module")
def torch():
    import torch as _torch
    if not _torch.cuda.is_available():
        pytest.skip("no GPU")
    return _torch


def gpu_scene(walls, dirs, alpha, scat):
    from sparrowpy_amd import DirectionalRadiosityFast
    r = DirectionalRadiosityFast.from_polygon(walls, .5)
    for n in range(len(walls)):
        br = B.create_from_scattering(
            dirs, dirs, FrequencyData(scat[n][None, :], FREQS),
            FrequencyData(alpha[n][None, :], FREQS))
        r.set_wall_brdf([n], br, dirs, dirs)
    r.set_air_attenuation(FrequencyData(AIR, FREQS))
    r.bake_geometry()
    return r


def _adjoint(torch, walls, dirs, alpha, scat, src, rcv, T, c, dt, K, every):
    from sparrowpy_amd import adjoint
    r = gpu_scene(walls, dirs, alpha, scat)
    adj = adjoint.ExchangeAdjoint(r, src, Coordinates(*rcv.T), c, dt,
                                  T * dt, K, checkpoint_every=every)
    return r, adj


def test_forward_matches_class_path(torch):
    """Histograms from the tables equal the drop-in class's (float32)."""
    from sparrowpy_amd import adjoint
    walls, dirs, ta, ts, alpha, scat, src, rcv, (T, c, dt, K) = _setup()
    r, adj = _adjoint(torch, walls, dirs, alpha, scat, src, rcv, T, c, dt, K,
                      2)
    sv, _ = spec_info(dirs)
    a = torch.tensor(alpha, dtype=torch.float64, device="cuda")
    s = torch.tensor(scat, dtype=torch.float64, device="cuda")
    A, C = adjoint.scattering_tables(a, s, sv)
    np.testing.assert_allclose(A.cpu().numpy(), adj.A0.cpu().numpy(),
                               rtol=1e-14)
    np.testing.assert_allclose(C.cpu().numpy(), adj.C0.cpu().numpy(),
                               rtol=1e-12)
    H = adj.forward(A, C).cpu().numpy()
    r.init_source_energy(Coordinates(*src))
    r.calculate_energy_exchange(c, dt, T * dt, K)
    ref = r.collect_energy_receiver_mono(Coordinates(*rcv.T)).time
    from conftest import rel_err_within_db
    assert rel_err_within_db(H, ref) <= 1e-5


@pytest.mark.parametrize("every", [1, 2, 5])
def test_gradient_matches_fp64_reference(torch, every):
    from sparrowpy_amd import adjoint
    walls, dirs, ta, ts, alpha, scat, src, rcv, (T, c, dt, K) = _setup()
    target = run(oracle_scene(walls, dirs, ta, ts), src, rcv, T, c, dt, K)
    # FP64 reference gradient
    ro = oracle_scene(walls, dirs, alpha, scat)
    Ho = run(ro, src, rcv, T, c, dt, K)
    L, gH = loss_and_grad(Ho, target)
    sv, m = spec_info(dirs)
    ga_ref, gs_ref = oadj.scattering_gradient(
        oadj.gradient(ro, rcv, gH, K), alpha, scat, sv, m)
    # GPU adjoint through torch autograd
    _, adj = _adjoint(torch, walls, dirs, alpha, scat, src, rcv, T, c, dt, K,
                      every)
    a = torch.tensor(alpha, dtype=torch.float64, device="cuda",
                     requires_grad=True)
    s = torch.tensor(scat, dtype=torch.float64, device="cuda",
                     requires_grad=True)
    A, C = adjoint.scattering_tables(a, s, sv)
    H = adjoint.receiver_histograms(adj, A, C)
    loss = adjoint.log_energy_loss(H, target)
    loss.backward()
    ga, gs = a.grad.cpu().numpy(), s.grad.cpu().numpy()
    gmax = max(np.abs(ga_ref).max(), np.abs(gs_ref).max())
    err = max(np.abs(ga - ga_ref).max(), np.abs(gs - gs_ref).max()) / gmax
    print(f"[adjoint every={every}] loss gpu {float(loss):.6e} fp64 {L:.6e}"
          f" max grad err / max grad {err:.2e}")
    assert abs(float(loss) - L) <= 1e-5 * L
    assert err <= 1e-4


def test_two_calls_in_one_graph(torch):
    """Two receiver_histograms calls with different tables in one loss,
    and a second backward through the same graph (retain_graph): each call
    keeps its own checkpoints and tables, so the gradient is the sum of the
    FP64 reference gradients at both points (twice that after the second
    backward)."""
    from sparrowpy_amd import adjoint
    walls, dirs, ta, ts, alpha, scat, src, rcv, (T, c, dt, K) = _setup()
    target = run(oracle_scene(walls, dirs, ta, ts), src, rcv, T, c, dt, K)
    sv, m = spec_info(dirs)
    delta = np.full_like(alpha, .03)
    ga_ref = np.zeros_like(alpha)
    gs_ref = np.zeros_like(scat)
    L_ref = 0.0
    for a_np in (alpha, alpha + delta):
        ro = oracle_scene(walls, dirs, a_np, scat)
        L, gH = loss_and_grad(run(ro, src, rcv, T, c, dt, K), target)
        ga, gs = oadj.scattering_gradient(oadj.gradient(ro, rcv, gH, K),
                                          a_np, scat, sv, m)
        ga_ref += ga
        gs_ref += gs
        L_ref += L
    _, adj = _adjoint(torch, walls, dirs, alpha, scat, src, rcv, T, c, dt, K,
                      2)
    kw = dict(dtype=torch.float64, device="cuda")
    a = torch.tensor(alpha, requires_grad=True, **kw)
    s = torch.tensor(scat, requires_grad=True, **kw)
    d = torch.tensor(delta, **kw)
    H1 = adjoint.receiver_histograms(adj, *adjoint.scattering_tables(a, s,
                                                                      sv))
    H2 = adjoint.receiver_histograms(adj, *adjoint.scattering_tables(a + d,
                                                                      s, sv))
    loss = adjoint.log_energy_loss(H1, target) + \
        adjoint.log_energy_loss(H2, target)
    loss.backward(retain_graph=True)
    ga, gs = a.grad.cpu().numpy().copy(), s.grad.cpu().numpy().copy()
    gmax = max(np.abs(ga_ref).max(), np.abs(gs_ref).max())
    err = max(np.abs(ga - ga_ref).max(), np.abs(gs - gs_ref).max()) / gmax
    print(f"[adjoint two calls] loss gpu {float(loss):.6e} fp64 "
          f"{L_ref:.6e}, max grad err / max grad {err:.2e}")
    assert abs(float(loss) - L_ref) <= 1e-5 * L_ref
    assert err <= 1e-4
    loss.backward()
    ga2, gs2 = a.grad.cpu().numpy(), s.grad.cpu().numpy()
    err2 = max(np.abs(ga2 - 2 * ga_ref).max(),
               np.abs(gs2 - 2 * gs_ref).max()) / (2 * gmax)
    assert err2 <= 1e-4
    # without a record, backward refuses instead of dropping orders
    adj.release()
    with pytest.raises(RuntimeError, match="no forward record"):
        adj.backward(torch.zeros((len(rcv), 2, T), **kw))


def test_gradient_matches_finite_differences(torch):
    """GPU gradient against central differences (h = 1e-4) of the FP64
    oracle loss on 50 random parameters."""
    from sparrowpy_amd import adjoint
    walls, dirs, ta, ts, alpha, scat, src, rcv, (T, c, dt, K) = _setup()
    target = run(oracle_scene(walls, dirs, ta, ts), src, rcv, T, c, dt, K)
    sv, _ = spec_info(dirs)
    _, adj = _adjoint(torch, walls, dirs, alpha, scat, src, rcv, T, c, dt, K,
                      2)
    a = torch.tensor(alpha, dtype=torch.float64, device="cuda",
                     requires_grad=True)
    s = torch.tensor(scat, dtype=torch.float64, device="cuda",
                     requires_grad=True)
    A, C = adjoint.scattering_tables(a, s, sv)
    adjoint.log_energy_loss(adjoint.receiver_histograms(adj, A, C),
                            target).backward()
    ga, gs = a.grad.cpu().numpy(), s.grad.cpu().numpy()
    gmax = max(np.abs(ga).max(), np.abs(gs).max())
    rng = np.random.default_rng(7)
    h = 1e-4
    for kind, n, b in zip(rng.choice(["alpha", "s"], 50),
                          rng.integers(0, len(walls), 50),
                          rng.integers(0, 2, 50)):
        vals = []
        for sign in (1, -1):
            a2, s2 = alpha.copy(), scat.copy()
            (a2 if kind == "alpha" else s2)[n, b] += sign * h
            Hs = run(oracle_scene(walls, dirs, a2, s2), src, rcv, T, c, dt, K)
            vals.append(loss_and_grad(Hs, target)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        g = (ga if kind == "alpha" else gs)[n, b]
        assert abs(g - fd) <= 1e-4 * gmax, (kind, n, b, g, fd)


def test_adam_recovers_materials(torch):
    """Inverse materials: 200 Adam steps through the adjoint towards the
    histograms of known per-wall absorption and scattering (52 one-patch
    walls, 2 bands, 4 receivers, 400 bins, 10 orders), starting from
    every wall's absorption 0.05 too high (plus noise).  Four receivers
    determine the room's decay, i.e. the per-band mean absorption, not
    each of the 208 parameters: the loss must fall by two orders of
    magnitude and the per-band mean absorption error to at most 0.3 of
    its start (0.045 / 0.041 -> 0.008 / 0.007 when measured); the
    per-wall absorption error only has to shrink (0.044 -> 0.024)."""
    from sparrowpy_amd import adjoint
    walls, dirs, ta, ts, alpha, scat, src, _, (T, c, dt, K) = _setup(seed=3)
    rcv = np.array([[1.4, 1.0, .6], [.9, .4, .7], [.4, 1.1, .3],
                    [1.6, .3, .8]])
    T, K = 400, 10
    sv, _ = spec_info(dirs)
    rng = np.random.default_rng(5)
    a0 = np.clip(ta + .05 + rng.normal(0, .03, ta.shape), .01, .99)
    s0 = np.clip(ts + rng.normal(0, .1, ts.shape), .01, .99)
    _, adj = _adjoint(torch, walls, dirs, a0, s0, src, rcv, T, c, dt, K, 5)
    kw = dict(dtype=torch.float64, device="cuda")
    A, C = adjoint.scattering_tables(torch.tensor(ta, **kw),
                                     torch.tensor(ts, **kw), sv)
    target = adj.forward(A, C).to(torch.float64)
    adj.release()
    a, s, losses = adjoint.recover_materials(adj, target, a0, s0, sv,
                                             steps=200, lr=0.02)
    a, s = a.cpu().numpy(), s.cpu().numpy()
    band0 = np.abs(a0.mean(0) - ta.mean(0))
    band1 = np.abs(a.mean(0) - ta.mean(0))
    print(f"[adam] loss {losses[0]:.3e} -> {losses[-1]:.3e}, band-mean "
          f"absorption error {band0.round(4)} -> {band1.round(4)}, "
          f"per-wall |alpha - true| {np.abs(a0 - ta).mean():.4f} -> "
          f"{np.abs(a - ta).mean():.4f}, |s - true| "
          f"{np.abs(s0 - ts).mean():.4f} -> {np.abs(s - ts).mean():.4f}")
    assert losses[-1] < 1e-2 * losses[0]
    assert (band1 <= 0.3 * band0).all()
    assert np.abs(a - ta).mean() < np.abs(a0 - ta).mean()


def test_gradient_without_visible_pairs(torch):
    """A floor split into patches (no patch sees another, so no transposed
    entries): the gradient flows only through injection and collection,
    and still matches the FP64 reference."""
    from sparrowpy_amd import adjoint, samplings, scenes, testing
    floor = [w for w in testing.shoebox_room_stub(2, 1.5, 1)
             if np.allclose(w.normal, [0, 0, 1])]
    walls = scenes._split_walls_into_patches(floor, .5)
    dirs = samplings.hemisphere(samplings.sph_gaussian(sh_order=3))
    n = len(walls)
    rng = np.random.default_rng(1)
    ta, ts = rng.uniform(.05, .4, (n, 2)), rng.uniform(.1, .9, (n, 2))
    alpha = np.clip(ta + rng.normal(0, .05, (n, 2)), .02, .6)
    scat = np.clip(ts + rng.normal(0, .1, (n, 2)), .05, .95)
    src = np.array([.6, .5, .4])
    rcv = np.array([[1.4, 1.0, .6], [.9, .4, .7]])
    T, c, dt, K = 120, 343., 1 / 2000, 3
    target = run(oracle_scene(walls, dirs, ta, ts), src, rcv, T, c, dt, K)
    ro = oracle_scene(walls, dirs, alpha, scat)
    assert len(ro.visible_patches) == 0
    L, gH = loss_and_grad(run(ro, src, rcv, T, c, dt, K), target)
    sv, m = spec_info(dirs)
    ga_ref, gs_ref = oadj.scattering_gradient(
        oadj.gradient(ro, rcv, gH, K), alpha, scat, sv, m)
    _, adj = _adjoint(torch, walls, dirs, alpha, scat, src, rcv, T, c, dt, K,
                      2)
    a = torch.tensor(alpha, dtype=torch.float64, device="cuda",
                     requires_grad=True)
    s = torch.tensor(scat, dtype=torch.float64, device="cuda",
                     requires_grad=True)
    A, C = adjoint.scattering_tables(a, s, sv)
    loss = adjoint.log_energy_loss(adjoint.receiver_histograms(adj, A, C),
                                   target)
    loss.backward()
    ga, gs = a.grad.cpu().numpy(), s.grad.cpu().numpy()
    gmax = max(np.abs(ga_ref).max(), np.abs(gs_ref).max())
    assert gmax > 0
    assert abs(float(loss) - L) <= 1e-5 * L
    assert max(np.abs(ga - ga_ref).max(), np.abs(gs - gs_ref).max()) <= \
        1e-4 * gmax
