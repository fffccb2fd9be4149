"""GPU adjoint for BRDF rows of any form (sparrowpy_amd.adjoint.
BRDFExchangeAdjoint, the grouped plan) against the FP64 reference gradient
with respect to the stored BRDF table (oracle/adjoint.py, pinned for such
tables by central differences in test_oracle_adjoint.py), against the
diffuse + specular adjoint where both apply, and against the drop-in class
path for the forward.  Tolerance: 1e-4 of the largest gradient component
(the GPU histograms are float32)."""
import numpy as np
import pytest

from conftest import rel_err_within_db
from oracle import adjoint as oadj
from test_oracle_adjoint import (FREQS, _setup, directional_tables,
                                 loss_and_grad, oracle_scene,
                                 oracle_scene_general, run, spec_info)
from sparrowpy_amd.coordinates import Coordinates, FrequencyData

pytestmark = pytest.mark.gpu

AIR = np.array([2e-3, 5e-3])


@pytest.fixture(scope="module")
def torch():
    import torch as _torch
    if not _torch.cuda.is_available():
        pytest.skip("no GPU")
    return _torch


def gpu_scene_general(walls, dirs, freqs):
    from sparrowpy_amd import DirectionalRadiosityFast
    r = DirectionalRadiosityFast.from_polygon(walls, .5)
    for n, f in enumerate(freqs):
        r.set_wall_brdf([n], FrequencyData(f, FREQS), dirs, dirs)
    r.set_air_attenuation(FrequencyData(AIR, FREQS))
    r.bake_geometry()
    return r


def _brdf_adjoint(r, src, rcv, T, c, dt, K, every):
    from sparrowpy_amd import adjoint
    return adjoint.BRDFExchangeAdjoint(r, src, Coordinates(*rcv.T), c, dt,
                                       T * dt, K, checkpoint_every=every)


def test_brdf_forward_matches_class_path(torch):
    """Histograms from the stored table equal the drop-in class's."""
    walls, dirs, _, _, _, _, src, rcv, (T, c, dt, K) = _setup()
    r = gpu_scene_general(walls, dirs, directional_tables(dirs, len(walls),
                                                          4))
    adj = _brdf_adjoint(r, src, rcv, T, c, dt, K, 2)
    assert adj.table0.shape == (len(walls) * dirs.csize, dirs.csize, 2)
    H = adj.forward(adj.table0).cpu().numpy()
    r.init_source_energy(Coordinates(*src))
    r.calculate_energy_exchange(c, dt, T * dt, K)
    ref = r.collect_energy_receiver_mono(Coordinates(*rcv.T)).time
    assert rel_err_within_db(H, ref) <= 1e-5


@pytest.mark.parametrize("every", [1, 2])
def test_brdf_gradient_matches_fp64_reference(torch, every):
    """dL/dtable through torch autograd against the FP64 reference, with
    directional-scattering BRDFs (rows of no diffuse + specular form)."""
    from sparrowpy_amd import adjoint
    walls, dirs, _, _, _, _, src, rcv, (T, c, dt, K) = _setup()
    target = run(oracle_scene_general(walls, dirs,
                                      directional_tables(dirs, len(walls), 3)),
                 src, rcv, T, c, dt, K)
    f0 = directional_tables(dirs, len(walls), 4)
    ro = oracle_scene_general(walls, dirs, f0)
    L, gH = loss_and_grad(run(ro, src, rcv, T, c, dt, K), target)
    g_ref = oadj.gradient(ro, rcv, gH, K)
    adj = _brdf_adjoint(gpu_scene_general(walls, dirs, f0), src, rcv, T, c,
                        dt, K, every)
    table = adj.table0.clone().requires_grad_(True)
    loss = adjoint.log_energy_loss(adjoint.receiver_histograms_brdf(adj,
                                                                    table),
                                   target)
    loss.backward()
    g = table.grad.cpu().numpy()
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    print(f"[brdf adjoint every={every}] loss gpu {float(loss):.6e} fp64 "
          f"{L:.6e}, max grad err / max grad {err:.2e}")
    assert abs(float(loss) - L) <= 1e-5 * L
    assert err <= 1e-4


@pytest.mark.parametrize("n_points", [(8, 8), (8, 16)])
def test_brdf_gradient_many_directions(torch, n_points):
    """32 and 64 directions (the matrix-core tiles of the table gradient and
    of the transposed spread at their larger sizes): dL/dtable against the
    FP64 reference."""
    from sparrowpy_amd import adjoint, samplings
    walls, _, _, _, _, _, src, rcv, (T, c, dt, K) = _setup()
    dirs = samplings.hemisphere(samplings.sph_gaussian(n_points=n_points))
    target = run(oracle_scene_general(walls, dirs,
                                      directional_tables(dirs, len(walls), 3)),
                 src, rcv, T, c, dt, K)
    f0 = directional_tables(dirs, len(walls), 4)
    ro = oracle_scene_general(walls, dirs, f0)
    L, gH = loss_and_grad(run(ro, src, rcv, T, c, dt, K), target)
    g_ref = oadj.gradient(ro, rcv, gH, K)
    adj = _brdf_adjoint(gpu_scene_general(walls, dirs, f0), src, rcv, T, c,
                        dt, K, 2)
    assert adj.D == dirs.csize
    table = adj.table0.clone().requires_grad_(True)
    loss = adjoint.log_energy_loss(adjoint.receiver_histograms_brdf(adj,
                                                                    table),
                                   target)
    loss.backward()
    err = np.abs(table.grad.cpu().numpy() - g_ref).max() / np.abs(g_ref).max()
    print(f"[brdf adjoint D={dirs.csize}] max grad err / max grad {err:.2e}")
    assert abs(float(loss) - L) <= 1e-5 * L
    assert err <= 1e-4


def test_brdf_adjoint_agrees_with_diffuse_specular_adjoint(torch):
    """On create_from_scattering BRDFs both adjoints apply: the table
    gradient chained to (alpha, s) equals the diffuse + specular adjoint's
    gradient, and both equal the FP64 reference."""
    from sparrowpy_amd import adjoint
    from test_gpu_adjoint import _adjoint
    walls, dirs, ta, ts, alpha, scat, src, rcv, (T, c, dt, K) = _setup()
    target = run(oracle_scene(walls, dirs, ta, ts), src, rcv, T, c, dt, K)
    sv, m = spec_info(dirs)
    ro = oracle_scene(walls, dirs, alpha, scat)
    L, gH = loss_and_grad(run(ro, src, rcv, T, c, dt, K), target)
    ga_ref, gs_ref = oadj.scattering_gradient(oadj.gradient(ro, rcv, gH, K),
                                              alpha, scat, sv, m)
    r, ds = _adjoint(torch, walls, dirs, alpha, scat, src, rcv, T, c, dt, K,
                     2)
    kw = dict(dtype=torch.float64, device="cuda")
    a = torch.tensor(alpha, requires_grad=True, **kw)
    s = torch.tensor(scat, requires_grad=True, **kw)
    adjoint.log_energy_loss(adjoint.receiver_histograms(
        ds, *adjoint.scattering_tables(a, s, sv)), target).backward()
    adj = _brdf_adjoint(r, src, rcv, T, c, dt, K, 2)
    table = adj.table0.clone().requires_grad_(True)
    adjoint.log_energy_loss(adjoint.receiver_histograms_brdf(adj, table),
                            target).backward()
    ga, gs = oadj.scattering_gradient(table.grad.cpu().numpy(), alpha, scat,
                                      sv, m)
    gmax = max(np.abs(ga_ref).max(), np.abs(gs_ref).max())
    for x, y in ((ga, a.grad.cpu().numpy()), (gs, s.grad.cpu().numpy()),
                 (ga, ga_ref), (gs, gs_ref)):
        assert np.abs(x - y).max() <= 1e-4 * gmax


def test_brdf_two_calls_in_one_graph(torch):
    """Two calls with different tables in one loss keep their own records;
    a second backward (retain_graph) doubles the gradient; without a
    record backward refuses."""
    from sparrowpy_amd import adjoint
    walls, dirs, _, _, _, _, src, rcv, (T, c, dt, K) = _setup()
    target = run(oracle_scene_general(walls, dirs,
                                      directional_tables(dirs, len(walls), 3)),
                 src, rcv, T, c, dt, K)
    f0 = directional_tables(dirs, len(walls), 4)
    f1 = [f * 0.9 for f in f0]     # the stored table times 0.9

    def ref_grad(f):
        ro = oracle_scene_general(walls, dirs, f)
        gH = loss_and_grad(run(ro, src, rcv, T, c, dt, K), target)[1]
        return oadj.gradient(ro, rcv, gH, K)

    # d/dtable [L(table) + L(0.9 table)] = g(f0) + 0.9 g(f1)
    g_exp = ref_grad(f0) + 0.9 * ref_grad(f1)
    adj = _brdf_adjoint(gpu_scene_general(walls, dirs, f0), src, rcv, T, c,
                        dt, K, 2)
    table = adj.table0.clone().requires_grad_(True)
    loss = adjoint.log_energy_loss(
        adjoint.receiver_histograms_brdf(adj, table), target) + \
        adjoint.log_energy_loss(
            adjoint.receiver_histograms_brdf(adj, 0.9 * table), target)
    loss.backward(retain_graph=True)
    g = table.grad.cpu().numpy().copy()
    gmax = np.abs(g_exp).max()
    assert np.abs(g - g_exp).max() <= 1e-4 * gmax
    loss.backward()
    assert np.abs(table.grad.cpu().numpy() - 2 * g_exp).max() <= \
        2e-4 * gmax
    adj.release()
    with pytest.raises(RuntimeError, match="no forward record"):
        adj.backward(torch.zeros((len(rcv), 2, T), dtype=torch.float64,
                                 device="cuda"))


def test_brdf_gradient_without_visible_pairs(torch):
    """Floor patches only (no visible pair, no group): the gradient flows
    through injection and collection alone."""
    from sparrowpy_amd import adjoint, samplings, scenes, testing
    floor = [w for w in testing.shoebox_room_stub(2, 1.5, 1)
             if np.allclose(w.normal, [0, 0, 1])]
    walls = scenes._split_walls_into_patches(floor, .5)
    dirs = samplings.hemisphere(samplings.sph_gaussian(sh_order=3))
    src = np.array([.6, .5, .4])
    rcv = np.array([[1.4, 1.0, .6], [.9, .4, .7]])
    T, c, dt, K = 120, 343., 1 / 2000, 3
    target = run(oracle_scene_general(walls, dirs,
                                      directional_tables(dirs, len(walls), 3)),
                 src, rcv, T, c, dt, K)
    f0 = directional_tables(dirs, len(walls), 4)
    ro = oracle_scene_general(walls, dirs, f0)
    assert len(ro.visible_patches) == 0
    L, gH = loss_and_grad(run(ro, src, rcv, T, c, dt, K), target)
    g_ref = oadj.gradient(ro, rcv, gH, K)
    adj = _brdf_adjoint(gpu_scene_general(walls, dirs, f0), src, rcv, T, c,
                        dt, K, 2)
    assert adj.plan.n_groups == 0
    table = adj.table0.clone().requires_grad_(True)
    loss = adjoint.log_energy_loss(
        adjoint.receiver_histograms_brdf(adj, table), target)
    loss.backward()
    gmax = np.abs(g_ref).max()
    assert gmax > 0
    assert abs(float(loss) - L) <= 1e-5 * L
    assert np.abs(table.grad.cpu().numpy() - g_ref).max() <= 1e-4 * gmax
