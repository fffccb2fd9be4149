"""The oracle and the GPU kernels against input/output vectors produced by
the reference's own code (tests/golden/make_reference_vectors.py): the
functions no closed form pins -- directional form factors and direction
maps, injection with BRDFs, the directional exchange, collection,
Stokes/Nusselt form factors and visibility on an occluded scene.

Tolerances: bit-exact for booleans and indices; float64 quantities to the
rounding the respective formula can accumulate (see each test).
"""
import os

import numpy as np
import pytest

import oracle as orc
from sparrowpy_amd import geometry
from conftest import rel_err_within_db

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"))


def _dir_inputs():
    """Patches, normals, centers, areas and wall ids of the directional
    scene, from this package's host geometry (checked bit-exact against
    the reference's subdivision in test_host_geometry)."""
    pts, nrm, _, ids = geometry.process_patches(
        G["dir_walls_points"], G["dir_walls_normal"], 1.0)
    return (pts, nrm, geometry.calculate_center(pts),
            geometry.calculate_area(pts), ids)


# ---- host geometry vs reference vectors (CPU) ---------------------------------
def test_host_geometry():
    pts, nrm, ctr, area, ids = _dir_inputs()
    np.testing.assert_array_equal(pts, G["dir_patches_points"])
    np.testing.assert_array_equal(ids, G["dir_patch_wall_ids"])
    vp, _, _, _ = geometry.process_patches(G["vis_walls_points"],
                                           G["vis_walls_normal"], 1.0)
    np.testing.assert_array_equal(vp, G["vis_patches_points"])


# ---- oracle vs reference vectors (CPU) ---------------------------------------
def test_oracle_visibility_and_direction_map():
    pts, nrm, ctr, area, ids = _dir_inputs()
    vis = orc.p2p_visibility(ctr, nrm, pts)
    np.testing.assert_array_equal(vis, G["dir_visibility"])
    P = pts.shape[0]
    D = G["dir_out_dirs"].shape[1]
    p2o = np.full((P, P), D, dtype=np.int64)
    vv = vis | vis.T
    for j in range(P):
        v = np.nonzero(vv[:, j])[0]
        p2o[v, j] = orc.receiver_index(ctr[v], ctr[j], G["dir_out_dirs"],
                                       ids[v])
    np.testing.assert_array_equal(p2o, G["dir_p2o"])


def test_oracle_occluded_visibility():
    wp, wn = G["vis_walls_points"], G["vis_walls_normal"]
    pts, nrm, _, _ = geometry.process_patches(wp, wn, 1.0)
    ctr = geometry.calculate_center(pts)
    vis = orc.p2p_visibility(ctr, nrm, pts)
    assert 0 < vis.sum() < vis.size  # the recesses occlude some pairs
    np.testing.assert_array_equal(vis, G["vis_matrix"])
    for q, ref in zip(G["vis_query_points"], G["vis_point_vis"]):
        np.testing.assert_array_equal(orc.point_visibility(q, ctr, wn, wp),
                                      ref)


def test_oracle_form_factors():
    """Stokes and Nusselt: ln/acos/pow may differ by an ulp between libms;
    the contour integral cancels ~1e4-fold, so 1e-11 relative."""
    pts, nrm, ctr, area, ids = _dir_inputs()
    pairs = np.argwhere(G["dir_visibility"]).astype(np.int32)
    ff = orc.form_factors(pts, nrm, area, pairs)
    np.testing.assert_allclose(ff, G["dir_form_factors"], rtol=1e-11,
                               atol=1e-16)
    for k in range(G["ff_values"].size):
        sp = G["ff_src_points"][k]
        v = orc.universal_ff(sp, G["ff_src_normals"][k],
                             geometry.polygon_area(sp), G["ff_rcv_points"][k],
                             G["ff_rcv_normals"][k])
        np.testing.assert_allclose(v, G["ff_values"][k], rtol=1e-11)


def test_oracle_directional_form_factor():
    """Q1: the sending wall's BRDF and incoming-direction set, exp(-air)
    on a unit length -- exactly the reference's F~."""
    pts, nrm, ctr, area, ids = _dir_inputs()
    fft = orc.fft(G["dir_visibility"], G["dir_form_factors"], 2, ctr, area,
                  G["dir_air"], ids, G["dir_brdf"], G["dir_brdf_index"],
                  G["dir_in_dirs"], G["dir_out_dirs"])
    np.testing.assert_allclose(fft, G["dir_fft"], rtol=1e-15, atol=0)


def test_oracle_injection():
    pts, nrm, ctr, area, ids = _dir_inputs()
    src = G["dir_source"]
    vis = orc.point_visibility(src, ctr, G["dir_walls_normal"],
                               G["dir_walls_points"])
    np.testing.assert_array_equal(vis, G["dir_source_visibility"])
    e0, d0 = orc.source_energy(src, ctr, pts, vis, G["dir_air"], 2)
    # The reference's source distance is np.linalg.norm of a 3-vector, i.e.
    # BLAS ddot (FMA-accumulated here; dnrm2 under numba): it may differ from
    # the oracle's (x0*x0 + x1*x1) + x2*x2 in the last ulp.  The delays
    # derived from it must agree exactly.
    np.testing.assert_allclose(d0, G["dir_distance0"], rtol=4.5e-16, atol=0)
    c, dt = float(G["dir_c"]), float(G["dir_dt"])
    assert [orc.delay_floor(d, c, dt) for d in d0] == \
        [orc.delay_floor(d, c, dt) for d in G["dir_distance0"]]
    np.testing.assert_allclose(e0, G["dir_energy0"], rtol=1e-11)
    e0d = orc.add_directional(G["dir_energy0"], src, ctr, ids,
                              G["dir_in_dirs"], G["dir_out_dirs"],
                              G["dir_brdf"], G["dir_brdf_index"])
    np.testing.assert_allclose(e0d, G["dir_energy0_dir"], rtol=1e-15, atol=0)


def _dist(ctr):
    d = ctr[:, None, :] - ctr[None, :, :]
    return np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1])
                   + d[..., 2] * d[..., 2])


def test_oracle_exchange():
    """Same inputs, same accumulation order: the FP64 restatement of
    _energy_exchange reproduces the reference's histogram to the last
    few ulps."""
    pts, nrm, ctr, area, ids = _dir_inputs()
    pairs = np.argwhere(G["dir_visibility"]).astype(np.int32)
    etc = orc.exchange(int(G["dir_T"]), G["dir_energy0_dir"],
                       G["dir_distance0"], _dist(ctr), G["dir_fft"],
                       G["dir_p2o"], float(G["dir_c"]), float(G["dir_dt"]),
                       int(G["dir_K"]), pairs)
    np.testing.assert_allclose(etc, G["dir_etc"], rtol=1e-13, atol=1e-300)


def test_oracle_collection():
    pts, nrm, ctr, area, ids = _dir_inputs()
    rcv = G["dir_receiver"]
    vis = orc.point_visibility(rcv, ctr, G["dir_walls_normal"],
                               G["dir_walls_points"])
    np.testing.assert_array_equal(vis, G["dir_receiver_visibility"])
    g = orc.receiver_energy(rcv, pts, vis)
    np.testing.assert_allclose(g, G["dir_receiver_factor"], rtol=1e-11)
    idx = orc.receiver_index(ctr, rcv, G["dir_out_dirs"], ids)
    np.testing.assert_array_equal(idx, G["dir_receiver_index"])
    P = pts.shape[0]
    E = G["dir_etc"][np.arange(P), idx] * G["dir_receiver_factor"][:, None,
                                                                    None]
    d = ctr - rcv
    dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    out = orc.collect_receiver_energy(E, dist, float(G["dir_c"]),
                                      float(G["dir_dt"]), G["dir_air"])
    np.testing.assert_allclose(out, G["dir_collected"], rtol=1e-14, atol=0)


# ---- GPU kernels vs reference vectors -----------------------------------------
@pytest.fixture(scope="module")
def ops():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sparrowpy_amd import ops as _ops
    return _ops


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.gpu
def test_gpu_against_reference_vectors(ops):
    pts, nrm, ctr, area, ids = _dir_inputs()
    vis = _np(ops.check_patch2patch_visibility(ctr, nrm, pts)).astype(bool)
    np.testing.assert_array_equal(vis, G["dir_visibility"])
    pairs = _np(ops.visible_pairs(vis))
    ff = _np(ops.patch2patch_ff_universal(pts, nrm, area, pairs))
    np.testing.assert_allclose(ff, G["dir_form_factors"], rtol=1e-11,
                               atol=1e-16)
    p2o = _np(ops.direction_map(vis, ctr, G["dir_out_dirs"], ids))
    np.testing.assert_array_equal(p2o, G["dir_p2o"])
    fft = _np(ops.form_factors_with_directivity_dim(
        vis, G["dir_form_factors"], 2, ctr, area, G["dir_air"], ids,
        G["dir_brdf"], G["dir_brdf_index"], G["dir_in_dirs"],
        G["dir_out_dirs"]))
    np.testing.assert_allclose(fft, G["dir_fft"], rtol=1e-13, atol=0)
    src = G["dir_source"]
    svis = _np(ops.check_point2patch_visibility(
        src, ctr, G["dir_walls_normal"], G["dir_walls_points"]))[0]
    np.testing.assert_array_equal(svis.astype(bool),
                                  G["dir_source_visibility"])
    e0, e0d, d0 = ops.source_energy(src, ctr, pts, svis, G["dir_air"], 2,
                                    ids, G["dir_in_dirs"], G["dir_brdf"],
                                    G["dir_brdf_index"], 16)
    np.testing.assert_allclose(_np(d0), G["dir_distance0"], rtol=4.5e-16,
                               atol=0)
    np.testing.assert_allclose(_np(e0d), G["dir_energy0_dir"], rtol=1e-10)
    T, c, dt, K = (int(G["dir_T"]), float(G["dir_c"]), float(G["dir_dt"]),
                   int(G["dir_K"]))
    for kind in ("generic", "grouped"):
        args = (vis, ctr, area, G["dir_form_factors"], G["dir_p2o"],
                G["dir_air"], ids, G["dir_in_dirs"], G["dir_brdf"],
                G["dir_brdf_index"], 16, 2, c, dt)
        plan = (ops.GroupedPlan if kind == "grouped"
                else ops.ExchangePlan).from_geometry(*args)
        st = ops.ExchangeState(pts.shape[0], 16, 2, T)
        st.init(G["dir_energy0_dir"], G["dir_distance0"], c, dt)
        for _ in range(K):
            st.step(plan)
        etc = _np(st.finalize())
        st.check()
        err = rel_err_within_db(etc, G["dir_etc"])
        print(f"[golden exchange {kind}] rel err {err:.3e}")
        assert err <= 1e-5
    rcv = G["dir_receiver"]
    rv = ops.check_point2patch_visibility(rcv, ctr, G["dir_walls_normal"],
                                          G["dir_walls_points"])
    np.testing.assert_array_equal(_np(rv)[0].astype(bool),
                                  G["dir_receiver_visibility"])
    g = ops.receiver_energy(rcv, pts, rv)
    np.testing.assert_allclose(_np(g)[0], G["dir_receiver_factor"],
                               rtol=1e-10)
    idx = ops.point_direction_index(rcv, ctr, G["dir_out_dirs"], ids)
    np.testing.assert_array_equal(_np(idx)[0], G["dir_receiver_index"])
    pw = _np(ops.collect(G["dir_etc"].astype(np.float32), rcv, ctr, idx, g,
                         G["dir_air"], c, dt, patchwise=True))[0]
    assert rel_err_within_db(pw, G["dir_collected"]) <= 1e-5


@pytest.mark.gpu
def test_gpu_occluded_visibility_against_reference(ops):
    wp, wn = G["vis_walls_points"], G["vis_walls_normal"]
    pts, nrm, _, _ = geometry.process_patches(wp, wn, 1.0)
    ctr = geometry.calculate_center(pts)
    vis = _np(ops.check_patch2patch_visibility(ctr, nrm, pts)).astype(bool)
    np.testing.assert_array_equal(vis, G["vis_matrix"])
    pv = _np(ops.check_point2patch_visibility(G["vis_query_points"], ctr, wn,
                                              wp)).astype(bool)
    np.testing.assert_array_equal(pv, G["vis_point_vis"])


@pytest.mark.gpu
def test_gpu_universal_form_factors_against_reference(ops):
    """Each pair as a two-patch scene: Stokes for the non-coplanar-edge
    pairs, Nusselt where the patches share an edge or vertex."""
    for k in range(G["ff_values"].size):
        pts = np.stack([G["ff_src_points"][k], G["ff_rcv_points"][k]])
        nrm = np.stack([G["ff_src_normals"][k], G["ff_rcv_normals"][k]])
        area = np.array([geometry.polygon_area(p) for p in pts])
        ff = _np(ops.patch2patch_ff_universal(pts, nrm, area,
                                              np.array([[0, 1]])))
        np.testing.assert_allclose(ff[0, 1], G["ff_values"][k], rtol=1e-11)
