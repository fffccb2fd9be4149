"""FP64 reference gradient of receiver histograms with respect to BRDF
tables and absorption/scattering coefficients.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference has no gradients.  This is the plain reverse-mode
derivative of its forward path on the dense arrays the reference itself
uses (``F~ [P,P,D,B]``, ``E [P,D,B,T]``), numpy glue around the C oracle's
per-order loops (``orc_exchange_step``, ``orc_exchange_adjoint_step``):

* exchange orders E_0..E_K by the loop of ``_energy_exchange``
  (``RadiosityFast.py:1114-1143``: per visible pair, both orientations,
  ``E_{k+1}[j, :, :, n:] += F~[i,j,:,:,None] * E_k[i, p2o[i,j], :, :T-n]``);
* total = sum_k E_k (``:1144``), receiver histograms by
  ``collect_energy_receiver_mono`` (``:570-602``, ``:686-752``,
  ``:1148-1185``: gather row p2o-direction, times G, air factor, circular
  roll by ceil(d/c/dt));
* the adjoint of each of these linear maps written out loop for loop,
  then the chain rule through F~[i,j,:,b] = w_ijb table[k_ij, :, b] and
  E0dir[i,:,b] = E0[i,b] table[k0_i, :, b] (``:1251-1270``, ``:988-1034``)
  to the diffuse/specular table entries (A, C), and through
  ``create_from_scattering`` (``brdf.py:104-118``) to (alpha, s).

Pinned by central finite differences of the FP64 oracle forward
(``tests/test_oracle_adjoint.py``).  Dense arrays bound the size: the
1,072-patch shoebox with 64 directions and 6 bands fits at a few hundred
bins (F~ and its gradient 3.5 GB each).
"""
from __future__ import annotations

import numpy as np

import oracle as orc


def _vis(P, pairs):
    """Upper-triangular visibility (uint8 [P, P]) of a visible-pair list."""
    vis = np.zeros((P, P), np.uint8)
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    vis[pairs[:, 0], pairs[:, 1]] = 1
    return vis


def exchange_orders(T, e0dir, d0, dist, fft, p2o, c, dt, K, pairs):
    """[E_0, ..., E_K] in FP64 (RadiosityFast.py:1037-1143 per order;
    each order by ``orc_exchange_step``: per destination, every visible
    source in ascending order, ``E_{k+1}[j, :, :, n:] += F~[i,j,:,:,None] *
    E_k[i, p2o[i,j], :, :T-n]``)."""
    e0dir = np.ascontiguousarray(e0dir, float)
    P, D, B = e0dir.shape
    vis = _vis(P, pairs)
    lib = orc._load()
    dist = np.ascontiguousarray(dist, float)
    fft = np.ascontiguousarray(fft, float)
    p2o = np.ascontiguousarray(p2o, np.int64)
    E = [orc.init_energy(T, e0dir, d0, c, dt)]
    for _ in range(K):
        cur = np.empty_like(E[-1])
        lib.orc_exchange_step(T, P, D, B, E[-1], cur, dist, fft, p2o, c, dt,
                              vis, orc.default_threads())
        E.append(cur)
    return E


def collect_adjoint(grad, receivers, r):
    """dL/dtotal [P,D,B,T] from dL/dH [R,B,T] for OracleRadiosity ``r``."""
    P, T = r.n_patches, grad.shape[-1]
    D = r.brdf_out.shape[1]
    lam = np.zeros((P, D, r.n_bins, T))
    centers = r.patches_center
    for q, rp in enumerate(np.asarray(receivers, float).reshape(-1, 3)):
        vis = orc.point_visibility(rp, centers, r.walls_normal, r.walls_points)
        g = orc.receiver_energy(rp, r.patches_points, vis)
        idx = orc.receiver_index(centers, rp, r.brdf_out, r.patch_to_wall_ids)
        diff = centers - rp
        dist = np.sqrt((diff[:, 0] * diff[:, 0] + diff[:, 1] * diff[:, 1])
                       + diff[:, 2] * diff[:, 2])
        for p in range(P):
            if g[p] == 0.0:
                continue
            n = orc.delay_ceil(dist[p], r.speed_of_sound,
                               r.etc_time_resolution) % T
            a = np.exp(-r.air * dist[p]) * g[p]            # [B]
            lam[p, idx[p]] += a[:, None] * np.roll(grad[q], -n, axis=-1)
    return lam


def exchange_adjoint(lam, E, dist, fft, p2o, c, dt, pairs):
    """(dL/dF~ [P,P,D,B], dL/dE_0 [P,D,B,T]) for dL/dtotal = lam.

    Order by order from the last (``orc_exchange_adjoint_step``): with
    mu_K = lam, mu_k = lam + A^T mu_{k+1}, where for every directed visible
    pair i -> j (delay n < T, d = p2o[i,j])
    ``mu_k[i, d, :, :T-n] += sum_e F~[i,j,e,:] mu_{k+1}[j, e, :, n:]`` and
    ``dL/dF~[i,j,e,:] += sum_t mu_{k+1}[j,e,:,t+n] E_k[i,d,:,t]``."""
    lam = np.ascontiguousarray(lam, float)
    P, D, B, T = lam.shape
    K = len(E) - 1
    vis = _vis(P, pairs)
    lib = orc._load()
    dist = np.ascontiguousarray(dist, float)
    fft = np.ascontiguousarray(fft, float)
    p2o = np.ascontiguousarray(p2o, np.int64)
    mu = lam.copy()                                        # mu_K
    dfft = np.zeros_like(fft)
    for k in range(K - 1, -1, -1):
        new = np.empty_like(lam)
        lib.orc_exchange_adjoint_step(T, P, D, B, lam, mu,
                                      np.ascontiguousarray(E[k]), dist, fft,
                                      p2o, c, dt, vis, new, dfft,
                                      orc.default_threads())
        mu = new
    return dfft, mu


def table_gradient(r, dfft, mu0, tables_shape):
    """dL/dtable[NK, D, B] from dL/dF~ and dL/dE_0 for OracleRadiosity
    ``r`` (tables as stored: brdf * pi, rows k = bi * S + s)."""
    P, D, B = r.n_patches, r.brdf_out.shape[1], r.n_bins
    S = r.brdf_in.shape[1]
    centers = r.patches_center
    w = orc.fft(r.visibility_matrix, r.form_factors, B, centers,
                r.patches_area, r.air, r.patch_to_wall_ids, None, None,
                None, None)[:, :, 0, :]                    # [P,P,B]
    gt = np.zeros(tables_shape)
    vp = np.asarray(r.visible_patches, np.int64).reshape(-1, 2)
    both = np.concatenate([vp, vp[:, ::-1]])               # i -> j, j -> i
    ii, jj = both[:, 0], both[:, 1]
    s_idx = np.zeros(len(both), np.int64)
    wall = np.ascontiguousarray(r.patch_to_wall_ids, np.int64)
    orc._load().orc_source_index_pairs(
        len(both), np.ascontiguousarray(both), np.ascontiguousarray(centers),
        S, np.ascontiguousarray(r.brdf_in, float), wall, s_idx)
    k = np.asarray(r.brdf_index)[wall[ii]] * S + s_idx
    np.add.at(gt, k, dfft[ii, jj] * w[ii, jj][:, None, :])
    d0 = r.distance_patches_to_source
    for i in range(P):
        wi = r.patch_to_wall_ids[i]
        k = r.brdf_index[wi] * S + orc.source_index(r.source, centers[i],
                                                    r.brdf_in, wi)
        n = orc.delay_floor(d0[i], r.speed_of_sound, r.etc_time_resolution)
        gt[k] += r.energy_0[i][None, :] * mu0[i, :, :, n]
    return gt


def gradient(r, receivers, grad_hist, K):
    """dL/dtable [NBRDF*S, D, B] for an OracleRadiosity ``r`` after
    ``calculate_energy_exchange`` (which sets speed of sound, resolution)."""
    c, dt = r.speed_of_sound, r.etc_time_resolution
    T = grad_hist.shape[-1]
    dist = r.pair_distances()
    E = exchange_orders(T, r.energy_init_source, r.distance_patches_to_source,
                        dist, r.fft, r.p2o, c, dt, K, r.visible_patches)
    lam = collect_adjoint(grad_hist, receivers, r)
    dfft, mu0 = exchange_adjoint(lam, E, dist, r.fft, r.p2o, c, dt,
                                 r.visible_patches)
    tables = np.array(r.brdf)
    shape = (tables.shape[0] * tables.shape[1],) + tables.shape[2:]
    return table_gradient(r, dfft, mu0, shape)


def scattering_gradient(gtable, alpha, scattering, spec_value, spec_dir):
    """Chain dL/dtable (stored brdf * pi, rows n*S + s) to dL/dalpha and
    dL/ds of ``create_from_scattering`` BRDFs (one per n):
    table[n*S+s, d, b] = (1-alpha)(s_nb + (1-s_nb) pi spec_value[s]
    [d = spec_dir[s]])."""
    N, B = alpha.shape
    S = len(spec_value)
    g = gtable.reshape(N, S, -1, B)
    gsum = g.sum(axis=2)                                   # [N,S,B]
    gspec = g[:, np.arange(S), spec_dir, :]                # [N,S,B]
    sv = np.pi * np.asarray(spec_value)[None, :, None]
    a, s = alpha[:, None, :], scattering[:, None, :]
    galpha = -(gsum * s + gspec * (1 - s) * sv).sum(axis=1)
    gs = ((gsum - gspec * sv) * (1 - a)).sum(axis=1)
    return galpha, gs
