"""Seeded synthetic scenes for the benchmark and parity workloads.

Each builder returns a :class:`Scene` holding walls, per-wall BRDF
assignments, air attenuation, sources, receivers and the histogram settings.
The wall layout and subdivision follow the reference conventions
(``sparrowpy/testing/stub_utils.py:5-48``, ``geometry.py:373-410``: every
wall is an axis-aligned rectangle, patches are a regular grid on it).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .brdf import create_from_directional_scattering, create_from_scattering
from .coordinates import Coordinates, FrequencyData
from .geometry import Polygon
from .samplings import hemisphere, sph_gaussian
from .testing import shoebox_room_stub

__all__ = ["Scene", "small_shoebox", "directional_shoebox", "street_canyon",
           "batched_shoeboxes", "OCTAVE_BANDS", "air_attenuation_octaves"]

OCTAVE_BANDS = np.array([63., 125., 250., 500., 1000., 2000., 4000., 8000.])

# Air attenuation in Np/m at 20 C, 50 % relative humidity (ISO 9613-1
# values rounded); only its magnitude matters for the workloads.
_AIR_NP_PER_M = {63.: 1.3e-5, 125.: 4.9e-5, 250.: 1.4e-4, 500.: 3.1e-4,
                 1000.: 5.7e-4, 2000.: 1.1e-3, 4000.: 3.3e-3, 8000.: 1.2e-2}


def air_attenuation_octaves(freqs):
    return np.array([_AIR_NP_PER_M[float(f)] for f in freqs])


@dataclasses.dataclass
class Scene:
    """A radiosity workload: geometry, materials, sources and receivers."""

    name: str
    walls: list
    patch_size: float
    frequencies: np.ndarray
    brdf_groups: list          # [(wall_indexes, FrequencyData brdf)]
    directions: Coordinates    # BRDF incoming == outgoing direction set
    air_attenuation: np.ndarray
    sources: np.ndarray        # (n_sources, 3)
    receivers: np.ndarray      # (n_receivers, 3)
    speed_of_sound: float
    time_resolution: float
    duration: float
    max_order: int
    # per-BRDF (absorption, scattering) [n_brdf, B] where the scene draws
    # them (None otherwise)
    materials: tuple = None

    @property
    def n_samples(self):
        return int(self.duration / self.time_resolution)


def _brdf(directions, s, alpha, freqs):
    return create_from_scattering(
        directions, directions,
        FrequencyData(np.asarray(s, float).reshape(1, -1), freqs),
        FrequencyData(np.asarray(alpha, float).reshape(1, -1), freqs))


def small_shoebox():
    """5x6x4 m shoebox, 1 m patches (148), diffuse BRDF on 16 directions,
    one band, alpha=0.1, order 20, 1 kHz bins over 1 s."""
    dirs = hemisphere(sph_gaussian(sh_order=3))
    freqs = np.array([1000.])
    walls = shoebox_room_stub(5, 6, 4)
    brdf = _brdf(dirs, [1.0], [0.1], freqs)
    return Scene(
        name="shoebox-5x6x4-p1.0", walls=walls, patch_size=1.0,
        frequencies=freqs, brdf_groups=[(np.arange(len(walls)), brdf)],
        directions=dirs, air_attenuation=np.zeros(1),
        sources=np.array([[2., 3., 1.5]]),
        receivers=np.array([[4., 2., 1.2]]),
        speed_of_sound=343., time_resolution=1 / 1000, duration=1.0,
        max_order=20)


def _split_walls_into_patches(walls, patch_size):
    """One wall per patch, so every patch can carry its own BRDF."""
    from .geometry import create_patches
    out = []
    for w in walls:
        for p in create_patches(w.pts, patch_size):
            out.append(Polygon(p, w.up_vector, w.normal))
    return out


def directional_shoebox(seed=2, n_receivers=4, size=(10., 8., 3.),
                        patch_size=0.5, sh_order=7, n_bands=6,
                        duration=1.5, rate=4000., max_order=30,
                        brdf="scattering"):
    """10x8x3 m shoebox at 0.5 m patches (1,072), one wall per patch with
    its own seeded BRDF (alpha in [0.02, 0.4], s in [0.1, 0.9] per band),
    64 directions, 6 octave bands, 4 receivers >= 1 m from walls, 4 kHz
    bins over 1.5 s, order 30.

    ``brdf="directional"``: the six walls unsplit (the same 1,072 patches
    once subdivided), each with one BRDF from a seeded directional
    scattering coefficient (``create_from_directional_scattering``, rows
    uniform in [0.2, 1] normalised over the outgoing directions, alpha in
    [0.02, 0.4]): rows of no diffuse + specular form, shared per wall;
    ``materials`` then holds the walls' alpha and directional scattering."""
    if brdf not in ("scattering", "directional"):
        raise ValueError('brdf must be "scattering" or "directional"')
    rng = np.random.default_rng(seed)
    X, Y, Z = size
    dirs = hemisphere(sph_gaussian(sh_order=sh_order))
    freqs = OCTAVE_BANDS[1:1 + n_bands] if n_bands <= 7 else \
        OCTAVE_BANDS[:n_bands]
    groups, alphas, scats = [], [], []
    split = _split_walls_into_patches(shoebox_room_stub(X, Y, Z), patch_size)
    for i in range(len(split)):
        alpha = rng.uniform(0.02, 0.4, n_bands)
        s = rng.uniform(0.1, 0.9, n_bands)
        if brdf == "scattering":
            groups.append(([i], _brdf(dirs, s, alpha, freqs)))
            alphas.append(alpha)
            scats.append(s)
    walls = split
    if brdf == "directional":
        # own generator: the draws above keep sources and receivers those
        # of the scattering form
        walls = shoebox_room_stub(X, Y, Z)
        drng = np.random.default_rng(seed + 1000)
        n = dirs.csize
        for i in range(len(walls)):
            alpha = drng.uniform(0.02, 0.4, n_bands)
            sd = drng.uniform(0.2, 1.0, (n, n, n_bands))
            sd /= sd.sum(axis=1, keepdims=True)
            groups.append(([i], create_from_directional_scattering(
                dirs, dirs, FrequencyData(sd, freqs),
                FrequencyData(alpha.reshape(1, -1), freqs))))
            alphas.append(alpha)
            scats.append(sd)
    recv = np.stack([rng.uniform(1.0, X - 1.0, n_receivers),
                     rng.uniform(1.0, Y - 1.0, n_receivers),
                     rng.uniform(1.0, Z - 1.0, n_receivers)], axis=-1)
    src = np.array([[X * 0.3, Y * 0.4, Z * 0.5]])
    return Scene(
        name=f"shoebox-{X:g}x{Y:g}x{Z:g}-p{patch_size:g}-dir"
             + ("-dirscat" if brdf == "directional" else ""),
        walls=walls, patch_size=patch_size, frequencies=freqs,
        brdf_groups=groups, directions=dirs,
        air_attenuation=air_attenuation_octaves(freqs), sources=src,
        receivers=recv, speed_of_sound=343., time_resolution=1 / rate,
        duration=duration, max_order=max_order,
        materials=(np.array(alphas), np.array(scats)))


def _facade_rects(x_len, z_len, holes):
    """Axis-aligned rectangles covering [0,x_len]x[0,z_len] minus holes.

    ``holes`` are (x0, x1, z0, z1); the result is a list of the same form.
    """
    xs = sorted({0.0, float(x_len)} | {h[0] for h in holes}
                | {h[1] for h in holes})
    rects = []
    for x0, x1 in zip(xs[:-1], xs[1:]):
        inside = sorted((h[2], h[3]) for h in holes
                        if h[0] <= x0 and h[1] >= x1)
        z = 0.0
        for z0, z1 in inside:
            if z0 > z:
                rects.append((x0, x1, z, z0))
            z = z1
        if z < z_len:
            rects.append((x0, x1, z, float(z_len)))
    return rects


def street_canyon(seed=3, length=100., width=20., height=15.,
                  n_recesses=10, recess_depth=1.0, patch_size=1.0,
                  n_dirs=(4, 8), n_bands=8, duration=1.5, rate=4000.,
                  max_order=40, n_receivers=16, brdf="scattering"):
    """Open street canyon: 100x20 m ground and two 100x15 m facades at 1 m
    patches, facades broken by 10 seeded rectangular recesses 2-4 m wide and
    high and ``recess_depth`` deep (each adds a back wall and four reveal
    walls and produces occlusion).  32 BRDF directions (4 colatitudes x 8
    azimuths), 8 octave bands, per-wall alpha in [0.05, 0.3] and s in
    [0.1, 0.9], 2 sources at 1.5 m height, receivers along the canyon axis,
    4 kHz bins over 1.5 s, order 40.

    ``brdf="directional"``: every wall's BRDF from a seeded directional
    scattering coefficient (``create_from_directional_scattering``, each
    incoming direction's row uniform in [0.2, 1] and normalised to sum to
    1 over the outgoing directions, per band) with the wall's absorption:
    rows that are neither a constant plus one specular entry nor shared
    between walls."""
    if brdf not in ("scattering", "directional"):
        raise ValueError('brdf must be "scattering" or "directional"')
    rng = np.random.default_rng(seed)
    L, Wd, H, dep = length, width, height, recess_depth
    holes = {0: [], 1: []}
    tries = 0
    w_max, h_max = int(min(4, L - 2)), int(min(4, H - 2))
    while (sum(len(v) for v in holes.values()) < n_recesses
           and tries < 10000 and w_max >= 2 and h_max >= 2):
        tries += 1
        side = int(rng.integers(0, 2))
        w = float(rng.integers(2, w_max + 1))
        h = float(rng.integers(2, h_max + 1))
        if L - w <= 1 or H - h <= 1:
            continue
        x0 = float(rng.integers(1, int(L - w)))
        z0 = float(rng.integers(1, int(H - h)))
        cand = (x0, x0 + w, z0, z0 + h)
        if all(cand[1] + 1 <= o[0] or o[1] + 1 <= cand[0]
               or cand[3] + 1 <= o[2] or o[3] + 1 <= cand[2]
               for o in holes[side]):
            holes[side].append(cand)
    walls = [Polygon([[0, 0, 0], [L, 0, 0], [L, Wd, 0], [0, Wd, 0]],
                     [1, 0, 0], [0, 0, 1])]
    for side, y, ny in ((0, 0.0, 1.0), (1, Wd, -1.0)):
        yb = y - ny * dep  # recess back wall plane
        for (x0, x1, z0, z1) in _facade_rects(L, H, holes[side]):
            walls.append(Polygon([[x0, y, z0], [x1, y, z0], [x1, y, z1],
                                  [x0, y, z1]], [1, 0, 0], [0, ny, 0]))
        for (x0, x1, z0, z1) in holes[side]:
            ylo, yhi = min(y, yb), max(y, yb)
            walls += [
                Polygon([[x0, yb, z0], [x1, yb, z0], [x1, yb, z1],
                         [x0, yb, z1]], [1, 0, 0], [0, ny, 0]),
                Polygon([[x0, ylo, z0], [x0, yhi, z0], [x0, yhi, z1],
                         [x0, ylo, z1]], [0, 0, 1], [1, 0, 0]),
                Polygon([[x1, ylo, z0], [x1, yhi, z0], [x1, yhi, z1],
                         [x1, ylo, z1]], [0, 0, 1], [-1, 0, 0]),
                Polygon([[x0, ylo, z0], [x1, ylo, z0], [x1, yhi, z0],
                         [x0, yhi, z0]], [1, 0, 0], [0, 0, 1]),
                Polygon([[x0, ylo, z1], [x1, ylo, z1], [x1, yhi, z1],
                         [x0, yhi, z1]], [1, 0, 0], [0, 0, -1]),
            ]
    dirs = hemisphere(sph_gaussian(n_points=(2 * n_dirs[0], n_dirs[1])))
    freqs = OCTAVE_BANDS[:n_bands]
    groups = []
    drng = np.random.default_rng(seed + 1000)
    for i in range(len(walls)):
        alpha = rng.uniform(0.05, 0.3, n_bands)
        s = rng.uniform(0.1, 0.9, n_bands)
        if brdf == "scattering":
            groups.append(([i], _brdf(dirs, s, alpha, freqs)))
            continue
        n = dirs.csize
        sd = drng.uniform(0.2, 1.0, (n, n, n_bands))
        sd /= sd.sum(axis=1, keepdims=True)
        groups.append(([i], create_from_directional_scattering(
            dirs, dirs, FrequencyData(sd, freqs),
            FrequencyData(alpha.reshape(1, -1), freqs))))
    src = np.array([[L * 0.25, Wd * 0.35, 1.5], [L * 0.7, Wd * 0.6, 1.5]])
    rx = np.linspace(L * 0.05, L * 0.95, n_receivers)
    recv = np.stack([rx, np.full_like(rx, Wd / 2), np.full_like(rx, 1.5)],
                    axis=-1)
    return Scene(
        name="street-canyon-100x20x15-p1.0" + (
            "-dirscat" if brdf == "directional" else ""), walls=walls,
        patch_size=patch_size, frequencies=freqs, brdf_groups=groups,
        directions=dirs, air_attenuation=air_attenuation_octaves(freqs),
        sources=src, receivers=recv, speed_of_sound=343.,
        time_resolution=1 / rate, duration=duration, max_order=max_order)


def batched_shoeboxes(n_scenes=256, seed=4, rate=2000., duration=1.0,
                      max_order=20):
    """Many small rooms for batched runs: shoeboxes with dimensions uniform
    in [3,20] x [3,20] x [2.5,6] m and patch size uniform in [0.5,1.5] m
    (28 to ~3,000 patches), 1-4 sources and 1-8 receivers at least 0.5 m
    from the walls, 16 or 64 BRDF directions, 1-8 octave bands, one seeded
    material per room (alpha in [0.05,0.3], s in [0.1,0.9] per band),
    2 kHz bins over 1 s, order 20."""
    rng = np.random.default_rng(seed)
    dirs = {16: hemisphere(sph_gaussian(sh_order=3)),
            64: hemisphere(sph_gaussian(sh_order=7))}
    out = []
    for x in range(n_scenes):
        X, Y = rng.uniform(3, 20, 2)
        Z = rng.uniform(2.5, 6)
        ps = float(rng.uniform(0.5, 1.5))
        n_src, n_rcv = int(rng.integers(1, 5)), int(rng.integers(1, 9))
        D = int(rng.choice([16, 64]))
        n_bands = int(rng.integers(1, 9))
        freqs = OCTAVE_BANDS[:n_bands]

        def inside(n):
            return np.stack([rng.uniform(.5, X - .5, n),
                             rng.uniform(.5, Y - .5, n),
                             rng.uniform(.5, Z - .5, n)], axis=-1)
        walls = shoebox_room_stub(X, Y, Z)
        brdf = _brdf(dirs[D], rng.uniform(0.1, 0.9, n_bands),
                     rng.uniform(0.05, 0.3, n_bands), freqs)
        out.append(Scene(
            name=f"room{x}-{X:.1f}x{Y:.1f}x{Z:.1f}-p{ps:.2f}", walls=walls,
            patch_size=ps, frequencies=freqs,
            brdf_groups=[(np.arange(len(walls)), brdf)], directions=dirs[D],
            air_attenuation=air_attenuation_octaves(freqs),
            sources=inside(n_src), receivers=inside(n_rcv),
            speed_of_sound=343., time_resolution=1 / rate, duration=duration,
            max_order=max_order))
    return out
