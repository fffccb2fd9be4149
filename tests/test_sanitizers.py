"""The C oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

``make -C oracle sanitize`` builds the oracle's sources with
``-fsanitize=address,undefined -fno-sanitize-recover=all``; a child Python
process with libasan preloaded loads that build (``ORACLE_LIB``) and runs
every C entry point on small scenes: visibility (full and by pairs), form
factors (Stokes and Nusselt pairs), injection, F~ (dense and by pairs),
the exchange (serial and threaded), one order on compact rows, the dense
order and its adjoint through the FP64 reference gradient, and the
receiver collection.  Any out-of-bounds access, use after free or
undefined behaviour aborts the child.  Its outputs must equal the
optimised build's bit for bit (the same IEEE operations in the same
order: -ffp-contract=off, no fast-math).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SANLIB = os.path.join(ROOT, "oracle", "liboracle_san.so")

CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
import oracle as orc
from oracle import adjoint as oadj
from conftest import make_oracle
from sparrowpy_amd import brdf as B, samplings, scenes, testing
from sparrowpy_amd.coordinates import FrequencyData

out = {}
sc = scenes.street_canyon(seed=5, length=8., width=5., height=4.,
                          n_recesses=2)
r = make_oracle(sc.walls, 1.)
ctr, nrm, pts = r.patches_center, r.patches_normal, r.patches_points
vis = orc.p2p_visibility(ctr, nrm, pts, nthreads=4)
ii, jj = np.triu_indices(len(ctr), 1)
out["vis"] = vis
out["vis_pairs"] = orc.p2p_visibility_pairs(
    ctr, nrm, pts, np.stack([ii, jj], -1)[::7], nthreads=4)

rng = np.random.default_rng(0)
room = make_oracle(testing.shoebox_room_stub(2, 1.5, 1.), .5)
c = samplings.hemisphere(samplings.sph_gaussian(sh_order=2))
freqs = np.array([500., 1000.])
for w in range(room.n_walls):
    br = B.create_from_scattering(
        c, c, FrequencyData(rng.uniform(.1, .9, (1, 2)), freqs),
        FrequencyData(rng.uniform(.05, .3, (1, 2)), freqs))
    room.set_wall_brdf([w], br.freq, c.cartesian, c.cartesian, freqs)
room.set_air_attenuation(np.array([1e-3, 3e-3]), freqs)
room.bake_geometry(nthreads=4)
out["ff"], out["fft"] = room.form_factors, room.fft
src = np.array([.7, .5, .4])
room.init_source_energy(src)
out["e0dir"] = room.energy_init_source
T, cs, dt = 120, 343., 1 / 2000
for nth in (1, 4):
    room.calculate_energy_exchange(cs, dt, T * dt, 3, nthreads=nth)
    out[f"etc{nth}"] = room.etc
rcv = np.array([[1.2, .9, .6], [.4, .3, .5]])
out["hist"] = room.collect_energy_receiver_mono(rcv)
sym = room.visibility_matrix | room.visibility_matrix.T
i2, j2 = np.nonzero(sym)
pairs = np.stack([i2, j2], -1)
ffu = room.form_factors[np.minimum(i2, j2), np.maximum(i2, j2)]
out["fft_pairs"] = orc.fft_pairs(pairs, ffu, 2, room.patches_center,
                                 room.patches_area, room.air,
                                 room.patch_to_wall_ids,
                                 np.array(room.brdf), room.brdf_index,
                                 room.brdf_in, room.brdf_out.shape[1])
e0 = orc.init_energy(T, room.energy_init_source,
                     room.distance_patches_to_source, cs, dt)
D = room.brdf_out.shape[1]
keys, row_src, lags, coefs, row_ptr = {}, [], [], [], [0]
dist = room.pair_distances()
for j in range(room.n_patches):
    for i in np.nonzero(sym[:, j])[0]:
        key = (int(i), int(room.p2o[i, j]))
        row_src.append(keys.setdefault(key, len(keys)))
        lags.append(orc.delay_floor(dist[i, j], cs, dt))
        coefs.append(room.fft[i, j])
    row_ptr.append(len(row_src))
prev_rows = np.stack([e0[i, d] for i, d in keys])
out["rows"] = orc.exchange_rows(T, np.array(row_ptr), np.array(row_src),
                                np.array(lags), np.array(coefs), prev_rows)
gH = rng.normal(size=out["hist"].shape)
out["grad"] = oadj.gradient(room, rcv, gH, 3)
out["lib"] = np.array(orc._LIB_PATH)
np.savez(sys.argv[2], **out)
"""

# A planted mistake: the visibility output buffer one byte short of P*P.
PLANTED = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import oracle as orc
ctr = np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.]])
pts = np.zeros((0, 4, 3))
out = np.zeros(3 * 3 - 1, np.uint8)
orc._load().orc_p2p_visibility(3, ctr, 0, np.zeros((0, 3)), pts, 4, out, 1)
"""


def _run(env_extra, outfile):
    env = dict(os.environ, **env_extra)
    return subprocess.run([sys.executable, "-c", CHILD, ROOT, outfile],
                          env=env, capture_output=True, text=True,
                          timeout=600)


@pytest.fixture(scope="module")
def sanlib():
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"],
                             capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(libasan) or not os.path.exists(libasan):
        pytest.skip("gcc has no AddressSanitizer runtime")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"),
                        "sanitize"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return libasan


def test_oracle_clean_under_asan_ubsan(sanlib, tmp_path):
    san = _run({"LD_PRELOAD": sanlib, "ORACLE_LIB": SANLIB,
                "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
                "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
                "ORACLE_THREADS": "4"}, str(tmp_path / "san.npz"))
    assert san.returncode == 0, san.stderr[-4000:]
    assert "AddressSanitizer" not in san.stderr
    assert "runtime error" not in san.stderr
    ref = _run({"ORACLE_THREADS": "4"}, str(tmp_path / "ref.npz"))
    assert ref.returncode == 0, ref.stderr[-4000:]
    with np.load(tmp_path / "san.npz") as a, \
            np.load(tmp_path / "ref.npz") as b:
        assert str(a["lib"]) == SANLIB
        assert str(b["lib"]).endswith("liboracle.so")
        assert set(a.files) == set(b.files)
        for k in a.files:
            if k != "lib":
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_sanitizer_catches_planted_overflow(sanlib):
    """The harness itself: an oracle call with an output buffer one byte
    short aborts the child with an AddressSanitizer report."""
    env = dict(os.environ, LD_PRELOAD=sanlib, ORACLE_LIB=SANLIB,
               ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, "-c", PLANTED, ROOT], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr
