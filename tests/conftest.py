"""Shared fixtures and helpers for the sparrowpy_amd test suite.

Tests marked ``gpu`` need an AMD GPU and the built HIP library; everything
else runs on the CPU (oracle pins, host-side logic, the C-ABI symbol table).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an AMD GPU and libsparrow_hip.so")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    import oracle
    oracle.build()


def make_oracle(walls, patch_size):
    """OracleRadiosity for a list of Polygons (host subdivision)."""
    from oracle.pipeline import OracleRadiosity
    from sparrowpy_amd import geometry
    wp = np.array([w.pts for w in walls])
    wn = np.array([w.normal for w in walls])
    wu = np.array([w.up_vector for w in walls])
    pts, _, _, ids = geometry.process_patches(wp, wn, patch_size)
    return OracleRadiosity(wp, wn, wu, pts, ids,
                           geometry.rotate_coords_to_normal)


def rel_err_within_db(test, ref, db=60.0, axis=-1):
    """Max relative error over bins within ``db`` of each histogram's peak."""
    test = np.asarray(test, dtype=float)
    ref = np.asarray(ref, dtype=float)
    peak = np.max(np.abs(ref), axis=axis, keepdims=True)
    mask = (np.abs(ref) >= peak * 10 ** (-db / 10)) & (peak > 0)
    if not mask.any():
        return 0.0
    return float(np.max(np.abs(test[mask] - ref[mask]) / np.abs(ref[mask])))
