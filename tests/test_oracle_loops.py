"""Oracle pins: Philox4x32-10 and the seeded unit-loop generator."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# Random123 known-answer vectors for philox4x32_10 (ctr, key -> out).
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_kat(orc, ctr, key, out):
    assert orc.philox(ctr, key) == out


def numpy_loops(config_id, Nl, N, orc):
    """Independent restatement of the generator in numpy (same Philox words)."""
    sigma = np.sqrt(2.0 / N)
    xs, ys = [], []
    for l in range(Nl):
        r = np.array([orc.philox([j, l, 0, 0], [0x5EED0000 + config_id, 0]) for j in range(N // 2)], np.float64)
        u = (r + 0.5) / 2.0**32
        ra = np.sqrt(-2 * np.log(u[:, 0])); aa = 2 * np.pi * u[:, 1]
        rb = np.sqrt(-2 * np.log(u[:, 2])); ab = 2 * np.pi * u[:, 3]
        dx = np.empty(N); dy = np.empty(N)
        dx[0::2] = sigma * ra * np.cos(aa); dy[0::2] = sigma * ra * np.sin(aa)
        dx[1::2] = sigma * rb * np.cos(ab); dy[1::2] = sigma * rb * np.sin(ab)
        px = np.cumsum(dx - dx.mean()); py = np.cumsum(dy - dy.mean())
        xs.append(px - px.mean()); ys.append(py - py.mean())
    return np.array(xs), np.array(ys)


def test_generator_matches_numpy_restatement(orc):
    x, y = orc.gen_loops(7, 6, 64)
    nx, ny = numpy_loops(7, 6, 64, orc)
    # numpy's transcendentals may differ from libm in the last ulp: agree to FP32 rounding
    assert np.max(np.abs(x - nx)) < 1e-6 and np.max(np.abs(y - ny)) < 1e-6


def test_generator_closure_and_moments(orc):
    Nl, N = 400, 128
    x, y = orc.gen_loops(11, Nl, N)
    # centre of mass at the origin (to FP32 rounding of each point)
    assert np.max(np.abs(x.mean(1))) < 1e-6 and np.max(np.abs(y.mean(1))) < 1e-6
    # increments (including the closing one) have variance 2/N per component
    dx = np.diff(np.concatenate([x, x[:, :1]], 1), axis=1)
    var = dx.var()
    assert abs(var / (2.0 / N) - 1.0) < 0.03
    # per-component variance of a Brownian bridge about its centre: 1/6 (continuum), (N^2-1)/(6 N^2) discrete
    for c in (x, y):
        m = (c.astype(np.float64) ** 2).mean(1)
        assert abs(m.mean() - 1.0 / 6.0) < 4 * m.std() / np.sqrt(Nl)


def test_levy_area_law(orc):
    """<cos(eBT A)> = eBT/sinh(eBT) for Brownian loops in a constant field:
    the step profile with every point inside the tube is a constant field, so
    the oracle's scalar phase is exactly eBT times the enclosed area."""
    Nl, N = 3000, 128
    x, y = orc.gen_loops(5, Nl, N)
    F, l2 = 50.0, 1.0            # eB = 2F/l2 = 100, T = 0.01: rho^2 = T |w|^2 < l2 for |w| < 10
    eB = 2 * F / l2
    T = 1.0 / eB                  # eBT = 1
    g = orc.Geom(orc.STEP, F=F, l2=l2, quantize=False)
    gs, pl = orc.eval_batch([g], orc.MODE_PHYSICAL, 2, 0, x, y, Nl, [0], [0.0], [0.0], [T])
    phase = pl[0, :, 0]           # N/2 sub-loop: every other point, a Brownian loop of N/2 points
    w = np.cos(phase)
    exact = 1.0 / np.sinh(1.0)
    assert abs(w.mean() - exact) < 4 * w.std() / np.sqrt(Nl) + 2e-3


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_generator_hash_fixture(orc, wfl):
    with open(os.path.join(GOLDEN, "loop_hashes.json")) as fh:
        fixtures = json.load(fh)
    for f in fixtures:
        x, y = orc.gen_loops(f["config_id"], f["Nl"], f["N"])
        assert _sha(x) == f["sha256_x"] and _sha(y) == f["sha256_y"], f
        px, py = wfl.gen_loops(f["config_id"], f["Nl"], f["N"])
        assert np.array_equal(px, x) and np.array_equal(py, y)
