"""The C-ABI library: it loads, exports every function include/*.h declares,
and its host-side argument checks and error paths work without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("wfl.h", "intT.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?[\s\*](\w+)\s*\([^;{]*\)\s*;", text, flags=re.M):
            if not m.group(0).lstrip().startswith("typedef"):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(wfl):
    names = declared_functions()
    assert len(names) > 30
    lib = C.CDLL(wfl.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names <= set(wfl.EXPORTS) | names   # binding covers them
    assert set(wfl.EXPORTS) >= names, sorted(names - set(wfl.EXPORTS))


def test_argument_errors_are_reported(wfl):
    b = wfl.Batch([wfl.Geometry()], np.array([0]), np.array([0.0]), np.array([0.0]), np.array([1.0]), nsub=3)
    h = C.c_void_p()
    cb, keep = b.c()
    rc = wfl.lib.wfl_plan_create(C.byref(h), C.byref(cb), C.c_void_p(1))
    assert rc == -1 and b"nsub" in wfl.lib.wfl_last_error()
    b2 = wfl.Batch([wfl.Geometry(kind=42)], np.array([0]), np.array([0.0]), np.array([0.0]), np.array([1.0]))
    cb, keep = b2.c()
    assert wfl.lib.wfl_plan_create(C.byref(h), C.byref(cb), C.c_void_p(1)) == -1
    b3 = wfl.Batch([wfl.Geometry()], np.array([0]), np.array([0.0]), np.array([0.0]), np.array([-1.0]))
    cb, keep = b3.c()
    assert wfl.lib.wfl_plan_create(C.byref(h), C.byref(cb), C.c_void_p(1)) == -1
    assert b"T > 0" in wfl.lib.wfl_last_error()
    b4 = wfl.Batch([wfl.Geometry(kind=wfl.SPLINE, spline_table=0)], np.array([0]), np.array([0.0]),
                   np.array([0.0]), np.array([1.0]))
    cb, keep = b4.c()
    assert wfl.lib.wfl_plan_create(C.byref(h), C.byref(cb), C.c_void_p(1)) == -1
    # an unused table still goes into the plan: it is checked too
    bad = (np.array([0.0, 1.0, 0.5], np.float32), np.zeros((3, 4), np.float32))
    b5 = wfl.Batch([wfl.Geometry()], np.array([0]), np.array([0.0]), np.array([0.0]), np.array([1.0]), splines=[bad])
    cb, keep = b5.c()
    assert wfl.lib.wfl_plan_create(C.byref(h), C.byref(cb), C.c_void_p(1)) == -1
    assert b"knots not increasing" in wfl.lib.wfl_last_error()
    # loop shapes: N must be a multiple of 4
    x = np.zeros((2, 6), np.float32)
    assert wfl.lib.wfl_loops_create(C.byref(h), x.ctypes.data, x.ctypes.data, 2, 6) == -1
    assert wfl.lib.wfl_gen_loops(1, 2, 6, x.ctypes.data, x.ctypes.data) == -1
    assert wfl.lib.wfl_set_numerics(0, 0, 3) == -1
    n, e = C.c_int(), C.c_double()
    cb, keep = b4.c()
    assert wfl.lib.wfl_autotune_nsub(C.byref(cb), C.c_void_p(1), -1.0, 64, 16, 16, C.byref(n), C.byref(e)) == -1
    assert wfl.lib.wfl_autotune_nsub(C.byref(cb), None, 1e-6, 64, 16, 16, C.byref(n), C.byref(e)) == -1
    with pytest.raises(wfl.WflError):
        wfl.spline_build([0.0, 1.0, 0.5], [0.0, 1.0, 1.0])


@pytest.mark.skipif(os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"),
                    reason="a GPU is present")
def test_no_device_path_fails_loudly(wfl):
    """Without a device every entry that needs one returns WFL_ERR_DEVICE
    with a message (there is no CPU fallback)."""
    if wfl.device_count() > 0:
        pytest.skip("device present")
    x = np.zeros((2, 8), np.float32)
    h = C.c_void_p()
    rc = wfl.lib.wfl_loops_create(C.byref(h), x.ctypes.data, x.ctypes.data, 2, 8)
    assert rc == -2 and wfl.lib.wfl_last_error()
    with pytest.raises(wfl.WflError):
        wfl.Loops(x, x)


def test_gen_loops_and_text_reader_roundtrip(wfl, tmp_path):
    """getwl.c text layout: per loop N x values then N y values."""
    x, y = wfl.gen_loops(4, 3, 8)
    p = tmp_path / "wl.dat"
    with open(p, "w") as fh:
        for l in range(3):
            fh.write(" ".join(f"{v:.9g}" for v in x[l]) + "\n")
            fh.write(" ".join(f"{v:.9g}" for v in y[l]) + "\n")
    rx = np.zeros((3, 8), np.float32)
    ry = np.zeros((3, 8), np.float32)
    wfl.check(wfl.lib.wfl_read_text_loops(str(p).encode(), 3, 8, rx.ctypes.data, ry.ctypes.data))
    assert np.array_equal(rx, x) and np.array_equal(ry, y)
    aos = np.zeros((3 * 8, 4), np.float32)
    wfl.lib.getwl(aos.ctypes.data, str(p).encode(), 3, 8, 2)
    assert np.array_equal(aos[:, 0].reshape(3, 8), x) and np.array_equal(aos[:, 1].reshape(3, 8), y)
    assert wfl.lib.wfl_read_text_loops(str(p).encode(), 4, 8, rx.ctypes.data, ry.ctypes.data) == -1


def test_binary_loop_file_roundtrip(wfl, tmp_path):
    """Binary loop file: header fields, exact round trip, a shape mismatch
    and a corrupted byte are refused."""
    Nl, N = 5, 16
    x, y = wfl.gen_loops(7, Nl, N)
    p = tmp_path / "wl.bin"
    wfl.check(wfl.lib.wfl_loops_write(str(p).encode(), x.ctypes.data, y.ctypes.data, Nl, N, 7))
    assert os.path.getsize(p) == 40 + 2 * 4 * Nl * N
    hn, hN, hc = C.c_int(), C.c_int(), C.c_uint32()
    wfl.check(wfl.lib.wfl_loops_read_header(str(p).encode(), C.addressof(hn), C.addressof(hN), C.addressof(hc)))
    assert (hn.value, hN.value, hc.value) == (Nl, N, 7)
    rx = np.zeros((Nl, N), np.float32)
    ry = np.zeros((Nl, N), np.float32)
    wfl.check(wfl.lib.wfl_loops_read(str(p).encode(), rx.ctypes.data, ry.ctypes.data, Nl, N))
    assert np.array_equal(rx, x) and np.array_equal(ry, y)
    assert wfl.lib.wfl_loops_read(str(p).encode(), rx.ctypes.data, ry.ctypes.data, Nl, N + 1) == -1
    raw = bytearray(p.read_bytes())
    raw[40 + 4 * N + 1] ^= 0x10
    q = tmp_path / "bad.bin"
    q.write_bytes(bytes(raw))
    assert wfl.lib.wfl_loops_read(str(q).encode(), rx.ctypes.data, ry.ctypes.data, Nl, N) == -1
    assert b"hash" in wfl.lib.wfl_last_error()
    q.write_bytes(b"NOTALOOPFILE" + bytes(60))
    assert wfl.lib.wfl_loops_read_header(str(q).encode(), None, None, None) == -1


def test_effective_parameters_match_oracle(wfl, orc):
    """The engine's parameter narrowing (FP32 in the Gaussian fast path's
    scaled units, plain FP32 otherwise) restated independently in the
    oracle's binding gives the same doubles."""
    for kind, l2, F, lat in [(wfl.GAUSSIAN, 0.25, 10.0, 0.0), (wfl.GAUSSIAN, 1.7, 0.1, 1.3), (wfl.SMOOTH, 0.7, 2.0, 0.0)]:
        g = wfl.Geometry(kind=kind, F=F, l2=l2, lattice=lat, images=1)
        og = orc.Geom(kind, F=F, l2=l2, lattice=lat, images=1)
        for cx, cy, T in [(0.0, 0.0, 1e-3), (0.37, -1.1, 0.77), (3.3, 0.2, 96.0)]:
            e = wfl.effective_tuple(g, cx, cy, T)
            assert (e[0], e[1], e[5]) == (og.c.l2, og.c.F, og.c.lattice)
            assert tuple(e[2:5]) == og.tuple(cx, cy, T)
            assert wfl.effective_T(g, T)[0] == e[4] ** 2
            assert abs(e[4] ** 2 / T - 1) < 2e-7 and abs(e[0] / l2 - 1) < 2e-7


def test_group_ev(wfl):
    g = np.array([3.0, -1.5, 0.6])
    assert wfl.group_ev_minus1(g, 3) == pytest.approx((8 / 3 * 3.0 + 2 * 1.5 + 0.2) / 3)


def test_leff_groups_matches_oracle(wfl, orc):
    """Host L_eff (Tfunc policy + log-Simpson) against the oracle's on
    synthetic group sums."""
    rng = np.random.default_rng(1)
    T = np.logspace(-3, 2, 24)
    gs = -np.abs(rng.normal(size=(24, 4, 3))) * 1e-2
    counts = np.full(4, 128, np.int32)
    for fermion in (0, 1):
        for policy in (0, 1):
            for rho in (0.0, 0.8):
                a = wfl.leff_groups(wfl.Geometry(kind=wfl.GAUSSIAN, F=1.0, l2=0.25), fermion, T, rho, 0.0, gs,
                                    counts, policy)
                b = orc.leff_groups(orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.25), fermion, T, rho, 0.0, gs, counts,
                                    policy)
                assert np.allclose(a, b, rtol=1e-12, atol=1e-300)


def test_leff_groups_lattice_matches_oracle(wfl, orc):
    """Lattice geometries: the field in the counterterm and the branch policy
    is the image-window sum, in the host L_eff and the oracle alike."""
    rng = np.random.default_rng(2)
    T = np.logspace(-3, 2, 20)
    gs = -np.abs(rng.normal(size=(20, 3, 3))) * 1e-2
    counts = np.full(3, 128, np.int32)
    for lat, M in ((1.5, 1), (0.4, 2)):
        for cx, cy in ((0.1, 0.2), (0.7, 1.3), (-2.2, 0.45)):
            for fermion in (0, 1):
                a = wfl.leff_groups(wfl.Geometry(kind=wfl.GAUSSIAN, F=1.0, l2=0.25, lattice=lat, images=M), fermion,
                                    T, cx, cy, gs, counts, 1)
                b = orc.leff_groups(orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.25, lattice=lat, images=M), fermion, T, cx,
                                    cy, gs, counts, 1)
                assert np.allclose(a, b, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("kind,F,l2", [("STEP", 1.0, 1.0), ("SMOOTH", 0.5, 2.0), ("QUADRATIC", 2.0, 0.5),
                                       ("GAUSSIAN", 1.0, 0.25)])
def test_class_action_is_field_energy(wfl, orc, kind, F, l2):
    """ClassAction (EffAct.c:231-263) is minus half the field energy,
    -1/2 int B^2 d^2x, with eB = 2 F f'(rho^2) (Integrand.c:137): pinned by
    quadrature of the oracle's field."""
    from scipy.integrate import quad
    wfl.set_profile(wfl.Geometry(kind=getattr(wfl, kind), F=F, l2=l2))
    g = orc.Geom(getattr(orc, kind), F=F, l2=l2, quantize=False)
    lam = np.sqrt(l2)
    pts = [lam] if kind in ("STEP", "QUADRATIC") else None
    val, _ = quad(lambda r: orc.lib.orc_magfield(g.ref, r * r) ** 2 * 2 * np.pi * r, 0.0, 40.0 * lam, points=pts,
                  limit=400)
    assert wfl.lib.ClassAction(F, l2) == pytest.approx(-0.5 * val, rel=1e-8)
    wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN))


def test_binary_loop_file_python_helpers(wfl, tmp_path):
    x, y = wfl.gen_loops(3, 4, 32)
    wfl.write_loops(tmp_path / "a.bin", x, y, config_id=3)
    rx, ry, cid = wfl.read_loops(tmp_path / "a.bin")
    assert cid == 3 and np.array_equal(rx, x) and np.array_equal(ry, y)


@pytest.mark.parametrize("fmt", ["text", "bin"])
def test_generator_cli(wfl, tmp_path, fmt):
    """wlgen.py writes the generator's loops in either format; the readers
    give back the same floats (the text layout at 9 significant digits)."""
    import subprocess
    import sys
    cli = os.path.join(ROOT, "worldline-flux-tube-lattice_amd", "wlgen.py")
    r = subprocess.run([sys.executable, cli, "--config-id", "5", "--Nl", "6", "--N", "64", "--format", fmt,
                        "--verify"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    name = r.stdout.strip()
    assert name == f"worldlines.6.64.2.{'dat' if fmt == 'text' else 'bin'}"
    x, y = wfl.gen_loops(5, 6, 64)
    if fmt == "text":
        rx, ry = wfl.read_text_loops(tmp_path / name, 6, 64)
    else:
        rx, ry, cid = wfl.read_loops(tmp_path / name)
        assert cid == 5
    assert np.array_equal(rx, x) and np.array_equal(ry, y)


def test_c_driver_links_and_agrees(wfl, tmp_path):
    """A C program written like the reference's drivers (intT.h, struct
    Wparams filled by field) compiles with gcc, links libwfl.so and gets the
    same values as the ctypes calls: the header's struct layout and the
    calling convention match the library's."""
    import subprocess
    import sys
    exe = tmp_path / "driver"
    pkg = os.path.dirname(wfl.LIB_PATH)
    cmd = ["gcc", "-std=c11", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
           "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "c", "driver.c"), "-L", pkg, "-lwfl",
           f"-Wl,-rpath,{pkg}", "-lm", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line)
    assert int(vals["sizeof_Wparams"]) == C.sizeof(wfl.Wparams)
    p = wfl.Wparams()
    p.Nl, p.Nppl, p.F, p.l2, p.ng = 256, 64, 1.5, 0.64, 2
    p.xcm.x = 0.3
    assert float(vals["MagField"]) == wfl.lib.MagField(0.25, C.byref(p))
    assert float(vals["Exact_scalar"]) == wfl.lib.Exact(0.7, 0.9, 0)
    assert float(vals["Exact_spinor"]) == wfl.lib.Exact(0.7, 0.9, 1)
    assert float(vals["smTscal"]) == wfl.lib.smTscal(0.9, 0.01)
    assert float(vals["ClassAction"]) == wfl.lib.ClassAction(1.5, 0.64)
    if wfl.device_count() == 0:
        assert vals["EV_without_device"] == "nan"


def test_evmean_matches_oracle(wfl, orc):
    """EVMean (Integrand.c:325-407), host code: the library's against the
    oracle's on random float4 phases and spin sums, scalar and spinor,
    including values past the reference's |.| > 600 cosh guard."""
    rng = np.random.default_rng(5)
    ws = rng.normal(scale=0.8, size=(300, 4)).astype(np.float32)
    wf = rng.normal(scale=0.5, size=(300, 4)).astype(np.float32)
    wf[7, 0] = 650.0            # past the guard
    WL = rng.permutation(300)[:128].astype(np.int32)
    for fermion in (0, 1):
        out = C.c_double(0.0)
        wfl.lib.EVMean(C.addressof(out), ws.ctypes.data, wf.ctypes.data, 128, WL.ctypes.data, 0.7, fermion)
        want = orc.evmean(ws, wf, WL, fermion)
        assert out.value == pytest.approx(want, rel=1e-13, abs=1e-300)
    WL[0] = 7                   # the guarded loop in the group
    out = C.c_double(0.0)
    wfl.lib.EVMean(C.addressof(out), ws.ctypes.data, wf.ctypes.data, 128, WL.ctypes.data, 0.7, 1)
    assert out.value == pytest.approx(orc.evmean(ws, wf, WL, 1), rel=1e-13)


def test_local_taylor_coefficients(wfl):
    """The local Taylor node's coefficients a_k = g^(k)(uc)/k! of
    g(u) = (1 - 2^-u)/u against high-precision derivatives (mpmath), and
    the series' truncation at k = 7 within 4e-9 of g(u) - g(uc) for
    |u - uc| <= 0.5 (the kernel's range) from uc = 0 to 9."""
    import mpmath as mp
    mp.mp.dps = 40
    L = mp.log(2)

    def g(u):
        u = mp.mpf(u)
        return (1 - mp.power(2, -u)) / u if u != 0 else L

    for uc in (0.0, 1e-3, 0.3, 1.0, 2.25, 4.0, 9.0):
        a = np.zeros(8)
        wfl.check(wfl.lib.wfl_local_taylor(uc, a.ctypes.data))
        u0 = mp.mpf(uc) if uc > 0 else mp.mpf("1e-30")
        for k in range(8):
            ref = mp.diff(g, u0, k) / mp.factorial(k)
            assert abs(float((a[k] - ref) / ref)) < 1e-11, (uc, k)
        worst = 0.0
        for d in np.linspace(-0.5, 0.5, 41):
            if d == 0.0 or uc + d <= 0.0:
                continue
            exact = g(uc + d) - g(uc)
            series = sum(mp.mpf(a[k]) * mp.mpf(d) ** k for k in range(1, 8))
            worst = max(worst, abs(float((series - exact) / exact)))
        assert worst < 4e-9, (uc, worst)
    assert wfl.lib.wfl_local_taylor(-1.0, np.zeros(8).ctypes.data) == -1


def test_interaction_energy_statistics(wfl, orc):
    """wfl_interaction_energy against interaction.sage's formula restated in
    the oracle (mean of E_cell - E_iso with E = -action, and
    sqrt(var + var - 2 cov)), and the formula's fixed points: equal inputs
    give 0 and 0, a constant shift of the cell actions gives minus that
    shift with zero spread."""
    rng = np.random.default_rng(5)
    iso = rng.normal(-0.3, 0.05, 40)
    cell = iso + rng.normal(0.01, 0.002, 40)
    got = wfl.interaction_energy(cell, iso)
    ref = orc.interaction_energy(cell, iso)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)
    # the difference's spread is the std of the per-group differences
    assert got[1] == pytest.approx(np.std(cell - iso, ddof=1), rel=1e-9)
    assert wfl.interaction_energy(iso, iso) == (0.0, 0.0, 0.0)
    m, e, s = wfl.interaction_energy(iso + 0.25, iso)
    assert m == pytest.approx(-0.25, rel=1e-12) and e < 1e-7 and s < 1e-7


def test_radial_and_cell_action_quadrature(wfl, orc):
    """The lattice output quadratures against the oracle restatement on
    oracle group sums (a Gaussian tube, 8 loops x 64 points, 2 groups), and
    the radial Simpson weights exact for cubic integrands (even and odd
    interval counts)."""
    x, y = orc.gen_loops(9, 8, 64)
    T = np.logspace(-2, 1, 7)
    counts = np.array([4, 4], np.int32)
    geom = wfl.Geometry(kind=wfl.GAUSSIAN, F=1.0, l2=1.0)
    og = orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0)
    for nrho in (5, 6):
        rho = np.linspace(0.0, 2.0, nrho)
        tg = np.zeros(nrho * len(T), np.int32)
        gs, _ = orc.eval_batch([og], orc.MODE_PHYSICAL, 2, 1, x, y, 4, tg, np.repeat(rho, len(T)), 0 * tg,
                               np.tile(T, nrho), perloop=False)
        gs = gs.reshape(nrho, len(T), 2, 3)
        got = wfl.radial_action_groups(geom, 1, T, rho, gs, counts)
        ref = orc.radial_action_groups(og, 1, T, rho, gs, counts)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-300)
    # lattice cell: 2 x 2 midpoint grid of a lattice geometry
    lat = wfl.Geometry(kind=wfl.GAUSSIAN, F=1.0, l2=1.0, lattice=2.0, images=1)
    olat = orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0, lattice=2.0, images=1)
    cx = np.array([0.5, 1.5, 0.5, 1.5]); cy = np.array([0.5, 0.5, 1.5, 1.5]); w = np.full(4, 1.0)
    tg = np.zeros(4 * len(T), np.int32)
    gs, _ = orc.eval_batch([olat], orc.MODE_PHYSICAL, 2, 0, x, y, 4, tg, np.repeat(cx, len(T)), np.repeat(cy, len(T)),
                           np.tile(T, 4), perloop=False)
    gs = gs.reshape(4, len(T), 2, 3)
    got = wfl.cell_action_groups(lat, 0, T, cx, cy, w, gs, counts)
    ref = orc.cell_action_groups(olat, 0, T, cx, cy, w, gs, counts)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-300)
    # Simpson (odd count) and Simpson + 3/8 (even count) are exact for cubics
    for n in (9, 10):
        h = 0.3
        f = np.array([(k * h) ** 3 - 2 * (k * h) + 1 for k in range(n)])
        b = (n - 1) * h
        assert orc.simpson_uniform(h, f) == pytest.approx(b ** 4 / 4 - b ** 2 + b, rel=1e-13)
    with pytest.raises(wfl.WflError):
        wfl.radial_action_groups(geom, 1, T, np.array([0.0, 0.5, 1.7]), np.zeros((3, len(T), 2, 3)), counts)


def _c_stdout(capfd, fn):
    """Run fn and return what the C library printed (stdio flushed)."""
    libc = C.CDLL(None)
    libc.fflush(None)
    capfd.readouterr()
    try:
        fn()
    finally:
        libc.fflush(None)
    return capfd.readouterr().out


def test_tfunc_verbosity_lines(wfl, orc, capfd):
    """Tfunc at verbosity 2 prints the reference's branch traces and its
    WLvconstExpression: line (Integrand.c:255-313): rho^2, T, Exact,
    smTscal, and the mean and meanigrand spread of the integrand over the
    groups (here every group takes the same analytic branch, so the spread
    is 0).  Small-T and large-T branches need no device."""
    wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN, F=1.0, l2=1.0))
    p = wfl.Wparams()
    p.Nl, p.Nppl, p.F, p.l2, p.ng, p.fermion = 256, 64, 1.0, 1.0, 2, 1
    p.xcm.x, p.xcm.y = 0.5, 0.0
    order = np.arange(256, dtype=np.int32)
    og = orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0, quantize="float")
    B = orc.lib.orc_magfield(og.ref, 0.25)
    wfl.lib.wfl_set_verbosity(2)
    try:
        for T, tag, ref_ig in ((0.005, "using small T expression:", orc.lib.orc_smTferm(B, 0.005)),
                               (60.0, "using large T expression", -np.exp(-60.0) / 180.0 * (0.30282212 * B) ** 2)):
            out = _c_stdout(capfd, lambda: wfl.lib.Tfunc(T, C.byref(p), order.ctypes.data))
            lines = out.strip().splitlines()
            assert sum(l.startswith(tag) for l in lines) == 2, out
            wl = [l for l in lines if l.startswith("WLvconstExpression:")]
            assert len(wl) == 1, out
            v = [float(t) for t in wl[0].split()[1:]]
            assert v[0] == pytest.approx(0.25, rel=1e-6) and v[1] == pytest.approx(T, rel=1e-6)
            assert v[2] == pytest.approx(orc.lib.orc_exact(T, B, 1), rel=1e-5, abs=1e-300)
            assert v[3] == pytest.approx(orc.lib.orc_smTscal(B, T), rel=1e-5, abs=1e-300)
            assert v[4] == pytest.approx(ref_ig, rel=1e-5) and v[5] == 0.0
        wfl.lib.wfl_set_verbosity(1)
        out = _c_stdout(capfd, lambda: wfl.lib.Tfunc(0.005, C.byref(p), order.ctypes.data))
        assert "WLvconstExpression" not in out and "using" not in out
    finally:
        wfl.lib.wfl_set_verbosity(1)


def test_fp_trap_modes(wfl, capfd):
    """EVMean's floating-point checks (Integrand.c:318-342): off by default
    (an invalid operation passes silently), mode 1 reports and counts, mode
    2 is the reference's trap and exit (status 1), run in a child process."""
    ws = np.zeros((4, 4), np.float32)
    wf = np.zeros((4, 4), np.float32)
    ws[1, 0] = np.inf                      # cos(inf): invalid
    WL = np.arange(4, dtype=np.int32)
    ev = C.c_double()
    run = lambda: wfl.lib.EVMean(C.byref(ev), ws.ctypes.data, wf.ctypes.data, 4, WL.ctypes.data, 1.0, 0)
    out = _c_stdout(capfd, run)
    assert "Floating point" not in out
    n0 = wfl.lib.wfl_fp_trap_count()
    wfl.check(wfl.lib.wfl_set_fp_traps(1))
    try:
        out = _c_stdout(capfd, run)
        assert "Error: Floating point signal detected." in out
        assert wfl.lib.wfl_fp_trap_count() == n0 + 1
        ws[1, 0] = 0.0
        out = _c_stdout(capfd, run)          # clean inputs raise nothing
        assert "Floating point" not in out and wfl.lib.wfl_fp_trap_count() == n0 + 1
    finally:
        wfl.check(wfl.lib.wfl_set_fp_traps(0))
    assert wfl.lib.wfl_set_fp_traps(3) == -1
    import subprocess
    import sys
    code = (
        "import ctypes as C, numpy as np, sys\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "import __graft_entry__ as ge\n"
        "w = ge._import_pkg()\n"
        "ws = np.zeros((4, 4), np.float32); ws[1, 0] = np.inf\n"
        "wf = np.zeros((4, 4), np.float32); WL = np.arange(4, dtype=np.int32); ev = C.c_double()\n"
        "w.lib.wfl_set_fp_traps(2)\n"
        "w.lib.EVMean(C.byref(ev), ws.ctypes.data, wf.ctypes.data, 4, WL.ctypes.data, 1.0, 0)\n"
        "print('not reached')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "Error: Floating point signal detected." in r.stdout and "not reached" not in r.stdout


def test_two_level_estimator(wfl, orc):
    """The two-level N -> infinity estimate 2 W_{N/2} - W_{N/4} removes a
    discretization error a/N_k exactly (W_{N/2} = W + 2a/N, W_{N/4} =
    W + 4a/N), where the reference's 8/3, -2, 1/3 on (W_N, W_{N/2}, W_{N/4})
    does only if W_N's error is a/N; L_eff with either estimator equals the
    oracle's."""
    W, a, N, cnt = -0.0123, 0.37, 1024.0, 128
    g3 = np.array([0.0, W + 2 * a / N, W + 4 * a / N]) * cnt
    assert wfl.lib.wfl_group_ev_minus1_est(g3.ctypes.data, cnt, 1) == pytest.approx(W, rel=1e-13)
    g3[0] = (W + a / N) * cnt
    assert wfl.lib.wfl_group_ev_minus1_est(g3.ctypes.data, cnt, 0) == pytest.approx(W, rel=1e-12)
    g3[0] = (W + 1.7 * a / N) * cnt           # a blend whose error is not a/N
    assert abs(wfl.lib.wfl_group_ev_minus1_est(g3.ctypes.data, cnt, 0) - W) > 1e-4
    x, y = orc.gen_loops(9, 8, 64)
    T = np.logspace(-2, 1, 7)
    og = orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0)
    gs, _ = orc.eval_batch([og], orc.MODE_PHYSICAL, 2, 0, x, y, 4, np.zeros(len(T), np.int32), np.full(len(T), 0.3),
                           np.zeros(len(T)), T, perloop=False)
    geom = wfl.Geometry(kind=wfl.GAUSSIAN, F=1.0, l2=1.0)
    for est in (0, 1):
        got = wfl.leff_groups(geom, 0, T, 0.3, 0.0, gs, np.array([4, 4]), 0, est)
        ref = orc.leff_groups(og, 0, T, 0.3, 0.0, gs, np.array([4, 4]), 0, est)
        assert np.allclose(got, ref, rtol=1e-12)


# analysis-scripts/EAscffvsconst.sage:63-74: the leading-order derivative
# expansion (locally constant field) of L_eff for the fixflux profile
# (F = 1, tube l = 0.3, tubedist a = 1, lmin = 0.1, q1 = 0.443991,
# q2 = 0.0742478), -(1/2 pi) int_0^100 e^-T/T^3 (eBT/sinh(eBT) - 1 +
# (eBT)^2/6) dT, tabulated by the script on rho = 0.005, 0.010, ... (its
# first 30 entries; past rho = l/2 only the background field remains).
FIXFLUX_LCF_TABLE = [
    -3293.28555160, -3268.32687001, -3227.08914982, -3169.55070950, -3095.55024139,
    -3005.60011341, -2899.58309952, -2778.08130014, -2641.45440100, -2490.31314204,
    -2325.39478220, -2147.84432326, -1958.80989950, -1760.12013070, -1553.99342531,
    -1343.23209559, -1131.33125294, -922.575232475, -722.092884924, -535.815812501,
    -370.244365454, -231.872091647, -126.085122294, -55.3961420686, -17.2828780744,
    -2.99638285276, -0.195028060671, -0.0101282156775, -0.00610544135068, -0.00610357265228]


def test_fixflux_field_and_exact_reproduce_the_reference_lcf_table(wfl):
    """The library's fixflux field (MagField, with the script's q1, q2) and
    its constant-field closed form (Exact, scalar) reproduce the reference
    analysis script's locally-constant-field table: the T-integral of
    Exact(T, B(rho)) over (0, 100], times -1/(2 pi), at each tabulated rho,
    within 5e-5 (the script's own numerical integration of the sharply
    peaked integrand at the tube centre, eB T ~ 1 at T ~ 0.005, is good to
    a few 1e-5 there and to ~1e-6 elsewhere).  A wrong field normalisation,
    core or background term, or a wrong sign or term in Exact fails it."""
    from scipy import integrate
    geom = wfl.Geometry(kind=wfl.FIXFLUX, F=1.0, l2=0.09, a=1.0, lam_min=0.1, fix_q1=0.443991, fix_q2=0.0742478)
    wfl.set_profile(geom)
    p = wfl.Wparams()
    p.F, p.l2 = 1.0, 0.09
    for i, want in enumerate(FIXFLUX_LCF_TABLE):
        rho = 0.005 * (i + 1)
        B = wfl.lib.MagField(rho * rho, C.byref(p))
        # (no break points: near T = 0 the closed form's 1 - 1 + (eBT)^2/6 cancels,
        # and an adaptive rule that is made to sample there reads the noise)
        val = -integrate.quad(lambda T: wfl.lib.Exact(T, B, 0), 0.0, 100.0, limit=800, epsabs=0.0,
                              epsrel=1e-11)[0] / (2.0 * np.pi)
        assert abs(val - want) <= 5e-5 * abs(want), (rho, val, want)


def test_exact_reproduces_the_reference_scripts_constant_field_values(wfl):
    """The reference's analysis scripts plot the constant-field L_eff at
    eB = m^2 = 1 as `exact`: 0.0139688 for scalar QED
    (analysis-scripts/Leffvsrho.sage:17) and -0.0164599 for spinor QED
    (Leffvsrho.F1l1.sage:30), the proper-time integrals int_0^inf of the
    closed forms.  The library's Exact (Integrand.c:20-37) integrated over
    T at B = 1/e gives both to the scripts' 6 printed digits (half a unit
    of the last digit, 5e-8, plus the quadrature's error)."""
    from scipy import integrate
    e = 0.30282212
    for fermion, want in ((0, 0.0139688), (1, -0.0164599)):
        val = integrate.quad(lambda T: wfl.lib.Exact(T, 1.0 / e, fermion), 0.0, np.inf, limit=500, epsrel=1e-10)[0]
        assert abs(val - want) <= 5.5e-8, (fermion, val, want)
