"""The station table of the device station kernels (m3d_station_desc),
checked on the CPU: Z and the adjoint right-hand side evaluated from the
table alone, by the rules the kernels follow (group sums by role, H from the
finite-difference curl, the contribution list per station edge), equal the
host restatement of extract_impedance_ / build_adjoint_rhs_from_residual
(forward_solver_fv.cpp:1490-1727, :1804-2013) on random fields."""
import numpy as np
import pytest

from maple3dmt_amd import configs

MU0 = 4e-7 * np.pi


@pytest.fixture(scope="module")
def model():
    wl = configs.random3d(7, min_level=3, max_level=5, n_freq=2, half_width=50e3, station_radius=5e3)
    return configs.build_model(wl)


def table(m):
    return {k: m.array("st_" + k) for k in ("valid", "edge_ptr", "edge", "role", "cell_size", "uedge",
                                            "contrib_ptr", "contrib", "contrib_we", "sigma")}


def z_from_table(t, w, e1, e2):
    ns = len(t["valid"])
    Z = np.zeros((ns, 4), complex)
    cache = []
    for s in range(ns):
        if not t["valid"][s]:
            cache.append(None)
            continue
        g = {}
        for k in range(t["edge_ptr"][s], t["edge_ptr"][s + 1]):
            e, r = t["edge"][k], t["role"][k]
            d = r & 3
            keys = [("x" if d == 0 else "y") + ("t" if r & 4 else "b")] if d < 2 else \
                ["zpx" if r & 4 else "zmx", "zpy" if r & 8 else "zmy"]
            for key in keys:
                a = g.setdefault(key, [0j, 0j, 0])
                a[0] += e1[e]
                a[1] += e2[e]
                a[2] += 1
        for a in g.values():
            a[0] /= a[2]
            a[1] /= a[2]
        get = lambda key, p: g[key][p] if key in g else 0j  # noqa: E731
        has = lambda key: key in g  # noqa: E731
        def centre(tk, bk, p):
            if has(tk) and has(bk):
                return 0.5 * (get(tk, p) + get(bk, p))
            return get(tk, p) if has(tk) else get(bk, p)
        dx, dy, dz = t["cell_size"][3 * s:3 * s + 3]
        Ex = [centre("xt", "xb", p) for p in (0, 1)]
        Ey = [centre("yt", "yb", p) for p in (0, 1)]
        diff = lambda a, b, h, p: (get(a, p) - get(b, p)) / h if has(a) and has(b) else 0j  # noqa: E731
        iwmu = 1j * w * MU0
        Hx = [(diff("zpy", "zmy", dy, p) - diff("yt", "yb", dz, p)) / iwmu for p in (0, 1)]
        Hy = [(diff("xt", "xb", dz, p) - diff("zpx", "zmx", dx, p)) / iwmu for p in (0, 1)]
        det = Hx[0] * Hy[1] - Hx[1] * Hy[0]
        Z[s] = [(Ex[0] * Hy[1] - Ex[1] * Hy[0]) / det, (Ex[1] * Hx[0] - Ex[0] * Hx[1]) / det,
                (Ey[0] * Hy[1] - Ey[1] * Hy[0]) / det, (Ey[1] * Hx[0] - Ey[0] * Hx[1]) / det]
        cache.append((Hx, Hy, det, Z[s].copy()))
    return Z, cache


def rhs_from_table(t, w, cache, wr, n):
    ns = len(t["valid"])
    coef = []
    for s in range(ns):
        if cache[s] is None:
            coef.append(None)
            continue
        (Hx1, Hx2), (Hy1, Hy2), det, (Zxx, Zxy, Zyx, Zyy) = cache[s]
        q = wr[s]
        sx1 = (q[0] * Hy2 - q[1] * Hx2) / det
        sy1 = (q[2] * Hy2 - q[3] * Hx2) / det
        sx2 = (-q[0] * Hy1 + q[1] * Hx1) / det
        sy2 = (-q[2] * Hy1 + q[3] * Hx1) / det
        aH = {"x1": -(Zxx * sx1 + Zyx * sy1), "y1": -(Zxy * sx1 + Zyy * sy1),
              "x2": -(Zxx * sx2 + Zyx * sy2), "y2": -(Zxy * sx2 + Zyy * sy2)}
        h = t["cell_size"][3 * s:3 * s + 3]
        iax = [1 / (1j * w * MU0 * h[2]), 1 / (1j * w * MU0 * h[0]), 1 / (1j * w * MU0 * h[1])]
        coef.append((sx1, sy1, sx2, sy2, aH, iax))
    r1, r2 = np.zeros(n, complex), np.zeros(n, complex)
    c = t["contrib"].reshape(-1, 4)
    for k, e in enumerate(t["uedge"]):
        for j in range(t["contrib_ptr"][k], t["contrib_ptr"][k + 1]):
            s, kind, ax, cnt = c[j]
            if coef[s] is None:
                continue
            sx1, sy1, sx2, sy2, aH, iax = coef[s]
            if kind < 2:
                we = t["contrib_we"][j]
                r1[e] += (sx1 if kind == 0 else sy1) * we
                r2[e] += (sx2 if kind == 0 else sy2) * we
            else:
                wh = (-1 if ax & 1 else 1) * iax[ax >> 1] / cnt
                comp = "y" if kind == 2 else "x"
                r1[e] += aH[comp + "1"] * wh
                r2[e] += aH[comp + "2"] * wh
    return r1, r2


def test_table_reproduces_host_impedance_and_adjoint_rhs(model):
    m = model
    t = table(m)
    ns = len(t["valid"])
    assert ns == 121 and t["valid"].all()
    # Contributions reference stations and station edges in range.
    c = t["contrib"].reshape(-1, 4)
    assert c[:, 0].max() < ns and set(c[:, 1]) <= {0, 1, 2, 3} and len(c) == t["contrib_ptr"][-1]
    assert np.array_equal(t["uedge"], np.sort(np.unique(t["edge"])))
    rng = np.random.default_rng(3)
    for w in (2 * np.pi * 1e-3, 2 * np.pi * 10.0):
        e1 = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
        e2 = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
        zt, cache = z_from_table(t, w, e1, e2)
        zh = m.extract_impedance(w, e1, e2)
        assert np.abs(zt - zh).max() <= 1e-12 * np.abs(zh).max()
        wr = rng.standard_normal((ns, 4)) + 1j * rng.standard_normal((ns, 4))
        r1, r2 = rhs_from_table(t, w, cache, wr, m.n)
        h1, h2 = m.adjoint_rhs(w, wr)
        assert np.abs(r1 - h1).max() <= 1e-12 * np.abs(h1).max()
        assert np.abs(r2 - h2).max() <= 1e-12 * np.abs(h2).max()
        # Nonzero exactly on the station edges.
        assert set(np.flatnonzero(h1)) <= set(t["uedge"])
