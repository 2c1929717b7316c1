"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no golden vectors, known-answer tests or fixtures for this
path (it has no tests at all), so each part of the oracle is checked against
a closed form, an invariant or a brute-force restatement chosen so that a
plausible transcription mistake (a wrong stencil offset, a dropped term, a
swapped coefficient, a wrong sign) fails:

  C-PML profile      closed form in double, invariants, left/right symmetry
  tapers             closed form of the sin^2 / cos^2 ramps
  reduction          the reference's fp32 block sum vs a double sum
  stencils           exact derivatives of cubic polynomials at the staggered
                     points, every increment of every kernel (fp64 oracle)
  reference quirks   the 13/12 slope of the dsxy/dx typo, the off-by-one
                     velocity C-PML predicate, the swapped adjoint C-PML
                     coefficients
  time stepping      reverse steps undo forward steps in the interior
                     (boundary-band restore included)
"""
import numpy as np
import pytest

from util import rel_l2

H = 10.0
DT = 1e-3
DTF = float(np.float32(DT))  # the grid carries dt as float


def _grid(orc, n=24, npml=6, nt=10, h=H, f0=10.0):
    return orc.Grid(n, n, n, npml, nt, h, h, h, DT, f0)


# ---------------------------------------------------------------------------
# C-PML profile (Src/utilities.cu:204-320)
# ---------------------------------------------------------------------------
def _cpml_closed_form(N, npml, h, f0, dt):
    """Closed form in double: d0 = -(N+1) v ln R / (2L), q = depth / L,
    damp = d0 (q/4 + 3 q^8 / 4), K = 1 + q^8, alpha = pi f0 (1 - q),
    b = exp(-(damp/K + alpha) dt), a = damp (b - 1) / (K (damp + K alpha))."""
    L = npml * h
    d0 = -9.0 * 3000.0 * np.log(0.0008) / (2 * L)
    out = np.zeros((6, N))
    out[0] = out[3] = 1.0
    out[2] = out[5] = 1.0
    for half, off in ((0, 0.0), (1, 0.5)):
        for i in range(N):
            dl = (npml - i - off) * h
            dr = (npml - N + i + off) * h
            d = dl if dl >= 0 else (dr if dr >= 0 else None)
            if dr >= 0:
                d = dr
            if d is None:
                continue
            q = d / L
            damp = d0 * (0.25 * q + 0.75 * q ** 8)
            K = 1 + q ** 8
            al = np.pi * f0 * (1 - q)
            b = np.exp(-(damp / K + al) * dt)
            a = damp * (b - 1) / (K * (damp + K * al)) if damp > 1e-6 else 0.0
            out[3 * half + 0, i] = K
            out[3 * half + 1, i] = a
            out[3 * half + 2, i] = b
    return out


@pytest.mark.parametrize("N,npml,h,f0", [(84, 10, 10.0, 10.0), (240, 20, 12.5, 8.0), (40, 8, 7.0, 25.0)])
def test_cpml_profile_closed_form(orc, N, npml, h, f0):
    got = orc.cpml_profile(N, npml, h, f0, DT).astype(np.float64)
    ref = _cpml_closed_form(N, npml, h, f0, DT)
    # float rounding of the intermediates only
    np.testing.assert_allclose(got[[0, 2, 3, 5]], ref[[0, 2, 3, 5]], rtol=2e-6, atol=0)
    np.testing.assert_allclose(got[[1, 4]], ref[[1, 4]], rtol=5e-5, atol=1e-9)


def test_cpml_profile_invariants(orc):
    N, npml = 84, 10
    p = orc.cpml_profile(N, npml, H, 10.0, DT)
    K, a, b, Kh, ah, bh = p
    inner = slice(npml + 1, N - npml)  # strictly outside both PMLs on the integer grid
    assert np.all(K[inner] == 1) and np.all(a[inner] == 0) and np.all(b[inner] == 1)
    assert np.all(Kh[npml:N - npml] == 1) and np.all(ah[npml:N - npml] == 0)
    assert np.all((b[:npml] > 0) & (b[:npml] < 1)) and np.all(a[:npml] < 0)
    assert np.all(K[:npml] >= 1) and np.all(K[:npml] <= 2) and np.all(np.diff(K[:npml]) <= 0)
    # integer grid: depth(i) on the left equals depth(N - i) on the right
    for i in range(1, npml + 1):
        assert K[i] == K[N - i] and a[i] == a[N - i] and b[i] == b[N - i]
    # half grid: left half point i-1/2... mirrors right i+1/2 about N/2
    for i in range(0, npml):
        assert Kh[i] == pytest.approx(Kh[N - 1 - i], rel=1e-6)
        assert ah[i] == pytest.approx(ah[N - 1 - i], rel=1e-5)
    # at the PML edge i = npml the depth is 0: K = 1, a = 0 (alpha = pi f0, harmless)
    assert K[npml] == 1 and a[npml] == 0


# ---------------------------------------------------------------------------
# Time taper (Src/utilities.cu:578-614) and the fp32 block reduction (:145-166)
# ---------------------------------------------------------------------------
def test_taper_closed_form(orc):
    # ratio 0.005 over 3000 steps of 1 ms: ramps of 15 ms -> w(t) = sin(pi/2 t/15ms)
    nt, dt = 3000, 1e-3
    w2 = orc.taper_w2(nt, dt, 0.005).astype(np.float64)
    t = np.arange(nt) * dt
    off = nt * dt * 0.005
    ref = np.ones(nt)
    m = t < off
    ref[m] = np.sin(np.pi / 2 * t[m] / off) ** 2
    m = t >= nt * dt - off
    ref[m] = np.cos(np.pi / 2 * (t[m] - (nt * dt - off)) / off) ** 2
    # float t = it*dt carries an absolute error of ~ulp(3.0) = 2.4e-7 s on a 15 ms ramp
    np.testing.assert_allclose(w2, ref, rtol=1e-5, atol=5e-5)
    # stf taper (ratio 0.001) over 500 steps: only the first sample is zeroed
    w2s = orc.taper_w2(500, 1e-3, 0.001)
    assert w2s[0] == 0 and np.all(w2s[1:] == 1)


def test_fp32_block_sum(orc):
    rng = np.random.default_rng(5)
    e = rng.standard_normal(3 * 4096 * 50).astype(np.float32)
    s32 = orc.fp32_block_sumsq(e)
    s64 = float(np.sum(e.astype(np.float64) ** 2))
    assert abs(s32 - s64) / s64 < 1e-5
    assert orc.fp32_block_sumsq(np.array([3.0, 4.0], np.float32)) == 25.0


# ---------------------------------------------------------------------------
# Stencils: exact derivatives of cubics at the staggered points (fp64 oracle)
# ---------------------------------------------------------------------------
def _cubic(c):
    return (lambda u: c[0] + c[1] * u + c[2] * u ** 2 + c[3] * u ** 3,
            lambda u: c[1] + 2 * c[2] * u + 3 * c[3] * u ** 2)


def _coords(n, h):
    i = np.arange(n) * h
    return np.meshgrid(i, i, i, indexing="ij")


@pytest.fixture
def poly():
    rng = np.random.default_rng(11)
    return [_cubic(rng.uniform(-1, 1, 4) * np.array([1, 1e-2, 1e-4, 1e-6])) for _ in range(9)]


def _interior_mask(n, npml):
    m = np.zeros((n, n, n), bool)
    m[npml:n - npml, npml:n - npml, npml:n - npml] = True
    return m


def test_stress_fwd_stencil_exact(orc, poly):
    g = _grid(orc)
    n = g.nz
    st = orc.Stepper(g, "fp64", cpml_off=True)
    Z, X, Y = _coords(n, H)
    (P, dP), (T, dT), (U, dU), (Q, dQ), (S, dS), (V, dV), (R, dR), (Xf, dX), (Yz, dYz) = poly
    st.set("vz", P(Z) + T(X) + U(Y))
    st.set("vx", Q(X) + S(Z) + V(Y))
    st.set("vy", R(Y) + Xf(X) + Yz(Z))
    lam, mu = 2.0, 3.0
    st.set("lam", lam)
    st.set("mu", mu)
    st.kernel("stress_fwd")
    h2 = H / 2
    ezz, exx, eyy = dP(Z - h2), dQ(X - h2), dR(Y - h2)
    m = np.zeros((n, n, n), bool)
    m[2:n - 2, 2:n - 2, 2:n - 2] = True
    for name, ref in (("szz", (lam + 2 * mu) * ezz + lam * exx + lam * eyy),
                      ("sxx", lam * ezz + (lam + 2 * mu) * exx + lam * eyy),
                      ("syy", lam * ezz + lam * exx + (lam + 2 * mu) * eyy),
                      ("sxz", mu * (dS(Z + h2) + dT(X + h2))),
                      ("sxy", mu * (dV(Y + h2) + dX(X + h2))),
                      ("syz", mu * (dYz(Z + h2) + dU(Y + h2)))):
        got = st.get(name) / DTF
        np.testing.assert_allclose(got[m], ref[m], rtol=1e-9, atol=1e-12, err_msg=name)
        assert np.all(got[~m] == 0), name  # only [2, n-3] is updated


@pytest.mark.parametrize("formulation", ["reference", "consistent"])
def test_velocity_fwd_stencil(orc, poly, formulation):
    g = _grid(orc)
    n = g.nz
    st = orc.Stepper(g, "fp64", formulation=formulation, cpml_off=True)
    Z, X, Y = _coords(n, H)
    (P, dP), (T, dT), (U, dU), (Q, dQ), (S, dS), (V, dV), (R, dR), (Xl, dXl), (Yz, dYz) = poly
    st.set("szz", P(Z))
    st.set("sxz", T(X) + S(Z))
    st.set("syz", U(Y) + Yz(Z))
    st.set("sxx", Q(X))
    st.set("syy", R(Y))
    slope = 0.37
    st.set("sxy", V(Y) + slope * X)  # linear in x: isolates the dsxy/dx stencil
    byc = 0.5
    st.set("byc", byc)
    st.kernel("velocity_fwd")
    h2 = H / 2
    m = np.zeros((n, n, n), bool)
    m[2:n - 2, 2:n - 2, 2:n - 2] = True
    # c1 - c2 = 13/12 on a linear function for the typo (Src/velocity.cu:79)
    fac = 13.0 / 12.0 if formulation == "reference" else 1.0
    refs = {"vz": dP(Z + h2) + dT(X - h2) + dU(Y - h2),
            "vx": dS(Z - h2) + dQ(X + h2) + dV(Y - h2),
            "vy": dYz(Z - h2) + fac * slope + dR(Y + h2)}
    for name, ref in refs.items():
        got = st.get(name) / (byc * DTF)
        np.testing.assert_allclose(got[m], ref[m], rtol=1e-9, atol=1e-12, err_msg=name)


def test_adjoint_stencils_exact(orc, poly):
    g = _grid(orc)
    n = g.nz
    Z, X, Y = _coords(n, H)
    h2 = H / 2
    lam, mu, byc = 2.0, 3.0, 0.5
    m = np.zeros((n, n, n), bool)
    m[2:n - 2, 2:n - 2, 2:n - 2] = True
    (P, dP), (T, dT), (U, dU), (Q, dQ), (S, dS), (V, dV), (R, dR), (Xl, dXl), (Yz, dYz) = poly
    # velocity_adj: v-dagger += C : (-D_f) s-dagger, dt, with psi = 0 and no PML
    st = orc.Stepper(g, "fp64", cpml_off=True)
    st.set("lam", lam); st.set("mu", mu); st.set("byc", byc)
    st.set("aszz", P(Z) + T(X) + U(Y))
    st.set("asxx", Q(X) + S(Z) + V(Y))
    st.set("asyy", R(Y) + Xl(X) + Yz(Z))
    st.set("asxz", Q(Z) + P(X))
    st.set("asxy", S(Y) + R(X))
    st.set("asyz", V(Z) + T(Y))
    st.kernel("velocity_adj")
    l2m = lam + 2 * mu
    ref_vx = -(l2m * dQ(X + h2) + lam * dT(X + h2) + lam * dXl(X + h2)) - mu * dQ(Z - h2) - mu * dS(Y - h2)
    ref_vz = -(l2m * dP(Z + h2) + lam * dS(Z + h2) + lam * dYz(Z + h2)) - mu * dT(Y - h2) - mu * dP(X - h2)
    ref_vy = -(l2m * dR(Y + h2) + lam * dV(Y + h2) + lam * dU(Y + h2)) - mu * dR(X - h2) - mu * dV(Z - h2)
    for name, ref in (("avx", ref_vx), ("avz", ref_vz), ("avy", ref_vy)):
        np.testing.assert_allclose(st.get(name)[m] / DTF, ref[m], rtol=1e-9, atol=1e-12, err_msg=name)
    # stress_adj: s-dagger += Byc (-D) v-dagger, dt
    st = orc.Stepper(g, "fp64", cpml_off=True)
    st.set("lam", lam); st.set("mu", mu); st.set("byc", byc)
    st.set("avz", P(Z) + T(X) + U(Y))
    st.set("avx", Q(X) + S(Z) + V(Y))
    st.set("avy", R(Y) + Xl(X) + Yz(Z))
    st.kernel("stress_adj")
    refs = {"asxz": -byc * (dT(X + h2) + dS(Z + h2)),
            "asxy": -byc * (dXl(X + h2) + dV(Y + h2)),
            "asyz": -byc * (dU(Y + h2) + dYz(Z + h2)),
            "asxx": -byc * dQ(X - h2), "aszz": -byc * dP(Z - h2), "asyy": -byc * dR(Y - h2)}
    for name, ref in refs.items():
        np.testing.assert_allclose(st.get(name)[m] / DTF, ref[m], rtol=1e-9, atol=1e-12, err_msg=name)


def test_reverse_kernels_imaging_terms(orc, poly):
    """stress_rev / velocity_rev subtract the forward update and accumulate the
    imaging terms of Src/stress.cu:130-134 and Src/velocity.cu:268-270."""
    g = _grid(orc)
    n, p = g.nz, g.npml
    Z, X, Y = _coords(n, H)
    h2 = H / 2
    lam, mu = 2.0, 3.0
    (P, dP), (T, dT), (U, dU), (Q, dQ), (S, dS), (V, dV), (R, dR), (Xl, dXl), (Yz, dYz) = poly
    st = orc.Stepper(g, "fp64")
    st.set("lam", lam); st.set("mu", mu)
    st.set("vz", P(Z) + T(X)); st.set("vx", Q(X) + S(Z)); st.set("vy", R(Y) + Xl(X))
    a = {"aszz": 0.3, "asxx": -0.2, "asyy": 0.7, "asxz": 0.11, "asxy": -0.5, "asyz": 0.9}
    for k, v in a.items():
        st.set(k, v)
    st.kernel("stress_rev")
    ezz, exx, eyy = dP(Z - h2), dQ(X - h2), dR(Y - h2)
    exz = dS(Z + h2) + dT(X + h2)
    exy = dXl(X + h2)
    eyz = 0.0 * Z
    m = _interior_mask(n, p)
    div = ezz + exx + eyy
    ref_gl = -(a["aszz"] + a["asxx"] + a["asyy"]) * div
    ref_gm = -(2 * a["aszz"] * ezz + 2 * a["asxx"] * exx + 2 * a["asyy"] * eyy + a["asxz"] * exz +
               a["asyz"] * eyz + a["asxy"] * exy)
    np.testing.assert_allclose(st.get("glam")[m] / DTF, ref_gl[m], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(st.get("gmu")[m] / DTF, ref_gm[m], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(-st.get("szz")[m] / DTF, ((lam + 2 * mu) * ezz + lam * exx + lam * eyy)[m],
                               rtol=1e-9, atol=1e-12)
    assert np.all(st.get("glam")[~m] == 0) and np.all(st.get("szz")[~m] == 0)
    # density term: grho += -sum_c v_c^dag (div s)_c dt (-Byc^2)
    st = orc.Stepper(g, "fp64", formulation="consistent")
    byc = 0.5
    st.set("byc", byc)
    st.set("szz", P(Z)); st.set("sxz", T(X)); st.set("sxx", Q(X)); st.set("syy", R(Y))
    av = {"avz": 0.4, "avx": -0.3, "avy": 0.8}
    for k, v in av.items():
        st.set(k, v)
    st.kernel("velocity_rev")
    divz, divx, divy = dP(Z + h2) + dT(X - h2), dQ(X + h2), dR(Y + h2)
    ref = byc ** 2 * (av["avz"] * divz + av["avx"] * divx + av["avy"] * divy)
    np.testing.assert_allclose(st.get("grho")[m] / DTF, ref[m], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(-st.get("vz")[m] / (byc * DTF), divz[m], rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------------------
# Reference quirks in the C-PML handling
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("formulation", ["reference", "consistent"])
def test_velocity_cpml_predicate_off_by_one(orc, formulation):
    """Src/velocity.cu:46-90 use i > n - nPml (the stress kernel i > n - nPml - 1):
    cell x = nx - nPml is a PML cell of the stress kernel only. Its half-grid
    coefficient a_half is non-zero there, so the velocity kernel's dsxx/dx
    memory variable stays 0 in the reference formulation only."""
    g = _grid(orc, n=28, npml=6)
    n, p = g.nz, g.npml
    Z, X, Y = _coords(n, H)
    st = orc.Stepper(g, "fp64", formulation=formulation)
    st.set("lam", 2.0); st.set("mu", 1.0); st.set("byc", 0.5)
    st.set("vz", 1e-3 * X ** 2)
    st.set("sxx", 1e-3 * X ** 2)
    st.kernel("stress_fwd")
    st.kernel("velocity_fwd")
    c = n // 2
    mv = st.get("m_dvz_dx")[c, :, c]   # stress kernel, half-grid x coefficients
    ms = st.get("m_dsxx_dx")[c, :, c]  # velocity kernel, half-grid x coefficients
    edge = n - p
    assert st.cpml(1)[4, edge] != 0
    assert mv[edge] != 0 and mv[edge - 1] == 0
    if formulation == "reference":
        assert ms[edge] == 0 and ms[edge + 1] != 0
    else:
        assert ms[edge] != 0
    assert ms[p] == 0 and ms[p - 1] != 0


@pytest.mark.parametrize("formulation", ["reference", "consistent"])
def test_adjoint_cpml_coefficient_swap(orc, formulation):
    """Src/velocity_adj.cu:285,288: psi_dsyz_dz decays with b_z_half and
    psi_dsyy_dy with b_y (the forward uses b_z and b_y_half)."""
    g = _grid(orc, n=28, npml=6)
    n = g.nz
    st = orc.Stepper(g, "fp64", formulation=formulation)
    st.set("byc", 0.5)
    st.set("m_dsyz_dz", 1.0)
    st.set("m_dsyy_dy", 1.0)
    st.kernel("velocity_adj")  # v-dagger stays 0: psi <- b psi
    bz = st.cpml(0)
    by = st.cpml(2)
    z, y = 3, 3
    c = n // 2
    got_z = st.get("m_dsyz_dz")[z, c, c]
    got_y = st.get("m_dsyy_dy")[c, c, y]
    if formulation == "reference":
        assert got_z == bz[5, z] and got_y == by[2, y]
    else:
        assert got_z == bz[2, z] and got_y == by[5, y]
    assert bz[5, z] != bz[2, z] and by[2, y] != by[5, y]


# ---------------------------------------------------------------------------
# Reverse reconstruction (Src/libCUFD.cu:532-560 with the band of Src/Boundary.cu)
# ---------------------------------------------------------------------------
def test_reverse_steps_undo_forward(orc):
    g = _grid(orc, n=30, npml=6, nt=40)
    n = g.nz
    rng = np.random.default_rng(2)
    st = orc.Stepper(g, "fp64")
    lam = 2.2 * (2500 ** 2 - 2 * 1450 ** 2) * (1 + 0.02 * rng.standard_normal((n,) * 3))
    mu = 2.2 * 1450 ** 2 * (1 + 0.02 * rng.standard_normal((n,) * 3))
    st.set("lam", lam); st.set("mu", mu); st.set("byc", 1 / 2.2)
    band = orc.band_mask(g)
    fields = ["szz", "sxz", "sxx", "syy", "sxy", "syz", "vz", "vx", "vy"]
    stf = rng.standard_normal(g.nt) * 1e-3
    src = (n // 2, n // 2, n // 2)
    saved = []
    for it in range(g.nt - 1):
        saved.append({f: st.get(f)[band] for f in fields})
        st.kernel("stress_fwd")
        st.add_source(*src, float(np.float32(stf[it])), +1)
        st.kernel("velocity_fwd")
    peak = max(np.abs(st.get(f)).max() for f in ("vz", "vx", "vy"))
    assert peak > 0
    for it in range(g.nt - 2, -1, -1):
        st.kernel("velocity_rev")
        for f in ("vz", "vx", "vy"):
            a = st.get(f); a[band] = saved[it][f]; st.set(f, a)
        st.add_source(*src, float(np.float32(stf[it])), -1)
        st.kernel("stress_rev")
        for f in fields[:6]:
            a = st.get(f); a[band] = saved[it][f]; st.set(f, a)
    m = _interior_mask(n, g.npml)
    for f in ("vz", "vx", "vy"):
        assert np.abs(st.get(f)[m]).max() < 1e-9 * peak, f


def test_band_mask_is_box_minus_inner_box(orc):
    """The union of the six five-layer faces of from_bnd (Src/utilities.cu:323-372,
    restated here from its index decomposition) is the box [p-2, n-p+1]^3 minus
    the inner box [p+3, n-p-4]^3."""
    g = orc.Grid(30, 34, 38, 6, 4, H, H, H, DT, 10.0)
    p = g.npml
    nzB, nxB, nyB = g.nz - 2 * p + 4, g.nx - 2 * p + 4, g.ny - 2 * p + 4
    L = 5
    ref = np.zeros((g.nz, g.nx, g.ny), bool)
    total = 2 * L * (nzB * nxB + nxB * nyB + nyB * nzB)
    for idx in range(total):
        A, Bq = L * nzB * nxB, L * (2 * nzB * nxB + nxB * nyB)
        if idx < A:
            k = idx // (nzB * nxB); j = (idx - k * nzB * nxB) // nzB; i = idx - k * nzB * nxB - j * nzB
            z, x, y = i + p - 2, j + p - 2, k + p - 2
        elif idx < 2 * A:
            r = idx - A
            k = r // (nzB * nxB); j = (r - k * nzB * nxB) // nzB; i = r - k * nzB * nxB - j * nzB
            z, x, y = i + p - 2, j + p - 2, g.ny - p - k - 1 + 2
        elif idx < Bq:
            r = idx - 2 * A
            i = r // (nxB * nyB); k = (r - i * nxB * nyB) // nxB; j = r - i * nxB * nyB - k * nxB
            z, x, y = i + p - 2, j + p - 2, k + p - 2
        elif idx < 2 * L * (nzB * nxB + nxB * nyB):
            r = idx - Bq
            i = r // (nxB * nyB); k = (r - i * nxB * nyB) // nxB; j = r - i * nxB * nyB - k * nxB
            z, x, y = g.nz - p - i - 1 + 2, j + p - 2, k + p - 2
        elif idx < L * (2 * (nzB * nxB + nxB * nyB) + nyB * nzB):
            r = idx - 2 * L * (nzB * nxB + nxB * nyB)
            j = r // (nyB * nzB); i = (r - j * nyB * nzB) // nyB; k = r - j * nyB * nzB - i * nyB
            z, x, y = i + p - 2, j + p - 2, k + p - 2
        else:
            r = idx - L * (2 * (nzB * nxB + nxB * nyB) + nyB * nzB)
            j = r // (nyB * nzB); i = (r - j * nyB * nzB) // nyB; k = r - j * nyB * nzB - i * nyB
            z, x, y = i + p - 2, g.nx - p - j - 1 + 2, k + p - 2
        ref[z, x, y] = True
    got = orc.band_mask(g)
    assert np.array_equal(got, ref)
    box = np.zeros_like(ref)
    box[p - 2:g.nz - p + 2, p - 2:g.nx - p + 2, p - 2:g.ny - p + 2] = True
    inner = np.zeros_like(ref)
    inner[p + 3:g.nz - p - 3, p + 3:g.nx - p - 3, p + 3:g.ny - p - 3] = True
    assert np.array_equal(got, box & ~inner)


# ---------------------------------------------------------------------------
# Source and receiver terms (Src/utilities.cu:442-495, Src/libCUFD.cu:319-351,
# Src/Model.cu:65): zero and linear response in the source time function, and
# the exact wavefield of an explosive point source in a homogeneous medium
# (tests/analytic.py), which fixes the amplitude 1500^2 s dt, the buoyancy
# 1/rho, the source and receiver cells and the it + 1 record slot.
# ---------------------------------------------------------------------------
from analytic import Explosion, components  # noqa: E402

_EXPL = dict(n=72, npml=12, nt=260, h=5.0, dt=1e-3, f0=25.0, vp=2000.0, rho=2.0)
_EXPL_SRC = (33, 38, 35)
_EXPL_OFFS = [(0, 10, 0), (0, 0, 14), (8, 8, 0), (-10, 6, 6), (0, -18, 0), (12, 0, -9), (0, 20, 5),
              (-15, -5, 0)]


def _homogeneous(orc, src, offs, stf, rho_scale=1.0, precision="fp64", nt=None):
    e = _EXPL
    nt = nt or e["nt"]
    n = e["n"]
    g = orc.Grid(n, n, n, e["npml"], nt, e["h"], e["h"], e["h"], e["dt"], e["f0"])
    vp, rho = e["vp"], e["rho"]
    vs = vp / 1.73
    lam, mu = (vp ** 2 - 2 * vs ** 2) * rho, vs ** 2 * rho
    shp = (n, n, n)
    rec = np.array([[_EXPL_SRC[0] + a, _EXPL_SRC[1] + b, _EXPL_SRC[2] + c] for a, b, c in offs])
    sh = orc.Shots(src=np.array([src]), stf=np.asarray(stf, np.float32)[None, :nt], rec=[rec])
    r = orc.run(g, sh, np.full(shp, lam, np.float32), np.full(shp, mu, np.float32),
                np.full(shp, rho * rho_scale, np.float32), calc_id=2, precision=precision)
    return r.syn[0]


def _stf(nt=None):
    from efwi import synth
    e = _EXPL
    s = synth.ricker_integrated(e["f0"], nt or e["nt"], e["dt"]).astype(np.float32)
    s[0] = 0.0  # the 0.001 stf taper (Src/Src_Rec.cu:146) zeroes sample 0 only; keep it identity
    return s


def _explosion_fits(syn, stf):
    e = _EXPL
    A = Explosion(stf, e["dt"], e["rho"], e["vp"], e["h"])
    return [(k, ci, A.fit(syn[ci, k], syn.shape[2], rel, axis)) for k, ci, rel, axis in components(_EXPL_OFFS)]


def _explosion_ok(fits):
    return all(abs(f["shift"]) <= 0.4 and 0.92 <= f["amp"] <= 1.08 and f["err"] <= 0.08 and f["err0"] <= 0.10
               for _, _, f in fits)


def test_zero_source_gives_zero_output(orc):
    syn = _homogeneous(orc, _EXPL_SRC, _EXPL_OFFS[:3], np.zeros(_EXPL["nt"], np.float32), nt=80)
    assert np.count_nonzero(syn) == 0


def test_output_linear_in_source(orc):
    nt = 120
    s1 = _stf(nt)
    rng = np.random.default_rng(5)
    s2 = (rng.standard_normal(nt) * np.abs(s1).max()).astype(np.float32)
    s2[0] = 0.0
    a, b = 0.75, -1.5
    y1 = _homogeneous(orc, _EXPL_SRC, _EXPL_OFFS[:4], s1, nt=nt)
    y2 = _homogeneous(orc, _EXPL_SRC, _EXPL_OFFS[:4], s2, nt=nt)
    y12 = _homogeneous(orc, _EXPL_SRC, _EXPL_OFFS[:4], (a * s1.astype(np.float64) + b * s2).astype(np.float32),
                       nt=nt)
    ref = a * y1 + b * y2
    assert np.abs(y1).max() > 0
    # the combined stf is rounded to float once: that is the only difference
    assert rel_l2(y12, ref) <= 1e-6


def test_point_source_matches_exact_explosion(orc):
    """fp64 oracle against the closed-form explosion at 8 receivers 10-21 cells
    away in every direction: best time shift within 0.4 samples of zero,
    amplitude within 8 %, shape within 8 % (4th-order grid dispersion at
    ~7 cells per shortest wavelength and the point source's discretisation)."""
    stf = _stf()
    fits = _explosion_fits(_homogeneous(orc, _EXPL_SRC, _EXPL_OFFS, stf), stf)
    for k, ci, f in fits:
        print(f"receiver {k} component {'xyz'[ci]}: shift {f['shift']:+.2f} amp {f['amp']:.3f} "
              f"err {f['err']:.3f} err0 {f['err0']:.3f}")
    assert len(fits) >= 12
    assert _explosion_ok(fits)


def test_explosion_check_catches_planted_mistakes(orc):
    """The same check fails for each mistake it is meant to catch, planted in
    the oracle's inputs or outputs: the amplitude 1500 instead of 1500^2, the
    source one cell off, the record slot it instead of it + 1, and rho in
    place of the buoyancy 1/rho."""
    stf = _stf()
    good = _homogeneous(orc, _EXPL_SRC, _EXPL_OFFS, stf)
    assert _explosion_ok(_explosion_fits(good, stf))
    # 1500 * s * dt
    assert not _explosion_ok(_explosion_fits(good / 1500.0, stf))
    # record slot it: every trace one sample earlier
    early = np.zeros_like(good)
    early[:, :, :-1] = good[:, :, 1:]
    assert not _explosion_ok(_explosion_fits(early, stf))
    # source one cell off along each axis
    for ax in range(3):
        src = list(_EXPL_SRC)
        src[ax] += 1
        assert not _explosion_ok(_explosion_fits(_homogeneous(orc, src, _EXPL_OFFS, stf), stf)), ax
    # buoyancy rho instead of 1/rho: the model's rho becomes 1/rho
    assert not _explosion_ok(_explosion_fits(_homogeneous(orc, _EXPL_SRC, _EXPL_OFFS, stf,
                                                          rho_scale=1.0 / _EXPL["rho"] ** 2), stf))
