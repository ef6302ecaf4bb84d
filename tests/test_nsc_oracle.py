"""Pinning of the compressible NS oracle (oracle.c or_nsc_*, a restatement of examples/ns-residual/physics.hpp and
reference.hpp:29-128). The reference cannot be built here (physics.hpp includes sycl/sycl.hpp), so each piece is checked
against what the mathematics fixes, chosen so that a plausible slip (a wrong sign, a swapped side or component, a
dropped term, a wrong neighbour index) fails one of them:

  limiters      closed forms: zero at extrema, minmod = the smaller magnitude, van Leer = the harmonic form
  HLLC          consistency F(w, w) = the Euler flux; supersonic upwinding; exact stationary contacts and shear layers
                (its defining property); rotation (dir 1 = dir 0 with u, v swapped); mirror symmetry x -> -x
  viscous       the Stokes stress law and Fourier heat flux on linear fields (closed form)
  diagnostics   closed forms for uniform flow
  residual      uniform periodic flow gives exactly 0 (the reference's own check); conservation (the sum of R over a
                periodic domain vanishes); periodic translation invariance (bitwise); mirror and transpose symmetries;
                walls = the lower half of a periodic domain of twice the height holding the mirrored field (bitwise)
"""
import numpy as np
import pytest

import oracle as O

G, MU, PR, CFL = O.NSC_GAMMA, O.NSC_MU, O.NSC_PR, O.NSC_CFL


def euler_flux(w, dir):
    r, u, v, p = w
    un, ut = (u, v) if dir == 0 else (v, u)
    E = p / (G - 1) + 0.5 * r * (u * u + v * v)
    f = np.array([r * un, r * un * un + p, r * un * ut, (E + p) * un])
    return f if dir == 0 else f[[0, 2, 1, 3]]


def rand_state(rng, umax=0.5):
    return np.array([rng.uniform(0.5, 2.0), rng.uniform(-umax, umax), rng.uniform(-umax, umax), rng.uniform(0.5, 2.0)])


# ---- limiters ---------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("vl", [False, True])
def test_limiters_closed_forms(vl):
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b = rng.normal(size=2)
        s = O.nsc_slope(a, b, vl)
        if a * b <= 0:
            assert s == 0.0
        elif vl:
            assert s == pytest.approx((a * abs(b) + abs(a) * b) / (abs(a) + abs(b)), rel=1e-14)
        else:
            assert s == (a if abs(a) < abs(b) else b)
    assert O.nsc_slope(0.25, 0.25, vl) == 0.25  # a linear field keeps its slope


# ---- HLLC -------------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dir", [0, 1])
def test_hllc_consistency(dir):
    rng = np.random.default_rng(2)
    for _ in range(100):
        w = rand_state(rng)
        np.testing.assert_allclose(O.nsc_hllc(w, w, dir), euler_flux(w, dir), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("dir", [0, 1])
def test_hllc_supersonic_upwinds(dir):
    rng = np.random.default_rng(3)
    for sign in (1.0, -1.0):
        for _ in range(50):
            wl, wr = rand_state(rng, 0.1), rand_state(rng, 0.1)
            for w in (wl, wr):
                w[1 + dir] = sign * (3.0 + rng.uniform(0, 1))  # |u_n| > c everywhere (c <= sqrt(1.4 * 4))
            want = euler_flux(wl if sign > 0 else wr, dir)
            np.testing.assert_array_equal(O.nsc_hllc(wl, wr, dir), want)


@pytest.mark.parametrize("dir", [0, 1])
def test_hllc_resolves_contacts_and_shear_exactly(dir):
    """A stationary contact (density jump) with a tangential velocity jump at equal pressure: the exact flux is pressure
    only. HLLC keeps it exactly (HLL would smear it)."""
    rng = np.random.default_rng(4)
    for _ in range(50):
        p = rng.uniform(0.5, 2.0)
        wl = np.array([rng.uniform(0.5, 2), 0.0, 0.0, p])
        wr = np.array([rng.uniform(0.5, 2), 0.0, 0.0, p])
        wl[2 - dir], wr[2 - dir] = rng.uniform(-1, 1), rng.uniform(-1, 1)  # tangential velocities
        want = np.zeros(4)
        want[1 + dir] = p
        np.testing.assert_allclose(O.nsc_hllc(wl, wr, dir), want, rtol=1e-13, atol=1e-13)


def cons(w):
    r, u, v, p = w
    return np.array([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)])


@pytest.mark.parametrize("dir", [0, 1])
def test_hllc_star_states_satisfy_rankine_hugoniot(dir):
    """Subsonic faces: the flux is the star flux of one side K, F = F_K + S_K (U*_K - U_K). Recovering U*_K from it, the
    star state must carry the Euler-form flux with one star pressure, F = (rho* S*, rho* S*^2 + p*, rho* S* v_K,
    (E* + p*) S*), and p* must be the OTHER side's star pressure p_J + rho_J (S_J - u_J)(S* - u_J): the jump conditions
    across the outer wave and the contact (Davis speeds S_L = min(u_L - c_L, u_R - c_R), S_R = max(u_L + c_L, u_R + c_R))."""
    rng = np.random.default_rng(10)
    n, t = 1 + dir, 2 - dir
    for _ in range(200):
        wl, wr = rand_state(rng, 0.4), rand_state(rng, 0.4)
        F = O.nsc_hllc(wl, wr, dir)
        cl, cr = np.sqrt(G * wl[3] / wl[0]), np.sqrt(G * wr[3] / wr[0])
        S = {"L": min(wl[n] - cl, wr[n] - cr), "R": max(wl[n] + cl, wr[n] + cr)}
        assert S["L"] < 0 < S["R"]
        W = {"L": wl, "R": wr}
        star = {}
        for K in "LR":
            Us = cons(W[K]) + (F - euler_flux(W[K], dir)) / S[K]
            star[K] = (Us, Us[n] / Us[0])
        K = "L" if star["L"][1] >= 0 else "R"
        J = "R" if K == "L" else "L"
        Us, sm = star[K]
        assert (sm >= 0) == (K == "L")
        ps = F[n] - Us[0] * sm * sm
        want = np.zeros(4)
        want[0], want[n], want[t], want[3] = Us[0] * sm, Us[0] * sm * sm + ps, Us[0] * sm * W[K][t], (Us[3] + ps) * sm
        np.testing.assert_allclose(F, want, rtol=1e-10, atol=1e-12)
        wj = W[J]
        assert ps == pytest.approx(wj[3] + wj[0] * (S[J] - wj[n]) * (sm - wj[n]), rel=1e-10)


def test_hllc_rotation_and_mirror():
    rng = np.random.default_rng(5)
    for _ in range(100):
        wl, wr = rand_state(rng, 1.5), rand_state(rng, 1.5)
        f0 = O.nsc_hllc(wl, wr, 0)
        # rotation: dir 1 of the states with u and v swapped is dir 0 with the momenta swapped
        f1 = O.nsc_hllc(wl[[0, 2, 1, 3]], wr[[0, 2, 1, 3]], 1)
        np.testing.assert_allclose(f1[[0, 2, 1, 3]], f0, rtol=1e-12, atol=1e-13)
        # mirror x -> -x: the right state becomes the left one with u negated; mass, y-momentum and energy fluxes flip
        m = np.array([1.0, -1.0, 1.0, 1.0])
        fm = O.nsc_hllc(wr * m, wl * m, 0)
        np.testing.assert_allclose(fm, f0 * np.array([-1.0, 1.0, -1.0, -1.0]), rtol=1e-11, atol=1e-12)


# ---- viscous flux -----------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dir", [0, 1])
def test_viscous_flux_on_linear_fields(dir):
    """u = a x + b y, v = c x + d y, T = e x + f y (unit spacing): the two cells differ by one cell along the normal and
    carry the exact gradients, so the flux is the Stokes stress and the Fourier heat flux in closed form."""
    rng = np.random.default_rng(6)
    k = MU * G / ((G - 1) * PR)
    for _ in range(50):
        a, b, c, d, e, f = rng.normal(size=6)
        r0, x0, y0 = rng.uniform(0.5, 2), rng.normal(), rng.normal()
        xl, yl = x0, y0
        xr, yr = (x0 + 1, y0) if dir == 0 else (x0, y0 + 1)

        def prim(x, y):
            T = 2.0 + e * x + f * y
            return np.array([r0, a * x + b * y, c * x + d * y, r0 * T])

        wl, wr = prim(xl, yl), prim(xr, yr)
        g = np.array([a, b, c, d, e, f])
        F = O.nsc_visc(wl, wr, g, g, dir)
        dv = a + d
        txx, tyy, txy = MU * (2 * a - 2 / 3 * dv), MU * (2 * d - 2 / 3 * dv), MU * (b + c)
        uf, vf = 0.5 * (wl[1] + wr[1]), 0.5 * (wl[2] + wr[2])
        want = [0.0, txx, txy, uf * txx + vf * txy + k * e] if dir == 0 else [0.0, txy, tyy, uf * txy + vf * tyy + k * f]
        np.testing.assert_allclose(F, want, rtol=1e-10, atol=1e-12)


def test_cell_diagnostics_closed_form():
    rng = np.random.default_rng(7)
    for _ in range(50):
        W = rand_state(rng, 2.0)
        c = np.sqrt(G * W[3] / W[0])
        nu = MU / W[0] * max(4 / 3, G / PR)
        dt, ws = O.nsc_cell_diag(W)
        assert dt == pytest.approx(CFL / (abs(W[1]) + abs(W[2]) + 2 * c + 4 * nu), rel=1e-14)
        assert ws == pytest.approx(max(abs(W[1]), abs(W[2])) + c, rel=1e-14)


def test_primitives_round_trip():
    rng = np.random.default_rng(8)
    for _ in range(50):
        W = rand_state(rng, 2.0)
        r, u, v, p = W
        U = np.array([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)])
        np.testing.assert_allclose(O.nsc_prim(U), W, rtol=1e-13, atol=1e-14)


# ---- the residual -----------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("vl", [False, True])
def test_uniform_periodic_flow_is_exactly_zero(vl):
    U = O.nsc_init(48, 32, 0.0)
    R, d = O.nsc_residual(U, walls=False, vanleer=vl)
    assert np.all(R == 0.0)
    W = O.nsc_prim(U[:, 0, 0].astype(np.float64))
    dt, ws = O.nsc_cell_diag(W)
    assert d["dt"] == dt and d["ws"] == ws and np.all(d["sq"] == 0.0)


@pytest.mark.parametrize("vl", [False, True])
def test_periodic_conservation(vl):
    """Every face flux enters its two cells with opposite signs: over a periodic domain the sum of R vanishes."""
    U = O.nsc_init(64, 48)
    R, _ = O.nsc_residual(U, walls=False, vanleer=vl)
    for q in range(4):
        assert abs(R[q].sum()) <= 1e-12 * np.abs(R[q]).sum(), q


def test_periodic_translation_is_bitwise():
    U = O.nsc_init(64, 48)
    R, _ = O.nsc_residual(U, walls=False)
    Us = np.roll(U, (5, 17), axis=(1, 2))
    Rs, _ = O.nsc_residual(Us, walls=False)
    np.testing.assert_array_equal(Rs, np.roll(R, (5, 17), axis=(1, 2)))


@pytest.mark.parametrize("walls", [False, True])
def test_mirror_and_transpose_symmetries(walls):
    rng = np.random.default_rng(9)
    U = O.nsc_init(48, 48)
    U = (U * (1 + 0.01 * rng.standard_normal(U.shape))).astype(np.float32)  # no symmetry of its own
    R, _ = O.nsc_residual(U, walls=walls)
    scale = np.abs(R).max()
    # x -> -x (x is periodic under both boundary conditions): x-momentum changes sign
    Um = U[:, :, ::-1].copy()
    Um[1] *= -1
    Rm, _ = O.nsc_residual(Um, walls=walls)
    Rm = Rm[:, :, ::-1]
    Rm[1] *= -1
    np.testing.assert_allclose(Rm, R, rtol=0, atol=1e-12 * scale)
    if not walls:  # transpose on a square periodic domain: x <-> y, the momenta swap
        Ut = np.transpose(U, (0, 2, 1))[[0, 2, 1, 3]].copy()
        Rt, _ = O.nsc_residual(Ut, walls=False)
        np.testing.assert_allclose(np.transpose(Rt, (0, 2, 1))[[0, 2, 1, 3]], R, rtol=0, atol=1e-12 * scale)


def test_walls_are_the_mirrored_periodic_domain():
    """A no-slip adiabatic wall's ghost rows mirror the interior with both momenta negated: the same cells a periodic
    domain of twice the height sees when its upper half is that mirror image."""
    U = O.nsc_init(32, 24)
    R, d = O.nsc_residual(U, walls=True)
    up = U[:, ::-1, :].copy()
    up[1:3] *= -1
    Rd, _ = O.nsc_residual(np.concatenate([U, up], axis=1), walls=False)
    np.testing.assert_array_equal(Rd[:, :24, :], R)
    np.testing.assert_allclose(d["sq"], (R ** 2).sum(axis=(1, 2)), rtol=1e-12)
