"""Pins of the oracle's synthetic NS solve (mode M1) by what the mathematics
fixes, independently of how the oracle computes it:

* one inner iteration's dU equals the dense LU-SGS solution
  dU = (D - U)^-1 D (D - L)^-1 Res built from the Euler flux Jacobians
  (complex-step derivatives of the textbook normal flux) and the oracle's own
  Lc, Lv and dt diagonals (du_LU_SGS_2D, src/sub_LU_SGS.f90:4-143);
* Lc is the spectral radius |u.n| + c of that Jacobian on the face-averaged
  normal, times the face length (comput_Lvc, src/sub_NS_singlegrid.f90:187-221);
* the NS residual of a parallel shear flow is zero in momentum for a linear
  profile, mu u'' vol for a quadratic one, its energy part the viscous
  heating mu u'^2 vol, and in a gas at rest with a temperature gradient the
  difference of two-point Fourier fluxes (Residual_FVM,
  src/sub_Residual.f90:125-202,262-338);
* the inner loop converges to the steady compressible Couette flow with
  mu = T^0.81/Re, second order in the mesh spacing (a boundary-value problem
  solved with scipy);
* the far-field ghost cells of the NS state average with their mirror cells
  to the kinetic wall state Swall (boundary_Farfield's GKUA branch,
  src/sub_boundary.f90:204-229), and periodic ghost cells copy the opposite
  interior cells (boundary_inlet, :248-298).

When the pins were written, each of these mistakes planted in a copy of the
oracle failed at least one of them: the backward sweep's i-term on the j-face
normal, or on si(i,j) instead of si(i+1,j); the forward sweep's Lv term
dropped; the j-face temperature difference taken across the wrong cells; the
face heat flux with the wrong sign; Lc without the sound speed; the viscous
face value from one cell; the local time step without the viscous radius.
The dense LU-SGS check also shows that a non-recursive backward pass (dU* of
the upper neighbours, as in the kinetic sweep) misses by a hundred times its
tolerance.
"""
import math

import numpy as np
import pytest

from helpers import GOLDEN, load_oracle, load_pkg


def _normal_flux(u, n1, n2, g):
    """F.n of the 2-D Euler equations for conserved u = (rho, rho u, rho v, E)."""
    rho, mx, my, e = u
    vn = (mx * n1 + my * n2) / rho
    p = (g - 1.0) * (e - 0.5 * (mx * mx + my * my) / rho)
    return np.array([rho * vn, mx * vn + p * n1, my * vn + p * n2, (e + p) * vn])


def _jacobian(u, n1, n2, g):
    """dF.n/dU by complex-step differentiation (exact to rounding)."""
    h = 1e-30
    a = np.zeros((4, 4))
    for k in range(4):
        up = np.array(u, dtype=complex)
        up[k] += 1j * h
        a[:, k] = _normal_flux(up, n1, n2, g).imag / h
    return a


def _skewed_case(n=9, nv=8):
    """A skewed, curved mesh (faces of every orientation), walls all round."""
    g = load_pkg().cases
    s = np.linspace(0.0, 1.0, n + 1)
    S, T = np.meshgrid(s, s, indexing="ij")
    xn = np.asfortranarray(S + 0.3 * T + 0.05 * np.sin(math.pi * T) + 0.04 * S * T)
    yn = np.asfortranarray(T + 0.1 * S ** 2 + 0.06 * S * T ** 2)
    nx = ny = n + 1
    W = load_pkg().capi.BC_WALL
    bcs = [g.BC(2, 1, nx, 1, 1, W), g.BC(3, nx, nx, 1, ny, W), g.BC(4, 1, nx, ny, ny, W), g.BC(1, 1, 1, 1, ny, W)]
    mesh = g.build_mesh(xn, yn, bcs)
    v, w = g.sinh_velocity_grid(nv, 6.0)
    const = g.constants_from_kn(0.2, ma=0.5, dt=0.5)
    return g.Case("skew", mesh, bcs, v, w, v.copy(), w.copy(), const, "rest")


def _smooth_state(case, amp):
    """U(4,-1:nx+1,-1:ny+1) of a moving gas with smooth perturbations of size amp."""
    c = case.const
    g, ma = c["gamma"], c["ma"]
    x1, y1 = case.mesh.x1y1()
    rho = 1.0 + amp * np.sin(2.1 * x1 + 0.7 * y1)
    u = 0.4 + amp * np.cos(1.3 * x1 - 2.0 * y1)
    v = -0.25 + amp * np.sin(0.9 * x1 + 1.7 * y1)
    t = 1.0 + amp * np.cos(1.1 * x1 + 0.4 * y1)
    p = rho * t / (g * ma * ma)
    out = np.zeros((4,) + x1.shape, order="F")
    out[0], out[1], out[2] = rho, rho * u, rho * v
    out[3] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v)
    return out


def test_ns_lusgs_equals_dense_solve():
    """dU of one inner iteration is the LU-SGS solution of the linearised
    system: the forward pass solves (D - L) dU* = Res, the backward pass
    (D - U) dU = D dU*, with D = vol/dt + Lci + Lcj + 2(Lvi + Lvj) and the
    off-diagonal blocks 0.5(A si + Lc) + Lv from the lower neighbours and
    -0.5(A s - Lc) + Lv from the upper ones, A the normal-flux Jacobian of the
    neighbour on this cell's face normal, s the face shared with the
    neighbour (src/sub_LU_SGS.f90:34-111).  comput_DFn is F(U+dU) - F(U), so
    with a residual of 1e-6 the two agree to the linearisation's 1e-6."""
    case = _skewed_case()
    g = case.const["gamma"]
    o = load_oracle().OracleSolver(case, mode=1)
    o.initialize()
    u0 = _smooth_state(case, 3e-3)
    o.upload(u=u0, u0=u0)
    assert o.ns_advance(1) > 0.0
    wk = o.ns_work()
    u1 = o.macro()["u"]
    m, nx, ny = case.mesh, case.nx, case.ny
    cells = [(i, j) for j in range(3, ny - 2) for i in range(3, nx - 2)]
    idx = {c: k for k, c in enumerate(cells)}
    nc = len(cells)
    D = np.zeros((4 * nc, 4 * nc))
    L = np.zeros_like(D)
    Up = np.zeros_like(D)
    res = np.zeros(4 * nc)

    def C(a, i, j):
        return a[i + 1, j + 1]

    def G(a, i, j):
        return a[i - 1, j - 1]

    eye = np.eye(4)
    for (i, j), k in idx.items():
        b = slice(4 * k, 4 * k + 4)
        alfa = (G(m.vol, i, j) / C(wk["dt"], i, j) + C(wk["lci"], i, j) + C(wk["lcj"], i, j)
                + 2.0 * (C(wk["lvi"], i, j) + C(wk["lvj"], i, j)))
        D[b, b] = alfa * eye
        res[b] = wk["res"][:, i + 1, j + 1]
        ni, nj = (G(m.ni1, i, j), G(m.ni2, i, j)), (G(m.nj1, i, j), G(m.nj2, i, j))
        for (ii, jj), nrm, s, lc, lv in (((i - 1, j), ni, G(m.si, i, j), "lci", "lvi"),
                                         ((i, j - 1), nj, G(m.sj, i, j), "lcj", "lvj")):
            if (ii, jj) in idx:
                a = _jacobian(u0[:, ii + 1, jj + 1], *nrm, g)
                blk = 0.5 * (a * s + C(wk[lc], ii, jj) * eye) + C(wk[lv], ii, jj) * eye
                L[b, slice(4 * idx[ii, jj], 4 * idx[ii, jj] + 4)] = blk
        for (ii, jj), nrm, s, lc, lv in (((i + 1, j), ni, G(m.si, i + 1, j), "lci", "lvi"),
                                         ((i, j + 1), nj, G(m.sj, i, j + 1), "lcj", "lvj")):
            if (ii, jj) in idx:
                a = _jacobian(u0[:, ii + 1, jj + 1], *nrm, g)
                blk = -0.5 * (a * s - C(wk[lc], ii, jj) * eye) + C(wk[lv], ii, jj) * eye
                Up[b, slice(4 * idx[ii, jj], 4 * idx[ii, jj] + 4)] = blk
    dstar = np.linalg.solve(D - L, res)
    du = np.linalg.solve(D - Up, D @ dstar)
    got = np.concatenate([wk["du"][:, i + 1, j + 1] for (i, j) in cells])
    scale = np.max(np.abs(du))
    assert scale > 1e-7
    assert np.max(np.abs(got - du)) <= 1e-5 * scale
    # the update: U + dU on cells 3..n-3, nothing else
    upd = np.concatenate([u1[:, i + 1, j + 1] - u0[:, i + 1, j + 1] for (i, j) in cells])
    assert np.max(np.abs(upd - got)) <= 1e-15 * np.max(np.abs(u0))
    mask = np.ones(u0.shape[1:], dtype=bool)
    for (i, j) in cells:
        mask[i + 1, j + 1] = False
    assert np.array_equal(u1[:, mask], u0[:, mask])
    # the non-recursive variant of the backward pass (reading dU* of the upper
    # neighbours, as the kinetic sweep does) differs visibly from both
    nonrec = dstar + np.linalg.solve(D, Up @ dstar)
    assert np.max(np.abs(nonrec - du)) > 1e-3 * scale


def test_ns_lc_is_the_flux_jacobian_spectral_radius():
    """Lci = max |eig(dF.n/dU)| * si on the face-averaged normal and length of
    the two i-faces of the cell (the sound speed taken along the unit normal);
    Lvi = gamma mu/(rho Pr) si^2/vol; dt = CFL vol / (sum of the four radii)
    clipped to [dtmin, dtmax] (comput_Lvc/comput_dt,
    src/sub_NS_singlegrid.f90:160-221)."""
    case = _skewed_case()
    c = case.const
    g = c["gamma"]
    o = load_oracle().OracleSolver(case, mode=1)
    o.initialize()
    u0 = _smooth_state(case, 0.05)
    o.upload(u=u0, u0=u0)
    o.ns_advance(1)
    wk = o.ns_work()
    m = case.mesh
    checked = 0
    for j in range(1, case.ny):
        for i in range(1, case.nx):
            si = 0.5 * (m.si[i - 1, j - 1] + m.si[i, j - 1])
            n1 = 0.5 * (m.ni1[i - 1, j - 1] + m.ni1[i, j - 1])
            n2 = 0.5 * (m.ni2[i - 1, j - 1] + m.ni2[i, j - 1])
            lam = np.max(np.abs(np.linalg.eigvals(_jacobian(u0[:, i + 1, j + 1], n1, n2, g))))
            # the averaged normal is not of unit length: the eigenvalues are
            # u.n +- c|n|, the reference's Lc takes c without |n|
            nn = math.hypot(n1, n2)
            rho, uu, vv = u0[0, i + 1, j + 1], u0[1, i + 1, j + 1] / u0[0, i + 1, j + 1], u0[2, i + 1, j + 1] / u0[0, i + 1, j + 1]
            cs = lam - abs(uu * n1 + vv * n2)
            np.testing.assert_allclose(wk["lci"][i + 1, j + 1], (abs(uu * n1 + vv * n2) + cs / nn) * si, rtol=1e-12)
            # viscous spectral radius gamma mu/(rho Pr) s^2/vol, and the local
            # step CFL vol / (Lci + Lcj + Lvi + Lvj) clipped to [dtmin, dtmax]
            t = g * c["ma"] ** 2 * (g - 1.0) * (u0[3, i + 1, j + 1] - 0.5 * rho * (uu * uu + vv * vv)) / rho
            mu = t ** float(np.float32(0.81)) / c["re"]
            vol = m.vol[i - 1, j - 1]
            np.testing.assert_allclose(wk["lvi"][i + 1, j + 1], g * mu / (rho * c["pr"]) * si * si / vol, rtol=1e-12)
            lsum = sum(wk[k][i + 1, j + 1] for k in ("lci", "lcj", "lvi", "lvj"))
            dt = min(max(c["cfl"] * vol / lsum, c["dtmin"]), c["dtmax"])
            np.testing.assert_allclose(wk["dt"][i + 1, j + 1], dt, rtol=1e-12)
            checked += 1
    assert checked == (case.nx - 1) * (case.ny - 1)


def _channel_case(ny_cells, nx_cells=8, kn=0.2, ma=0.5):
    """Uniform Cartesian channel, periodic in x, walls at y = 0, 1."""
    g = load_pkg().cases
    cap = load_pkg().capi
    nx, ny = nx_cells + 1, ny_cells + 1
    xs = np.linspace(0.0, nx_cells / ny_cells, nx)
    ys = np.linspace(0.0, 1.0, ny)
    xn, yn = np.meshgrid(xs, ys, indexing="ij")
    bcs = [g.BC(2, 1, nx, 1, 1, cap.BC_WALL), g.BC(3, nx, nx, 1, ny, cap.BC_PERIODIC),
           g.BC(4, 1, nx, ny, ny, cap.BC_WALL), g.BC(1, 1, 1, 1, ny, cap.BC_PERIODIC)]
    mesh = g.build_mesh(xn, yn, bcs)
    v, w = g.sinh_velocity_grid(8, 6.0)
    const = g.constants_from_kn(kn, ma=ma, dt=0.5)
    return g.Case(f"channel{ny_cells}", mesh, bcs, v, w, v.copy(), w.copy(), const, "rest")


def _shear_state(case, uy, ty=lambda y: 1.0 + 0.0 * y, p0=None):
    """U of a parallel flow u(y), v = 0, T(y) at uniform pressure."""
    c = case.const
    g, ma = c["gamma"], c["ma"]
    _, y1 = case.mesh.x1y1()
    t = ty(y1)
    p = (1.0 / (g * ma * ma)) if p0 is None else p0
    rho = g * ma * ma * p / t
    u = uy(y1)
    out = np.zeros((4,) + y1.shape, order="F")
    out[0], out[1] = rho, rho * u
    out[3] = p / (g - 1.0) + 0.5 * rho * u * u
    return out


@pytest.mark.parametrize("kind", ["linear", "quadratic", "heated"])
def test_ns_viscous_residual_of_shear_flows(kind):
    """Closed forms of the NS residual (src/sub_Residual.f90:125-202,262-338)
    for parallel flows u(y), v = 0 at uniform pressure on a uniform mesh:
    * u = a y at constant T: NND reconstructs linear data exactly, so the Van
      Leer flux is the physical flux (no jump at the faces) and cancels; the
      momentum residual is 0 and the energy residual the viscous heating
      mu a^2 vol;
    * u = b y^2 at constant T: the x-momentum residual is the stress
      difference mu u'' vol (the Van Leer dissipation of NND's constant jump
      b h^2 at every face cancels between the faces);
    * u = a y with a linear density at uniform pressure (so T and mu vary):
      all primitive fields are linear, the inviscid fluxes cancel, the
      x-momentum residual is a h (mu_f(j+1/2) - mu_f(j-1/2)) and the energy
      residual the difference of the faces' work u_f mu_f a plus two-point
      Fourier flux (cp/Pr) mu_f (T_j - T_{j-1})/h, times h, with mu_f the
      mean of the two cells' T^0.81/Re."""
    case = _channel_case(16)
    c = case.const
    o = load_oracle().OracleSolver(case, mode=1)
    o.initialize()
    a, b = 0.3, 0.8
    g, ma = c["gamma"], c["ma"]
    if kind == "linear":
        u0 = _shear_state(case, lambda y: a * y)
    elif kind == "quadratic":
        u0 = _shear_state(case, lambda y: b * y * y)
    else:
        p0 = 1.0 / (g * ma * ma)
        u0 = _shear_state(case, lambda y: a * y, lambda y: g * ma * ma * p0 / (1.0 + 0.4 * y), p0)
    o.upload(u=u0, u0=u0)
    o.ns_advance(1)
    res = o.ns_work()["res"]
    mu = 1.0 / c["re"]  # T = 1
    h = 1.0 / 16
    vol = h * h
    om = float(np.float32(0.81))

    def temp(j):
        d, vx = u0[0, 4, j + 1], u0[1, 4, j + 1] / u0[0, 4, j + 1]
        return g * ma * ma * (g - 1.0) * (u0[3, 4, j + 1] - 0.5 * d * vx * vx) / d

    for j in range(3, case.ny - 2):
        for i in range(3, case.nx - 2):
            r = res[:, i + 1, j + 1]
            assert abs(r[0]) < 1e-14 and abs(r[2]) < 1e-13, (i, j, r)
            if kind == "linear":
                assert abs(r[1]) < 1e-14
                np.testing.assert_allclose(r[3], mu * a * a * vol, rtol=1e-9)
            elif kind == "quadratic":
                np.testing.assert_allclose(r[1], mu * 2.0 * b * vol, rtol=1e-9)
            else:
                t = [temp(jj) for jj in (j - 1, j, j + 1)]
                mf = [0.5 * (t[0] ** om + t[1] ** om) / c["re"], 0.5 * (t[1] ** om + t[2] ** om) / c["re"]]
                yf = [(j - 1) * h, j * h]  # the faces j-1/2, j+1/2
                np.testing.assert_allclose(r[1], a * h * (mf[1] - mf[0]), rtol=1e-9)
                k = c["cp"] / c["pr"]
                e = [a * yf[q] * mf[q] * a + k * mf[q] * (t[q + 1] - t[q]) / h for q in range(2)]
                np.testing.assert_allclose(r[3], (e[1] - e[0]) * h, rtol=1e-9)


def _couette_bvp(uw, c):
    """Steady compressible Couette flow between walls at y = 0 (u = 0) and
    y = 1 (u = uw), both at T = 1, of the NS equations as the synthetic solve
    writes them: mu = T^0.81/Re (the REAL(4) 0.81), stress tau = mu u',
    heat flux (cp/Pr) mu T'; tau is constant and
    (tau u + (cp/Pr) mu T')' = 0."""
    from scipy.integrate import solve_bvp
    om = float(np.float32(0.81))
    re, cp, pr = c["re"], c["cp"], c["pr"]

    def f(y, z, p):
        tau = p[0]
        u, t, q = z  # q = (cp/Pr) mu T'
        mu = t ** om / re
        return np.vstack([tau / mu, q / ((cp / pr) * mu), -tau * tau / mu])

    def bc(za, zb, p):
        return np.array([za[0], za[1] - 1.0, zb[0] - uw, zb[1] - 1.0])

    y = np.linspace(0.0, 1.0, 201)
    z = np.vstack([uw * y, np.ones_like(y), np.zeros_like(y)])
    sol = solve_bvp(f, bc, y, z, p=[uw / re], tol=1e-10, max_nodes=100000)
    assert sol.success
    return sol


def test_inner_loop_converges_to_compressible_couette_second_order():
    """Cells 1-2 and n-2..n-1 of the synthetic state are held by the inner loop
    (it updates cells 3..n-3 only), so filling them with the exact Couette
    solution and iterating to a tiny residual gives the discrete steady
    solution of that Dirichlet problem.  Against the boundary-value problem
    its errors in velocity and in the viscous temperature rise fall at least
    like h^2 from 16 to 32 cells across the channel (by 5.5 and 8 here)."""
    uw = 0.6
    errs = []
    for n in (16, 32):
        case = _channel_case(n, nx_cells=8)
        c = case.const
        sol = _couette_bvp(uw, c)
        _, y1 = case.mesh.x1y1()
        yc = np.clip(y1, 0.0, 1.0)
        zu, zt = sol.sol(yc.ravel())[0].reshape(yc.shape), sol.sol(yc.ravel())[1].reshape(yc.shape)
        exact = _shear_state(case, lambda y: zu, lambda y: zt)
        start = _shear_state(case, lambda y: uw * np.clip(y, 0.0, 1.0))
        keep = np.ones(y1.shape, dtype=bool)
        keep[4:case.nx - 1, 4:case.ny - 1] = False  # cells 3..n-3 start from the guess
        start[:, keep] = exact[:, keep]
        o = load_oracle().OracleSolver(case, mode=1)
        o.initialize()
        o.upload(u=start, u0=start)
        r = 1.0
        for _ in range(200):
            r = o.ns_advance(50)
            if r < 1e-15:
                break
        assert 0.0 <= r < 1e-13, r
        u = o.macro()["u"]
        inner = (slice(4, case.nx - 1), slice(4, case.ny - 1))
        g, ma = c["gamma"], c["ma"]

        def vel_temp(a):
            d, vx = a[0][inner], a[1][inner] / a[0][inner]
            t = g * ma * ma * (g - 1.0) * (a[3][inner] - 0.5 * d * vx * vx) / d
            return vx, t

        (gu, gt), (eu, et) = vel_temp(u), vel_temp(exact)
        errs.append((np.max(np.abs(gu - eu)) / uw, np.max(np.abs(gt - et)) / np.max(np.abs(et - 1.0))))
    for k in range(2):  # velocity, temperature rise: at least second order
        assert errs[0][k] < 2e-2
        assert errs[0][k] / errs[1][k] > 3.0, (k, errs)


def test_ns_farfield_and_periodic_ghost_cells():
    """The NS ghost cells at initialisation (Boundary_condition_onemesh,
    src/sub_boundary.f90:10-39), checked by the property each fill must have:
    on far-field faces (boundary_Farfield with the kinetic wall state,
    :204-229) each ghost layer k and its mirror interior layer k average to
    Swall's density, velocity and temperature at the face node (T by linear
    extrapolation unless it would turn negative); on the periodic faces of a
    channel each ghost layer equals the interior layer one period away.  A
    Swall taken one node off, or the layers paired the other way round, fails
    the far-field check."""
    g = load_pkg()
    case = g.cases.shipped_case(GOLDEN / "cylinder", "cylinder40")
    c = case.const
    gam, ma = c["gamma"], c["ma"]
    sq = math.sqrt(2.0 / gam)
    o = load_oracle().OracleSolver(case)
    o.initialize()
    sw = o.macro()["swall"]
    # a non-uniform interior, so that every ghost cell has its own partner
    u = g.cases.ns_boundary_init(case, _smooth_state(case, 0.1), sw)

    def prim(i, j):
        d = u[0, i + 1, j + 1]
        uu, vv = u[1, i + 1, j + 1] / d, u[2, i + 1, j + 1] / d
        p = (u[3, i + 1, j + 1] - 0.5 * d * (uu * uu + vv * vv)) * (gam - 1.0)
        return np.array([d, uu, vv, gam * ma * ma * p / d])

    def check(shift_node=0, swap_layers=False):
        worst, off = 0.0, 0
        for b in case.bcs:
            nk = (b.jend - b.jst + 1) if b.face in (1, 3) else (b.iend - b.ist + 1)
            if b.code not in (g.capi.BC_WALL, g.capi.BC_FARFIELD):
                continue
            w = sw[off:off + 10 * nk].reshape((10, nk), order="F")
            off += 10 * nk
            if b.code != g.capi.BC_FARFIELD:
                continue
            for kk in range(nk - 1):  # cells along the face
                wk = w[:, min(kk + shift_node, nk - 1)]
                face = np.array([wk[0], wk[1] / ma * sq, wk[2] / ma * sq, wk[3]])
                for layer in (1, 2):
                    li = 3 - layer if swap_layers else layer
                    if b.face == 1:
                        ins, gh = (li, b.jst + kk), (1 - layer, b.jst + kk)
                    elif b.face == 3:
                        ins, gh = (b.iend - li, b.jst + kk), (b.iend - 1 + layer, b.jst + kk)
                    elif b.face == 2:
                        ins, gh = (b.ist + kk, li), (b.ist + kk, 1 - layer)
                    else:
                        ins, gh = (b.ist + kk, b.jend - li), (b.ist + kk, b.jend - 1 + layer)
                    a, q = prim(*ins), prim(*gh)
                    avg = 0.5 * (a + q)
                    t_ghost_ok = q[3] > 0 and abs(avg[3] - face[3]) <= 1e-12 * abs(face[3])
                    t_clamped = abs(q[3] - face[3]) <= 1e-12 * abs(face[3])
                    worst = max(worst, np.max(np.abs(avg[:3] - face[:3]) / (1.0 + np.abs(face[:3]))))
                    if not (t_ghost_ok or t_clamped):
                        worst = max(worst, 1.0)
        return worst

    assert any(b.code == g.capi.BC_FARFIELD for b in case.bcs)
    assert check() < 1e-12
    assert check(shift_node=1) > 1e-6
    assert check(swap_layers=True) > 1e-6
    # periodic faces of a channel: ghost layer = interior layer one period away
    ch = _channel_case(8, nx_cells=6)
    oc = load_oracle().OracleSolver(ch)
    oc.initialize()
    uc = _smooth_state(ch, 0.1)
    uc2 = g.cases.ns_boundary_init(ch, uc, oc.macro()["swall"])
    nx = ch.nx
    for j in range(1, ch.ny):
        for layer in (1, 2):
            np.testing.assert_array_equal(uc2[:, 1 - layer + 1, j + 1], uc[:, nx - layer + 1, j + 1])
            np.testing.assert_array_equal(uc2[:, nx - 1 + layer + 1, j + 1], uc[:, layer + 1, j + 1])


def test_mass_correct_scales_density_to_init_mass():
    """MassCorrect (src/GKUA_DETFP.f90:153-172): after the moments of a step
    every interior S_GKUA(1) is the uncorrected value times
    Init_mass/out_mass, so the interior mass equals Init_mass; Init_mass is
    the out_mass of the state when the correction is switched on."""
    from helpers import load_oracle, load_pkg
    g = load_pkg()
    case = g.cases.cavity_case(10, 8, kn=0.2, lid=0.15, dt=0.5)
    orc = load_oracle()
    a, b = orc.OracleSolver(case), orc.OracleSolver(case)
    for o in (a, b):
        o.initialize()
    m = a.macro()
    s = m["s"].copy()
    s[0] *= 1.3  # a state whose mass the next moments will not reproduce
    a.upload(s=s)
    vol = case.mesh.vol[:-1, :-1]
    want = float((s[0, 4:14, 4:14] * vol).sum())
    a.set_mass_correct(True)
    assert a.out_mass() == pytest.approx(want, rel=1e-14)
    a.step(2)
    b.step(2)
    assert a.out_mass() == pytest.approx(want, rel=1e-13)
    sb = b.macro()["s"][0, 4:14, 4:14]
    out_b = b.out_mass()
    assert out_b == pytest.approx(float((sb * vol).sum()), rel=1e-14)
    assert abs(out_b / want - 1.0) > 0.1  # the uncorrected mass is far from Init_mass
    # step 1 of two runs from the same state: the uncorrected run's density
    # times the ratio is the corrected run's, cell by cell
    a2, b2 = orc.OracleSolver(case), orc.OracleSolver(case)
    for o in (a2, b2):
        o.initialize()
        o.upload(s=s)
    a2.set_mass_correct(True)
    a2.step(1)
    b2.step(1)
    s_a = a2.macro()["s"][0, 4:14, 4:14]
    s_b = b2.macro()["s"][0, 4:14, 4:14]
    ratio = want / b2.out_mass()
    assert np.allclose(s_a, s_b * ratio, rtol=1e-14, atol=0)
    for o in (a, b, a2, b2):
        o.close()
