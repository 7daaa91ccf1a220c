"""The oracle's pointwise LES models (les_real_calc.jl:74-114, VremanLESModel.jl,
ProductionViscosityLESModel.jl, silvismodel.jl, stablenonlinearmodel.jl) against
independent constructions from the velocity gradient G_ij = d_j u_i: S = (G + G^T)/2,
w = curl u, W = (G - G^T)/2.

Each closed form below is computed with numpy's matrix routines, not with the
oracle's own helpers, so a dropped term, a transposed operand or a wrong sign in
the oracle fails one of them.  The tensor conventions of the unvendored
FluidTensors package (AntiSymTen, Lie) are the oracle's documented choice; the
gradient-model identity is what pins them to the physics."""
import numpy as np
import pytest

from oracle import les_models as lm

N = 2000


def gradients(seed, n=N):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, 3, 3))
    G -= np.trace(G, axis1=1, axis2=2)[:, None, None] * np.eye(3) / 3  # incompressible: tr G = 0
    S = 0.5 * (G + G.transpose(0, 2, 1))
    W = 0.5 * (G - G.transpose(0, 2, 1))
    w = np.stack([G[:, 2, 1] - G[:, 1, 2], G[:, 0, 2] - G[:, 2, 0], G[:, 1, 0] - G[:, 0, 1]])
    s5 = np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2]])
    return G, S, W, w, s5


def full(t5):
    xx, xy, xz, yy, yz = t5
    return np.stack([np.stack([xx, xy, xz], -1), np.stack([xy, yy, yz], -1), np.stack([xz, yz, -xx - yy], -1)], -2)


def ddot(a, b):
    return np.einsum("nij,nij->n", a, b)


def test_conventions_rotation_tensor():
    """AntiSymTen(-w/2) is minus the rotation tensor, so S + AntiSymTen(-w/2) = G^T."""
    G, S, W, w, s5 = gradients(1)
    A = np.moveaxis(np.array(lm.cross_matrix([-0.5 * x for x in w])), -1, 0)
    assert np.allclose(A, -W, atol=1e-14)
    assert np.allclose(full(s5) + A, G.transpose(0, 2, 1), atol=1e-14)
    v = np.random.default_rng(2).standard_normal((3, N))
    M = np.moveaxis(np.array(lm.cross_matrix(list(w))), -1, 0)
    assert np.allclose(np.einsum("nij,jn->in", M, v), np.cross(w.T, v.T).T, atol=1e-13)


@pytest.mark.parametrize("c,d2", [(1.0, 1.0), (0.1, 0.37)])
def test_stable_nonlinear_is_the_gradient_model(c, d2):
    """W^2 - Lie(S, W) - S^2 with W = AntiSymTen(-w/2) equals -(G G^T), so t = c Delta^2 traceless(-G G^T)
    + 2 mnu S, mnu = -min(0, c Delta^2 (W^2:S - S^2:S) / (2 S:S))."""
    G, S, W, w, s5 = gradients(3)
    t = full(lm.stress("stableNonLinear", s5, w, c, d2))
    GG = -np.einsum("nik,njk->nij", G, G)
    T = GG - np.trace(GG, axis1=1, axis2=2)[:, None, None] * np.eye(3) / 3
    ss = ddot(S, S)
    x = c * d2 * (np.einsum("ni,nij,nj->n", w.T, S, w.T) / 4 - 3 * np.linalg.det(S)) / (2 * ss)
    mnu = -np.minimum(0.0, x)
    ref = c * d2 * T + 2 * mnu[:, None, None] * S
    assert np.allclose(t, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert (mnu > 0).any() and (mnu == 0).any()


@pytest.mark.parametrize("c,d2", [(0.4, 1.0), (0.1, 2.5)])
def test_production_viscosity_closed_form(c, d2):
    """W^2:S - S^2:S = w.S.w / 4 - 3 det S for traceless S (W^2 = (w w^T - |w|^2 I)/4, S^3 = tr(S^2)/2 S
    + det(S) I), so nu = max(0, c Delta^2 (w.S.w/4 - 3 det S) / (2 S:S)) and t = 2 nu S."""
    G, S, W, w, s5 = gradients(4)
    t = full(lm.stress("productionViscosity", s5, w, c, d2))
    ss = ddot(S, S)
    nu = np.maximum(0.0, c * d2 * (np.einsum("ni,nij,nj->n", w.T, S, w.T) / 4 - 3 * np.linalg.det(S)) / (2 * ss))
    ref = 2 * nu[:, None, None] * S
    assert np.allclose(t, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())
    assert (nu > 0).any() and (nu == 0).any()


@pytest.mark.parametrize("c,d2", [(0.032, 1.0), (0.05, 0.3)])
def test_vreman_closed_form(c, d2):
    """beta = Delta^2 G G^T, B_beta = ((tr beta)^2 - tr(beta^2)) / 2 (the sum of its principal 2x2 minors),
    nu = c sqrt(B_beta / G:G), t = 2 nu S; nu >= 0 and nu = 0 for a pure shear."""
    G, S, W, w, s5 = gradients(5)
    t = full(lm.stress("Vreman", s5, w, c, d2))
    beta = d2 * np.einsum("nik,njk->nij", G, G)
    tr = np.trace(beta, axis1=1, axis2=2)
    B = 0.5 * (tr * tr - np.einsum("nij,nji->n", beta, beta))
    nu = c * np.sqrt(B / ddot(G, G))
    ref = 2 * nu[:, None, None] * S
    assert np.allclose(t, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
    # u = (gamma y, 0, 0): G_xy = gamma only
    Gs = np.zeros((1, 3, 3))
    Gs[0, 0, 1] = 1.7
    Ss = 0.5 * (Gs + Gs.transpose(0, 2, 1))
    ws = np.array([[Gs[0, 2, 1] - Gs[0, 1, 2]], [Gs[0, 0, 2] - Gs[0, 2, 0]], [Gs[0, 1, 0] - Gs[0, 0, 1]]])
    s5s = np.array([[Ss[0, 0, 0]], [Ss[0, 0, 1]], [Ss[0, 0, 2]], [Ss[0, 1, 1]], [Ss[0, 1, 2]]])
    assert np.all(lm.stress("Vreman", s5s, ws, c, d2) == 0)


@pytest.mark.parametrize("c,cp,d2", [(np.sqrt(0.17), 5.0, 1.0), (0.3, 2.0, 0.5)])
def test_silvis_closed_form(c, cp, d2):
    """f = |S w| / (|S| |w|) in [0, 1], nu_t = c^2 Delta^2 f^3 |S| / 2, nu_p = cp Delta^2 f^4 / 4 and
    t = 2 nu_t S + nu_p (S W - W S) with W the rotation tensor (G - G^T)/2."""
    G, S, W, w, s5 = gradients(6)
    t = full(lm.stress("Silvis", s5, w, c, d2, cp=cp))
    Sw = np.einsum("nij,jn->ni", S, w)
    f = np.linalg.norm(Sw, axis=1) / (np.sqrt(ddot(S, S)) * np.linalg.norm(w, axis=0))
    assert np.all((f >= 0) & (f <= 1 + 1e-15))
    nut = c * c * d2 * f**3 * np.sqrt(ddot(S, S)) / 2
    nup = cp * d2 * f**4 / 4
    P = S @ W - W @ S
    ref = 2 * nut[:, None, None] * S + nup[:, None, None] * P
    assert np.allclose(t, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("smodel", lm.S_MODELS)
def test_sandp_adds_beta_p(smodel):
    """SandP: t = t_S + cb Delta^2 (S W - W S); with fakeSmagorinsky only the P term."""
    G, S, W, w, s5 = gradients(7)
    c, cb, d2 = 0.17, 0.09, 0.8
    t = full(lm.stress("SandP", s5, w, c, d2, cp=cb, smodel=smodel))
    bp = cb * d2 * (S @ W - W @ S)
    base = 0 if smodel == "fakeSmagorinsky" else full(lm.stress(smodel, s5, w, c, d2))
    assert np.allclose(t, base + bp, rtol=1e-13, atol=1e-14 * np.abs(t).max())


def test_symmetric_traceless_results():
    """Every model's t is symmetric and traceless, so the five stored components carry all of it:
    the full products are checked against the stored zz = -xx - yy."""
    G, S, W, w, s5 = gradients(8, 50)
    P = S @ W - W @ S
    assert np.allclose(P, P.transpose(0, 2, 1), atol=1e-14)
    assert np.allclose(np.trace(P, axis1=1, axis2=2), 0, atol=1e-14)
    T = W @ W - P - S @ S  # the stable model's tensor, before traceless(): symmetric
    assert np.allclose(T, T.transpose(0, 2, 1), atol=1e-14)


def test_defaults():
    assert lm.default_constants("Vreman") == (0.032, None)
    assert lm.default_constants("Silvis") == (np.sqrt(0.17), 5.0)
    assert lm.default_constants("SandP", "productionViscosity") == (0.1, 0.09)
    assert lm.default_constants("productionViscosity") == (0.4, None)
    d = (2 * np.pi / 32) / (2 / 3)
    assert lm.default_delta("SandP", "Smagorinsky", 1.0, 32, 2 / 3) == 2 * d
    assert lm.default_delta("Vreman", None, 1.0, 32, 2 / 3) == d


# ---- dynamicSmagorinsky (dynamic_smagorinsky.jl) ------------------------------------------------------
from oracle.ffs_oracle import Grid  # noqa: E402


def _dyn_inputs(grid, seed, shift=(0.0, 0.0, 0.0), amp=1.0):
    """A random dealiased solenoidal field (+ a constant velocity) with S, S real and u real."""
    from ffs.initial import random_spectrum
    u0 = random_spectrum(grid.NRX, seed=seed, k0=3.0)
    uh = np.array([grid.myfourier(amp * c) for c in u0])
    uh = grid.project(uh)
    for c in range(3):
        uh[c][0, 0, 0] = shift[c]
    kv = (grid.kx, grid.ky, grid.kz)
    pairs = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2))
    S = np.array([np.where(grid.RANGE, 0.5 * (1j * kv[a] * uh[b] + 1j * kv[b] * uh[a]), 0) for a, b in pairs])
    return uh, S, np.array([grid.irfft(c) for c in S]), np.array([grid.irfft(c) for c in uh])


def test_test_filter_on_one_mode():
    """The test filter multiplies cos(k.x) by exp(-td2 |k|^2 / 24) and keeps a constant."""
    g = Grid(16, 16, 16)
    x, y, z = g.coords()
    td2 = 0.7
    f = np.cos(3 * x + 2 * y - z) + 0.25
    out = lm.filter_real(g, f, td2)
    want = np.exp(-td2 * (9 + 4 + 1) / 24) * np.cos(3 * x + 2 * y - z) + 0.25
    assert np.allclose(out, want, atol=1e-13)
    # the Nyquist planes stay real and filtered: cos(8 x) -> exp(-td2 64/24) cos(8 x)
    assert np.allclose(lm.filter_real(g, np.cos(8 * x), td2), np.exp(-td2 * 64 / 24) * np.cos(8 * x), atol=1e-13)


def test_dynamic_coefficient_invariances():
    """The dynamic procedure is scale invariant (c unchanged, t x lambda^2 when u -> lambda u) and
    Galilean invariant (a constant velocity leaves L, M, c and t unchanged)."""
    g = Grid(16, 16, 16)
    d2, td2 = 0.09, 0.36
    uh, S, Sr, ur = _dyn_inputs(g, 5)
    t1, c1, _ = lm.dynamic_stress(g, uh, S, Sr, ur, d2, td2, 0.0)
    uh2, S2, Sr2, ur2 = _dyn_inputs(g, 5, amp=3.0)
    t2, c2, _ = lm.dynamic_stress(g, uh2, S2, Sr2, ur2, d2, td2, 0.0)
    assert np.allclose(c2, c1, atol=1e-12) and np.allclose(t2, 9 * t1, rtol=1e-10, atol=1e-12 * np.abs(t2).max())
    uh3, S3, Sr3, ur3 = _dyn_inputs(g, 5, shift=(0.7, -0.4, 1.1))
    assert np.allclose(ur3, ur + np.array([0.7, -0.4, 1.1])[:, None, None, None], atol=1e-13)
    t3, c3, _ = lm.dynamic_stress(g, uh3, S3, Sr3, ur3, d2, td2, 0.0)
    assert np.allclose(c3, c1, atol=1e-11) and np.allclose(t3, t1, atol=1e-11 * np.abs(t1).max())
    assert np.abs(c1).max() <= 1.0


def test_dynamic_stress_limits():
    """cmin floors the coefficient and 0.4 caps it: with cmin = 0.4 the model is Smagorinsky with
    c^2 = 0.4 (t = 0.8 Delta^2 |S| S) whatever the coefficient."""
    g = Grid(16, 16, 16)
    d2, td2 = 0.09, 0.36
    uh, S, Sr, ur = _dyn_inputs(g, 6)
    t, _, _ = lm.dynamic_stress(g, uh, S, Sr, ur, d2, td2, 0.4)
    smag = lm.stress("Smagorinsky", Sr, None, np.sqrt(0.4), d2)
    assert np.allclose(t, smag, rtol=1e-12, atol=1e-14 * np.abs(smag).max())


def test_piomelli_extrapolation():
    """calc_cstar!: c* = c + (dt/dt_prev)(c - c_prev), with dt = 0 read as 1, a zero previous dt read as
    dt (so r = 1), and c_prev = c at iteration 0 (c* = c)."""
    rng = np.random.default_rng(9)
    c, cm1 = rng.standard_normal(50), rng.standard_normal(50)
    cs, new = lm.piomelli_extrapolate(c, cm1, 0.02, 0.01, 5)
    assert np.allclose(cs, 3 * c - 2 * cm1, atol=1e-15) and np.array_equal(new, c)
    cs, _ = lm.piomelli_extrapolate(c, cm1, 0.02, 0.0, 5)
    assert np.allclose(cs, 2 * c - cm1, atol=1e-15)
    cs, _ = lm.piomelli_extrapolate(c, cm1, 0.0, 0.5, 5)
    assert np.allclose(cs, c + 2 * (c - cm1), atol=1e-15)
    cs, new = lm.piomelli_extrapolate(c, cm1, 0.02, 0.01, 0)
    assert np.array_equal(cs, c) and np.array_equal(new, c)


def test_piomelli_coefficient_invariances():
    """With cstar the coefficient is (sa + L):M/(M:M), M = 2 (td2 + Delta^2)|M| M: under u -> lambda u it is
    unchanged (sa and L scale as lambda^2, M as lambda^2), and it is Galilean invariant."""
    g = Grid(16, 16, 16)
    d2, td2 = 0.09, 0.36
    cstar = 0.02 + 0.01 * np.cos(g.coords()[0])
    uh, S, Sr, ur = _dyn_inputs(g, 7)
    t1, c1, _ = lm.dynamic_stress(g, uh, S, Sr, ur, d2, td2, -1.0, cstar=cstar)
    uh2, S2, Sr2, ur2 = _dyn_inputs(g, 7, amp=2.0)
    t2, c2, _ = lm.dynamic_stress(g, uh2, S2, Sr2, ur2, d2, td2, -1.0, cstar=cstar)
    assert np.allclose(c2, c1, atol=1e-12) and np.allclose(t2, 4 * t1, atol=1e-10 * np.abs(t2).max())
    uh3, S3, Sr3, ur3 = _dyn_inputs(g, 7, shift=(0.3, 0.9, -0.5))
    t3, c3, _ = lm.dynamic_stress(g, uh3, S3, Sr3, ur3, d2, td2, -1.0, cstar=cstar)
    assert np.allclose(c3, c1, atol=1e-11)


@pytest.mark.parametrize("pmode,smodel", [("dyn", None), ("dyn", "pio"), ("pio", None), ("pio", "pio")])
def test_dynamic_sandp_cp_invariances(pmode, smodel):
    """dynamicSandP / piomelliSandP: cp is scale invariant (L, P, Ph all scale as lambda^2) and Galilean
    invariant (the vorticity and the Leonard stress ignore a constant velocity), and the P part of t is
    clamp(cp, -0.8, 0.8) Delta^2 (S W - W S) with W the rotation tensor."""
    g = Grid(16, 16, 16)
    d2, td2 = 0.09, 0.36
    x = g.coords()[0]
    cstar = 0.02 + 0.01 * np.cos(x) if smodel == "pio" else None
    cpstar = 0.09 + 0.02 * np.sin(x) if pmode == "pio" else None

    def run(amp=1.0, shift=(0.0, 0.0, 0.0)):
        uh, S, Sr, ur = _dyn_inputs(g, 11, shift=shift, amp=amp)
        kv = (g.kx, g.ky, g.kz)
        w = np.array([1j * (kv[1] * uh[2] - kv[2] * uh[1]), 1j * (kv[2] * uh[0] - kv[0] * uh[2]),
                      1j * (kv[0] * uh[1] - kv[1] * uh[0])])
        w = np.where(g.RANGE, w, 0)
        wr = np.array([g.irfft(c) for c in w])
        return lm.dynamic_stress(g, uh, S, Sr, ur, d2, td2, -1.0, cstar=cstar, wh=w, wr=wr, pmode=pmode,
                                 cpstar=cpstar), Sr, wr

    (t1, c1, cp1), Sr, wr = run()
    (t2, c2, cp2), _, _ = run(amp=1.5)
    (t3, c3, cp3), _, _ = run(shift=(0.4, -0.2, 0.9))
    assert np.allclose(cp2, cp1, atol=1e-11) and np.allclose(cp3, cp1, atol=1e-11)
    assert np.allclose(c2, c1, atol=1e-11) and np.allclose(c3, c1, atol=1e-11)
    assert np.allclose(t2, 2.25 * t1, atol=1e-10 * np.abs(t2).max())
    # the P part from the rotation tensor W = (grad u - grad u^T)/2 = AntiSymTen(w/2)
    eddy = np.array([2 * np.maximum(-1.0, np.minimum(0.4, c1)) * d2 * np.sqrt(
        2 * lm.ddot(lm.sym_full(Sr), lm.sym_full(Sr))) * s for s in Sr])
    Sm = np.moveaxis(np.array(lm.sym_full(Sr)), (0, 1), (-2, -1))
    Wm = np.zeros_like(Sm)
    Wm[..., 0, 1], Wm[..., 0, 2], Wm[..., 1, 2] = -0.5 * wr[2], 0.5 * wr[1], -0.5 * wr[0]
    Wm[..., 1, 0], Wm[..., 2, 0], Wm[..., 2, 1] = 0.5 * wr[2], -0.5 * wr[1], 0.5 * wr[0]
    P = Sm @ Wm - Wm @ Sm
    f = np.clip(cp1, -0.8, 0.8) * d2
    want = np.array([P[..., 0, 0], P[..., 0, 1], P[..., 0, 2], P[..., 1, 1], P[..., 1, 2]]) * f
    assert np.allclose(t1 - eddy, want, atol=1e-12 * np.abs(t1).max())
