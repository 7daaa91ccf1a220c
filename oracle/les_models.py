"""Test infrastructure: CPU restatement of the pointwise LES stresses of calc_les!
(LESmodels/les_real_calc.jl:19-139) for the eddy-viscosity models other than
Smagorinsky.  Only tests/, smoke() and bench.py's cpu_baseline leg use it, as the
checker.

Inputs are real-space fields: S = symouter(ik, u) in its five stored components
(xx xy xz yy yz, zz = -xx - yy: the SymTrTen layout) and the vorticity w = ik x u
(3 components).  The result t is stored into the same five components.

The tensor algebra comes from the FluidTensors package, which is not vendored
with the reference.  The conventions used here, each a standard one:
  * AntiSymTen(v) is the cross-product matrix [v]x, AntiSymTen(v) . x = v x x.
    Then AntiSymTen(-w/2) = -(grad u - grad u^T)/2 and S + AntiSymTen(-w/2) is
    the transposed velocity gradient (grad u)^T.
  * Lie(a, b) = a . b - b . a;  square(a) = a . a;  a:b = sum_ij a_ij b_ij;
    traceless(a) = a - tr(a)/3 I;  a' is the transpose;  norm of a vector is its
    Euclidean length.
With them the stable non-linear model's tensor W^2 - Lie(S, W) - S^2 equals
-(grad u)(grad u)^T (Clark's gradient model), which tests/test_les_models.py
checks from the velocity gradient itself.  These conventions are parity
unpinned: no reference test or fixture fixes them.
"""
import numpy as np

MODELS = ("Smagorinsky", "Vreman", "Silvis", "productionViscosity", "stableNonLinear", "SandP")
S_MODELS = ("Smagorinsky", "Vreman", "productionViscosity", "fakeSmagorinsky")


def sym_full(s5):
    """The 3x3 tensor (as [i][j] arrays) of the five stored components, zz = -xx - yy."""
    xx, xy, xz, yy, yz = s5
    zz = -xx - yy
    return [[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]


def cross_matrix(v):
    """AntiSymTen(v): the matrix M with M . x = v x x, built from the cross products of the basis."""
    m = [[None] * 3 for _ in range(3)]
    for j in range(3):
        e = [np.zeros_like(v[0]) for _ in range(3)]
        e[j] = np.ones_like(v[0])
        col = (v[1] * e[2] - v[2] * e[1], v[2] * e[0] - v[0] * e[2], v[0] * e[1] - v[1] * e[0])
        for i in range(3):
            m[i][j] = col[i]
    return m


def mat(a, b):
    return [[sum(a[i][k] * b[k][j] for k in range(3)) for j in range(3)] for i in range(3)]


def add(a, b, fb=1.0):
    return [[a[i][j] + fb * b[i][j] for j in range(3)] for i in range(3)]


def ddot(a, b):
    return sum(a[i][j] * b[i][j] for i in range(3) for j in range(3))


def five(m):
    return np.array([m[0][0], m[0][1], m[0][2], m[1][1], m[1][2]])


def scale(a, f):
    return [[a[i][j] * f for j in range(3)] for i in range(3)]


def traceless(a):
    t = (a[0][0] + a[1][1] + a[2][2]) / 3
    return [[a[i][j] - (t if i == j else 0) for j in range(3)] for i in range(3)]


def vreman_nu(S, w, c, d2):
    """Vreman_eddy_viscosity (VremanLESModel.jl:1-7): alpha = S + AntiSymTen(-w/2), beta = Delta^2
    sym(alpha' alpha), B = b11 (b22 + b33) + b22 b33 - (b12^2 + b13^2 + b23^2) clipped at 0,
    nu = c sqrt(B / alpha:alpha)."""
    a = add(S, cross_matrix([-0.5 * x for x in w]))
    at = [[a[j][i] for j in range(3)] for i in range(3)]
    b = scale(mat(at, a), d2)
    B = b[0][0] * (b[1][1] + b[2][2]) + b[1][1] * b[2][2] - (b[0][1] ** 2 + b[0][2] ** 2 + b[1][2] ** 2)
    B = np.where(B < 0.0, 0.0, B)
    return c * np.sqrt(B / ddot(a, a))


def production_x(S, W, c, d2):
    """c Delta^2 (W^2:S - S^2:S) / (2 S:S) (ProductionViscosityLESModel.jl, stablenonlinearmodel.jl:21-27),
    and S:S; the models clip it at 0 (0 where S:S = 0)."""
    a = ddot(S, S)
    with np.errstate(divide="ignore", invalid="ignore"):
        x = c * d2 * (ddot(mat(W, W), S) - ddot(mat(S, S), S)) / (2 * a)
    return x, a


def stress(model, s5, w3, c, d2, cp=None, smodel=None):
    """calc_les! (les_real_calc.jl:74-114) for one model: returns the five stored components of t.
    model in MODELS; c the model constant (smagorinsky/vreman/silvis/production/nl constant); cp the
    Silvis P constant or SandP's pConstant; smodel SandP's sLesModel."""
    S = sym_full(s5)
    if model == "SandP":
        t = None if smodel == "fakeSmagorinsky" else stress(smodel, s5, w3, c, d2)
        P = add(mat(S, cross_matrix([0.5 * x for x in w3])), mat(cross_matrix([0.5 * x for x in w3]), S), -1.0)
        bp = five(scale(P, cp * d2))  # beta P, beta = cb Delta^2
        return bp if t is None else t + bp
    if model == "Smagorinsky":
        nu = c * c * d2 * np.sqrt(2 * ddot(S, S))  # alpha norm(S), norm(S) = sqrt(2 S:S)
        return five(scale(S, 2 * nu))
    if model == "Vreman":
        return five(scale(S, 2 * vreman_nu(S, w3, c, d2)))
    if model == "productionViscosity":
        x, a = production_x(S, cross_matrix([-0.5 * v for v in w3]), c, d2)
        nu = np.where(a == 0.0, 0.0, np.maximum(0.0, x))
        return five(scale(S, 2 * nu))
    if model == "Silvis":
        Sw = [sum(S[i][j] * w3[j] for j in range(3)) for i in range(3)]
        nsw = np.sqrt(Sw[0] ** 2 + Sw[1] ** 2 + Sw[2] ** 2)
        nw = np.sqrt(w3[0] ** 2 + w3[1] ** 2 + w3[2] ** 2)
        ss = ddot(S, S)
        f = nsw / (np.sqrt(ss) * nw)  # fvs (silvismodel.jl:24)
        nut = c * c * d2 * f * f * f * np.sqrt(ss) / 2
        nup = cp * d2 * f * f * f * f / 4
        half = cross_matrix([0.5 * x for x in w3])
        P = add(mat(S, half), mat(half, S), -1.0)  # Lie(S, AntiSymTen(w/2))
        return five(add(scale(S, 2 * nut), scale(P, nup)))
    if model == "stableNonLinear":
        W = cross_matrix([-0.5 * x for x in w3])
        lie = add(mat(S, W), mat(W, S), -1.0)
        T = traceless(add(add(mat(W, W), lie, -1.0), mat(S, S), -1.0))
        x, a = production_x(S, W, c, d2)
        mnu = np.where(a == 0.0, 0.0, -np.minimum(0.0, x))
        return five(add(scale(T, c * d2), scale(S, 2 * mnu)))
    raise ValueError(f"LES model {model!r}")


def default_constants(model, smodel=None):
    """(c, cp) defaults of les_types (les_types.jl:364-404); cp is None where the model has none."""
    base = {"Smagorinsky": 0.17, "Vreman": 0.032, "Silvis": np.sqrt(0.17), "productionViscosity": 0.4,
            "stableNonLinear": 0.1, "fakeSmagorinsky": 0.0}
    if model == "SandP":
        # the S model's own constant; SandP's production default is 0.1 (les_types.jl:398)
        c = 0.1 if smodel == "productionViscosity" else base[smodel]
        return c, 0.09
    return base[model], (5.0 if model == "Silvis" else None)


def default_delta(model, smodel, lx, nx, cutoff):
    """filterWidth default (les_types.jl:366-401): (2 pi lx / nx) / cutoff; SandP's Smagorinsky twice that."""
    d = (lx * 2 * np.pi / nx) / cutoff
    return 2 * d if (model == "SandP" and smodel == "Smagorinsky") else d


def gauss(grid, td2):
    """Gaussfilter(td2, k^2) = exp(-td2 k^2 / 24) on the half spectrum (utils.jl:220-225)."""
    return np.exp((-td2 / 24) * grid.K2)


def filter_real(grid, a, td2):
    """fullfourier! (rfft, every mode scaled by 1/N^3, none dealiased), x G, real! (utils.jl:108-112,
    dynamic_smagorinsky.jl:79-117, average_c! dynamic_piomelli.jl:13-26)."""
    return grid.irfft(grid.rfft(a) * grid.scale * gauss(grid, td2))


def piomelli_extrapolate(c, cm1, dt, dtm1, iteration):
    """calc_cstar! as realspace_dynamic_les_calculation! calls it (dynamic_smagorinsky.jl:9-21,
    dynamic_piomelli.jl:1-11): dt = 0 reads 1, a zero previous dt reads dt; at iteration 0 cm1 = c first.
    Returns (cstar, the new cm1) = (muladd(r, c - cm1, c), c), r = dt / dtm1."""
    dtf = 1.0 if dt == 0.0 else dt
    dtm1f = dtf if dtm1 == 0.0 else dtm1
    if iteration == 0:
        cm1 = c.copy()
    r = dtf / dtm1f
    return r * (c - cm1) + c, c.copy()


def dynamic_stress(grid, uh, S, Sr, ur, d2, td2, cmin, cstar=None, wh=None, wr=None, pmode=None, cpstar=None):
    """dynamicSmagorinsky with the averaged coefficient (avg = true): fourierspacep1! forms M = G S and
    uf = G u on the loop range (models.jl:98-108); realspace_dynamic_les_calculation! sa = Delta^2 |S| S
    and L = symouter(u, u) (dynamic_smagorinsky.jl:44-76), both test-filtered; the coefficient
    c = clamp(0.5 (L:M)/(M:M), -1, 1) with L = traceless(uf uf - L), M = (td2 + Delta^2) |M| M - sa
    (:131-200), averaged by one more filter; calc_les! t = 2 max(cmin, min(0.4, c)) Delta^2 |S| S
    (les_real_calc.jl:75-77).  |X| = sqrt(2 X:X).  With cstar (piomelliSmagorinsky, the extrapolated
    coefficient field) sa = 2 cstar Delta^2 |S| S and c = clamp((sa + L):M / (M:M), -1, 1) with
    M = 2 (td2 + Delta^2) |M| M (:62-66, :168-182).

    pmode "dyn" (dynamicSandP) or "pio" (piomelliSandP, with cpstar its extrapolated coefficient) adds
    the P part (les_types.jl:196-218, 302-337): P = [cpstar] Delta^2 Lie(S, AntiSymTen(w/2)), filtered;
    wf = G w; Ph = Lie(M, AntiSymTen(wf/2)); "dyn": Mp = (td2 + Delta^2) Ph - P, cp = clamp(L:Mp / Mp:Mp);
    "pio": Mp = (td2 + Delta^2) Ph, cp = clamp((tau + L):Mp / Mp:Mp) with tau = sa + P under
    piomelliSmagorinsky, P otherwise (and then c's tau is sa + P too); cp averaged; calc_les! adds
    clamp(cp, -0.8, 0.8) Delta^2 Lie(S, AntiSymTen(w/2)) (les_real_calc.jl:116-118).
    Returns (the five stored components of t, c, cp or None)."""
    G = gauss(grid, td2)
    Mr = np.array([grid.irfft(np.where(grid.RANGE, G * c, 0)) for c in S])
    ufr = np.array([grid.irfft(np.where(grid.RANGE, G * c, 0)) for c in uh])
    Sf = sym_full(Sr)
    nS = np.sqrt(2 * ddot(Sf, Sf))
    if cstar is None:
        sa = np.array([filter_real(grid, d2 * nS * x, td2) for x in Sr])
    else:
        sa = np.array([filter_real(grid, 2 * cstar * d2 * nS * x, td2) for x in Sr])
    pairs = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))
    Lf = {ab: filter_real(grid, ur[ab[0]] * ur[ab[1]], td2) for ab in pairs}
    L = [[ufr[i] * ufr[j] - Lf[(min(i, j), max(i, j))] for j in range(3)] for i in range(3)]
    L = traceless(L)
    Mh = sym_full(Mr)
    nM = np.sqrt(2 * ddot(Mh, Mh))
    pa = None
    if pmode is not None:
        half = cross_matrix([0.5 * x for x in wr])
        lie = add(mat(Sf, half), mat(half, Sf), -1.0)  # Lie(S, AntiSymTen(w/2))
        fac = cpstar * d2 if pmode == "pio" else d2
        pa = np.array([filter_real(grid, x, td2) for x in five(scale(lie, fac))])
    with np.errstate(divide="ignore", invalid="ignore"):
        if cstar is None:
            M = add(scale(Mh, (td2 + d2) * nM), sym_full(sa), -1.0)
            c = np.clip(0.5 * (ddot(L, M) / ddot(M, M)), -1.0, 1.0)
        else:
            tau = sym_full(sa + pa) if pmode == "pio" else sym_full(sa)
            M = scale(Mh, 2 * (td2 + d2) * nM)
            Mn = scale(M, 1.0 / ddot(M, M))
            c = np.clip(ddot(add(tau, L), Mn), -1.0, 1.0)
        cp = None
        if pmode is not None:
            wfr = np.array([grid.irfft(np.where(grid.RANGE, G * x, 0)) for x in wh])
            hf = cross_matrix([0.5 * x for x in wfr])
            Ph = add(mat(Mh, hf), mat(hf, Mh), -1.0)  # Lie(Sh, AntiSymTen(wf/2))
            if pmode == "dyn":
                Mp = add(scale(Ph, td2 + d2), sym_full(pa), -1.0)
                cp = np.clip(ddot(L, Mp) / ddot(Mp, Mp), -1.0, 1.0)
            else:
                tau = sym_full(sa + pa) if cstar is not None else sym_full(pa)
                Mp = scale(Ph, td2 + d2)
                Mpn = scale(Mp, 1.0 / ddot(Mp, Mp))
                cp = np.clip(ddot(add(tau, L), Mpn), -1.0, 1.0)
    c = filter_real(grid, c, td2)
    nut = np.maximum(cmin, np.minimum(0.4, c)) * d2 * nS
    t = np.array([2 * nut * x for x in Sr])
    if cp is not None:
        cp = filter_real(grid, cp, td2)
        t = t + five(scale(lie, np.maximum(-0.8, np.minimum(0.8, cp)) * d2))
    return t, c, cp


def eddy_nu(model, s5, w3, c, d2, smodel=None):
    """The eddy viscosity nu_t that calc_les! leaves for the scalars' LES fluxes (les_real_calc.jl:74-91,
    120-136): Smagorinsky alpha |S| (alpha = c c Delta^2), Vreman's, productionViscosity's, Silvis's nu_t
    (not its P coefficient), stableNonLinear's nu (not mnu); SandP its S model's, 0 for fakeSmagorinsky."""
    S = sym_full(s5)
    if model == "SandP":
        return 0.0 * s5[0] if smodel == "fakeSmagorinsky" else eddy_nu(smodel, s5, w3, c, d2)
    if model == "Smagorinsky":
        return (c * c * d2) * np.sqrt(2 * ddot(S, S))
    if model == "Vreman":
        return vreman_nu(S, w3, c, d2)
    if model in ("productionViscosity", "stableNonLinear"):
        x, a = production_x(S, cross_matrix([-0.5 * v for v in w3]), c, d2)
        return np.where(a == 0.0, 0.0, np.maximum(0.0, x))
    if model == "Silvis":
        Sw = [sum(S[i][j] * w3[j] for j in range(3)) for i in range(3)]
        ss = ddot(S, S)
        f = np.sqrt(Sw[0] ** 2 + Sw[1] ** 2 + Sw[2] ** 2) / (np.sqrt(ss) * np.sqrt(w3[0] ** 2 + w3[1] ** 2 + w3[2] ** 2))
        return c * c * d2 * f * f * f * np.sqrt(ss) / 2
    raise ValueError(f"LES model {model!r} has no pointwise eddy viscosity")


def scalar_les_flux(nut, grad, w3, cv):
    """EddyDiffusion nu_t grad s, plus (c Delta^2 / 2) w x grad s for VorticityDiffusion (cv = c Delta^2,
    None for EddyDiffusion) (les_real_calc.jl:120-136)."""
    f = [nut * g for g in grad]
    if cv is not None:
        cr = (w3[1] * grad[2] - w3[2] * grad[1], w3[2] * grad[0] - w3[0] * grad[2], w3[0] * grad[1] - w3[1] * grad[0])
        f = [f[i] + (cv / 2) * cr[i] for i in range(3)]
    return np.array(f)
