"""Point-particle source of the s = -2 Teukolsky equation, derived symbolically.

Test infrastructure: the independent derivation that pins the particle-source
coefficients of TD_GPU_code/HighOrder/kern.cu:770-2683 (`sourced`).  It is
redone from the published equations rather than taken from the reference's
generated code:

* Teukolsky (1973), ApJ 185, 635, eq. (2.15): the s = -2 source
      T_4 = (D + 3g - g* + 4mu + mu*) [ (db - 2 tau* + 2 al) T_nmb - (D + 2g - 2g* + mu*) T_mbmb ]
          + (db - tau* + beta* + 3 al + 4 pi) [ (D + 2g + 2mu*) T_nmb - (db - tau* + 2 beta* + 2 al) T_nn ]
  with D = n^a d_a, db = mbar^a d_a, Kinnersley tetrad and spin
  coefficients (Teukolsky 1973 eqs. 4.4-4.5), signature (+,-,-,-).
* Point particle: T^{ab} = mu u^a u^b delta^3(x - z(t)) / (Sigma sin(theta) u^t),
  with the 4-velocity represented as the reference does (kern.cu:1258-1300):
  covariant u_t = E, u_phi = -L_z, u_r = -(Sigma/Delta) dr/dt u^t,
  u_theta = -Sigma dtheta/dt u^t and U = Sigma u^t = E[(r^2+a^2)^2/Delta -
  a^2 sin^2] + a L_z [1 - (r^2+a^2)/Delta] (kern.cu:1260); E and L_z are
  constants of the step; third time derivatives of r, theta are zero
  (body.cu:124 passes 0).
* Azimuthal mode: delta(phi - phi_p(t)) -> e^{-i m phi_p}/pi and d_phi -> i m,
  the normalisation of kern.cu:1270-1271.  (The reference's generated T_4 as
  a whole is half of the result: teuk_t4.h applies that factor.)

The operators act on the delta functions through the particle's motion,
d_t delta(r - r_p(t)) = -r_p' delta'(r - r_p), and the result is reduced to
the unique canonical form
      T_4 = sum_{i+j<=2} c_ij(t) delta^(i)(r - r_p) delta^(j)(theta - theta_p) e^{-i m phi_p}/pi
with every smooth factor evaluated on the worldline
(f(r) delta^(n)(r - r_p) = sum_k (-1)^k C(n,k) f^(k)(r_p) delta^(n-k)(r - r_p)).
The reference evaluates the same canonical form (its coefficients are
functions of rp, tp only and multiply DelR, DelR1, DelR2 x DelTH, DelTH1,
DelTH2; kern.cu:1211-1227, 2597-2662).

To keep the result manageable the field-point functions (tetrad components,
spin-coefficient combinations, 1/(Sigma sin theta)) stay abstract during the
derivation: the coefficients come out as polynomials in their values and
first and second (r, theta) derivatives at the particle, named
<f>_<i>_<j>; `field_jets()` gives each of those as an explicit expression.
"""
import functools

import sympy as sp

t = sp.Symbol("t", real=True)
r, th = sp.symbols("r theta", real=True)
M, a, E, Lz, mu, m = sp.symbols("mass a E lz nmu mm", real=True)
I = sp.I
SQ2 = sp.sqrt(2)

# particle worldline as functions of t
Rt = sp.Function("R", real=True)(t)
Tt = sp.Function("TH", real=True)(t)
Pt = sp.Function("PH", real=True)(t)
Ut = sp.Function("U", real=True)(t)       # Sigma u^t along the worldline

# symbols the coefficients are finally expressed in (kern.cu:770-774 names)
rp, tp, phip = sp.symbols("rp tp phip", real=True)
drdt, d2rdt2, dthdt, d2thdt2, dphidt, d2phidt2 = sp.symbols(
    "drdt d2rdt2 dthdt d2thdt2 dphidt d2phidt2", real=True)
U0, U1, U2 = sp.symbols("U0 U1 U2", real=True)


class Dist:
    """sum_{(i,j)} G_ij(t, r, theta) delta^(i)(r - R(t)) delta^(j)(theta - TH(t)) e^{-i m PH(t)}/pi."""

    def __init__(self, terms=None):
        self.t = {k: v for k, v in (terms or {}).items() if v != 0}

    def __add__(self, o):
        out = dict(self.t)
        for k, v in o.t.items():
            out[k] = out.get(k, 0) + v
        return Dist(out)

    def __sub__(self, o):
        return self + o.scale(-1)

    def scale(self, f):
        return Dist({k: sp.expand(f * v) for k, v in self.t.items()})

    def d_r(self):
        out = {}
        for (i, j), g in self.t.items():
            out[(i, j)] = out.get((i, j), 0) + sp.diff(g, r)
            out[(i + 1, j)] = out.get((i + 1, j), 0) + g
        return Dist(out)

    def d_th(self):
        out = {}
        for (i, j), g in self.t.items():
            out[(i, j)] = out.get((i, j), 0) + sp.diff(g, th)
            out[(i, j + 1)] = out.get((i, j + 1), 0) + g
        return Dist(out)

    def d_t(self):
        rdot, tdot, pdot = sp.diff(Rt, t), sp.diff(Tt, t), sp.diff(Pt, t)
        out = {}
        for (i, j), g in self.t.items():
            out[(i, j)] = out.get((i, j), 0) + sp.diff(g, t) - I * m * pdot * g
            out[(i + 1, j)] = out.get((i + 1, j), 0) - rdot * g
            out[(i, j + 1)] = out.get((i, j + 1), 0) - tdot * g
        return Dist(out)

    def d_ph(self):
        return self.scale(I * m)


# ---- explicit field-point functions (Teukolsky 1973 eqs. 4.4-4.5) ---------
def _explicit():
    Sig = r**2 + a**2 * sp.cos(th)**2
    Dlt = r**2 - 2 * M * r + a**2
    rho = -1 / (r - I * a * sp.cos(th))
    rhob = -1 / (r + I * a * sp.cos(th))
    zb = SQ2 * (r - I * a * sp.cos(th))
    betab = -rho * sp.cos(th) / sp.sin(th) / (2 * SQ2)
    pi_ = I * a * rho**2 * sp.sin(th) / SQ2
    taub = I * a * rho * rhob * sp.sin(th) / SQ2
    mu_ = rho**2 * rhob * Dlt / 2
    mub = rhob**2 * rho * Dlt / 2
    gam = mu_ + rho * rhob * (r - M) / 2
    gamb = mub + rho * rhob * (r - M) / 2
    alp = pi_ - betab
    return {
        # tetrad components used by D = n^a d_a and db = mbar^a d_a
        "nT": (r**2 + a**2) / (2 * Sig), "nR": -Dlt / (2 * Sig), "nP": a / (2 * Sig),
        "bT": -I * a * sp.sin(th) / zb, "bH": 1 / zb, "bP": -I / (sp.sin(th) * zb),
        # 1/(Sigma sin theta) of the point-particle stress tensor
        "w": 1 / (Sig * sp.sin(th)),
        # spin-coefficient combinations of eq. (2.15)
        "s1": -2 * taub + 2 * alp,
        "s2": 2 * gam - 2 * gamb + mub,
        "s3": 2 * gam + 2 * mub,
        "s4": -taub + 2 * betab + 2 * alp,
        "s5": 3 * gam - gamb + 4 * mu_ + mub,
        "s6": -taub + betab + 3 * alp + 4 * pi_,
    }


FIELDS = list(_explicit())
F = {name: sp.Function(name)(r, th) for name in FIELDS}


def D_op(X):
    """n^a d_a with coordinates (t, r, theta, phi)."""
    return X.d_t().scale(F["nT"]) + X.d_r().scale(F["nR"]) + X.d_ph().scale(F["nP"])


def db_op(X):
    """mbar^a d_a."""
    return X.d_t().scale(F["bT"]) + X.d_th().scale(F["bH"]) + X.d_ph().scale(F["bP"])


def particle_projections():
    """F_nn, F_nmb, F_mbmb: T_ab = F_ab delta^3 with delta(phi) -> e^{-im phi}/pi."""
    Sp = Rt**2 + a**2 * sp.cos(Tt)**2
    Dp = Rt**2 - 2 * M * Rt + a**2
    # covariant velocity, signature (+,-,-,-), Sigma u^t = U
    u_t, u_r, u_h, u_p = E, -sp.diff(Rt, t) * Ut / Dp, -sp.diff(Tt, t) * Ut, -Lz
    nu = F["nT"] * u_t + F["nR"] * u_r + F["nP"] * u_p
    mbu = F["bT"] * u_t + F["bH"] * u_h + F["bP"] * u_p
    norm = mu * Sp / Ut * F["w"]
    return norm * nu * nu, norm * nu * mbu, norm * mbu * mbu


@functools.lru_cache(maxsize=None)
def t4_dist():
    Fnn, Fnm, Fmm = particle_projections()
    Tnn, Tnm, Tmm = Dist({(0, 0): Fnn}), Dist({(0, 0): Fnm}), Dist({(0, 0): Fmm})
    A = (db_op(Tnm) + Tnm.scale(F["s1"])) - (D_op(Tmm) + Tmm.scale(F["s2"]))
    B = (D_op(Tnm) + Tnm.scale(F["s3"])) - (db_op(Tnn) + Tnn.scale(F["s4"]))
    return (D_op(A) + A.scale(F["s5"])) + (db_op(B) + B.scale(F["s6"]))


def jet_symbol(name, i, j):
    return sp.Symbol(f"{name}_{i}_{j}")


def _worldline_subs(expr):
    """Name worldline derivatives (third ones are 0) and field-point jets."""
    rep = {}
    for d in expr.atoms(sp.Derivative):
        f = d.expr
        if isinstance(f, sp.core.function.AppliedUndef) and f.func.__name__ in FIELDS:
            i = sum(c for v, c in d.variable_count if v == r)
            j = sum(c for v, c in d.variable_count if v == th)
            rep[d] = jet_symbol(f.func.__name__, i, j)
    expr = expr.xreplace(rep)
    expr = expr.xreplace({F[n]: jet_symbol(n, 0, 0) for n in FIELDS})
    R1, R2, R3 = [sp.diff(Rt, t, k) for k in (1, 2, 3)]
    T1, T2, T3 = [sp.diff(Tt, t, k) for k in (1, 2, 3)]
    P1, P2 = [sp.diff(Pt, t, k) for k in (1, 2)]
    V1, V2 = [sp.diff(Ut, t, k) for k in (1, 2)]
    rep2 = [(V2, U2), (V1, U1), (Ut, U0), (R3, 0), (T3, 0), (R2, d2rdt2), (T2, d2thdt2), (P2, d2phidt2),
            (R1, drdt), (T1, dthdt), (P1, dphidt), (Rt, rp), (Tt, tp), (Pt, phip)]
    for old, new in rep2:
        expr = expr.subs(old, new)
    return expr


@functools.lru_cache(maxsize=None)
def derive():
    """Canonical coefficients {(i, j): c_ij} of T_4 (without e^{-i m phi_p}/pi)."""
    T4 = t4_dist()
    canon = {}
    for (i, j), g in T4.t.items():
        for kr in range(i + 1):
            gr = sp.diff(g, r, kr) if kr else g
            for kt in range(j + 1):
                gg = sp.diff(gr, th, kt) if kt else gr
                coef = (-1) ** (kr + kt) * sp.binomial(i, kr) * sp.binomial(j, kt) * gg
                key = (i - kr, j - kt)
                canon[key] = canon.get(key, 0) + coef
    return {key: sp.expand(_worldline_subs(c)) for key, c in canon.items()}


def used_jets(coeffs):
    names = set()
    for c in coeffs.values():
        for s in c.free_symbols:
            parts = s.name.rsplit("_", 2)
            if len(parts) == 3 and parts[0] in FIELDS:
                names.add((parts[0], int(parts[1]), int(parts[2])))
    return sorted(names)


def field_jets(jets):
    """Explicit expressions (in rp, tp, a, mass) of the jet symbols."""
    ex = _explicit()
    out = {}
    for name, i, j in jets:
        e = ex[name]
        if i:
            e = sp.diff(e, r, i)
        if j:
            e = sp.diff(e, th, j)
        out[jet_symbol(name, i, j)] = e.subs({r: rp, th: tp})
    return out


def u_series():
    """(U, dU/dt, d^2U/dt^2) on the worldline."""
    Dp = Rt**2 - 2 * M * Rt + a**2
    U = E * ((Rt**2 + a**2)**2 / Dp - a**2 * sp.sin(Tt)**2) + a * Lz * (1 - (Rt**2 + a**2) / Dp)
    return tuple(_worldline_subs(sp.diff(U, t, k)) for k in range(3))
