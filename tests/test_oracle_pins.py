"""Pins of the oracle's He, energy and deposit terms by checks that do not
go through its own formulas: the reference's void rate table (every rate
zero, src/atomic-rate.c:58-60 synthesizes Brem), collisional equilibria and
cooling built from the columns of the rate file itself, single-particle
deposits from Beer-Lambert absorption shared by opacity, the limits of the
secondary-ionization fit, and the He fixture against LSODA measured from a
tight solution."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

from conftest import GOLDEN

HUI = os.path.join(GOLDEN, "atomic_rates_Hui.txt")
VOID = os.path.join(GOLDEN, "atomic_rates_void.txt")


def rate_file(path):
    """columns of an atomic_rates_*.txt file by header name (the first
    column, 'Log T', as logT); values as printed, CGS"""
    with open(path) as f:
        lines = f.read().splitlines()
    names = ["logT"] + lines[1].split()[2:]
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[3:] if ln.strip()])
    return {n: rows[:, i] for i, n in enumerate(names)}


def one_step(**kw):
    s = np.zeros(1, po.STEP_DTYPE)
    for k, v in kw.items():
        s[k] = v
    return s


def cgs_rate(u):
    return u.U_CM ** 3 / u.U_SEC


def cgs_cool(u):
    return u.U_ERG * u.U_CM ** 3 / u.U_SEC


# ------------------------------------------------------------- void table
@pytest.mark.parametrize("ots", [True, False])
def test_void_table_step_removes_exactly_the_deposit(ots):
    """with every rate zero the fractions obey dx/dt = -yGdep/time only
    (src/solve.c:84-104), so one step gives x = x0 - yGdep exactly (up to
    rounding), He lines included; a neutral particle (no free electrons,
    nothing to cool) gains exactly its heat"""
    u = po.units(hmf=0.76)
    orc = po.Oracle(u, po.flags(isothermal=False, h_only=False, on_the_spot=ots), rates=VOID)
    cols = rate_file(VOID)
    assert cols["logT"].tolist() == [0.0, 9.0] and all(
        (v == 0).all() for k, v in cols.items() if k != "logT")
    nH = 1e-3 / u.U_CM ** 3
    s = one_step(lambdaH=0.7, lambdaHeI=0.5, lambdaHeII=0.3, yGdepHI=0.2, yGdepHeI=0.1,
                 yGdepHeII=0.05, ie=orc.Tye2ie(1e4, 0.5), heat=0.0, time=3.0, nH=nH, T=1e4)
    o, bad = orc.step_evolve(s)
    assert bad == 0
    assert o["lambdaH"][0] == pytest.approx(0.5, abs=1e-14)
    assert o["lambdaHeI"][0] == pytest.approx(0.4, abs=1e-14)
    assert o["lambdaHeII"][0] == pytest.approx(0.25, abs=1e-14)
    if not ots:
        # alpha_A - alpha_B = 0: no recombination photons accumulate
        assert o["yGrecHII"][0] == 0 and o["yGrecHeII"][0] == 0 and o["yGrecHeIII"][0] == 0
    # neutral: ye = 0, so L = 0 (src/solve.c:113-117) and ie = ie0 + heat
    ie0 = orc.Tye2ie(1e4, 0.0)
    heat = 0.37 * ie0
    s = one_step(lambdaH=1.0, lambdaHeI=1.0, lambdaHeII=0.0, ie=ie0, heat=heat, time=3.0, nH=nH,
                 T=1e4)
    o, bad = orc.step_evolve(s)
    assert bad == 0
    assert o["ie"][0] == pytest.approx(ie0 + heat, rel=1e-14)
    assert o["lambdaH"][0] == 1.0 and o["lambdaHeI"][0] == 1.0


def test_void_table_brem_is_the_synthesized_column():
    """the void file has no Brem column: ar_init fills it from
    1.42e-27 * 1.5 * sqrt(v) evaluated at the table's x column, which is
    log10 T (src/atomic-rate.c:58-60, tabfun_setup_col), and interpolates
    linearly between its two rows; Compton is 0"""
    u = po.units(hmf=1.0)
    orc = po.Oracle(u, po.flags(), rates=VOID)
    for logT in (0.0, 2.0, 4.0, 4.5, 8.9):
        expect = logT / 9.0 * 1.42e-27 * 1.5 * np.sqrt(9.0) * cgs_cool(u)
        assert orc.ar_get(13, logT) == pytest.approx(expect, rel=1e-12)
        assert orc.ar_get(14, logT) == 0.0
        for col in range(13):
            assert orc.ar_get(col, logT) == 0.0


def test_void_table_cooling_is_bremsstrahlung():
    """ionized hydrogen, no photo-ionization, no rates: the only energy loss
    is the synthesized Brem column, L = C_H_PER_MASS * beta * nH * ye * xHII;
    over a short step ie drops by L * time"""
    u = po.units(hmf=1.0)
    orc = po.Oracle(u, po.flags(isothermal=False, h_only=True), rates=VOID)
    nH = 1.0 / u.U_CM ** 3
    xHII = 0.6
    ye = xHII
    T = 1e4
    ie0 = orc.Tye2ie(T, ye)
    beta = np.log10(T) / 9.0 * 1.42e-27 * 1.5 * 3.0 * cgs_cool(u) * nH
    L = u.C_H_PER_MASS * beta * ye * xHII
    time = 1e-6 * ie0 / L
    s = one_step(lambdaH=1 - xHII, lambdaHeI=1.0, ie=ie0, time=time, nH=nH, T=T)
    o, bad = orc.step_evolve(s, rtol=1e-10)
    assert bad == 0
    assert (o["ie"][0] - ie0) / time == pytest.approx(-L, rel=1e-5)


# ------------------------------------------------------ He equilibria
def _node(cols, logT):
    k = int(np.argmin(np.abs(cols["logT"] - logT)))
    return k, cols["logT"][k]


@pytest.mark.parametrize("logT", [4.8, 5.0, 5.2])
def test_collisional_equilibrium_case_a(logT):
    """case A, no photons, fixed T at a row of the table: the steady state of
    src/solve.c:84-92 as written is x_HII/x_HI = gamma_HI/alpha_HII_A,
    x_HeII/x_HeI = gamma_HeI/alpha_HeII_A and x_HeIII/x_HeII =
    (gamma_HeII + alpha_HeII_A)/alpha_HeIII_A (the HeII line carries no
    gamma_HeI x_HeI source term); the ratios are taken from the file's
    columns and do not depend on ye or on the unit factors"""
    cols = rate_file(HUI)
    k, lt = _node(cols, logT)
    u = po.units(hmf=0.76)
    orc = po.Oracle(u, po.flags(isothermal=True, h_only=False, on_the_spot=False), rates=HUI)
    g = {n: cols[n][k] for n in ("HIci", "HIIrcA", "HeIci", "HeIIci", "HeIIrcA", "HeIIIrcA")}
    nH = 1.0 / u.U_CM ** 3
    slow = min(g.values()) * cgs_rate(u) * nH
    s = one_step(lambdaH=0.5, lambdaHeI=0.4, lambdaHeII=0.3, nH=nH, T=10 ** lt, time=200 / slow,
                 ie=1.0)
    o, bad = orc.step_evolve(s, rtol=1e-10)
    assert bad == 0
    xHI, xHeI, xHeII = o["lambdaH"][0], o["lambdaHeI"][0], o["lambdaHeII"][0]
    xHeIII = 1 - xHeI - xHeII
    assert (1 - xHI) / xHI == pytest.approx(g["HIci"] / g["HIIrcA"], rel=1e-5)
    assert xHeII / xHeI == pytest.approx(g["HeIci"] / g["HeIIrcA"], rel=1e-5)
    assert xHeIII / xHeII == pytest.approx((g["HeIIci"] + g["HeIIrcA"]) / g["HeIIIrcA"], rel=1e-5)


@pytest.mark.parametrize("logT", [4.6, 4.8])
def test_collisional_equilibrium_on_the_spot(logT):
    """on-the-spot (case B), no photons, fixed T at a table row: the He lines
    of src/solve.c:103-104 take their ionization from x_HI (not x_HeI), so
    the steady state is x_HI = alpha_HII_B/(gamma_HI + alpha_HII_B),
    x_HeII = gamma_HeI x_HI / alpha_HeII_B and
    x_HeIII = (gamma_HeII x_HI + alpha_HeII_B x_HeII)/alpha_HeIII_B, from the
    file's columns"""
    cols = rate_file(HUI)
    k, lt = _node(cols, logT)
    u = po.units(hmf=0.76)
    orc = po.Oracle(u, po.flags(isothermal=True, h_only=False, on_the_spot=True), rates=HUI)
    g = {n: cols[n][k] for n in ("HIci", "HIIrcB", "HeIci", "HeIIci", "HeIIrcB", "HeIIIrcB")}
    xHI = g["HIIrcB"] / (g["HIci"] + g["HIIrcB"])
    xHeII = g["HeIci"] * xHI / g["HeIIrcB"]
    xHeIII = (g["HeIIci"] * xHI + g["HeIIrcB"] * xHeII) / g["HeIIIrcB"]
    assert 0 < xHeII + xHeIII < 1, "choose a temperature with a physical steady state"
    nH = 1.0 / u.U_CM ** 3
    slow = min(g.values()) * cgs_rate(u) * nH
    s = one_step(lambdaH=0.5, lambdaHeI=0.4, lambdaHeII=0.3, nH=nH, T=10 ** lt, time=200 / slow,
                 ie=1.0)
    o, bad = orc.step_evolve(s, rtol=1e-10)
    assert bad == 0
    assert o["lambdaH"][0] == pytest.approx(xHI, rel=1e-5)
    assert o["lambdaHeII"][0] == pytest.approx(xHeII, rel=1e-5)
    assert 1 - o["lambdaHeI"][0] - o["lambdaHeII"][0] == pytest.approx(xHeIII, rel=1e-5)


# ------------------------------------------------------------- cooling
@pytest.mark.parametrize("ots", [True, False])
@pytest.mark.parametrize("logT", [4.2, 4.6, 5.4])
def test_cooling_rate_from_the_file_columns(ots, logT):
    """a short non-isothermal step at a table row: (ie1 - ie0 - heat)/time
    approaches -L with L = (nH / m_p) [(HIcic + HIcec) ye x_HI + HIIrcc ye
    x_HII + Brem ye x_HII] (src/solve.c:113-117), HIIrcc the case-B column
    on the spot and case A otherwise, Brem synthesized at load; ie is the
    ideal-gas energy 3/2 kT/m_p (X (1 + ye) + Y/4) per unit mass"""
    cols = rate_file(HUI)
    k, lt = _node(cols, logT)
    X, Y = 0.76, 0.24
    u = po.units(hmf=X)
    orc = po.Oracle(u, po.flags(isothermal=False, h_only=False, on_the_spot=ots), rates=HUI)
    xHI, xHeI, xHeII = 0.3, 0.2, 0.5
    ye = (1 - xHI) + Y / X * 0.25 * (xHeII + 2 * (1 - xHeI - xHeII))
    T = 10 ** lt
    ie0 = 1.5 * u.C_BOLTZMANN * T / u.U_MPROTON * (X * (1 + ye) + Y / 4)
    nH = 1e-2 / u.U_CM ** 3
    rcc = cols["HIIrccB" if ots else "HIIrccA"][k]
    assert "Brem" not in cols and "Compton" not in cols
    # neither file column exists: ar_init fills Brem from 1.42e-27 * 1.5 *
    # sqrt(v) at the x column (log10 T) and Compton with 0
    # (src/atomic-rate.c:58-64)
    brem = 1.42e-27 * 1.5 * np.sqrt(cols["logT"][k])
    L = u.C_H_PER_MASS * nH * cgs_cool(u) * (
        (cols["HIcic"][k] + cols["HIcec"][k]) * ye * xHI + rcc * ye * (1 - xHI)
        + brem * ye * (1 - xHI))
    time = 1e-7 * ie0 / L
    heat = 0.5 * L * time
    s = one_step(lambdaH=xHI, lambdaHeI=xHeI, lambdaHeII=xHeII, ie=ie0, heat=heat, time=time,
                 nH=nH, T=T)
    o, bad = orc.step_evolve(s, rtol=1e-10)
    assert bad == 0
    assert (o["ie"][0] - ie0 - heat) / time == pytest.approx(-L, rel=2e-5)


# ------------------------------------------------------------- deposit
def _particle(orc, xHI, xHeI, xHeII, mass=1.0):
    ps = po_chain(1, mass)
    ps.lambdaH[:] = xHI
    ps.lambdaHeI[:] = xHeI
    ps.lambdaHeII[:] = xHeII
    return ps


def po_chain(n, mass):
    class PS:
        pass
    ps = PS()
    ps.npar = n
    ps.pos = np.zeros((n, 3), np.float32)
    ps.sml = np.full(n, 1.0, np.float32)
    ps.flag = np.zeros(n, np.int8)
    for f in "mass rho ie yeMET yGdepHI yGdepHeI yGdepHeII yGrecHII yGrecHeII yGrecHeIII heat".split():
        setattr(ps, f, np.zeros(n, np.float32))
    ps.mass[:] = mass
    ps.lambdaH = np.ones(n)
    ps.lambdaHeI = np.ones(n)
    ps.lambdaHeII = np.zeros(n)
    ps.lasthit = np.zeros(n, np.int64)
    ps.hits = np.zeros(n, np.int32)
    return ps


def _one_hit(orc, ps, Nph, freq, b=0.3):
    hits = np.zeros(1, po.HIT_DTYPE)
    hits["ipar"] = 0
    hits["d"] = 1.0
    hits["b"] = b
    rays = np.zeros(1, po.RAY_DTYPE)
    rays["Nph"] = Nph
    rays["freq"] = freq
    return po.deposit(orc, ps, rays, np.array([0, 1], np.int64), hits)


@pytest.mark.parametrize("saturate", [False, True])
def test_single_h_he_particle_deposit_without_secondary_ionization(saturate):
    """one particle, photons above every threshold: Beer-Lambert absorption
    Nph (1 - e^-tau) with tau = sum of sigma_s x_s N_s K(b) (K the SPH
    column table over sml^2), shared between species by their opacity and
    capped at the atoms present (src/loop.c:479-504); each photo-ionization
    heats by the photon's excess energy over that species' threshold, and the
    ray keeps Nph minus what was absorbed"""
    X, Y = 0.76, 0.24
    u = po.units(hmf=X)
    orc = po.Oracle(u, po.flags(isothermal=False, h_only=False, secondary_ionization=False))
    xs = {"HI": 0.4, "HeI": 0.3, "HeII": 0.6}
    eng = 60.0 * u.U_EV
    K = po.reference_kernel_table()[int(4096 * 0.3)]   # sml = 1: K(b) / sml^2 = K
    sig = {"HI": orc.xs_get(0, eng), "HeI": orc.xs_get(1, eng), "HeII": orc.xs_get(2, eng)}
    per_mass = {"HI": xs["HI"] * X, "HeI": xs["HeI"] * Y / 4, "HeII": xs["HeII"] * Y / 4}
    # the mass that makes the total optical depth 1.5
    mass = 1.5 * u.U_MPROTON / (K * sum(sig[s] * per_mass[s] for s in sig))
    ps = _particle(orc, xs["HI"], xs["HeI"], xs["HeII"], mass=mass)
    mass = float(ps.mass[0])   # as stored (f32)
    NH = mass / u.U_MPROTON * X
    NHe = mass / u.U_MPROTON * Y / 4
    N = {"HI": xs["HI"] * NH, "HeI": xs["HeI"] * NHe, "HeII": xs["HeII"] * NHe}
    tau = {s: sig[s] * N[s] * K for s in N}
    ttot = sum(tau.values())
    assert ttot == pytest.approx(1.5, rel=1e-6) and min(tau.values()) > 0.01
    # Nph where the species' shares stay below their atoms, and 100x the atoms
    Nph = (1e3 * (NH + NHe)) if saturate else 0.05 * min(N.values())
    absorbed = Nph * -np.expm1(-ttot)
    a = {s: min(tau[s] / ttot * absorbed, N[s]) for s in N}
    if saturate:
        assert all(a[s] == N[s] for s in N)
    ln, tm, st = _one_hit(orc, ps, Nph, eng)
    assert ps.yGdepHI[0] == pytest.approx(a["HI"] / NH, rel=1e-6)
    assert ps.yGdepHeI[0] == pytest.approx(a["HeI"] / NHe, rel=1e-6)
    assert ps.yGdepHeII[0] == pytest.approx(a["HeII"] / NHe, rel=1e-6)
    thr = {"HI": 13.60569253, "HeI": 24.587, "HeII": 54.416}
    heat = sum(a[s] * (eng - thr[s] * u.U_EV) for s in N) / mass
    assert ps.heat[0] == pytest.approx(heat, rel=1e-4)
    assert tm[0] == pytest.approx(Nph - sum(a.values()), rel=1e-9, abs=1e-12 * Nph)
    assert st.saturated_deposit_count == (3 if saturate else 0)
    assert ps.hits[0] == 1


def test_secondary_ionization_fit_limits():
    """the fit of Ricotti, Gnedin & Shull (2002) that src/secion.c tabulates
    over E in [10, 200] eV and log10 x in [-8, 0]: phi_s(E, x) = y1_s(x) -
    y2_s(x) (28 eV / E)^0.4 and Eh = 1 - y1_h(x) + y2_h(x) (11 eV / E)^0.7,
    y1 the energy-independent high-energy forms of Shull & van Steenberg
    (1985), C (1 - x^b)^c with (C, b, c) = (0.3908, 0.4092, 1.7592) for HI,
    (0.0554, 0.4614, 1.6660) for HeI and (1, 0.2663, 1.3163) for heat.  The
    gate E >= gate - 0.1 compares internal energy units with 0.1, so it
    always passes (src/secion.c:48-49, :59-60) and the low-energy term
    applies down to 10 eV.  Checked at table rows: the y1 values, the E^0.4
    and E^0.7 laws of the remainder; in fully ionized gas (x -> 1) no energy
    goes to secondary ionization and all of it to heat; more electrons mean
    less secondary ionization of HI"""
    u = po.units(hmf=0.76)
    orc = po.Oracle(u, po.flags(isothermal=False, h_only=False), rates=HUI)
    e_min, e_step = 10 * u.U_EV, (200 - 10) * u.U_EV / 4096
    y_step = 8.0 / 1024

    def e_row(ev):
        return e_min + round((ev * u.U_EV - e_min) / e_step) * e_step

    svs = ((0.3908, 0.4092, 1.7592), (0.0554, 0.4614, 1.6660), (1.0, 0.2663, 1.3163))
    for j in (128, 512, 768, 960, 1000):
        lx = -8.0 + j * y_step
        x = 10.0 ** lx
        y1 = [C * (1 - x ** b) ** c for C, b, c in svs]
        Es = [e_row(v) for v in (12.0, 20.0, 60.0, 150.0)]
        for i in (0, 1):
            rem = [y1[i] - orc.secion_get(i, E, lx) for E in Es]
            assert all(r > 0 for r in rem), (i, j)     # the gate never cuts the low-energy term
            for E, r in zip(Es[1:], rem[1:]):
                assert r / rem[0] == pytest.approx((Es[0] / E) ** 0.4, rel=1e-9)
        rem = [orc.secion_get(2, E, lx) - (1 - y1[2]) for E in Es]
        for E, r in zip(Es[1:], rem[1:]):
            assert r / rem[0] == pytest.approx((Es[0] / E) ** 0.7, rel=1e-9)
    for ev in (20.0, 60.0, 150.0):
        E = e_row(ev)
        lxs = -8.0 + np.arange(1, 1023) * y_step
        phi = np.array([orc.secion_get(0, E, v) for v in lxs])
        if ev > 28.0:
            # above the gate, where the fit is meant to hold (below it the
            # always-on low-energy term bends phi up near x ~ 0.4)
            assert (np.diff(phi) <= 1e-12).all()
        assert orc.secion_get(0, E, -7.99) > orc.secion_get(1, E, -7.99) > 0
        # x -> 1: the table's last row (log10 x = -8/1024) is within 1e-3 of the limit
        assert abs(orc.secion_get(0, E, 0.0)) < 1e-3
        assert abs(orc.secion_get(1, E, 0.0)) < 1e-3
        assert abs(orc.secion_get(2, E, 0.0) - 1.0) < 1e-3


def test_secondary_ionization_drops_the_heii_terms():
    """src/loop.c:530-531 and :545-546 end the statement before the HeII
    term (a stray ';'), so with secondary ionization on a photo-ionization of
    HeII adds neither to yGdepHI nor to the heat; with it off the same hit
    heats by the HeII excess energy"""
    X, Y = 0.76, 0.24
    u = po.units(hmf=X)
    eng = 80.0 * u.U_EV
    out = {}
    for sec in (True, False):
        orc = po.Oracle(u, po.flags(isothermal=False, h_only=False, secondary_ionization=sec))
        ps = _particle(orc, 0.0, 0.0, 1.0, mass=1.0)
        NHe = 1.0 / u.U_MPROTON * Y / 4
        ln, tm, st = _one_hit(orc, ps, 0.01 * NHe, eng)
        out[sec] = (float(ps.yGdepHI[0]), float(ps.yGdepHeI[0]), float(ps.yGdepHeII[0]),
                    float(ps.heat[0]))
    assert out[True][2] > 0 and out[True][2] == out[False][2]
    assert out[True][0] == 0.0 and out[True][1] == 0.0 and out[True][3] == 0.0
    assert out[False][3] == pytest.approx(out[False][2] * (eng - 54.416 * u.U_EV) * NHe / 1.0,
                                          rel=1e-4)


# ------------------------------------------------------- He against LSODA
def test_helium_outliers_against_lsoda_are_lsodas_error(golden):
    """the reference's LSODA (tests/golden/chem_lsoda.npz, rtol 1e-4) and the
    oracle's ROS3 differ by up to ~8e-3 in x_HeI on two of 300 H + He steps.
    Measured against a tight solution (ROS3 at rtol 1e-10), the oracle's
    own error stays below 1e-3 on every step, and wherever the two differ by
    more than 1e-3 LSODA carries the larger error"""
    c = np.load(f"{golden}/chem_lsoda.npz")
    orc = po.Oracle(po.units(hmf=0.76), po.flags(isothermal=False, h_only=False))
    ours, bad = orc.step_evolve(c["he_in"])
    assert bad == 0
    tight, bad = orc.step_evolve(c["he_in"], rtol=1e-10)
    assert bad == 0
    lsoda = c["he_out"]
    worst_lsoda = 0.0
    for f in ("lambdaH", "lambdaHeI", "lambdaHeII"):
        e_ours = np.abs(ours[f] - tight[f])
        e_ls = np.abs(lsoda[f] - tight[f])
        assert e_ours.max() < 1e-3, (f, e_ours.max())
        far = np.abs(ours[f] - lsoda[f]) > 1e-3
        assert (e_ls[far] > e_ours[far]).all(), f
        worst_lsoda = max(worst_lsoda, e_ls.max())
    assert worst_lsoda > 1e-3   # the outliers exist (else this test pins nothing)


@pytest.mark.parametrize("ionized", [False, True])
def test_single_h_he_particle_deposit_with_secondary_ionization(ionized):
    """secondary ionization on (src/loop.c:516-552): each photo-electron of
    excess energy dE makes dE/I_s phi_s(dE, x) further ionizations of HI and
    HeI and heats by dE Eh(dE, x), x = ye NH/(NH + NHe) capped at 1; the
    HeII photo-electrons count for neither (the dropped terms).  Nearly
    neutral gas: the deposits follow from the pinned fit; fully ionized gas
    (x = 1): phi -> 0 and Eh -> 1, so the deposits equal those without
    secondary ionization, less the HeII heat"""
    X, Y = 0.76, 0.24
    u = po.units(hmf=X)
    eng = 70.0 * u.U_EV
    fr = (0.01, 0.01, 0.01) if ionized else (0.999, 0.998, 0.001)
    res = {}
    for sec in (True, False):
        orc = po.Oracle(u, po.flags(isothermal=False, h_only=False, secondary_ionization=sec))
        ps = _particle(orc, *fr, mass=1.0)
        NH = 1.0 / u.U_MPROTON * X
        NHe = 1.0 / u.U_MPROTON * Y / 4
        Nph = 1e-4 * NH * min(fr)
        ln, tm, st = _one_hit(orc, ps, Nph, eng)
        res[sec] = (ps, orc, tm[0])
    on, off = res[True][0], res[False][0]
    orc = res[True][1]
    dHI = float(off.yGdepHI[0])
    dHeI = float(off.yGdepHeI[0])
    dHeII = float(off.yGdepHeII[0])
    assert dHI > 0 and dHeI > 0 and dHeII > 0
    assert res[True][2] == res[False][2]     # the ray loses the same photons
    dE = {"HI": eng - u.C_HI_ENERGY, "HeI": eng - u.C_HEI_ENERGY, "HeII": eng - u.C_HEII_ENERGY}
    ye = (1 - fr[0]) + Y / X * 0.25 * (fr[2] + 2 * (1 - fr[1] - fr[2]))
    x = min(ye * X / (X + Y / 4), 1.0)
    lx = np.log10(x)
    if ionized:
        assert x == 1.0
        assert on.yGdepHI[0] == pytest.approx(dHI, rel=1e-2)
        assert on.yGdepHeI[0] == pytest.approx(dHeI, rel=1e-2)
        assert on.yGdepHeII[0] == dHeII
        heat_off_no_heii = (dHI * NH * dE["HI"] + dHeI * NHe * dE["HeI"]) / 1.0
        assert on.heat[0] == pytest.approx(heat_off_no_heii, rel=1e-2)
        assert off.heat[0] > on.heat[0] * 1.05
    else:
        sHI = (dHI * dE["HI"] * orc.secion_get(0, dE["HI"], lx)
               + dHeI * dE["HeI"] * orc.secion_get(0, dE["HeI"], lx)) / u.C_HI_ENERGY
        sHeI = (dHI * dE["HI"] * orc.secion_get(1, dE["HI"], lx)
                + dHeI * dE["HeI"] * orc.secion_get(1, dE["HeI"], lx)) / u.C_HEI_ENERGY
        assert sHI > 0.5 * dHI      # the secondaries matter in neutral gas
        assert on.yGdepHI[0] == pytest.approx(dHI + sHI, rel=1e-5)
        assert on.yGdepHeI[0] == pytest.approx(dHeI + sHeI, rel=1e-5)
        assert on.yGdepHeII[0] == dHeII
        heat = (dHI * NH * dE["HI"] * orc.secion_get(2, dE["HI"], lx)
                + dHeI * NHe * dE["HeI"] * orc.secion_get(2, dE["HeI"], lx))
        assert on.heat[0] == pytest.approx(heat, rel=1e-5)
