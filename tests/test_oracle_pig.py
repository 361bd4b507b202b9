"""PIG_EOS in the oracle (pig_rhoT_given_full / given1 / given2 / rhoP /
rhoe, source/source.F90:25457-26731) on the generated partially ionised
hydrogen table (phlegethon_amd/piggen.py), pinned by the closed forms of
the model and by what the interpolation fixes:

- P = (1 + x) rho k T / m_H + a T^4 / 3 with the Saha fraction x, and
  E = (3/2 (1 + x) k T + x chi) / m_H + a T^4 / rho, at random off-node
  points across the ionisation zone (a Helmholtz-style ion term or a
  Ye != 1 scaling would be off by O(1));
- c_v equals dE/dT of the closed form and the sound speed the relativistic
  Gamma_1 form of the closed-form P(rho, T), E(rho, T) derivatives;
- at the nodes the interpolation returns the table's own values;
- T -> e -> T and T -> P -> T Newton round trips from a 1 % warm start.
"""
import numpy as np
import pytest

from phlegethon_amd import piggen
from phlegethon_amd.config import EOS_PIG, ConfigError, PhlxConfig

ARAD = 7.565767381646406e-15
CL = 2.99792458e10


@pytest.fixture(scope="module")
def pig(oracle_mod):
    tab = piggen.build(**piggen.DEFAULT)
    o = oracle_mod.OracleHelm(tab)
    yield tab, o
    o.close()


def _closed(rho, T):
    x = piggen.ionised_fraction(rho, T)
    P = (1.0 + x) * rho * piggen.KB * T / piggen.MH + ARAD * T ** 4 / 3.0
    E = (1.5 * (1.0 + x) * piggen.KB * T + x * piggen.CHI) / piggen.MH + ARAD * T ** 4 / rho
    return P, E, x


def _points(n, seed=1):
    rng = np.random.default_rng(seed)
    return 10 ** rng.uniform(-11.5, -0.5, n), 10 ** rng.uniform(3.6, 7.9, n)


def test_closed_forms(pig):
    tab, o = pig
    rho, T = _points(4000)
    one = np.ones_like(rho)
    P, E, cs, cv = o.full(rho, T, one, one)
    Pe, Ee, x = _closed(rho, T)
    assert np.max(np.abs(P / Pe - 1.0)) < 2e-4      # interpolation, 0.1 x 0.05 dex cells
    assert np.max(np.abs(E / Ee - 1.0)) < 5e-4
    assert 0.05 < np.mean((x > 0.05) & (x < 0.95)) < 0.6   # the ionisation zone is sampled
    # abar / zbar are ignored by the PIG evaluation
    P2, E2, _, _ = o.full(rho, T, 4.0 * one, 2.0 * one)
    assert np.array_equal(P, P2) and np.array_equal(E, E2)


def test_cv_and_sound(pig):
    tab, o = pig
    rho, T = _points(500, seed=2)
    one = np.ones_like(rho)
    P, E, cs, cv = o.full(rho, T, one, one)
    h = 1e-5
    Pp, Ep, _ = _closed(rho, T * (1 + h))
    Pm, Em, _ = _closed(rho, T * (1 - h))
    dEdT = (Ep - Em) / (2 * h * T)
    dPdT = (Pp - Pm) / (2 * h * T)
    Pr, _, _ = _closed(rho * (1 + h), T)
    Pl, _, _ = _closed(rho * (1 - h), T)
    dPdr = (Pr - Pl) / (2 * h * rho)
    Pc, Ec, _ = _closed(rho, T)
    # second derivatives of f: the interpolation error peaks in the steep
    # part of the ionisation zone
    rel = np.abs(cv / dEdT - 1.0)
    assert np.median(rel) < 1e-4 and rel.max() < 1e-2, (np.median(rel), rel.max())
    zz = Pc / rho
    chiT = T / Pc * dPdT
    chir = dPdr / zz
    gam1 = chiT * (zz * chiT / (T * dEdT)) + chir
    sound = CL * np.sqrt(gam1 / (1.0 + (Ec + CL ** 2) / zz))
    rel = np.abs(cs / sound - 1.0)
    assert np.median(rel) < 1e-4 and rel.max() < 1e-2, (np.median(rel), rel.max())


def test_nodes_exact(pig):
    tab, o = pig
    i = np.array([3, 40, 77, 100])
    j = np.array([5, 20, 50, 80])
    rho = tab.rho_nodes[i]
    T = tab.T_nodes[j]
    one = np.ones_like(rho)
    P, E, cs, cv = o.full(rho, T, one, one)
    f9 = tab.f9[j, i]       # f, fd, ft, ...
    P_tab = rho * rho * f9[:, 1] + ARAD * T ** 4 / 3.0
    E_tab = f9[:, 0] - T * f9[:, 2] + ARAD * T ** 4 / rho
    np.testing.assert_allclose(P, P_tab, rtol=1e-13)
    np.testing.assert_allclose(E, E_tab, rtol=1e-13)


def test_newton_round_trips(pig):
    tab, o = pig
    rho, T = _points(3000, seed=3)
    one = np.ones_like(rho)
    P, E, _, _ = o.full(rho, T, one, one)
    for fn, target in ((o.rhoe_given, E), (o.rhoP_given, P)):
        Tb = fn(rho, target, one, one, T * 1.01)[0]
        ok = np.abs(Tb / T - 1.0) < 1e-8
        assert np.mean(ok) > 0.995, np.mean(ok)


def test_config():
    c = PhlxConfig.from_opts(["sdims_make=2", "nx1_make=8", "nx2_make=8", "nas_make=2",
                              "PIG_EOS", "USE_FASTEOS"])
    assert c.eos == EOS_PIG and c.fasteos
    PhlxConfig(sdims=2, nx1=8, nx2=8, eos=EOS_PIG, nas=0, fasteos=True).validate()
    with pytest.raises(ConfigError):
        PhlxConfig(sdims=2, nx1=8, nx2=8, eos=EOS_PIG, nas=1).validate()
    with pytest.raises(ConfigError):
        PhlxConfig(sdims=2, nx1=8, nx2=8, eos=EOS_PIG, nas=2, precision=1).validate()


def test_hydro_runs_and_conserves(oracle_mod, pig):
    """The oracle's hydro path on the PIG box (FASTEOS and the per-face
    Newton): mass and energy conserved on the periodic box."""
    from phlegethon_amd import problems as P
    tab, o = pig
    for fasteos, nas in ((True, 2), (False, 2), (True, 0), (False, 0)):
        pb = P.pig_box(o, tab, n=16, fasteos=fasteos, nas=nas)
        r = oracle_mod.OracleRun(pb)
        r.init()
        s0 = r.state()
        r.run(5)
        s1 = r.state()
        r.close()

        def tot(s):
            p = s["prim"]
            ek = 0.5 * p[0] * (p[1] ** 2 + p[2] ** 2)
            return p[0].sum(), (s["eint"] + ek).sum()

        m0, e0 = tot(s0)
        m1, e1 = tot(s1)
        assert abs(m1 / m0 - 1) < 1e-13 and abs(e1 / e0 - 1) < 1e-12
        assert np.abs(s1["temp"] - s0["temp"]).max() > 1.0
