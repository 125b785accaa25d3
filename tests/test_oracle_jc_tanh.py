"""The oracle's tanh-modified Johnson-Cook flow stress (sigma_yield_tanh, plasticity.cuh:48-70,
after Calamaz 2008), pinned against its closed form with numpy's libm in distinct regimes, and
through its ratio to the plain Johnson-Cook stress, which shares the strain-rate and thermal
factors: sigma_tanh / sigma_JC = (A + B e^n exp(-e^a)) / (A + B e^n) * D(e, T)."""
from __future__ import annotations

import math

import pytest

import oracle_binding as ob

# A, B, C, n, m, Tmelt, Tref, eps_dot_ref (Ti6Al4V, the reference's units) and tanh a, b, c, d
JC = [0.0086200, 0.0033100, 0.0100000, 0.34, 0.8, 1836.0, 300.0, 1e-6]
TANH = (2.0, 5.0, 1.0, 0.5)


def D(eps, T, tanh):
    a, b, c, d = tanh
    tD = 1.0 - (T / JC[5]) ** d
    tS = (T / JC[5]) ** b
    return tD + (1.0 - tD) * math.tanh(1.0 / (eps + tS) ** c)


@pytest.mark.parametrize("eps,epsd,T", [(0.0, 0.0, 300.0), (0.05, 1e-7, 500.0), (0.3, 2e-5, 900.0),
                                        (1.5, 3e-3, 1200.0)])
def test_ratio_to_plain_jc(eps, epsd, T):
    A, B, n = JC[0], JC[1], JC[3]
    a = TANH[0]
    want = (A + B * eps ** n * math.exp(-eps ** a)) / (A + B * eps ** n) * D(eps, T, TANH)
    got = ob.sigma_yield_tanh(JC, TANH, eps, epsd, T) / ob.sigma_yield(JC, eps, epsd, T)
    assert got == pytest.approx(want, rel=1e-13)


def test_closed_form():
    eps, epsd, T = 0.2, 5e-6, 700.0
    A, B, C, n, m, Tm, Tr, e0 = JC
    theta = (T - Tr) / (Tm - Tr)
    ed = epsd / e0
    TB = 1.0 + C * math.log(ed) if ed > 1.0 else (1.0 + ed) ** C
    want = (A + B * eps ** n / math.exp(eps ** TANH[0])) * TB * (1.0 - theta ** m) * D(eps, T, TANH)
    assert ob.sigma_yield_tanh(JC, TANH, eps, epsd, T) == pytest.approx(want, rel=1e-13)


def test_softening_term_vanishes_at_the_melting_point():
    # T = Tmelt: tanh_D = 0, tanh_S = 1, and the thermal factor 1 - theta^m = 0
    assert ob.sigma_yield_tanh(JC, TANH, 0.3, 1e-6, JC[5]) == 0.0
