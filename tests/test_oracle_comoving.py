"""Pins the oracle's comoving pieces: expand_terms (the C restatement against
the reference's expand_terms.F built with flang, bitwise, and closed forms),
the interpolated pressure of Grid_ComputePressure.C (its special cases), and
grid::ComovingExpansionTerms as a whole (port against reference build)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from enzo_ppm import problems

needs_ref = pytest.mark.skipif(not oracle.available("ref"), reason="reference build absent")


def expand_terms(kind, rank, idual, coef, imethod, gamma, fields):
    """Call the oracle's (or the reference's) expand_terms on copies of
    `fields` = dict(p, d, e, ge, u, v, w, dold) arrays; returns the copies."""
    L = oracle.lib(kind)
    f = getattr(L, "orc_expand_terms" if kind == "port" else "expand_terms_")
    f.restype = None
    a = {k: np.array(v, dtype=np.float64) for k, v in fields.items()}
    n = len(a["d"])
    dp = C.POINTER(C.c_double)
    i64 = lambda v: C.byref(C.c_int64(v))  # noqa: E731
    f64 = lambda v: C.byref(C.c_double(v))  # noqa: E731
    ptr = lambda k: a[k].ctypes.data_as(dp)  # noqa: E731
    ecr = np.zeros(n)
    f(i64(rank), i64(n), i64(idual), f64(coef), i64(imethod), f64(gamma), ptr("p"), ptr("d"),
      ptr("e"), ptr("ge"), ptr("u"), ptr("v"), ptr("w"), ptr("dold"), ptr("e"), ptr("ge"),
      ptr("u"), ptr("v"), ptr("w"), i64(0), ecr.ctypes.data_as(dp), ecr.ctypes.data_as(dp))
    return a


def random_fields(rng, n):
    return dict(p=np.exp(rng.normal(0, 1, n)), d=np.exp(rng.normal(0, 1, n)),
                e=np.exp(rng.normal(1, 1, n)), ge=np.exp(rng.normal(0, 1, n)),
                u=rng.normal(0, 1, n), v=rng.normal(0, 1, n), w=rng.normal(0, 1, n),
                dold=np.exp(rng.normal(0, 1, n)))


@needs_ref
@pytest.mark.parametrize("rank,idual,imethod,gamma", [
    (3, 1, 0, 5.0 / 3.0), (3, 0, 0, 1.4), (2, 1, 0, 1.4), (1, 0, 0, 5.0 / 3.0), (3, 1, 2, 1.4)])
def test_expand_terms_port_matches_reference(rank, idual, imethod, gamma):
    rng = np.random.default_rng(rank * 10 + idual + imethod)
    f = random_fields(rng, 2000)
    for coef in (0.0, 1e-3, 0.05, 0.9):
        a = expand_terms("port", rank, idual, coef, imethod, gamma, f)
        b = expand_terms("ref", rank, idual, coef, imethod, gamma, f)
        for k in a:
            assert np.array_equal(a[k], b[k]), (k, coef)


def test_expand_terms_closed_form():
    rng = np.random.default_rng(1)
    f = random_fields(rng, 500)
    a = expand_terms("port", 3, 1, 0.0, 0, 1.4, f)        # coef 0: identity
    for k in f:
        assert np.array_equal(a[k], f[k]), k
    c = 0.02
    a = expand_terms("port", 3, 1, c, 0, 1.4, f)
    # velocities and gas energy: semi-implicit decay factors
    assert np.allclose(a["u"], f["u"] * (1 - c / 2) / (1 + c / 2), rtol=1e-15, atol=0)
    assert np.allclose(a["ge"], f["ge"] * (1 - c) / (1 + c), rtol=1e-15, atol=0)
    # total energy: decay, then the 3 - 2/(gamma-1) pressure term (gamma 1.4:
    # -2), floored at half
    e1 = f["e"] * (1 - c) / (1 + c)
    expect = np.maximum(e1 + c * 2.0 * f["p"] / f["d"], 0.5 * e1)
    assert np.allclose(a["e"], expect, rtol=1e-14, atol=0)
    # rank 1: v and w untouched
    a = expand_terms("port", 1, 1, c, 0, 1.4, f)
    assert np.array_equal(a["v"], f["v"]) and np.array_equal(a["w"], f["w"])


def test_pressure_at_special_cases():
    g, glob, _ = problems.turbulent(8, seed=3)
    old = g.copy()
    rng = np.random.default_rng(2)
    for f in g.fields:
        f *= np.exp(rng.normal(0, 0.1, f.shape))
    # weight 1: the current state's pressure; weight 0: the old state's
    assert np.array_equal(oracle.compute_pressure_at("port", g, old, glob, 1.0),
                          oracle.compute_pressure("port", g, glob))
    assert np.array_equal(oracle.compute_pressure_at("port", g, old, glob, 0.0),
                          oracle.compute_pressure("port", old, glob))
    # weight 1/2 with dual energy: (gamma-1) * mean(rho) * mean(ge)
    p = oracle.compute_pressure_at("port", g, old, glob, 0.5)
    k = g.GENum
    expect = (glob["Gamma"] - 1.0) * (0.5 * g.fields[0] + 0.5 * old.fields[0]) * \
        (0.5 * g.fields[k] + 0.5 * old.fields[k])
    assert np.allclose(p, np.maximum(expect, 1e-20), rtol=1e-15, atol=0)


@needs_ref
@pytest.mark.parametrize("dual", [1, 0])
def test_comoving_expansion_terms_port_matches_reference(dual):
    g, glob, _ = problems.turbulent(10, seed=5, ncolours=2)
    glob = dict(glob, DualEnergyFormalism=dual)
    old = g.copy()
    rng = np.random.default_rng(6)
    for f in g.fields:
        f *= np.exp(rng.normal(0, 0.05, f.shape))
    a, b = g.copy(), g.copy()
    oracle.comoving_expansion_terms("port", a, old, glob, 0.5000000000000001, 0.013)
    oracle.comoving_expansion_terms("ref", b, old, glob, 0.5000000000000001, 0.013)
    for x, y in zip(a.fields, b.fields):
        assert np.array_equal(x, y)
    assert not np.array_equal(a.fields[2], g.fields[2])     # velocities changed


@needs_ref
def test_comoving_steps_port_matches_reference():
    """Whole comoving steps (Courant with a(t), a(t+dt/2)-scaled widths,
    expansion terms) on a cosmological box with gravity and 6 species."""
    from enzo_ppm import comoving
    from helpers import field_report, oracle_comoving_steps
    g, glob, _ = problems.cosmological(12, seed=8)
    a, b = g.copy(), g.copy()
    model = comoving.EdS(0.1, 1.0)
    _, da = oracle_comoving_steps("port", a, glob, model, 1.0, 3, 0.4)
    _, db = oracle_comoving_steps("ref", b, glob, model, 1.0, 3, 0.4)
    assert da == db
    assert all(r[1] for r in field_report(a, b))
