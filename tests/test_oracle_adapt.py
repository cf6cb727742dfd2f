"""The oracle's restatement of the adaptive driver's refinement indicator,
checkconv and cal_conv (src/adapt.f:1260-1332, 1122-1192):

- against golden vectors written by the reference Fortran itself
  (tests/golden/make_goldens.py, calconv_p<p>.npz), bit for bit: the three
  indicators and every decision, for threshold sets that put fac1 and fac2
  exactly on indicator values (the tests are strict), ask for a roll-back,
  and zero a field (its indicator then vanishes);
- against what the mathematics fixes: a field that is the same constant on
  every element has no jumps; scaling a field by a power of two leaves the
  normalised indicator bit-identical; a jump planted on one element shows
  on that element and its face neighbours only, and nowhere else."""
import numpy as np
import pytest

from conftest import gold, oracle_for
from spoc_amd import configs

CASES = [(3, 12, 4), (6, 33, 7)]


def _case(nc, ng, ngf):
    g = gold(f"calconv_p{nc}.npz")
    O = oracle_for(nc, ng, ngf)
    t = O.tables()
    t.update(nc=nc, ng=ng, ngf=ngf, nm=O.nm)
    c = configs.random_case(t, int(g["nx"]), int(g["ny"]), seed=int(g["seed"]), scale=0.01)
    nm = O.nm
    st = [a[:nm] for a in c["state"]]
    for k, a in zip(("u0", "v0", "h0"), st):
        assert np.array_equal(a, g[k]), k  # the generator reproduces the inputs
    return O, c["mesh"], st, g


@pytest.mark.parametrize("nc,ng,ngf", CASES)
def test_checkconv_bit_exact(nc, ng, ngf):
    O, m, st, g = _case(nc, ng, ngf)
    for f in range(3):
        assert np.array_equal(O.checkconv(m, st[f]), g["amax"][f]), f


@pytest.mark.parametrize("nc,ng,ngf", CASES)
@pytest.mark.parametrize("which", ["mid", "rollback", "flat"])
def test_cal_conv_decisions_bit_exact(nc, ng, ngf, which):
    O, m, st, g = _case(nc, ng, ngf)
    if which == "flat":
        st = [st[0], st[1], np.zeros_like(st[2])]
    f1, f2, f3 = g[f"{which}_fac"]
    r = O.cal_conv(m, *st, f1, f2, f3)
    assert np.array_equal(r["fconv"], g[f"{which}_fconv"])
    assert np.array_equal(r["fderaff"], g[f"{which}_fderaff"])
    assert [r["conv"], r["deraff"], r["timedir"]] == list(g[f"{which}_flags"])
    if which == "flat":
        assert np.array_equal(r["amax"][2], g["flat_amax_h"])
        assert not r["amax"][2].any()
    else:
        assert np.array_equal(r["amax"], g["amax"])
    assert r["supmax"] == r["amax"].max()


def _random_mesh(nc, ng, ngf, seed=5):
    O = oracle_for(nc, ng, ngf)
    t = O.tables()
    t.update(nc=nc, ng=ng, ngf=ngf, nm=O.nm)
    c = configs.random_case(t, 8, 7, seed=seed, scale=0.01)
    return O, c["mesh"], [a[:O.nm] for a in c["state"]]


@pytest.mark.parametrize("nc,ng,ngf", CASES)
def test_constant_field_has_no_jump(nc, ng, ngf):
    O, m, st = _random_mesh(nc, ng, ngf)
    a = np.zeros_like(st[0])
    a[0] = 0.37  # the same mean on every element, no other mode
    assert not O.checkconv(m, a).any()


@pytest.mark.parametrize("nc,ng,ngf", CASES)
def test_power_of_two_scaling_is_exact(nc, ng, ngf):
    O, m, st = _random_mesh(nc, ng, ngf)
    ref = O.checkconv(m, st[0])
    assert ref.any()
    for s in (2.0 ** 7, 2.0 ** -9):
        assert np.array_equal(O.checkconv(m, st[0] * s), ref)


@pytest.mark.parametrize("nc,ng,ngf", CASES)
def test_planted_jump_is_local(nc, ng, ngf):
    O, m, st = _random_mesh(nc, ng, ngf)
    ne = st[0].shape[1]
    a = np.zeros_like(st[0])
    a[0] = 1.0
    k = ne // 2
    a[1:, k] = 0.25  # element k alone gets higher modes
    amax = O.checkconv(m, a)
    tseg = np.asarray(m["tseg"])
    near = {k}
    for i in range(tseg.shape[1]):
        e1, e2 = tseg[0, i] - 1, tseg[1, i] - 1
        if e2 >= 0 and k in (e1, e2):
            near |= {e1, e2}
    assert len(near) >= 2
    assert all(amax[e] > 0 for e in near)
    assert not np.delete(amax, sorted(near)).any()
    # the planted element sees every one of its interior faces: its value is
    # the largest
    assert amax[k] == amax.max()


def test_tiny_mean_gives_zero_indicator():
    """max |a(1,e)| <= 1e-6 zeroes the indicator (src/adapt.f:1321-1329)."""
    O, m, st = _random_mesh(3, 12, 4)
    a = st[0].copy()
    a[0] = 1e-7
    assert not O.checkconv(m, a).any()
    a[0, 3] = 2e-6
    assert O.checkconv(m, a).any()
