"""Pins the oracle's numerics: Philox against published / independent KAT
vectors, exp/log/sqrt/log k! against the mathematics, the Poisson sampler
against exact quantiles and goodness of fit."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

from oracle import ffi

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123 kat_vectors (Salmon et al. 2011), philox4x32_10
RANDOM123_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def test_philox_published_kat():
    for ctr, key, out in RANDOM123_KAT:
        assert ffi.philox(ctr, key) == out


def test_philox_matches_rocrand_vectors():
    with open(os.path.join(HERE, "golden", "philox_kat.json")) as f:
        kat = json.load(f)
    assert len(kat) >= 20
    for k in kat:
        assert ffi.philox(k["ctr"], k["key"]) == k["out"]


def test_exp_log_sqrt_accuracy():
    L = ffi.lib()
    xs = np.concatenate([-np.linspace(0, 12, 4001), -np.geomspace(1e-15, 700, 800), np.linspace(0, 5, 101)])
    for x in xs:
        e = L.orc_exp(float(x))
        assert abs(e - math.exp(x)) <= 4.5e-16 * math.exp(x), x
    ys = np.concatenate([np.geomspace(1e-300, 1e300, 4000), np.linspace(0.25, 4, 2001)])
    for y in ys:
        l = L.orc_log(float(y))
        assert abs(l - math.log(y)) <= 4.5e-16 * max(abs(math.log(y)), 1e-2), y
    for y in np.geomspace(1.0, 1e15, 1000):
        assert abs(L.orc_sqrt(float(y)) - math.sqrt(y)) <= 4.5e-16 * math.sqrt(y)
    assert L.orc_log(1.0) == 0.0 and L.orc_exp(0.0) == 1.0


def test_logfact_matches_lgamma():
    L = ffi.lib()
    for k in list(range(0, 40)) + [100, 1000, 10**5, 10**8]:
        assert abs(L.orc_logfact(k) - math.lgamma(k + 1)) <= 1e-10 * max(1.0, math.lgamma(k + 1))


def test_zero_event_shortcut_bound():
    """The kernels skip the exp when u <= (1 - lam) - 2^-50; that is exact only
    if this bound never exceeds the oracle's computed e^-lam."""
    L = ffi.lib()
    lams = np.concatenate([np.geomspace(1e-300, 1.0, 20000), np.linspace(0, 1, 20001)[1:]])
    for lam in lams:
        assert (1.0 - lam) - 2.0 ** -50 <= L.orc_exp(-float(lam)), lam


def _poisson_from_word(lam, u0):
    L = ffi.lib()
    ctr = (C.c_uint32 * 4)(0, 0, 0, 0)
    key = (C.c_uint32 * 2)(0, 0)
    L.orc_poisson.argtypes = [C.c_double, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                              C.c_uint32]
    L.orc_poisson.restype = C.c_int64
    return L.orc_poisson(lam, u0, ctr, key, 0)


@pytest.mark.parametrize("lam", [0.01, 0.37, 1.0, 2.5, 6.0, 9.99])
def test_inversion_equals_exact_quantile(lam):
    """Small means: the draw is the Poisson quantile of u = (w + 0.5) 2^-32."""
    rng = np.random.default_rng(int(lam * 1000))
    words = rng.integers(0, 2**32, size=3000, dtype=np.uint64)
    cdf = stats.poisson.cdf(np.arange(0, 200), lam)
    for w in words:
        u = (float(w) + 0.5) * 2.0 ** -32
        k = _poisson_from_word(lam, int(w))
        # skip words within 1e-12 of a CDF step (rounding of the running sum)
        if np.min(np.abs(cdf - u)) < 1e-12:
            continue
        assert k == int(np.searchsorted(cdf, u)), (lam, u)


SQ_DELTA = 2.0 ** -15
F10_BITS = int(np.array(10.0, dtype=np.float32).view(np.uint32))


def test_squeeze_float_exp():
    """The squeeze's float e^-l is within 4 float ulps of e^-l."""
    L = ffi.lib()
    for l in np.concatenate([np.geomspace(1e-30, 9.999, 4000), np.linspace(0, 9.999, 4001)]):
        lf = float(np.float32(l))
        g = L.orc_sq_exp_neg(lf)
        assert abs(g - math.exp(-lf)) <= 4 * 2.0 ** -24 * math.exp(-lf), lf


def test_squeeze_cdf_error_bound_every_float_mean():
    """The bound the kernels' float squeeze rests on, over EVERY float mean in
    (0, 10) and k <= 15: |G_k - F_k| < 2^-20, F_k the double search's CDF.
    With the mean's rounding (<= 2^-21), u's rounding and G -+ delta's
    rounding (2^-25 each) the guard band 2^-15 leaves a factor > 25."""
    worst = ffi.squeeze_maxerr(1, F10_BITS, 1)
    assert worst < 2.0 ** -20, worst
    assert worst + 2.0 ** -21 + 2 * 2.0 ** -25 < SQ_DELTA / 25


def _boundary_words(lam, n_steps=12, spread=40):
    """Words w whose uniform (w + 0.5) 2^-32 lies within `spread` words of
    the oracle's CDF steps F_0..F_{n_steps-1} at mean lam."""
    L = ffi.lib()
    p = L.orc_exp(-lam)
    F = p
    words = []
    for k in range(n_steps):
        c = int(F * 2.0 ** 32)
        words.extend(w for w in range(c - spread, c + spread) if 0 <= w < 2**32)
        p = (p * lam) / (k + 1)
        F = F + p
    return words


@pytest.mark.parametrize("lam", [1e-6, 0.0137, 0.37, 1.0, 2.5, 6.0, 9.99])
def test_squeeze_decisions_equal_double_search(lam):
    """Draw by draw, the kernels' fast path (zero shortcut + float squeeze)
    either leaves the draw to the double search or returns its result:
    random words, and words right at each CDF step where a band-less float
    comparison does go wrong."""
    rng = np.random.default_rng(int(lam * 1e6) + 7)
    words = [int(w) for w in rng.integers(0, 2**32, size=20000, dtype=np.uint64)]
    decided = 0
    for w in words:
        k = ffi.squeeze_poisson(lam, w)
        if k >= 0:
            decided += 1
            assert k == ffi.poisson_inversion(lam, w), (lam, w)
    # left to the double search: k > 15, or u within the band (~6e-5 per comparison)
    assert decided >= (stats.poisson.cdf(15, lam) - 0.003) * len(words)
    bw = _boundary_words(lam)
    lf = float(np.float32(lam))
    wrong_without_band = 0
    for w in bw:
        exact = ffi.poisson_inversion(lam, w)
        k = ffi.squeeze_poisson(lam, w)
        assert k < 0 or k == exact, (lam, w)
        uf = float(np.float32((w + 0.5) * 2.0 ** -32))
        k0 = ffi.squeeze_decide(lf, uf, 0.0)
        wrong_without_band += k0 >= 0 and k0 != exact
    if lam > 1e-3:
        # the guard band is needed: without it the float search errs here
        assert wrong_without_band > 0, lam


def test_float_mean_hop_draws_equal_double_search():
    """The leap kernel forms hop means in single precision (tau, d, x to
    float, three float products: within 8 2^-24 of the double mean) and its
    uniform from the word rounded to float; its zero test and squeeze keep
    margins for both.  Draw by draw over random operands (x across 2^24,
    where x itself rounds) and over words placed at the CDF steps of each
    mean, a settled draw equals the oracle's draw from the double mean."""
    rng = np.random.default_rng(77)
    n = 200000
    tau = 10.0 ** rng.uniform(-9, 0, n)
    d = 10.0 ** rng.uniform(-2, 1, n)
    x = np.minimum(10.0 ** rng.uniform(0, 9.6, n), 2**32 - 1).astype(np.uint32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    # keep means where the fast path is meant to work, mostly below 10
    lam = tau * ((d * x.astype(float)) * w.astype(float))
    keep = lam < 12
    tau, d, x, w, lam = tau[keep], d[keep], x[keep], w[keep], lam[keep]
    u0 = rng.integers(0, 2**32, size=len(tau), dtype=np.uint64).astype(np.uint32)
    fast, exact = ffi.squeeze_hop_check(tau, d, x, w, u0)
    ok = fast >= 0
    assert np.array_equal(fast[ok], exact[ok])
    assert (exact[ok] >= 0).all()  # never settles a mean >= 10
    assert ok[lam < 5].mean() > 0.999  # above, k > 15 and means >= 10 are left to the double search
    # words at the CDF steps of 3000 of the means (where a looser margin errs)
    L = ffi.lib()
    sel = rng.choice(np.flatnonzero(lam < 9.99), 3000, replace=False)
    tt, dd, xx, ww, uu = [], [], [], [], []
    for i in sel:
        p = L.orc_exp(-lam[i])
        F = p
        for k in range(6):
            c = int(F * 2.0**32)
            for wd in (c - 3, c - 1, c, c + 1, c + 3):
                if 0 <= wd < 2**32:
                    tt.append(tau[i]); dd.append(d[i]); xx.append(x[i]); ww.append(w[i]); uu.append(wd)
            p = (p * lam[i]) / (k + 1)
            F = F + p
    fast, exact = ffi.squeeze_hop_check(np.array(tt), np.array(dd), np.array(xx, dtype=np.uint32),
                                        np.array(ww, dtype=np.float32), np.array(uu, dtype=np.uint32))
    ok = fast >= 0
    assert np.array_equal(fast[ok], exact[ok])


@pytest.mark.parametrize("lam", [0.05, 0.8, 4.0, 9.9, 10.0, 12.5, 40.0, 300.0, 20000.0])
def test_poisson_goodness_of_fit(lam):
    n = 200000
    s = ffi.poisson_many(lam, 12345 + int(lam * 7), n)
    se = math.sqrt(lam / n)
    assert abs(s.mean() - lam) < 4.5 * se
    var_se = lam * math.sqrt(2.0 / n) * math.sqrt(1 + 1 / (2 * lam))
    assert abs(s.var() - lam) < 4.5 * var_se
    lo, hi = int(max(0, lam - 6 * math.sqrt(lam))), int(lam + 6 * math.sqrt(lam) + 6)
    ks = np.arange(lo, hi + 1)
    obs = np.array([(s == k).sum() for k in ks], dtype=float)
    exp = stats.poisson.pmf(ks, lam) * n
    keep = exp > 20
    obs, exp = obs[keep], exp[keep]
    chi2 = ((obs - exp) ** 2 / exp).sum()
    p = stats.chi2.sf(chi2, keep.sum() - 1)
    assert p > 1e-4, (lam, chi2, p)


def _coarse_bound(lf):
    """The leap walk's bound on a draw's top byte (float32 operations in the
    kernel's order): (255 - 2^-12) - lf (256 + 2^-10) for lf > 0."""
    lf = np.float32(lf)
    return np.where(lf > 0, np.float32(255.0 - 2.0 ** -12) - lf * np.float32(256.0 + 2.0 ** -10),
                    np.float32(256.0)).astype(np.float32)


def test_coarse_bound_settles_only_zero_draws():
    """A draw whose top byte hb satisfies hb <= coarse_bound(lf) is settled as
    0 without its fine bits; that is sound if every word with that byte gives
    0 for every double mean the float mean lf can stand for (lf within 8 2^-24
    relative).  Over 200,000 float means (log-uniform 1e-9..2, and dense just
    below each byte's threshold) and every byte: (hb + 1) 2^-8 <= 1 - lam_max
    holds wherever the bound settles, and the oracle's inversion of the
    largest word with that byte at lam_max is 0 on a sample; the bound also
    settles almost every draw it can (it gives away at most one byte value
    against the exact threshold)."""
    rng = np.random.default_rng(5)
    lf = np.concatenate([10.0 ** rng.uniform(-9, np.log10(2.0), 150000),
                         # means right at the byte thresholds 1 - (hb + 1)/256
                         (1.0 - (rng.integers(0, 256, 50000) + 1) / 256.0) * (1 + rng.uniform(-1e-6, 1e-6, 50000))])
    lf = lf[lf > 0].astype(np.float32)
    g = _coarse_bound(lf)
    hb = np.arange(256, dtype=np.float32)
    settled = hb[None, :] <= g[:, None]
    lam_max = lf.astype(np.float64) * (1.0 + 8 * 2.0 ** -24)
    sound = (hb[None, :].astype(np.float64) + 1.0) / 256.0 <= 1.0 - lam_max[:, None]
    assert not (settled & ~sound).any()
    # the exact threshold settles hb iff (hb + 1)/256 <= 1 - lam for the float mean itself
    exact = (hb[None, :].astype(np.float64) + 1.0) / 256.0 <= 1.0 - lf.astype(np.float64)[:, None]
    assert ((exact.sum(1) - settled.sum(1)) <= 1).all()
    i, j = np.nonzero(settled)
    pick = rng.choice(len(i), 2000, replace=False)
    for a, b in zip(i[pick], j[pick]):
        word = (int(b) << 24) | 0xFFFFFF
        assert ffi.poisson_inversion(float(lam_max[a]), word) == 0
    # a mean above 1 is never settled
    assert not settled[lf > 1.0].any()


def test_float_mean_ptrs_threshold_is_conservative():
    """The leap walk sends a tile to the redo pass when a hop's float mean
    tau_f ((d_f x_f) w) exceeds 9.9999f; a tile it keeps is drawn with the
    inversion, which needs every double mean tau ((d x) w) below 10.  Check
    that a float mean at or below the threshold never has a double mean of
    10 or more, on means packed around 10 (random d, w, counts and the tau
    that puts the double mean within 2e-4 of 10), including the largest
    counts."""
    F = np.float32
    thr = F(9.9999)
    rng = np.random.default_rng(7)
    n = 400_000
    d = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), n))
    w = rng.uniform(0.5, 1.5, n).astype(F)  # weights are floats
    x = rng.integers(1, 2**32, n, dtype=np.uint64)
    x[: n // 4] = rng.integers(1, 1000, n // 4, dtype=np.uint64)
    xd = x.astype(np.float64)
    target = 10.0 * (1.0 + rng.uniform(-2e-4, 2e-4, n))
    tau = target / ((d * xd) * w.astype(np.float64))
    lam = tau * ((d * xd) * w.astype(np.float64))
    lf = tau.astype(F) * ((d.astype(F) * xd.astype(F)) * w)
    kept = lf <= thr
    assert kept.sum() > n // 10 and (~kept).sum() > n // 10  # both sides of the threshold are exercised
    assert not np.any(lam[kept] >= 10.0), "a kept tile would draw a mean >= 10 with the inversion"
    # the float mean's relative error stays inside the 8 * 2^-24 the threshold assumes
    rel = np.abs(lf.astype(np.float64) - lam) / lam
    assert rel.max() < 8 * 2.0 ** -24
