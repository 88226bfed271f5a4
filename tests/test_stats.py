"""The post-eigendecomposition statistics of the drop-in (eig_amd/host/stats.c), on the
CPU: Tracy-Widom, F and chi-square tails, anova, kurtosis and Fst with its block
jackknife.

Pins, none of which restates stats.c: the reference's own twstats output for its
example eigenvalues (POPGEN/twexample.out, byte for byte), its published TW table
(POPGEN/twtable), scipy's F, chi-square, one-way anova and kurtosis, and Fst values
worked by hand or by a direct delete-one-block computation.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
from scipy import stats as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "eig_amd", "libeig_hoststats.so")
TW = os.path.join(ROOT, "eig_amd", "bin", "twstats")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


@pytest.fixture(scope="module")
def L():
    lib = C.CDLL(LIB)
    for name, res, args in (
            ("st_ftail", C.c_double, [C.c_int, C.c_int, C.c_double]),
            ("st_chitail", C.c_double, [C.c_int, C.c_double]),
            ("st_twtail", C.c_double, [C.c_double]),
            ("st_twtable", C.c_int, [f64p, f64p, f64p]),
            ("st_kurtosis", C.c_double, [f64p, C.c_int]),
            ("st_anova", C.c_double, [f64p, C.c_int, i32p, C.c_int]),
            ("st_anova1", C.c_double, [f64p, C.c_int, i32p, C.c_int]),
            ("st_setblocks", C.c_int, [i32p, f64p, u8p, C.c_int64, C.c_double, i32p]),
            ("st_fst", None, [i32p, u8p, i32p, C.c_int64, C.c_int, C.c_int, f64p, f64p, f64p,
                              i32p])):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def test_twstats_reproduces_the_reference_example(golden, tmp_path):
    out = tmp_path / "out.txt"
    subprocess.run([TW, "-t", "twtable", "-i", os.path.join(golden, "popgen", "twexample.eval"),
                    "-o", str(out)], check=True, cwd=tmp_path)
    assert out.read_text() == open(os.path.join(golden, "popgen", "twexample.out")).read()


def test_tracy_widom_table_from_painleve_ii(L, golden):
    x = np.zeros(161); t = np.zeros(161); p = np.zeros(161)
    assert L.st_twtable(x, t, p) == 161
    ref = np.loadtxt(os.path.join(golden, "popgen", "twtable"), comments="#")
    assert np.array_equal(x, ref[:, 0])
    # both are rounded to 9 decimals; the reference's own solver sits within 1.5e-10
    assert np.max(np.abs(t - ref[:, 1])) <= 1.0001e-9
    assert np.max(np.abs(p - ref[:, 2])) <= 1.0001e-9
    assert np.count_nonzero(t != ref[:, 1]) <= 10
    # known values of the TW1 distribution: mean -1.2065, F1(0) = 0.8319 (Tracy-Widom 1996)
    assert abs((1 - L.st_twtail(0.0)) - 0.831908) < 2e-6
    xs = np.linspace(-8, 8, 16001)
    dens = -np.gradient(np.array([L.st_twtail(v) for v in xs]), xs)
    assert abs(np.trapezoid(xs * dens, xs) - (-1.2065335745820)) < 1e-4


@pytest.mark.parametrize("d1,d2", [(1, 1), (1, 7), (2, 9), (3, 40), (9, 4), (24, 4975)])
def test_f_tail_matches_scipy(L, d1, d2):
    for F in (0.01, 0.3, 1.0, 2.5, 9.0, 60.0):
        want = S.f.sf(F, d1, d2)
        got = L.st_ftail(d1, d2, F)
        assert abs(got - want) <= 1e-12 + 1e-10 * want, (F, got, want)


@pytest.mark.parametrize("df", [1, 2, 3, 4, 10, 11, 30])
def test_chisquare_tail_matches_scipy(L, df):
    for z in (0.05, 1.0, 3.7, 9.554, 40.0, 300.0):
        want = S.chi2.sf(z, df)
        got = L.st_chitail(df, z)
        assert abs(got - want) <= 1e-300 + 1e-10 * want, (z, got, want)


def test_anova_matches_scipy_and_the_reference_log(L):
    rng = np.random.default_rng(1)
    v = rng.standard_normal(90)
    xt = np.repeat(np.arange(3, dtype=np.int32), 30)
    v[xt == 2] += 0.4
    want = S.f_oneway(v[xt == 0], v[xt == 1], v[xt == 2]).pvalue
    assert abs(L.st_anova(v, 90, xt, 3) - want) <= 1e-12
    # anova1: the likelihood ratio statistic 0.5 n log(SS_total / SS_within)
    sst = np.sum((v - v.mean()) ** 2)
    ssw = sum(np.sum((v[xt == k] - v[xt == k].mean()) ** 2) for k in range(3))
    assert abs(L.st_anova1(v, 90, xt, 3) - 0.5 * 90 * np.log(sst / ssw)) <= 1e-9


def test_kurtosis_matches_scipy(L):
    a = np.random.default_rng(2).standard_normal(5000) ** 3
    assert abs(L.st_kurtosis(a, 5000) - S.kurtosis(a, fisher=False)) <= 1e-10


def test_fst_hand_worked_and_jackknife(L):
    # two populations, three SNPs, counts (first allele, second allele)
    cnt = np.array([[[6, 4], [2, 8]],
                    [[5, 5], [5, 5]],
                    [[10, 0], [0, 10]]], np.int32)
    ign = np.zeros(3, np.uint8)
    tag = np.array([0, 1, 2], np.int32)
    mean = np.zeros(4); est = np.zeros(4); sig = np.zeros(4); num = np.zeros(4, np.int32)
    L.st_fst(cnt.reshape(-1), ign, tag, 3, 2, 3, mean, est, sig, num)

    def nd(a, b, c, d):  # Hudson-style numerator and denominator of one SNP
        z, zz = a + b, c + d
        p1, p2 = a / z, c / zz
        h1, h2 = a * b / (z * (z - 1)), c * d / (zz * (zz - 1))
        n_ = (p1 - p2) ** 2 - h1 / z - h2 / zz
        return n_, n_ + h1 + h2
    parts = [nd(*cnt[m].reshape(-1)) for m in range(3)]
    N = np.array([p[0] for p in parts]); D = np.array([p[1] for p in parts])
    fst = N.sum() / (D.sum() + 1e-10)
    assert abs(mean[1] - fst) <= 1e-12 and num[1] == 3
    # delete-one-block estimates and the weighted jackknife (equal block sizes here)
    jm = np.array([(N.sum() - N[b]) / (D.sum() - D[b] + 1e-10) for b in range(3)])
    g = 3
    jk = g * fst - (g - 1) * jm.mean()  # the textbook delete-one jackknife
    var = np.sum((g * fst - (g - 1) * jm - jk) ** 2 / (g - 1)) / g
    assert abs(est[1] - jk) <= 1e-12 and abs(sig[1] - np.sqrt(var)) <= 1e-12
    assert est[2] == est[1] and mean[0] == 0.0


def test_setblocks_by_chromosome_and_distance(L):
    chrom = np.array([1, 1, 1, 1, 2, 2, 2], np.int32)
    gpos = np.array([0.0, 0.01, 0.06, 0.07, 0.0, 0.049, 0.051])
    ign = np.array([0, 0, 0, 1, 0, 0, 0], np.uint8)
    tag = np.zeros(7, np.int32)
    assert L.st_setblocks(chrom, gpos, ign, 7, 0.05, tag) == 4
    assert list(tag) == [0, 0, 1, -1, 2, 2, 3]
