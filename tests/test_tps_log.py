"""The natural logarithm used by the TPS kernels (ln_pos in
fluidimage_amd/csrc/piv_tps.hip): its reduction and series, evaluated in
numpy with the kernel's operation order (fused multiply-adds through 80-bit
long doubles), against numpy's log on values spanning the kernel's inputs
(squared distances of 1e-6 .. 1e9 px^2) and the reduction's edges.

A mistake in a coefficient, in the two-part ln 2 or in the interval
reduction gives errors far above the 4 ulp bar."""

import numpy as np

LN2_HI = float.fromhex("0x1.62e42fefa3800p-1")
LN2_LO = float.fromhex("0x1.ef35793c76730p-45")


def _fma(a, b, c):
    ld = np.longdouble
    return np.float64(ld(a) * ld(b) + ld(c))


def ln_pos(x):
    m, e = np.frexp(x)
    low = m < 0.70710678118654752440
    m = np.where(low, 2.0 * m, m)
    e = np.where(low, e - 1, e)
    s = (m - 1.0) * (1.0 / (m + 1.0))
    z = s * s
    p = np.full_like(x, 1.0 / 19.0)
    for k in (17, 15, 13, 11, 9, 7, 5, 3):
        p = _fma(p, z, 1.0 / k)
    t = 2.0 * s
    de = e.astype(np.float64)
    return _fma(de, LN2_HI, _fma(de, LN2_LO, _fma(t * z, p, t)))


def test_ln2_split():
    from decimal import Decimal, getcontext

    getcontext().prec = 50
    assert abs(Decimal(LN2_HI) + Decimal(LN2_LO) - Decimal(2).ln()) < Decimal(2) ** -100
    # 42 significant bits: e * LN2_HI is exact for |e| < 2048
    mant = int(float.hex(LN2_HI).split(".")[1].split("p")[0], 16)
    assert mant & 0x7FF == 0


def test_ln_pos_within_4_ulp():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        10.0 ** rng.uniform(-6, 9, 200_000),
        np.arange(1, 20001, dtype=np.float64),  # integer squared distances
        np.nextafter(np.sqrt(0.5), 0) * 2.0 ** np.arange(-20, 30),
        np.sqrt(0.5) * 2.0 ** np.arange(-20, 30),
        1.0 + np.linspace(-1e-3, 1e-3, 2001),
    ])
    got = ln_pos(x)
    ref = np.log(x)
    ulp = np.spacing(np.abs(ref))
    err = np.abs(got - ref) / ulp
    # at x = 1 both are 0
    err[ref == 0] = np.abs(got[ref == 0]) / np.spacing(1.0)
    assert err.max() <= 4.0, (err.max(), x[np.argmax(err)])
