"""GaussOnSphere spline tables of the SCBragg model (CPU checker; the GPU
process builds bit-identical tables, tests/test_gpu_scbragg.py) pinned by
their definitions (ncrystal_core/src/phys_utils/NCGaussOnSphere.cc:51-62,
149-189, 255-378):

* normalisation 1/(2 pi int_0^min(T, 20 sigma) sin(t) exp(-t^2/(2 sigma^2)) dt)
  against adaptive quadrature;
* knot values of the density table, norm exp(-acos(x)^2 / (2 sigma^2)), and
  of the circle-integral table, norm sqrt(2 pi) sigma exp(-d^2/(2 sigma^2))
  erf(sqrt((T^2 - d^2) / (2 sigma^2))) with d = acos(x), in closed form;
* second derivatives: the interior equations of the cubic spline, and the
  interpolated values between knots against the closed forms.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

DEMI = np.array([[2, 2, 0]], dtype=np.int32)


def _gos(oracle, mos_arcmin, prec):
    a = 5.65735
    d = a / math.sqrt(8.0)
    mos = mos_arcmin * math.pi / (180 * 60)
    o = oracle.scbragg(a, a, a, a ** 3 * 8, np.array([d]), np.array([40.0]), np.array([0, 1]), DEMI, np.eye(3),
                       mos, prec)
    return o.info()["gos"], o.tables()


@pytest.mark.parametrize("mos_arcmin,prec", [(30.0, 1e-3), (60.0, 1e-2), (5.0, 1e-4), (120.0, 3e-3)])
def test_gaussonsphere_tables(oracle, mos_arcmin, prec):
    gos, t = _gos(oracle, mos_arcmin, prec)
    sigma, trunc, expfact, norm, nlt = gos[0], gos[1], gos[5], gos[6], int(gos[10])
    assert expfact == pytest.approx(-0.5 / sigma ** 2, rel=1e-15)
    upper = min(trunc, 20 * sigma)
    val, err = integrate.quad(lambda x: math.sin(x) * math.exp(-x * x / (2 * sigma * sigma)), 0.0, upper,
                              epsabs=0.0, epsrel=1e-13, limit=200)
    assert norm == pytest.approx(1.0 / (2 * math.pi * val), rel=1e-10)

    def f_dens(x):
        return norm * np.exp(expfact * np.arccos(np.clip(x, -1, 1)) ** 2)

    def f_circ(x):
        d = np.arccos(np.clip(x, -1, 1))
        return (norm * math.sqrt(2 * math.pi) * sigma * np.exp(expfact * d * d)
                * special.erf(np.sqrt(np.maximum(0.0, -expfact * (trunc * trunc - d * d)))))

    for tab, a, invd, f in ((t["lt_evalcosx"], gos[11], gos[12], f_dens), (t["lt_sofcosd"], gos[13], gos[14], f_circ)):
        y, y2 = tab[:, 0], tab[:, 1]
        assert len(y) == nlt
        x = a + np.arange(nlt) / invd
        ref = f(x)
        scale = np.abs(ref).max()
        assert np.max(np.abs(y - ref)) <= 1e-12 * scale
        # the table stores y'' h^2 (its evaluation has no h): interior spline
        # equations y2[i-1] + 4 y2[i] + y2[i+1] = 6 (y[i+1] - 2 y[i] + y[i-1])
        h = 1.0 / invd
        lhs = y2[:-2] + 4 * y2[1:-1] + y2[2:]
        rhs = 6 * (y[2:] - 2 * y[1:-1] + y[:-2])
        assert np.max(np.abs(lhs - rhs)) <= 1e-8 * np.abs(rhs).max()
        # the spline between knots follows the function
        xm = x[:-1] + 0.5 * h
        b = 0.5
        sm = 0.5 * y[:-1] + 0.5 * y[1:] + ((b ** 3 - b) * y2[:-1] + (b ** 3 - b) * y2[1:]) / 6
        assert np.max(np.abs(sm - f(xm))) <= max(prec, 1e-4) * scale
