"""Writes tests/golden/decimal_laws.json: high-precision (Decimal) values of the point laws the
tissue step uses, from their published formulas, independent of both the oracle and the kernels.

- constant-field (GHK) route flux, positive from -> to, fmol/s: A P / 1e-15 x (c_from e^x - c_to)
  / (e^x - 1), x = z F dv / RT (the formula of validation/betse/junction_transport/check.py:10-23
  in the reference);
- junction gate rates (BETSE v1.5.0's rate equations as validation/betse/junction_gate/
  reference.py:9-20 of the reference states them): alpha = 1000 * 0.0013 exp(-0.077 s),
  beta = 1000 b / (1 + 50 b), b = 0.0013 exp(0.14 s), s = (|v| - threshold) in mV;
- the Na/K-ATPase mass-action magnitude cap x saturation x (1 - e^a).
Run: python tests/golden/make_decimal_vectors.py
"""

import json
import os
from decimal import Decimal, localcontext

T = 310.0
FARADAY = 6.02214076e23 * 1.602176634e-19
GAS = 6.02214076e23 * 1.380649e-23


def D(v):
    return Decimal(repr(float(v)))


def ghk(ca, cb, dv, z, pa):
    with localcontext() as c:
        c.prec = 70
        x = D(z) * Decimal(FARADAY) * D(dv) / (Decimal(GAS) * D(T))
        if x == 0:
            drive = D(ca) - D(cb)
        else:
            e = x.exp()
            drive = x * (D(ca) * e - D(cb)) / (e - 1)
        return float(D(pa) * drive * Decimal("1e15"))


def gate(v, threshold):
    with localcontext() as c:
        c.prec = 65
        off = (abs(D(v)) - D(threshold)) * 1000
        a = Decimal("0.0013") * (-Decimal("0.077") * off).exp()
        b = Decimal("0.0013") * (Decimal("0.14") * off).exp()
        return float(1000 * a), float(1000 * b / (1 + 50 * b))


def main():
    flux = []
    for z in (1.0, -1.0, 2.0):
        for dv in (-0.12, -0.05, -1e-3, -3e-6, 0.0, 2e-6, 4e-3, 0.07, 0.12):
            for ca, cb in ((12.0, 145.0), (140.0, 5.0), (1e-4, 2.0)):
                flux.append(dict(c_from=ca, c_to=cb, dv=dv, z=z, pa=4.7e-18, flux=ghk(ca, cb, dv, z, 4.7e-18)))
    gates = []
    for thr in (0.0, 0.015, 0.06, 0.12):
        for i in range(25):
            v = -0.12 + 0.01 * i
            a, b = gate(v, thr)
            gates.append(dict(v=v, threshold=thr, alpha=a, beta=b))
    out = dict(temperature=T, ghk=flux, gate=gates)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decimal_laws.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(path, len(flux), len(gates))


if __name__ == "__main__":
    main()
