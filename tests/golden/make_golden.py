"""Writes tests/golden/closed_forms.json: known answers derived from
mathematics only (exact rational arithmetic), used to pin the oracle.

The reference repository ships no tests, fixtures or data, and its sources
do not build in this image (header.h:7-12 includes GSL and FFTW3 headers
that are not installed), so every fixture here comes from first principles:

  * free-surface one-sided weights (fwi_2d.c:805-807, 848): derivatives of
    the interpolating polynomial through the nodes the rows use, with
    p(-1/2) = 0 imposed at the surface for the v rows;
  * CPML coefficient closed forms (fwi_2d.c:120-131) in exact decimal
    arithmetic for the config-1 geometry;
  * the staggered-derivative consistency constant c0 + 3 c1 (fwi_2d.c:307-308).

Run:  python tests/golden/make_golden.py
"""
from fractions import Fraction as F
import json
import math
import os


def lagrange_derivative_weights(nodes, at):
    """Weights w_k with q'(at) = sum_k w_k f(nodes[k]) for the interpolant q."""
    ws = []
    for k, xk in enumerate(nodes):
        # derivative of the k-th Lagrange basis polynomial at `at`
        others = [xj for j, xj in enumerate(nodes) if j != k]
        denom = F(1)
        for xj in others:
            denom *= (xk - xj)
        tot = F(0)
        for m in range(len(others)):
            term = F(1)
            for j, xj in enumerate(others):
                if j != m:
                    term *= (at - xj)
            tot += term
        ws.append(tot / denom)
    return ws


def main():
    out = {}
    half = F(1, 2)
    # v rows: p nodes at z = 0,1,2,3 (cell units), plus p(-1/2) = 0 at the surface.
    nodes = [-half, F(0), F(1), F(2), F(3)]
    w0 = lagrange_derivative_weights(nodes, -half)[1:]  # drop the zero node
    w1 = lagrange_derivative_weights(nodes, half)[1:]
    out["v_row0_times_840"] = [str(w * 840) for w in w0]
    out["v_row1_times_840"] = [str(w * 840) for w in w1]
    # vz row 0: v nodes at z = -1/2, 1/2, 3/2, 5/2, 7/2, derivative at z = 0
    vnodes = [-half, half, F(3, 2), F(5, 2), F(7, 2)]
    wz = lagrange_derivative_weights(vnodes, F(0))
    out["vz_row0_times_24"] = [str(w * 24) for w in wz]
    # staggered 4th order Taylor weights for comparison and c0 + 3 c1
    out["c0_plus_3c1_times_dx"] = str(F("1.129042") + 3 * F("-0.04301412"))
    # CPML closed forms for config 1: npml=20, dx=10, dt=1e-3, Vmax=6000, R=1e-3
    npml, dx, dt, vmax, R, f0 = 20, 10.0, 1e-3, 6000.0, 1e-3, 15.0
    L = npml * dx
    d0 = -(2 + 1) * vmax * math.log(R) / (2.0 * L)
    out["cpml_config1"] = {
        "d0": d0,
        "b_outer": math.exp(-d0 * dt),
        "a_outer": (d0 / d0) * (math.exp(-d0 * dt) - 1.0),
        "b_inner": math.exp(-(math.pi * f0) * dt),
        "a_inner": 0.0,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "closed_forms.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(path)


if __name__ == "__main__":
    main()
