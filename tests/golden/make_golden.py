"""Writes the golden fixtures of tests/golden/ from the compiled reference hot path
(oracle/_ref/libovlref.so, built by oracle/Makefile from the reference's own source files).

Every fixture holds inputs and the reference's outputs only.  Run from the repository root:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import orc  # noqa: E402
from overlap_amd import configs as cf  # noqa: E402


def main():
    if not orc.ref_available():
        raise SystemExit("oracle/_ref/libovlref.so is not built (make -C oracle ref)")
    n = (4, 4, 4)
    V = 64
    g = cf.rng(20240601)
    links = cf.su2_links(n, 0.3, g)
    phi = cf.spinors(n, 2, 2, g)
    # Wilson stencil, two parameter sets (WilsonOperator.cpp:30)
    L1 = orc.lat(2, n, 1.0, 0.1, 1.4)
    L2 = orc.lat(2, n, 1.0, 0.0, 0.0)
    ucc = orc.links_to_ucc(L1, links)
    np.savez(os.path.join(HERE, "wilson_su2_4.npz"), n=np.array(n), links=links, ucc=ucc, phi=phi,
             params=np.array([[1.0, 0.1, 1.4], [1.0, 0.0, 0.0]]),
             psi=np.stack([orc.ref_wilson_evo(L1, ucc, phi[0]), orc.ref_wilson_evo(L2, ucc, phi[1])]))
    # Zolotarev coefficients (ZolotarevCoefficients.cpp:71)
    ranges = np.array([[0.1, 25.0], [0.01, 30.0], [0.5, 4.0], [0.05, 23.0], [1e-3, 40.0]])
    eps, bzs, dzs = [], [], []
    for amin, amax in ranges:
        e, bz, dz = orc.ref_zolotarev(amin, amax)
        eps.append(e); bzs.append(bz); dzs.append(dz)
    np.savez(os.path.join(HERE, "zolotarev.npz"), ranges=ranges, eps=np.array(eps), bz=np.array(bzs), dz=np.array(dzs))
    # overlap operator (OverlapOperator.cpp:25 minus the unused random vector), M = 1.0, m = 0.1
    L3 = orc.lat(2, n, 1.0, 0.1, 1.0)
    lmin, lmax = 0.2, 30.0
    np.savez(os.path.join(HERE, "overlap_su2_4.npz"), n=np.array(n), ucc=ucc, phi=phi, params=np.array([1.0, 0.1, 1.0]),
             interval=np.array([lmin, lmax]),
             psi=np.stack([orc.ref_overlap_evo(L3, ucc, p, lmin, lmax) for p in phi]))
    # Ritz extreme eigenvalues of +-h_W^2 (RitzSolver.cpp:92)
    start = orc.aos_to_flat(2, V, phi[1])
    np.savez(os.path.join(HERE, "ritz_su2_4.npz"), n=np.array(n), ucc=ucc, start=start, params=np.array([1.0, 0.1, 1.0]),
             eig=np.array([orc.ref_ritz(L3, ucc, start, sign=+1.0), orc.ref_ritz(L3, ucc, start, sign=-1.0)]))
    # SU(2) exponential and product (SU2/AlgebraOperations.cpp:37, SU2/GroupOperations.cpp:103)
    cs = g.normal(size=8) * 0.3
    al = g.normal(size=(8, 3))
    us = np.stack([links[4 * k:4 * k + 4] for k in range(8)])
    np.savez(os.path.join(HERE, "su2_algebra.npz"), c=cs, alpha=al, u=us,
             iexp=np.stack([orc.ref_su2_iexp(c, a) for c, a in zip(cs, al)]),
             uu=np.stack([orc.ref_su2_uu(orc.ref_su2_iexp(c, a), u) for c, a, u in zip(cs, al, us)]))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
