"""Synthetic inputs of the benchmark configurations (host-side input generation).

Every draw comes from numpy's Philox counter-based generator keyed by the configuration
seed, in a fixed order (links, then E field, then spinors), so tests, the bench and the
oracle see identical inputs.  Layouts are the reference's:

* SU(2) links: 4 reals per (site, mu) at ``4*(mu + 3*site)``, ``U = u3 + i u_a sigma_a``
  (SU2/GroupOperations.cpp:4-8); ``U = exp(i eps n.sigma/2)`` gives
  ``u = (sin(eps/2) n, cos(eps/2))`` (SU2/AlgebraOperations.cpp:37-55 with c = 1).
* SU(3) links: 9 complex per (site, mu), element (i, j) at ``i + 3j``
  (SU3/GroupOperations.cpp:28-30); ``U = exp(i eps sum_a n_a lambda_a/2)``.
* E field: ``E[a + (Nc^2-1)*(mu + 3*site)]`` (ElectricFields.cpp:36-41).
* spinors: per mode AoS ``[site][colour][spin]`` complex (FermionField.cpp:8-9, 537-545).
"""
from __future__ import annotations

import dataclasses

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def gell_mann_generators() -> np.ndarray:
    """t^a = lambda_a / 2, a = 0..7 (SU3/AlgebraOperations.cpp:10-25)."""
    t = np.zeros((8, 3, 3), complex)
    t[0][0, 1] = t[0][1, 0] = 1
    t[1][0, 1], t[1][1, 0] = -1j, 1j
    t[2][0, 0], t[2][1, 1] = 1, -1
    t[3][0, 2] = t[3][2, 0] = 1
    t[4][0, 2], t[4][2, 0] = -1j, 1j
    t[5][1, 2] = t[5][2, 1] = 1
    t[6][1, 2], t[6][2, 1] = -1j, 1j
    t[7][0, 0] = t[7][1, 1] = 1 / np.sqrt(3)
    t[7][2, 2] = -2 / np.sqrt(3)
    return t / 2


def pauli_generators() -> np.ndarray:
    s = np.zeros((3, 2, 2), complex)
    s[0][0, 1] = s[0][1, 0] = 1
    s[1][0, 1], s[1][1, 0] = -1j, 1j
    s[2][0, 0], s[2][1, 1] = 1, -1
    return s / 2


def su2_links(n, sigma, g: np.random.Generator) -> np.ndarray:
    V = int(np.prod(n))
    eps = g.normal(0.0, sigma, size=(V, 3))
    v = g.normal(size=(V, 3, 3))
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    u = np.empty((V, 3, 4))
    u[..., :3] = np.sin(eps / 2)[..., None] * v
    u[..., 3] = np.cos(eps / 2)
    return u.reshape(-1)


def su3_exp_i(alpha: np.ndarray) -> np.ndarray:
    """exp(i alpha_a t^a) for a batch of 8-vectors, as row-major 3x3 matrices."""
    t = gell_mann_generators()
    h = np.einsum("...a,aij->...ij", alpha, t)
    w, vec = np.linalg.eigh(h)
    return np.einsum("...ij,...j,...kj->...ik", vec, np.exp(1j * w), vec.conj())


def su3_links(n, sigma, g: np.random.Generator) -> np.ndarray:
    V = int(np.prod(n))
    eps = g.normal(0.0, sigma, size=(V, 3))
    v = g.normal(size=(V, 3, 8))
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    m = su3_exp_i(eps[..., None] * v)  # (V,3,3,3) row-major
    cm = np.swapaxes(m, -1, -2).reshape(V, 3, 9)  # element (i,j) at i+3j
    out = np.empty((V, 3, 18))
    out[..., 0::2] = cm.real
    out[..., 1::2] = cm.imag
    return out.reshape(-1)


def links_to_matrices(nc: int, links: np.ndarray) -> np.ndarray:
    """Reference link format -> (V, 3, nc, nc) row-major complex matrices."""
    if nc == 2:
        u = links.reshape(-1, 3, 4)
        m = np.empty(u.shape[:2] + (2, 2), complex)
        m[..., 0, 0] = u[..., 3] + 1j * u[..., 2]
        m[..., 0, 1] = u[..., 1] + 1j * u[..., 0]
        m[..., 1, 0] = -u[..., 1] + 1j * u[..., 0]
        m[..., 1, 1] = u[..., 3] - 1j * u[..., 2]
        return m
    c = links.reshape(-1, 3, 9, 2)
    c = c[..., 0] + 1j * c[..., 1]
    return np.swapaxes(c.reshape(-1, 3, 3, 3), -1, -2)


def efield(n, nc, sigma, g: np.random.Generator) -> np.ndarray:
    V = int(np.prod(n))
    return g.normal(0.0, sigma, size=V * 3 * (nc * nc - 1))


def spinors(n, nc, nmodes, g: np.random.Generator) -> np.ndarray:
    """nmodes unit-norm complex Gaussian mode functions, AoS, as interleaved float64."""
    V = int(np.prod(n))
    z = g.normal(size=(nmodes, V * nc * 4)) + 1j * g.normal(size=(nmodes, V * nc * 4))
    z /= np.linalg.norm(z, axis=1, keepdims=True)
    return z


@dataclasses.dataclass
class Config:
    name: str
    nc: int
    n: tuple
    nmodes: int
    steps: int
    dt: float
    link_sigma: float
    e_sigma: float
    seed: int
    kind: str  # "wilson" or "overlap"
    m_fermion: float = 0.0
    r_wilson: float = 1.0
    m_dw: float = 0.0
    nb: int = 0
    q: float = 1.0
    cheb_n: int = 0
    cheb_cut: float = 0.05

    @property
    def volume(self) -> int:
        return int(np.prod(self.n))

    def generate(self, nmodes: int | None = None):
        g = rng(self.seed)
        links = su2_links(self.n, self.link_sigma, g) if self.nc == 2 else su3_links(self.n, self.link_sigma, g)
        e = efield(self.n, self.nc, self.e_sigma, g)
        psi = spinors(self.n, self.nc, self.nmodes if nmodes is None else nmodes, g)
        return links, e, psi


# The five workloads.  m_dw is the kernel's M (MDWHeight); for the plain Wilson method M = 0
# so that the kernel mass is m (the reference folds -2M into the Wilson mass,
# WilsonOperator.cpp:39).  The overlap rho is m_dw.
CONFIGS = {
    "su2_wilson_8": Config("su2_wilson_8", 2, (8, 8, 8), 16, 100, 0.05, 0.3, 0.1, 1, "wilson",
                           m_fermion=0.1, m_dw=0.0),
    "su2_overlap_16": Config("su2_overlap_16", 2, (16, 16, 16), 64, 50, 0.02, 0.3, 0.1, 2, "overlap",
                             m_dw=1.0, cheb_n=256, cheb_cut=0.05),
    "su3_wilson_32": Config("su3_wilson_32", 3, (32, 32, 32), 512, 200, 0.05, 0.25, 0.1, 3, "wilson",
                            nb=4, q=1.0, m_dw=0.0),
    "su3_overlap_24": Config("su3_overlap_24", 3, (24, 24, 24), 256, 100, 0.02, 0.25, 0.1, 3, "overlap",
                             nb=8, q=1.0, m_dw=1.2, cheb_n=384, cheb_cut=0.05),
}


@dataclasses.dataclass(frozen=True)
class EnsembleConfig:
    """Many independent SU(2) overlap systems of uneven sizes, one seed each.

    System k uses seed ``seed0 + k``: its generator first draws the lattice size L from
    ``sizes`` and the flux quanta Nb from ``0..nb_max``, then links, E field and spinors in the
    order of Config.generate.
    """
    name: str
    n_systems: int
    seed0: int
    sizes: tuple
    nb_max: int
    nmodes: int
    steps: int
    dt: float
    link_sigma: float
    e_sigma: float
    m_dw: float
    cheb_n: int
    cheb_cut: float
    nc: int = 2
    m_fermion: float = 0.0
    r_wilson: float = 1.0
    q: float = 1.0

    def system(self, k: int, nmodes: int | None = None):
        """(dims, nb, links, e, psi) of system k."""
        g = rng(self.seed0 + k)
        L = int(self.sizes[int(g.integers(len(self.sizes)))])
        nb = int(g.integers(0, self.nb_max + 1))
        n = (L, L, L)
        links = su2_links(n, self.link_sigma, g) if self.nc == 2 else su3_links(n, self.link_sigma, g)
        e = efield(n, self.nc, self.e_sigma, g)
        psi = spinors(n, self.nc, self.nmodes if nmodes is None else nmodes, g)
        return n, nb, links, e, psi

    def layout(self):
        """(dims, nb) of every system without drawing the fields."""
        out = []
        for k in range(self.n_systems):
            g = rng(self.seed0 + k)
            L = int(self.sizes[int(g.integers(len(self.sizes)))])
            out.append(((L, L, L), int(g.integers(0, self.nb_max + 1))))
        return out


# 32 SU(2) overlap systems, L in {12, 16, 20, 24}, seeds 100-131, 32 modes each, Nb in 0..6,
# 50 steps at dt = 0.02, links and E as the 8^3 Wilson workload, kernel as the 16^3 overlap one
ENSEMBLES = {
    "su2_overlap_ensemble": EnsembleConfig("su2_overlap_ensemble", 32, 100, (12, 16, 20, 24), 6, 32, 50, 0.02,
                                           0.3, 0.1, 1.0, 256, 0.05),
}
