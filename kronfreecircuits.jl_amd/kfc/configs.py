"""The benchmark workloads, seeded and with every parameter stated.

two_transmons   : test/fixtures/two_transmons.json at E_J1=30, E_C1=1.2, n_g1=0.25,
                  E_J2=25, E_C2=1.0, n_g2=0, g=0.3 (test/helpers.jl:111); ChargeBasis(30)
                  leaves keep 8, the top keeps 20.
three_fluxonia  : three PhaseGrid(-6pi, 6pi, 512; order=2) modes in one flat node,
                  4 E_C n^2 + E_L phi^2/2 - E_J cos(phi - phi_ext) per mode, coupling
                  0.05 n_i n_j per pair; E_C ~ U[0.8,1.2], E_L ~ U[0.3,1.0], E_J ~ U[3,8],
                  phi_ext ~ U[0, 2pi); Lanczos, lowest 10.
zero_pi_tree    : the full zero-pi spec (test/fixtures/full_zero_pi.json) with
                  symmetric capacitances (dC_J = dC = 0, so every node is real),
                  E_J = 10, E_L = 0.05, E_CJ = 20, E_CS ~ U[0.02, 0.2] (the charging
                  energy that ranges), E_Cz = 1/(1/E_CS - 1/E_CJ), n_g ~ U[0,1],
                  dE_J = 0.05, dE_L = 0.1, varphi_ext = 2pi*0.23; leaves theta
                  ChargeBasis(40), phi PhaseGrid(-6pi, 6pi, 256; order=8), zeta
                  FockBasis(30; phi_zpf=(2E_Cz/E_L)^(1/4)) keep 40, 40, 30; the parent
                  [40, 40, 30] runs LOBPCG for the lowest 16 (guard 4).
"""
from __future__ import annotations

import math

import numpy as np

from .bases import ChargeBasis, FockBasis, PeriodicGrid, PhaseGrid
from .operator import KronFreeOperator
from .solvers import LOBPCG, Lanczos
from .tree import Node, NumericSpec, circuit_terms, constant_term, node_axisterms

TWO_TRANSMONS_THETA = dict(E_J1=30.0, E_C1=1.2, n_g1=0.25, E_J2=25.0, E_C2=1.0, n_g2=0.0, g=0.3)


def two_transmons_spec(E_J1, E_C1, n_g1, E_J2, E_C2, n_g2, g):
    """test/fixtures/two_transmons.json."""
    return NumericSpec(["varphi_1", "varphi_2"], [[E_C1, g / 8], [g / 8, E_C2]], np.zeros((2, 2)), [0.0, 0.0],
                       [n_g1, n_g2], [(E_J1, [1, 0], 0.0), (E_J2, [0, 1], 0.0)])


def two_transmons_tree(ncut=30, keep_leaf=8, keep_top=20, solver=None):
    leaves = [Node([(f"varphi_{i}", ChargeBasis(ncut))], keep=keep_leaf) for i in (1, 2)]
    return Node(leaves, keep=keep_top, solver=solver)


def three_fluxonia_theta(seed):
    r = np.random.default_rng(seed)
    return dict(E_C=r.uniform(0.8, 1.2, 3), E_L=r.uniform(0.3, 1.0, 3), E_J=r.uniform(3.0, 8.0, 3),
                phi_ext=r.uniform(0.0, 2 * math.pi, 3))


def three_fluxonia_spec(E_C, E_L, E_J, phi_ext, g=0.05):
    EC = np.full((3, 3), g / 8)
    for i in range(3):
        EC[i, i] = E_C[i]
    return NumericSpec(["phi1", "phi2", "phi3"], EC, np.diag(E_L), [0.0] * 3, [0.0] * 3,
                       [(E_J[i], [int(i == j) for j in range(3)], -phi_ext[i]) for i in range(3)])


def three_fluxonia_node(N=512, order=2, span=6 * math.pi, m=10, solver=None):
    return Node([(f"phi{i}", PhaseGrid(-span, span, N, order)) for i in (1, 2, 3)], keep=m, compute=m,
                solver=solver or Lanczos())


def three_fluxonia_operator(theta, N=512, order=2, span=6 * math.pi, device=0, ctx=None):
    """The flat three-fluxonia node's real operator (one KronFreeOperator), on
    device `device` or on the context `ctx`."""
    spec = three_fluxonia_spec(theta["E_C"], theta["E_L"], theta["E_J"], theta["phi_ext"])
    node = three_fluxonia_node(N, order, span)
    named = circuit_terms(spec)
    ts = node_axisterms(node, [None] * 3, named, real=True)
    if ctx is not None:
        return KronFreeOperator([N] * 3, ts, constant=constant_term(spec), ctx=ctx)
    return KronFreeOperator([N] * 3, ts, constant=constant_term(spec), device=device)


def zero_pi_theta(seed):
    r = np.random.default_rng(seed)
    return dict(E_CS=float(r.uniform(0.02, 0.2)), n_g=float(r.uniform(0.0, 1.0)))


def zero_pi_spec(E_CS, n_g, E_J=10.0, E_L=0.05, E_CJ=20.0, dE_J=0.05, dE_L=0.1, varphi_ext=2 * math.pi * 0.23):
    """test/fixtures/full_zero_pi.json with dC_J = dC = 0."""
    E_Cz = 1.0 / (1.0 / E_CS - 1.0 / E_CJ)
    s2 = math.sqrt(2.0)
    E_C = [[E_CS / 2, 0.0, 0.0], [0.0, E_CJ / 2, 0.0], [0.0, 0.0, E_Cz]]
    E_Lm = [[0.0, 0.0, 0.0], [0.0, 2 * E_L, E_L * dE_L / s2], [0.0, E_L * dE_L / s2, E_L]]
    x = varphi_ext / 2
    cos = [(E_J, [1, 1, 0], -x), (E_J, [1, -1, 0], x), (E_J * dE_J / 2, [1, 1, 0], -x),
           (-E_J * dE_J / 2, [1, -1, 0], x)]
    const = 2 * E_J - math.sqrt(2 * E_Cz * E_L)
    return NumericSpec(["theta", "phi", "zeta"], E_C, E_Lm, [0.0] * 3, [n_g, 0.0, 0.0], cos, const), E_Cz


def zero_pi_tree(E_Cz, E_L=0.05, ncut=40, Nphi=256, Nzeta=30, keep=40, m=16, solver=None):
    th = Node([("theta", ChargeBasis(ncut))], keep=keep)
    ph = Node([("phi", PhaseGrid(-6 * math.pi, 6 * math.pi, Nphi, 8))], keep=keep)
    ze = Node([("zeta", FockBasis(Nzeta, (2 * E_Cz / E_L) ** 0.25))], keep=min(keep, Nzeta))
    return Node([th, ph, ze], keep=m, compute=m, solver=solver or LOBPCG())

def periodic_transmon_leaf(N, order=8):
    return PeriodicGrid(N, order)
