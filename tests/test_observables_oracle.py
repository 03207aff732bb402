"""Pins the oracle's log-step observables: calculate_F2_self!
(src/AuxiliaryFunctions.jl:968-987) and the Steinhardt order parameters
calculate_ql / calculate_ql_averaged (:687-711, :832-902).

Spherical harmonics come from SphericalHarmonics.jl in the reference (an
unversioned dependency absent here).  Its published definition --
orthonormal Y_l^m with the Condon-Shortley phase -- is pinned against
scipy.special.sph_harm_y for polar angles in [0, pi].  The reference's
angles (theta = atan(rho/z), phi = atan(y/x)) put neighbours below the
particle at theta < 0; there the oracle continues the formula with the
signed sine, Y(theta, phi) = (-1)^l Y(theta + pi, phi) (parity unpinned:
the package's value at theta < 0 is not held by any reference test).  The
full q_l is then checked against an independent numpy implementation that
evaluates the reference's formulas with scipy's harmonics and that
continuation.
"""
import warnings

import numpy as np
import pytest
import scipy.special as sp

import oracle
from tests import helpers as H


def test_ylm_matches_scipy():
    rng = np.random.default_rng(1)
    for l in range(1, 9):
        for m in range(-l, l + 1):
            for th, ph in zip(rng.uniform(0.0, np.pi, 4), rng.uniform(-np.pi, np.pi, 4)):
                assert abs(oracle.ylm(l, m, th, ph) - sp.sph_harm_y(l, m, th, ph)) <= 1e-13


def test_ylm_continuation_below_the_particle():
    rng = np.random.default_rng(2)
    for l in (3, 4, 6):
        for m in range(-l, l + 1):
            th, ph = rng.uniform(0.05, np.pi / 2), rng.uniform(-np.pi, np.pi)
            assert abs(oracle.ylm(l, m, th - np.pi, ph) - (-1) ** l * oracle.ylm(l, m, th, ph)) <= 1e-13


def _ylm_ref(l, m, th, ph):
    """Y at the reference's angles: scipy on [0, pi], continued below."""
    return np.where(th >= 0.0, sp.sph_harm_y(l, m, th, ph), (-1) ** l * sp.sph_harm_y(l, m, th + np.pi, ph))


def np_ql(r, L, qc, l, averaged):
    N = len(r)
    d = r[:, None, :] - r[None, :, :]
    d -= np.round(d / L) * L
    d2 = np.sum(d * d, axis=2)
    mask = (d2 <= qc * qc) & ~np.eye(N, dtype=bool)
    dx, dy, dz = d[..., 0], d[..., 1], d[..., 2]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        th = np.arctan(np.sqrt(dx * dx + dy * dy) / dz)
        ph = np.arctan(dy / np.where(dx == 0.0, dx + 1.0000000000000006e-10, dx))
    nq = mask.sum(1)
    q = np.zeros((N, 2 * l + 1), dtype=complex)
    for a, m in enumerate(range(-l, l + 1)):
        Y = np.where(mask, _ylm_ref(l, m, th, ph), 0.0)
        q[:, a] = np.where(nq > 0, Y.sum(1) / np.maximum(nq, 1), 0.0)
    if averaged:
        q = (q + mask.astype(float) @ q) / (nq + 1)[:, None]
    return float(np.mean(np.sqrt(4 * np.pi / (2 * l + 1) * np.sum(np.abs(q) ** 2, axis=1))))


@pytest.mark.parametrize("l", [4, 6])
@pytest.mark.parametrize("averaged", [False, True])
@pytest.mark.parametrize("qc", [1.45, 2.0])
def test_steinhardt_oracle_matches_numpy(l, averaged, qc):
    op, r, D, v, rc = H.make_system("berthier", 343, 3, seed=61)
    lists = oracle.build_lists(op, r, D, cells=False)
    got = oracle.steinhardt(op, r, lists, l, qc, averaged)
    want = np_ql(r, op.box_size, qc, l, averaged)
    assert abs(got - want) <= 1e-12 * want


def test_steinhardt_isolated_particles_are_zero():
    op, r, D, v, rc = H.make_system("berthier", 125, 3, seed=62)
    lists = oracle.build_lists(op, r, D, cells=False)
    assert oracle.steinhardt(op, r, lists, 6, 0.1, False) == 0.0


def test_f2_self_oracle():
    op, r, D, v, rc = H.make_system("berthier", 512, 3, seed=63)
    rng = np.random.default_rng(3)
    r2 = r + rng.normal(0.0, 0.2, r.shape)
    r2 -= np.floor(r2 / op.box_size) * op.box_size
    k = 7.2
    want = np.mean(np.cos(k * (r2[:, 0] - r[:, 0])))  # cos(a-b) = cos a cos b + sin a sin b
    assert abs(oracle.f2_self(op, r, r2, k) - want) <= 1e-13
    assert abs(oracle.f2_self(op, r, r, k) - 1.0) <= 1e-15
