"""Materials of the hot path, with the reference's constructors and accessors.

SVK                     -- src/Materials/SVKMaterial.jl:25-100
NeoHookean              -- src/Materials/NeoHookeanMaterial.jl:127-193
IsotropicLinearElastic  -- src/Materials/IsotropicLinearElasticMaterial.jl:264-318

A shear modulus G may be a per-element array (one value per element of the set
the material is assigned to); the reference would need one material per
element for that.
"""
from __future__ import annotations

import numpy as np

from . import _capi


def _label(label):
    return "no_label" if label is None else str(label)


class AbstractMaterial:
    kind = 0
    label = "no_label"

    def params(self):
        raise NotImplementedError


class SVK(AbstractMaterial):
    """Saint-Venant-Kirchhoff, Psi(E) = lambda/2 tr(E)^2 + G tr(E^2).

    SVK(lam, G) or SVK(E=..., nu=...) (SVKMaterial.jl:34-54)."""

    kind = _capi.MAT_SVK

    def __init__(self, lam=None, G=None, label=None, *, E=None, nu=None, rho=None):
        if E is not None or nu is not None:
            if lam is not None or G is not None:
                raise ValueError("give either (lam, G) or (E, nu)")
            lam = E * nu / ((1 + nu) * (1 - 2 * nu))
            G = E / (2 * (1 + nu))
        if lam is None or G is None:
            raise ValueError("SVK needs (lam, G) or (E=, nu=)")
        if rho is not None and not rho > 0:
            raise ValueError("Density must be positive.")
        if not lam >= 0:
            raise AssertionError("The first Lamé parameter `λ` must be positive.")
        if not np.all(np.asarray(G) >= 0):
            raise AssertionError("The second Lamé parameter or shear modulus `G` must be positive.")
        self.lam = float(lam)
        self.G = G if np.ndim(G) else float(G)
        self.rho = rho
        self.label = _label(label)

    def lame_parameters(self):
        return self.lam, self.G

    def shear_modulus(self):
        return self.G

    def poisson_ratio(self):
        lam, G = self.lam, self.G
        return lam / (2 * (lam + G))

    def elasticity_modulus(self):
        lam, G = self.lam, self.G
        return G * (3 * lam + 2 * G) / (lam + G)

    def bulk_modulus(self):
        return self.lam + 2 * self.G / 3

    def params(self):
        return (self.lam, self.G)


class NeoHookean(AbstractMaterial):
    """Psi = G/2 (tr C - 2 log J) + K/2 (J - 1)^2.

    NeoHookean(K, G) or NeoHookean(E=..., nu=...) (NeoHookeanMaterial.jl:136-158)."""

    kind = _capi.MAT_NEOHOOKEAN

    def __init__(self, K=None, G=None, label=None, *, E=None, nu=None, rho=None):
        if E is not None or nu is not None:
            lam = E * nu / ((1 + nu) * (1 - 2 * nu))
            G = E / (2 * (1 + nu))
            K = lam + 2 * G / 3
        if K is None or G is None:
            raise ValueError("NeoHookean needs (K, G) or (E=, nu=)")
        if rho is not None and not rho > 0:
            raise ValueError("Density must be positive.")
        if not K >= 0:
            raise AssertionError("The bulk modulus `K` must be positive.")
        if not np.all(np.asarray(G) >= 0):
            raise AssertionError("The shear modulus or second Lamé parameter `μ` must be positive.")
        self.K = float(K)
        self.G = G if np.ndim(G) else float(G)
        self.rho = rho
        self.label = _label(label)

    def lame_parameters(self):
        G = self.G
        return self.K - 2 * G / 3, G

    def shear_modulus(self):
        return self.G

    def bulk_modulus(self):
        return self.K

    def poisson_ratio(self):
        lam, G = self.lame_parameters()
        return lam / (2 * (lam + G))

    def elasticity_modulus(self):
        lam, G = self.lame_parameters()
        return G * (3 * lam + 2 * G) / (lam + G)

    def params(self):
        return (self.K, self.G)


class IsotropicLinearElastic(AbstractMaterial):
    """sigma = lambda tr(eps) I + 2 G eps.

    IsotropicLinearElastic(E, nu) or IsotropicLinearElastic(lam=..., G=...)
    (IsotropicLinearElasticMaterial.jl:273-291)."""

    kind = _capi.MAT_LINEAR

    def __init__(self, E=None, nu=None, label=None, *, lam=None, G=None, rho=None):
        if lam is not None or G is not None:
            E = G * (3 * lam + 2 * G) / (lam + G)
            nu = lam / (2 * (lam + G))
        if E is None or nu is None:
            raise ValueError("IsotropicLinearElastic needs (E, nu) or (lam=, G=)")
        self.E = float(E)
        self.nu = float(nu)
        self.rho = rho
        self.label = _label(label)

    def elasticity_modulus(self):
        return self.E

    def poisson_ratio(self):
        return self.nu

    def shear_modulus(self):
        return self.E / (2 * (1 + self.nu))

    def bulk_modulus(self):
        return self.E / (3 * (1 - 2 * self.nu))

    def lame_parameters(self):
        E, nu = self.E, self.nu
        return E * nu / ((1 + nu) * (1 - 2 * nu)), self.shear_modulus()

    def params(self):
        return (self.E, self.nu)


def elasticity_modulus(m):
    return m.elasticity_modulus()


def shear_modulus(m):
    return m.shear_modulus()


def bulk_modulus(m):
    return m.bulk_modulus()


def poisson_ratio(m):
    return m.poisson_ratio()


def lame_parameters(m):
    return m.lame_parameters()
