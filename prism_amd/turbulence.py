"""turbulence::IncompressibleKEpsilon (KEpsilon.h, KEpsilon.cpp, k_production.cpp) on the C ABI.

``equations()`` refreshes the per-cell coefficient fields (nu_t, Pk, the effective
viscosities and the epsilon source and sink) in one device pass (prism_kepsilon_update) and
returns the k and epsilon transport equations built as KEpsilon.cpp:41-77 builds them.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _capi as A
from .core import Scalar, Transport, Vector, _dp, convection, diffusion, source, temporal


class IncompressibleKEpsilon:
    """Standard k-epsilon model for incompressible flow (KEpsilon.h:11-40).

    k and epsilon are fields named "k" and "epsilon" on the viscosity's mesh (their boundary
    conditions come from that mesh), initialised to 0 as the reference does: set their
    values before the first ``equations()`` call.  With ``dt`` > 0 both keep two previous
    time levels and their equations gain AdamMoulton(dt)."""

    c_mu, c_eps1, c_eps2, sigma_k, sigma_eps = 0.09, 1.44, 1.92, 1.0, 1.3

    def __init__(self, mdot: Vector, U: Vector, mu: Scalar, dt: Optional[float] = None):
        mesh = mu.mesh
        self._mdot, self._U, self._nu = mdot, U, mu
        self._k = mesh.scalar("k", 0.0)
        self._epsilon = mesh.scalar("epsilon", 0.0)
        self._nu_t = mesh.scalar("mu_t", 0.0)
        self._transient = dt is not None and dt > 0.0
        self._dt = float(dt) if self._transient else 0.0
        if self._transient:
            self._k.set_history_size(2)
            self._epsilon.set_history_size(2)
        # coefficient fields of the two equations, rewritten by every equations() call
        self._pk = mesh.scalar("Pk", 0.0)
        self._nu_eff_k = mesh.scalar("mu_eff_k", 0.0)
        self._nu_eff_eps = mesh.scalar("mu_eff_eps", 0.0)
        self._eps_source = mesh.scalar("eps_source", 0.0)
        self._eps_sink = mesh.scalar("eps_sink", 0.0)
        self._eqns: Optional[Tuple[Transport, Transport]] = None

    def k(self) -> Scalar:
        return self._k

    def epsilon(self) -> Scalar:
        return self._epsilon

    def turbulent_viscosity(self) -> Scalar:
        return self._nu_t

    def production(self) -> Scalar:
        """Pk of the last equations() call (kProductionIncompressible)."""
        return self._pk

    def update_fields(self):
        out = (C.c_void_p * 6)(*[f.h.value for f in (self._nu_t, self._pk, self._nu_eff_k, self._nu_eff_eps,
                                                     self._eps_source, self._eps_sink)])
        consts = np.array([self.c_mu, self.c_eps1, self.c_eps2, self.sigma_k, self.sigma_eps])
        self._nu.mesh.backend("kepsilon_update", self._k.h, self._epsilon.h, self._nu.h, self._U.h, _dp(consts), out)

    def equations(self) -> Tuple[Transport, Transport]:
        """(k equation, epsilon equation) of KEpsilon.cpp:35-77.  The same two Transport
        objects are returned by every call, with refreshed coefficient fields and the
        under-relaxation factor reset to 1, as the reference's freshly built ones have."""
        self.update_fields()
        if self._eqns is None:
            k_schemes = [convection.LinearUpwind(self._mdot, self._k),
                         diffusion.NonCorrected(self._nu_eff_k, self._k),
                         source.ConstantScalar(self._pk, "+"),
                         source.ConstantScalar(self._epsilon, "-")]
            e_schemes = [convection.LinearUpwind(self._mdot, self._epsilon),
                         diffusion.NonCorrected(self._nu_eff_eps, self._epsilon),
                         source.ConstantScalar(self._eps_source, "+"),
                         source.ConstantScalar(self._eps_sink, "-")]
            if self._transient:
                k_schemes.append(temporal.AdamMoulton(self._k, self._dt))
                e_schemes.append(temporal.AdamMoulton(self._epsilon, self._dt))
            self._eqns = (Transport(*k_schemes), Transport(*e_schemes))
        else:
            for e in self._eqns:
                e.set_under_relax_factor(1.0)
        return self._eqns
