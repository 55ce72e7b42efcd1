"""Interaction energy functionals as quadratic forms (drop-in for hamiltonian.py).

The reference builds w, W' and the filter with sympy (hamiltonian.py:34-486).
All three functionals are quadratic in the filtered densities, so each is held
here in closed form as

    w(phi) = 1/2 phi^T A phi + b.phi + c0,    v_t = dw/dphi_t = (A phi + b)_t,

which is what the HIP k-space pass consumes.  ``H``, ``v_ext``, ``w_0``,
``w_elec`` keep the reference's call signatures (numpy or torch arrays).
"""
import numpy as np

from .config import chi_matrix


class Hamiltonian:
    """Superclass: rho0 / a (hamiltonian.py:41-47) and the Gaussian filter
    H(k) = exp(-sigma^2 |k|^2 / 2) (hamiltonian.py:57-71)."""

    def __init__(self, config):
        self.config = config
        self._setup()

    def _setup(self):
        c = self.config
        if getattr(c, "simulation_volume", None) is None:
            c.simulation_volume = float(np.prod(np.asarray(c.box_size)))
        if not c.barostat:
            c.rho0 = c.n_particles / c.simulation_volume
            c.a = c.rho0
        if not c.rho0:
            c.rho0 = c.n_particles / c.simulation_volume
        if not c.self_energy:
            c.self_energy = 0.0
        sigma = float(c.sigma)

        def H(k, v):
            k = [np.asarray(kk, dtype=np.float64) for kk in k]
            return v * np.exp(-0.5 * sigma ** 2 * (k[0] ** 2 + k[1] ** 2 + k[2] ** 2))

        self.H = H

    # quadratic form ---------------------------------------------------------
    def quadratic(self):
        """(A, b, c0) with A (T,T), b (T,), c0 scalar."""
        raise NotImplementedError

    def _install(self):
        A, b, c0 = self.quadratic()
        T = len(b)
        V = float(self.config.simulation_volume)
        self_energy = float(self.config.self_energy or 0.0)

        def make_v(t):
            def v(phi):
                out = b[t] + 0.0 * phi[0]
                for s in range(T):
                    if A[t, s] != 0.0:
                        out = out + A[t, s] * phi[s]
                return out
            return v

        def w(phi):
            out = c0 + 0.0 * phi[0]
            for t in range(T):
                out = out + b[t] * phi[t]
                for s in range(T):
                    out = out + 0.5 * A[t, s] * phi[t] * phi[s]
            return out

        def w_elec(args):
            phi_q, psi = args
            return 0.5 * phi_q * psi - self_energy / V

        self.v_ext = [make_v(t) for t in range(T)]
        self.w_0 = w
        self.w_elec = w_elec
        self.w = w


class SquaredPhi(Hamiltonian):
    """w = (sum phi)^2 / (2 kappa rho0)  (hamiltonian.py:145-146)."""

    def __init__(self, config):
        super().__init__(config)
        self._install()

    def quadratic(self):
        c = self.config
        T = c.n_types
        A = np.full((T, T), 1.0 / (c.kappa * c.rho0))
        return A, np.zeros(T), 0.0


class DefaultNoChi(Hamiltonian):
    """w = (sum phi - a)^2 / (2 kappa rho0)  (hamiltonian.py:259-260)."""

    def __init__(self, config):
        super().__init__(config)
        self._install()

    def quadratic(self):
        c = self.config
        T = c.n_types
        inv = 1.0 / (c.kappa * c.rho0)
        return np.full((T, T), inv), np.full(T, -c.a * inv), 0.5 * c.a * c.a * inv


class DefaultWithChi(Hamiltonian):
    """DefaultNoChi plus sum_{i<j} chi_ij phi_i phi_j / rho0 (hamiltonian.py:393-412);
    the chi diagonal is ignored as in the reference (:403-410, :440-443)."""

    def __init__(self, config, unique_names=None, type_to_name_map=None):
        if unique_names is not None:
            config.unique_names = list(unique_names)
        if type_to_name_map is not None:
            config.type_to_name_map = dict(type_to_name_map)
        super().__init__(config)
        self._install()

    def quadratic(self):
        c = self.config
        T = c.n_types
        inv = 1.0 / (c.kappa * c.rho0)
        A = np.full((T, T), inv) + chi_matrix(c) / c.rho0
        return A, np.full(T, -c.a * inv), 0.5 * c.a * c.a * inv


def get_hamiltonian(config):
    """hamiltonian.py:489-512."""
    name = (config.hamiltonian or "DefaultNoChi").lower()
    if name == "defaultnochi":
        return DefaultNoChi(config)
    if name == "defaultwithchi":
        return DefaultWithChi(config, config.unique_names, config.type_to_name_map)
    if name == "squaredphi":
        return SquaredPhi(config)
    raise ValueError(f"unknown hamiltonian {config.hamiltonian!r}")


def quadratic_form(hamiltonian, n_types):
    """(A, b, c0) of any quadratic Hamiltonian.

    Objects of this module return their closed form.  Any other object with
    ``v_ext[t](phi)`` and ``w_0(phi)`` callables (e.g. the reference's
    sympy-lambdified classes) is probed at phi = 0 and unit vectors, then
    checked for linearity of v at a random point; non-quadratic functionals
    raise ValueError (the fused k-space pass needs a linear W').
    """
    if hasattr(hamiltonian, "quadratic"):
        A, b, c0 = hamiltonian.quadratic()
        return np.asarray(A, dtype=np.float64), np.asarray(b, dtype=np.float64), float(c0)
    T = int(n_types)

    def ev(f, phi):
        return float(np.asarray(f([np.asarray(p, dtype=np.float64) for p in phi])).reshape(-1)[0])

    zero = [np.zeros(1) for _ in range(T)]
    b = np.array([ev(hamiltonian.v_ext[t], zero) for t in range(T)])
    A = np.zeros((T, T))
    for s in range(T):
        e = [np.full(1, 1.0 if u == s else 0.0) for u in range(T)]
        for t in range(T):
            A[t, s] = ev(hamiltonian.v_ext[t], e) - b[t]
    rng = np.random.default_rng(0)
    x = rng.uniform(0.1, 2.0, T)
    probe = [np.full(1, x[u]) for u in range(T)]
    for t in range(T):
        got = ev(hamiltonian.v_ext[t], probe)
        want = b[t] + A[t] @ x
        if not np.isclose(got, want, rtol=1e-10, atol=1e-12):
            raise ValueError("Hamiltonian is not quadratic in phi: W' is not linear")
    c0 = ev(hamiltonian.w_0, zero)
    return A, b, c0
