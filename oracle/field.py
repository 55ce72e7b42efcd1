"""FP64 restatement of HyMD's particle-field path (oracle; test infrastructure only).

Follows, function by function (reference file:line):

* CIC paint / readout: pmesh's ``ParticleMesh.paint``/``RealField.readout``
  with the default cloud-in-cell window, as called from field.py:200,363,402,574.
  Convention: grid node ``i`` sits at ``x = i*h`` (h = L/M); a particle at grid
  coordinate ``g = x*M/L`` gives weight ``1-(g-floor g)`` to node ``floor g``
  and ``g-floor g`` to node ``floor g + 1`` (periodic).
* r2c / c2r: pmesh normalisation, ``r2c(x) = rfftn(x)/G`` and
  ``c2r(X) = G*irfftn(X)`` so that ``c2r(r2c(x)) == x`` (field.py:576-578).
* Filter: ``H(k) = exp(-sigma^2 |k|^2 / 2)`` (hamiltonian.py:57-71).
* W'(phi): the three Hamiltonians are quadratic, ``w = 1/2 phi^T A phi + b.phi
  + c0`` (hamiltonian.py:145-146, 259-260, 393-412); ``v_t = (A phi + b)_t``
  (hamiltonian.py:188-191, 303-306, 470-473).
* update_field: field.py:570-616 step by step (6T transforms, 7T with
  ``compute_potential``).
* compute_field_force: field.py:197-200.
* compute_field_and_kinetic_energy: field.py:688-703.
* compute_self_energy_q / update_field_force_q: field.py:231-238, 356-403.

Spectral-gradient convention at the Nyquist index (pmesh's own choice cannot be
checked without pmesh): for an even mesh edge ``M_d`` the derivative transfer
``-i k_d`` uses ``k_d = 0`` on the plane ``n_d = M_d/2``.  The filter and
``|k|^2`` use ``(pi/h)^2`` there.  The HIP path uses the same convention.

Pinning (checks that do not retype these formulas; see tests/):
* paint: mass conservation (test_hamiltonian.py:106-109), brute force on tiny
  inputs, integer-cell translation => permuted grid.
* paint -> filter -> energy: Gaussian-core closed forms at 160^3 / 15 nm
  (test_hamiltonian.py:172-227, 270-360) and their Mathematica values.
* field forces: particles on grid nodes reproduce the continuum Gaussian-core
  pair forces of compute_gaussian_core.f90:60-89 generalised to a coupling
  matrix; sum of forces = 0; 90-degree rotation permutes forces.
* PME: self-energy closed form; two opposite charges reproduce the
  Gaussian-smeared Coulomb force; neutral system sum of forces = 0.
"""
import numpy as np

try:  # scipy's pocketfft with worker threads (CPU baseline speed)
    import scipy.fft as _fft

    def _rfftn(x, workers):
        return _fft.rfftn(x, workers=workers)

    def _irfftn(x, s, workers):
        return _fft.irfftn(x, s=s, workers=workers)
except ImportError:  # pragma: no cover
    def _rfftn(x, workers):
        return np.fft.rfftn(x)

    def _irfftn(x, s, workers):
        return np.fft.irfftn(x, s=s)

#: fixed-point fraction bits of the deterministic paint (HIP path replica)
FIXED_SHIFT = 32
_WORKERS = -1


def set_workers(n):
    """Number of FFT worker threads used by r2c/c2r (-1 = all cores)."""
    global _WORKERS
    _WORKERS = int(n)


def _mesh3(mesh):
    m = np.broadcast_to(np.asarray(mesh, dtype=np.int64), (3,))
    return tuple(int(v) for v in m)


def _box3(box):
    return np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)).copy()


# ---------------------------------------------------------------------------
# CIC window
# ---------------------------------------------------------------------------
def cic_axis(x, M, L):
    """Left node index (in [0, M)) and fraction for one axis.

    ``g = x * (M/L)`` in float64, ``i = floor(g)``, ``f = g - i``.
    """
    inv_h = float(M) / float(L)
    g = np.asarray(x, dtype=np.float64) * inv_h
    i0 = np.floor(g)
    f = g - i0
    return (i0.astype(np.int64) % M), f


def _corners(x, mesh, box):
    M = _mesh3(mesh)
    L = _box3(box)
    x = np.asarray(x)
    ix, fx = cic_axis(x[:, 0], M[0], L[0])
    iy, fy = cic_axis(x[:, 1], M[1], L[1])
    iz, fz = cic_axis(x[:, 2], M[2], L[2])
    wx = (1.0 - fx, fx)
    wy = (1.0 - fy, fy)
    wz = (1.0 - fz, fz)
    jx = (ix, (ix + 1) % M[0])
    jy = (iy, (iy + 1) % M[1])
    jz = (iz, (iz + 1) % M[2])
    out = []
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = (wx[a] * wy[b]) * wz[c]
                idx = (jx[a] * M[1] + jy[b]) * M[2] + jz[c]
                out.append((idx, w))
    return M, out


def paint(x, mesh, box, mass=1.0):
    """CIC paint (pmesh ``paint``), returning sum of mass*weights (not /dV)."""
    M, corners = _corners(x, mesh, box)
    G = M[0] * M[1] * M[2]
    grid = np.zeros(G, dtype=np.float64)
    n = len(np.asarray(x))
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,))
    for idx, w in corners:
        grid += np.bincount(idx, weights=w * m, minlength=G)
    return grid.reshape(M)


def paint_fixed(x, mesh, box, mass=1.0, shift=FIXED_SHIFT):
    """Deterministic int64 fixed-point CIC paint, the exact integer replica of
    the HIP paint kernel.

    Per particle and corner ``c`` (corners ordered x-major: c = 4a+2b+c),
    ``q_c = rint(((wx*wy)*wz)*m * 2^shift)`` for the first seven corners and
    ``q_7 = rint(m*2^shift) - sum(q_0..q_6)``, so every particle deposits
    exactly its fixed-point mass.  Returns the int64 grid.
    """
    M, corners = _corners(x, mesh, box)
    G = M[0] * M[1] * M[2]
    n = len(np.asarray(x))
    m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,))
    scale = float(2 ** shift)
    total = np.rint(m * scale).astype(np.int64)
    grid = np.zeros(G, dtype=np.int64)
    acc = np.zeros(n, dtype=np.int64)
    for c, (idx, w) in enumerate(corners):
        if c < 7:
            q = np.rint((w * m) * scale).astype(np.int64)
            acc += q
        else:
            q = total - acc
        np.add.at(grid, idx, q)
    return grid.reshape(M)


def readout(grid, x, mesh, box):
    """CIC readout (pmesh ``RealField.readout``)."""
    M, corners = _corners(x, mesh, box)
    g = np.asarray(grid, dtype=np.float64).reshape(-1)
    val = np.zeros(len(np.asarray(x)), dtype=np.float64)
    for idx, w in corners:
        val += w * g[idx]
    return val


# ---------------------------------------------------------------------------
# Transforms and k-space
# ---------------------------------------------------------------------------
def r2c(real):
    G = real.size
    return _rfftn(real, _WORKERS) / G


def c2r(cplx, mesh):
    M = _mesh3(mesh)
    G = M[0] * M[1] * M[2]
    return _irfftn(cplx, M, _WORKERS) * G


def kvectors(mesh, box):
    """k_x, k_y (full axes, fftfreq order) and k_z (half axis), broadcastable."""
    M = _mesh3(mesh)
    L = _box3(box)
    kx = 2.0 * np.pi / L[0] * np.fft.fftfreq(M[0], 1.0 / M[0])
    ky = 2.0 * np.pi / L[1] * np.fft.fftfreq(M[1], 1.0 / M[1])
    kz = 2.0 * np.pi / L[2] * np.arange(M[2] // 2 + 1, dtype=np.float64)
    return kx[:, None, None], ky[None, :, None], kz[None, None, :]


def k_squared(mesh, box):
    kx, ky, kz = kvectors(mesh, box)
    return kx * kx + ky * ky + kz * kz


def derivative_k(mesh, box, d):
    """k_d for the odd (derivative) transfer, zero on the Nyquist plane."""
    M = _mesh3(mesh)
    ks = list(kvectors(mesh, box))
    k = ks[d].copy()
    if M[d] % 2 == 0:
        idx = [slice(None)] * 3
        idx[d] = M[d] // 2
        k[tuple(idx)] = 0.0
    return k


def filter_H(mesh, box, sigma):
    """Gaussian filter H(k) on the half spectrum (hamiltonian.py:57-71)."""
    return np.exp(-0.5 * sigma * sigma * k_squared(mesh, box))


def window_H(k, sigma):
    """H(k) at explicit k vectors (hamiltonian.py:64-69), k shape (3,) or (n,3)."""
    k = np.asarray(k, dtype=np.float64)
    return np.exp(-0.5 * sigma ** 2 * np.sum(k * k, axis=-1))


# ---------------------------------------------------------------------------
# Hamiltonians as quadratic forms  w = 1/2 phi.A.phi + b.phi + c0
# ---------------------------------------------------------------------------
class Quadratic:
    """Quadratic interaction energy density in the filtered densities.

    ``A`` (T,T), ``b`` (T,), ``c0`` scalar.  ``v_ext(t, phi) = (A phi + b)_t``.
    """

    def __init__(self, A, b, c0):
        self.A = np.asarray(A, dtype=np.float64)
        self.b = np.asarray(b, dtype=np.float64)
        self.c0 = float(c0)

    def v_ext(self, t, phi):
        v = self.b[t] + 0.0 * phi[0]
        for s in range(len(phi)):
            if self.A[t, s] != 0.0:
                v = v + self.A[t, s] * phi[s]
        return v

    def w(self, phi):
        T = len(phi)
        out = self.c0 + 0.0 * phi[0]
        for t in range(T):
            out = out + self.b[t] * phi[t]
            for s in range(T):
                out = out + 0.5 * self.A[t, s] * phi[t] * phi[s]
        return out


def squared_phi(T, kappa, rho0):
    """SquaredPhi: w = (sum phi)^2 / (2 kappa rho0) (hamiltonian.py:145-146)."""
    A = np.full((T, T), 1.0 / (kappa * rho0))
    return Quadratic(A, np.zeros(T), 0.0)


def default_no_chi(T, kappa, rho0, a):
    """DefaultNoChi: w = (sum phi - a)^2 / (2 kappa rho0) (hamiltonian.py:259-260)."""
    A = np.full((T, T), 1.0 / (kappa * rho0))
    b = np.full(T, -a / (kappa * rho0))
    return Quadratic(A, b, 0.5 * a * a / (kappa * rho0))


def default_with_chi(T, kappa, rho0, a, chi):
    """DefaultWithChi (hamiltonian.py:393-412): adds sum_{i<j} chi_ij phi_i phi_j / rho0.

    ``chi`` is (T,T) symmetric; its diagonal is ignored as in the reference
    (hamiltonian.py:403-410, 440-443).
    """
    q = default_no_chi(T, kappa, rho0, a)
    chi = np.array(chi, dtype=np.float64)
    np.fill_diagonal(chi, 0.0)
    q.A = q.A + chi / rho0
    return q


# ---------------------------------------------------------------------------
# update_field / compute_field_force / energies
# ---------------------------------------------------------------------------
def update_field(x, types, n_types, mesh, box, sigma, ham, m=None,
                 compute_potential=False):
    """field.py:570-616, in the reference's order (6T transforms, +T).

    Returns dict with ``phi`` (filtered real densities), ``phi_fourier``
    (filtered spectra, pmesh normalisation), ``force_mesh`` [T][3] and, if
    ``compute_potential``, ``v_ext`` (real potentials).
    """
    M = _mesh3(mesh)
    L = _box3(box)
    V = float(np.prod(L))
    dV = V / float(M[0] * M[1] * M[2])
    x = np.asarray(x)
    types = np.asarray(types)
    if m is None:
        m = [1.0] * n_types
    H = filter_H(M, L, sigma)
    phi, phi_fourier = [], []
    for t in range(n_types):
        p = paint(x[types == t], M, L, mass=m[t]) / dV
        pf = r2c(p) * H
        phi_fourier.append(pf)
        phi.append(c2r(pf, M))
    kd = [derivative_k(M, L, d) for d in range(3)]
    force_mesh, v_ext = [], []
    for t in range(n_types):
        v = ham.v_ext(t, phi)
        vf = r2c(v) * H
        force_mesh.append([c2r(-1j * kd[d] * vf, M) for d in range(3)])
        if compute_potential:
            v_ext.append(c2r(vf, M))
    return {"phi": phi, "phi_fourier": phi_fourier, "force_mesh": force_mesh,
            "v_ext": v_ext}


def compute_field_force(x, types, force_mesh, mesh, box):
    """field.py:197-200: per-type CIC readout of the 3 force grids."""
    x = np.asarray(x)
    types = np.asarray(types)
    f = np.zeros((len(x), 3), dtype=np.float64)
    for t in range(len(force_mesh)):
        ind = types == t
        for d in range(3):
            f[ind, d] = readout(force_mesh[t][d], x[ind], mesh, box)
    return f


def field_energy(phi, ham, mesh, box):
    """E_field = dV * sum_cells w(phi~) (field.py:692-693)."""
    M = _mesh3(mesh)
    dV = float(np.prod(_box3(box))) / float(M[0] * M[1] * M[2])
    return float(np.sum(ham.w(phi)) * dV)


def kinetic_energy(v, mass):
    """K = 1/2 m sum v^2 (field.py:695)."""
    return 0.5 * mass * float(np.sum(np.asarray(v, dtype=np.float64) ** 2))


def self_energy_q(charges, sigma, coulomb_constant, dielectric):
    """field.py:231-238."""
    elec = coulomb_constant / dielectric
    prefac = elec * np.sqrt(1.0 / (2.0 * np.pi * sigma * sigma))
    q = np.asarray(charges, dtype=np.float64)
    return float(prefac * np.sum(q * q))


def update_field_force_q(x, charges, mesh, box, sigma, coulomb_constant,
                         dielectric):
    """field.py:356-403 (filter applied once; |k|^2 := 1 at k = 0)."""
    M = _mesh3(mesh)
    L = _box3(box)
    dV = float(np.prod(L)) / float(M[0] * M[1] * M[2])
    elec = coulomb_constant / dielectric
    q = np.asarray(charges, dtype=np.float64)
    phi_q = paint(x, M, L, mass=q) / dV
    phq = r2c(phi_q) * filter_H(M, L, sigma)
    k2 = k_squared(M, L)
    k2[0, 0, 0] = 1.0
    psi_f = 4.0 * np.pi * elec * phq / k2
    psi = c2r(psi_f, M)
    elec_field = [c2r(-1j * derivative_k(M, L, d) * psi_f, M) for d in range(3)]
    forces = np.stack([q * readout(elec_field[d], x, M, L) for d in range(3)],
                      axis=1)
    return {"phi_q": phi_q, "phi_q_fourier": phq, "psi": psi,
            "psi_fourier": psi_f, "elec_field": elec_field, "forces": forces}


def field_q_energy(phi_q, psi, self_energy, mesh, box):
    """E_q = dV sum (1/2 phi_q psi - E_self/V) (field.py:697-699, hamiltonian.py:414-421)."""
    M = _mesh3(mesh)
    L = _box3(box)
    V = float(np.prod(L))
    dV = V / float(M[0] * M[1] * M[2])
    return float(np.sum(0.5 * phi_q * psi - self_energy / V) * dV)


# ---------------------------------------------------------------------------
# Continuum (h -> 0) Gaussian-core model: closed forms used to pin the grid
# ---------------------------------------------------------------------------
def gaussian_core_pairs(x, types, A, sigma, box):
    """Energy and forces of the continuum limit of a quadratic Hamiltonian.

    With phi~_t = sum_i G_sigma(r - r_i) the grid energy tends to
    ``sum_{i<j} A_{t_i t_j} g(r_ij) + 1/2 sum_i A_{t_i t_i} g(0)`` plus the
    linear and constant terms, with g(r) = (4 pi sigma^2)^{-3/2}
    exp(-r^2/(4 sigma^2)); this is compute_gaussian_core.f90:43-89 with the
    coupling matrix in place of (1 + kappa chi_ij)/(kappa rho0).  Minimum
    image, no periodic images beyond it.
    """
    x = np.asarray(x, dtype=np.float64)
    L = _box3(box)
    types = np.asarray(types)
    n = len(x)
    norm = (4.0 * np.pi * sigma * sigma) ** -1.5
    e = 0.0
    f = np.zeros((n, 3))
    for i in range(n):
        e += 0.5 * A[types[i], types[i]] * norm
        for j in range(i + 1, n):
            rij = x[i] - x[j]
            rij -= L * np.round(rij / L)
            r2 = float(np.dot(rij, rij))
            g = A[types[i], types[j]] * norm * np.exp(-r2 / (4.0 * sigma * sigma))
            e += g
            fij = g * rij / (2.0 * sigma * sigma)
            f[i] += fij
            f[j] -= fij
    return e, f
