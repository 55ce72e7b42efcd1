/*
 * hymd_amd.h -- C-ABI of the MI355X-native HyMD particle-field hot path.
 *
 * Plain pointers and sizes only.  "device" pointers are HIP device memory
 * owned by the caller (e.g. torch tensors' data_ptr()); "host" pointers are
 * ordinary host memory.  Every entry point returns HYMD_OK (0) or a negative
 * status; hymd_last_error() returns a message for the calling thread.
 * Work is enqueued on the stream given at creation (or per call) and is
 * asynchronous unless an energy/virial is requested into host memory, in
 * which case the call synchronises that stream before returning.  Nothing
 * prints or aborts.  Calls on distinct handles may run from distinct threads;
 * calls on one handle must not overlap.
 *
 * Reference interfaces replaced (Cascella-Group-UiO/HyMD v2.2.0):
 *   field.py   initialize_pm :10-149, compute_field_force :152-200,
 *              compute_self_energy_q :203-238, update_field_force_q :241-403,
 *              update_field :428-616, compute_field_and_kinetic_energy :619-703
 *   force.py   f2py module force_kernels cbf/caf/cdf(+_d) :8-23
 *              (compute_bond_forces*.f90, compute_angle_forces*.f90,
 *              compute_dihedral_forces*.f90), dipole_forces_redistribution :855-877
 *   integrator.py integrate_velocity :9-41, integrate_position :44-73 as
 *              composed by the rRESPA loop main.py:802-887,1143-1169
 *   thermostat.py csvr_thermostat :111-219, cancel_com_momentum :12-15
 *   force.py   prepare_bonds :573-728 (over prepare_bonds_old :325-505 and
 *              prepare_index_based_bonds :508-570)
 *   pressure.py comp_pressure :9-179 (field terms, with comp_laplacian :182-200)
 *   barostat.py / barostat_scr.py isotropic, semiisotropic (box and position rescaling)
 */
#ifndef HYMD_AMD_H
#define HYMD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HYMD_OK 0
#define HYMD_ERR_INVALID (-1)     /* bad argument or shape */
#define HYMD_ERR_HIP (-2)         /* a HIP runtime call failed */
#define HYMD_ERR_UNSUPPORTED (-3) /* e.g. a mesh edge that is not 2^a 3^b 5^c */
#define HYMD_ERR_NOMEM (-4)       /* device allocation failed */

#define HYMD_MAX_TYPES 16
/* fraction bits of the deterministic int64 fixed-point density grid */
#define HYMD_FIXED_SHIFT 32

/* precision selector of the bonded kernels (f2py cbf vs cbf_d) */
#define HYMD_F32 0
#define HYMD_F64 1

const char *hymd_last_error(void);
int hymd_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Particle-mesh handle: mesh, box, FFT plans, internal scratch.            */
/* Replaces pmesh.ParticleMesh + the grids of initialize_pm (field.py:45-86). */
/* Mesh edges: even, of the form 2^a 3^b 5^c, 4 <= M <= 1024.               */
/* ------------------------------------------------------------------------ */
typedef struct hymd_pm hymd_pm;

int hymd_pm_create(hymd_pm **pm, const int32_t mesh[3], const double box[3], int32_t max_types,
                   void *stream);
int hymd_pm_destroy(hymd_pm *pm);
int hymd_pm_set_stream(hymd_pm *pm, void *stream);
/* new box at fixed mesh (barostats re-create pmesh objects, barostat.py:163-164) */
int hymd_pm_set_box(hymd_pm *pm, const double box[3]);
/* Launch shapes of the handle's kernels (no reference counterpart: the
 * autotuner's knobs; paint and readout shapes leave every result unchanged
 * bit for bit, another x-line width may round the FP32 x transform
 * differently):
 *   HYMD_TUNE_PAINT_ROWS     mesh rows per paint workgroup, 1 (default) or 4
 *                            (one per wavefront)
 *   HYMD_TUNE_READOUT_BLOCK  threads per readout workgroup, 64, 128 or 256 (default)
 *   HYMD_TUNE_XLINE_WIDTH    lines per x-line FFT + k-space workgroup, 0 (default:
 *                            the widest that fits) or 2, 4, 8, 16 (used when it
 *                            fits the update's type count, else the default)
 * A change makes captured MD steps capture again. */
#define HYMD_TUNE_PAINT_ROWS 1
#define HYMD_TUNE_READOUT_BLOCK 2
#define HYMD_TUNE_XLINE_WIDTH 3
int hymd_pm_set_tuning(hymd_pm *pm, int32_t knob, int32_t value);
int hymd_pm_get_tuning(hymd_pm *pm, int32_t knob, int32_t *value);

/* update_field (field.py:428-616) for a quadratic Hamiltonian
 *   w(phi) = 1/2 phi^T A phi + b.phi + c0   (hamiltonian.py:145,259,393-412)
 * x          device float32 (n,3) row-major positions (any real value; wrapped periodically)
 * types      device int32 (n) in [0, n_types); other values are skipped
 * coupling   host double (n_types, n_types) A;  shift  host double (n_types) b
 * mass       host double (n_types) paint weight per type (config.m, field.py:574)
 * force_mesh device float32 (n_types,3,Mx,My,Mz) out: F_{t,d} = c2r(-i k_d H r2c(v_t))
 * phi        device float32 (n_types,Mx,My,Mz) out or NULL: filtered densities (field.py:578)
 * phi_fourier device complex64 (n_types,Mx,My,Mz/2+1) out or NULL: H*r2c(phi) (pmesh norm.)
 * v_ext      device float32 (n_types,Mx,My,Mz) out or NULL: potentials (field.py:615-616)
 * energy     host double out or NULL: E_field = dV sum w(phi~) (field.py:692-693)
 */
int hymd_update_field(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                      const double *coupling, const double *shift, double constant, const double *mass,
                      double sigma, float *force_mesh, float *phi, void *phi_fourier, float *v_ext,
                      double *energy);

/* compute_field_force (field.py:152-200): force[i] = CIC readout of force_mesh[types[i]] */
int hymd_compute_field_force(hymd_pm *pm, const float *x, const int32_t *types, int64_t n,
                             int32_t n_types, const float *force_mesh, float *force);

/* update_field_force_q (field.py:241-403): PME/Poisson forces on charges
 * charges    device float32 (n); elec_factor = coulomb_constant / dielectric_const
 * elec_forces device float32 (n,3) out
 * phi_q, psi device float32 (Mx,My,Mz) out or NULL; elec_field float32 (3,Mx,My,Mz) out or NULL
 * energy_q   host double out or NULL: dV sum(1/2 phi_q psi) - self_energy (field.py:697-699)
 */
int hymd_update_field_force_q(hymd_pm *pm, const float *x, const float *charges, int64_t n, double sigma,
                              double elec_factor, double self_energy, float *elec_forces, float *phi_q,
                              float *psi, float *elec_field, double *energy_q);

/* update_field_force_q_GPE (field.py:764-1112): electrostatics with the
 * dielectric eps(r) = sum_t eps_t phi~_t / sum_t phi~_t of the filtered
 * densities phi (device float32 (n_types,Mx,My,Mz), as update_field writes
 * them), solved by the reference's polarisation-charge iteration (mixing
 * pol_mixing, at most max_iter iterations while the convergence measure of
 * |change| exceeds conv_crit: conv_type 0 its maximum (max_diff), 1 its sum
 * (csum), 2 its sum of squares (euclidean_norm)).
 *   charges device float32 (n) painted, types device int32 (n) read out
 *   dielectric_type, type_charges host double (n_types)
 *   phi_eps, elec_field_contrib device float32 grids in/out: kept where the
 *     density (resp. its sum) is <= 1e-6, as the reference's masked divisions
 *   phi_pol_prev device grid in or NULL (zero): the iteration's start
 *   phi_pol, psi, elec_dot device grids out or NULL; vbar (n_types grids) out or NULL
 *   elec_forces device float32 (n,3) out: readout of c2r(-i k H r2c(V_t)) per type
 *   energy host out or NULL: compute_field_energy_q_GPE (field.py:706-761)
 *   iterations, delta host out or NULL
 * Synchronises the handle's stream once per iteration. */
int hymd_update_field_force_q_gpe(hymd_pm *pm, const float *x, const int32_t *types, const float *charges, int64_t n,
                                  int32_t n_types, const float *phi, const double *dielectric_type,
                                  const double *type_charges, double sigma, double coulomb_constant,
                                  double conv_crit, double pol_mixing, int32_t conv_type, int32_t max_iter,
                                  float *phi_eps, const float *phi_pol_prev, float *phi_pol,
                                  float *elec_field_contrib, float *psi, float *elec_dot, float *vbar,
                                  float *elec_forces, double *energy, int32_t *iterations, double *delta);

/* compute_field_and_kinetic_energy (field.py:688-703), real-space sums in FP64:
 * field energy  dV sum_cells w(phi~) from filtered densities phi (n_types,Mx,My,Mz);
 * charge energy dV sum_cells (1/2 phi_q psi - self_energy/V). host double out. */
int hymd_field_energy(hymd_pm *pm, const float *phi, int32_t n_types, const double *coupling, const double *shift,
                      double constant, double *energy);
int hymd_field_q_energy(hymd_pm *pm, const float *phi_q, const float *psi, double self_energy, double *energy);

/* Spectral moments of the filtered densities phi~_t (n_types <= 8), for the
 * energy and its chi / kappa / sigma derivatives (no reference counterpart:
 * the differentiable extension of update_field):
 *   moments = [P_ts (t<=s, row-major), Q_ts (t<=s), phi~_t(k=0)],
 *   P_ts = sum_k Re(phi~_t^* phi~_s), Q_ts = sum_k |k|^2 Re(phi~_t^* phi~_s)
 * over the full spectrum (so dV sum_cells phi~_t phi~_s = V P_ts).  With
 * force_mesh != NULL the force grids of (coupling, shift) are written too.
 * Synchronises the stream. */
int hymd_field_moments(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                       const double *mass, double sigma, double *moments, const double *coupling,
                       const double *shift, float *force_mesh);

/* Field forces of the quadratic form (coupling, shift) and sigma from the
 * density spectra and binning held by the last hymd_field_moments call on
 * this handle (frozen positions, e.g. a parameter sweep; n_types <= 8 and
 * equal to that call's): force_mesh device float32 (n_types,3,Mx,My,Mz) out,
 * force device float32 (n,3) out or NULL, rows as that call's positions.
 * The k-space pass and inverse transforms of update_field (field.py:580-613)
 * and the readout of compute_field_force (:152-200), without binning or
 * paint.  Asynchronous. */
int hymd_field_forces_held(hymd_pm *pm, int32_t n_types, const double *coupling, const double *shift,
                           double sigma, float *force_mesh, float *force);

/* Adjoint of the force chain compute_field_force o update_field with respect
 * to the coupling A and sigma (no reference counterpart: the backward pass of
 * the differentiable field forces).  For a loss L of the field forces
 * F = readout(c2r(-i k H A H rho^)) of (coupling, sigma), and its gradient
 * grad_force = dL/dF (device float32 (n,3), the rows of x):
 *   d_coupling host double (n_types, n_types) out: dL/dA_ts (t: the type of
 *              the particles whose forces are read out, s: the density)
 *   d_sigma    host double out or NULL: dL/dsigma at fixed A
 * Chain: CIC paint of each gradient component per type (the readout's
 * adjoint, int64 fixed point scaled to the largest |g|), r2c of the 3 n_types
 * grids, one reduction against the density spectra.  n_types <= 8.
 * reuse != 0: x and types are not read; the binning and density spectra of
 * the last hymd_field_moments call on this handle are used (same particles,
 * types and masses: frozen positions).  Synchronises the handle's stream. */
int hymd_field_adjoint(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                       const double *mass, double sigma, const double *coupling, const float *grad_force,
                       int32_t reuse, double *d_coupling, double *d_sigma);

/* comp_pressure's field terms (pressure.py:9-179, comp_laplacian
 * pressure.py:182-200) from the filtered spectra of update_field:
 *   phi_fourier device complex64 (n_types,Mx,My,Mz/2+1), pmesh normalisation (n_types <= 8)
 *   coupling, shift, constant: the quadratic form of w (A, b, c0), host
 *   psi, phi_q device float32 (Mx,My,Mz) grids of the PME path, or NULL: with psi,
 *     V_bar_t carries type_charges[t] * psi (host, n_types); with phi_q too,
 *     w carries 1/2 phi_q psi - self_energy / V (coulombtype PIC_Spectral)
 * out host double[5]: p0, p1, p2_x, p2_y, p2_z (the 18-vector's entries 1-5).
 * Parseval sums over the spectra in FP64; no Laplacian grids are formed.
 * Synchronises the handle's stream. */
int hymd_field_pressure(hymd_pm *pm, const void *phi_fourier, int32_t n_types, const double *coupling,
                        const double *shift, double constant, double sigma, const float *psi, const float *phi_q,
                        const double *type_charges, double self_energy, double *out);

/* A fixed dihedral term list (compute_dihedral_forces__double.f90 +
 * dipole_reconstruction__double.f90, force.py:855-877) held on the device with
 * its per-particle lists built once: the loop-side form of hymd_cdf (double)
 * and hymd_dipole_redistribution, without their per-call list build and
 * synchronisation.  Term arrays are referenced, not copied.  forces: f (N,3)
 * float64 overwritten; dipoles (M,4,3) and tmat (M,6,3,3) float64 written as
 * cdf_d does (zero where a term has none); energy (host) optional, its
 * reduction synchronises.  redistribute: f (N,3) float64 overwritten with the
 * dipole forces (M,4,3) projected through tmat onto the beads. */
typedef struct hymd_dihedrals hymd_dihedrals;
int hymd_dihedrals_create(hymd_dihedrals **out, int64_t n, int64_t m, const int32_t *a, const int32_t *b,
                          const int32_t *c, const int32_t *d, const float *coeff, const int32_t *dtype,
                          const int32_t *last, void *stream);
int hymd_dihedrals_destroy(hymd_dihedrals *h);
int hymd_dihedrals_forces(hymd_dihedrals *h, const double *x, double *f, double *dipoles, double *tmat,
                          const double box[3], int32_t dipole_flag, double *energy, void *stream);
int hymd_dihedrals_redistribute(hymd_dihedrals *h, const double *f_dipoles, const double *tmat, double *f,
                                void *stream);
/* The dipoles (M,4,3) and transfer matrices (M,6,3,3) alone at positions x:
 * the dipole outputs of hymd_dihedrals_forces with dipole_flag 1
 * (compute_dihedral_forces__plus__dipoles.f90 via dipole_reconstruction.f90
 * :148-213), without the forces; the fused inner loop's dipoles of a protein
 * come from here after its launch. */
int hymd_dihedrals_dipoles(hymd_dihedrals *h, const double *x, double *dipoles, double *tmat, const double box[3],
                           void *stream);

/* Batched replicas (no reference counterpart: the reference runs one system
 * per process).  n_replicas independent systems, replica r owning particles
 * [offsets[r], offsets[r+1]) of the caller's arrays, mesh[3r..3r+2],
 * box[3r..3r+2] and one mass per type (host arrays, fixed for the handle);
 * n_types x the largest mesh edge along z at most 2048.
 * hymd_batch_moments paints, transforms and reduces every replica to the
 * moments of hymd_field_moments (n_types <= 8; row r of `moments` holds
 * replica r's 2 T(T+1)/2 + T doubles, bit-identical to hymd_field_moments on
 * that replica alone) for per-replica filter widths sigma[r].  With use_graph
 * the launch sequence is captured on the first call (and again when x or
 * types change) and replayed.  The work runs on the handle's own stream after
 * the caller's `stream`; the call synchronises before returning. */
typedef struct hymd_batch hymd_batch;
int hymd_batch_create(hymd_batch **out, int32_t n_replicas, const int32_t *mesh, const double *box,
                      const int64_t *offsets, int32_t n_types, const double *mass);
int hymd_batch_destroy(hymd_batch *b);
int hymd_batch_moments(hymd_batch *b, const float *x, const int32_t *types, const double *sigma, double *moments,
                       int32_t use_graph, void *stream);
/* hymd_batch_moments plus every replica's field forces in the same launch
 * sequence: coupling host double (n_replicas, n_types, n_types), replica r's
 * quadratic form A_r; forces device float32 (n,3) out, the forces of
 * compute_field_force (field.py:152-200) of update_field with A_r and
 * sigma[r] (field.py:580-613), bit-identical to hymd_field_moments with
 * force grids + hymd_compute_field_force on that replica alone.  With
 * use_graph the sequence is captured again when x, types or forces change. */
int hymd_batch_forces(hymd_batch *b, const float *x, const int32_t *types, const double *sigma,
                      const double *coupling, double *moments, float *forces, int32_t use_graph, void *stream);

/* compute_self_energy_q (field.py:203-238): host double out */
int hymd_self_energy_q(const float *charges, int64_t n, double sigma, double elec_factor, double *out,
                       void *stream);

/* Building blocks (used by the above; exported for parity tests and callers
 * that assemble their own pipelines). */
/* CIC paint into int64 fixed-point grids (value = round(w*m*2^32)); grid is
 * (n_types or 1, Mx,My,Mz); if types==NULL all particles go to grid 0 with weight
 * weights[i] (device float32) or mass[0].  accumulate=0 overwrites the grid
 * (cell sort + per-cell gather, the update_field path); accumulate=1 adds to
 * it (one 64-bit integer atomic per particle corner).  Both give identical
 * integers. */
int hymd_paint_fixed(hymd_pm *pm, const float *x, const int32_t *types, const float *weights, int64_t n,
                     int32_t n_types, const double *mass, int64_t *grid, int32_t accumulate);
/* unnormalised batched 3-D r2c / c2r on (batch,Mx,My,Mz) real <-> (batch,Mx,My,Mz/2+1) complex64 */
int hymd_fft_r2c(hymd_pm *pm, const float *in, void *out, int32_t batch);
int hymd_fft_c2r(hymd_pm *pm, const void *in, float *out, int32_t batch);
/* CIC readout of one real grid at n positions */
int hymd_readout(hymd_pm *pm, const float *grid, const float *x, int64_t n, float *out);

/* ------------------------------------------------------------------------ */
/* Bonded kernels: the f2py calls of force.py:8-23 (f2py zeroes nothing but  */
/* the Fortran zeroes its outputs on entry, as these do).  precision selects */
/* cbf (HYMD_F32: f,r float32) or cbf_d (HYMD_F64: f,r float64).  r, f are   */
/* device (n,3) row-major; index/parameter arrays are device arrays.          */
/* energy / pr are host doubles (may be NULL: skips the reduction and sync).  */
/* ------------------------------------------------------------------------ */
int hymd_cbf(void *f, const void *r, int64_t n, const double box[3], const int32_t *a, const int32_t *b,
             const double *r0, const double *k, int64_t m, int32_t precision, double *energy, double *bond_pr,
             void *stream);
/* Allocation-free variant of hymd_cbf, with no counterpart in the reference:
 * for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results as
 * hymd_cbf; the term lists, partial sums and reduced sums live in one static
 * device array of about 16.1 MB per device, held until the process ends.
 * Rejects n > HYMD_NOALLOC_MAX_PARTICLES or m > HYMD_NOALLOC_MAX_BONDS with
 * HYMD_ERR_INVALID.  Every call on a device shares that array, so calls that
 * overlap on one device (other threads, other streams) race. */
#define HYMD_NOALLOC_MAX_PARTICLES (1 << 20)
#define HYMD_NOALLOC_MAX_BONDS (1 << 20)
int hymd_cbf_noalloc(void *f, const void *r, int64_t n, const double box[3], const int32_t *a, const int32_t *b,
                     const double *r0, const double *k, int64_t m, int32_t precision, double *energy,
                     double *bond_pr, void *stream);
int hymd_caf(void *f, const void *r, int64_t n, const double box[3], const int32_t *a, const int32_t *b,
             const int32_t *c, const double *t0, const double *k, int64_t m, int32_t precision, double *energy,
             double *angle_pr, void *stream);
/* coeff device float32 (m,6,5); dipoles device (m,4,3), transfer_matrix device
 * (m,6,3,3) in the precision of r (zeroed, then filled where dtype==1 and
 * dipole_flag!=0); n_clamped host int64 out or NULL: count of sin(gamma) clamps */
int hymd_cdf(void *f, const void *r, void *dipoles, void *transfer_matrix, int64_t n, const double box[3],
             const int32_t *a, const int32_t *b, const int32_t *c, const int32_t *d, const float *coeff,
             const int32_t *dtype, const int32_t *bb_index, int64_t m, int32_t dipole_flag, int32_t precision,
             double *energy, int64_t *n_clamped, void *stream);
/* dipole_forces_redistribution (force.py:855-877): f_on_bead (n,3) out */
int hymd_dipole_redistribution(void *f_on_bead, int64_t n, const void *f_dipoles, const void *transfer_matrix,
                               const int32_t *a, const int32_t *b, const int32_t *c, const int32_t *d,
                               const int32_t *dtype, const int32_t *last, int64_t m, int32_t precision,
                               void *stream);

/* Persistent topology: bonds + angles with the per-particle term lists built
 * once (the rRESPA loop calls the kernels respa_inner times per step on an
 * unchanged topology, main.py:830-881).  The term arrays are referenced, not
 * copied: they must stay allocated and unchanged while the handle is used. */
typedef struct hymd_topology hymd_topology;
int hymd_topology_create(hymd_topology **top, int64_t n, int64_t m2, const int32_t *a2, const int32_t *b2,
                         const double *r0, const double *k2, int64_t m3, const int32_t *a3, const int32_t *b3,
                         const int32_t *c3, const double *t0, const double *k3, void *stream);
int hymd_topology_destroy(hymd_topology *top);
/* total bonded force (bonds + angles) from float64 positions into float32 f;
 * energies (host double[2]: bond, angle) and virials (host double[6]) or NULL */
int hymd_topology_forces(hymd_topology *top, const double *r, float *f, const double box[3], double *energies,
                         double *virials, void *stream);

/* Molecule-blocked fused inner loop: all respa_inner sub-steps (kicks, drift,
 * wrap, bonds + angles (+ dihedrals)) of main.py:830-887 in one launch, one
 * workgroup per block of whole molecules.  Creation fails with
 * HYMD_ERR_UNSUPPORTED when a molecule is not a contiguous index range or does
 * not fit a workgroup.  The term arrays are copied at creation. */
typedef struct hymd_fused hymd_fused;
int hymd_fused_create(hymd_fused **fused, int64_t n, int64_t m2, const int32_t *a2, const int32_t *b2,
                      const double *r0, const double *k2, int64_t m3, const int32_t *a3, const int32_t *b3,
                      const int32_t *c3, const double *t0, const double *k3);
/* The same with the m4 dihedral terms of hymd_cdf / hymd_dihedrals_create
 * (indices a4..d4, coeff (m4,6,5), dtype, last) evaluated every sub-step as
 * main.py:863-881 does: each sub-step's dihedral forces (FP64 arithmetic of
 * hymd_dihedrals_forces, stored in the gather's precision) join the fast
 * forces after the bonds and angles.  The dipoles of a protein are not
 * written by the loop: hymd_dihedrals_dipoles at the final positions gives
 * those of its last sub-step. */
int hymd_fused_create_dihedrals(hymd_fused **fused, int64_t n, int64_t m2, const int32_t *a2, const int32_t *b2,
                                const double *r0, const double *k2, int64_t m3, const int32_t *a3, const int32_t *b3,
                                const int32_t *c3, const double *t0, const double *k3, int64_t m4, const int32_t *a4,
                                const int32_t *b4, const int32_t *c4, const int32_t *d4, const float *coeff,
                                const int32_t *dtype, const int32_t *last);
int hymd_fused_destroy(hymd_fused *fused);
/* v += [pending ? h_out f_slow : 0] + h_out f_slow; then n_inner times
 * { v += h_in f_fast; x += dt v; wrap; f_fast = bonds + angles; v += h_in f_fast }
 * (the second half kick of one sub-step and the first of the next are separate
 * additions); x32 = float(x).  energies host double[2], virials host double[6]
 * of the last sub-step, or NULL. */
/* bond / angle terms of the fused inner loop in FP32 arithmetic (HYMD_F32;
 * the reference's real(4) kernels compute_bond_forces.f90 /
 * compute_angle_forces.f90 difference positions in FP32 and evaluate the
 * terms in FP64) or FP64 (HYMD_F64, as the __double kernels; the default);
 * kicks, drifts, positions and the energy pass stay FP64 */
int hymd_fused_set_precision(hymd_fused *fused, int32_t precision);
int hymd_fused_inner(hymd_fused *fused, double *x, double *v, float *x32, float *f_fast, const float *f_slow,
                     double h_in, double h_out, double dt, int32_t n_inner, int32_t pending, const double box[3],
                     double *energies, double *virials, void *stream);
/* hymd_fused_inner with energies host double[4]: bond, angle, dihedral energy
 * and the count of clamped sin(gamma) of the last sub-step (zero without
 * dihedral terms), or NULL. */
int hymd_fused_inner_dihedrals(hymd_fused *fused, double *x, double *v, float *x32, float *f_fast,
                               const float *f_slow, double h_in, double h_out, double dt, int32_t n_inner,
                               int32_t pending, const double box[3], double *energies, double *virials,
                               void *stream);

/* ------------------------------------------------------------------------ */
/* rRESPA integrator (FP64 x, v): main.py:802-887,1143-1169.                 */
/* ------------------------------------------------------------------------ */
/* v += c_fast*f_fast + c_slow*f_slow (f_slow may be NULL); x += dt*v;
 * x = mod(x, box) (np.mod semantics, main.py:837); x32 = float(x) if non-NULL */
int hymd_respa_kick_drift(double *x, double *v, float *x32, int64_t n, const float *f_fast, double c_fast,
                          const float *f_slow, double c_slow, double dt, const double box[3], void *stream);
/* v += c1*f1 + c2*f2 (f2 may be NULL) */
int hymd_respa_kick(double *v, int64_t n, const float *f1, double c1, const float *f2, double c2, void *stream);
/* Box rescaling of the barostats (barostat.py:145-151, 279-302; barostat_scr.py:125-129):
 * x[r, d] *= factors[d] for the rows [row_begin, row_end) of x (n,3), float32 or float64 per
 * precision (each product in FP64, rounded to the array's type); x32 (float32 (n,3)) receives
 * float(x) of those rows, or NULL */
int hymd_scale_positions(void *x, int32_t precision, float *x32, int64_t n, int64_t row_begin, int64_t row_end,
                         const double factors[3], void *stream);
/* 1/2 m sum v^2 (field.py:695) into host double; v (n,3) float32 or float64 per precision */
int hymd_kinetic_energy(const void *v, int64_t n, int32_t precision, double mass, double *out, void *stream);

/* ------------------------------------------------------------------------ */
/* Thermostat and momentum removal on device FP64 velocities v (n,3).        */
/* Both take a caller-owned device workspace of                              */
/* hymd_thermostat_workspace_bytes() bytes, zero-filled before its first use */
/* (calls that share one workspace must be ordered on one stream), and do    */
/* not synchronise.  hymd_csvr_thermostat keeps each COM group's velocity in */
/* it as the shift of the next call's one-pass sums: give every thermostat   */
/* its own workspace so its results depend on its own history only.         */
/* ------------------------------------------------------------------------ */
int64_t hymd_thermostat_workspace_bytes(void);
/* csvr_thermostat (thermostat.py:111-219).  Coupling group g (0 <= g <
 * n_groups <= 16, in the reference's order) holds the particles whose type t
 * has type_group[t] == g (host int32[n_types], -1: no group); group_n[g]
 * (host) its particle count (all > 0), gaussians[g] and chi_squares[g]
 * (host) the N(0,1) and chi^2(3 group_n[g] - 1) numbers drawn for it;
 * c = exp(-time_step * respa_inner / tau).  With remove_com and
 * group_n[g] > 1 the group's COM-free velocities are rescaled, otherwise
 * every velocity is rescaled from the kinetic energy of all particles (the
 * reference's branches).  work (device double) += the thermostat work
 * (config.thermostat_work). */
int hymd_csvr_thermostat(double *v, const int32_t *types, int64_t n, int32_t n_types, const int32_t *type_group,
                         int32_t n_groups, const int64_t *group_n, const double *gaussians,
                         const double *chi_squares, double mass, double gas_constant, double target_temperature,
                         double c, int32_t remove_com, double *work, void *workspace, int64_t workspace_bytes,
                         void *stream);
/* cancel_com_momentum (thermostat.py:12-15): v -= sum(v) / n_particles */
int hymd_cancel_com_momentum(double *v, int64_t n, int64_t n_particles, void *workspace, int64_t workspace_bytes,
                             void *stream);

/* ------------------------------------------------------------------------ */
/* Bonded term arrays from the connectivity table: prepare_bonds            */
/* (force.py:573-728 over prepare_bonds_old :325-505), terms in the         */
/* reference's order (networkx traversal orders of the molecule graphs).    */
/* ------------------------------------------------------------------------ */
typedef struct hymd_bond_builder hymd_bond_builder;
/* Rows mol_start[m] .. mol_start[m+1]-1 (device int64, n_mol+1) of the
 * device arrays below are molecule m's particles, molecules in np.unique
 * order, rows in the reference's row order:
 *   index  int64 (n)  global index (the graph's node keys)
 *   row    int64 (n)  local index written into the terms, or NULL (= position)
 *   name   int32 (n)  name id
 *   bonds  int64 (n, max_bonds) global indices of bond partners, -1 = none
 * Type tables (host): bond_names (n_bond_types,2) + bond_params (r0, k);
 * angle_names (n_angle_types,3) + angle_params (theta0 in radians, k);
 * dihedral_names (n_dihedral_types,4), dihedral_kind (dih_type),
 * dihedral_coeffs (n_dihedral_types,6,5) as np.resize gives them; name ids
 * of -1 never match.  At most 256 types of each kind.  Builds and walks the
 * molecule graphs, writes the term counts (bonds, angles, dihedrals) into
 * counts (host) and synchronises.  Fails with HYMD_ERR_INVALID and a message
 * starting "KeyError" (a bond partner outside its molecule) or "ValueError"
 * (a dihedral whose reversed copy was stored with another type first), the
 * reference's exceptions. */
int hymd_bond_builder_create(hymd_bond_builder **builder, int64_t n, int64_t n_mol, const int64_t *mol_start,
                             const int64_t *index, const int64_t *row, const int32_t *name, const int64_t *bonds,
                             int32_t max_bonds, const int32_t *bond_names, const double *bond_params,
                             int32_t n_bond_types, const int32_t *angle_names, const double *angle_params,
                             int32_t n_angle_types, const int32_t *dihedral_names, const int32_t *dihedral_kind,
                             const double *dihedral_coeffs, int32_t n_dihedral_types, int64_t counts[3],
                             void *stream);
/* The 16 arrays of prepare_bonds (device, sized by counts; indices int32):
 * bonds a2, b2, equilibrium, strength; angles a3, b3, c3, equilibrium,
 * strength; dihedrals a4, b4, c4, d4, coeff (D,6,5) float64, dih_type, last.
 * type2 / type3 / type4 (optional, NULL = not written): the config entry of
 * each term.  Enqueued on the builder's stream. */
int hymd_bond_builder_fill(hymd_bond_builder *builder, int32_t *a2, int32_t *b2, double *eq2, double *k2, int32_t *a3,
                           int32_t *b3, int32_t *c3, double *eq3, double *k3, int32_t *a4, int32_t *b4, int32_t *c4,
                           int32_t *d4, double *coeff4, int32_t *kind4, int32_t *last4, int32_t *type2,
                           int32_t *type3, int32_t *type4);
int hymd_bond_builder_destroy(hymd_bond_builder *builder);

/* prepare_bonds with a topology (prepare_index_based_bonds, force.py:508-570):
 * molecule m (n_mol, np.unique order) copies the term list of residue
 * template template_of[m] (device int32; -1 or out of range: no terms) with
 * its indices offset by first_row[m] (device int64).  Templates (host):
 * offsets int32 [3][n_templates + 1] (term ranges of bonds, angles,
 * dihedrals, each from 0), idx2 [B][2] / par2 (r0, k), idx3 [A][3] / par3
 * (theta0 in radians, k), idx4 [D][4] / coeff4 [D][6][5] / kind4 (dih_type),
 * 0-based within the molecule.  counts (host) receives the term totals;
 * synchronises.  fill writes prepare_bonds' 16 arrays (last = 0). */
typedef struct hymd_index_bonds hymd_index_bonds;
int hymd_index_bonds_create(hymd_index_bonds **builder, int64_t n_mol, const int64_t *first_row,
                            const int32_t *template_of, int32_t n_templates, const int32_t *offsets,
                            const int32_t *idx2, const double *par2, const int32_t *idx3, const double *par3,
                            const int32_t *idx4, const double *coeff4, const int32_t *kind4, int64_t counts[3],
                            void *stream);
int hymd_index_bonds_fill(hymd_index_bonds *builder, int32_t *a2, int32_t *b2, double *eq2, double *k2, int32_t *a3,
                          int32_t *b3, int32_t *c3, double *eq3, double *k3, int32_t *a4, int32_t *b4, int32_t *c4,
                          int32_t *d4, double *coeff4, int32_t *kind4, int32_t *last4);
int hymd_index_bonds_destroy(hymd_index_bonds *builder);

/* ------------------------------------------------------------------------ */
/* Whole rRESPA outer steps enqueued from C (the MD loop of main.py:764-1169 */
/* restricted to the hot path: kicks, drift + wrap, bonds + angles every     */
/* inner step, update_field + compute_field_force every outer step).  The    */
/* handle references the caller's arrays, the pm handle and the topology.    */
/* ------------------------------------------------------------------------ */
typedef struct hymd_md hymd_md;
int hymd_md_create(hymd_md **md, hymd_pm *pm, hymd_topology *topo, int64_t n, int32_t n_types, double *x, double *v,
                   float *x32, const int32_t *types, float *f_field, float *f_fast, float *force_mesh,
                   const double *coupling, const double *shift, double constant, const double *type_mass,
                   double sigma, double mass, double dt, int32_t respa_inner, const double box[3]);
/* n_steps outer steps ordered after the work queued on `stream`, which then
 * waits for them; use_graph replays one captured step as a hipGraph;
 * energies host double[5] (field, bond, angle, kinetic, 0) after the last step, or NULL */
int hymd_md_step(hymd_md *md, int32_t n_steps, int32_t use_graph, double *energies, void *stream);
/* Adds the peptide terms of main.py:508-588, 802-887, 1044-1095 to the outer
 * step (either may be NULL): PME of `charges` (device float32 (N,), with
 * elec_factor = k_e/eps_r and the self energy) and the dihedral term list
 * `dih` (n_dihedrals terms), whose dipoles, when dipole_charges (device
 * float32 (4 n_dihedrals,)) is given, are rebuilt on the last inner step and
 * go through PME with those charges, their forces redistributed onto the
 * backbone.  f_slow (device float32 (N,3)) receives the summed slow forces
 * the outer kicks use.  Turns the fused inner loop off (it has no dihedral
 * terms).  peptide_state copies the dipoles (M,4,3) and transfer matrices
 * (M,6,3,3) (device float64, optional) and the (dihedral, PME) energies of the
 * last energies step (host double[2], optional). */
int hymd_md_set_peptide(hymd_md *md, const float *charges, double elec_factor, double self_energy,
                        hymd_dihedrals *dih, int64_t n_dihedrals, const float *dipole_charges, float *f_slow);
int hymd_md_peptide_state(hymd_md *md, double *dipoles, double *tmat, double *extra_energies);
int hymd_md_destroy(hymd_md *md);
/* use the fused inner loop (1) or the per-kernel inner loop (0); returns
 * HYMD_ERR_UNSUPPORTED if the topology does not allow fusing */
/* One half of an outer step, launched eagerly, for a coupling between the
 * halves (the barostat, main.py:889-932): part 0 = the pending and the first
 * outer half kick and the inner loop (velocities complete through the last
 * inner half kick); virials host double[6] or NULL = bond then angle virials
 * (bond_pr, angle_pr) of the last inner step.  part 1 = the slow forces at
 * the current positions in the pm's current box, the second outer half kick,
 * then the thermostat / COM coupling of hymd_md_step (one outer step
 * counted); velocities are complete when it returns.  energies as hymd_md_step (part 0: bond, angle; part 1: field,
 * kinetic).  Synchronises only for host outputs. */
int hymd_md_step_half(hymd_md *md, int32_t part, double *virials, double *energies, void *stream);
/* Field updates of the step write the filtered spectra phi_fourier (device
 * complex64 (n_types,Mx,My,Mz/2+1), the pmesh normalisation, or NULL) and the
 * PME psi, phi_q grids (or NULL): the inputs of comp_pressure. */
int hymd_md_set_spectra(hymd_md *md, void *phi_fourier, float *psi, float *phi_q);
int hymd_md_set_fused(hymd_md *md, int32_t enable);
/* Coupling after outer steps, as main.py:1314-1321: after outer step s
 * (counted by the handle; hymd_md_set_step sets the count of steps done) the
 * CSVR thermostat when s % n_b == 0 (hymd_csvr_thermostat's arguments; mass
 * is the handle's; work and workspace as there, kept by the handle; n_b = 0
 * turns it off), then COM momentum removal when s % every == 0
 * (hymd_cancel_com_momentum; every = 0 turns it off).  The thermostat folds
 * the step's pending outer half kick into its passes (same FP64 additions).
 * push_draws queues `count` calls' numbers (gaussians, chi_squares: count x
 * n_groups each, in call order); hymd_md_step fails before enqueuing anything
 * when its steps need more thermostat calls than are queued.  Energies of an
 * energies step are taken after the coupling (the kinetic energy). */
int hymd_md_set_thermostat(hymd_md *md, int32_t n_types, const int32_t *type_group, int32_t n_groups,
                           const int64_t *group_n, double gas_constant, double target_temperature, double c,
                           int32_t remove_com, int32_t n_b, double *work, void *workspace, int64_t workspace_bytes);
int hymd_md_push_draws(hymd_md *md, const double *gaussians, const double *chi_squares, int32_t count);
int hymd_md_set_com_removal(hymd_md *md, int32_t every, int64_t n_particles, void *workspace,
                            int64_t workspace_bytes);
int hymd_md_set_step(hymd_md *md, int64_t step);
/* term precision of the fused inner loop (hymd_fused_set_precision); synchronises */
int hymd_md_set_term_precision(hymd_md *md, int32_t precision);
/* 1 if the fused inner loop is active */
int hymd_md_is_fused(hymd_md *md);
/* Every `every` outer steps (0: never), put the trailing one-particle
 * molecules (rows [hymd_md_single_tail, n), which carry no bonded terms) in
 * the mesh-row order of their current positions, so that the field path's
 * binning writes and force writes stay local as they diffuse.  types must be
 * the array given to hymd_md_create; order (state row -> input row) and where
 * (input row -> state row) are the caller's int64 device arrays; all three
 * are permuted in place with x, v, x32 and the forces.  Fused path only
 * (HYMD_ERR_UNSUPPORTED otherwise).  No counterpart in the reference, which
 * keeps each rank's particles in input order (file_io.py:781-874). */
int hymd_md_set_reorder(hymd_md *md, int32_t every, int32_t *types, int64_t *order, int64_t *where);
/* first row of the one-particle tail that hymd_md_set_reorder permutes */
int64_t hymd_md_single_tail(hymd_md *md);
/* HIP-event timing of one stage (1 bonded forces, 2 kick+drift, 3 update_field,
 * 4 readout, 5 PME of the charges and dipoles; 0 off) in the eager steps that follow, recorded on the launch
 * stream around each launch; at most max_launches are recorded */
int hymd_md_time_stage(hymd_md *md, int32_t stage, int32_t max_launches);
/* total milliseconds and count of the timed launches so far (synchronises) */
int hymd_md_stage_time(hymd_md *md, double *total_ms, int32_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* HYMD_AMD_H */
