// rRESPA outer step enqueued from C++ (main.py:802-887, 975-1004, 1143-1169),
// optionally captured once into a hipGraph and replayed.
//
// Per outer step:  kick_drift (outer half kick + inner half kick + drift + wrap)
//                  -> bonded forces, repeated respa_inner times (later inner
//                  steps kick with the previous and the current half of F_fast),
//                  -> update_field -> readout -> final kick (inner + outer halves).
// The graph holds kernel nodes only: the paint grid is cleared by the z pass
// of the forward FFT, so nothing in the step is a memset.
#include <cstring>
#include <vector>

#include "common.h"

using namespace hymd;

struct hymd_md {
    hymd_pm *pm = nullptr;
    hymd_topology *topo = nullptr;
    hipStream_t work = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // per-kernel path: the last inner step's bonded forces run on `side`
    // beside the field update (both read the positions after the last drift)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // captured outer steps: [1] starts with the last step's outer half kick
    // pending (fused path, replayed back to back), [0] without (per-kernel
    // path, or the fused path after the velocities were completed)
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    // pm_generation(pm) when the graphs were captured: they hold the pm's
    // buffer pointers and its mesh by value
    uint64_t pm_gen = 0;
    int64_t n = 0;
    int T = 0;
    double *x = nullptr, *v = nullptr;
    float *x32 = nullptr, *f_field = nullptr, *f_fast = nullptr, *force_mesh = nullptr;
    const int32_t *types = nullptr;
    double A[HYMD_MAX_TYPES * HYMD_MAX_TYPES], b[HYMD_MAX_TYPES], c0 = 0.0, mass_t[HYMD_MAX_TYPES];
    double sigma = 0.0, mass = 1.0, dt = 0.0, box[3] = {0, 0, 0};
    int respa_inner = 1;
    hymd_fused *fused = nullptr;  // fused inner loop (null: per-kernel path)
    int pending = 0;              // fused: outer half kick of the last step not yet applied
    // outputs for a pressure between the step's halves (hymd_md_set_spectra):
    // every field update writes the filtered spectra (and PME psi, phi_q)
    void *spectra = nullptr;
    float *psi = nullptr, *phi_q = nullptr;
    int term_precision = HYMD_F64; // fused inner loop: bond / angle terms in FP64 or FP32
    // coupling after outer steps (main.py:1314-1321): CSVR thermostat every
    // n_b steps, COM momentum removal every com_every steps; `step` counts the
    // outer steps done, draws holds 2 n_groups numbers (R, then chi^2) per call
    int64_t step = 0;
    int n_b = 0;
    CsvrSetup thermo;
    double *thermo_work = nullptr, *thermo_ws = nullptr;
    std::vector<double> draws;
    size_t draw_next = 0;
    int com_every = 0;
    int64_t com_n = 0;
    double *com_ws = nullptr;
    // peptide terms (optional): PME of the charges, dihedrals with dipoles
    // whose charges go through PME and back onto the backbone
    const float *charges = nullptr;
    double elec = 0.0, e_self = 0.0;
    hymd_dihedrals *dih = nullptr;
    int64_t m4 = 0;
    bool protein = false;
    const float *q_dip = nullptr;
    // a second mesh handle (same mesh and box, own buffers) for the dipole
    // charges' PME on the side stream, beside the field update on `pm`
    hymd_pm *pm2 = nullptr;
    float *f_slow = nullptr;  // caller's: summed slow forces
    DevBuf f_ba, f_dih, dipoles, tmat, f_dip, dip32, f_dip64, f_recon;
    double extra_e[2] = {0.0, 0.0};  // dihedral, field_q of the last energies step
    // periodic reorder of the one-particle molecules (fused path): every
    // reorder_every outer steps the rows [single0, n) are put in the mesh-row
    // order of their current positions, so the binning scatter and the
    // readout's writes stay local as the solvent diffuses
    int reorder_every = 0, since_reorder = 0;
    int64_t single0 = 0;
    int32_t *types_rw = nullptr;
    int64_t *order = nullptr, *where = nullptr;  // caller's: state row -> input row, input row -> state row
    DevBuf rperm, rtmp;
    // optional HIP-event timing of one stage in eager steps
    int timed = 0;  // 0 none, 1 bonded forces, 2 kick+drift, 3 update_field, 4 readout, 5 PME (+ dipoles)
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
};

// rows s0 + j <- rows s0 + perm[j] (j < ns) of the fused-path state, through
// tmp: gather, then store back and point where[] at the new rows.
struct ReorderArrays {
    double *x, *v;
    float *x32, *ff, *fs;
    int32_t *types;
    int64_t *order, *where;
};

// One row of the reordered fields, loaded whole before it is stored (the
// arrays are distinct, but the compiler cannot know it: loads first keep all
// of a row's loads in flight together).
struct ReorderRow {
    double x[3], v[3];
    float x32[3], ff[3], fs[3];
    int32_t type;
    int64_t order;
};

__device__ __forceinline__ ReorderRow load_row(const ReorderArrays &a, int64_t i) {
    ReorderRow r;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        r.x[d] = a.x[3 * i + d];
        r.v[d] = a.v[3 * i + d];
        r.x32[d] = a.x32[3 * i + d];
        r.ff[d] = a.ff[3 * i + d];
        r.fs[d] = a.fs[3 * i + d];
    }
    r.type = a.types[i];
    r.order = a.order[i];
    return r;
}

__device__ __forceinline__ void store_row(const ReorderArrays &a, int64_t i, const ReorderRow &r) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        a.x[3 * i + d] = r.x[d];
        a.v[3 * i + d] = r.v[d];
        a.x32[3 * i + d] = r.x32[d];
        a.ff[3 * i + d] = r.ff[d];
        a.fs[3 * i + d] = r.fs[d];
    }
    a.types[i] = r.type;
    a.order[i] = r.order;
}

__global__ void __launch_bounds__(kBlock) reorder_gather_kernel(int64_t ns, int64_t s0, const int *__restrict__ perm,
                                                                ReorderArrays a, ReorderArrays t) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < ns; j += stride)
        store_row(t, j, load_row(a, s0 + perm[j]));
}

__global__ void __launch_bounds__(kBlock) reorder_store_kernel(int64_t ns, int64_t s0, ReorderArrays t,
                                                               ReorderArrays a) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < ns; j += stride) {
        const ReorderRow r = load_row(t, j);
        store_row(a, s0 + j, r);
        a.where[r.order] = s0 + j;
    }
}

static int reorder_singles(hymd_md *m) {
    const int64_t s0 = m->single0, ns = m->n - s0;
    if (ns <= 1) return HYMD_OK;
    hipStream_t s = m->work;
    HYMD_TRY(m->rperm.reserve(sizeof(int) * ns));
    HYMD_TRY(m->rtmp.reserve((sizeof(double) * 6 + sizeof(float) * 9 + sizeof(int32_t) + sizeof(int64_t)) * ns));
    HYMD_TRY(hymd_pm_set_stream(m->pm, s));
    HYMD_TRY(row_order(m->pm, m->x32 + 3 * s0, ns, (int *)m->rperm.ptr, s));
    char *p = (char *)m->rtmp.ptr;
    ReorderArrays t;
    t.x = (double *)p;
    t.v = t.x + 3 * ns;
    t.order = (int64_t *)(t.v + 3 * ns);
    t.x32 = (float *)(t.order + ns);
    t.ff = t.x32 + 3 * ns;
    t.fs = t.ff + 3 * ns;
    t.types = (int32_t *)(t.fs + 3 * ns);
    t.where = nullptr;
    ReorderArrays a{m->x, m->v, m->x32, m->f_field, m->f_fast, m->types_rw, m->order, m->where};
    reorder_gather_kernel<<<grid_for(ns), kBlock, 0, s>>>(ns, s0, (const int *)m->rperm.ptr, a, t);
    HYMD_HIP(hipGetLastError());
    reorder_store_kernel<<<grid_for(ns), kBlock, 0, s>>>(ns, s0, t, a);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

static void mark(hymd_md *m, int stage) {
    if (m->timed == stage && m->ev_used < m->ev.size()) (void)hipEventRecord(m->ev[m->ev_used++], m->work);
}

extern "C" int hymd_md_set_fused(hymd_md *m, int32_t enable);

// The slow forces the outer kicks apply: the field forces alone, or the
// caller's sum with the PME (and redistributed dipole) forces.
static float *slow_force(hymd_md *m) { return (m->charges || m->dih) ? m->f_slow : m->f_field; }

// Slow forces beyond the field ones (main.py:1044-1095): f_slow = f_field
// + PME forces of the charges + the dipole-charge PME forces redistributed
// onto the backbone, in the order (and rounding) of the Python step.
// The field readout, then f_slow = f_field (a coalesced copy: a second
// scattered store from the readout costs more).
static int slow_readout(hymd_md *m) {
    HYMD_TRY(field_readout_last(m->pm, m->T, m->force_mesh, m->f_field));
    if (m->charges || m->dih) HYMD_TRY(force_combine(m->f_slow, m->f_field, nullptr, 3 * m->n, 3, m->work));
    return HYMD_OK;
}

// Part one: f_slow += the charges' PME forces (read out into f_slow, each
// a float32 addition to the field force as a separate sum pass would make).
static int charge_slow_forces(hymd_md *m, bool energy) {
    if (m->charges)
        HYMD_TRY(update_field_force_q_impl(m->pm, m->x32, m->charges, m->n, m->sigma, m->elec, m->e_self, m->f_slow,
                                           m->phi_q, m->psi, nullptr, energy ? &m->extra_e[1] : nullptr, true));
    return HYMD_OK;
}

// Part two: the dipole charges' PME forces redistributed onto the backbone
// (reads the dipoles of the last inner step) into f_recon, on stream s with
// the mesh handle pmd (m->pm on the work stream, or m->pm2 on the side
// stream), then f_slow += f_recon on the work stream.
static int dipole_recon_forces(hymd_md *m, hymd_pm *pmd, hipStream_t s) {
    if (!m->protein) return HYMD_OK;
    const int64_t nd = 4 * m->m4;
    HYMD_TRY(convert_f64_f32((float *)m->dip32.ptr, (double *)m->dipoles.ptr, 3 * nd, true, s));
    HYMD_TRY(hymd_update_field_force_q(pmd, (const float *)m->dip32.ptr, m->q_dip, nd, m->sigma, m->elec, 0.0,
                                       (float *)m->f_dip.ptr, nullptr, nullptr, nullptr, nullptr));
    HYMD_TRY(convert_f64_f32((float *)m->f_dip.ptr, (double *)m->f_dip64.ptr, 3 * nd, false, s));
    return hymd_dihedrals_redistribute(m->dih, (const double *)m->f_dip64.ptr, (const double *)m->tmat.ptr,
                                       (double *)m->f_recon.ptr, s);
}

static int dipole_combine(hymd_md *m) {
    if (!m->protein) return HYMD_OK;
    return force_combine(m->f_slow, nullptr, (const double *)m->f_recon.ptr, 3 * m->n, 2, m->work);
}

static int dipole_slow_forces(hymd_md *m) {
    HYMD_TRY(dipole_recon_forces(m, m->pm, m->work));
    return dipole_combine(m);
}

static int peptide_slow_forces(hymd_md *m, bool energy) {
    HYMD_TRY(charge_slow_forces(m, energy));
    return dipole_slow_forces(m);
}

// The same slow forces beside the field update: on the side stream with the
// second mesh handle, f_recon = the dipoles' redistributed PME forces and,
// with `charges` (the fused path, whose side stream has nothing else),
// f_slow = the charges' PME forces (zero where uncharged).  After the field
// readout into f_field, side_join_slow_forces forms on the work stream
// f_slow = elec + field (the float32 sum field + elec: addition commutes),
// or field + elec with the charges' PME on the work stream, then waits for
// the side stream and adds f_recon.
static int side_slow_forces(hymd_md *m, bool charges) {
    if (charges && m->charges)
        HYMD_TRY(update_field_force_q_impl(m->pm2, m->x32, m->charges, m->n, m->sigma, m->elec, m->e_self, m->f_slow,
                                           m->phi_q, m->psi, nullptr, nullptr, false));
    return dipole_recon_forces(m, m->pm2, m->side);
}

static int side_join_slow_forces(hymd_md *m, bool charges_on_side) {
    if (charges_on_side) {
        HYMD_HIP(hipStreamWaitEvent(m->work, m->ev_join, 0));
        HYMD_TRY(force_combine(m->f_slow, m->f_field, nullptr, 3 * m->n, m->charges ? 1 : 3, m->work));
    } else {
        HYMD_TRY(force_combine(m->f_slow, m->f_field, nullptr, 3 * m->n, 3, m->work));
        HYMD_TRY(charge_slow_forces(m, false));
        HYMD_HIP(hipStreamWaitEvent(m->work, m->ev_join, 0));
    }
    return dipole_combine(m);
}

// Bonded forces of one inner step at the current x on stream s: bonds +
// angles, then with dihedrals (dipoles on the last inner step, for a
// protein) f_fast = f_ba + f_dih.  eb (host double[2]) with energies: the
// bond / angle energies, the dihedral one into extra_e[0].
static int bonded_forces(hymd_md *m, bool last_inner, double *eb, hipStream_t s) {
    HYMD_TRY(hymd_topology_forces(m->topo, m->x, m->dih ? (float *)m->f_ba.ptr : m->f_fast, m->box, eb, nullptr, s));
    if (m->dih) {
        HYMD_TRY(hymd_dihedrals_forces(m->dih, m->x, (double *)m->f_dih.ptr, (double *)m->dipoles.ptr,
                                       (double *)m->tmat.ptr, m->box, (last_inner && m->protein) ? 1 : 0,
                                       eb ? &m->extra_e[0] : nullptr, s));
        HYMD_TRY(force_combine(m->f_fast, (const float *)m->f_ba.ptr, (const double *)m->f_dih.ptr, 3 * m->n, 0, s));
    }
    return HYMD_OK;
}

// The fused inner loop of one outer step; with dihedrals also their energy of
// the last sub-step (when energies) and, for a protein, the dipoles and
// transfer matrices of the final positions (main.py:863-881 reconstructs them
// on the last inner step; the kernel leaves them to hymd_dihedrals_dipoles).
static int fused_inner(hymd_md *m, int pending, double *energies, double *virials, const RowKeyTarget *key) {
    const double half_in = 0.5 * m->dt / m->mass;
    const double half_out = 0.5 * m->respa_inner * m->dt / m->mass;
    double eb[2], ed[2];
    FusedDihOut dout{ed};
    HYMD_TRY(fused_inner_impl(m->fused, m->x, m->v, m->x32, m->f_fast, slow_force(m), half_in, half_out, m->dt,
                              m->respa_inner, pending, m->box, energies ? eb : nullptr, virials, m->work, key,
                              (m->dih && energies) ? &dout : nullptr));
    if (m->protein)
        HYMD_TRY(hymd_dihedrals_dipoles(m->dih, m->x, (double *)m->dipoles.ptr, (double *)m->tmat.ptr, m->box,
                                        m->work));
    if (energies) {
        energies[1] = eb[0];
        energies[2] = eb[1];
        if (m->dih) m->extra_e[0] = ed[0];
    }
    return HYMD_OK;
}

// energies (host, 5): field, bond, angle, kinetic, (unused = 0)
static int enqueue_step(hymd_md *m, double *energies, int pending) {
    const double half_in = 0.5 * m->dt / m->mass;
    const double half_out = 0.5 * m->respa_inner * m->dt / m->mass;
    HYMD_TRY(hymd_pm_set_stream(m->pm, m->work));
    if (m->fused) {
        // fused: the step ends after the slow forces; its outer half kick is
        // applied by the next step's kernel (or by finish_pending)
        mark(m, 1);
        // the inner loop also writes the field path's row keys of the new positions
        RowKeyTarget key;
        HYMD_TRY(field_key_target(m->pm, m->n, &key));
        HYMD_TRY(fused_inner(m, pending, energies, nullptr, &key));
        mark(m, 1);
        // the PME of the charges and of the dipoles (second mesh handle) on
        // the side stream, beside the field update, except in steps with
        // energies (their reductions synchronise) or stage timing
        const bool side = m->pm2 && !energies && !m->timed;
        if (side) {
            HYMD_HIP(hipEventRecord(m->ev_fork, m->work));
            HYMD_HIP(hipStreamWaitEvent(m->side, m->ev_fork, 0));
            HYMD_TRY(side_slow_forces(m, true));
            HYMD_HIP(hipEventRecord(m->ev_join, m->side));
        }
        mark(m, 3);
        HYMD_TRY(update_field_keyed(m->pm, m->x32, m->types, m->n, m->T, m->A, m->b, m->c0, m->mass_t, m->sigma,
                                    m->force_mesh, energies ? &energies[0] : nullptr, m->spectra));
        mark(m, 3);
        if (side) {
            HYMD_TRY(field_readout_last(m->pm, m->T, m->force_mesh, m->f_field));
            HYMD_TRY(side_join_slow_forces(m, true));
            return HYMD_OK;
        }
        mark(m, 4);
        HYMD_TRY(slow_readout(m));
        mark(m, 4);
        // PME of the charges (field.py:241-403) into f_slow = f_field + f_elec
        if (m->charges || m->dih) {
            mark(m, 5);
            HYMD_TRY(peptide_slow_forces(m, energies != nullptr));
            mark(m, 5);
        }
        return HYMD_OK;
    }
    const bool extra = m->charges || m->dih;
    float *f_slow = extra ? m->f_slow : m->f_field;
    // the last inner step's bonded forces beside the field update and the
    // charges' PME (all read the positions after the last drift; the
    // dipoles' PME waits for them), except in steps with energies (their
    // reductions synchronise) or stage timing
    const bool overlap = !energies && !m->timed;
    for (int i = 0; i < m->respa_inner; ++i) {
        const bool first = (i == 0), last_inner = (i == m->respa_inner - 1);
        mark(m, 2);
        HYMD_TRY(hymd_respa_kick_drift(m->x, m->v, m->x32, m->n, m->f_fast, half_in, first ? f_slow : m->f_fast,
                                       first ? half_out : half_in, m->dt, m->box, m->work));
        mark(m, 2);
        if (last_inner && overlap) {
            HYMD_HIP(hipEventRecord(m->ev_fork, m->work));
            HYMD_HIP(hipStreamWaitEvent(m->side, m->ev_fork, 0));
            HYMD_TRY(bonded_forces(m, true, nullptr, m->side));
            // the dipoles' PME (second mesh handle) follows there; the charges'
            // PME stays on the work stream after the field update (measured:
            // the side stream already carries the dihedral kernel)
            if (m->pm2) HYMD_TRY(side_slow_forces(m, false));
            HYMD_HIP(hipEventRecord(m->ev_join, m->side));
            break;
        }
        const bool last_e = energies && last_inner;
        double eb[2];
        mark(m, 1);
        HYMD_TRY(bonded_forces(m, last_inner, last_e ? eb : nullptr, m->work));
        mark(m, 1);
        if (last_e) {
            energies[1] = eb[0];
            energies[2] = eb[1];
        }
    }
    mark(m, 3);
    HYMD_TRY(hymd_update_field(m->pm, m->x32, m->types, m->n, m->T, m->A, m->b, m->c0, m->mass_t, m->sigma,
                               m->force_mesh, nullptr, m->spectra, nullptr, energies ? &energies[0] : nullptr));
    mark(m, 3);
    if (overlap && m->pm2) {
        HYMD_TRY(field_readout_last(m->pm, m->T, m->force_mesh, m->f_field));
        HYMD_TRY(side_join_slow_forces(m, false));
    } else {
        mark(m, 4);
        HYMD_TRY(slow_readout(m));
        mark(m, 4);
        if (extra) {
            mark(m, 5);
            HYMD_TRY(charge_slow_forces(m, energies != nullptr));
        }
        if (overlap) HYMD_HIP(hipStreamWaitEvent(m->work, m->ev_join, 0));
        if (extra) {
            HYMD_TRY(dipole_slow_forces(m));
            mark(m, 5);
        }
    }
    HYMD_TRY(hymd_respa_kick(m->v, m->n, m->f_fast, half_in, f_slow, half_out, m->work));
    return HYMD_OK;
}

// The two halves of an outer step, eager, for a coupling between them (the
// barostat of main.py:889-932 acts after the inner loop, before the slow
// forces).  part 0: the pending outer half kick, the first outer half kick and
// the inner loop (the per-kernel path also applies the last inner half kick
// here); virials (host double[6] or NULL): bond and angle virials of the last
// inner step.  part 1: the slow forces at the current positions (no row keys
// from the inner loop: the barostat may have moved the positions), with the
// filtered spectra when hymd_md_set_spectra asked for them, then the second
// outer half kick (left pending on the fused path).  energies as enqueue_step.
static int enqueue_half(hymd_md *m, int part, double *virials, double *energies) {
    const double half_in = 0.5 * m->dt / m->mass;
    const double half_out = 0.5 * m->respa_inner * m->dt / m->mass;
    const bool extra = m->charges || m->dih;
    float *f_slow = extra ? m->f_slow : m->f_field;
    const int64_t n3 = 3 * m->n;
    HYMD_TRY(hymd_pm_set_stream(m->pm, m->work));
    if (part == 0) {
        if (m->fused) {
            HYMD_TRY(fused_inner(m, m->pending, energies, virials, nullptr));
            m->pending = 0;
            return HYMD_OK;
        }
        for (int i = 0; i < m->respa_inner; ++i) {
            const bool first = (i == 0), last_inner = (i == m->respa_inner - 1);
            HYMD_TRY(hymd_respa_kick_drift(m->x, m->v, m->x32, m->n, m->f_fast, half_in, first ? f_slow : m->f_fast,
                                           first ? half_out : half_in, m->dt, m->box, m->work));
            const bool last_e = (energies || virials) && last_inner;
            double eb[2];
            HYMD_TRY(hymd_topology_forces(m->topo, m->x, m->dih ? (float *)m->f_ba.ptr : m->f_fast, m->box,
                                          last_e ? eb : nullptr, last_inner ? virials : nullptr, m->work));
            if (m->dih) {
                HYMD_TRY(hymd_dihedrals_forces(m->dih, m->x, (double *)m->f_dih.ptr, (double *)m->dipoles.ptr,
                                               (double *)m->tmat.ptr, m->box, (last_inner && m->protein) ? 1 : 0,
                                               (energies && last_inner) ? &m->extra_e[0] : nullptr, m->work));
                HYMD_TRY(force_combine(m->f_fast, (const float *)m->f_ba.ptr, (const double *)m->f_dih.ptr, n3, 0,
                                       m->work));
            }
            if (energies && last_inner) {
                energies[1] = eb[0];
                energies[2] = eb[1];
            }
        }
        return hymd_respa_kick(m->v, m->n, m->f_fast, half_in, nullptr, 0.0, m->work);
    }
    HYMD_TRY(hymd_update_field(m->pm, m->x32, m->types, m->n, m->T, m->A, m->b, m->c0, m->mass_t, m->sigma,
                               m->force_mesh, nullptr, m->spectra, nullptr, energies ? &energies[0] : nullptr));
    HYMD_TRY(slow_readout(m));
    if (extra) HYMD_TRY(peptide_slow_forces(m, energies != nullptr));
    if (m->fused) {
        m->pending = 1;
        return HYMD_OK;
    }
    return hymd_respa_kick(m->v, m->n, f_slow, half_out, nullptr, 0.0, m->work);
}

extern "C" int hymd_md_create(hymd_md **out, hymd_pm *pm, hymd_topology *topo, int64_t n, int32_t n_types,
                              double *x, double *v, float *x32, const int32_t *types, float *f_field, float *f_fast,
                              float *force_mesh, const double *coupling, const double *shift, double constant,
                              const double *type_mass, double sigma, double mass, double dt, int32_t respa_inner,
                              const double box[3]) {
    HYMD_REQUIRE(out && pm && topo && box && coupling && n >= 0, "bad argument");
    HYMD_REQUIRE(n == 0 || (x && v && x32 && types && f_field && f_fast && force_mesh), "null array");
    HYMD_REQUIRE(n_types >= 1 && n_types <= HYMD_MAX_TYPES, "n_types out of range");
    HYMD_REQUIRE(respa_inner >= 1 && mass > 0.0 && dt > 0.0, "respa_inner, mass and dt must be positive");
    hymd_md *m = new hymd_md();
    m->pm = pm;
    m->topo = topo;
    m->n = n;
    m->T = n_types;
    m->x = x;
    m->v = v;
    m->x32 = x32;
    m->types = types;
    m->f_field = f_field;
    m->f_fast = f_fast;
    m->force_mesh = force_mesh;
    for (int i = 0; i < n_types * n_types; ++i) m->A[i] = coupling[i];
    for (int t = 0; t < n_types; ++t) {
        m->b[t] = shift ? shift[t] : 0.0;
        m->mass_t[t] = type_mass ? type_mass[t] : 1.0;
    }
    m->c0 = constant;
    m->sigma = sigma;
    m->mass = mass;
    m->dt = dt;
    m->respa_inner = respa_inner;
    for (int d = 0; d < 3; ++d) m->box[d] = box[d];
    m->pm_gen = pm_generation(pm);
    hipError_t e = hipStreamCreateWithFlags(&m->work, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev_out, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
        set_error("stream/event creation failed: %s", hipGetErrorString(e));
        (void)hymd_md_destroy(m);
        return HYMD_ERR_HIP;
    }
    // fuse the inner loop when the topology allows it; otherwise keep the
    // per-kernel path (an unsupported topology is not an error)
    const int st = hymd_md_set_fused(m, 1);
    if (st == HYMD_ERR_HIP || st == HYMD_ERR_NOMEM) {
        (void)hymd_md_destroy(m);
        return st;
    }
    *out = m;
    return HYMD_OK;
}

static void drop_graphs(hymd_md *m) {
    for (int i = 0; i < 2; ++i) {
        if (m->exec[i]) (void)hipGraphExecDestroy(m->exec[i]);
        if (m->graph[i]) (void)hipGraphDestroy(m->graph[i]);
        m->exec[i] = nullptr;
        m->graph[i] = nullptr;
    }
}

extern "C" int hymd_md_destroy(hymd_md *m) {
    if (!m) return HYMD_OK;
    if (m->work) (void)hipStreamSynchronize(m->work);
    if (m->side) (void)hipStreamSynchronize(m->side);
    for (hipEvent_t e : m->ev) (void)hipEventDestroy(e);
    if (m->fused) (void)hymd_fused_destroy(m->fused);
    if (m->pm2) (void)hymd_pm_destroy(m->pm2);
    drop_graphs(m);
    if (m->ev_in) (void)hipEventDestroy(m->ev_in);
    if (m->ev_out) (void)hipEventDestroy(m->ev_out);
    if (m->ev_fork) (void)hipEventDestroy(m->ev_fork);
    if (m->ev_join) (void)hipEventDestroy(m->ev_join);
    if (m->side) (void)hipStreamDestroy(m->side);
    if (m->work) (void)hipStreamDestroy(m->work);
    delete m;
    return HYMD_OK;
}

// capture one outer step starting with (pending = 1) or without a pending
// outer half kick into graph[pending]
static int capture(hymd_md *m, int pending) {
    const int timed = m->timed;
    m->timed = 0;  // no event records inside the graph
    HYMD_HIP(hipStreamBeginCapture(m->work, hipStreamCaptureModeThreadLocal));
    int st = enqueue_step(m, nullptr, pending);
    m->timed = timed;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(m->work, &g);
    if (st != HYMD_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
        return HYMD_ERR_HIP;
    }
    m->graph[pending] = g;
    HYMD_HIP(hipGraphInstantiate(&m->exec[pending], g, nullptr, nullptr, 0));
    m->pm_gen = pm_generation(m->pm);
    return HYMD_OK;
}

// The pm handle moved a buffer or took a new box since the graphs were
// captured (a call on the exposed pm with more particles, a barostat's
// hymd_pm_set_box): follow its box and capture again.
static void check_pm(hymd_md *m) {
    const uint64_t g = pm_generation(m->pm);
    if (g == m->pm_gen) return;
    drop_graphs(m);
    pm_box(m->pm, m->box);
    if (m->pm2) (void)hymd_pm_set_box(m->pm2, m->box);  // the dipoles' handle follows the box
    m->pm_gen = g;
}

// Runs n_steps outer steps ordered after the work already queued on `stream`,
// and makes `stream` wait for them.  use_graph != 0 replays a captured graph of
// one step.  energies (host double[5]) or NULL: energies after the last step
// (that step runs eagerly).
// fused path: apply the outer half kick left pending by the last step
static int finish_pending(hymd_md *m) {
    if (!m->fused || !m->pending) return HYMD_OK;
    const double half_out = 0.5 * m->respa_inner * m->dt / m->mass;
    HYMD_TRY(hymd_respa_kick(m->v, m->n, slow_force(m), half_out, nullptr, 0.0, m->work));
    m->pending = 0;
    return HYMD_OK;
}

// Runs n_steps outer steps ordered after the work already queued on `stream`,
// and makes `stream` wait for them.  use_graph != 0 replays a captured graph of
// one step.  energies (host double[5]) or NULL: energies after the last step
// (that step runs eagerly).  Velocities are complete when the call returns.
// after every outer step: thermostat (the pending outer half kick folded into
// its passes), then COM momentum removal, each on its period
static int couple_step(hymd_md *m) {
    ++m->step;
    const bool th = m->n_b > 0 && m->step % m->n_b == 0;
    const bool com = m->com_every > 0 && m->step % m->com_every == 0;
    if (th) {
        const int G = m->thermo.gp.n_groups;
        const double *R = m->draws.data() + m->draw_next * 2 * G;
        const bool pend = m->fused && m->pending;
        const double half_out = 0.5 * m->respa_inner * m->dt / m->mass;
        HYMD_TRY(csvr_enqueue(m->thermo, m->v, m->types, m->n, R, R + G, pend ? slow_force(m) : nullptr,
                              pend ? half_out : 0.0, m->thermo_work, m->thermo_ws, m->work));
        m->pending = 0;
        if (++m->draw_next * 2 * G >= m->draws.size()) {
            m->draws.clear();
            m->draw_next = 0;
        }
    }
    if (com) {
        HYMD_TRY(finish_pending(m));
        HYMD_TRY(com_enqueue(m->v, m->n, m->com_n, m->com_ws, m->work));
    }
    return HYMD_OK;
}

extern "C" int hymd_md_step(hymd_md *m, int32_t n_steps, int32_t use_graph, double *energies, void *stream) {
    HYMD_REQUIRE(m && n_steps >= 0, "bad argument");
    if (m->n_b > 0) {
        const int64_t need = (m->step + n_steps) / m->n_b - m->step / m->n_b;
        const int G = m->thermo.gp.n_groups;
        const int64_t have = (int64_t)(m->draws.size() / (2 * G)) - (int64_t)m->draw_next;
        HYMD_REQUIRE(need <= have, "the thermostat needs %lld draw sets for these steps, %lld were pushed",
                     (long long)need, (long long)have);
    }
    hipStream_t caller = (hipStream_t)stream;
    HYMD_HIP(hipEventRecord(m->ev_in, caller));
    HYMD_HIP(hipStreamWaitEvent(m->work, m->ev_in, 0));
    const int plain = energies ? n_steps - 1 : n_steps;
    for (int s = 0; s < plain; ++s) {
        check_pm(m);  // an eager step may have grown a pm buffer too
        if (use_graph) {
            const int gi = m->fused ? m->pending : 0;
            if (!m->exec[gi]) HYMD_TRY(capture(m, gi));
            HYMD_HIP(hipGraphLaunch(m->exec[gi], m->work));
        } else {
            HYMD_TRY(enqueue_step(m, nullptr, m->pending));
        }
        if (m->fused) m->pending = 1;
        if (m->fused && m->reorder_every > 0 && ++m->since_reorder >= m->reorder_every) {
            HYMD_TRY(reorder_singles(m));
            m->since_reorder = 0;
        }
        HYMD_TRY(couple_step(m));
    }
    if (energies && n_steps > 0) {
        check_pm(m);
        std::memset(energies, 0, 5 * sizeof(double));
        HYMD_TRY(enqueue_step(m, energies, m->pending));
        if (m->fused) m->pending = 1;
        HYMD_TRY(finish_pending(m));
        HYMD_TRY(couple_step(m));
        HYMD_TRY(hymd_kinetic_energy(m->v, m->n, HYMD_F64, m->mass, &energies[3], m->work));
    }
    HYMD_TRY(finish_pending(m));
    HYMD_HIP(hipEventRecord(m->ev_out, m->work));
    HYMD_HIP(hipStreamWaitEvent(caller, m->ev_out, 0));
    return HYMD_OK;
}

extern "C" int hymd_md_step_half(hymd_md *m, int32_t part, double *virials, double *energies, void *stream) {
    HYMD_REQUIRE(m && (part == 0 || part == 1), "bad argument");
    HYMD_REQUIRE(part == 0 || !virials, "virials come with part 0");
    if (part == 1 && m->n_b > 0) {
        const int G = m->thermo.gp.n_groups;
        const int64_t have = (int64_t)(m->draws.size() / (2 * G)) - (int64_t)m->draw_next;
        const int64_t need = (m->step + 1) % m->n_b == 0 ? 1 : 0;
        HYMD_REQUIRE(need <= have, "the thermostat needs a draw set for this step, none was pushed");
    }
    hipStream_t caller = (hipStream_t)stream;
    HYMD_HIP(hipEventRecord(m->ev_in, caller));
    HYMD_HIP(hipStreamWaitEvent(m->work, m->ev_in, 0));
    check_pm(m);  // a barostat between the halves moves the pm's box
    if (energies) std::memset(energies, 0, 5 * sizeof(double));
    HYMD_TRY(enqueue_half(m, part, virials, energies));
    if (part == 1) {  // velocities are complete when part 1 returns, as after hymd_md_step
        HYMD_TRY(couple_step(m));
        HYMD_TRY(finish_pending(m));
        if (energies) HYMD_TRY(hymd_kinetic_energy(m->v, m->n, HYMD_F64, m->mass, &energies[3], m->work));
    }
    HYMD_HIP(hipEventRecord(m->ev_out, m->work));
    HYMD_HIP(hipStreamWaitEvent(caller, m->ev_out, 0));
    return HYMD_OK;
}

extern "C" int hymd_md_set_spectra(hymd_md *m, void *phi_fourier, float *psi, float *phi_q) {
    HYMD_REQUIRE(m, "null handle");
    HYMD_REQUIRE(!phi_q || psi, "phi_q comes with psi");
    HYMD_HIP(hipStreamSynchronize(m->work));
    m->spectra = phi_fourier;
    m->psi = psi;
    m->phi_q = phi_q;
    drop_graphs(m);
    return HYMD_OK;
}

extern "C" int hymd_md_set_fused(hymd_md *m, int32_t enable) {
    HYMD_REQUIRE(m, "null handle");
    HYMD_HIP(hipStreamSynchronize(m->work));
    drop_graphs(m);
    if (!enable) {
        if (m->fused) (void)hymd_fused_destroy(m->fused);
        m->fused = nullptr;
        return HYMD_OK;
    }
    if (m->fused) return HYMD_OK;
    const TopoTerms t = topology_terms(m->topo);
    HYMD_REQUIRE(t.n == m->n, "topology and system sizes differ");
    if (m->dih) {  // bonds, angles and the dihedrals in one block layout
        const DihTerms d = dihedral_terms(m->dih);
        HYMD_TRY(hymd_fused_create_dihedrals(&m->fused, t.n, t.m2, t.a2, t.b2, t.r0, t.k2, t.m3, t.a3, t.b3, t.c3,
                                             t.t0, t.k3, d.m, d.a, d.b, d.c, d.d, d.coeff, d.dtype, d.last));
    } else {
        HYMD_TRY(hymd_fused_create(&m->fused, t.n, t.m2, t.a2, t.b2, t.r0, t.k2, t.m3, t.a3, t.b3, t.c3, t.t0, t.k3));
    }
    return hymd_fused_set_precision(m->fused, m->term_precision);
}

extern "C" int hymd_md_is_fused(hymd_md *m) { return (m && m->fused) ? 1 : 0; }

extern "C" int hymd_md_set_thermostat(hymd_md *m, int32_t n_types, const int32_t *type_group, int32_t n_groups,
                                      const int64_t *group_n, double gas_constant, double target_temperature,
                                      double c, int32_t remove_com, int32_t n_b, double *work, void *workspace,
                                      int64_t workspace_bytes) {
    HYMD_REQUIRE(m && n_b >= 0, "bad argument");
    m->draws.clear();
    m->draw_next = 0;
    if (n_b == 0) {
        m->n_b = 0;
        return HYMD_OK;
    }
    HYMD_REQUIRE(n_types == m->T, "n_types differs from the handle's");
    HYMD_REQUIRE(work && workspace && workspace_bytes >= thermostat_workspace_doubles() * (int64_t)sizeof(double),
                 "the thermostat needs its work accumulator and a workspace of hymd_thermostat_workspace_bytes()");
    HYMD_TRY(csvr_setup(&m->thermo, n_types, type_group, n_groups, group_n, m->mass, gas_constant,
                        target_temperature, c, remove_com));
    m->thermo_work = work;
    m->thermo_ws = (double *)workspace;
    m->n_b = n_b;
    return HYMD_OK;
}

extern "C" int hymd_md_push_draws(hymd_md *m, const double *gaussians, const double *chi_squares, int32_t count) {
    HYMD_REQUIRE(m && m->n_b > 0, "no thermostat is set");
    HYMD_REQUIRE(count >= 0 && (count == 0 || (gaussians && chi_squares)), "bad draws");
    const int G = m->thermo.gp.n_groups;
    for (int c = 0; c < count; ++c) {
        m->draws.insert(m->draws.end(), gaussians + (size_t)c * G, gaussians + (size_t)(c + 1) * G);
        m->draws.insert(m->draws.end(), chi_squares + (size_t)c * G, chi_squares + (size_t)(c + 1) * G);
    }
    return HYMD_OK;
}

extern "C" int hymd_md_set_com_removal(hymd_md *m, int32_t every, int64_t n_particles, void *workspace,
                                       int64_t workspace_bytes) {
    HYMD_REQUIRE(m && every >= 0, "bad argument");
    if (every > 0)
        HYMD_REQUIRE(n_particles > 0 && workspace &&
                         workspace_bytes >= thermostat_workspace_doubles() * (int64_t)sizeof(double),
                     "COM removal needs n_particles and a workspace of hymd_thermostat_workspace_bytes()");
    m->com_every = every;
    m->com_n = n_particles;
    m->com_ws = (double *)workspace;
    return HYMD_OK;
}

extern "C" int hymd_md_set_step(hymd_md *m, int64_t step) {
    HYMD_REQUIRE(m && step >= 0, "bad argument");
    m->step = step;
    return HYMD_OK;
}

extern "C" int hymd_md_set_term_precision(hymd_md *m, int32_t precision) {
    HYMD_REQUIRE(m, "null handle");
    HYMD_REQUIRE(precision == HYMD_F32 || precision == HYMD_F64, "precision must be HYMD_F32 or HYMD_F64");
    HYMD_HIP(hipStreamSynchronize(m->work));
    m->term_precision = precision;
    drop_graphs(m);  // the captured steps hold the other kernel
    if (m->fused) HYMD_TRY(hymd_fused_set_precision(m->fused, precision));
    return HYMD_OK;
}

// Every `every` outer steps (0: never) the trailing one-particle molecules
// are put in the mesh-row order of their current positions.  types (the
// array the handle was created with), order (state row -> input row) and
// where (input row -> state row) are the caller's int arrays, permuted in
// place with the state.  Needs the fused path.
extern "C" int hymd_md_set_reorder(hymd_md *m, int32_t every, int32_t *types, int64_t *order, int64_t *where) {
    HYMD_REQUIRE(m && every >= 0, "bad argument");
    if (every == 0) {
        m->reorder_every = 0;
        return HYMD_OK;
    }
    if (!m->fused) {
        set_error("reordering needs the fused inner loop (one-particle molecules without terms)");
        return HYMD_ERR_UNSUPPORTED;
    }
    if (m->charges) {
        set_error("reordering would leave the caller's charges in the old order");
        return HYMD_ERR_UNSUPPORTED;
    }
    HYMD_REQUIRE(types == m->types && order && where, "types must be the handle's array; order and where non-null");
    m->types_rw = types;
    m->order = order;
    m->where = where;
    m->single0 = fused_single_tail(m->fused);
    m->reorder_every = every;
    m->since_reorder = 0;
    return HYMD_OK;
}

// First row of the reordered one-particle tail (n when there is none).
extern "C" int64_t hymd_md_single_tail(hymd_md *m) { return (m && m->fused) ? fused_single_tail(m->fused) : 0; }

// Time one stage of the eager steps that follow with HIP events recorded on
// the launch stream around every launch of that stage (graph replays are not
// timed).  stage: 0 off, 1 bonded forces, 2 kick+drift, 3 update_field (all its
// kernels), 4 readout, 5 PME of the charges (and dipoles).  Up to max_launches
// launches are recorded.
extern "C" int hymd_md_time_stage(hymd_md *m, int32_t stage, int32_t max_launches) {
    HYMD_REQUIRE(m && stage >= 0 && stage <= 5 && max_launches >= 0, "bad argument");
    HYMD_HIP(hipStreamSynchronize(m->work));
    for (hipEvent_t e : m->ev) (void)hipEventDestroy(e);
    m->ev.clear();
    m->ev_used = 0;
    m->timed = stage;
    if (stage == 0) return HYMD_OK;
    m->ev.resize(2 * (size_t)max_launches, nullptr);
    for (auto &e : m->ev) HYMD_HIP(hipEventCreate(&e));
    return HYMD_OK;
}

// Total milliseconds and number of timed launches so far (synchronises).
extern "C" int hymd_md_stage_time(hymd_md *m, double *total_ms, int32_t *launches) {
    HYMD_REQUIRE(m && total_ms && launches, "bad argument");
    HYMD_HIP(hipStreamSynchronize(m->work));
    double tot = 0.0;
    const size_t pairs = m->ev_used / 2;
    for (size_t i = 0; i < pairs; ++i) {
        float ms = 0.f;
        HYMD_HIP(hipEventElapsedTime(&ms, m->ev[2 * i], m->ev[2 * i + 1]));
        tot += ms;
    }
    *total_ms = tot;
    *launches = (int32_t)pairs;
    return HYMD_OK;
}

extern "C" int hymd_md_set_peptide(hymd_md *m, const float *charges, double elec_factor, double self_energy,
                                   hymd_dihedrals *dih, int64_t n_dihedrals, const float *dipole_charges,
                                   float *f_slow) {
    HYMD_REQUIRE(m, "null handle");
    HYMD_REQUIRE(f_slow || (!charges && !dih), "f_slow is required with charges or dihedrals");
    HYMD_REQUIRE(!dih || n_dihedrals >= 0, "bad dihedral count");
    HYMD_HIP(hipStreamSynchronize(m->work));
    // With dihedrals the per-kernel inner loop is the default: its last
    // sub-step's bonded and dihedral kernels run beside the field update and
    // the charges' PME, which the fused loop cannot do (its FP64 dihedral
    // terms hold a block at two waves per SIMD); hymd_md_set_fused(m, 1)
    // builds the fused loop with the dihedral terms.
    if (dih) HYMD_TRY(hymd_md_set_fused(m, 0));
    // the periodic reorder permutes the handle's own per-particle arrays,
    // not the caller's charges or summed slow forces
    if (charges) m->reorder_every = 0;
    drop_graphs(m);
    m->charges = charges;
    m->elec = elec_factor;
    m->e_self = self_energy;
    m->dih = dih;
    m->m4 = dih ? n_dihedrals : 0;
    m->q_dip = dipole_charges;
    m->protein = dih && dipole_charges;
    m->f_slow = f_slow;
    const size_t n3 = 3 * (size_t)std::max<int64_t>(m->n, 1), nd = 4 * (size_t)std::max<int64_t>(m->m4, 1);
    if (dih) {
        HYMD_TRY(m->f_ba.reserve(sizeof(float) * n3));
        HYMD_TRY(m->f_dih.reserve(sizeof(double) * n3));
        HYMD_TRY(m->dipoles.reserve(sizeof(double) * 3 * nd));
        HYMD_TRY(m->tmat.reserve(sizeof(double) * 54 * (nd / 4)));
    }
    if (m->pm2) (void)hymd_pm_destroy(m->pm2);
    m->pm2 = nullptr;
    if (m->protein) {
        HYMD_TRY(m->f_dip.reserve(sizeof(float) * 3 * nd));
        HYMD_TRY(m->dip32.reserve(sizeof(float) * 3 * nd));
        HYMD_TRY(m->f_dip64.reserve(sizeof(double) * 3 * nd));
        HYMD_TRY(m->f_recon.reserve(sizeof(double) * n3));
    }
    if (charges || m->protein) {
        // the side stream's mesh handle; eager PME calls (the charges at the
        // current positions, the dipole charges at zero) size its buffers
        // before any step is captured
        HYMD_TRY(pm_clone(m->pm, 1, m->side, &m->pm2));
        HYMD_TRY(m->f_recon.reserve(sizeof(double) * n3));  // (float) scratch of the warm-up
        if (charges)
            HYMD_TRY(hymd_update_field_force_q(m->pm2, m->x32, charges, m->n, m->sigma, m->elec, 0.0,
                                               (float *)m->f_recon.ptr, nullptr, nullptr, nullptr, nullptr));
        if (m->protein) {
            HYMD_HIP(hipMemsetAsync(m->dip32.ptr, 0, sizeof(float) * 3 * nd, m->side));
            HYMD_TRY(hymd_update_field_force_q(m->pm2, (const float *)m->dip32.ptr, m->q_dip, (int64_t)nd, m->sigma,
                                               m->elec, 0.0, (float *)m->f_dip.ptr, nullptr, nullptr, nullptr,
                                               nullptr));
        }
        HYMD_HIP(hipStreamSynchronize(m->side));
    }
    return HYMD_OK;
}

extern "C" int hymd_md_peptide_state(hymd_md *m, double *dipoles, double *tmat, double *extra_energies) {
    HYMD_REQUIRE(m, "null handle");
    HYMD_HIP(hipStreamSynchronize(m->work));
    if (dipoles && m->dih)
        HYMD_HIP(hipMemcpy(dipoles, m->dipoles.ptr, sizeof(double) * 12 * m->m4, hipMemcpyDeviceToDevice));
    if (tmat && m->dih) HYMD_HIP(hipMemcpy(tmat, m->tmat.ptr, sizeof(double) * 54 * m->m4, hipMemcpyDeviceToDevice));
    if (extra_energies) {
        extra_energies[0] = m->extra_e[0];
        extra_energies[1] = m->extra_e[1];
    }
    return HYMD_OK;
}
