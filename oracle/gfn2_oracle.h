/*
 * GFN2-xTB CPU oracle: a plain-C FP64 restatement of the reference CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  The product (xtbloom_amd) never links it.
 *
 * Dense linear algebra (dpotrf, dpocon, dsyevd) goes through the same LP64
 * LAPACKE provider family the reference loads at run time (scipy-openblas,
 * reference src/model/gfn2/eigensolver.cpp:308-330); call oracle_init() with
 * the absolute path of that library first.
 */
#ifndef XTB_GFN2_ORACLE_H
#define XTB_GFN2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_options {
  int32_t max_iterations;   /* reference default 250 (api.cpp:159) */
  int32_t history;          /* Broyden memory, default 8 */
  double charge_tolerance;  /* default 1e-6 */
  double energy_tolerance;  /* default 1e-8 */
  double kt;                /* k_B*T in Hartree, default 300 K */
  double damping;           /* default 0.4 */
  double efield[3];         /* uniform electric field, Hartree/(e bohr); default 0 (none) */
  /* external charge operator v = b + A q, E = q.b + q.A.q/2 on the atomic
   * charges (periodic QM/MM embedding): b [nat] and A [nat*nat, row-major,
   * symmetric], each NULL when absent (default) */
  const double* potential_shifts;
  const double* charge_response;
  int32_t spin_channels;    /* 1 restricted (default), 2 unrestricted */
  int32_t reserved;
} oracle_options;

/* Optional diagnostics for term-level pinning; every pointer may be NULL. */
typedef struct oracle_trace {
  int32_t capacity;        /* number of iteration slots in the arrays below */
  int32_t recorded;        /* output: iterations recorded */
  double* free_energy;     /* [capacity] F per iteration */
  double* residual_rms;    /* [capacity] mixer residual rms per iteration */
  double* overlap;         /* [nao*nao] S, row-major */
  double* h0;              /* [nao*nao] core Hamiltonian */
  double* dipole;          /* [3*nao*nao] component-major D */
  double* quadrupole;      /* [6*nao*nao] component-major traceless Q */
  double* density;         /* [nao*nao] converged P */
  double* shell_charges;   /* [nsh] converged raw q_sh */
  double* dipoles;         /* [3*nat] converged atomic dipoles */
  double* quadrupoles;     /* [6*nat] converged atomic quadrupoles */
  double* coordination;    /* [nat] GFN2 CN */
  double* components;      /* [8]: core, es2, es3, aes2, d4, pc, rep, atm */
} oracle_trace;

int oracle_init(const char* lapack_library_path);
void oracle_default_options(oracle_options* options);

/* Counts for one system; returns 0 or an xtbloom status code. */
int oracle_basis_counts(int32_t nat, const int32_t* Z, int32_t* nshell, int32_t* nao,
                        double* electrons_neutral);

/*
 * One restricted GFN2 single point.  grad receives dE/dR (Hartree/bohr),
 * pc_grad receives dE/dR_point, dedq dE/dQ_point.  Any output may be NULL.
 * Returns an xtbloom status: 0 success, 7 SCC not converged, 8 eigensolver
 * failed, 1 invalid input.
 */
int oracle_singlepoint(int32_t nat, const int32_t* Z, const double* xyz, double charge,
                       int32_t unpaired, int32_t npc, const double* pc_xyz, const double* pc_q,
                       const double* pc_gamma, const oracle_options* options, double* energy,
                       double* grad, double* atomic_charges, double* pc_grad, double* dedq,
                       int32_t* iterations, oracle_trace* trace);

/*
 * oracle_singlepoint with the SCC multipole state exposed: state_in (NULL =
 * FRESH SAD start) is the state a WARM start begins from, state_out receives
 * the final state.  The state holds nshell shell charges, then 3 nat atomic
 * dipole and 6 nat atomic quadrupole components ([xx, xy, yy, xz, yz, zz]).
 */
int oracle_singlepoint_state(int32_t nat, const int32_t* Z, const double* xyz, double charge,
                             int32_t unpaired, int32_t npc, const double* pc_xyz, const double* pc_q,
                             const double* pc_gamma, const oracle_options* options, double* energy,
                             double* grad, double* atomic_charges, double* pc_grad, double* dedq,
                             int32_t* iterations, oracle_trace* trace, const double* state_in,
                             double* state_out);

/*
 * Ragged batch with one system per worker thread (threads <= 0: all cores).
 * Layout follows xtbloom_batch_t: atom_offsets[B+1], point_offsets[B+1] or NULL.
 * Failed systems get NaN energy/gradient/charge slices.
 */
int oracle_batch(int64_t batch, const int64_t* atom_offsets, const int32_t* Z, const double* xyz,
                 const double* charges, const int32_t* unpaired, const int64_t* point_offsets,
                 const double* pc_xyz, const double* pc_q, const double* pc_gamma,
                 const oracle_options* options, int32_t threads, double* energies, double* grad,
                 double* atomic_charges, double* pc_grad, int32_t* iterations, int32_t* status,
                 const double* efield /* [batch][3] uniform fields, or NULL */,
                 const double* shifts /* [total atoms] b, or NULL */,
                 const double* response /* per system nat*nat A packed in system order, or NULL */,
                 const int32_t* spin_channels /* [batch] 1 or 2, or NULL (all 1) */);

/*
 * One SCC step like oracle_scc_replay with 1 or 2 spin channels: mixed and
 * raw in the reference's state packing ([qsh, dip, quad], each field holding
 * its channels in turn); hamiltonian and density [nch][nao*nao] (alpha,
 * beta), eigenvalues [nch][nao], occupations [2][nao] (alpha, beta).
 */
int oracle_scc_replay_channels(int32_t nat, const int32_t* Z, const double* xyz, double charge,
                               int32_t unpaired, int32_t spin_channels, int32_t npc, const double* pc_xyz,
                               const double* pc_q, const double* pc_gamma, double kt, const double* mixed,
                               double* raw, double* hamiltonian, double* density, double* eigenvalues,
                               double* occupations, double* free_energy);

/*
 * Replay exactly one SCC step from a given mixed state [qsh, dip(3N), quad(6N)]
 * (mirrors the reference's tests/scc_trace_replay.cpp contract): returns the
 * raw Mulliken state, H, P and the free energy of that step.
 */
int oracle_scc_replay(int32_t nat, const int32_t* Z, const double* xyz, double charge,
                      int32_t unpaired, int32_t npc, const double* pc_xyz, const double* pc_q,
                      const double* pc_gamma, double kt, const double* mixed, double* raw,
                      double* hamiltonian, double* density, double* free_energy);

/* Modified-Broyden mixer driven by a residual sequence; mixed_out is [steps*dim]. */
int oracle_broyden(int32_t dim, int32_t memory, double damping, const double* initial,
                   int32_t steps, const double* residuals, double* mixed_out);

/* Term-level entry points (single system) used by the pinning tests. */
int oracle_coordination(int32_t nat, const int32_t* Z, const double* xyz, double* cn);
int oracle_integrals(int32_t nat, const int32_t* Z, const double* xyz, double* S, double* D,
                     double* Q);
int oracle_h0(int32_t nat, const int32_t* Z, const double* xyz, const double* S, double* H0);
double oracle_repulsion(int32_t nat, const int32_t* Z, const double* xyz, double* grad);
int oracle_d4_two_body(int32_t nat, const int32_t* Z, const double* xyz, const double* q,
                       double* energy, double* potential, double* grad);
int oracle_d4_atm(int32_t nat, const int32_t* Z, const double* xyz, double* energy, double* grad);
int oracle_fermi(int32_t n, const double* eps, double electrons, double kt, double* occ,
                 double* mu, double* entropy);

#ifdef __cplusplus
}
#endif

#endif
