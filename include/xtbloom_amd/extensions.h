/*
 * MI355X-backend extensions to the xtbloom C ABI.
 *
 * No entry point here has a counterpart in the reference ABI.  The extended
 * compute call and the packed-matrix size exist for callers of this backend
 * that need more than xtbloom_batch_result_t carries: the differentiable
 * PyTorch binding (dE/dQ of external point charges) and the parity tests
 * (packed intermediate matrices and energy components).  The no-allocation
 * compute call exists for callers that cannot allocate.
 */
#ifndef XTBLOOM_AMD_EXTENSIONS_H
#define XTBLOOM_AMD_EXTENSIONS_H

#include "xtbloom/xtbloom.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel families of one call, the index space of kernel_stats below. */
typedef enum xtbloom_amd_kernel {
  XTBLOOM_AMD_KERNEL_GEOMETRY = 0,   /* CN, D4 CN, ES2 Gamma, point-charge potential, SAD */
  XTBLOOM_AMD_KERNEL_INTEGRALS = 1,  /* S, D, Q, H0 per shell pair */
  XTBLOOM_AMD_KERNEL_CHOLESKY = 2,   /* S = L L^T, X = L^-T */
  XTBLOOM_AMD_KERNEL_POTENTIALS = 3, /* ES2/ES3/AES2/D4 potentials */
  XTBLOOM_AMD_KERNEL_FOCK = 4,       /* H = H0 + multipole potential terms */
  XTBLOOM_AMD_KERNEL_REDUCE = 5,     /* H' = X^T H X (two GEMM launches) */
  XTBLOOM_AMD_KERNEL_TRIDIAG = 6,    /* Householder tridiagonalization */
  XTBLOOM_AMD_KERNEL_FERMI = 7,      /* occupations (computed inside TRIDIAG_EIGEN: no launches) */
  XTBLOOM_AMD_KERNEL_BACK = 8,       /* C = X Y */
  XTBLOOM_AMD_KERNEL_DENSITY = 9,    /* P = C f C^T */
  XTBLOOM_AMD_KERNEL_MULLIKEN = 10,  /* shell charges and atomic multipoles */
  XTBLOOM_AMD_KERNEL_ENERGY_MIX = 11,/* SCC energy, Broyden step, convergence */
  XTBLOOM_AMD_KERNEL_ENERGY_WEIGHTED = 12, /* W = C f eps C^T */
  XTBLOOM_AMD_KERNEL_FORCE_PAIRS = 13,     /* derivative integrals x adjoints */
  XTBLOOM_AMD_KERNEL_FORCE_ATOMS = 14,     /* AES2/D4/ES2/CN/repulsion/point-charge gradients */
  XTBLOOM_AMD_KERNEL_FORCE_ATM = 15,       /* D4 three-body energy and gradient */
  XTBLOOM_AMD_KERNEL_TRIDIAG_EIGEN = 16,   /* eigenvalues of T, occupations, occupied eigenvectors */
  XTBLOOM_AMD_KERNEL_BACKTRANSFORM = 17,   /* eigenvectors of H' = Q Z */
  XTBLOOM_AMD_KERNEL_COUNT = 18
} xtbloom_amd_kernel_t;

/* Device time of one kernel family, summed over the call's launches (HIP
 * events recorded on the context stream around each launch), and the
 * algorithmic work of those launches: FP64-equivalent flops for the dense
 * linear algebra, minimal HBM bytes for the streaming kernels (0 when the
 * family has no figure). */
typedef struct xtbloom_amd_kernel_stat {
  int32_t launches;
  int32_t reserved;
  double total_ms;
  double flops;
  double bytes;
} xtbloom_amd_kernel_stat_t;

/* Host buffers; every pointer may be NULL.  Matrices are packed batch-major
 * (system s occupies sum_{r<s} n_r^2 .. + n_s^2, row-major), sized by
 * xtbloom_amd_packed_matrix_elements(). */
typedef struct xtbloom_amd_extension {
  uint32_t struct_size;
  uint32_t api_version;
  double* point_charge_potential_gradient; /* [P]  dE/dQ_p at the converged state */
  double* overlap;                         /* [sum n^2] S */
  double* core_hamiltonian;                /* [sum n^2] H0 */
  double* density;                         /* [sum n^2] converged P */
  double* coordination_numbers;            /* [N] GFN2 CN */
  double* energy_components;               /* [8B] core, es2, es3, aes2, d4, pc, rep, atm */
  xtbloom_amd_kernel_stat_t* kernel_stats; /* [XTBLOOM_AMD_KERNEL_COUNT]; NULL = no per-launch events */
  /* device timing of the call, milliseconds (outputs) */
  float setup_ms;
  float scc_ms;
  float force_ms;
  float total_ms;
  int32_t scc_rounds;
  /* input: bit f set = per-launch events for kernel family f only; 0 = every
   * family (fewer events perturb concurrent lanes less) */
  uint32_t kernel_stats_mask;
  /* [sum n] orbital energies of the final SCC round, ascending per system,
   * and the total (alpha + beta) occupation of each level */
  double* orbital_energies;
  double* occupations;
} xtbloom_amd_extension_t;

/* xtbloom_compute plus the extension outputs above. */
XTBLOOM_API xtbloom_status_t xtbloom_amd_compute_extended(xtbloom_context_t* context,
                                                          const xtbloom_batch_t* batch,
                                                          const xtbloom_compute_options_t* options,
                                                          xtbloom_batch_result_t* result,
                                                          xtbloom_amd_extension_t* extension);

/* Size of the fixed per-GPU device workspace of xtbloom_amd_compute_no_alloc. */
#define XTBLOOM_AMD_NO_ALLOC_WORKSPACE_BYTES ((uint64_t)64 << 20)

/* xtbloom_compute that allocates no device memory and no pinned host memory.
 * No counterpart in the reference ABI: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The caller passes inputs and outputs as for xtbloom_compute;
 * the device workspace is a fixed array of XTBLOOM_AMD_NO_ALLOC_WORKSPACE_BYTES
 * compiled into the library, one per GPU, held until the process ends.  A
 * batch whose workspace exceeds it returns XTBLOOM_STATUS_INVALID_ARGUMENT
 * before any output is written (the message states the bytes it needs).
 * Every no-allocation call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other contexts, other threads) race.  The call runs on
 * the context's stream without concurrent lanes; host staging uses the
 * context's pinned block when it is large enough and pageable copies
 * otherwise.  Results, statuses, WARM checkpoints and errors are those of
 * xtbloom_compute. */
XTBLOOM_API xtbloom_status_t xtbloom_amd_compute_no_alloc(xtbloom_context_t* context,
                                                          const xtbloom_batch_t* batch,
                                                          const xtbloom_compute_options_t* options,
                                                          xtbloom_batch_result_t* result);

/* sum over systems of n_ao^2 for a host batch; -1 on invalid input. */
XTBLOOM_API int64_t xtbloom_amd_packed_matrix_elements(const xtbloom_batch_t* batch);

#ifdef __cplusplus
}
#endif

#endif
