/*
 * MI355X-backend extensions to the xtbloom C ABI.
 *
 * Neither entry point has a counterpart in the reference ABI.  They exist for
 * callers of this backend that need more than xtbloom_batch_result_t carries:
 * the differentiable PyTorch binding (dE/dQ of external point charges) and the
 * parity tests (packed intermediate matrices and energy components).
 */
#ifndef XTBLOOM_AMD_EXTENSIONS_H
#define XTBLOOM_AMD_EXTENSIONS_H

#include "xtbloom/xtbloom.h"

#ifdef __cplusplus
extern "C" {
#endif

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
  /* device timing of the call, milliseconds (outputs) */
  float setup_ms;
  float scc_ms;
  float force_ms;
  float total_ms;
  int32_t scc_rounds;
  uint32_t reserved;
} xtbloom_amd_extension_t;

/* xtbloom_compute plus the extension outputs above. */
XTBLOOM_API xtbloom_status_t xtbloom_amd_compute_extended(xtbloom_context_t* context,
                                                          const xtbloom_batch_t* batch,
                                                          const xtbloom_compute_options_t* options,
                                                          xtbloom_batch_result_t* result,
                                                          xtbloom_amd_extension_t* extension);

/* sum over systems of n_ao^2 for a host batch; -1 on invalid input. */
XTBLOOM_API int64_t xtbloom_amd_packed_matrix_elements(const xtbloom_batch_t* batch);

#ifdef __cplusplus
}
#endif

#endif
