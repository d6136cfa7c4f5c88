// MI355X (gfx950) GFN2-xTB batched single-point executor: host orchestration.
//
// One call = one ragged batch.  The host packs the topology once (basis,
// offsets, per-atom/shell parameters), uploads it into a grow-only device arena
// owned by the context, and drives the device kernels:
//   geometry -> integrals/H0 -> Cholesky/inverse -> SCC rounds -> forces.
// Each SCC round runs over all still-active systems; converged systems drop
// out through a device-side active mask.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../data/gfn2_tables.h"
#include "executor.hpp"
#include "gfn2_device.hpp"
#include "gfn2_integrals.hpp"
#include "gfn2_linalg.hpp"
#include "gfn2_mix_force.hpp"
#include "gfn2_scc.hpp"

static_assert(xtb::kMaxSystemAOs == xtb::kMaxSystemAO, "AO limit of the executor and the kernels");

namespace xtb {

namespace {

constexpr double kEvToHartree = 1.0 / 27.21138505;
constexpr double kAngstromToBohr = 1.8897261246204404;

#define HIP_TRY(expr)                                                           \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      error = std::string("HIP error ") + hipGetErrorString(_e) + " at " #expr; \
      return 6;                                                                 \
    }                                                                           \
  } while (0)

// Stewart STO-nG selection (reference common/sto.cpp:20-68).
bool sto_row(int n, int l, int ng, const double** a, const double** c) {
  if (n == 6 && ng == 6) {
    if (l == 0) { *a = xtbp_sto_alpha6s; *c = xtbp_sto_coeff6s; return true; }
    if (l == 1) { *a = xtbp_sto_alpha6p; *c = xtbp_sto_coeff6p; return true; }
    return false;
  }
  if (n <= 0 || n > 5 || l >= n || l > 4) return false;
  const int type = l == 0 ? n - 1 : l == 1 ? n + 3 : l == 2 ? n + 6 : l == 3 ? n + 8 : n + 9;
  if (type >= 15) return false;
  if (ng == 3) { *a = xtbp_sto_alpha3 + 3 * type; *c = xtbp_sto_coeff3 + 3 * type; return true; }
  if (ng == 4) { *a = xtbp_sto_alpha4 + 4 * type; *c = xtbp_sto_coeff4 + 4 * type; return true; }
  if (ng == 6) { *a = xtbp_sto_alpha6 + 6 * type; *c = xtbp_sto_coeff6 + 6 * type; return true; }
  return false;
}

// Normalized primitives of every shell parameter set, built once per process:
// the exponents are zeta^2 alpha_g and the contraction coefficients carry the
// primitive normalization (2 alpha/pi)^{3/4} (4 alpha)^{l/2} / sqrt((2l-1)!!).
struct ShellPrims {
  int ng = 0;     // 0: no STO-nG row for this parameter set
  double alpha[6];
  double coef[6];
};

const ShellPrims* shell_prims() {
  static const std::vector<ShellPrims> table = [] {
    static const double dfac[3] = {1.0, 1.0, 3.0};
    std::vector<ShellPrims> t(XTBP_NSHELLPAR);
    for (int p = 0; p < XTBP_NSHELLPAR; ++p) {
      const int l = xtbp_shell_l[p], ng = xtbp_shell_ngauss[p];
      const double* ta;
      const double* tc;
      if (l > 2 || ng > 6 || !sto_row(xtbp_shell_n[p], l, ng, &ta, &tc)) continue;
      const double z2 = xtbp_shell_slater[p] * xtbp_shell_slater[p];
      t[p].ng = ng;
      for (int g = 0; g < ng; ++g) {
        const double al = ta[g] * z2;
        const double norm = std::pow(2.0 / kPi * al, 0.75) *
                            std::pow(std::sqrt(4.0 * al), (double)l) / std::sqrt(dfac[l]);
        t[p].alpha[g] = al;
        t[p].coef[g] = tc[g] * norm;
      }
    }
    return t;
  }();
  return table.data();
}

template <typename T>
size_t align_bytes(size_t n) {
  return ((n * sizeof(T)) + 255) & ~size_t(255);
}

// Host-side packed topology; every vector is uploaded verbatim.
struct HostTopo {
  int B = 0, N = 0, NS = 0, NAO = 0, NP = 0, NPC = 0;
  std::vector<int> atom_off, shell_off, ao_off, vec_off, pc_off, order;
  std::vector<int64_t> mat_off, gam_off, apair_off, sp_off;
  std::vector<double> nalpha, nbeta, charges;
  std::vector<int> Z, atom_sys, atom_sh0, at_d4ref, at_d4nref, perm;
  std::vector<double> xyz, at_cnrad, at_h0rad, at_en, at_dk, at_qk, at_mprad, at_mpvcn,
      at_sqrt_arep, at_zeff, at_d4rcov, at_d4en, at_d4zeff, at_d4hard, at_d4r4r2;
  std::vector<int> sh_atom, sh_l, sh_ao, sh_prim, sh_nprim;
  std::vector<double> sh_level, sh_kcn, sh_kpoly, sh_zeta, sh_refocc, sh_eta, sh_gam3;
  std::vector<int> ao_sh, ao_atom;
  std::vector<double> alpha, coef;
  std::vector<double> pc_xyz, pc_q, pc_gam;
  std::vector<double> field;  // [3B] when the call carries a uniform field, else empty
  std::vector<double> pshift, presp;  // external charge operator b [N], A [apair_off[B]], or empty
  // unrestricted systems (empty when there are none): the partner of each
  // system (-1: restricted) and the spin couplings of each shell
  std::vector<int> partner;
  std::vector<double> sh_wspin;
  std::vector<int> setup_status;
  int max_ao = 0;
  int max_atoms = 0;
};

// partner (optional, [B]): the alpha system of an unrestricted member and
// its beta copy name each other; the alpha system holds alpha electrons only
// and a state of both channels, the copy beta electrons only.
int build_topology(const HostBatch& hb, HostTopo& T, std::string& error, const int* partner) {
  const int B = (int)hb.batch;
  T.B = B;
  T.atom_off.assign(B + 1, 0);
  T.shell_off.assign(B + 1, 0);
  T.ao_off.assign(B + 1, 0);
  T.vec_off.assign(B + 1, 0);
  T.pc_off.assign(B + 1, 0);
  T.mat_off.assign(B + 1, 0);
  T.gam_off.assign(B + 1, 0);
  T.apair_off.assign(B + 1, 0);
  T.sp_off.assign(B + 1, 0);
  T.nalpha.assign(B, 0.0);
  T.nbeta.assign(B, 0.0);
  T.charges.assign(hb.charges, hb.charges + B);
  T.setup_status.assign(B, 0);
  const ShellPrims* prims = shell_prims();
  {
    // upper bounds: 3 shells (s, p, d), 9 AOs and 18 primitives per atom
    const size_t N = (size_t)hb.atom_offsets[B];
    for (auto* v : {&T.Z, &T.atom_sys, &T.at_d4ref, &T.at_d4nref}) v->reserve(N);
    T.atom_sh0.reserve(N + 1);
    T.xyz.reserve(3 * N);
    for (auto* v : {&T.at_cnrad, &T.at_h0rad, &T.at_en, &T.at_dk, &T.at_qk, &T.at_mprad,
                    &T.at_mpvcn, &T.at_sqrt_arep, &T.at_zeff, &T.at_d4rcov, &T.at_d4en,
                    &T.at_d4zeff, &T.at_d4hard, &T.at_d4r4r2})
      v->reserve(N);
    for (auto* v : {&T.sh_atom, &T.sh_l, &T.sh_ao, &T.sh_prim, &T.sh_nprim}) v->reserve(3 * N);
    for (auto* v : {&T.sh_level, &T.sh_kcn, &T.sh_kpoly, &T.sh_zeta, &T.sh_refocc, &T.sh_eta,
                    &T.sh_gam3})
      v->reserve(3 * N);
    T.ao_sh.reserve(9 * N);
    T.ao_atom.reserve(9 * N);
    T.alpha.reserve(18 * N);
    T.coef.reserve(18 * N);
  }
  for (int s = 0; s < B; ++s) {
    const int64_t a0 = hb.atom_offsets[s], a1 = hb.atom_offsets[s + 1];
    T.atom_off[s] = (int)a0;
    int nsh = 0, nao = 0;
    double refel = 0.0;
    for (int64_t a = a0; a < a1; ++a) {
      const int z = hb.Z[a];
      if (z < 1 || z > XTBP_NELEM) {
        error = "atomic number outside the GFN2 parameter domain (1..86)";
        return 1;
      }
      const int off = xtbp_elem_shell_offset[z - 1], cnt = xtbp_elem_shell_count[z - 1];
      const int a_idx = (int)a;
      if ((int)T.Z.size() != a_idx) { error = "internal atom ordering error"; return 6; }
      T.Z.push_back(z);
      T.atom_sys.push_back(s);
      T.atom_sh0.push_back((int)T.sh_atom.size());
      for (int k = 0; k < 3; ++k) T.xyz.push_back(hb.xyz[3 * a + k]);
      T.at_cnrad.push_back(xtbp_covalent_radius_aa[z - 1] * (4.0 / 3.0) * kAngstromToBohr);
      T.at_h0rad.push_back(xtbp_atomic_radius_aa[z - 1] * kAngstromToBohr);
      T.at_en.push_back(xtbp_elem_en[z - 1]);
      T.at_dk.push_back(xtbp_elem_dkernel[z - 1]);
      T.at_qk.push_back(xtbp_elem_qkernel[z - 1]);
      T.at_mprad.push_back(xtbp_elem_mprad[z - 1]);
      T.at_mpvcn.push_back(xtbp_elem_mpvcn[z - 1]);
      T.at_sqrt_arep.push_back(std::sqrt(xtbp_elem_arep[z - 1]));
      T.at_zeff.push_back(xtbp_elem_zeff[z - 1]);
      T.at_d4rcov.push_back(xtbp_d4_covalent_radius[z - 1]);
      T.at_d4en.push_back(xtbp_d4_en[z - 1]);
      T.at_d4zeff.push_back(xtbp_d4_zeff[z - 1]);
      T.at_d4hard.push_back(xtbp_d4_hardness[z - 1]);
      T.at_d4r4r2.push_back(xtbp_d4_r4r2[z - 1]);
      T.at_d4ref.push_back(xtbp_d4_ref_offset[z - 1]);
      T.at_d4nref.push_back(xtbp_d4_ref_count[z - 1]);
      int seen[5] = {0, 0, 0, 0, 0};
      for (int k = 0; k < cnt; ++k) {
        const int p = off + k, l = xtbp_shell_l[p], ng = xtbp_shell_ngauss[p];
        if (l > 2 || seen[l]) {
          error = "element with f shells or repeated angular momenta is not supported";
          return 3;
        }
        seen[l] = 1;
        const ShellPrims& sp = prims[p];
        if (sp.ng == 0) {
          error = "missing STO-nG table row";
          return 6;
        }
        const int gsh = (int)T.sh_atom.size();
        T.sh_atom.push_back(a_idx);
        T.sh_l.push_back(l);
        T.sh_ao.push_back((int)T.ao_sh.size());
        T.sh_prim.push_back((int)T.alpha.size());
        T.sh_nprim.push_back(ng);
        T.sh_level.push_back(xtbp_shell_level_ev[p] * kEvToHartree);
        T.sh_kcn.push_back(xtbp_shell_kcn_ev[p] * kEvToHartree);
        T.sh_kpoly.push_back(xtbp_shell_kpoly[p]);
        T.sh_zeta.push_back(xtbp_shell_slater[p]);
        T.sh_refocc.push_back(xtbp_shell_refocc[p]);
        T.sh_eta.push_back(xtbp_elem_gam[z - 1] * xtbp_shell_hubbard_scale[p]);
        const double s3[3] = {XTBP_ES3_SCALE_S, XTBP_ES3_SCALE_P, XTBP_ES3_SCALE_D};
        T.sh_gam3.push_back(xtbp_elem_gam3[z - 1] * s3[l]);
        T.alpha.insert(T.alpha.end(), sp.alpha, sp.alpha + ng);
        T.coef.insert(T.coef.end(), sp.coef, sp.coef + ng);
        for (int m = 0; m < 2 * l + 1; ++m) {
          T.ao_sh.push_back(gsh);
          T.ao_atom.push_back(a_idx);
        }
        nsh += 1;
        nao += 2 * l + 1;
        refel += xtbp_shell_refocc[p];
      }
    }
    const int N = (int)(a1 - a0);
    if (N <= 0) {
      error = "every batch member needs at least one atom";
      return 1;
    }
    if (nao > kMaxSystemAO) {
      error = "system exceeds the 8192-AO limit of the MI355X executor";
      return 3;
    }
    T.max_ao = std::max(T.max_ao, nao);
    T.shell_off[s + 1] = T.shell_off[s] + nsh;
    T.ao_off[s + 1] = T.ao_off[s] + nao;
    T.vec_off[s + 1] = T.vec_off[s] + (partner && partner[s] > s ? 2 : 1) * (nsh + 9 * N);
    T.mat_off[s + 1] = T.mat_off[s] + (int64_t)nao * nao;
    T.gam_off[s + 1] = T.gam_off[s] + (int64_t)nsh * nsh;
    T.apair_off[s + 1] = T.apair_off[s] + (int64_t)N * N;
    T.sp_off[s + 1] = T.sp_off[s] + (int64_t)nsh * nsh;
    // electron counts (reference wavefunction.cpp:614-636)
    const double nel = refel - hb.charges[s];
    const long nint = std::lround(nel);
    if (!(nel >= 0.0) || ((nint & 1) != (hb.unpaired[s] & 1)) || hb.unpaired[s] < 0) {
      error = "total electron count and unpaired-electron count have incompatible parity";
      return 1;
    }
    T.nalpha[s] = 0.5 * (nel + hb.unpaired[s]);
    T.nbeta[s] = 0.5 * (nel - hb.unpaired[s]);
    if (T.nbeta[s] < 0.0 || T.nalpha[s] > nao || T.nbeta[s] > nao) {
      error = "alpha or beta electron count cannot be represented by the system orbital space";
      return 1;
    }
    if (partner && partner[s] >= 0) {
      if (partner[s] > s)
        T.nbeta[s] = 0.0;
      else
        T.nalpha[s] = 0.0;
    }
    T.pc_off[s + 1] = hb.point_offsets ? (int)hb.point_offsets[s + 1] : 0;
  }
  T.atom_off[B] = (int)hb.atom_offsets[B];
  T.atom_sh0.push_back((int)T.sh_atom.size());
  if (partner) {
    // spin constants of each shell with the shells of its atom
    // (reference spin.cpp:200-263: W_ll' by element, symmetric in l, l')
    T.partner.assign(partner, partner + B);
    T.sh_wspin.assign(3 * T.sh_atom.size(), 0.0);
    for (size_t sh = 0; sh < T.sh_atom.size(); ++sh) {
      const int a = T.sh_atom[sh], z = T.Z[a];
      for (int k = 0; T.atom_sh0[a] + k < T.atom_sh0[a + 1]; ++k) {
        int l1 = T.sh_l[sh], l2 = T.sh_l[T.atom_sh0[a] + k];
        if (l1 > l2) std::swap(l1, l2);
        const int idx = l1 == 0 ? (l2 == 0 ? 0 : (l2 == 1 ? 1 : 3)) : (l1 == 1 ? (l2 == 1 ? 2 : 4) : 5);
        T.sh_wspin[3 * sh + k] = xtbp_spin_w[6 * (z - 1) + idx];
      }
    }
  }
  T.N = (int)T.Z.size();
  T.NS = (int)T.sh_atom.size();
  T.NAO = (int)T.ao_sh.size();
  T.NP = (int)T.alpha.size();
  T.NPC = T.pc_off[B];
  if (T.NPC > 0) {
    T.pc_xyz.assign(hb.pc_xyz, hb.pc_xyz + 3 * (size_t)T.NPC);
    T.pc_q.assign(hb.pc_q, hb.pc_q + T.NPC);
    T.pc_gam.assign(hb.pc_gamma, hb.pc_gamma + T.NPC);
  }
  // atoms of each system grouped by D4 element, stable inside a group
  T.perm.resize(T.N);
  for (int s = 0; s < B; ++s) {
    const int a0 = T.atom_off[s], a1 = T.atom_off[s + 1];
    T.max_atoms = std::max(T.max_atoms, a1 - a0);
    std::iota(T.perm.begin() + a0, T.perm.begin() + a1, a0);
    std::stable_sort(T.perm.begin() + a0, T.perm.begin() + a1,
                     [&](int a, int b) { return T.at_d4ref[a] < T.at_d4ref[b]; });
  }
  // size-sorted visiting order (largest AO count first)
  T.order.resize(B);
  std::iota(T.order.begin(), T.order.end(), 0);
  std::stable_sort(T.order.begin(), T.order.end(), [&](int a, int b) {
    return (T.ao_off[a + 1] - T.ao_off[a]) > (T.ao_off[b + 1] - T.ao_off[b]);
  });
  return 0;
}

}  // namespace

struct TopoCache {
  HostTopo topo;
  bool valid = false;
  int64_t tag = -1;         // which slice of which lane split the topology belongs to
  uint64_t generation = 0;  // process-unique id of the current contents
};
std::atomic<uint64_t> g_topo_generation{0};

TopoCache* topo_cache_create() { return new TopoCache(); }
void topo_cache_destroy(TopoCache* cache) { delete cache; }
bool topo_cache_retag(TopoCache* cache, int64_t tag) {
  if (cache->tag != tag) {
    cache->valid = false;
    cache->tag = tag;
  }
  return cache->valid;
}

// Device workspace of RunOptions::no_alloc calls: one copy of this array per
// GPU, held from module load until the process ends.  Host code never passes
// `scratch` itself to a kernel or a copy (in host code the name is only a
// stand-in for the symbol): each device's address comes from
// hipGetSymbolAddress, resolved once per device with that device current,
// and every no_alloc call on a device uses that address.  Calls on one GPU
// share the array, so no_alloc calls that overlap on one GPU (two contexts,
// two threads) race; each call rewrites every byte it reads before reading it.
static __device__ __align__(256) char scratch[kNoAllocWorkspaceBytes];

std::mutex g_scratch_lock;
std::vector<char*> g_scratch_addr;  // device address of `scratch` per device id

// With device `id` current: the device address of `scratch` on that device.
hipError_t scratch_address(int id, char** out) {
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  if ((int)g_scratch_addr.size() <= id) g_scratch_addr.resize(id + 1, nullptr);
  if (g_scratch_addr[id] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[id] = static_cast<char*>(p);
  }
  *out = g_scratch_addr[id];
  return hipSuccess;
}

struct Device {
  int id = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // persistent parameter tables
  double* d4_c6 = nullptr;
  double* d4_refcn = nullptr;
  double* d4_refq = nullptr;
  int* d4_refng = nullptr;
  double* kpair = nullptr;
  // grow-only arena
  char* arena = nullptr;
  size_t arena_bytes = 0;
  // this device's address of the no_alloc workspace array
  char* scratch = nullptr;
  // captured SCC rounds of the batched path and what they were captured for
  hipGraphExec_t graph_exec = nullptr;
  std::vector<char> graph_key;

  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t round_ev[2] = {nullptr, nullptr};  // end of the SCC rounds of each parity
  std::vector<hipEvent_t> probe_pool;  // per-launch timing events, grow-only
  size_t lds_limit = 65536;            // dynamic LDS granted to the back-transformation
  // grow-only pinned staging: [0, 256) the per-round active count, then the
  // call's host-to-device uploads (pinned, so each copy is a plain async DMA)
  char* pinned = nullptr;
  size_t pinned_bytes = 0;
  // the arena holds the uploaded topology of cache generation resident_gen
  // (0: none), laid out for a call of resident_bytes arena bytes
  uint64_t resident_gen = 0;
  size_t resident_bytes = 0;
};

constexpr size_t kPinnedInitialBytes = 4096;
// SCC rounds per captured graph on the batched path: after convergence at
// most this many rounds run with no active system (each a few microseconds)
constexpr int kGraphRounds = 4;

Device* device_create(int device_id, void* stream, std::string& error) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    error = "no ROCm device is visible";
    return nullptr;
  }
  if (device_id < 0) device_id = 0;
  if (device_id >= count) {
    error = "device_id is out of range";
    return nullptr;
  }
  if (hipSetDevice(device_id) != hipSuccess) {
    error = "hipSetDevice failed";
    return nullptr;
  }
  Device* d = new Device();
  d->id = device_id;
  if (stream) {
    d->stream = static_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
      error = "hipStreamCreate failed";
      delete d;
      return nullptr;
    }
    d->own_stream = true;
  }

  bool ok = hipMalloc(&d->d4_c6, sizeof(double) * 262 * 262) == hipSuccess &&
            hipMalloc(&d->d4_refcn, sizeof(double) * 262) == hipSuccess &&
            hipMalloc(&d->d4_refq, sizeof(double) * 262) == hipSuccess &&
            hipMalloc(&d->d4_refng, sizeof(int) * 262) == hipSuccess &&
            hipMalloc(&d->kpair, sizeof(double) * 9) == hipSuccess;
  ok = ok && hipMemcpy(d->d4_c6, xtbp_d4_c6, sizeof(double) * 262 * 262, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(d->d4_refcn, xtbp_d4_ref_cn, sizeof(double) * 262, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(d->d4_refq, xtbp_d4_ref_q, sizeof(double) * 262, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(d->d4_refng, xtbp_d4_ref_ngauss, sizeof(int) * 262, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(d->kpair, xtbp_kpair, sizeof(double) * 9, hipMemcpyHostToDevice) == hipSuccess;
  // the per-round active counts live in pinned memory from the start, so that
  // a no_alloc call never needs to allocate it
  ok = ok && hipHostMalloc((void**)&d->pinned, kPinnedInitialBytes) == hipSuccess;
  if (ok) d->pinned_bytes = kPinnedInitialBytes;
  if (ok && scratch_address(device_id, &d->scratch) != hipSuccess) {
    error = "hipGetSymbolAddress failed for the no-allocation device workspace";
    device_destroy(d);
    return nullptr;
  }
  for (int i = 0; i < 4 && ok; ++i) ok = hipEventCreate(&d->ev[i]) == hipSuccess;
  for (int i = 0; i < 2 && ok; ++i)
    ok = hipEventCreateWithFlags(&d->round_ev[i], hipEventDisableTiming) == hipSuccess;
  // gfx950 grants up to 160 KB of LDS to one workgroup; opt the back-transformations in
  int lds = 0;
  if (ok && hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device_id) == hipSuccess &&
      lds > 65536 &&
      hipFuncSetAttribute((const void*)k_backtransform, hipFuncAttributeMaxDynamicSharedMemorySize,
                          lds - (int)kBackStaticLds) == hipSuccess &&
      hipFuncSetAttribute((const void*)k_backtransform_wy<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          lds) == hipSuccess)
    d->lds_limit = (size_t)lds;
  if (!ok) {
    error = "failed to allocate or upload the device parameter tables";
    device_destroy(d);
    return nullptr;
  }
  return d;
}

void device_destroy(Device* d) {
  if (!d) return;
  (void)hipSetDevice(d->id);
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  (void)hipFree(d->d4_c6);
  (void)hipFree(d->d4_refcn);
  (void)hipFree(d->d4_refq);
  (void)hipFree(d->d4_refng);
  (void)hipFree(d->kpair);
  (void)hipFree(d->arena);
  if (d->pinned) (void)hipHostFree(d->pinned);
  for (auto& e : d->round_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : d->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : d->probe_pool) (void)hipEventDestroy(e);
  if (d->graph_exec) (void)hipGraphExecDestroy(d->graph_exec);
  if (d->own_stream && d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
}

int device_id(const Device* d) { return d ? d->id : -1; }


int64_t batch_matrix_elements(const HostBatch& hb, int32_t* max_ao, std::string& error,
                              int32_t* system_aos, int32_t* system_state) {
  int64_t total = 0;
  int32_t mx = 0;
  for (int64_t s = 0; s < hb.batch; ++s) {
    int64_t nao = 0, nsh = 0;
    for (int64_t a = hb.atom_offsets[s]; a < hb.atom_offsets[s + 1]; ++a) {
      const int z = hb.Z[a];
      if (z < 1 || z > XTBP_NELEM) {
        error = "atomic number outside the GFN2 parameter domain (1..86)";
        return -1;
      }
      const int off = xtbp_elem_shell_offset[z - 1], cnt = xtbp_elem_shell_count[z - 1];
      for (int k = 0; k < cnt; ++k) nao += 2 * xtbp_shell_l[off + k] + 1;
      nsh += cnt;
    }
    if (system_state)
      system_state[s] = (int32_t)((hb.spin && hb.spin[s] == 2 ? 2 : 1) *
                                  (nsh + 9 * (hb.atom_offsets[s + 1] - hb.atom_offsets[s])));
    mx = std::max<int32_t>(mx, (int32_t)nao);
    if (system_aos) system_aos[s] = (int32_t)nao;
    total += nao * nao;
  }
  if (max_ao) *max_ao = mx;
  return total;
}

namespace {

// Carves typed sub-buffers out of the device arena.
struct Carver {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t n) {
    T* p = reinterpret_cast<T*>(base + off);
    off += align_bytes<T>(std::max<size_t>(n, 1));
    return p;
  }
};

template <typename T>
size_t plan_bytes(size_t n) { return align_bytes<T>(std::max<size_t>(n, 1)); }

__global__ void k_init_active(int B, const SysState* st, unsigned char* active, int* iters_fix,
                              SysState* stw) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const bool ok = st[s].status == 0;
  active[s] = ok ? 1 : 0;
  if (!ok) stw[s].iterations = 1;  // a failed overlap factorization is reported after one round
  (void)iters_fix;
}

__global__ void k_valid(int B, const SysState* st, unsigned char* valid) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < B) valid[s] = (st[s].status == 0 && st[s].converged) ? 1 : 0;
}

}  // namespace

// The whole pipeline over a batch whose unrestricted members are already
// split into alpha systems and beta copies (partner, or null).
static int run_batch_core(Device* dev, const HostBatch& hb, const RunOptions& opt, HostOutputs& out,
                          RunTiming* timing, std::string& error, TopoCache* cache, const int* partner) {
  if (hipSetDevice(dev->id) != hipSuccess) {
    error = "hipSetDevice failed";
    return 6;
  }
  HostTopo local;
  HostTopo& T = cache ? cache->topo : local;
  const bool cache_hit = cache && cache->valid;
  if (cache_hit) {
    // same topology: only the geometry changes
    T.xyz.assign(hb.xyz, hb.xyz + 3 * (size_t)hb.atoms);
    T.pc_xyz.assign(hb.pc_xyz, hb.pc_xyz + 3 * (size_t)T.NPC);
    T.pc_q.assign(hb.pc_q, hb.pc_q + T.NPC);
  } else {
    if (cache) {
      cache->valid = false;
      T = HostTopo();  // build_topology appends to empty vectors
    }
    const int st = build_topology(hb, T, error, partner);
    if (st != 0) return st;
    if (cache) {
      cache->valid = true;
      cache->generation = ++g_topo_generation;
    }
  }
  // the field and the charge operator are per-call inputs like the geometry
  if (hb.field)
    T.field.assign(hb.field, hb.field + 3 * (size_t)hb.batch);
  else
    T.field.clear();
  if (hb.pshift)
    T.pshift.assign(hb.pshift, hb.pshift + (size_t)hb.atoms);
  else
    T.pshift.clear();
  if (hb.presp)
    T.presp.assign(hb.presp, hb.presp + (size_t)T.apair_off[T.B]);
  else
    T.presp.clear();
  // alpha systems of the unrestricted members (one block each in the spin kernels)
  // (those above kMaxAO AOs first)
  std::vector<int> alpha_sys;
  for (size_t s = 0; s < T.partner.size(); ++s)
    if (T.partner[s] > (int)s) alpha_sys.push_back((int)s);
  const auto xl_end = std::stable_partition(alpha_sys.begin(), alpha_sys.end(),
                                            [&](int s) { return T.ao_off[s + 1] - T.ao_off[s] > kMaxAO; });
  const int U = (int)alpha_sys.size();
  const int U_xl = (int)(xl_end - alpha_sys.begin());
  const int B = T.B, N = T.N, NS = T.NS, NAO = T.NAO, NPC = T.NPC;
  const int64_t NN = T.mat_off[B];
  const int64_t NG = T.gam_off[B];
  const int64_t NAP = T.apair_off[B];
  const int64_t NSP = T.sp_off[B];
  const int NV = T.vec_off[B];
  const int mem = std::max(1, opt.history);

  // GEMM tile queues and eigenvector back-transformation column blocks
  std::vector<TileTask> full, lower;
  std::vector<BackTask> back;
  std::vector<FockTile> chunks;
  std::vector<int64_t> part_off(B + 1, 0);
  size_t back_lds = 0, back_lds_xl = 0;  // dynamic LDS of k_backtransform_wy, <false> and <true>
  int n_large = 0;  // systems above 256 AOs (a prefix of the size order)
  int n_xl = 0;     // systems above kMaxAO AOs (a prefix of that prefix)
  size_t n_back_large = 0, n_back_xl = 0;  // their back-transformation tasks (prefixes of `back`)
  for (int i = 0; i < B; ++i) {
    const int s = T.order[i];
    const int n = T.ao_off[s + 1] - T.ao_off[s];
    n_large += n > 256;
    n_xl += n > kMaxAO;
    const int nt = (n + kTile - 1) / kTile;
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) {
        full.push_back({s, a, b});
        if (a >= b) lower.push_back({s, a, b});
      }
    int cw = 64;  // columns per task: register path (n <= 256), WY path in place (n > kMaxAO)
    if (n > 256 && n <= kMaxAO) {
      while (cw > 4 && backtransform_wy_lds(n, cw, false) > dev->lds_limit) cw /= 2;
      if (backtransform_wy_lds(n, cw, false) > dev->lds_limit) {
        error = "the device grants too little LDS for the back-transformation";
        return 6;
      }
      back_lds = std::max(back_lds, backtransform_wy_lds(n, cw, false));
    } else if (n > kMaxAO) {
      back_lds_xl = backtransform_wy_lds(n, cw, true);
    }
    for (int j0 = 0; j0 < n; j0 += cw) back.push_back({s, j0, cw, 0});
    if (n > kMaxAO) n_back_xl = back.size();
    if (n > 256) n_back_large = back.size();
    for (int r0 = 0; r0 < n; r0 += kFockTile)
      for (int c0 = r0; c0 < n; c0 += kFockTile) chunks.push_back({s, r0, c0});
  }
  // systems split by whether any shell is a d shell: [0] s/p only, [1] with d;
  // per class the prefix of ordered shell pairs and of ordered atom pairs
  std::vector<int> cls_sys[2];
  std::vector<int64_t> cls_sp[2], cls_ap[2];
  for (int c = 0; c < 2; ++c) { cls_sp[c].push_back(0); cls_ap[c].push_back(0); }
  for (int i = 0; i < B; ++i) {
    const int s = T.order[i];
    bool has_d = false;
    for (int sh = T.shell_off[s]; sh < T.shell_off[s + 1]; ++sh) has_d = has_d || T.sh_l[sh] == 2;
    const int c = has_d ? 1 : 0;
    const int64_t ns = T.shell_off[s + 1] - T.shell_off[s];
    const int64_t na = T.atom_off[s + 1] - T.atom_off[s];
    cls_sys[c].push_back(s);
    cls_sp[c].push_back(cls_sp[c].back() + (int64_t)ns * (ns + 1) / 2);
    cls_ap[c].push_back(cls_ap[c].back() + na * (na + 1) / 2);
  }
  // Mulliken partials (one per 32-row chunk and column), system by system;
  // k_mulliken_rows runs over the Fock tile list
  for (int s = 0; s < B; ++s) {
    const int n = T.ao_off[s + 1] - T.ao_off[s];
    const int nch = (n + kMullRows - 1) / kMullRows;
    part_off[s + 1] = part_off[s] + (int64_t)nch * n * 11;
  }

  // ---- arena layout ----
  size_t bytes = 0;
  auto addi = [&](size_t n) { bytes += plan_bytes<int>(n); };
  auto addl = [&](size_t n) { bytes += plan_bytes<int64_t>(n); };
  auto addd = [&](size_t n) { bytes += plan_bytes<double>(n); };
  auto addb = [&](size_t n) { bytes += plan_bytes<unsigned char>(n); };
  // topology
  addi(B + 1); addi(B + 1); addi(B + 1); addi(B + 1); addi(B + 1); addi(B);
  addl(B + 1); addl(B + 1); addl(B + 1); addl(B + 1);
  addd(B); addd(B); addd(B);
  addi(N); addi(N); addi(N + 1); addi(N); addi(N);
  addd(3 * (size_t)N);
  for (int k = 0; k < 15; ++k) addd(N);
  for (int k = 0; k < 5; ++k) addi(NS);
  for (int k = 0; k < 7; ++k) addd(NS);
  addi(NAO); addi(NAO);
  addd(T.NP); addd(T.NP);
  addd(3 * (size_t)NPC); addd(NPC); addd(NPC);
  addd(3 * (size_t)B);  // uniform field per system
  addd(T.pshift.size()); addd(T.presp.size());  // charge operator b, A
  addi(T.partner.size()); addi(alpha_sys.size()); addd(T.sh_wspin.size());  // unrestricted systems
  bytes += plan_bytes<TileTask>(full.size()) + plan_bytes<TileTask>(lower.size());
  bytes += plan_bytes<BackTask>(back.size());
  bytes += plan_bytes<FockTile>(chunks.size());
  for (int c = 0; c < 2; ++c) {
    addi(cls_sys[c].size());
    addl(cls_sp[c].size());
    addl(cls_ap[c].size());
  }
  addl(B + 1);
  addd((size_t)part_off[B]);  // Mulliken chunk partials
  // work arrays
  const int kMat = 20;  // S, D3, Q6, H0, X, H, A, Z, Y, C, P, W, T
  addd((size_t)kMat * NN);
  addd(NG);                    // gamma
  for (int k = 0; k < 7; ++k) addd(N);  // cn cnd4 mrad dmrad qat vat dedcn
  addd(N);                     // dedcn4
  addd((size_t)kD4MaxRef * N); addd((size_t)kD4MaxRef * N);
  addd(NS); addd(NS);          // vpc, vsh
  addd(3 * (size_t)N); addd(6 * (size_t)N);
  for (int k = 0; k < 5; ++k) addd(NAO);  // eps fa fb occ fe
  addd(NAO); addd(NAO);                   // tridiagonal d and e
  addd((size_t)kBackChunk * NAO + (size_t)kBackChunk * kBackChunk * B);  // WY factors (wy_offset)
  addd(B);                                // per-system eigenvalue change of the last round
  addd(NAO);                              // eigenvector normalization factors
  addd((size_t)NAO * 10);      // mulliken column partials
  for (int k = 0; k < 5; ++k) addd(NV); // x raw cur prev prev_res
  addd((size_t)mem * NV); addd((size_t)mem * NV);  // df, u
  addd((size_t)mem * B);       // omega
  bytes += plan_bytes<SysState>(B);
  addb(B); addb(B);            // active, valid
  addi(2);                     // counter
  addi(1);                     // device round counter (batched path)
  addi((size_t)opt.max_iterations + kGraphRounds);  // active count after each round (batched path)
  addi(B);                     // kocc
  addd(3 * (size_t)N);         // grad
  addd(3 * (size_t)NPC); addd(NPC);
  addd(B); addd(B);            // erep, eatm
  addd(kAtmFields * (size_t)NAP);  // ATM per-pair factors
  addi(N);                     // element-grouped atom order
  addd((size_t)kPairStageDoubles * N); addi((size_t)kPairStageInts * N);  // pair staging (large systems)
  addd((size_t)mem * mem * B);  // Broyden overlaps
  // work vectors of the systems above kMaxAO AOs (big_offset)
  const size_t xl_doubles = n_xl > 0 ? (size_t)kBigPerAO * NAO + (size_t)kBigPerSys * B : 0;
  if (xl_doubles > 0) addd(xl_doubles);
  addi(B); addi(NAO); addd(NAO); addi(NAO);  // eigenvalue clusters handed to k_cluster_vectors
  if (opt.no_alloc) {
    if (bytes > kNoAllocWorkspaceBytes) {
      error = "the batch needs " + std::to_string(bytes) +
              " bytes of device workspace; the no-allocation entry point holds at most " +
              std::to_string(kNoAllocWorkspaceBytes) + " bytes";
      return 1;
    }
  } else if (bytes > dev->arena_bytes) {
    (void)hipFree(dev->arena);
    dev->arena = nullptr;
    dev->arena_bytes = 0;
    dev->resident_gen = 0;
    const size_t want = bytes + bytes / 4;
    if (hipMalloc(&dev->arena, want) != hipSuccess) {
      error = "device arena allocation failed";
      return 4;
    }
    dev->arena_bytes = want;
  }
  if (timing) timing->workspace_bytes = bytes;
  if (opt.reserve_only) return 0;
  Carver cv{opt.no_alloc ? dev->scratch : dev->arena};
  hipStream_t sm = dev->stream;
  // uploads are collected here and issued together from pinned staging right
  // before the first kernel (the source vectors live until the call returns)
  struct Upload {
    void* dst;
    const void* src;
    size_t bytes;
  };
  std::vector<Upload> uploads;
  auto up = [&](auto* dst, const auto& vec) -> hipError_t {
    if (!vec.empty()) uploads.push_back({dst, vec.data(), vec.size() * sizeof(vec[0])});
    return hipSuccess;
  };
  // Topology arrays stay in the arena between calls: a call with the same
  // cached topology and the same arena layout uploads only the geometry.
  // (the no_alloc workspace is another memory region: it never holds them)
  const bool resident = !opt.no_alloc && cache_hit && cache->generation == dev->resident_gen &&
                        bytes == dev->resident_bytes;
  auto up_topo = [&](auto* dst, const auto& vec) -> hipError_t {
    return resident ? hipSuccess : up(dst, vec);
  };
  Topo t{};
  t.B = B;
  int* d_atom_off = cv.take<int>(B + 1);
  int* d_shell_off = cv.take<int>(B + 1);
  int* d_ao_off = cv.take<int>(B + 1);
  int* d_vec_off = cv.take<int>(B + 1);
  int* d_pc_off = cv.take<int>(B + 1);
  int* d_order = cv.take<int>(B);
  int64_t* d_mat_off = cv.take<int64_t>(B + 1);
  int64_t* d_gam_off = cv.take<int64_t>(B + 1);
  int64_t* d_apair_off = cv.take<int64_t>(B + 1);
  int64_t* d_sp_off = cv.take<int64_t>(B + 1);
  double* d_nalpha = cv.take<double>(B);
  double* d_nbeta = cv.take<double>(B);
  double* d_charges = cv.take<double>(B);
  int* d_Z = cv.take<int>(N);
  int* d_atom_sys = cv.take<int>(N);
  int* d_atom_sh0 = cv.take<int>(N + 1);
  int* d_d4ref = cv.take<int>(N);
  int* d_d4nref = cv.take<int>(N);
  double* d_xyz = cv.take<double>(3 * (size_t)N);
  double* at[15];
  for (int k = 0; k < 15; ++k) at[k] = cv.take<double>(N);
  int* shi[5];
  for (int k = 0; k < 5; ++k) shi[k] = cv.take<int>(NS);
  double* shd[7];
  for (int k = 0; k < 7; ++k) shd[k] = cv.take<double>(NS);
  int* d_ao_sh = cv.take<int>(NAO);
  int* d_ao_atom = cv.take<int>(NAO);
  double* d_alpha = cv.take<double>(T.NP);
  double* d_coef = cv.take<double>(T.NP);
  double* d_pcxyz = cv.take<double>(3 * (size_t)NPC);
  double* d_pcq = cv.take<double>(NPC);
  double* d_pcg = cv.take<double>(NPC);
  double* d_field = cv.take<double>(3 * (size_t)B);
  double* d_pshift = cv.take<double>(T.pshift.size());
  double* d_presp = cv.take<double>(T.presp.size());
  int* d_partner = cv.take<int>(T.partner.size());
  int* d_alpha_sys = cv.take<int>(alpha_sys.size());
  double* d_wspin = cv.take<double>(T.sh_wspin.size());
  TileTask* d_full = cv.take<TileTask>(full.size());
  TileTask* d_lower = cv.take<TileTask>(lower.size());
  BackTask* d_back = cv.take<BackTask>(back.size());
  FockTile* d_chunks = cv.take<FockTile>(chunks.size());
  SysList sp_list[2], ap_list[2];
  for (int c = 0; c < 2; ++c) {
    int* ls = cv.take<int>(cls_sys[c].size());
    int64_t* lsp = cv.take<int64_t>(cls_sp[c].size());
    int64_t* lap = cv.take<int64_t>(cls_ap[c].size());
    HIP_TRY(up_topo(ls, cls_sys[c]));
    HIP_TRY(up_topo(lsp, cls_sp[c]));
    HIP_TRY(up_topo(lap, cls_ap[c]));
    sp_list[c] = SysList{ls, lsp, (int)cls_sys[c].size()};
    ap_list[c] = SysList{ls, lap, (int)cls_sys[c].size()};
  }
  int64_t* d_part_off = cv.take<int64_t>(B + 1);
  double* d_mpart = cv.take<double>((size_t)part_off[B]);
  HIP_TRY(up_topo(d_atom_off, T.atom_off));
  HIP_TRY(up_topo(d_shell_off, T.shell_off));
  HIP_TRY(up_topo(d_ao_off, T.ao_off));
  HIP_TRY(up_topo(d_vec_off, T.vec_off));
  HIP_TRY(up_topo(d_pc_off, T.pc_off));
  HIP_TRY(up_topo(d_order, T.order));
  HIP_TRY(up_topo(d_mat_off, T.mat_off));
  HIP_TRY(up_topo(d_gam_off, T.gam_off));
  HIP_TRY(up_topo(d_apair_off, T.apair_off));
  HIP_TRY(up_topo(d_sp_off, T.sp_off));
  HIP_TRY(up_topo(d_nalpha, T.nalpha));
  HIP_TRY(up_topo(d_nbeta, T.nbeta));
  HIP_TRY(up_topo(d_charges, T.charges));
  HIP_TRY(up_topo(d_Z, T.Z));
  HIP_TRY(up_topo(d_atom_sys, T.atom_sys));
  HIP_TRY(up_topo(d_atom_sh0, T.atom_sh0));
  HIP_TRY(up_topo(d_d4ref, T.at_d4ref));
  HIP_TRY(up_topo(d_d4nref, T.at_d4nref));
  HIP_TRY(up(d_xyz, T.xyz));
  const std::vector<double>* atv[15] = {&T.at_cnrad, &T.at_h0rad, &T.at_en, &T.at_dk, &T.at_qk,
                                        &T.at_mprad, &T.at_mpvcn, &T.at_sqrt_arep, &T.at_zeff,
                                        &T.at_d4rcov, &T.at_d4en, &T.at_d4zeff, &T.at_d4hard,
                                        &T.at_d4r4r2, &T.at_cnrad};
  for (int k = 0; k < 15; ++k) HIP_TRY(up_topo(at[k], *atv[k]));
  const std::vector<int>* shiv[5] = {&T.sh_atom, &T.sh_l, &T.sh_ao, &T.sh_prim, &T.sh_nprim};
  for (int k = 0; k < 5; ++k) HIP_TRY(up_topo(shi[k], *shiv[k]));
  const std::vector<double>* shdv[7] = {&T.sh_level, &T.sh_kcn, &T.sh_kpoly, &T.sh_zeta,
                                        &T.sh_refocc, &T.sh_eta, &T.sh_gam3};
  for (int k = 0; k < 7; ++k) HIP_TRY(up_topo(shd[k], *shdv[k]));
  HIP_TRY(up_topo(d_ao_sh, T.ao_sh));
  HIP_TRY(up_topo(d_ao_atom, T.ao_atom));
  HIP_TRY(up_topo(d_alpha, T.alpha));
  HIP_TRY(up_topo(d_coef, T.coef));
  HIP_TRY(up(d_pcxyz, T.pc_xyz));
  HIP_TRY(up(d_pcq, T.pc_q));
  HIP_TRY(up_topo(d_pcg, T.pc_gam));
  HIP_TRY(up(d_field, T.field));
  HIP_TRY(up(d_pshift, T.pshift));
  HIP_TRY(up(d_presp, T.presp));
  HIP_TRY(up_topo(d_partner, T.partner));
  HIP_TRY(up_topo(d_alpha_sys, alpha_sys));
  HIP_TRY(up_topo(d_wspin, T.sh_wspin));
  HIP_TRY(up_topo(d_full, full));
  HIP_TRY(up_topo(d_lower, lower));
  HIP_TRY(up_topo(d_back, back));
  HIP_TRY(up_topo(d_chunks, chunks));
  HIP_TRY(up_topo(d_part_off, part_off));

  t.atom_off = d_atom_off; t.shell_off = d_shell_off; t.ao_off = d_ao_off; t.vec_off = d_vec_off;
  t.pc_off = d_pc_off; t.mat_off = d_mat_off; t.gam_off = d_gam_off; t.apair_off = d_apair_off;
  t.nalpha = d_nalpha; t.nbeta = d_nbeta;
  t.Z = d_Z; t.xyz = d_xyz; t.atom_sys = d_atom_sys; t.atom_sh0 = d_atom_sh0;
  t.at_cnrad = at[0]; t.at_h0rad = at[1]; t.at_en = at[2]; t.at_dk = at[3]; t.at_qk = at[4];
  t.at_mprad = at[5]; t.at_mpvcn = at[6]; t.at_sqrt_arep = at[7]; t.at_zeff = at[8];
  t.at_d4rcov = at[9]; t.at_d4en = at[10]; t.at_d4zeff = at[11]; t.at_d4hard = at[12];
  t.at_d4r4r2 = at[13]; t.at_d4ref = d_d4ref; t.at_d4nref = d_d4nref;
  t.sh_atom = shi[0]; t.sh_l = shi[1]; t.sh_ao = shi[2]; t.sh_prim = shi[3]; t.sh_nprim = shi[4];
  t.sh_level = shd[0]; t.sh_kcn = shd[1]; t.sh_kpoly = shd[2]; t.sh_zeta = shd[3];
  t.sh_refocc = shd[4]; t.sh_eta = shd[5]; t.sh_gam3 = shd[6];
  t.ao_sh = d_ao_sh; t.ao_atom = d_ao_atom;
  t.alpha = d_alpha; t.coef = d_coef;
  t.pc_xyz = d_pcxyz; t.pc_q = d_pcq; t.pc_gam = d_pcg;
  t.field = T.field.empty() ? nullptr : d_field;
  t.pshift = T.pshift.empty() ? nullptr : d_pshift;
  t.presp = T.presp.empty() ? nullptr : d_presp;
  t.partner = U > 0 ? d_partner : nullptr;
  t.sh_wspin = U > 0 ? d_wspin : nullptr;
  t.d4_c6 = dev->d4_c6; t.d4_refcn = dev->d4_refcn; t.d4_refq = dev->d4_refq;
  t.d4_refng = dev->d4_refng; t.kpair = dev->kpair;

  double* mats = cv.take<double>((size_t)kMat * NN);
  double* dS = mats;
  double* dD = mats + NN;
  double* dQ = mats + 4 * NN;
  double* dH0 = mats + 10 * NN;
  double* dX = mats + 11 * NN;
  double* dH = mats + 12 * NN;
  double* dA = mats + 13 * NN;
  double* dZ = mats + 14 * NN;
  double* dY = mats + 15 * NN;
  double* dC = mats + 16 * NN;
  double* dP = mats + 17 * NN;
  double* dW = mats + 18 * NN;
  double* dT = mats + 19 * NN;
  double* d_gam = cv.take<double>(NG);
  double* d_cn = cv.take<double>(N);
  double* d_cnd4 = cv.take<double>(N);
  double* d_mrad = cv.take<double>(N);
  double* d_dmrad = cv.take<double>(N);
  double* d_qat = cv.take<double>(N);
  double* d_vat = cv.take<double>(N);
  double* d_dedcn = cv.take<double>(N);
  double* d_dedcn4 = cv.take<double>(N);
  double* d_w4 = cv.take<double>((size_t)kD4MaxRef * N);
  double* d_dw4 = cv.take<double>((size_t)kD4MaxRef * N);
  double* d_vpc = cv.take<double>(NS);
  double* d_vsh = cv.take<double>(NS);
  double* d_vd = cv.take<double>(3 * (size_t)N);
  double* d_vqq = cv.take<double>(6 * (size_t)N);
  double* d_eps = cv.take<double>(NAO);
  double* d_fa = cv.take<double>(NAO);
  double* d_fb = cv.take<double>(NAO);
  double* d_occ = cv.take<double>(NAO);
  double* d_fe = cv.take<double>(NAO);
  double* d_shift = cv.take<double>(B);
  double* d_zscale = cv.take<double>(NAO);
  double* d_td = cv.take<double>(NAO);
  double* d_wyT = cv.take<double>((size_t)kBackChunk * NAO + (size_t)kBackChunk * kBackChunk * B);
  double* d_te = cv.take<double>(NAO);
  double* d_colsum = cv.take<double>((size_t)NAO * 10);
  MixerArrays mx{};
  mx.x = cv.take<double>(NV);
  mx.raw = cv.take<double>(NV);
  mx.cur = cv.take<double>(NV);
  mx.prev = cv.take<double>(NV);
  mx.prev_res = cv.take<double>(NV);
  mx.df = cv.take<double>((size_t)mem * NV);
  mx.u = cv.take<double>((size_t)mem * NV);
  mx.omega = cv.take<double>((size_t)mem * B);
  mx.mem = mem;
  mx.damping = opt.damping;
  // WARM start: the state arrives in raw (free until the first Mulliken
  // step) and replaces the FRESH state right after the geometry kernel
  if (opt.warm_state) uploads.push_back({mx.raw, opt.warm_state, sizeof(double) * (size_t)NV});
  mx.qtol = opt.charge_tolerance;
  mx.etol = opt.energy_tolerance;
  mx.kt = opt.kt;
  mx.max_iter = opt.max_iterations;
  SysState* d_st = cv.take<SysState>(B);
  unsigned char* d_active = cv.take<unsigned char>(B);
  unsigned char* d_valid = cv.take<unsigned char>(B);
  int* d_count = cv.take<int>(2);
  int* d_round = cv.take<int>(1);
  int* d_cnt_round = cv.take<int>((size_t)opt.max_iterations + kGraphRounds);
  int* d_kocc = cv.take<int>(B);
  double* d_grad = cv.take<double>(3 * (size_t)N);
  double* d_pcgrad = cv.take<double>(3 * (size_t)NPC);
  double* d_dedq = cv.take<double>(NPC);
  double* d_erep = cv.take<double>(B);
  double* d_eatm = cv.take<double>(B);
  double* d_c6pair = cv.take<double>(kAtmFields * (size_t)NAP);
  int* d_perm = cv.take<int>(N);
  double* d_gstage = cv.take<double>((size_t)kPairStageDoubles * N);
  int* d_gstagei = cv.take<int>((size_t)kPairStageInts * N);
  mx.ovl = cv.take<double>((size_t)mem * mem * B);
  double* d_xl = n_xl > 0 ? cv.take<double>(xl_doubles) : nullptr;
  ClusterOut co{};
  co.flag = cv.take<int>(B);
  co.end = cv.take<int>(NAO);
  co.lam = cv.take<double>(NAO);
  co.rank = cv.take<int>(NAO);
  HIP_TRY(up_topo(d_perm, T.perm));
  t.perm = d_perm;
  const bool big_systems = T.max_atoms > kPairStageAtoms;
  auto launch_potentials = [&](const unsigned char* mask) {
    hipLaunchKernelGGL(k_potentials<true>, dim3(B), dim3(512), 0, sm, t, d_order, mask, mx.x,
                       d_cnd4, d_mrad, d_gam, d_vpc, d_qat, d_w4, d_dw4, d_vat, d_vsh, d_vd, d_vqq,
                       d_gstage, d_gstagei);
    if (big_systems)
      hipLaunchKernelGGL(k_potentials<false>, dim3(B), dim3(512), 0, sm, t, d_order, mask, mx.x,
                         d_cnd4, d_mrad, d_gam, d_vpc, d_qat, d_w4, d_dw4, d_vat, d_vsh, d_vd,
                         d_vqq, d_gstage, d_gstagei);
  };

  HIP_TRY(hipMemsetAsync(d_st, 0, sizeof(SysState) * B, sm));
  HIP_TRY(hipMemsetAsync(co.flag, 0, sizeof(int) * B, sm));
  HIP_TRY(hipMemsetAsync(co.end, 0, sizeof(int) * (size_t)NAO, sm));
  HIP_TRY(hipMemsetAsync(mx.df, 0, sizeof(double) * (size_t)mem * NV, sm));
  HIP_TRY(hipMemsetAsync(mx.u, 0, sizeof(double) * (size_t)mem * NV, sm));
  HIP_TRY(hipMemsetAsync(mx.omega, 0, sizeof(double) * (size_t)mem * B, sm));

  // ---- optional per-launch timing (HIP events on this stream) ----
  KernelStat* ks = opt.no_alloc ? nullptr : out.kernel_stats;
  struct ProbeRec {
    int kind;
    double flops, bytes;
  };
  std::vector<ProbeRec> probes;
  // families outside out.kernel_stats_mask (0 = every family) get no events
  bool rec = false;  // the current begin/end pair records events
  auto probe_begin = [&](int kind) -> hipError_t {
    rec = ks && (out.kernel_stats_mask == 0u || ((out.kernel_stats_mask >> kind) & 1u));
    if (!rec) return hipSuccess;
    while (dev->probe_pool.size() < 2 * (probes.size() + 1)) {
      hipEvent_t e;
      const hipError_t err = hipEventCreate(&e);
      if (err != hipSuccess) return err;
      dev->probe_pool.push_back(e);
    }
    return hipEventRecord(dev->probe_pool[2 * probes.size()], sm);
  };
  auto probe_end = [&](int kind, double flops, double bytes) -> hipError_t {
    if (!rec) return hipSuccess;
    const hipError_t err = hipEventRecord(dev->probe_pool[2 * probes.size() + 1], sm);
    probes.push_back({kind, flops, bytes});
    return err;
  };
  // algorithmic work of a system subset: sum n^2 and sum n^3
  std::vector<unsigned char> h_active(B, 1);
  auto work = [&](const std::vector<unsigned char>& mask, double& n2, double& n3) {
    n2 = n3 = 0.0;
    for (int s = 0; s < B; ++s) {
      if (!mask[s]) continue;
      const double n = T.ao_off[s + 1] - T.ao_off[s];
      n2 += n * n;
      n3 += n * n * n;
    }
  };
  // sum n^2 kocc of a subset (the eigenvector-limited products)
  std::vector<int> h_kocc(B, 0);
  auto work_occ = [&](const std::vector<unsigned char>& mask) {
    double w = 0.0;
    for (int s = 0; s < B; ++s) {
      if (!mask[s]) continue;
      const double n = T.ao_off[s + 1] - T.ao_off[s];
      w += n * n * h_kocc[s];
    }
    return w;
  };
  double all2, all3;
  work(h_active, all2, all3);

  {
    size_t need = 256 + 2 * align_bytes<unsigned char>(B) + 2 * align_bytes<int>(B);
    for (const Upload& u : uploads) need += align_bytes<char>(u.bytes);
    // copies of an earlier call that failed midway may still read the staging
    HIP_TRY(hipStreamSynchronize(sm));
    if (need > dev->pinned_bytes && !opt.no_alloc) {
      if (dev->pinned) (void)hipHostFree(dev->pinned);
      dev->pinned = nullptr;
      dev->pinned_bytes = 0;
      HIP_TRY(hipHostMalloc((void**)&dev->pinned, need + need / 4));
      dev->pinned_bytes = need + need / 4;
    }
    if (need <= dev->pinned_bytes) {
      size_t off = 256 + 2 * align_bytes<unsigned char>(B) + 2 * align_bytes<int>(B);
      for (const Upload& u : uploads) {
        std::memcpy(dev->pinned + off, u.src, u.bytes);
        HIP_TRY(hipMemcpyAsync(u.dst, dev->pinned + off, u.bytes, hipMemcpyHostToDevice, sm));
        off += align_bytes<char>(u.bytes);
      }
    } else {
      // no_alloc with a smaller pinned block: copies from pageable memory
      // (the runtime stages them before the call returns); the pinned block
      // still holds the per-round counts (256 bytes, allocated with the device)
      for (const Upload& u : uploads)
        HIP_TRY(hipMemcpyAsync(u.dst, u.src, u.bytes, hipMemcpyHostToDevice, sm));
    }
  }
  // the arena now holds this call's topology (stream order: every later call
  // on this device reads it after these copies)
  if (!opt.no_alloc) {
    dev->resident_gen = cache ? cache->generation : 0;
    dev->resident_bytes = bytes;
  }
  // pinned per-parity round results: active counts, and (with kernel stats)
  // snapshots of the active mask and kocc at the end of each round
  int* h_count = reinterpret_cast<int*>(dev->pinned);
  unsigned char* h_act_snap[2];
  int* h_kocc_snap[2];
  {
    char* q = dev->pinned + 256;
    for (int k = 0; k < 2; ++k) {
      h_act_snap[k] = reinterpret_cast<unsigned char*>(q);
      q += align_bytes<unsigned char>(B);
    }
    for (int k = 0; k < 2; ++k) {
      h_kocc_snap[k] = reinterpret_cast<int*>(q);
      q += align_bytes<int>(B);
    }
  }
  HIP_TRY(hipEventRecord(dev->ev[0], sm));
  // ---- geometry, integrals, overlap factorization ----
  HIP_TRY(probe_begin(kKGeometry));
  hipLaunchKernelGGL(k_geometry, dim3(B), dim3(256), 0, sm, t, d_order, d_cn, d_cnd4, d_mrad,
                     d_dmrad, d_gam, d_vpc, mx, d_st, d_charges);
  HIP_TRY(probe_end(kKGeometry, 0.0, 0.0));
  if (opt.warm_state) {
    HIP_TRY(hipMemcpyAsync(mx.x, mx.raw, sizeof(double) * (size_t)NV, hipMemcpyDeviceToDevice, sm));
    HIP_TRY(hipMemcpyAsync(mx.cur, mx.raw, sizeof(double) * (size_t)NV, hipMemcpyDeviceToDevice, sm));
  }
  if (NSP > 0) {
    HIP_TRY(probe_begin(kKIntegrals));
    if (cls_sp[0].back() > 0)
      hipLaunchKernelGGL((k_integrals<false>), dim3((unsigned)((cls_sp[0].back() + 127) / 128)), dim3(128), 0,
                         sm, t, sp_list[0], cls_sp[0].back(), d_cn, dS, dD, dQ, dH0);
    if (cls_sp[1].back() > 0)
      hipLaunchKernelGGL((k_integrals<true>), dim3((unsigned)((cls_sp[1].back() + 127) / 128)), dim3(128), 0,
                         sm, t, sp_list[1], cls_sp[1].back(), d_cn, dS, dD, dQ, dH0);
    HIP_TRY(probe_end(kKIntegrals, 0.0, 88.0 * all2));  // S + 3 D + 6 Q + H0 written
  }
  HIP_TRY(probe_begin(kKCholesky));
  // systems are ordered largest first: the n_large above 256 AOs take the
  // blocked factorization, the rest the register one
  if (n_large > 0)
    hipLaunchKernelGGL(k_cholesky_blocked, dim3(n_large), dim3(512), 0, sm, t, d_order, dS, dA, dX, d_st);
  if (n_large < B)
    hipLaunchKernelGGL(k_cholesky_inverse, dim3(B - n_large), dim3(512), 0, sm, t, d_order + n_large, dS, dA,
                       dX, d_st);
  if (n_xl > 0)
    hipLaunchKernelGGL(k_overlap_rcond<true>, dim3(n_xl), dim3(kRcondThreads), 0, sm, t, d_order, dS, dX, d_st,
                       d_xl);
  if (n_xl < B)
    hipLaunchKernelGGL(k_overlap_rcond<false>, dim3(B - n_xl), dim3(kRcondThreads), 0, sm, t, d_order + n_xl,
                       dS, dX, d_st, d_xl);
  HIP_TRY(probe_end(kKCholesky, all3 / 3.0, 0.0));
  hipLaunchKernelGGL(k_init_active, dim3((B + 255) / 256), dim3(256), 0, sm, B, d_st, d_active,
                     (int*)nullptr, d_st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(dev->ev[1], sm));
  if (ks) {
    HIP_TRY(hipMemcpyAsync(h_active.data(), d_active, B, hipMemcpyDeviceToHost, sm));
    HIP_TRY(hipStreamSynchronize(sm));
  }

  // ---- SCC rounds ----
  // Round it+1 is enqueued before the host waits for round it's active count,
  // so the device never idles on the host.  Once a round reports no active
  // system the loop stops; the round already enqueued behind it finds every
  // system inactive and its blocks return at once.  Per-round work figures
  // of the kernel stats are filled in when the round's active set and kocc
  // are known (one round later).
  const unsigned nfull = (unsigned)full.size(), nlower = (unsigned)lower.size();
  int rounds = 0;
  struct RoundProbes {
    size_t fock, red1, red2, tri, mull, bt, back, dens;  // probe index + 1 (0: none)
  };
  std::vector<RoundProbes> rprobes;
  std::vector<std::vector<unsigned char>> start_set;  // active set at the start of each round
  if (ks) start_set.push_back(h_active);
  auto last_probe = [&]() -> size_t { return rec ? probes.size() : 0; };
  auto patch_round = [&](int r, const std::vector<int>& kocc_after) {
    const RoundProbes& rp = rprobes[r];
    double a2 = 0.0, a3 = 0.0;
    work(start_set[r], a2, a3);
    h_kocc = kocc_after;
    const double w = work_occ(start_set[r]);
    auto set = [&](size_t idx, double flops, double bytes) {
      if (idx) { probes[idx - 1].flops = flops; probes[idx - 1].bytes = bytes; }
    };
    set(rp.fock, 0.0, 52.0 * a2);      // reads half of S, D(3), Q(6), H0 (H symmetric); writes H
    set(rp.red1, a3, 0.0);             // X^T H, X upper triangular: n^3
    set(rp.red2, a3 / 3.0, 0.0);       // (X^T H) X, symmetric: n^3/3
    set(rp.tri, (4.0 / 3.0) * a3, 0.0);
    set(rp.mull, 0.0, 48.0 * a2);      // reads half of P, S, D(3), Q(6), H0 (P symmetric)
    // back-transformation 2 n^2 kocc, C = X Y n^2 kocc (X upper triangular),
    // P (lower half) n^2 kocc
    set(rp.bt, 2.0 * w, 0.0);
    set(rp.back, w, 0.0);
    set(rp.dens, w, 0.0);
  };
  // waits for round r and records its completion; `count` receives the number
  // of systems still active after it (the return value is a status code)
  auto finish_round = [&](int r, int& count) -> int {
    const int par = r & 1;
    HIP_TRY(hipEventSynchronize(dev->round_ev[par]));
    ++rounds;
    if (ks) {
      start_set.emplace_back(h_act_snap[par], h_act_snap[par] + B);
      patch_round(r, std::vector<int>(h_kocc_snap[par], h_kocc_snap[par] + B));
    }
    count = h_count[par];
    return 0;
  };
  // One SCC round's launches.  rptr: null in the host-driven loop (round `it`,
  // active count into cnt), else the device-resident round counter (the
  // eigen kernel reads the round from it and k_energy_mix counts the systems
  // still active into cnt[round]).
  auto enqueue_round = [&](int it, int* cnt, const int* rptr, RoundProbes* rp) -> int {
    HIP_TRY(probe_begin(kKPotentials));
    launch_potentials(d_active);
    if (U > 0)
      hipLaunchKernelGGL(k_spin_potentials, dim3(U), dim3(256), 0, sm, t, d_alpha_sys, d_active, mx.x, d_vsh, d_vd,
                         d_vqq);
    HIP_TRY(probe_end(kKPotentials, 0.0, 0.0));
    HIP_TRY(probe_begin(kKFock));
    hipLaunchKernelGGL(k_fock, dim3((unsigned)chunks.size()), dim3(256), 0, sm, t, d_chunks,
                       d_active, dS, dD, dQ, dH0, d_vsh, d_vd, d_vqq, dH);
    HIP_TRY(probe_end(kKFock, 0.0, 0.0));
    if (rp) rp->fock = last_probe();
    // H' = X^T H X
    HIP_TRY(probe_begin(kKReduce));
    hipLaunchKernelGGL((k_gemm<true, false, false, 1>), dim3(nfull), dim3(256), 0, sm, t, d_full,
                       dX, dH, dT, (const double*)nullptr, (const int*)nullptr,
                       (const int*)nullptr, d_active);
    HIP_TRY(probe_end(kKReduce, 0.0, 0.0));
    if (rp) rp->red1 = last_probe();
    HIP_TRY(probe_begin(kKReduce));
    hipLaunchKernelGGL((k_gemm<false, false, true, 2>), dim3(nlower), dim3(256), 0, sm, t, d_lower,
                       dT, dX, dA, (const double*)nullptr, (const int*)nullptr,
                       (const int*)nullptr, d_active);
    HIP_TRY(probe_end(kKReduce, 0.0, 0.0));
    if (rp) rp->red2 = last_probe();
    // eigenpairs of H' (A is overwritten by the reflectors; T, H, C, P, W serve as scratch)
    HIP_TRY(probe_begin(kKTridiag));
    // systems are ordered largest first: the n_large above 256 AOs take the
    // blocked reduction (its first n_xl, above kMaxAO, with global work
    // vectors), the rest the register one
    if (n_xl > 0)
      hipLaunchKernelGGL(k_tridiag_blocked<true>, dim3(n_xl), dim3(512), 0, sm, t, d_order, dA, d_td,
                         d_te, dT, d_active, d_xl);
    if (n_large > n_xl)
      hipLaunchKernelGGL(k_tridiag_blocked<false>, dim3(n_large - n_xl), dim3(512), 0, sm, t, d_order + n_xl,
                         dA, d_td, d_te, dT, d_active, d_xl);
    if (n_large < B)
      hipLaunchKernelGGL(k_tridiag, dim3(B - n_large), dim3(512), 0, sm, t, d_order + n_large, dA,
                         d_td, d_te, d_wyT, d_active);
    HIP_TRY(probe_end(kKTridiag, 0.0, 0.0));
    if (rp) rp->tri = last_probe();
    HIP_TRY(probe_begin(kKTridiagEigen));
    // eigenvalues, occupations (kocc), eigenvectors of the kocc lowest levels
    if (n_xl > 0)
      hipLaunchKernelGGL(k_tridiag_eigen<true>, dim3(n_xl), dim3(256), 0, sm, t, d_order, d_active, d_td,
                         d_te, dZ, EigenScratch{dT, dH, dC, dP, dW}, d_eps, d_shift, it, rptr, d_zscale,
                         FermiOut{d_fa, d_fb, d_occ, d_fe, d_st, d_kocc, opt.kt}, d_xl, co);
    if (n_xl < B)
      hipLaunchKernelGGL(k_tridiag_eigen<false>, dim3(B - n_xl), dim3(256), 0, sm, t, d_order + n_xl,
                         d_active, d_td, d_te, dZ, EigenScratch{dT, dH, dC, dP, dW}, d_eps, d_shift, it,
                         rptr, d_zscale, FermiOut{d_fa, d_fb, d_occ, d_fe, d_st, d_kocc, opt.kt}, d_xl, co);
    // clusters of more than kSerialCluster eigenvalues (rare: near-dependent bases)
    hipLaunchKernelGGL(k_cluster_vectors, dim3(B), dim3(256), 0, sm, t, d_order, d_active, d_td, d_te, dZ,
                       EigenScratch{dT, dH, dC, dP, dW}, d_zscale, co);
    HIP_TRY(probe_end(kKTridiagEigen, 0.0, 0.0));
    HIP_TRY(probe_begin(kKBacktransform));
    // systems above 256 AOs: the WY factors of their reflectors, then the WY
    // back-transformation (in place in Y above kMaxAO); the rest: registers
    if (n_xl > 0)
      hipLaunchKernelGGL(k_wy_factors<true>, dim3(n_xl), dim3(512), 0, sm, t, d_order, dA, d_wyT, d_active,
                         d_xl);
    if (n_large > n_xl)
      hipLaunchKernelGGL(k_wy_factors<false>, dim3(n_large - n_xl), dim3(512), 0, sm, t, d_order + n_xl, dA,
                         d_wyT, d_active, d_xl);
    if (n_back_xl > 0)
      hipLaunchKernelGGL(k_backtransform_wy<true>, dim3((unsigned)n_back_xl), dim3(kWyThreads), back_lds_xl, sm,
                         t, d_back, dA, dZ, d_wyT, dY, d_kocc, d_zscale, d_active);
    if (n_back_large > n_back_xl)
      hipLaunchKernelGGL(k_backtransform_wy<false>, dim3((unsigned)(n_back_large - n_back_xl)),
                         dim3(kWyThreads), back_lds, sm, t, d_back + n_back_xl, dA, dZ, d_wyT, dY, d_kocc,
                         d_zscale, d_active);
    if (back.size() > n_back_large)
      hipLaunchKernelGGL(k_backtransform, dim3((unsigned)(back.size() - n_back_large)), dim3(kBackThreads), 0,
                         sm, t, d_back + n_back_large, dA, dZ, d_wyT, dY, d_kocc, d_zscale, d_active);
    HIP_TRY(probe_end(kKBacktransform, 0.0, 0.0));
    if (rp) rp->bt = last_probe();
    // C = X Y (the kocc leading columns), P = C diag(f) C^T
    HIP_TRY(probe_begin(kKBack));
    hipLaunchKernelGGL((k_gemm<false, false, false, 3>), dim3(nfull), dim3(256), 0, sm, t, d_full,
                       dX, dY, dC, (const double*)nullptr, (const int*)nullptr, d_kocc, d_active);
    HIP_TRY(probe_end(kKBack, 0.0, 0.0));
    if (rp) rp->back = last_probe();
    HIP_TRY(probe_begin(kKDensity));
    hipLaunchKernelGGL((k_gemm<false, true, true>), dim3(nlower), dim3(256), 0, sm, t, d_lower, dC,
                       dC, dP, d_occ, d_kocc, (const int*)nullptr, d_active,
                       opt.split32 ? (const SysState*)d_st : nullptr);
    if (opt.split32)
      hipLaunchKernelGGL(k_gemm_split32, dim3(nlower), dim3(256), 0, sm, t, d_lower, dC, dP, d_occ, d_kocc,
                         (const SysState*)d_st, d_active);
    HIP_TRY(probe_end(kKDensity, 0.0, 0.0));
    if (rp) rp->dens = last_probe();
    HIP_TRY(probe_begin(kKMulliken));
    hipLaunchKernelGGL(k_mulliken_rows, dim3((unsigned)chunks.size()), dim3(256), 0, sm, t,
                       d_chunks, d_active, d_part_off, dP, dS, dD, dQ, dH0, d_mpart);
    hipLaunchKernelGGL(k_mulliken_finish, dim3(B), dim3(256), 0, sm, t, d_order, d_active, d_part_off,
                       d_mpart, d_colsum, mx.raw, d_st);
    if (U > 0)
      hipLaunchKernelGGL(k_spin_combine, dim3(U), dim3(256), 0, sm, t, d_alpha_sys, d_active, mx.raw, d_st);
    HIP_TRY(probe_end(kKMulliken, 0.0, 0.0));
    if (rp) rp->mull = last_probe();
    HIP_TRY(probe_begin(kKEnergyMix));
    hipLaunchKernelGGL(k_energy_mix<true>, dim3(B), dim3(512), 0, sm, t, d_order, d_active,
                       d_cnd4, d_mrad, d_gam, d_vpc, d_qat, d_w4, mx, d_st, cnt, d_gstage,
                       d_gstagei, rptr);
    if (big_systems)
      hipLaunchKernelGGL(k_energy_mix<false>, dim3(B), dim3(512), 0, sm, t, d_order, d_active,
                         d_cnd4, d_mrad, d_gam, d_vpc, d_qat, d_w4, mx, d_st, cnt, d_gstage,
                         d_gstagei, rptr);
    HIP_TRY(probe_end(kKEnergyMix, 0.0, 0.0));
    HIP_TRY(hipGetLastError());
    return 0;
  };
  int enqueued = 0;
  bool stopped = false;
  // a stream that cannot be captured (the legacy null stream) runs the
  // host-driven loop
  bool graph_rounds = opt.batched && !ks;
  if (graph_rounds) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (sm == nullptr || hipStreamIsCapturing(sm, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
      graph_rounds = false;
  }
  if (graph_rounds) {
    // Batched path: kGraphRounds rounds per captured graph, the round number
    // and the per-round active counts on the device.  The counters are reset
    // here, outside the graph (no memset node is captured).  The instantiated
    // graph is kept on the device while the arena, the topology and the policy
    // it was captured for are unchanged (every kernel argument is a function
    // of those).
    HIP_TRY(hipMemsetAsync(d_round, 0, sizeof(int), sm));
    HIP_TRY(hipMemsetAsync(d_cnt_round, 0, sizeof(int) * ((size_t)opt.max_iterations + kGraphRounds), sm));
    std::vector<char> key;
    auto put = [&](const auto& v) {
      const char* p = reinterpret_cast<const char*>(&v);
      key.insert(key.end(), p, p + sizeof(v));
    };
    const char* arena_base = opt.no_alloc ? dev->scratch : dev->arena;
    put(arena_base); put(bytes); put(B); put(N); put(NS); put(NAO); put(NPC);
    put(cache ? cache->generation : ~uint64_t(0)); put(mem); put(opt.damping);
    put(opt.charge_tolerance); put(opt.energy_tolerance); put(opt.kt); put(opt.max_iterations);
    put(opt.split32); put(t.field); put(t.pshift); put(t.presp); put(U);
    const bool reuse = cache != nullptr && dev->graph_exec != nullptr && dev->graph_key == key;
    if (!reuse) {
      if (dev->graph_exec) (void)hipGraphExecDestroy(dev->graph_exec);
      dev->graph_exec = nullptr;
      dev->graph_key.clear();
      HIP_TRY(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
      int st_cap = 0;
      for (int r = 0; r < kGraphRounds && st_cap == 0; ++r) {
        st_cap = enqueue_round(0, d_cnt_round, d_round, nullptr);
        if (st_cap == 0) hipLaunchKernelGGL(k_round_advance, dim3(1), dim3(64), 0, sm, d_round);
      }
      hipGraph_t graph = nullptr;
      const hipError_t ce = hipStreamEndCapture(sm, &graph);
      if (st_cap != 0) {
        if (graph) (void)hipGraphDestroy(graph);
        return st_cap;
      }
      HIP_TRY(ce);
      const hipError_t ie = hipGraphInstantiate(&dev->graph_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIP_TRY(ie);
      dev->graph_key = key;
    }
    int launched = 0, g = 0;
    while (launched < opt.max_iterations) {
      HIP_TRY(hipGraphLaunch(dev->graph_exec, sm));
      launched += kGraphRounds;
      HIP_TRY(hipMemcpyAsync(h_count + (g & 1), d_cnt_round + launched - 1, sizeof(int),
                             hipMemcpyDeviceToHost, sm));
      HIP_TRY(hipEventRecord(dev->round_ev[g & 1], sm));
      ++g;
      if (g >= 2) {
        HIP_TRY(hipEventSynchronize(dev->round_ev[g & 1]));  // graph g-2
        if (h_count[g & 1] == 0) break;
      }
    }
    HIP_TRY(hipEventSynchronize(dev->round_ev[(g - 1) & 1]));
    // rounds that started with an active system: up to the first round after
    // which none was left
    std::vector<int> cnt((size_t)std::min(launched, opt.max_iterations));
    HIP_TRY(hipMemcpyAsync(cnt.data(), d_cnt_round, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost, sm));
    HIP_TRY(hipStreamSynchronize(sm));
    rounds = (int)cnt.size();
    for (size_t r = 0; r < cnt.size(); ++r)
      if (cnt[r] == 0) { rounds = (int)r + 1; break; }
  }
  for (int it = 0; it < opt.max_iterations && !graph_rounds; ++it) {
    const int par = it & 1;
    int* d_cnt = d_count + par;
    RoundProbes rp{};
    HIP_TRY(hipMemsetAsync(d_cnt, 0, sizeof(int), sm));
    if (const int st_round = enqueue_round(it, d_cnt, nullptr, &rp)) return st_round;
    HIP_TRY(hipMemcpyAsync(h_count + par, d_cnt, sizeof(int), hipMemcpyDeviceToHost, sm));
    if (ks) {
      HIP_TRY(hipMemcpyAsync(h_act_snap[par], d_active, B, hipMemcpyDeviceToHost, sm));
      HIP_TRY(hipMemcpyAsync(h_kocc_snap[par], d_kocc, sizeof(int) * B, hipMemcpyDeviceToHost, sm));
    }
    HIP_TRY(hipEventRecord(dev->round_ev[par], sm));
    rprobes.push_back(rp);
    enqueued = it + 1;
    if (it >= 1) {
      int still_active = 0;
      if (const int st_round = finish_round(it - 1, still_active)) return st_round;
      if (still_active == 0) {
        stopped = true;
        break;
      }
    }
  }
  // the last enqueued round: either the speculative one behind a finished
  // batch (no active system) or the final round at the iteration limit
  if (enqueued > 0) {
    const int last = enqueued - 1;
    HIP_TRY(hipEventSynchronize(dev->round_ev[last & 1]));
    if (ks) patch_round(last, std::vector<int>(h_kocc_snap[last & 1], h_kocc_snap[last & 1] + B));
    if (!stopped) ++rounds;
  }
  HIP_TRY(hipEventRecord(dev->ev[2], sm));

  // ---- stationary energy and forces ----
  hipLaunchKernelGGL(k_valid, dim3((B + 255) / 256), dim3(256), 0, sm, B, d_st, d_valid);
  std::vector<unsigned char> h_valid(B, 1);
  double v2 = 0.0, v3 = 0.0;
  if (ks) {
    HIP_TRY(hipMemcpyAsync(h_valid.data(), d_valid, B, hipMemcpyDeviceToHost, sm));
    HIP_TRY(hipStreamSynchronize(sm));
    work(h_valid, v2, v3);
  }
  HIP_TRY(hipMemsetAsync(d_grad, 0, sizeof(double) * 3 * (size_t)N, sm));
  HIP_TRY(hipMemsetAsync(d_dedcn, 0, sizeof(double) * (size_t)N, sm));
  HIP_TRY(hipMemsetAsync(d_erep, 0, sizeof(double) * (size_t)B, sm));
  HIP_TRY(hipMemsetAsync(d_eatm, 0, sizeof(double) * (size_t)B, sm));
  if (NPC > 0) {
    HIP_TRY(hipMemsetAsync(d_pcgrad, 0, sizeof(double) * 3 * (size_t)NPC, sm));
    HIP_TRY(hipMemsetAsync(d_dedq, 0, sizeof(double) * (size_t)NPC, sm));
  }
  // potentials at the converged raw multipoles (reference gfn2_cpu_execution.cpp:1358-1441)
  HIP_TRY(probe_begin(kKPotentials));
  launch_potentials(d_valid);
  HIP_TRY(probe_end(kKPotentials, 0.0, 0.0));
  if (opt.forces) {
    // W = C diag(f eps) C^T
    HIP_TRY(probe_begin(kKEnergyWeighted));
    if (opt.split32)
      hipLaunchKernelGGL(k_gemm_split32, dim3(nlower), dim3(256), 0, sm, t, d_lower, dC, dW, d_fe, d_kocc,
                         (const SysState*)nullptr, d_valid);
    else
      hipLaunchKernelGGL((k_gemm<false, true, true>), dim3(nlower), dim3(256), 0, sm, t, d_lower, dC,
                         dC, dW, d_fe, d_kocc, (const int*)nullptr, d_valid);
    HIP_TRY(probe_end(kKEnergyWeighted, work_occ(h_valid), 0.0));  // lower half, K = kocc
    // unrestricted: total P and W with the magnetization term, beta copies out
    if (U_xl > 0)
      hipLaunchKernelGGL(k_spin_fold<true>, dim3(U_xl), dim3(256), 0, sm, t, d_alpha_sys, d_valid, mx.x, dP,
                         dW, d_xl);
    if (U > U_xl)
      hipLaunchKernelGGL(k_spin_fold<false>, dim3(U - U_xl), dim3(256), 0, sm, t, d_alpha_sys + U_xl, d_valid,
                         mx.x, dP, dW, d_xl);
    if (NAP > 0) {
      HIP_TRY(probe_begin(kKForcePairs));
      for (int c = 0; c < 2; ++c) {
        const int64_t np = cls_ap[c].back();
        if (np == 0) continue;
        const dim3 grid((unsigned)((np + 127) / 128));
        if (c == 0 && !opt.reproducible)
          hipLaunchKernelGGL((k_force_pairs<false, false>), grid, dim3(128), 0, sm, t, ap_list[c], np,
                             d_valid, d_cn, dP, dW, dS, dD, d_vsh, d_vd, d_vqq, d_grad, d_dedcn);
        else if (c == 0)
          hipLaunchKernelGGL((k_force_pairs<false, true>), grid, dim3(128), 0, sm, t, ap_list[c], np,
                             d_valid, d_cn, dP, dW, dS, dD, d_vsh, d_vd, d_vqq, d_grad, d_dedcn);
        else if (!opt.reproducible)
          hipLaunchKernelGGL((k_force_pairs<true, false>), grid, dim3(128), 0, sm, t, ap_list[c], np,
                             d_valid, d_cn, dP, dW, dS, dD, d_vsh, d_vd, d_vqq, d_grad, d_dedcn);
        else
          hipLaunchKernelGGL((k_force_pairs<true, true>), grid, dim3(128), 0, sm, t, ap_list[c], np,
                             d_valid, d_cn, dP, dW, dS, dD, d_vsh, d_vd, d_vqq, d_grad, d_dedcn);
      }
      if (opt.reproducible) {
        hipLaunchKernelGGL(k_fix_to_double, dim3((unsigned)((3 * (int64_t)N + 255) / 256)), dim3(256), 0, sm,
                           d_grad, 3 * (int64_t)N);
        hipLaunchKernelGGL(k_fix_to_double, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, sm, d_dedcn,
                           (int64_t)N);
      }
      HIP_TRY(probe_end(kKForcePairs, 0.0, 24.0 * v2));  // reads P, W, S
    }
  }
  HIP_TRY(probe_begin(kKForceAtoms));
  hipLaunchKernelGGL(k_force_atoms, dim3(B), dim3(512), 0, sm, t, d_order, d_valid, mx.x, d_qat,
                     d_cnd4, d_mrad, d_dmrad, d_gam, d_w4, d_dw4, d_dedcn, d_dedcn4, d_grad,
                     d_erep, d_pcgrad, d_dedq);
  HIP_TRY(probe_end(kKForceAtoms, 0.0, 0.0));
  HIP_TRY(probe_begin(kKForceAtm));
  if (opt.reproducible) {
    if (n_xl > 0)
      hipLaunchKernelGGL((k_force_atm<true, true>), dim3(n_xl), dim3(512), 0, sm, t, d_order, d_valid, d_cnd4,
                         d_w4, d_dw4, d_c6pair, d_dedcn4, d_grad, d_eatm, d_xl);
    if (n_xl < B)
      hipLaunchKernelGGL((k_force_atm<true, false>), dim3(B - n_xl), dim3(512), 0, sm, t, d_order + n_xl,
                         d_valid, d_cnd4, d_w4, d_dw4, d_c6pair, d_dedcn4, d_grad, d_eatm, d_xl);
  } else {
    if (n_xl > 0)
      hipLaunchKernelGGL((k_force_atm<false, true>), dim3(n_xl), dim3(512), 0, sm, t, d_order, d_valid, d_cnd4,
                         d_w4, d_dw4, d_c6pair, d_dedcn4, d_grad, d_eatm, d_xl);
    if (n_xl < B)
      hipLaunchKernelGGL((k_force_atm<false, false>), dim3(B - n_xl), dim3(512), 0, sm, t, d_order + n_xl,
                         d_valid, d_cnd4, d_w4, d_dw4, d_c6pair, d_dedcn4, d_grad, d_eatm, d_xl);
  }
  HIP_TRY(probe_end(kKForceAtm, 0.0, 0.0));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(dev->ev[3], sm));

  // ---- publication ----
  std::vector<SysState> hst(B);
  std::vector<double> herep(B), heatm(B), hx(NV), hqat(N), hgrad(3 * (size_t)N);
  HIP_TRY(hipMemcpyAsync(hst.data(), d_st, sizeof(SysState) * B, hipMemcpyDeviceToHost, sm));
  HIP_TRY(hipMemcpyAsync(herep.data(), d_erep, sizeof(double) * B, hipMemcpyDeviceToHost, sm));
  HIP_TRY(hipMemcpyAsync(heatm.data(), d_eatm, sizeof(double) * B, hipMemcpyDeviceToHost, sm));
  HIP_TRY(hipMemcpyAsync(hx.data(), mx.x, sizeof(double) * NV, hipMemcpyDeviceToHost, sm));
  HIP_TRY(hipMemcpyAsync(hqat.data(), d_qat, sizeof(double) * N, hipMemcpyDeviceToHost, sm));
  HIP_TRY(hipMemcpyAsync(hgrad.data(), d_grad, sizeof(double) * 3 * (size_t)N, hipMemcpyDeviceToHost, sm));
  std::vector<double> hpcg, hdq;
  if (NPC > 0) {
    hpcg.resize(3 * (size_t)NPC);
    hdq.resize(NPC);
    HIP_TRY(hipMemcpyAsync(hpcg.data(), d_pcgrad, sizeof(double) * 3 * (size_t)NPC, hipMemcpyDeviceToHost, sm));
    HIP_TRY(hipMemcpyAsync(hdq.data(), d_dedq, sizeof(double) * NPC, hipMemcpyDeviceToHost, sm));
  }
  if (out.dbg_overlap)
    HIP_TRY(hipMemcpyAsync(out.dbg_overlap, dS, sizeof(double) * NN, hipMemcpyDeviceToHost, sm));
  if (out.dbg_h0)
    HIP_TRY(hipMemcpyAsync(out.dbg_h0, dH0, sizeof(double) * NN, hipMemcpyDeviceToHost, sm));
  if (out.dbg_density)
    HIP_TRY(hipMemcpyAsync(out.dbg_density, dP, sizeof(double) * NN, hipMemcpyDeviceToHost, sm));
  if (out.dbg_cn)
    HIP_TRY(hipMemcpyAsync(out.dbg_cn, d_cn, sizeof(double) * N, hipMemcpyDeviceToHost, sm));
  if (out.dbg_levels)
    HIP_TRY(hipMemcpyAsync(out.dbg_levels, d_eps, sizeof(double) * NAO, hipMemcpyDeviceToHost, sm));
  if (out.dbg_occ)
    HIP_TRY(hipMemcpyAsync(out.dbg_occ, d_occ, sizeof(double) * NAO, hipMemcpyDeviceToHost, sm));
  HIP_TRY(hipStreamSynchronize(sm));
  if (out.scc_state) std::copy(hx.begin(), hx.end(), out.scc_state);
  if (timing) {
    HIP_TRY(hipEventElapsedTime(&timing->setup_ms, dev->ev[0], dev->ev[1]));
    HIP_TRY(hipEventElapsedTime(&timing->scc_ms, dev->ev[1], dev->ev[2]));
    HIP_TRY(hipEventElapsedTime(&timing->force_ms, dev->ev[2], dev->ev[3]));
    HIP_TRY(hipEventElapsedTime(&timing->total_ms, dev->ev[0], dev->ev[3]));
    timing->scc_rounds = rounds;
  }
  if (ks) {
    for (int k = 0; k < kKernelCount; ++k) ks[k] = KernelStat{0, 0, 0.0, 0.0, 0.0};
    for (size_t i = 0; i < probes.size(); ++i) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, dev->probe_pool[2 * i], dev->probe_pool[2 * i + 1]));
      KernelStat& k = ks[probes[i].kind];
      k.launches += 1;
      k.total_ms += ms;
      k.flops += probes[i].flops;
      k.bytes += probes[i].bytes;
    }
  }
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (int s = 0; s < B; ++s) {
    const SysState& z = hst[s];
    int status = z.status;
    if (status == 0 && !z.converged) status = 7;
    const bool ok = status == 0;
    const int a0 = T.atom_off[s], a1 = T.atom_off[s + 1];
    const int ns = T.shell_off[s + 1] - T.shell_off[s];
    const double* xs = hx.data() + T.vec_off[s];
    if (out.status) out.status[s] = status;
    if (out.iterations) out.iterations[s] = z.iterations;
    if (out.converged) out.converged[s] = ok ? 1 : 0;
    if (out.energies) out.energies[s] = ok ? z.free_energy + herep[s] + heatm[s] : nan;
    for (int a = a0; a < a1; ++a) {
      if (out.charges) out.charges[a] = ok ? hqat[a] : nan;
      if (out.gradient)
        for (int c = 0; c < 3; ++c) out.gradient[3 * a + c] = ok ? hgrad[3 * a + c] : nan;
    }
    if (out.dipoles) {
      for (int c = 0; c < 3; ++c) {
        double m = 0.0;
        for (int a = a0; a < a1; ++a) m += T.xyz[3 * a + c] * hqat[a] + xs[ns + 3 * (a - a0) + c];
        out.dipoles[3 * s + c] = ok ? m : nan;
      }
    }
    for (int p = T.pc_off[s]; p < T.pc_off[s + 1]; ++p) {
      if (out.pc_gradient)
        for (int c = 0; c < 3; ++c) out.pc_gradient[3 * p + c] = ok ? hpcg[3 * p + c] : nan;
      if (out.pc_dedq) out.pc_dedq[p] = ok ? hdq[p] : nan;
    }
    if (out.dbg_components) {
      double* cp = out.dbg_components + 8 * s;
      cp[0] = z.e_core; cp[1] = z.e_es2; cp[2] = z.e_es3; cp[3] = z.e_aes2;
      cp[4] = z.e_d4; cp[5] = z.e_pc; cp[6] = herep[s]; cp[7] = heatm[s];
    }
  }
  error.clear();
  return 0;
}

// Unrestricted members (spin channel count 2) run as an alpha system in
// their own place and a beta copy appended after all members: the copy has
// the member's atoms and geometry, no point charges, and zero field, b and A
// (its potentials are the alpha system's, set by the spin kernels).  The
// outputs of the members are the leading slices of the expanded run's.
int run_batch(Device* dev, const HostBatch& hb, const RunOptions& opt, HostOutputs& out,
              RunTiming* timing, std::string& error, TopoCache* cache) {
  int64_t U = 0;
  if (hb.spin)
    for (int64_t s = 0; s < hb.batch; ++s) U += hb.spin[s] == 2;
  if (U == 0) return run_batch_core(dev, hb, opt, out, timing, error, cache, nullptr);
  const int64_t B = hb.batch, N = hb.atoms, B2 = B + U;
  std::vector<int64_t> aoff(hb.atom_offsets, hb.atom_offsets + B + 1), poff;
  std::vector<int32_t> Z(hb.Z, hb.Z + N), unpaired(hb.unpaired, hb.unpaired + B);
  std::vector<double> xyz(hb.xyz, hb.xyz + 3 * N), chg(hb.charges, hb.charges + B), field, pshift, presp;
  std::vector<int> partner((size_t)B2, -1);
  if (hb.point_offsets) poff.assign(hb.point_offsets, hb.point_offsets + B + 1);
  if (hb.field) field.assign(hb.field, hb.field + 3 * B);
  if (hb.pshift) pshift.assign(hb.pshift, hb.pshift + N);
  if (hb.presp) {
    int64_t nresp = 0;
    for (int64_t s = 0; s < B; ++s) nresp += (aoff[s + 1] - aoff[s]) * (aoff[s + 1] - aoff[s]);
    presp.assign(hb.presp, hb.presp + nresp);
  }
  std::vector<int32_t> state_dim((size_t)B);
  if (batch_matrix_elements(hb, nullptr, error, nullptr, state_dim.data()) < 0) return 1;
  int64_t nstate = 0, beta_state = 0;
  for (int64_t s = 0; s < B; ++s) {
    nstate += state_dim[s];
    if (hb.spin[s] != 2) continue;
    const int64_t a0 = hb.atom_offsets[s], a1 = hb.atom_offsets[s + 1], n = a1 - a0;
    const int b = (int)aoff.size() - 1;
    partner[s] = b;
    partner[b] = (int)s;
    Z.insert(Z.end(), hb.Z + a0, hb.Z + a1);
    xyz.insert(xyz.end(), hb.xyz + 3 * a0, hb.xyz + 3 * a1);
    aoff.push_back(aoff.back() + n);
    chg.push_back(hb.charges[s]);
    unpaired.push_back(hb.unpaired[s]);
    if (!poff.empty()) poff.push_back(poff.back());
    if (hb.field) field.insert(field.end(), 3, 0.0);
    if (hb.pshift) pshift.insert(pshift.end(), (size_t)n, 0.0);
    if (hb.presp) presp.insert(presp.end(), (size_t)(n * n), 0.0);
    beta_state += state_dim[s] / 2;
  }
  HostBatch e = hb;
  e.batch = B2;
  e.atoms = (int64_t)Z.size();
  e.atom_offsets = aoff.data();
  e.Z = Z.data();
  e.xyz = xyz.data();
  e.charges = chg.data();
  e.unpaired = unpaired.data();
  e.point_offsets = poff.empty() ? nullptr : poff.data();
  e.field = hb.field ? field.data() : nullptr;
  e.pshift = hb.pshift ? pshift.data() : nullptr;
  e.presp = hb.presp ? presp.data() : nullptr;
  e.spin = nullptr;
  const int64_t N2 = e.atoms;
  // a WARM state covers the members; the copies' slots are scratch
  RunOptions eopt = opt;
  std::vector<double> warm;
  if (opt.warm_state) {
    warm.assign(opt.warm_state, opt.warm_state + nstate);
    warm.resize((size_t)(nstate + beta_state), 0.0);
    eopt.warm_state = warm.data();
  }
  // expanded outputs; point-charge outputs are shared (the copies have none)
  std::vector<int32_t> nao2((size_t)B2);
  int64_t nn2 = 0, nao_total = 0;
  {
    std::string err;
    nn2 = batch_matrix_elements(e, nullptr, err, nao2.data(), nullptr);
    if (nn2 < 0) { error = err; return 1; }
    for (int32_t v : nao2) nao_total += v;
  }
  std::vector<double> en, grad, q, dip, ov, h0, den, cn, comp, lev, occ, state;
  std::vector<int32_t> iters, status;
  std::vector<uint8_t> conv;
  HostOutputs eo = out;
  auto own = [](auto& vec, const void* want, int64_t count) {
    if (want == nullptr) return (decltype(vec.data()))nullptr;
    vec.assign((size_t)count, {});
    return vec.data();
  };
  eo.energies = own(en, out.energies, B2);
  eo.gradient = own(grad, out.gradient, 3 * N2);
  eo.charges = own(q, out.charges, N2);
  eo.dipoles = own(dip, out.dipoles, 3 * B2);
  eo.iterations = own(iters, out.iterations, B2);
  eo.converged = own(conv, out.converged, B2);
  eo.status = own(status, out.status, B2);
  eo.dbg_overlap = own(ov, out.dbg_overlap, nn2);
  eo.dbg_h0 = own(h0, out.dbg_h0, nn2);
  eo.dbg_density = own(den, out.dbg_density, nn2);
  eo.dbg_cn = own(cn, out.dbg_cn, N2);
  eo.dbg_components = own(comp, out.dbg_components, 8 * B2);
  eo.dbg_levels = own(lev, out.dbg_levels, nao_total);
  eo.dbg_occ = own(occ, out.dbg_occ, nao_total);
  eo.scc_state = own(state, out.scc_state, nstate + beta_state);
  const int rc = run_batch_core(dev, e, eopt, eo, timing, error, cache, partner.data());
  if (rc != 0) return rc;
  // the members' slices lead every packed output
  int64_t nn = 0, nao_members = 0;
  for (int64_t s = 0; s < B; ++s) {
    nn += (int64_t)nao2[s] * nao2[s];
    nao_members += nao2[s];
  }
  auto give = [](const auto& vec, auto* dst, int64_t count) {
    if (dst) std::copy(vec.begin(), vec.begin() + count, dst);
  };
  give(en, out.energies, B);
  give(grad, out.gradient, 3 * N);
  give(q, out.charges, N);
  give(dip, out.dipoles, 3 * B);
  give(iters, out.iterations, B);
  give(conv, out.converged, B);
  give(status, out.status, B);
  give(ov, out.dbg_overlap, nn);
  give(h0, out.dbg_h0, nn);
  give(den, out.dbg_density, nn);
  give(cn, out.dbg_cn, N);
  give(comp, out.dbg_components, 8 * B);
  give(lev, out.dbg_levels, nao_members);
  give(occ, out.dbg_occ, nao_members);
  give(state, out.scc_state, nstate);
  return 0;
}

}  // namespace xtb
