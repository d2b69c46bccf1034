// Internal Hamiltonian state shared by qbh_api.hip and qbh_nl.hip.
#pragma once
#include <vector>

#include "../../include/qbh.h"
#include "qbh_internal.h"

namespace qbh {

struct Ham {
  Ctx* ctx = nullptr;
  Basis* b = nullptr;
  void* b_handle = nullptr;  // the qbh_basis handle b came from
  double* v = nullptr;    // v_r on the grid (owned copy)
  double* kin = nullptr;  // 0.5 |k+G-A|^2 per plane wave
  double A[3] = {0, 0, 0};
  // projectors
  int nsp = 0;
  int P = 0;                      // total projector columns sum_s npr[s]*na[s]
  std::vector<int> na, npr;
  double omega = 0.0;
  int* col_l = nullptr;           // per column: l
  int* col_tw = nullptr;          // per column: row of the twnl table
  int* col_atom = nullptr;        // per column: atom index into tau
  double* col_fac = nullptr;      // per column: wt/Omega
  std::vector<int> col_atom_host; // col_atom on the host (force accumulation)
  double* twnl = nullptr;         // concatenated tables, row = species projector
  // species the tables were made from (qbh_ham_set_species; empty when the caller passed
  // the tables): their rows are re-tabled at |k+G-A| whenever A changes
  std::vector<qbh_species> sps;
  double* tau = nullptr;          // 3 * total atoms
  void* anl64 = nullptr;          // Γ: 2ngw x P real; complex: ngw x P complex
  void* anl32 = nullptr;
  bool anl64_ok = false, anl32_ok = false;
  // complex bases, this library's matrix-core products: 2P rows of 2ngw reals, row 2p the
  // anl row of projector p, row 2p + 1 the row times i (built from anl when it changes)
  void* anlc64 = nullptr;
  void* anlc32 = nullptr;
  bool anlc64_ok = false, anlc32_ok = false;
  // scratch (grown on demand)
  void* f64 = nullptr;            // F (P x nst), real (Γ) or complex
  size_t f64_bytes = 0;
  void* f32 = nullptr;
  size_t f32_bytes = 0;
  void* psi32 = nullptr;          // FP32 copy of psi / FP32 product
  size_t psi32_bytes = 0;
  void* tmp32 = nullptr;
  size_t tmp32_bytes = 0;
  void* dpsi = nullptr;           // forces: i (k+G-A)_j psi (FP64)
  size_t dpsi_bytes = 0;
  void* df = nullptr;             // forces: dfnl partials
  size_t df_bytes = 0;
  void* fpart = nullptr;          // forces: per-column partial sums over band chunks
  size_t fpart_bytes = 0;
  void* fg = nullptr;             // forces: per-column sums, 3 x P
  size_t fg_bytes = 0;
  double* red = nullptr;          // reduction partials, then occupations
  size_t red_n = 0;
  std::vector<double> occ_host;   // occupations currently on the device
  std::vector<void*> expo;        // exponential buffers
  size_t expo_bytes = 0;
  // tuning (qbh_ham_set_tuning / qbh_ham_autotune): 0 / -1 = built-in choice
  int nsplit = 0;    // K slices of the projector contraction
  int overlap = -1;  // projector GEMMs on the side stream (1), after the transforms (0)
  // projector GEMM implementation: kOwnGemm = this library's matrix-core kernels (real
  // view of Γ, or the combined table of a complex basis; F over K slices); 0 = rocBLAS (F
  // over nsplit K slices); 1..16 = hipBLASLt with the n-th algorithm its heuristic lists
  // for the shape (F over the full K; n past the end of the list: the last one); shapes
  // hipBLASLt lists nothing for (complex) use rocBLAS; -1 = built-in choice (nl_builtin_algo)
  int fnl_algo = -1, enl_algo = -1;
  void* etrs = nullptr;           // ETRS step: psi(t+dt/2), psi~(t+dt), rho, rhog
  size_t etrs_bytes = 0;
};

// Hψ flags of the entry points that apply H as a whole (exponential, ETRS step, TD
// propagator, SCF loop, autotuner, batch): the terms named in flags, all three when none
// is named, and the precision bit
inline int whole_h_flags(int flags) {
  const int terms = flags & QBH_ALL;
  return (terms ? terms : QBH_ALL) | (flags & QBH_FP32);
}

// Hψ of qbh_hpsi without argument checks beyond leading dimensions (qbh_api.hip)
int hpsi_impl(Ham* h, const double* psi, long ld, int nst, const double* occ, double* hpsi,
              long ldh, int flags, double* enl);
// Taylor exponential of qbh_exponential; krylov (nullable) receives H^k in, k < order
int expo_impl(Ham* h, const double* in, long ldin, double* outp, long ldout, int nst, double dt,
              int order, int flags, double* krylov);
// defer_add32 (nullable, FP32 only): instead of adding the FP32 product to hpsi, return
// it (compact, 2*ngw floats per band) for the forward z pass to add.  overwrite (FP64):
// hpsi = beta F instead of hpsi += beta F (rows 0..2*ngw-1 of every band)
int nl_apply(Ham* h, const double* psi, long ld, int nst, const double* occ, double* hpsi,
             long ldh, int fp32, double* enl, const float** defer_add32 = nullptr,
             bool overwrite = false);
double nl_read_enl(Ham* h, hipStream_t s);
// non-local ionic forces of qbh_nl_forces into fion (host, 3 per atom, overwritten)
int nl_forces(Ham* h, const double* psi, long ld, int nst, const double* occ, int fp32,
              double* fion, double* enl);
// number of hipBLASLt algorithms listed for the fnl (which 0) or enl (which 1) GEMM of an
// Hψ call with these arguments (0: hipBLASLt does not cover it)
int nl_lt_candidates(Ham* h, int which, int nst, long ld, long ldh, int fp32);
// fnl_algo / enl_algo value of the matrix-core kernels of qbh_gemm.hip
constexpr int kOwnGemm = 17;
// the value the built-in choice (-1) of fnl_algo (which 0) / enl_algo (which 1) stands
// for: kOwnGemm, except rocBLAS (0) for the enl product of a complex basis
int nl_builtin_algo(Ham* h, int which);
int blas_handle(Ctx* c, void** out);
// frees a context's rocBLAS handle and hipBLASLt state (qbh_nl.hip)
void blas_destroy(Ctx* c);
void* grow(void*& p, size_t& have, size_t need);
// twnl rows of a species at |k+G-A| (grad 0: rows x ngw into out) or their gradient with
// respect to k+G-A (grad 1: component d at out + d * dstride), on c's stream (qbh_species.hip)
int species_table(qbh_species sp, Ctx* c, Basis* b, const double* A, int grad, double* out,
                  long dstride);
int species_nproj(qbh_species sp);
// combined table of complex projector rows (rows x ngw complex, ld ngw) for the matrix-core
// products: 2 rows x 2ngw reals, row 2p = row p, row 2p + 1 = i row p (qbh_nl.hip)
int anl_pair_c64(Ctx* c, int ngw, int rows, const double* anl, double* out);
// re-tables the Hamiltonian's twnl rows from its species at its vector potential
int retable_twnl(Ham* h);
// complex projector table (ngw x P, ld ngw) of the Hamiltonian's columns built from the rows
// `rows` (row r of ngw, in the layout of twnl) with phases exp(-i (k+G-A).tau), on the
// context's stream: anl_p = (-i)^l rows[col_tw[p]] exp(...) (qbh_nl.hip)
int build_anl_c64(Ham* h, const double* rows, const double* A, double* out);

}  // namespace qbh

// the opaque qbh_ham handle of the C ABI
struct qbh_ham_s {
  qbh::Ham h;
};
