// Internal (non-ABI) declarations shared by the qbh HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "fft_device.h"

namespace qbh {

// ------------------------------------------------------------------------------------
// Device-side description of one wavefunction basis on one FFT grid.
//
// Layout of the transform pipeline (per "slot" = one Γ band pair or one complex band):
//   T1[slot][col][z]      z-pencils of the nonzero columns (col order: sorted by x-line,
//                         then y) -- written by the z pass, read by the y pass
//   T2[slot][z][lx][y]    y-transformed x-lines (lx = index of a distinct x position
//                         of the G sphere) -- written by the y pass, read by the x pass
// The real-space grid f[z][y][x] (x fastest, the reference's order) only exists in the
// x pass (LDS) unless a caller asks for it.
struct GridDev {
  int np0, np1, np2;
  int ngw;
  int nrods;
  int ncol;     // nonzero z-columns (Γ: 2*nrods-1, complex: nrods)
  int nlx;      // distinct x positions among the columns
  int real;     // Γ half-sphere storage
  int ngroups;  // rod groups for the z passes
  int xpos_max; // largest x position of a line with h >= 0 (lines cover [0, xpos_max])
  int xneg_min; // smallest x position of a line with h < 0 (lines cover [xneg_min, np0))
  // rods (caller's order)
  const int* rod_first;
  const int* rod_lmin;
  const int* rod_size;
  // z-pass work list: group g covers entries [grp_start[g], grp_start[g+1]); entry e
  // is one LDS line = one column: column index, rod, kind (0 = rod(0,0) carrying both
  // +G and -G, 1 = +G image, 2 = -G image, 3 = complex rod)
  const int* grp_start;
  const int* ent_col;
  const int* ent_rod;
  const int* ent_kind;
  // y pass: columns sorted by line; line lx covers columns [line_start[lx], line_start[lx+1])
  const int* col_y;
  const int* line_start;
  const int* line_x;
  // twiddles exp(+2 pi i m / n) per axis, double and float copies
  const cplx<double>* tw64[3];
  const cplx<float>* tw32[3];
  AxisPlan ax[3];
};

// Host-side owner of a basis description.
struct Basis {
  int device = 0;
  GridDev g{};
  std::vector<int> rod_h, rod_k, rod_lmin, rod_size;
  double b[9] = {0};      // reciprocal vectors (rows)
  double kpoint[3] = {0};
  std::vector<void*> allocations;
  // per-G tables (device): |k+G|^2 and (k+G)_{x,y,z} (3*ngw, component-major)
  double* kpg2 = nullptr;
  double* kpgx = nullptr;
  int max_lines_z = 0, max_lines_y = 0;
  // launch choices (qbh_ham_set_tuning / qbh_ham_autotune)
  struct Tune {
    int zy_reg = 1;  // register two-factor z/y transforms where instantiated (else Stockham)
    int x_reg = 1;   // register x pass where instantiated
    int x_spw = 0;   // slots per x-pass workgroup (0 = occupancy heuristic)
    int radix = 0;   // Stockham radix plan variant (0: 8s first, 1: 4s, 2: odd radices first)
  } tune;
};

// Stockham radix plan of length n for a plan variant (qbh_basis.hip)
bool make_axis_variant(int n, int variant, AxisPlan& p);

// Kernel timing registry (optional; enabled per context).  Each launch of a timed
// kernel is bracketed by two events on the context's stream; totals are folded in when
// the caller reads them (after a synchronisation).
struct KernelTimer {
  bool enabled = false;
  struct Rec {
    hipEvent_t a, b;
    std::string name;
  };
  std::vector<Rec> pending;
  std::map<std::string, std::pair<double, long>> totals;  // ms, count
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* blas = nullptr;  // rocblas_handle
  void* lt = nullptr;    // hipBLASLt handle, workspace and per-shape plans (qbh_nl.hip)
  std::string last_error;
  KernelTimer timer;
  // second stream (default priority) for the projector GEMMs, forked from and joined to
  // `stream` with events so they overlap the transform pipeline
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool overlap = true;  // run the projector GEMMs concurrently with the transforms
  // workspace (grown on demand, reused across calls)
  void* ws = nullptr;
  size_t ws_bytes = 0;
  size_t ws_cap = size_t(8) << 30;  // cap on pipeline workspace
  // autotuner results per problem shape (key -> knob values)
  std::map<std::string, std::vector<int>> tune_cache;
  // this device's allocation-free transform scratch (device address of the file-scope
  // array in qbh_api.hip), or null with the reason in scratch_err
  void* scratch = nullptr;
  std::string scratch_err;
};

// Looks up (once per device, with `device` current) the device address of the
// allocation-free scratch array and stores it in c->scratch (qbh_api.hip).
void noalloc_scratch_setup(Ctx* c);

// Records the named kernel's launch window when timing is enabled.
struct TimedLaunch {
  Ctx* ctx;
  KernelTimer::Rec rec;
  bool on;
  TimedLaunch(Ctx* c, const char* name);
  ~TimedLaunch();
};

void* ws_get(Ctx* ctx, size_t bytes);

// Set (per host thread) while the multi-cell Hψ graph is being captured: every workspace,
// table and handle must then already exist, so the growth paths (ws_get, grow), the
// occupation upload and the Enl read-back refuse to allocate or synchronise and fail with
// an error instead.
extern thread_local bool t_capturing;

// FFT pipeline entry points (qbh_fft.hip)
enum SlotMode { SLOT_PAIR = 0, SLOT_SINGLE_REAL = 1, SLOT_COMPLEX = 2 };
enum Prec { PREC_FP64 = 0, PREC_FP32 = 1 };

struct PipelineArgs {
  const double* psi;   // complex coefficients (interleaved), column-major, ld complex
  long ld;
  int nst;             // number of bands
  int band0;           // first band of the chunk
  int nslots;          // slots in this chunk
  int mode;            // SlotMode
  int prec;            // Prec
  void* t1;            // workspace
  void* t2;
  hipEvent_t wait_z = nullptr;  // the forward z pass (which writes the output) waits on it
  const float* add32 = nullptr; // FP32 term (2*ngw floats per band) the forward z pass adds
};

int pipeline_bwd_yz(Ctx* ctx, Basis* b, const PipelineArgs& a);
int pipeline_fwd_zy_accumulate(Ctx* ctx, Basis* b, const PipelineArgs& a, double* hpsi,
                               long ldh, const double* kin /*0.5|k+G|^2 or null*/,
                               int store /*0: hpsi += kinetic, FP32 term, value; 1: hpsi =
                                           value; 2: hpsi = kinetic + FP32 term + value*/);
int pipeline_x_vmul(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* v);
int pipeline_x_store(Ctx* ctx, Basis* b, const PipelineArgs& a, double* f /*complex grid*/);
int pipeline_x_load(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* f, int real_input,
                    double scale);
int pipeline_x_density(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* facs,
                       double* rho);
size_t slot_bytes(const Basis* b, int prec);
// S (M x N, ld M) = alpha * A^T B - a0 b0^T over the real view with K rows (real != 0, the
// Γ metric when alpha = 2) or A^H B (complex, K rows); K-split dgemm/zgemm (qbh_dense.hip)
int band_product(Ctx* c, void* blas, bool real, int M, int N, long K, const double* A, long lda,
                 const double* B, long ldb, double alpha, double* S, double*& scratch,
                 size_t& scratch_bytes);
// a two-factor register transform is instantiated for edge length n
bool register_edge(int n);

}  // namespace qbh
