// qbh_qball.h -- C++ host adaptor presenting the qbh C ABI (include/qbh.h) behind
// classes shaped like qball's SlaterDet / FourierTransform / NonLocalPotential /
// EnergyFunctional / ChargeDensity / ExponentialWavefunctionStepper interfaces.
//
// Header-only; needs the HIP runtime API for host<->device copies and libqbh.so.
// Wavefunction blocks live on the host in the reference's ComplexMatrix layout
// (column-major, mloc rows, element n*mloc + ig; src/math/matrix.h:374-380) and are
// mirrored on the device; every numerical operation runs in libqbh.so.
//
// Reference interfaces mirrored (qball src/qball/):
//   FourierTransform::backward/forward           FourierTransform.h:144-153
//   SlaterDet::rs_mul_add / compute_density / gram  SlaterDet.h:113-126
//   NonLocalPotential::energy (Enl, Hψ term, fion)  NonLocalPotential.h:93-96
//   EnergyFunctional::energy(psi, compute_hpsi, dwf, ...), v_r   EnergyFunctional.h:104-113
//   ChargeDensity::update_density                 ChargeDensity.h:92
//   ExponentialWavefunctionStepper::preupdate/update  ExponentialWavefunctionStepper.h
#pragma once
#include <hip/hip_runtime_api.h>

#include <complex>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qbh.h"

namespace qbh {
namespace qball {

using cd = std::complex<double>;

inline void check(int rc, qbh_ctx ctx, const char* what) {
  if (rc != QBH_OK)
    throw std::runtime_error(std::string(what) + " failed: " + (ctx ? qbh_last_error(ctx) : ""));
}

// Device buffer (RAII)
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { resize(bytes); }
  ~DevBuf() { if (p_) (void)hipFree(p_); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  void resize(size_t bytes) {
    if (bytes <= n_) return;
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    if (hipMalloc(&p_, bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    n_ = bytes;
  }
  template <typename T> T* as() const { return static_cast<T*>(p_); }
  void upload(const void* h, size_t bytes) {
    resize(bytes);
    if (hipMemcpy(p_, h, bytes, hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("upload failed");
  }
  void download(void* h, size_t bytes) const {
    if (hipMemcpy(h, p_, bytes, hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("download failed");
  }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Context: one device, one stream
class Context {
 public:
  explicit Context(int device = 0) { check(qbh_ctx_create(device, nullptr, &h_), nullptr, "qbh_ctx_create"); }
  ~Context() { if (h_) qbh_ctx_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  qbh_ctx get() const { return h_; }
  void sync() const { check(qbh_sync(h_), h_, "qbh_sync"); }

 private:
  qbh_ctx h_ = nullptr;
};

// Unit cell (UnitCell::set, UnitCell.cc:36-90): rows of a are a0, a1, a2
struct UnitCell {
  double a[9];
  double b[9];
  double volume;
  explicit UnitCell(const double* arows) {
    std::memcpy(a, arows, sizeof(a));
    const double *a0 = a, *a1 = a + 3, *a2 = a + 6;
    auto cross = [](const double* x, const double* y, double* o) {
      o[0] = x[1] * y[2] - x[2] * y[1];
      o[1] = x[2] * y[0] - x[0] * y[2];
      o[2] = x[0] * y[1] - x[1] * y[0];
    };
    double c[3];
    cross(a1, a2, c);
    volume = a0[0] * c[0] + a0[1] * c[1] + a0[2] * c[2];
    const double f = 1.0 / volume;
    double fa0[3], fa1[3], fa2[3];
    for (int i = 0; i < 3; ++i) { fa0[i] = f * a0[i]; fa1[i] = f * a1[i]; fa2[i] = f * a2[i]; }
    cross(fa1, a2, b);
    cross(fa2, a0, b + 3);
    cross(fa0, a1, b + 6);
    for (int i = 0; i < 9; ++i) b[i] = 2.0 * M_PI * b[i];
  }
};

// Norm-conserving KB species from a species XML file (Species::initialize, Species.cc:74-673);
// rcps <= 0 takes the width AtomSet::set_rcps picks for the cell and cutoff (AtomSet.cc:1102)
class Species {
 public:
  Species(const std::string& path, double rcps) {
    if (qbh_species_create_xml(path.c_str(), rcps, &h_) != QBH_OK)
      throw std::runtime_error("qbh_species_create_xml: " + path);
    info();
  }
  Species(const std::string& path, const UnitCell& cell, double ecut) {
    double z = 0.0;
    qbh_species probe = nullptr;
    if (qbh_species_create_xml(path.c_str(), 1.0, &probe) != QBH_OK)
      throw std::runtime_error("qbh_species_create_xml: " + path);
    qbh_species_info(probe, &z, nullptr, nullptr, nullptr, nullptr, nullptr);
    qbh_species_destroy(probe);
    double rc = 0.0;
    if (qbh_species_auto_rcps(z, cell.a, ecut, &rc) != QBH_OK ||
        qbh_species_create_xml(path.c_str(), rc, &h_) != QBH_OK)
      throw std::runtime_error("species set-up failed: " + path);
    info();
  }
  ~Species() { if (h_) qbh_species_destroy(h_); }
  Species(const Species&) = delete;
  Species& operator=(const Species&) = delete;
  qbh_species get() const { return h_; }
  double zval = 0.0, rcps = 0.0, eself = 0.0;
  int nproj = 0;

 private:
  void info() { qbh_species_info(h_, &zval, &rcps, &eself, &nproj, nullptr, nullptr); }
  qbh_species h_ = nullptr;
};

// Plane-wave basis: rods in the caller's order (Basis::rod_h/k/lmin/size); when built
// from a cell and cutoff it restates Basis::resize through qbh_rods_build.
class Basis {
 public:
  Basis(const UnitCell& cell, double ecut, bool real)
      : cell_(cell), real_(real) {
    int n = 0, np[3];
    const double k0[3] = {0, 0, 0};
    check(qbh_rods_build(cell.a, ecut, real, k0, 0, &n, nullptr, nullptr, nullptr, nullptr, np),
          nullptr, "qbh_rods_build");
    h_.resize(n); k_.resize(n); lmin_.resize(n); size_.resize(n);
    check(qbh_rods_build(cell.a, ecut, real, k0, n, &n, h_.data(), k_.data(), lmin_.data(),
                         size_.data(), np), nullptr, "qbh_rods_build");
    for (int i = 0; i < 3; ++i) np_[i] = np[i];
    size = 0;
    for (int s : size_) size += s;
  }
  const UnitCell& cell() const { return cell_; }
  bool real() const { return real_; }
  int np(int i) const { return np_[i]; }
  int localsize() const { return size; }
  const std::vector<int>& rod_h() const { return h_; }
  const std::vector<int>& rod_k() const { return k_; }
  const std::vector<int>& rod_lmin() const { return lmin_; }
  const std::vector<int>& rod_size() const { return size_; }
  int size = 0;

 private:
  UnitCell cell_;
  bool real_;
  std::vector<int> h_, k_, lmin_, size_;
  int np_[3];
};

// ChargeDensity grid rule (ChargeDensity.cc:76-99): density basis at 4*ecut, np(i)+2
// bumped by 2 until factorizable
inline void density_grid(const UnitCell& cell, double ecut, int np[3]) {
  Basis vb(cell, 4.0 * ecut, true);
  for (int i = 0; i < 3; ++i) {
    int n = vb.np(i) + 2;
    while (!qbh_factorizable(n)) n += 2;
    np[i] = n;
  }
}

// FourierTransform(basis, np0, np1, np2) on host arrays
class FourierTransform {
 public:
  FourierTransform(Context& ctx, const Basis& basis, int np0, int np1, int np2)
      : ctx_(ctx), basis_(basis) {
    np_[0] = np0; np_[1] = np1; np_[2] = np2;
    const double k0[3] = {0, 0, 0};
    check(qbh_basis_create(ctx.get(), basis.real(), (int)basis.rod_h().size(),
                           basis.rod_h().data(), basis.rod_k().data(), basis.rod_lmin().data(),
                           basis.rod_size().data(), np0, np1, np2, basis.cell().b, k0, &h_),
          ctx.get(), "qbh_basis_create");
  }
  ~FourierTransform() { if (h_) qbh_basis_destroy(h_); }
  FourierTransform(const FourierTransform&) = delete;
  FourierTransform& operator=(const FourierTransform&) = delete;
  int np0() const { return np_[0]; }
  int np1() const { return np_[1]; }
  int np2() const { return np_[2]; }
  int np012() const { return np_[0] * np_[1] * np_[2]; }
  qbh_basis handle() const { return h_; }
  Context& context() const { return ctx_; }
  const Basis& basis() const { return basis_; }

  // backward(c, f): f = sum_G c(G) exp(+iG.r) (Γ: conj(c) at -G)
  void backward(const cd* c, cd* f) { bwd(c, nullptr, f); }
  // backward(c1, c2, f): f = psi1 + i psi2 (Γ)
  void backward(const cd* c1, const cd* c2, cd* f) { bwd(c1, c2, f); }
  // forward(f, c): c = zvec(+G) of exp(-iG.r)/N transform; f is not modified here
  void forward(cd* f, cd* c) { fwd(f, c, nullptr); }
  void forward(cd* f, cd* c1, cd* c2) { fwd(f, c1, c2); }

 private:
  void bwd(const cd* c1, const cd* c2, cd* f) {
    const int ngw = basis_.localsize();
    const int nb = c2 ? 2 : 1;
    std::vector<cd> tmp((size_t)nb * ngw);
    std::memcpy(tmp.data(), c1, sizeof(cd) * ngw);
    if (c2) std::memcpy(tmp.data() + ngw, c2, sizeof(cd) * ngw);
    dc_.upload(tmp.data(), sizeof(cd) * tmp.size());
    df_.resize(sizeof(cd) * np012());
    check(qbh_fft_backward(h_, dc_.as<double>(), ngw, nb, c2 ? 1 : 0, df_.as<double>(), 0),
          ctx_.get(), "qbh_fft_backward");
    ctx_.sync();
    df_.download(f, sizeof(cd) * np012());
  }
  void fwd(const cd* f, cd* c1, cd* c2) {
    const int ngw = basis_.localsize();
    const int nb = c2 ? 2 : 1;
    df_.upload(f, sizeof(cd) * np012());
    dc_.resize(sizeof(cd) * nb * ngw);
    check(qbh_fft_forward(h_, df_.as<double>(), nb, c2 ? 1 : 0, dc_.as<double>(), ngw, 0),
          ctx_.get(), "qbh_fft_forward");
    ctx_.sync();
    std::vector<cd> tmp((size_t)nb * ngw);
    dc_.download(tmp.data(), sizeof(cd) * tmp.size());
    std::memcpy(c1, tmp.data(), sizeof(cd) * ngw);
    if (c2) std::memcpy(c2, tmp.data() + ngw, sizeof(cd) * ngw);
  }
  Context& ctx_;
  const Basis& basis_;
  int np_[3];
  qbh_basis h_ = nullptr;
  DevBuf dc_, df_;
};

// Slater determinant: host coefficients c (mloc x nst, column-major) + device mirror
class SlaterDet {
 public:
  SlaterDet(const Basis& basis, int nst) : basis_(basis), nst_(nst), occ_(nst, 0.0) {
    mloc_ = basis.localsize();
    c_.assign((size_t)mloc_ * nst, cd(0.0, 0.0));
  }
  const Basis& basis() const { return basis_; }
  int nst() const { return nst_; }
  int mloc() const { return mloc_; }
  std::vector<cd>& c() { return c_; }
  const std::vector<cd>& c() const { return c_; }
  std::vector<double>& occ() { return occ_; }
  const std::vector<double>& occ() const { return occ_; }
  // device mirror of c (uploaded on demand)
  const double* dev_c() const {
    dev_.upload(c_.data(), sizeof(cd) * c_.size());
    return dev_.as<double>();
  }
  double* dev_buffer() const {
    dev_.resize(sizeof(cd) * c_.size());
    return dev_.as<double>();
  }
  void pull() { dev_.download(c_.data(), sizeof(cd) * c_.size()); }

  // SlaterDet::rs_mul_add (SlaterDet.cc:971-1040): sdp += FFT^-1[v FFT c]
  void rs_mul_add(FourierTransform& ft, const double* v, SlaterDet& sdp) const;
  // SlaterDet::compute_density (SlaterDet.cc:839-932): rho += weight/Omega occ |psi|^2
  void compute_density(FourierTransform& ft, double weight, double* rho) const {
    const double* dc = dev_c();
    DevBuf drho(sizeof(double) * ft.np012());
    drho.upload(rho, sizeof(double) * ft.np012());
    check(qbh_density(ft.handle(), dc, mloc_, nst_, occ_.data(), weight,
                      basis_.cell().volume, drho.as<double>(), 0),
          ft.context().get(), "qbh_density");
    ft.context().sync();
    drho.download(rho, sizeof(double) * ft.np012());
  }
  // SlaterDet::gram (SlaterDet.cc:1043-1143)
  void gram(FourierTransform& ft) {
    double* d = const_cast<double*>(dev_c());
    check(qbh_gram(ft.handle(), d, mloc_, nst_), ft.context().get(), "qbh_gram");
    ft.context().sync();
    pull();
  }

 private:
  const Basis& basis_;
  int nst_;
  int mloc_;
  std::vector<cd> c_;
  std::vector<double> occ_;
  mutable DevBuf dev_;
};

// Hamiltonian pieces: v_r, KB projectors, vector potential; energy(...) applies Hψ
class EnergyFunctional {
 public:
  EnergyFunctional(FourierTransform& ft) : ft_(ft) {
    check(qbh_ham_create(ft.context().get(), ft.handle(), &h_), ft.context().get(),
          "qbh_ham_create");
    v_r.assign(ft.np012(), 0.0);
  }
  ~EnergyFunctional() { if (h_) qbh_ham_destroy(h_); }
  EnergyFunctional(const EnergyFunctional&) = delete;
  EnergyFunctional& operator=(const EnergyFunctional&) = delete;

  // EnergyFunctional::v_r (EnergyFunctional.h:104); push after changing it
  std::vector<double> v_r;
  void update_vr() {
    dv_.upload(v_r.data(), sizeof(double) * v_r.size());
    check(qbh_ham_set_vloc(h_, dv_.as<double>()), ft_.context().get(), "qbh_ham_set_vloc");
  }
  // NonLocalPotential::init/update_twnl tables, concatenated over species
  void set_projectors(const std::vector<int>& na, const std::vector<int>& npr,
                      const std::vector<int>& lproj, const std::vector<double>& wt,
                      const std::vector<double>& twnl, const std::vector<double>& tau) {
    check(qbh_ham_set_projectors(h_, (int)na.size(), na.data(), npr.data(), lproj.data(),
                                 wt.data(), twnl.data(), tau.data(),
                                 ft_.basis().cell().volume),
          ft_.context().get(), "qbh_ham_set_projectors");
    na_ = na;
  }
  // the same tables made on the device from species (update_twnl), re-tabled at |k+G-A|
  // by set_vector_potential; tau host, 3 per atom, species after species
  void set_species(const std::vector<const Species*>& sps, const std::vector<int>& na,
                   const std::vector<double>& tau) {
    std::vector<qbh_species> h;
    for (const Species* s : sps) h.push_back(s->get());
    check(qbh_ham_set_species(h_, (int)h.size(), h.data(), na.data(), tau.data(),
                              ft_.basis().cell().volume),
          ft_.context().get(), "qbh_ham_set_species");
    na_ = na;
  }
  // non-local ionic forces, NonLocalPotential::energy(sd, ..., compute_forces=true, fion)
  // (NonLocalPotential.cc:2175-2323): fion[is][3*ia+j] overwritten; returns Enl
  double nl_forces(const SlaterDet& psi, std::vector<std::vector<double>>& fion,
                   int flags = 0) {
    size_t natoms = 0;
    for (int n : na_) natoms += (size_t)n;
    std::vector<double> f(3 * natoms);
    double enl = 0.0;
    check(qbh_nl_forces(h_, psi.dev_c(), psi.mloc(), psi.nst(), psi.occ().data(), f.data(),
                        flags, &enl),
          ft_.context().get(), "qbh_nl_forces");
    fion.resize(na_.size());
    size_t a0 = 0;
    for (size_t is = 0; is < na_.size(); ++is) {
      fion[is].assign(f.begin() + 3 * a0, f.begin() + 3 * (a0 + na_[is]));
      a0 += (size_t)na_[is];
    }
    return enl;
  }
  void set_vector_potential(const double A[3]) {
    check(qbh_ham_set_vector_potential(h_, A), ft_.context().get(), "set_vector_potential");
  }
  // energy(psi, compute_hpsi=true, dwf): dwf = Hψ (NL, T, V_loc); returns Enl
  double energy(const SlaterDet& psi, bool compute_hpsi, SlaterDet& dwf, int flags = QBH_ALL) {
    if (!compute_hpsi) return enl_only(psi);
    double enl = 0.0;
    const double* dc = psi.dev_c();
    double* dh = dwf.dev_buffer();
    check(qbh_hpsi(h_, dc, psi.mloc(), psi.nst(), psi.occ().data(), dh, dwf.mloc(), flags, &enl),
          ft_.context().get(), "qbh_hpsi");
    ft_.context().sync();
    dwf.pull();
    return enl;
  }
  qbh_ham handle() const { return h_; }
  FourierTransform& ft() const { return ft_; }

 private:
  double enl_only(const SlaterDet& psi) {
    DevBuf tmp(sizeof(cd) * psi.c().size());
    double enl = 0.0;
    check(qbh_hpsi(h_, psi.dev_c(), psi.mloc(), psi.nst(), psi.occ().data(), tmp.as<double>(),
                   psi.mloc(), QBH_NONLOCAL, &enl),
          ft_.context().get(), "qbh_hpsi");
    return enl;
  }

  FourierTransform& ft_;
  qbh_ham h_ = nullptr;
  DevBuf dv_;
  std::vector<int> na_;
};

inline void SlaterDet::rs_mul_add(FourierTransform& ft, const double* v, SlaterDet& sdp) const {
  // a local-only Hamiltonian with v, accumulating into sdp
  qbh_ham h = nullptr;
  check(qbh_ham_create(ft.context().get(), ft.handle(), &h), ft.context().get(), "qbh_ham_create");
  DevBuf dv;
  dv.upload(v, sizeof(double) * ft.np012());
  check(qbh_ham_set_vloc(h, dv.as<double>()), ft.context().get(), "qbh_ham_set_vloc");
  const double* dc = dev_c();
  const double* dp = sdp.dev_c();
  check(qbh_hpsi(h, dc, mloc_, nst_, nullptr, const_cast<double*>(dp), sdp.mloc(),
                 QBH_LOCAL | QBH_ACCUMULATE, nullptr),
        ft.context().get(), "qbh_hpsi");
  ft.context().sync();
  sdp.pull();
  qbh_ham_destroy(h);
}

// ChargeDensity::update_density (ChargeDensity.cc:276-556) for one spin / k-point:
// rho = band sum, nelectrons = Omega/N sum rho, rhog = vft.forward(Omega rho)
struct ChargeDensity {
  FourierTransform& ft;    // wavefunction basis on the density grid
  FourierTransform& vft;   // density basis on the same grid
  std::vector<double> rhor;
  std::vector<cd> rhog;
  double nelectrons = 0.0;
  ChargeDensity(FourierTransform& f, FourierTransform& v, int ngv) : ft(f), vft(v), rhog(ngv) {}
  void update_density(const SlaterDet& sd, double weight = 1.0) {
    const double omega = sd.basis().cell().volume;
    rhor.assign(ft.np012(), 0.0);
    sd.compute_density(ft, weight, rhor.data());
    double s = 0.0;
    for (double r : rhor) s += r;
    nelectrons = s * omega / ft.np012();
    DevBuf dr, dg(sizeof(cd) * rhog.size());
    dr.upload(rhor.data(), sizeof(double) * rhor.size());
    check(qbh_fft_forward_real(vft.handle(), dr.as<double>(), omega, dg.as<double>(), 0),
          vft.context().get(), "qbh_fft_forward_real");
    vft.context().sync();
    dg.download(rhog.data(), sizeof(cd) * rhog.size());
  }
};

// ExponentialWavefunctionStepper (ETRS, ExponentialWavefunctionStepper.cc:153-276):
// preupdate() propagates psi(t) -> psi(t+dt/2) (kept) -> psi~(t+dt) with H(t); the caller
// then rebuilds H(t+dt) from psi~(t+dt) (density, v_r) and calls update(), which
// propagates psi(t+dt/2) -> psi(t+dt) with H(t+dt).
class ExponentialWavefunctionStepper {
 public:
  ExponentialWavefunctionStepper(EnergyFunctional& ef, SlaterDet& wf, double dt)
      : ef_(ef), wf_(wf), dt_(dt) {}
  void preupdate() {
    exponential(wf_, 0.5 * dt_);
    half_ = wf_.c();
    exponential(wf_, 0.5 * dt_);
  }
  void update() {
    wf_.c() = half_;
    exponential(wf_, 0.5 * dt_);
  }

 private:
  void exponential(SlaterDet& sd, double dt) {
    double* d = const_cast<double*>(sd.dev_c());
    check(qbh_exponential(ef_.handle(), d, sd.mloc(), sd.nst(), dt, 4, QBH_ALL),
          ef_.ft().context().get(), "qbh_exponential");
    ef_.ft().context().sync();
    sd.pull();
  }
  EnergyFunctional& ef_;
  SlaterDet& wf_;
  double dt_;
  std::vector<cd> half_;
};

}  // namespace qball
}  // namespace qbh
