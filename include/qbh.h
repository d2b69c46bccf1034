/*
 * qbh -- MI355X (gfx950) Hamiltonian-application engine for plane-wave DFT/TDDFT.
 *
 * C ABI.  Plain pointers and sizes only; every call returns an int status (0 = ok,
 * negative = error kind, message via qbh_last_error) and no exception crosses it.
 *
 * Memory: arguments named *_dev are device pointers (hipMalloc'd by the caller, or any
 * device allocation on the context's device); everything else is host memory that is
 * copied during the call.  Coefficient arrays are complex<double> stored as interleaved
 * (re, im) doubles, band-major with a leading dimension `ld` in complex elements: band n,
 * plane wave ig lives at element n*ld + ig -- the layout of qball's ComplexMatrix
 * SlaterDet::c() (src/math/matrix.h:374-380), whose local leading dimension is mloc.
 * Real-space grids are x-fastest, index x + np0*(y + np1*z) (FourierTransform.h:165-166).
 *
 * All work is asynchronous on the context's stream; qbh_sync() waits for it.
 */
#ifndef QBH_H
#define QBH_H

#ifdef __cplusplus
extern "C" {
#endif

#define QBH_OK 0
#define QBH_ERR_ARG (-1)
#define QBH_ERR_ALLOC (-2)
#define QBH_ERR_DEVICE (-3)
#define QBH_ERR_UNSUPPORTED (-4)
#define QBH_ERR_BLAS (-5)

/* qbh_hpsi flags */
#define QBH_NONLOCAL 0x1    /* NonLocalPotential::energy(compute_hpsi=true) term */
#define QBH_KINETIC 0x2     /* 0.5*|k+G-A|^2 c */
#define QBH_LOCAL 0x4       /* SlaterDet::rs_mul_add with v_r */
#define QBH_ALL 0x7
#define QBH_ACCUMULATE 0x8  /* add into hpsi instead of clearing it first */
#define QBH_FP32 0x10       /* FFTs and projector GEMMs in FP32 (FP32 accumulate) */

typedef struct qbh_ctx_s* qbh_ctx;
typedef struct qbh_basis_s* qbh_basis;
typedef struct qbh_ham_s* qbh_ham;
typedef struct qbh_td_s* qbh_td;
typedef struct qbh_batch_s* qbh_batch;
typedef struct qbh_psda_s* qbh_psda;
typedef struct qbh_mixer_s* qbh_mixer;
typedef struct qbh_species_s* qbh_species;

/* ---- context ------------------------------------------------------------------------ */
/* stream: a hipStream_t to run on, or NULL for a stream owned by the context. */
int qbh_ctx_create(int device, void* stream, qbh_ctx* out);
int qbh_ctx_destroy(qbh_ctx ctx);
const char* qbh_last_error(qbh_ctx ctx);
int qbh_sync(qbh_ctx ctx);
/* Cap (bytes) on the transform workspace the library keeps per context. */
int qbh_ctx_set_workspace_cap(qbh_ctx ctx, double bytes);
/* Per-kernel timing with HIP events on the context's stream (0 = off). */
int qbh_ctx_set_timing(qbh_ctx ctx, int on);
/* Run the projector GEMMs of qbh_hpsi on a second stream, concurrently with the local-
 * potential transforms (1, default), or after them on the context's stream (0).  Results
 * are identical either way. */
int qbh_ctx_set_overlap(qbh_ctx ctx, int on);
/* Copies up to `cap` "name total_ms count" lines (newline separated) into buf and
 * resets the accumulated totals.  Synchronises the stream. */
int qbh_ctx_timing_report(qbh_ctx ctx, char* buf, int cap);
/* Versions of the HIP runtime, rocBLAS and hipBLASLt this process has loaded, and which
 * library the built-in projector GEMM choice uses with them (no counterpart in the
 * reference; diagnostics). */
int qbh_ctx_library_info(qbh_ctx ctx, char* buf, int cap);
/* Stream bandwidth of the context's device (no counterpart in the reference; diagnostics):
 * grid-stride read-only, write-only and copy kernels over two temporary buffers of `bytes`
 * (1 MiB .. 64 GiB), each the best over 1024, 4096 and 16384 workgroups with 1, 4 or 8
 * independent 16-byte accesses in flight per thread and over `reps` timed launches;
 * gbps[0..2] = read, write and copy rates in GB/s, the copy rate counting read plus
 * written bytes.  The practical ceiling for the HBM rate of a streaming pass.
 * Synchronises the stream. */
int qbh_ctx_stream_bandwidth(qbh_ctx ctx, double bytes, int reps, double* gbps);

/* ---- plane-wave basis ------------------------------------------------------------------
 * Replaces the index maps that FourierTransform(basis, np0, np1, np2) builds from a qball
 * Basis (FourierTransform.cc:144-526, Basis.cc:302-700).  The caller passes its rods in
 * its own order (Basis::rod_h/rod_k/rod_lmin/rod_size of the local process row); plane
 * wave ig of rod r is (h, k, lmin + ig - first(r)).  real != 0 selects Γ half-sphere
 * storage (Basis::real()); the first rod must then be rod (0,0) starting at l = 0.
 * b: reciprocal lattice vectors as rows (UnitCell::b(0..2)); kpoint in reduced units. */
int qbh_basis_create(qbh_ctx ctx, int real, int nrods, const int* rod_h, const int* rod_k,
                     const int* rod_lmin, const int* rod_size, int np0, int np1, int np2,
                     const double* b, const double* kpoint, qbh_basis* out);
int qbh_basis_destroy(qbh_basis basis);
/* ngw = number of plane waves (Basis::localsize()); ncol = nonzero z-columns (nvec_). */
int qbh_basis_info(qbh_basis basis, int* ngw, int* ncol, int* nlx);

/* Restates Basis::resize (Basis.cc:302-700) for one process row: rods of the sphere
 * |k+G|^2 < 2*ecut for cell rows a[0..2] (bohr), ecut in Hartree.  Writes at most
 * max_rods rods; *nrods receives the count (call with max_rods = 0 to size).  np receives
 * the Basis::np() FFT sizes.  Integer output, identical to the reference's rods. */
int qbh_rods_build(const double* a, double ecut, int real, const double* kpoint, int max_rods,
                   int* nrods, int* rod_h, int* rod_k, int* rod_lmin, int* rod_size, int* np);
/* Basis::factorizable (Basis.cc:126-154). */
int qbh_factorizable(int n);

/* ---- transforms (FourierTransform::backward / forward, FourierTransform.h:144-153) ------
 * backward: f_dev[s] (complex grid, np0*np1*np2) = sum_G z(G) exp(+iG.r) for nf transforms.
 *   pair != 0 (Γ only): transform s carries bands (2s, 2s+1) of c_dev, as in
 *   backward(c1, c2, f) (FourierTransform.cc:555-567); odd last band alone.
 *   pair == 0: transform s carries band s (backward(c, f)).
 * forward: inverse with 1/N; pair != 0 unpacks two bands (forward(f, c1, c2)),
 *   pair == 0 writes zvec(+G) (forward(f, c)). */
int qbh_fft_backward(qbh_basis basis, const double* c_dev, long ld, int nbands, int pair,
                     double* f_dev, int flags);
int qbh_fft_forward(qbh_basis basis, const double* f_dev, int nbands, int pair, double* c_dev,
                    long ld, int flags);
/* qbh_fft_backward without any device allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments and results as qbh_fft_backward; its
 * transform workspace is a 256 MiB array each GPU holds for the life of the process, so a
 * call fails (QBH_ERR_ARG) when one transform slot of the grid needs more (FP64: 16 bytes x
 * (columns x np2 + np2 x x-lines x np1)); larger band counts run in chunks.  Calls that
 * overlap on one GPU race on that array. */
int qbh_fft_backward_noalloc(qbh_basis basis, const double* c_dev, long ld, int nbands, int pair,
                             double* f_dev, int flags);
/* forward of a real grid (ChargeDensity.cc:523-550): c = forward(scale * rho + 0i). */
int qbh_fft_forward_real(qbh_basis basis, const double* rho_dev, double scale, double* c_dev,
                         int flags);

/* ---- Hamiltonian ------------------------------------------------------------------------ */
int qbh_ham_create(qbh_ctx ctx, qbh_basis basis, qbh_ham* out);
int qbh_ham_destroy(qbh_ham ham);
/* v_r on the grid of the basis (EnergyFunctional::v_r, EnergyFunctional.h:104). */
int qbh_ham_set_vloc(qbh_ham ham, const double* v_dev);
/* Norm-conserving KB projectors (NonLocalPotential::init / update_twnl,
 * NonLocalPotential.cc:76-258, :261-1521): nsp species; species s has na[s] atoms,
 * npr[s] projectors with angular momentum lproj[...] and weight wt[...] (Species::wsg);
 * twnl: per species npr[s] x ngw table (index ig + ngw*ipr); tau: 3*na[s] positions per
 * species (bohr).  Arrays are concatenated over species.  omega = cell volume. */
int qbh_ham_set_projectors(qbh_ham ham, int nsp, const int* na, const int* npr,
                           const int* lproj, const double* wt, const double* twnl,
                           const double* tau, double omega);
/* Vector potential A (VectorPotential::value_): kinetic |k+G-A|^2 in the order of
 * VectorPotential::get_kpgpa2 (VectorPotential.h:84-93) and projector phases with
 * (k+G-A).tau (NonLocalPotential.cc:1966-1969).  The caller re-tables twnl at |k+G-A|
 * (vector_potential_changed, EnergyFunctional.cc:2047-2057) with set_projectors. */
int qbh_ham_set_vector_potential(qbh_ham ham, const double* A);

/* Hψ (EnergyFunctional::energy(psi, compute_hpsi=true, dwf, ...) Hψ part,
 * EnergyFunctional.cc:1142-1153, 1500, 1639-1701): hpsi = [NL] + [T] + [V_loc] psi.
 * occ (host, nst entries) feeds Enl (NonLocalPotential.cc:2106-2148); enl may be NULL. */
int qbh_hpsi(qbh_ham ham, const double* psi_dev, long ld, int nst, const double* occ,
             double* hpsi_dev, long ldh, int flags, double* enl);

/* Non-local ionic forces (NonLocalPotential::energy(..., compute_forces=true, fion, ...),
 * norm-conserving branch, NonLocalPotential.cc:2175-2323; fion[is][3*ia+j] = tmpfion at
 * :2617-2622): fion (host, 3 per atom, atoms in the species order of set_projectors) is
 * overwritten with -dEnl/dtau = -sum_n 2 occ_n (wt/Omega) Re(fnl conj(dfnl_j)) (Γ: x2 for
 * G, -G in dfnl), dfnl_j the contraction with the (k+G-A)_j (-i)^(l+1) table.  enl
 * (nullable, host) receives Enl.  flags: QBH_FP32 takes both projector products in FP32
 * (FP64 sums); 0 = FP64. */
int qbh_nl_forces(qbh_ham ham, const double* psi_dev, long ld, int nst, const double* occ,
                  double* fion, int flags, double* enl);

/* ρ(r) += (weight/omega) * occ[n] * |psi_n(r)|^2 over bands with positive factor
 * (SlaterDet::compute_density, SlaterDet.cc:839-932).  rho_dev is np0*np1*np2 doubles;
 * the caller zeroes it (ChargeDensity::update_density, ChargeDensity.cc:287-290). */
int qbh_density(qbh_basis basis, const double* psi_dev, long ld, int nst, const double* occ,
                double weight, double omega, double* rho_dev, int flags);

/* Gram orthonormalisation in place (SlaterDet::gram, SlaterDet.cc:1043-1143) in FP64:
 * Γ: S = 2 C^T C - c0 c0^T, complex: S = C^H C; S = L L^T; C <- C L^-T (L^-H). */
int qbh_gram(qbh_basis basis, double* psi_dev, long ld, int nst);

/* Subspace diagonalisation in FP64 (Wavefunction::diag, Wavefunction.cc:1510-1715):
 * H = 2 C^T (HC) - c0 (hc)0^T (Γ) or C^H (HC); eigenvalues (ascending) into eig (host,
 * nst entries); if eigvec != 0, C <- C Z in place.  hpsi_dev is left unchanged (the
 * reference uses dwf as scratch). */
int qbh_subspace_diag(qbh_basis basis, double* psi_dev, long ld, const double* hpsi_dev,
                      long ldh, int nst, double* eig, int eigvec);

/* Self-consistent local potential on the density grid (EnergyFunctional::update_vhxc,
 * EnergyFunctional.cc:353-977; LDA PZ81 of LDAFunctional::xc_unpolarized, nspin = 1):
 *   v_r = v_xc(rho) + FFT^-1[v_ion(G) + 4 pi (rhog/omega + rho_ps(G)) / G^2]
 * vbasis: the density basis (ChargeDensity::vbasis_) on the same grid as rho.  rhog_dev is
 * ChargeDensity::rhog (forward transform of omega*rho, as qbh_fft_forward_real gives);
 * hamil_rhog_dev (nullable) replaces rhog in the Hartree term (TDDFT Hamiltonian density,
 * EnergyFunctional.cc:473-487); vion_g_dev, rhops_g_dev (nullable) are the ionic local
 * potential and pseudo-charge on the density basis.  energies (nullable, host) receives
 * {eps, ehart, exc}. */
int qbh_update_vhxc(qbh_basis vbasis, const double* rhor_dev, const double* rhog_dev,
                    const double* hamil_rhog_dev, const double* vion_g_dev,
                    const double* rhops_g_dev, double omega, double* v_r_dev, double* energies);

/* ---- setup from a pseudopotential (Species, update_twnl, cell_moved / atoms_moved) ------
 * A species holds the radial tables of Species::initialize for a norm-conserving
 * Kleinman-Bylander pseudopotential (Species.cc:74-673): Vloc(G) of the local potential
 * minus the long-range part of the Gaussian ion charge exp(-r^2/rcps^2), the KB weights
 * wsg and the projector transforms, each on the reference's cubic spline in |G|.
 * create: vps and phi hold (lmax+1) radial arrays of npts points on the mesh deltar (the
 * row of llocal in phi is not read); create_xml reads them from a species XML file (the
 * <norm_conserving_pseudopotential> element; semilocal species are QBH_ERR_UNSUPPORTED).
 * Species need no context: the device tables are made on first use on a device. */
int qbh_species_create(double zval, int lmax, int llocal, double deltar, int npts,
                       const double* vps, const double* phi, double rcps, qbh_species* out);
int qbh_species_create_xml(const char* path, double rcps, qbh_species* out);
/* The Gaussian width AtomSet::set_rcps (AtomSet.cc:1102-1163) picks for a species of
 * valence zval in the cell a (3 x 3 host, rows = cell vectors) at the wavefunction cutoff
 * ecut (Hartree): the smallest multiple of 0.05 above 0.3 whose reciprocal-space tail is
 * below 1e-7 rcps / zval, capped at min_wsdist / 2 pi and reduced until the charge outside
 * the cell is below 1e-8. */
int qbh_species_auto_rcps(double zval, const double* a, double ecut, double* rcps);
int qbh_species_destroy(qbh_species sp);
/* zval, rcps, eself = zval^2 / (sqrt(2 pi) rcps) (Species.h:159-160), the number of KB
 * projectors (sum of 2l+1 over l != llocal) and, per projector in the order of
 * qbh_species_twnl, its l and weight wsg (each output nullable; lproj / wt sized nproj). */
int qbh_species_info(qbh_species sp, double* zval, double* rcps, double* eself, int* nproj,
                     int* lproj, double* wt);
/* Host evaluation of the splines at n values g: which = -1 Vloc(G) (Species::dvlocg), else
 * the projector of angular momentum l = which (Species::dvnlg); v, dv (nullable). */
int qbh_species_eval(qbh_species sp, int which, int n, const double* g, double* v, double* dv);
/* NonLocalPotential::update_twnl (NonLocalPotential.cc:261-1521, KB branch): twnl_dev
 * (device, nproj rows of ngw) at |k+G-A| of the basis, A (nullable = 0) the vector
 * potential, rows and angular factors in the reference's order (l ascending; p: x, y, z;
 * d: y4..y8; f: y9..y15).  The rows, qbh_species_info's lproj and wt, and the atom
 * positions are the arguments of qbh_ham_set_projectors. */
int qbh_species_twnl(qbh_species sp, qbh_basis basis, const double* A, double* twnl_dev);
/* The gradient of qbh_species_twnl's rows with respect to q = k+G-A: dtwnl_dev (device, 3
 * blocks of nproj rows of ngw, d = x, y, z) = d twnl / dq_d; zero where q = 0.  The
 * derivative with respect to A is its negative (the stress tables of update_twnl,
 * NonLocalPotential.cc:261-1521, differentiate the same rows with respect to strain). */
int qbh_species_dtwnl(qbh_species sp, qbh_basis basis, const double* A, double* dtwnl_dev);
/* NonLocalPotential::update_twnl + EnergyFunctional::vector_potential_changed
 * (EnergyFunctional.cc:2047-2057) behind qbh_ham_set_projectors: the projector tables of
 * the Hamiltonian from species (rows, lproj, wt as qbh_species_twnl / qbh_species_info give
 * them; na atoms of each species, tau host 3 per atom species after species), tabled at
 * |k+G-A| of the Hamiltonian's vector potential and re-tabled by every later
 * qbh_ham_set_vector_potential (so by the TD propagator at each step).  The species must
 * outlive the Hamiltonian or its next qbh_ham_set_projectors / qbh_ham_set_species. */
int qbh_ham_set_species(qbh_ham ham, int nsp, const qbh_species* sps, const int* na,
                        const double* tau, double omega);
/* EnergyFunctional::cell_moved + atoms_moved (EnergyFunctional.cc:1889-1915, 2001-2021)
 * on a density basis: vion_dev = sum_s S_s(G) Vloc_s(|G|) / omega and rhops_dev =
 * sum_s S_s(G) (-zval_s exp(-rcps_s^2 G^2 / 4)) / omega, S_s(G) = sum_a exp(-i G.tau_a)
 * over the na[s] atoms of species s (tau_dev: device, 3 per atom, species after species);
 * complex per G (overwritten), the inputs of qbh_update_vhxc. */
int qbh_local_tables(qbh_basis vbasis, int nsp, const qbh_species* sps, const int* na,
                     const double* tau_dev, double omega, double* vion_dev, double* rhops_dev);
/* The pseudo-charge repulsion energy esr of EnergyFunctional::atoms_moved
 * (EnergyFunctional.cc:1918-1988): pairs of distinct atoms, each separation folded into
 * the Wigner-Seitz cell and summed over (2 n + 1)^3 images, n = int(2.5 rckj / (shortest
 * of the 13 Wigner-Seitz translations) + 0.5), rckj^2 = rcps_k^2 + rcps_j^2.  a: 3 x 3
 * host, rows = cell vectors; tau: host, 3 per atom, species after species. */
int qbh_ewald_esr(const double* a, int nsp, const qbh_species* sps, const int* na,
                  const double* tau, double* esr);

/* One exponential of the ETRS propagator (ExponentialWavefunctionStepper::exponential,
 * ExponentialWavefunctionStepper.cc:51-148): psi <- sum_{N<=order} (-i dt)^N/N! H^N psi
 * in place, complex basis.  flags: the Hψ terms as for qbh_hpsi, all three when none is
 * named, and QBH_FP32 for the FP32-bulk Hψ. */
int qbh_exponential(qbh_ham ham, double* psi_dev, long ld, int nst, double dt, int order,
                    int flags);

/* Electronic current (CurrentDensity::update_current, CurrentDensity.cc:39-106):
 *   j_d(r) = -Im sum_n (weight/omega) occ_n conj(psi_n(r)) (FFT^-1[i (k+G)_d c_n])(r)
 * over bands with positive factor (two-state SlaterDet::compute_density, SlaterDet.cc:
 * 935-966), d = x, y, z with the basis' k+G.  j_dev (nullable) receives 3 grids (d-major,
 * np0*np1*np2 each, overwritten); total (nullable, host, 3) receives
 * (omega/N) sum_r j_d(r) + A_d * nelectrons (A nullable: no vector potential). */
int qbh_current(qbh_basis basis, const double* psi_dev, long ld, int nst, const double* occ,
                double weight, double omega, const double* A, double nelectrons, double* j_dev,
                double* total);

/* One ETRS step with a self-consistent H(t + dt) (ExponentialWavefunctionStepper::preupdate
 * and update, ExponentialWavefunctionStepper.cc:151-276, merge_exp_ false, ETRS not AETRS),
 * complex basis:
 *   half  = T(dt/2) psi under H(t);  tilde = T(dt/2) half under H(t)   (preupdate)
 *   rho(tilde) and rhog = forward(omega rho)            (hamil_cd_->update_density())
 *   v_r of ham <- update_vhxc(vbasis, rho, rhog, hamil_rhog = rhog, vion_g, rhops_g)
 *                                     (update_hamiltonian(); update_vhxc(), :259-261)
 *   psi   = T(dt/2) half under H(t + dt)                                (update)
 * T = the order-`order` Taylor exponential of qbh_exponential.  vbasis must share the
 * wavefunction grid (ChargeDensity builds its wavefunction transforms on the density
 * grid, ChargeDensity.cc:145).  occ (host) and weight feed the density; energies
 * (nullable, host) receives {eps, ehart, exc} of the H(t + dt) density.  On return ham
 * holds H(t + dt)'s v_r. */
int qbh_etrs_step(qbh_ham ham, qbh_basis vbasis, double* psi_dev, long ld, int nst,
                  const double* occ, double weight, double dt, int order,
                  const double* vion_g_dev, const double* rhops_g_dev, double omega,
                  double* energies, int flags);

/* ---- SCF driver pieces --------------------------------------------------------------------
 * PSDAWavefunctionStepper (PSDAWavefunctionStepper.cc:38-375) with the Preconditioner of
 * Preconditioner.cc:47-103 (K_G = 0.5 / max(|k+G|^2/2, ecutprec); qball's ecutprec = 0
 * means the wavefunction ecut).  update: A = C^T(HC) in the basis metric (Γ: 2 C^T HC -
 * c0 hc0^T), hpsi_dev <- -K (HC - C A) (the reference also overwrites dwf), then
 * psi <- psi + descent with the Anderson extrapolation over the previous step (theta,
 * clamped as the reference does, is returned).  The caller orthonormalises (qbh_gram). */
int qbh_psda_create(qbh_basis basis, int nst, double ecutprec, qbh_psda* out);
int qbh_psda_destroy(qbh_psda p);
int qbh_psda_preprocess(qbh_psda p);  /* clears the extrapolation history (.h:53-58) */
int qbh_psda_update(qbh_psda p, double* psi_dev, long ld, double* hpsi_dev, long ldh,
                    const double* occ, double* theta);
/* sum_n occ_n <c_n| |k+G|^2/2 |c_n> in the basis metric (the ekin sum of
 * EnergyFunctional::energy, EnergyFunctional.cc:1155-1359). */
int qbh_kinetic_energy(qbh_basis basis, const double* psi_dev, long ld, int nst,
                       const double* occ, double* ekin);
/* AndersonMixer (AndersonMixer.cc) over m doubles with nmax (<= 8) previous pairs:
 * xbar/fbar = x_k/f_k - sum_i theta_i (x_k - x_{k-i-1}) with theta from the Tikhonov-
 * regularised least squares (regularisation doubled until sum theta^2 <= 1).  theta
 * (nullable, host) receives min(updates-1, nmax) values. */
int qbh_mixer_create(qbh_ctx ctx, long m, int nmax, qbh_mixer* out);
int qbh_mixer_destroy(qbh_mixer mixer);
int qbh_mixer_restart(qbh_mixer mixer);
int qbh_mixer_update(qbh_mixer mixer, const double* x_dev, const double* f_dev, double* xbar_dev,
                     double* fbar_dev, double* theta);
/* The BOSampleStepper SCF loop for fixed ions (BOSampleStepper.cc:880-1130) with
 * wf_dyn = PSDA and Anderson charge mixing: per cycle update_density, mixing of rhog
 * (Kerker rcut rc_kerker, wls rc1 = 3, coefficient mix_alpha, mix_ndim previous pairs,
 * BOSampleStepper.cc:394-447, 987-1046), update_rhor, update_vhxc (into ham's v_r), then
 * nite x (Hψ -> PSDA update -> gram).  The last inner iteration of each cycle uses
 * flags_last, the others flags_inner (e.g. QBH_FP32 then 0: one FP64 residual-correction
 * pass per cycle; the Hψ terms named in the flags, all three when none is named).  history (nullable, host) receives
 * nitscf*max(nite,1) rows of {etotal, ekin, enl, eps, ehart, exc, theta}; etotal excludes
 * the ion-ion terms esr - eself, which do not depend on the orbitals. */
int qbh_scf_run(qbh_ham ham, qbh_basis vbasis, double* psi_dev, long ld, int nst,
                const double* occ, double omega, const double* vion_g_dev,
                const double* rhops_g_dev, int nitscf, int nite, double ecutprec,
                double mix_alpha, int mix_ndim, double rc_kerker, int flags_inner,
                int flags_last, double* history);
/* qbh_scf_run with the subspace diagonalisation each cycle and, for smearing > 0,
 * fractional occupations (Fermi smearing of width `smearing`, Rydberg, as the reference's
 * smearing_width; BOSampleStepper.cc:1150-1185): after each cycle's iterations Hψ
 * (flags_last), the subspace diagonalisation with eigenvectors (Wavefunction::diag) and
 * Wavefunction::update_occ (Wavefunction.cc:1214-1361: mu by step-halving from 0 in 10 eV
 * steps until |sum occ - nel| <= 1e-8, occ = 2 f(eig, mu)).  smearing = 0: occupations
 * stay as given (the reference's wf_diag = T, BOSampleStepper.cc:197), nel is not read,
 * mu receives 0.
 * occ (host, nst) holds the starting occupations and receives the final ones; etotal in
 * history includes ets = -entropy * smearing / 2 (EnergyFunctional.cc:1602-1609); eig
 * (nullable, nitscf x nst) and mu (nullable, nitscf) receive each cycle's eigenvalues and
 * chemical potential (Hartree). */
int qbh_scf_run_smeared(qbh_ham ham, qbh_basis vbasis, double* psi_dev, long ld, int nst,
                        double* occ, double nel, double smearing, double omega,
                        const double* vion_g_dev, const double* rhops_g_dev, int nitscf,
                        int nite, double ecutprec, double mix_alpha, int mix_ndim,
                        double rc_kerker, int flags_inner, int flags_last, double* history,
                        double* eig, double* mu);

/* ---- launch tuning, autotuner, multi-cell graph (no counterpart in the reference) ------
 * Knobs (every value computes the same Hψ; only speed differs):
 *   "zy_reg"  1/0  register two-factor z/y transforms where instantiated (basis-wide)
 *   "x_reg"   1/0  register x pass where instantiated (basis-wide)
 *   "x_spw"   0..64 slots per x-pass workgroup, 0 = occupancy heuristic (basis-wide)
 *   "radix"   0..2 Stockham radix plan variant for the other lengths (basis-wide)
 *   "nsplit"  0..64 K slices of the projector contraction, 0 = heuristic
 *   "overlap" -1/0/1 projector GEMMs concurrent with the transforms, -1 = context default
 *   "fnl_algo", "enl_algo"  -1..17 projector GEMMs F = anl^T psi / hpsi += anl F by
 *             this library's matrix-core kernels (17; F over K slices), by rocBLAS (0) or
 *             by the n-th algorithm hipBLASLt lists for the shape (1..16; F over the full
 *             K; n past the end of the list: the last; complex bases, which hipBLASLt
 *             does not cover, use rocBLAS for 1..16); -1 = built-in: 17, except rocBLAS
 *             for the enl product of a complex basis */
int qbh_ham_set_tuning(qbh_ham ham, const char* knob, int value);
int qbh_ham_get_tuning(qbh_ham ham, const char* knob, int* value);
/* Times candidate knob values for this Hψ shape (reps applications each, HIP events on the
 * context's stream), keeps the fastest, and caches the choice per shape in the context
 * (force != 0 re-tunes).  flags: the Hψ terms (all three when none is named) and
 * QBH_FP32.  knob_mask: bit k set = tune knob k in the order listed above (0 = all); the
 * others keep their current values.  hpsi_dev is overwritten.  report (nullable) receives
 * the timings. */
int qbh_ham_autotune(qbh_ham ham, const double* psi_dev, long ld, int nst, const double* occ,
                     double* hpsi_dev, long ldh, int flags, int reps, int force, int knob_mask,
                     char* report, int cap);
/* Hψ of ncells cells (each: hams[i], psi_dev[i] with ld[i] and nst[i] bands, into
 * hpsi_dev[i] with ldh[i]; flags as qbh_hpsi without ACCUMULATE) captured into one
 * hipGraph over nbranch concurrent branches.  The context's stream must be a non-default
 * stream (qbh_ctx_create with stream NULL).  The graph reads the Hamiltonians' current v_r,
 * kinetic and projector tables at each launch; re-create it after changing a cell's shape,
 * band count or projector set.  Cells sharing a Hamiltonian run on one branch. */
int qbh_batch_create(qbh_ctx ctx, int ncells, const qbh_ham* hams, const double* const* psi_dev,
                     const long* ld, const int* nst, double* const* hpsi_dev, const long* ldh,
                     int flags, int nbranch, qbh_batch* out);
int qbh_batch_launch(qbh_batch batch);
int qbh_batch_destroy(qbh_batch batch);

/* ---- wavefunction files ------------------------------------------------------------------
 * raw: the per-task block of Wavefunction::write_fast / write_dump (Wavefunction.cc:
 * 2221-2404) for one process row: ld x nst complex doubles (the c() block, mloc = ld), then
 * eig[nst] and occ[nst] when both are given (qball writes them with empty states).  append
 * != 0 appends (several spins / k-points in one file, in the caller's order).  read_raw
 * starts at byte `offset` and returns the offset after the block.  psi_dev must be complete
 * (its producing stream synchronised) before a write. */
int qbh_wf_write_raw(const char* path, int append, const double* psi_dev, long ld, int nst,
                     const double* eig, const double* occ);
int qbh_wf_read_raw(const char* path, long offset, double* psi_dev, long ld, int nst,
                    double* eig, double* occ, long* next_offset);
/* xml: SlaterDet::write (SlaterDet.cc:2925-3343) for one process row with base64 encoding:
 * <slater_determinant [spin] kpoint weight size>, the diagonal <density_matrix> (occ), and
 * per state a <grid_function> of the orbital on the basis grid (backward transform; real
 * part at Γ, complex values otherwise; native byte order, 72 characters a line).  The basis
 * must carry the wavefunction basis' own grid (qbh_rods_build's np).  read_xml parses one
 * such element at or after `offset`, transforms each grid function back to the sphere and
 * returns the offset after </slater_determinant>; occ (nullable) receives the occupations. */
int qbh_sd_write_xml(qbh_basis basis, const char* path, int append, const double* psi_dev,
                     long ld, int nst, const double* occ, double weight, int ispin, int nspin);
int qbh_sd_read_xml(qbh_basis basis, const char* path, long offset, double* psi_dev, long ld,
                    int nst, double* occ, long* next_offset);

/* ---- real-time TDDFT over many steps and its adjoint -----------------------------------
 * The EhrenSampleStepper iteration (EhrenSampleStepper.cc:280-1152) for fixed ions, one
 * spin and k-point, complex basis, per step n:
 *   v_a = update_vhxc(rho(psi_n)) + v_ext                                  (:323-330)
 *   preupdate under H[v_a, A_n]; A -> A_{n+1} (:477-480); update under
 *   H[v_b, A_{n+1}] with v_b = update_vhxc(rho(psi~)) + v_ext    (qbh_etrs_step halves)
 * and the objective J = sum_{n<=nsteps} w_n e.j(psi_n, A_n) over the total current of
 * qbh_current (CurrentDensity.cc:39-106, j_d = -sum f_b (k+G)_d |c|^2 + A_d N_e).  The KB
 * radial tables of ham stay as set (not re-tabled at |k+G-A|); the projector phases at
 * k+G-A cancel in beta D beta^H, so A enters H through 0.5|k+G-A|^2.  No counterpart in the
 * reference for the adjoint: qbh_td_backward is the exact derivative of this discrete map
 * (dJ/dv_ext on the grid, dJ/dA_n), recomputing segments between checkpoints.
 * ham (projectors set) and vbasis (complex density basis on the same grid) must outlive td;
 * vion_g_dev / rhops_g_dev (nullable) are borrowed.  ham's v_r and A are restored on
 * return. */
int qbh_td_create(qbh_ham ham, qbh_basis vbasis, int nst, const double* occ, double weight,
                  double omega, const double* vion_g_dev, const double* rhops_g_dev, qbh_td* out);
int qbh_td_destroy(qbh_td td);
/* psi0_dev (not modified); A: host 3*(nsteps+1); w: host nsteps+1; e: host 3 (direction);
 * vext_dev: nullable, np0*np1*np2 doubles, borrowed until the matching backward;
 * checkpoint: steps between stored psi; flags: QBH_FP32 selects the FP32-bulk Hψ (all
 * three terms of H either way).
 * J receives the objective; psi_out_dev (nullable) psi_nsteps. */
int qbh_td_forward(qbh_td td, const double* psi0_dev, long ld, int nsteps, double dt, int order,
                   const double* A, const double* w, const double* e, const double* vext_dev,
                   int checkpoint, int flags, double* J, double* psi_out_dev, long ldo);
/* Gradients of the last forward's J: dvext_dev (nullable, device, N doubles, overwritten)
 * and dA (nullable, host, 3*(nsteps+1)). */
int qbh_td_backward(qbh_td td, double* dvext_dev, double* dA);

#ifdef __cplusplus
}
#endif
#endif /* QBH_H */
