"""Hψ, density, Gram and the ETRS exponential -- oracle restatement.

Test infrastructure only (see ``oracle/__init__.py``).  Wavefunction blocks are numpy
arrays ``c[ngw, nst]`` (complex128), i.e. the reference's column-major ComplexMatrix
``c()`` with mloc = ngw (math/matrix.h:374-380).
"""
import numpy as np

from . import fft as offt


# --------------------------------------------------------------------------------------
# local potential
def _re(x):
    """Contiguous copy of the real part: numpy's matmul only hands contiguous operands to
    BLAS (a strided .real view takes a plain loop, hundreds of times slower)."""
    return np.ascontiguousarray(x.real)


def _im(x):
    return np.ascontiguousarray(x.imag)


def rs_mul_add(maps, c, v, cp):
    """SlaterDet::rs_mul_add (SlaterDet.cc:971-1040): cp[:, n] += FFT^-1[v * FFT c_n].

    Γ: bands in pairs through one complex transform (:989-1006), odd last band alone with
    the imaginary part of v*psi zeroed (:1007-1023).  Complex basis: one band per
    transform (:1025-1037).  ``v`` is the real potential on the grid, flat in grid order.
    """
    vg = v.reshape(maps.np2, maps.np1, maps.np0)
    nst = c.shape[1]
    if maps.basis.real:
        n = 0
        while n < nst - 1:
            f = offt.backward(maps, c[:, n], c[:, n + 1])
            f = vg * f.real + 1j * (vg * f.imag)
            c1, c2 = offt.forward(maps, f, pair=True)
            cp[:, n] += c1
            cp[:, n + 1] += c2
            n += 2
        if nst % 2 != 0:
            n = nst - 1
            f = offt.backward(maps, c[:, n])
            f = (vg * f.real).astype(np.complex128)
            cp[:, n] += offt.forward(maps, f, pair=False)
    else:
        for n in range(nst):
            f = offt.backward(maps, c[:, n])
            cp[:, n] += offt.forward(maps, vg * f, pair=False)
    return cp


def rs_mul_add_batched(maps, c, v, cp, batch=8, plan=None):
    """rs_mul_add through the pruned pencil transforms of oracle.fft.PrunedPlan, `batch`
    transforms at a time.  Γ: bands in pairs; an odd last band alone with the imaginary
    part of v*psi zeroed (SlaterDet.cc:1007-1023).  Complex basis: one band per transform."""
    plan = plan or offt.PrunedPlan(maps)
    vt = np.ascontiguousarray(v.reshape(maps.np2, maps.np1, maps.np0).transpose(2, 1, 0))
    vt = vt[:, :, :, None]  # (np0, np1, np2, 1): the plan's grid layout
    nst = c.shape[1]
    real = maps.basis.real
    step = 2 * batch if real else batch
    npair = 2 * (nst // 2) if real else nst
    for n0 in range(0, npair, step):
        n1 = min(npair, n0 + step)
        f = plan.bwd(c[:, n0:n1], pair=real)
        f.real *= vt
        f.imag *= vt
        cp[:, n0:n1] += plan.fwd(f, n1 - n0, pair=real)
    if real and nst % 2:
        n = nst - 1
        f = plan.bwd(c[:, n:n + 1], pair=False)
        f = (vt * f.real).astype(np.complex128)
        cp[:, n:n + 1] += plan.fwd(f, 1, pair=False)
    return cp


def kinetic_add(kpg2, c, cp):
    """EnergyFunctional.cc:1677-1693: cp += 0.5*|k+G|^2 c (kpg2 may carry -A,
    VectorPotential.h:84-93)."""
    cp += (0.5 * kpg2)[:, None] * c
    return cp


def kpgpa2(basis, A):
    """VectorPotential::get_kpgpa2 (VectorPotential.h:84-93), same evaluation order."""
    A = np.asarray(A, dtype=np.float64)
    value2 = A[0] * A[0] + A[1] * A[1] + A[2] * A[2]
    out = basis.kpg2 + value2
    out = out - 2 * A[0] * basis.kpgx[:, 0]
    out = out - 2 * A[1] * basis.kpgx[:, 1]
    out = out - 2 * A[2] * basis.kpgx[:, 2]
    return out


# --------------------------------------------------------------------------------------
# non-local (Kleinman-Bylander) projectors
class Projectors:
    """Per-species KB data as NonLocalPotential holds it (NonLocalPotential.cc:76-258):
    ``lproj[ipr]``, ``wt[ipr]`` (=Species::wsg), ``twnl[ipr, ig]`` (update_twnl table),
    ``tau[ia, 3]`` atomic positions (bohr)."""

    def __init__(self, lproj, wt, twnl, tau):
        self.lproj = np.asarray(lproj, dtype=np.int64)
        self.wt = np.asarray(wt, dtype=np.float64)
        self.twnl = np.asarray(twnl, dtype=np.float64)
        self.tau = np.asarray(tau, dtype=np.float64).reshape(-1, 3)

    @property
    def npr(self):
        return len(self.lproj)

    @property
    def na(self):
        return self.tau.shape[0]


def anl_block(sp, kpgx, ia0, ia1):
    """comp_eigr + comp_anl (NonLocalPotential.cc:1959-2037) for atoms [ia0, ia1):
    returns complex anl[ngw, nprna] with column ipra = ia + ipr*nblk and
    anl = (-i)^l * twnl * exp(-i (k+G).tau)."""
    tau = sp.tau[ia0:ia1]
    nblk = ia1 - ia0
    # kpgr[ig, ia] = -(k+G).tau  (dgemm with alpha=-1, :1969)
    kpgr = -(kpgx[:, 0:1] * tau[None, :, 0] + kpgx[:, 1:2] * tau[None, :, 1]
             + kpgx[:, 2:3] * tau[None, :, 2])
    cs = np.cos(kpgr)
    sn = np.sin(kpgr)
    ngw = kpgx.shape[0]
    anl = np.empty((ngw, sp.npr * nblk), dtype=np.complex128)
    re, im = anl.real, anl.imag   # views: each block is written in place
    for ipr in range(sp.npr):
        t = sp.twnl[ipr][:, None]
        l = int(sp.lproj[ipr])
        cols = slice(ipr * nblk, (ipr + 1) * nblk)
        if l == 0:      # twnl * (cos, sin)
            np.multiply(t, cs, out=re[:, cols])
            np.multiply(t, sn, out=im[:, cols])
        elif l == 1:    # twnl * (sin, -cos)
            np.multiply(t, sn, out=re[:, cols])
            np.multiply(-t, cs, out=im[:, cols])
        elif l == 2:    # twnl * (-cos, -sin)
            np.multiply(-t, cs, out=re[:, cols])
            np.multiply(-t, sn, out=im[:, cols])
        elif l == 3:    # twnl * (-sin, cos)
            np.multiply(-t, sn, out=re[:, cols])
            np.multiply(t, cs, out=im[:, cols])
        else:
            raise ValueError("l > 3")
    return anl


def nonlocal_energy(basis, species, c, occ, omega, cp=None, kpgx=None, na_block=None):
    """NonLocalPotential::energy, norm-conserving branch (NonLocalPotential.cc:1909-2171).

    Returns enl; if ``cp`` is given, accumulates cp += anl * (wt/Omega * fnl) (:2150-2171).
    Γ: fnl = 2*(Re(anl)^T Re(c) + Im(anl)^T Im(c)) - anl0*c0 (dgemm over 2*ngw rows
    :2053, dger(-0.5) on row 0 :2071-2082, x2 :2101).  Complex: fnl = anl^H c (:2064).
    """
    if kpgx is None:
        kpgx = basis.kpgx
    omega_inv = 1.0 / omega
    namax = max(sp.na for sp in species)
    if na_block is None:
        na_block = 128 if namax > 128 else namax     # :1535-1541
    enl = 0.0
    if basis.real:
        cr, ci = _re(c), _im(c)
    for sp in species:
        if sp.npr == 0:
            continue
        nblocks = (sp.na + na_block - 1) // na_block
        for ib in range(nblocks):
            ia0 = ib * na_block
            ia1 = min(sp.na, ia0 + na_block)
            nblk = ia1 - ia0
            anl = anl_block(sp, kpgx, ia0, ia1)
            if basis.real:
                ar, ai = _re(anl), _im(anl)
                fnl = ar.T @ cr + ai.T @ ci
                fnl = fnl - 0.5 * np.outer(anl[0].real, c[0].real)
                fnl = 2.0 * fnl
            else:
                fnl = anl.conj().T @ c
            for ipr in range(sp.npr):
                fac = sp.wt[ipr] * omega_inv
                rows = slice(ipr * nblk, (ipr + 1) * nblk)
                t = fnl[rows]
                if basis.real:
                    enl += float(np.sum((fac * occ)[None, :] * t * t))
                else:
                    enl += float(np.sum((fac * occ)[None, :] * (t.real ** 2 + t.imag ** 2)))
                fnl[rows] = fac * t
            if cp is not None:
                if basis.real:
                    cp.real += ar @ fnl
                    cp.imag += ai @ fnl
                else:
                    cp += anl @ fnl
    return enl


def danl_block(sp, kpgx, j, ia0, ia1):
    """The force table of NonLocalPotential.cc:2188-2236 for direction j and atoms
    [ia0, ia1): column ipra = ia + ipr*nblk holds (k+G)_j twnl (-i)^(l+1) exp(-i (k+G).tau),
    written with the per-l real/imaginary pattern of :2204-2233 from cos/sin of
    kpgr = -(k+G).tau (comp_eigr, :1966-1975)."""
    tau = sp.tau[ia0:ia1]
    nblk = ia1 - ia0
    kpgr = -(kpgx[:, 0:1] * tau[None, :, 0] + kpgx[:, 1:2] * tau[None, :, 1]
             + kpgx[:, 2:3] * tau[None, :, 2])
    cs = np.cos(kpgr)
    sn = np.sin(kpgr)
    danl = np.empty((kpgx.shape[0], sp.npr * nblk), dtype=np.complex128)
    for ipr in range(sp.npr):
        tt = (kpgx[:, j] * sp.twnl[ipr])[:, None]
        l = int(sp.lproj[ipr])
        if l == 0:
            d = tt * sn + 1j * (-tt * cs)
        elif l == 1:
            d = -tt * cs + 1j * (-tt * sn)
        elif l == 2:
            d = -tt * sn + 1j * (tt * cs)
        elif l == 3:
            d = tt * cs + 1j * (tt * sn)
        else:
            raise ValueError("l > 3")
        danl[:, ipr * nblk:(ipr + 1) * nblk] = d
    return danl


def nonlocal_forces(basis, species, c, occ, omega, kpgx=None, na_block=None):
    """Non-local ionic forces, NonLocalPotential::energy(compute_forces=true),
    norm-conserving branch (NonLocalPotential.cc:2175-2323; fnl scaled by wt/Omega at
    :2106-2148 first; fion[is][3*ia+j] = tmpfion at :2617-2622).

    For each direction j the table anl_j = (k+G)_j twnl (-i)^(l+1) exp(-i (k+G).tau)
    (:2199-2234) gives dfnl = 2 anl_j^T c over the real view at Γ (no G=0 term, :2262-2267)
    or anl_j^H c (:2254-2259), and fion[ia, j] -= 2 occ_n fnl dfnl (Γ, :2283-2288) or
    2 occ_n Re(fnl conj(dfnl)) (:2307-2311), fnl already multiplied by wt/Omega.
    ``kpgx`` may carry -A (VectorPotential::get_kpgpax, :2187).  Returns (enl, fion) with
    fion a list of (na, 3) arrays, one per species."""
    if kpgx is None:
        kpgx = basis.kpgx
    omega_inv = 1.0 / omega
    namax = max(sp.na for sp in species)
    if na_block is None:
        na_block = 128 if namax > 128 else namax     # :1535-1541
    enl = 0.0
    fion = []
    for sp in species:
        f_sp = np.zeros((sp.na, 3))
        fion.append(f_sp)
        if sp.npr == 0:
            continue
        nblocks = (sp.na + na_block - 1) // na_block
        for ib in range(nblocks):
            ia0 = ib * na_block
            ia1 = min(sp.na, ia0 + na_block)
            nblk = ia1 - ia0
            anl = anl_block(sp, kpgx, ia0, ia1)
            if basis.real:
                fnl = _re(anl).T @ _re(c) + _im(anl).T @ _im(c)
                fnl = fnl - 0.5 * np.outer(anl[0].real, c[0].real)
                fnl = 2.0 * fnl
            else:
                fnl = anl.conj().T @ c
            for ipr in range(sp.npr):
                fac = sp.wt[ipr] * omega_inv
                rows = slice(ipr * nblk, (ipr + 1) * nblk)
                t = fnl[rows]
                if basis.real:
                    enl += float(np.sum((fac * occ)[None, :] * t * t))
                else:
                    enl += float(np.sum((fac * occ)[None, :] * (t.real ** 2 + t.imag ** 2)))
                fnl[rows] = fac * t
            for j in range(3):
                danl = danl_block(sp, kpgx, j, ia0, ia1)
                if basis.real:
                    dfnl = 2.0 * (_re(danl).T @ _re(c) + _im(danl).T @ _im(c))
                    prod = fnl * dfnl
                else:
                    dfnl = danl.conj().T @ c
                    prod = (fnl * dfnl.conj()).real
                facn = 2.0 * np.asarray(occ, dtype=np.float64)
                contrib = (prod * facn[None, :]).sum(axis=1)      # per column ipra
                for ipr in range(sp.npr):
                    f_sp[ia0:ia1, j] -= contrib[ipr * nblk:(ipr + 1) * nblk]
    return enl, fion


# --------------------------------------------------------------------------------------
def hpsi(maps, c, v, species, occ, omega, kpg2=None, kpgx=None):
    """EnergyFunctional::energy(compute_hpsi=true) Hψ part (EnergyFunctional.cc:1142-1153
    clear, :1500 non-local, :1677-1693 kinetic, :1695 local).  Returns (hpsi, enl)."""
    basis = maps.basis
    cp = np.zeros_like(c)
    enl = nonlocal_energy(basis, species, c, occ, omega, cp=cp, kpgx=kpgx) if species else 0.0
    kinetic_add(basis.kpg2 if kpg2 is None else kpg2, c, cp)
    rs_mul_add(maps, c, v, cp)
    return cp, enl


# --------------------------------------------------------------------------------------
def compute_density(maps, c, occ, weight, omega, rho=None):
    """SlaterDet::compute_density (SlaterDet.cc:839-932): one backward transform per band
    (the pair branch is disabled at :858); rho += (weight/Omega)*occ_n*|psi_n(r)|^2 for
    bands with positive factor.  Returns rho (flat, grid order)."""
    if rho is None:
        rho = np.zeros(maps.N, dtype=np.float64)
    prefac = weight / omega
    for n in range(c.shape[1]):
        fac = prefac * occ[n]
        if fac > 0.0:
            f = offt.backward(maps, c[:, n]).ravel()
            rho += fac * (f.real * f.real + f.imag * f.imag)
    return rho


def update_density(maps, vmaps, c, occ, weight, omega):
    """ChargeDensity::update_density (ChargeDensity.cc:276-556) for one spin/k-point:
    zero rho, band sum, nelectrons = Omega/N * sum(rho), rhog = vft.forward(Omega*rho)."""
    rho = compute_density(maps, c, occ, weight, omega)
    nel = float(np.sum(rho)) * omega / rho.size
    f = (omega * rho).astype(np.complex128).reshape(vmaps.np2, vmaps.np1, vmaps.np0)
    rhog = offt.forward(vmaps, f, pair=False)
    return rho, nel, rhog


# --------------------------------------------------------------------------------------
def gram(c, real):
    """SlaterDet::gram (SlaterDet.cc:1043-1143).  Γ: S = 2 C^T C - c0 c0^T over the real
    view (syrk 'l','t',2.0 then syr -1.0 on row 0), S = L L^T, C <- C L^-T.
    Complex: S = C^H C, C <- C L^-H."""
    if real:
        cr = np.concatenate([c.real, c.imag], axis=0)
        s = 2.0 * (_re(c).T @ _re(c) + _im(c).T @ _im(c)) - np.outer(c[0].real, c[0].real)
        L = np.linalg.cholesky(s)
        x = np.linalg.solve(L, cr.T).T        # X L^T = C  <=>  L X^T = C^T
        ngw = c.shape[0]
        return x[:ngw] + 1j * x[ngw:]
    s = c.conj().T @ c
    L = np.linalg.cholesky(s)
    return np.linalg.solve(L.conj(), c.T).T


def subspace_diag(c, hc, real):
    """Wavefunction::diag (Wavefunction.cc:1510-1715) with eigvec=true.  Γ:
    H = 2 C^T (HC) - c0 (hc)0^T over the real view (gemm 't','n',2.0 then ger -1.0 on
    row 0, :1546-1548), lower triangle to syevd; complex: H = C^H (HC), heevd.  Returns
    (eigenvalues ascending, C Z)."""
    if real:
        h = 2.0 * (_re(c).T @ _re(hc) + _im(c).T @ _im(hc)) - np.outer(c[0].real, hc[0].real)
        w, z = np.linalg.eigh(h, UPLO="L")
    else:
        h = c.conj().T @ hc
        w, z = np.linalg.eigh(h, UPLO="L")
    return w, c @ z


# --------------------------------------------------------------------------------------
def exponential(apply_h, c, dt, order=4):
    """ExponentialWavefunctionStepper::exponential, one exponential (num_exp=1)
    (ExponentialWavefunctionStepper.cc:51-148): c <- sum_{N<=order} (-i dt)^N/N! H^N c,
    with factor *= -i*dt/N and H applied to the previous (unscaled) H^{N-1} c."""
    expwf = c.copy()
    factor = 1.0 + 0.0j
    w = c
    for N in range(1, order + 1):
        factor *= -1j * dt / float(N)
        w = apply_h(w)
        expwf = expwf + factor * w
    return expwf


def etrs_step(apply_h_t, apply_h_tdt, c, dt):
    """ETRS step as preupdate + update (ExponentialWavefunctionStepper.cc:153-276) with
    ``merge_exp_`` false and ETRS (not AETRS): psi(t+dt/2) = exp(-i dt/2 H(t)) psi(t);
    psi~(t+dt) = exp(-i dt/2 H(t)) psi(t+dt/2) (used by the caller to build H(t+dt));
    psi(t+dt) = exp(-i dt/2 H(t+dt)) psi(t+dt/2).  ``apply_h_tdt`` may be a callable
    taking psi~(t+dt) and returning the H(t+dt) operator."""
    half = exponential(apply_h_t, c, 0.5 * dt)
    tilde = exponential(apply_h_t, half, 0.5 * dt)
    h2 = apply_h_tdt(tilde) if apply_h_tdt is not None else apply_h_t
    return exponential(h2, half, 0.5 * dt), tilde
