"""The library's matrix-core projector products (qbh_gemm.hip: F = anl^T psi over K slices,
hpsi += anl F taken transposed) against the oracle's non-local term
(NonLocalPotential::energy, NonLocalPotential.cc:2041-2171) on shapes that exercise their
edges: projector counts that are odd (element pairs of F not aligned) or not a multiple of
the 128-wide tile, band counts past one tile and not a multiple of it, K slices from one to
many, FP64 and FP32, overwriting and accumulating outputs; on Γ (the real view) and on a
complex basis (the combined table of rows beta_p and i beta_p).

FP64 bar: relative 2-norm error <= 1e-12 (Enl <= 1e-10 relative); FP32 form <= 2e-6.
"""
import numpy as np
import pytest

from helpers import from_dev, oracle_for, real_to_dev, rel_err, to_dev

pytestmark = pytest.mark.gpu

OWN = 17  # fnl_algo / enl_algo value of the matrix-core kernels


def _species(sysm, natoms, lproj, seed):
    """natoms atoms (jittered copies of the cell's sites) with one projector per entry of
    lproj and smooth radial tables on this basis' |k+G|."""
    from oracle import hpsi as oh
    ob, _, _ = oracle_for(sysm)
    g = np.sqrt(ob.kpg2)
    rng = np.random.default_rng(seed)
    tau = sysm.tau[np.arange(natoms) % sysm.na] + 0.2 * rng.standard_normal((natoms, 3))
    twnl = np.stack([(1.0 + 0.2 * i) * g ** l * np.exp(-(0.3 + 0.05 * i) * g * g)
                     for i, l in enumerate(lproj)])
    wt = np.array([1.1, -0.6, 0.45, 0.3, -0.2][:len(lproj)])
    return oh.Projectors(list(lproj), wt, twnl, tau)


def _run(ctx, sysm, species, flags, nst=None, accumulate_from=None, nsplit=0):
    import torch
    from helpers import make_gpu_basis
    from qball_amd import Hamiltonian
    basis = make_gpu_basis(ctx, sysm)
    ham = Hamiltonian(ctx, basis)
    vd = real_to_dev(sysm.v)
    ham.set_vloc(vd.data_ptr())
    ham.set_projectors([sp.na for sp in species], [sp.npr for sp in species],
                       np.concatenate([sp.lproj for sp in species]),
                       np.concatenate([sp.wt for sp in species]),
                       np.concatenate([sp.twnl.ravel() for sp in species]),
                       np.concatenate([sp.tau.ravel() for sp in species]), sysm.omega)
    ham.set_tuning("fnl_algo", OWN)
    ham.set_tuning("enl_algo", OWN)
    ham.set_tuning("nsplit", nsplit)
    nst = sysm.nst if nst is None else nst
    psi = sysm.psi[:, :nst]
    occ = np.full(nst, 2.0)
    cd = to_dev(psi)
    hd = to_dev(accumulate_from) if accumulate_from is not None else torch.full_like(cd, np.nan)
    enl = ham.hpsi(cd.data_ptr(), basis.ngw, nst, occ, hd.data_ptr(), basis.ngw, flags=flags,
                   want_enl=True)
    ctx.sync()
    got = from_dev(hd, basis.ngw, nst)
    ham.close()
    basis.close()
    return got, enl, psi, occ


def _ref(sysm, species, psi, occ):
    from oracle import hpsi as oh
    ob, _, _ = oracle_for(sysm)
    cp = np.zeros_like(psi)
    enl = oh.nonlocal_energy(ob, species, psi, occ, sysm.omega, cp=cp)
    return cp, enl


@pytest.fixture(scope="module")
def si64():
    from qball_amd.workload import make_system
    return make_system(ncell=2, ecut_ry=25.0, nbands=300, seed=31)


@pytest.mark.parametrize("natoms,lproj", [(5, (0, 1, 2)), (3, (0,)), (47, (0, 1, 1)),
                                          (131, (0, 2))])
@pytest.mark.parametrize("fp32", [False, True])
def test_projector_products_odd_and_ragged_projector_counts(gpu_ctx, si64, natoms, lproj, fp32):
    """P = 15, 3, 141 and 262 projector columns (odd ones read F element by element), 300
    bands (three tiles, the last partial)."""
    sp = _species(si64, natoms, lproj, seed=natoms)
    flags = 0x1 | (0x10 if fp32 else 0)  # non-local term only
    got, enl, psi, occ = _run(gpu_ctx, si64, [sp], flags)
    ref, enl_ref = _ref(si64, [sp], psi, occ)
    tol = 2e-6 if fp32 else 1e-12
    assert rel_err(got, ref) <= tol, (natoms, lproj, fp32)
    assert abs(enl - enl_ref) <= (1e-5 if fp32 else 1e-10) * abs(enl_ref)


@pytest.mark.parametrize("nst", [1, 2, 129, 256])
@pytest.mark.parametrize("nsplit", [0, 1, 7, 64])
def test_projector_products_band_counts_and_slices(gpu_ctx, si64, nst, nsplit):
    """One band to two full tiles of bands, the reduction over K in one slice or many."""
    from qball_amd.workload import make_system  # noqa: F401
    _, _, sp = oracle_for(si64)
    got, enl, psi, occ = _run(gpu_ctx, si64, [sp], 0x1, nst=nst, nsplit=nsplit)
    ref, enl_ref = _ref(si64, [sp], psi, occ)
    assert rel_err(got, ref) <= 1e-12, (nst, nsplit)
    assert abs(enl - enl_ref) <= 1e-10 * abs(enl_ref)


@pytest.mark.parametrize("fp32", [False, True])
def test_projector_products_accumulate(gpu_ctx, si64, fp32):
    """QBH_ACCUMULATE adds the non-local term to what the output holds."""
    _, _, sp = oracle_for(si64)
    rng = np.random.default_rng(5)
    base = rng.standard_normal(si64.psi.shape) + 1j * rng.standard_normal(si64.psi.shape)
    got, _, psi, occ = _run(gpu_ctx, si64, [sp], 0x1 | 0x8 | (0x10 if fp32 else 0),
                            accumulate_from=base)
    ref, _ = _ref(si64, [sp], psi, occ)
    assert rel_err(got - base, ref) <= (2e-6 if fp32 else 1e-12)


# ---- complex bases: the same kernels over the combined table (rows beta_p and i beta_p) ----

@pytest.fixture(scope="module")
def si8c():
    from qball_amd.workload import make_system
    return make_system(ncell=1, ecut_ry=25.0, nbands=300, seed=33, complex_basis=True)


@pytest.mark.parametrize("natoms,lproj", [(5, (0, 1, 2)), (47, (0, 1, 1)), (3, (0,))])
@pytest.mark.parametrize("fp32", [False, True])
def test_complex_projector_products_ragged(gpu_ctx, si8c, natoms, lproj, fp32):
    """Complex basis, both products on the matrix-core kernels: 2P = 30, 282 and 6 table
    rows (not multiples of the 128-row tile), 300 bands (the last band tile partial)."""
    assert not si8c.real
    sp = _species(si8c, natoms, lproj, seed=100 + natoms)
    flags = 0x1 | (0x10 if fp32 else 0)
    got, enl, psi, occ = _run(gpu_ctx, si8c, [sp], flags)
    ref, enl_ref = _ref(si8c, [sp], psi, occ)
    assert rel_err(got, ref) <= (2e-6 if fp32 else 1e-12), (natoms, lproj, fp32)
    assert abs(enl - enl_ref) <= (1e-5 if fp32 else 1e-10) * abs(enl_ref)


@pytest.mark.parametrize("nst", [1, 129])
@pytest.mark.parametrize("nsplit", [0, 1, 7])
def test_complex_projector_products_band_counts_and_slices(gpu_ctx, si8c, nst, nsplit):
    """Complex basis: one band and one band past a tile, K in one slice or several."""
    _, _, sp = oracle_for(si8c)
    got, enl, psi, occ = _run(gpu_ctx, si8c, [sp], 0x1, nst=nst, nsplit=nsplit)
    ref, enl_ref = _ref(si8c, [sp], psi, occ)
    assert rel_err(got, ref) <= 1e-12, (nst, nsplit)
    assert abs(enl - enl_ref) <= 1e-10 * abs(enl_ref)


@pytest.mark.parametrize("fp32", [False, True])
def test_complex_projector_products_accumulate(gpu_ctx, si8c, fp32):
    """Complex basis: QBH_ACCUMULATE adds the non-local term to what the output holds."""
    _, _, sp = oracle_for(si8c)
    rng = np.random.default_rng(7)
    base = rng.standard_normal(si8c.psi.shape) + 1j * rng.standard_normal(si8c.psi.shape)
    got, _, psi, occ = _run(gpu_ctx, si8c, [sp], 0x1 | 0x8 | (0x10 if fp32 else 0),
                            accumulate_from=base)
    ref, _ = _ref(si8c, [sp], psi, occ)
    assert rel_err(got - base, ref) <= (2e-6 if fp32 else 1e-12)
