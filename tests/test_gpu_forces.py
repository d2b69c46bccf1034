"""Non-local ionic forces through the C ABI (qbh_nl_forces) against the oracle restatement
of NonLocalPotential::energy(compute_forces=true) (NonLocalPotential.cc:2175-2323), whose
own check is -dEnl/dtau by finite differences (test_oracle_hpsi.py).

FP64 bar: relative 2-norm error of the (natoms, 3) force array <= 1e-12, Enl <= 1e-10 Ha.
FP32 form (both projector products in FP32, FP64 sums): <= 1e-5.
"""
import math

import numpy as np
import pytest

from helpers import oracle_for, real_to_dev, rel_err, to_dev

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


@pytest.fixture(scope="module")
def si8():
    from qball_amd.workload import make_system
    return make_system(ncell=1, ecut_ry=25.0, nbands=15, seed=21)


@pytest.fixture(scope="module")
def si8c():
    from qball_amd.workload import make_system
    return make_system(ncell=1, ecut_ry=25.0, nbands=8, seed=22, complex_basis=True)


def _ham(ctx, sysm, species):
    """GPU Hamiltonian with the given oracle species (concatenated in order)."""
    from helpers import make_gpu_basis
    from qball_amd import Hamiltonian
    basis = make_gpu_basis(ctx, sysm)
    ham = Hamiltonian(ctx, basis)
    vd = real_to_dev(sysm.v)
    ham.set_vloc(vd.data_ptr())
    ham._v = vd
    ham.set_projectors([sp.na for sp in species], [sp.npr for sp in species],
                       np.concatenate([sp.lproj for sp in species]),
                       np.concatenate([sp.wt for sp in species]),
                       np.concatenate([sp.twnl.ravel() for sp in species]),
                       np.concatenate([sp.tau.ravel() for sp in species]), sysm.omega)
    return basis, ham


def _oracle(sysm, species, occ, kpgx=None, psi=None):
    from oracle import hpsi as oh
    ob, _, _ = oracle_for(sysm)
    enl, f = oh.nonlocal_forces(ob, species, sysm.psi if psi is None else psi, occ,
                                sysm.omega, kpgx=kpgx)
    return enl, np.concatenate(f)


def _all_l_species(sysm, tau):
    """One projector per l = 0..3 with smooth radial tables on this basis' |k+G|."""
    from oracle import hpsi as oh
    ob, _, _ = oracle_for(sysm)
    g = np.sqrt(ob.kpg2)
    twnl = np.stack([(1.0 + 0.3 * l) * g ** l * np.exp(-0.35 * g * g) for l in range(4)])
    return oh.Projectors([0, 1, 2, 3], [1.3, -0.7, 0.4, 0.25], twnl, tau)


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_nl_forces_match_oracle(gpu_ctx, which, request):
    sysm = request.getfixturevalue(which)
    _, _, sp = oracle_for(sysm)
    occ = np.linspace(2.0, 0.5, sysm.nst)
    basis, ham = _ham(gpu_ctx, sysm, [sp])
    cd = to_dev(sysm.psi)
    f64, enl = ham.nl_forces(cd.data_ptr(), basis.ngw, sysm.nst, occ, want_enl=True)
    f32 = ham.nl_forces(cd.data_ptr(), basis.ngw, sysm.nst, occ, fp32=True)
    ham.close()
    basis.close()
    enl_ref, ref = _oracle(sysm, [sp], occ)
    print(which, "max|F|", np.max(np.abs(ref)), "fp64", rel_err(f64, ref), "fp32",
          rel_err(f32, ref))
    assert rel_err(f64, ref) <= FP64_TOL
    assert abs(enl - enl_ref) <= 1e-10 * max(1.0, abs(enl_ref))
    assert rel_err(f32, ref) <= FP32_TOL


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_nl_forces_two_species_all_l(gpu_ctx, which, request):
    """Columns of two species map to their own atoms; l = 0..3 phase patterns."""
    from oracle import hpsi as oh
    sysm = request.getfixturevalue(which)
    _, _, sp = oracle_for(sysm)
    s1 = oh.Projectors(sp.lproj, sp.wt, sp.twnl, sysm.tau[:3])
    s2 = _all_l_species(sysm, sysm.tau[3:])
    occ = np.full(sysm.nst, 2.0)
    basis, ham = _ham(gpu_ctx, sysm, [s1, s2])
    cd = to_dev(sysm.psi)
    f, enl = ham.nl_forces(cd.data_ptr(), basis.ngw, sysm.nst, occ, want_enl=True)
    # the Hψ of the same two species (projector columns of both in one product)
    import torch
    from helpers import from_dev
    hd = torch.zeros_like(cd)
    ham.hpsi(cd.data_ptr(), basis.ngw, sysm.nst, occ, hd.data_ptr(), basis.ngw)
    gpu_ctx.sync()
    got = from_dev(hd, basis.ngw, sysm.nst)
    ham.close()
    basis.close()
    ob, maps, _ = oracle_for(sysm)
    href = oh.hpsi(maps, sysm.psi, sysm.v, [s1, s2], occ, sysm.omega)[0]
    assert rel_err(got, href) <= FP64_TOL
    enl_ref, ref = _oracle(sysm, [s1, s2], occ)
    assert f.shape == (sysm.na, 3)
    assert rel_err(f, ref) <= FP64_TOL
    assert abs(enl - enl_ref) <= 1e-10 * max(1.0, abs(enl_ref))


@pytest.mark.parametrize("which", ["si8", "si8c"])
def test_nl_forces_vector_potential(gpu_ctx, which, request):
    """Phases and derivative factors at k+G-A (VectorPotential::get_kpgpax,
    NonLocalPotential.cc:2187)."""
    sysm = request.getfixturevalue(which)
    ob, _, sp = oracle_for(sysm)
    A = np.array([0.012, -0.02, 0.007])
    occ = np.full(sysm.nst, 2.0)
    basis, ham = _ham(gpu_ctx, sysm, [sp])
    ham.set_vector_potential(A)
    cd = to_dev(sysm.psi)
    f = ham.nl_forces(cd.data_ptr(), basis.ngw, sysm.nst, occ)
    ham.close()
    basis.close()
    _, ref = _oracle(sysm, [sp], occ, kpgx=ob.kpgx - A[None, :])
    assert rel_err(f, ref) <= FP64_TOL


@pytest.mark.parametrize("algo,nsplit", [(0, 0), (0, 3), (1, 0), (2, 0), (16, 0), (17, 0),
                                         (17, 3)])
def test_nl_forces_padded_ld_every_algorithm(gpu_ctx, si8, algo, nsplit):
    """A leading dimension past ngw, an odd band count, rocBLAS over K slices (nsplit),
    hipBLASLt algorithms and the library's matrix-core kernels (fnl_algo) give the same
    forces."""
    import torch
    _, _, sp = oracle_for(si8)
    occ = np.full(si8.nst, 2.0)
    basis, ham = _ham(gpu_ctx, si8, [sp])
    ham.set_tuning("fnl_algo", algo)
    ham.set_tuning("nsplit", nsplit)
    ngw, ld = basis.ngw, basis.ngw + 5
    pad = np.zeros((ld, si8.nst), dtype=np.complex128)
    pad[:ngw] = si8.psi
    pad[ngw:] = np.nan                        # rows past ngw are never read
    cd = to_dev(pad)
    f64 = ham.nl_forces(cd.data_ptr(), ld, si8.nst, occ)
    f32 = ham.nl_forces(cd.data_ptr(), ld, si8.nst, occ, fp32=True)
    ham.close()
    basis.close()
    _, ref = _oracle(si8, [sp], occ)
    assert rel_err(f64, ref) <= FP64_TOL
    assert rel_err(f32, ref) <= FP32_TOL
    torch.cuda.synchronize()


def test_nl_forces_rejects_bad_arguments(gpu_ctx, si8):
    import torch
    _, _, sp = oracle_for(si8)
    basis, ham = _ham(gpu_ctx, si8, [sp])
    cd = to_dev(si8.psi)
    with pytest.raises(RuntimeError, match="ld < ngw"):
        ham.nl_forces(cd.data_ptr(), basis.ngw - 1, si8.nst, si8.occ)
    with pytest.raises(RuntimeError, match="aligned"):
        ham.nl_forces(cd.data_ptr() + 8, basis.ngw, si8.nst - 1, si8.occ)
    ham.close()
    basis.close()
    torch.cuda.synchronize()
