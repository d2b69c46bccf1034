"""TD propagation and its adjoint on the GPU: parity with the oracle's discrete adjoint,
checkpoint independence and central finite differences."""
import numpy as np
import pytest

from helpers import make_gpu_basis, make_gpu_ham, oracle_for, rel_err, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def td_setup():
    import torch
    from oracle import hpsi as oh
    from oracle.basis import Basis as OB, FFTMaps, UnitCell
    from oracle.tddft import TDSystem
    from qball_amd import Basis, Context, TDPropagator, capi
    from qball_amd.workload import make_system
    sysm = make_system(ncell=1, ecut_ry=25.0, nbands=6, seed=31, complex_basis=True)
    ob, maps, sp = oracle_for(sysm)
    vob = OB(UnitCell(sysm.a), 4.0 * sysm.ecut, force_complex=True)
    vmaps = FFTMaps(vob, *sysm.np)
    rng = np.random.default_rng(8)
    vion = (rng.standard_normal(vob.size) + 1j * rng.standard_normal(vob.size)) * 0.5 * np.exp(-0.3 * vob.g2)
    rhops = (0.05 * np.exp(-0.5 * vob.g2)).astype(np.complex128)
    psi = oh.gram(sysm.psi, False)
    occ = np.array([2.0, 2.0, 2.0, 2.0, 1.0, 0.0])
    tds = TDSystem(maps, vmaps, [sp], occ, sysm.omega, vion, rhops)
    x = np.arange(maps.N)
    vext = 0.05 * np.cos(x * 2 * np.pi / sysm.np[0]) * np.sin(x * 0.001)
    stream = torch.cuda.current_stream()
    ctx = Context(0, stream.cuda_stream)
    basis = make_gpu_basis(ctx, sysm)
    ham = make_gpu_ham(ctx, basis, sysm)
    r = capi.rods_build(sysm.a, 4.0 * sysm.ecut, False)
    vb = Basis(ctx, False, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], sysm.np, sysm.b)
    cplx = lambda z: torch.from_numpy(np.ascontiguousarray(z).view(np.float64).copy()).cuda()
    dev = dict(vion=cplx(vion), rhops=cplx(rhops), psi=to_dev(psi),
               vext=torch.from_numpy(vext.copy()).cuda())
    td = TDPropagator(ham, vb, len(occ), occ, 1.0, sysm.omega, dev["vion"].data_ptr(),
                      dev["rhops"].data_ptr())
    yield dict(sysm=sysm, tds=tds, psi=psi, vext=vext, ctx=ctx, basis=basis, ham=ham, td=td,
               dev=dev)
    td.close()
    ham.close()
    vb.close()
    basis.close()
    ctx.close()


def _schedule(nsteps, dt, A0=0.05):
    t = dt * np.arange(nsteps + 1)
    A = np.zeros((nsteps + 1, 3))
    A[:, 0] = A0 * np.sin(0.3 * t + 0.4)
    A[:, 1] = 0.3 * A0 * np.cos(0.7 * t)
    w = dt * dt * (nsteps - np.arange(nsteps + 1) + 0.5)
    return A, w


def test_td_adjoint_matches_oracle(td_setup):
    import torch
    from oracle.tddft import td_gradient
    s = td_setup
    nsteps, dt = 2, 0.1
    A, w = _schedule(nsteps, dt)
    e = np.array([1.0, 0.2, 0.0])
    J = s["td"].forward(s["dev"]["psi"].data_ptr(), s["basis"].ngw, nsteps, dt, A, w, e,
                        vext_ptr=s["dev"]["vext"].data_ptr(), checkpoint=1)
    gv = torch.zeros(s["tds"].maps.N, dtype=torch.float64, device="cuda")
    gA = s["td"].backward(gv.data_ptr())
    s["ctx"].sync()
    Jr, gvr, gAr = td_gradient(s["tds"], s["psi"], s["vext"], A, w, e, dt, nsteps)
    assert abs(J - Jr) <= 1e-12 * abs(Jr)
    assert rel_err(gv.cpu().numpy(), gvr) <= 1e-10
    assert rel_err(gA, gAr) <= 1e-10


def test_td_checkpoint_spacing_does_not_change_results(td_setup):
    import torch
    s = td_setup
    nsteps, dt = 5, 0.1
    A, w = _schedule(nsteps, dt)
    out = []
    for ck in (1, 2, 5):
        J = s["td"].forward(s["dev"]["psi"].data_ptr(), s["basis"].ngw, nsteps, dt, A, w,
                            vext_ptr=s["dev"]["vext"].data_ptr(), checkpoint=ck)
        gv = torch.zeros(s["tds"].maps.N, dtype=torch.float64, device="cuda")
        gA = s["td"].backward(gv.data_ptr())
        s["ctx"].sync()
        out.append((J, gv.cpu().numpy(), gA))
    for J, gv, gA in out[1:]:
        assert abs(J - out[0][0]) <= 1e-13 * abs(out[0][0])
        assert rel_err(gv, out[0][1]) <= 1e-12 and rel_err(gA, out[0][2]) <= 1e-12


def test_td_adjoint_matches_central_differences(td_setup):
    import torch
    s = td_setup
    nsteps, dt, h = 6, 0.1, 1e-5
    A, w = _schedule(nsteps, dt)
    psi, ngw = s["dev"]["psi"].data_ptr(), s["basis"].ngw
    vext = s["dev"]["vext"]

    def J_of(v, AA):
        return s["td"].forward(psi, ngw, nsteps, dt, AA, w, vext_ptr=v.data_ptr(), checkpoint=3)
    J_of(vext, A)
    gv = torch.zeros_like(vext)
    gA = s["td"].backward(gv.data_ptr())
    s["ctx"].sync()
    x = np.arange(s["tds"].maps.N)
    dv = torch.from_numpy(np.sin(x * 0.013) * np.cos(x * 0.0007)).cuda()
    fd = (J_of(vext + h * dv, A) - J_of(vext - h * dv, A)) / (2 * h)
    ad = float((gv * dv).sum())
    print("dJ/dvext.dv", fd, ad)
    assert abs(fd - ad) <= 1e-6 * abs(fd)
    # amplitude of the field: A_n = A0 * s_n along x
    shape = A[:, 0] / 0.05
    fdA = (J_of(vext, A + h * np.outer(shape, [1, 0, 0])) - J_of(vext, A - h * np.outer(shape, [1, 0, 0]))) / (2 * h)
    adA = float(gA[:, 0] @ shape)
    print("dJ/dA0", fdA, adA)
    assert abs(fdA - adA) <= 1e-6 * abs(fdA)


def test_torch_autograd_function(td_setup):
    """The libtorch autograd Function gives the adjoint's gradients for J(v_ext, A0)."""
    import torch
    from qball_amd import torch_td
    s = td_setup
    nsteps, dt = 3, 0.1
    A, w = _schedule(nsteps, dt)
    shape = torch.from_numpy(A / 0.05)
    vext = s["dev"]["vext"].clone().requires_grad_(True)
    A0 = torch.tensor(0.05, dtype=torch.float64, requires_grad=True)
    J = torch_td.td_objective(s["td"], s["dev"]["psi"].data_ptr(), s["basis"].ngw, vext, A0, nsteps,
                              dt, shape, torch.from_numpy(w), checkpoint=2)
    J.backward()
    J2 = s["td"].forward(s["dev"]["psi"].data_ptr(), s["basis"].ngw, nsteps, dt, A, w,
                         vext_ptr=s["dev"]["vext"].data_ptr(), checkpoint=2)
    gv = torch.zeros_like(s["dev"]["vext"])
    gA = s["td"].backward(gv.data_ptr())
    s["ctx"].sync()
    assert abs(float(J.detach()) - J2) <= 1e-13 * abs(J2)
    assert rel_err(vext.grad.cpu().numpy(), gv.cpu().numpy()) <= 1e-13
    assert abs(float(A0.grad) - float((gA * (A / 0.05)).sum())) <= 1e-12 * abs(float(A0.grad))


def test_fp32_exponential_applies_the_whole_hamiltonian(td_setup):
    """QBH_FP32 alone selects the FP32-bulk Hψ with all three terms of H (the FP64 form's
    exponential to the FP32 bar), not an empty Hamiltonian."""
    import torch
    from qball_amd import QBH_FP32
    s = td_setup
    ham, ngw = s["ham"], s["basis"].ngw
    nst = s["sysm"].nst
    out = {}
    for fl in (0, QBH_FP32):
        p = s["dev"]["psi"].clone()
        ham.exponential(p.data_ptr(), ngw, nst, 0.1, 4, flags=fl)
        s["ctx"].sync()
        out[fl] = p.cpu().numpy()
    p0 = s["dev"]["psi"].cpu().numpy()
    assert rel_err(out[0], p0) > 1e-3                  # H is applied
    assert rel_err(out[QBH_FP32], out[0]) <= 2e-6


def test_td_fp32_forward_tracks_fp64(td_setup):
    """The FP32-bulk propagation stays within FP32 rounding of the FP64 one over a few
    steps, objective and gradients."""
    import torch
    from qball_amd import QBH_FP32
    s = td_setup
    nsteps, dt = 4, 0.1
    A, w = _schedule(nsteps, dt)
    e = np.array([1.0, 0.2, 0.0])
    res = {}
    for fl in (0, QBH_FP32):
        J = s["td"].forward(s["dev"]["psi"].data_ptr(), s["basis"].ngw, nsteps, dt, A, w, e,
                            vext_ptr=s["dev"]["vext"].data_ptr(), checkpoint=2, flags=fl)
        gv = torch.zeros(s["tds"].maps.N, dtype=torch.float64, device="cuda")
        gA = s["td"].backward(gv.data_ptr())
        s["ctx"].sync()
        res[fl] = (J, gv.cpu().numpy(), gA)
    (J64, gv64, gA64), (J32, gv32, gA32) = res[0], res[QBH_FP32]
    print("FP32 vs FP64: J", abs(J32 - J64) / abs(J64), "dJ/dv", rel_err(gv32, gv64), "dJ/dA",
          rel_err(gA32, gA64))
    assert abs(J32 - J64) <= 1e-5 * abs(J64)
    assert rel_err(gv32, gv64) <= 1e-4 and rel_err(gA32, gA64) <= 1e-4
