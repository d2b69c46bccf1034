"""The differentiation form at its benchmark size: real-time TDDFT on diamond Si64 in a
complex basis (18277 plane waves, 128 bands, 72^3 grid), 200 ETRS steps of dt = 0.05 a.u.
under the velocity-gauge field A(t) = A0 sin(0.3 t) exp(-(t - t0)^2 / 2 tau^2) along x with
A0 = 0.01, objective the time-integrated dipole along x, psi stored every 20 steps.

Through the torch autograd function: dJ/dA0 and dJ/dv_ext (along one smooth direction)
agree with central differences (h = 1e-5) of the FP64 forward to 1e-6 relative; the
FP32-bulk form's J and gradients are compared with the FP64 ones.  Two Hamiltonians: fixed
analytic s+p projector tables with smooth Gaussian ionic inputs, and the Si HSCV LDA
pseudopotential (projectors re-tabled at |k+G-A| every step, v_ion and rho_ps from its
local part)."""
import os
import time

import numpy as np
import pytest

from helpers import make_gpu_basis, make_gpu_ham, rel_err

pytestmark = pytest.mark.gpu

NSTEPS, DT, CHECKPOINT, H = 200, 0.05, 20, 1e-5


SI_XML = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "Si_HSCV_LDA-1.0.xml")


@pytest.fixture(scope="module", params=["analytic", "pseudopotential"])
def si64td(request):
    import torch
    from oracle.basis import Basis as OB, UnitCell
    from qball_amd import Basis, Context, Species, TDPropagator, capi, local_tables
    from qball_amd.workload import make_system, td_ionic_inputs
    sysm = make_system(ncell=2, ecut_ry=25.0, nbands=128, seed=41, complex_basis=True)
    stream = torch.cuda.current_stream()
    ctx = Context(0, stream.cuda_stream)
    basis = make_gpu_basis(ctx, sysm)
    r = capi.rods_build(sysm.a, 4.0 * sysm.ecut, False)
    vb = Basis(ctx, False, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], sysm.np, sysm.b)
    vob = OB(UnitCell(sysm.a), 4.0 * sysm.ecut, force_complex=True)
    assert vob.size == vb.ngw
    cplx = lambda z: torch.from_numpy(np.ascontiguousarray(z).view(np.float64).copy()).cuda()
    species = None
    if request.param == "analytic":
        ham = make_gpu_ham(ctx, basis, sysm)
        # ionic inputs of physical size (Gaussian ion charges neutral with the 256 electrons
        # and a short-range local potential), so that dt * |H| stays inside the Taylor
        # propagator's range
        vion, rhops = td_ionic_inputs(sysm, r)
    else:
        species = Species(SI_XML, rcps=1.5)
        ham = make_gpu_ham(ctx, basis, sysm, with_proj=False)
        ham.set_species([species], [sysm.na], sysm.tau, sysm.omega)
        tau_d = torch.from_numpy(np.ascontiguousarray(sysm.tau).ravel().copy()).cuda()
        vg = torch.zeros(2 * vb.ngw, dtype=torch.float64, device="cuda")
        rg = torch.zeros_like(vg)
        local_tables(vb, [species], [sysm.na], tau_d.data_ptr(), sysm.omega, vg.data_ptr(),
                     rg.data_ptr())
        ctx.sync()
        vion = vg.cpu().numpy().view(np.complex128)
        rhops = rg.cpu().numpy().view(np.complex128)
    psi = cplx(np.ascontiguousarray(sysm.psi.T))
    basis.gram(psi.data_ptr(), basis.ngw, sysm.nst)        # orthonormal start (FP64, GPU)
    n0, n1, n2 = sysm.np
    z, y, x = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    vext = (0.02 * np.cos(2 * np.pi * x / n0) * np.cos(2 * np.pi * y / n1)
            + 0.01 * np.sin(2 * np.pi * z / n2)).ravel()
    dv = (np.sin(2 * np.pi * (x + 2 * y) / n0) * np.cos(2 * np.pi * z / n2)).ravel()
    occ = np.full(sysm.nst, 2.0)
    dev = dict(vion=cplx(vion), rhops=cplx(rhops), psi=psi,
               vext=torch.from_numpy(vext.copy()).cuda(), dv=torch.from_numpy(dv.copy()).cuda())
    td = TDPropagator(ham, vb, sysm.nst, occ, 1.0, sysm.omega, dev["vion"].data_ptr(),
                      dev["rhops"].data_ptr())
    print("Si64 TD (", request.param, "): ngw", basis.ngw, "bands", sysm.nst, "grid", sysm.np,
          flush=True)
    yield dict(sysm=sysm, ctx=ctx, basis=basis, td=td, dev=dev)
    td.close()
    ham.close()
    if species is not None:
        species.close()
    vb.close()
    basis.close()
    ctx.close()


def _objective(s, vext, A0, flags=0):
    import torch
    from qball_amd import torch_td
    shape = torch_td.field_shape(NSTEPS, DT, omega=0.3)
    w = torch_td.dipole_weights(NSTEPS, DT)
    if not torch.is_tensor(A0):
        A0 = torch.tensor(float(A0), dtype=torch.float64)
    return torch_td.td_objective(s["td"], s["dev"]["psi"].data_ptr(), s["basis"].ngw, vext, A0,
                                 NSTEPS, DT, shape, w, checkpoint=CHECKPOINT, flags=flags)


def _gradients(s, flags):
    import torch
    vext = s["dev"]["vext"].clone().requires_grad_(True)
    A0 = torch.tensor(0.01, dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    J = _objective(s, vext, A0, flags)
    t1 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    print(f"flags {flags:#x}: J {float(J.detach()):.12e}  forward {t1 - t0:.2f} s  backward "
          f"{t2 - t1:.2f} s  dJ/dA0 {float(A0.grad):.12e}", flush=True)
    return float(J.detach()), vext.grad.detach().clone(), float(A0.grad)


def test_td_gradients_at_benchmark_size(si64td):
    import torch
    from qball_amd import QBH_FP32
    s = si64td
    J, gv, gA0 = _gradients(s, 0)
    v0, dv = s["dev"]["vext"], s["dev"]["dv"]
    with torch.no_grad():
        fdA = (float(_objective(s, v0, 0.01 + H)) - float(_objective(s, v0, 0.01 - H))) / (2 * H)
        fdv = (float(_objective(s, v0 + H * dv, 0.01))
               - float(_objective(s, v0 - H * dv, 0.01))) / (2 * H)
    adv = float((gv * dv).sum())
    print("dJ/dA0 fd", fdA, "adjoint", gA0, "| dJ/dvext.dv fd", fdv, "adjoint", adv, flush=True)
    assert abs(fdA - gA0) <= 1e-6 * abs(fdA)
    assert abs(fdv - adv) <= 1e-6 * abs(fdv)

    # FP32-bulk form against FP64 (same propagator, Hψ with QBH_FP32)
    J32, gv32, gA032 = _gradients(s, QBH_FP32)
    ej, egv, egA = abs(J32 - J) / abs(J), rel_err(gv32.cpu().numpy(), gv.cpu().numpy()), \
        abs(gA032 - gA0) / abs(gA0)
    print("FP32-bulk vs FP64: J", ej, "dJ/dvext", egv, "dJ/dA0", egA, flush=True)
    assert ej <= 1e-5 and egA <= 1e-4 and egv <= 1e-3
