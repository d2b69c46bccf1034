"""The qball-shaped C++ adaptor (qball_amd/host/qbh_qball.h) end to end on the GPU:
FourierTransform::backward, EnergyFunctional::energy(compute_hpsi=true), the non-local
forces and ChargeDensity::update_density against the CPU oracle on the driver's own inputs."""
import math
import os
import subprocess

import numpy as np
import pytest

from helpers import rel_err
from oracle import fft as offt
from oracle import hpsi as oh
from oracle.basis import Basis, FFTMaps, UnitCell, density_grid

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_adaptor_matches_oracle(tmp_path):
    exe = os.path.join(ROOT, "qball_amd", "host", "adaptor_check")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    alat, ecut, nst = 10.26, 12.5, 9
    cell = UnitCell.cubic(alat)
    b = Basis(cell, ecut)
    vb, npd = density_grid(cell, ecut)
    maps = FFTMaps(b, *npd)
    vmaps = FFTMaps(vb, *npd)
    rd = lambda n, dt: np.fromfile(os.path.join(tmp_path, n), dtype=dt)
    psi = rd("psi.bin", np.complex128).reshape(nst, b.size).T.copy()
    v = rd("v.bin", np.float64)
    twnl = rd("twnl.bin", np.float64).reshape(4, b.size)
    tau = rd("tau.bin", np.float64).reshape(8, 3)
    occ = np.array([2.0] * (nst - 1) + [0.0])
    # FourierTransform::backward(c1, c2, f)
    f = rd("f.bin", np.complex128)
    assert rel_err(f, offt.backward(maps, psi[:, 0], psi[:, 1]).ravel()) <= 1e-12
    # Hψ and Enl
    sp = oh.Projectors([0, 1, 1, 1], [1.2, -0.5, -0.5, -0.5], twnl, tau)
    ref, enl = oh.hpsi(maps, psi, v, [sp], occ, cell.volume)
    h = rd("hpsi.bin", np.complex128).reshape(nst, b.size).T
    assert rel_err(h, ref) <= 1e-12
    sc = rd("scalars.bin", np.float64)
    assert abs(sc[0] - enl) <= 1e-10
    # non-local ionic forces
    _, fref = oh.nonlocal_forces(b, [sp], psi, occ, cell.volume)
    assert rel_err(rd("fion.bin", np.float64).reshape(8, 3), fref[0]) <= 1e-12
    # update_density: rho, nelectrons, rhog
    rho, nel, rhog = oh.update_density(maps, vmaps, psi, occ, 1.0, cell.volume)
    assert rel_err(rd("rho.bin", np.float64), rho) <= 1e-12
    assert abs(sc[1] - nel) <= 1e-10
    assert rel_err(rd("rhog.bin", np.complex128), rhog) <= 1e-12
