"""Wavefunction files: the raw per-task block and the XML/base64 SlaterDet element, checked
byte by byte / value by value against numpy and the oracle transforms, and read back."""
import base64
import re

import numpy as np
import pytest

from helpers import from_dev, oracle_for, rel_err, to_dev

pytestmark = pytest.mark.gpu


def test_raw_block_layout_and_round_trip(gpu_ctx, tmp_path):
    import torch
    from qball_amd import capi
    from qball_amd.workload import make_system
    s = make_system(ncell=1, ecut_ry=25.0, nbands=5, seed=2)
    cd = to_dev(s.psi)
    eig = np.linspace(-0.5, 0.3, s.nst)
    occ = np.array([2.0, 2.0, 2.0, 1.0, 0.0])
    p = tmp_path / "wf000000"
    capi.wf_write_raw(p, cd.data_ptr(), s.ngw, s.nst, eig, occ)
    capi.wf_write_raw(p, cd.data_ptr(), s.ngw, s.nst, append=True)
    raw = np.fromfile(p, dtype=np.float64)
    blk = 2 * s.ngw * s.nst
    assert raw.size == 2 * blk + 2 * s.nst
    assert np.array_equal(raw[:blk].view(np.complex128).reshape(s.nst, s.ngw).T, s.psi)
    assert np.array_equal(raw[blk:blk + s.nst], eig) and np.array_equal(raw[blk + s.nst:blk + 2 * s.nst], occ)
    back = torch.zeros_like(cd)
    nxt, e2, o2 = capi.wf_read_raw(p, back.data_ptr(), s.ngw, s.nst, with_eig_occ=True)
    assert nxt == (blk + 2 * s.nst) * 8 and np.array_equal(e2, eig) and np.array_equal(o2, occ)
    assert torch.equal(back, cd)
    nxt2, _, _ = capi.wf_read_raw(p, back.data_ptr(), s.ngw, s.nst, offset=nxt)
    assert nxt2 == raw.size * 8 and torch.equal(back, cd)


@pytest.mark.parametrize("complex_basis", [False, True])
def test_xml_slater_determinant(gpu_ctx, tmp_path, complex_basis):
    import torch
    from oracle import fft as offt
    from oracle.basis import Basis as OB, FFTMaps, UnitCell
    from qball_amd import Basis, capi
    from qball_amd.workload import make_system
    s = make_system(ncell=1, ecut_ry=25.0, nbands=11, seed=5, complex_basis=complex_basis)
    r = s.rods
    grid = tuple(int(x) for x in r["np"])             # the wavefunction basis' own grid
    basis = Basis(gpu_ctx, s.real, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], grid, s.b)
    ob = OB(UnitCell(s.a), s.ecut, force_complex=complex_basis)
    maps = FFTMaps(ob, *grid)
    cd = to_dev(s.psi)
    occ = np.array([2.0] * 9 + [1.5, 0.25])
    p = tmp_path / "sd.xml"
    capi.sd_write_xml(basis, p, cd.data_ptr(), basis.ngw, s.nst, occ, weight=0.5)
    text = p.read_text()
    head = ('<slater_determinant kpoint="0 0 0"\n  weight="0.5" size="11">\n'
            '<density_matrix form="diagonal" size="11">\n'
            + " 2.00000000" * 9 + " 1.50000000\n 0.25000000\n</density_matrix>\n")
    assert text.startswith(head)
    assert text.endswith("</slater_determinant>\n")
    blocks = re.findall(r'<grid_function type="(\w+)" nx="(\d+)" ny="(\d+)" nz="(\d+)" '
                        r'encoding="base64">\n(.*?)</grid_function>\n', text, re.S)
    assert len(blocks) == s.nst
    for n, (typ, nx, ny, nz, body) in enumerate(blocks):
        assert typ == ("complex" if complex_basis else "double")
        assert (int(nx), int(ny), int(nz)) == grid
        lines = body.split("\n")[:-1]
        assert all(len(ln) == 72 for ln in lines[:-1]) and len(lines[-1]) <= 72
        vals = np.frombuffer(base64.b64decode("".join(lines)), dtype=np.float64)
        ref = offt.backward(maps, s.psi[:, n]).ravel()
        got = vals.view(np.complex128) if complex_basis else vals
        want = ref if complex_basis else ref.real
        assert rel_err(got, want) <= 1e-12, n
    back = torch.zeros_like(cd)
    nxt, occ2 = capi.sd_read_xml(basis, p, back.data_ptr(), basis.ngw, s.nst)
    gpu_ctx.sync()
    assert nxt == len(text.encode()) and np.array_equal(occ2, occ)
    assert rel_err(from_dev(back, basis.ngw, s.nst), s.psi) <= 1e-13
    basis.close()


def test_xml_reader_rejects_malformed_grid_functions(gpu_ctx, tmp_path):
    """A grid function whose base64 text holds more data than the basis grid, or one with no
    closing tag, is an error (the decoder never writes past the grid buffer and never reads
    past the file); the well-formed file still reads back."""
    import torch
    from qball_amd import Basis, QbhError, capi
    from qball_amd.workload import make_system
    s = make_system(ncell=1, ecut_ry=25.0, nbands=2, seed=3)
    r = s.rods
    grid = tuple(int(x) for x in r["np"])
    basis = Basis(gpu_ctx, s.real, r["rod_h"], r["rod_k"], r["rod_lmin"], r["rod_size"], grid, s.b)
    cd = to_dev(s.psi)
    p = tmp_path / "sd.xml"
    capi.sd_write_xml(basis, p, cd.data_ptr(), basis.ngw, s.nst, s.occ)
    text = p.read_text()
    back = torch.zeros_like(cd)
    # extra base64 quanta ahead of the first grid function's data: more bytes than the grid
    m = re.search(r"(<grid_function[^>]*>\n)", text)
    longer = text[:m.end(1)] + "AAAA" * 3000 + "\n" + text[m.end(1):]
    (tmp_path / "long.xml").write_text(longer)
    with pytest.raises(QbhError, match="size differs"):
        capi.sd_read_xml(basis, tmp_path / "long.xml", back.data_ptr(), basis.ngw, s.nst)
    # no closing tag on any grid function
    (tmp_path / "cut.xml").write_text(text.replace("</grid_function>", ""))
    with pytest.raises(QbhError, match="unterminated"):
        capi.sd_read_xml(basis, tmp_path / "cut.xml", back.data_ptr(), basis.ngw, s.nst)
    capi.sd_read_xml(basis, p, back.data_ptr(), basis.ngw, s.nst)
    gpu_ctx.sync()
    assert rel_err(from_dev(back, basis.ngw, s.nst), s.psi) <= 1e-13
    basis.close()
