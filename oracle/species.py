"""Oracle (test infrastructure only): the setup row of the Hψ path -- the norm-conserving
Kleinman-Bylander species tables, the projector tables twnl, the local pseudopotential and
pseudo-charge on the density basis, and the Ewald real-space energy.

Restates, in numpy:
  * the species XML reader (SpeciesReader, the <norm_conserving_pseudopotential> element);
  * Species::initialize, KB branch (src/qball/Species.cc:74-673): radial mesh extension
    (:140-260), long-range subtraction (:1749-1759), local potential transform (:280-330),
    KB weights wsg (:340-360), projector transforms bessel_trans (:773-1000), splines;
  * Species::vlocg / dvnlg / rhopsg / eself (:1451-1504, :1598-1602, Species.h:159-160);
  * the radial transforms sinft / cosft1 (sinft.cc, Numerical Recipes) as the discrete sine
    and cosine transforms they compute, and simpsn (Species.cc:45-62);
  * the cubic spline of pseudo/spline.cc (spline/splint/splintd, Numerical Recipes);
  * NonLocalPotential::update_twnl, KB branch (NonLocalPotential.cc:261-1521) for l = 0..3;
  * StructureFactor::update (StructureFactor.cc:85-173) and EnergyFunctional::cell_moved /
    atoms_moved (EnergyFunctional.cc:1889-2044): v_ion(G), rho_ps(G), esr.
"""
import math
import xml.etree.ElementTree as ET

import numpy as np
from scipy import fft as sfft
from scipy.special import erf, erfc


# ---------------------------------------------------------------------------------------
def read_species_xml(path):
    """The norm-conserving species element of a pseudopotential XML file (the quantities
    SpeciesReader::readSpecies reads for a KB pseudopotential)."""
    root = ET.parse(path).getroot()

    def child(node, tag):
        for c in node.iter():
            if c.tag == tag or c.tag.endswith("}" + tag):
                return c
        return None

    nc = child(root, "norm_conserving_pseudopotential")
    if nc is None:
        raise ValueError("not a norm-conserving pseudopotential")
    get = lambda t: child(nc, t).text.strip()
    out = {"symbol": child(root, "symbol").text.strip(),
           "zval": float(get("valence_charge")), "lmax": int(get("lmax")),
           "llocal": int(get("llocal")), "nquad": int(get("nquad")),
           "rquad": float(get("rquad")), "deltar": float(get("mesh_spacing"))}
    vps, phi = {}, {}
    for p in nc:
        if not (p.tag == "projector" or p.tag.endswith("}projector")):
            continue
        l = int(p.attrib["l"])
        size = int(p.attrib["size"])
        for c in p:
            vals = np.array(c.text.split(), dtype=np.float64)
            assert vals.size == size, (l, c.tag)
            if c.tag.endswith("radial_potential"):
                vps[l] = vals
            elif c.tag.endswith("radial_function"):
                phi[l] = vals
    out["vps"] = [vps[l] for l in range(out["lmax"] + 1)]
    out["phi"] = [phi.get(l) for l in range(out["lmax"] + 1)]
    return out


# ---------------------------------------------------------------------------------------
def simpsn(t):
    """Species.cc:45-62 (extended Simpson rule on overlapping segments)."""
    n = t.size
    s = (17.0 / 48.0) * (t[0] + t[n - 1]) + (59.0 / 48.0) * (t[1] + t[n - 2]) \
        + (43.0 / 48.0) * (t[2] + t[n - 3]) + (49.0 / 48.0) * (t[3] + t[n - 4])
    return float(s + np.sum(t[4:n - 4]))


def sinft(y):
    """sinft (sinft.cc): F_k = sum_{j=1}^{n-1} y_j sin(pi j k / n), k = 0..n-1."""
    n = y.size
    out = np.zeros(n)
    out[1:] = 0.5 * sfft.dst(y[1:], type=1)
    return out


def cosft1(y):
    """cosft1 (sinft.cc), y of n+1 points: F_k = (y_0 + (-1)^k y_n) / 2
    + sum_{j=1}^{n-1} y_j cos(pi j k / n), k = 0..n."""
    return 0.5 * sfft.dct(y, type=1)


class Spline:
    """Cubic spline of pseudo/spline.cc (spline, splint, splintd): first derivative yp1 at
    x[0] (or natural when yp1 is None), natural at x[n-1]."""

    def __init__(self, x, y, yp1):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        n = x.size
        y2 = np.zeros(n)
        u = np.zeros(n)
        if yp1 is None:
            y2[0] = 0.0
            u[0] = 0.0
        else:
            y2[0] = -0.5
            u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1)
        for i in range(1, n - 1):
            sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
            p = sig * y2[i - 1] + 2.0
            y2[i] = (sig - 1.0) / p
            u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) /
                           (x[i] - x[i - 1])) / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p
        qn = un = 0.0
        y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0)
        for k in range(n - 2, -1, -1):
            y2[k] = y2[k] * y2[k + 1] + u[k]
        self.x, self.y, self.y2 = x, y, y2

    def _seg(self, q):
        x = self.x
        khi = np.searchsorted(x, q, side="right")
        khi = np.clip(khi, 1, x.size - 1)
        klo = khi - 1
        return klo, khi

    def value(self, q):
        q = np.asarray(q, dtype=np.float64)
        klo, khi = self._seg(q)
        x, y, y2 = self.x, self.y, self.y2
        h = x[khi] - x[klo]
        a = (x[khi] - q) / h
        b = (q - x[klo]) / h
        return a * y[klo] + b * y[khi] + h * h * (1.0 / 6.0) * ((a * a * a - a) * y2[klo] +
                                                                 (b * b * b - b) * y2[khi])

    def derivative(self, q):
        q = np.asarray(q, dtype=np.float64)
        klo, khi = self._seg(q)
        x, y, y2 = self.x, self.y, self.y2
        h = x[khi] - x[klo]
        a = (x[khi] - q) / h
        b = (q - x[klo]) / h
        v = a * y[klo] + b * y[khi] + h * h * (1.0 / 6.0) * ((a * a * a - a) * y2[klo] +
                                                              (b * b * b - b) * y2[khi])
        dv = (y[khi] - y[klo]) / h + h * ((1.0 / 6.0 - 0.5 * a * a) * y2[klo] +
                                          (0.5 * b * b - 1.0 / 6.0) * y2[khi])
        return v, dv


# ---------------------------------------------------------------------------------------
class Species:
    """Species::initialize, Kleinman-Bylander branch (Species.cc:74-673), one channel."""

    def __init__(self, sp, rcps):
        self.zval = sp["zval"]
        self.lmax, self.llocal = sp["lmax"], sp["llocal"]
        self.deltar = dr = sp["deltar"]
        self.rcps = rcps
        if sp.get("nquad", 0) != 0:
            raise ValueError("semilocal species are not on this path")
        npts = sp["vps"][0].size
        ndft = 1
        while ndft < npts:
            ndft *= 2
        ndft *= 2
        while dr * ndft < 40.0:
            ndft *= 2
        self.ndft = ndft
        r = np.arange(ndft) * dr
        self.rps = r
        zv = self.zval
        vps, phi = [], []
        for l in range(self.lmax + 1):
            v = np.empty(ndft)
            v[:npts] = sp["vps"][l]
            v[npts:] = -zv / r[npts:]
            vps.append(v)
            f = np.zeros(ndft)
            if sp["phi"][l] is not None:
                f[:npts] = sp["phi"][l]
            phi.append(f)
        # local potential minus the long-range part of the smeared ion charge (:1749-1759)
        vl = vps[self.llocal]
        vsr = np.empty(ndft)
        vsr[0] = vl[0] + (zv / rcps) * (2.0 / math.sqrt(math.pi))
        vsr[1:] = vl[1:] + (zv / r[1:]) * erf(r[1:] / rcps)
        vlocr = vsr * (4.0 * math.pi) * r * dr
        v0 = simpsn(vlocr * r)
        vlocg = sinft(vlocr)
        fac = math.pi / (ndft * dr)
        g = np.arange(ndft) * fac
        g[0] = 0.0
        vlocg[1:] /= g[1:]
        vlocg[0] = v0
        self.gspl = g
        self.vlocg_tab = vlocg
        self.local_spline = Spline(g, vlocg, 0.0)
        # KB weights and projector transforms
        self.wsg = np.zeros(self.lmax + 1)
        self.proj = {}
        for l in range(self.lmax + 1):
            if l == self.llocal:
                continue
            t = phi[l] * r
            w = simpsn((vps[l] - vl) * t * t)
            self.wsg[l] = 1.0 / (dr * w)
            vnlr = 4.0 * math.pi * dr * (vps[l] - vl) * phi[l]
            vnlg = self.bessel_trans(l, vnlr)
            self.proj[l] = Spline(g, vnlg[:ndft], None if l == 1 else 0.0)

    def bessel_trans(self, l, fnr):
        """Species::bessel_trans (Species.cc:773-1000): fng(G) = sum_r r^2 j_l(G r) fnr(r)
        on the spline mesh G_i = i pi / (ndft dr), written with the sine and cosine
        transforms of the reference; fng(0) by Simpson for l = 0, zero otherwise."""
        n = self.ndft
        r = self.rps
        g = self.gspl
        fng = np.zeros(n + 1)
        if l == 0:
            f0 = simpsn(fnr * r * r)
            fng[:n] = sinft(fnr * r)
            fng[0] = f0
            fng[1:n] /= g[1:]
        elif l == 1:
            s = sinft(fnr.copy())
            fnint = np.zeros(n)
            fnint[1:] = s[1:] / (g[1:] * g[1:])
            y = np.zeros(n + 1)
            y[:n] = fnr * r
            c = cosft1(y)
            fng[0] = 0.0
            fng[1:n] = fnint[1:] - c[1:n] / g[1:]
        elif l == 2:
            y = np.zeros(n)
            y[1:] = fnr[1:] / r[1:]
            s = sinft(y)
            fnint = np.zeros(n)
            fnint[1:] = 3.0 * s[1:] / (g[1:] ** 3)
            s = sinft(fnr * r)
            fnint[1:] += -s[1:] / g[1:]
            y = np.zeros(n + 1)
            y[:n] = fnr
            c = cosft1(y)
            fnint[1:] += -3.0 * c[1:n] / (g[1:] * g[1:])
            fng[0] = 0.0
            fng[1:n] = fnint[1:]
        elif l == 3:
            y = np.zeros(n)
            y[1:] = fnr[1:] / (r[1:] * r[1:])
            s = sinft(y)
            fnint = np.zeros(n)
            fnint[1:] = 15.0 * s[1:] / (g[1:] ** 4)
            s = sinft(fnr.copy())
            fnint[1:] += -6.0 * s[1:] / (g[1:] * g[1:])
            y = np.zeros(n + 1)
            y[1:n] = fnr[1:] / r[1:]
            c = cosft1(y)
            fnint[1:] += -15.0 * c[1:n] / (g[1:] ** 3)
            y = np.zeros(n + 1)
            y[:n] = fnr * r
            c = cosft1(y)
            fnint[1:] += c[1:n] / g[1:]
            fng[0] = 0.0
            fng[1:n] = fnint[1:]
        else:
            raise ValueError("l > 3")
        return fng

    # Species::vlocg / dvnlg / rhopsg / eself
    def vlocg(self, q):
        q = np.asarray(q, dtype=np.float64)
        v = self.local_spline.value(q)
        return np.where(q > self.gspl[-1], 0.0, v)

    def dvnlg(self, l, q):
        q = np.asarray(q, dtype=np.float64)
        if l == self.llocal:
            return np.zeros_like(q), np.zeros_like(q)
        v, dv = self.proj[l].derivative(q)
        out = q > self.gspl[-1]
        return np.where(out, 0.0, v), np.where(out, 0.0, dv)

    def rhopsg(self, q):
        q = np.asarray(q, dtype=np.float64)
        return -self.zval * np.exp(-0.25 * self.rcps * self.rcps * q * q)

    def eself(self):
        return self.zval * self.zval / (math.sqrt(2.0 * math.pi) * self.rcps)

    def projector_ls(self):
        """l of each KB projector in the reference's order (l ascending, m inside)."""
        return [l for l in range(self.lmax + 1) if l != self.llocal for _ in range(2 * l + 1)]


# ---------------------------------------------------------------------------------------
def twnl_table(sp, kpgx):
    """NonLocalPotential::update_twnl, KB branch, for one species on |k+G(-A)| given by the
    Cartesian components kpgx (ngw, 3) (NonLocalPotential.cc:290-1000): one row per
    projector (l ascending, m as the reference orders it), plus the per-projector l and
    KB weight wsg."""
    kx, ky, kz = kpgx[:, 0], kpgx[:, 1], kpgx[:, 2]
    k2 = kx * kx + ky * ky + kz * kz
    k = np.sqrt(k2)
    ki = np.where(k > 0.0, 1.0 / np.where(k > 0.0, k, 1.0), 0.0)
    fpi = 4.0 * math.pi
    s14pi = math.sqrt(1.0 / fpi)
    s34pi = math.sqrt(3.0 / fpi)
    s54pi = math.sqrt(5.0 / fpi)
    s3 = math.sqrt(3.0)
    s74pi = math.sqrt(7.0 / fpi)
    s2132pi = math.sqrt(21.0 / (32.0 * math.pi))
    s3532pi = math.sqrt(35.0 / (32.0 * math.pi))
    s1054pi = math.sqrt(105.0 / fpi)
    rows, ls, wts = [], [], []
    for l in range(sp.lmax + 1):
        if l == sp.llocal:
            continue
        v, _ = sp.dvnlg(l, k)
        w = sp.wsg[l]
        if l == 0:
            rows.append(s14pi * v)
        elif l == 1:
            ux, uy, uz = kx * ki, ky * ki, kz * ki
            rows += [s34pi * ux * v, s34pi * uy * v, s34pi * uz * v]
        elif l == 2:
            ux, uy, uz = kx * ki, ky * ki, kz * ki
            y4 = s54pi * 0.5 * (3.0 * uz * uz - 1.0)
            y5 = s54pi * 0.5 * s3 * (ux * ux - uy * uy)
            y6 = s54pi * s3 * ux * uy
            y7 = s54pi * s3 * uy * uz
            y8 = s54pi * s3 * uz * ux
            rows += [y4 * v, y5 * v, y6 * v, y7 * v, y8 * v]
        else:
            ux, uy, uz = kx * ki, ky * ki, kz * ki
            y9 = s74pi * 0.5 * uz * (5.0 * uz * uz - 3.0)
            y10 = s2132pi * ux * (5.0 * uz * uz - 1.0)
            y11 = s2132pi * uy * (5.0 * uz * uz - 1.0)
            y12 = s1054pi * ux * uy * uz
            y13 = s1054pi * 0.5 * uz * (ux * ux - uy * uy)
            y14 = s3532pi * ux * (ux * ux - 3.0 * uy * uy)
            y15 = s3532pi * uy * (3.0 * ux * ux - uy * uy)
            rows += [y * v for y in (y9, y10, y11, y12, y13, y14, y15)]
        ls += [l] * (2 * l + 1)
        wts += [w] * (2 * l + 1)
    return np.array(rows), np.array(ls), np.array(wts)


def _harmonic_polys(l, ux, uy, uz):
    """The angular factors of twnl_table's rows of angular momentum l as polynomials Y(u)
    of the direction components, with their partial derivatives (dY/dux, dY/duy, dY/duz)
    taken with the components independent."""
    fpi = 4.0 * math.pi
    z = np.zeros_like(ux)
    if l == 0:
        c = math.sqrt(1.0 / fpi)
        return [(c + z, z, z, z)]
    if l == 1:
        c = math.sqrt(3.0 / fpi)
        o = c + z
        return [(c * ux, o, z, z), (c * uy, z, o, z), (c * uz, z, z, o)]
    if l == 2:
        c = math.sqrt(5.0 / fpi)
        s3 = math.sqrt(3.0)
        return [(c * 0.5 * (3.0 * uz * uz - 1.0), z, z, c * 3.0 * uz),
                (c * 0.5 * s3 * (ux * ux - uy * uy), c * s3 * ux, -c * s3 * uy, z),
                (c * s3 * ux * uy, c * s3 * uy, c * s3 * ux, z),
                (c * s3 * uy * uz, z, c * s3 * uz, c * s3 * uy),
                (c * s3 * uz * ux, c * s3 * uz, z, c * s3 * ux)]
    a = math.sqrt(7.0 / fpi)
    b = math.sqrt(21.0 / (32.0 * math.pi))
    c = math.sqrt(35.0 / (32.0 * math.pi))
    d = math.sqrt(105.0 / fpi)
    return [(a * 0.5 * uz * (5.0 * uz * uz - 3.0), z, z, a * 0.5 * (15.0 * uz * uz - 3.0)),
            (b * ux * (5.0 * uz * uz - 1.0), b * (5.0 * uz * uz - 1.0), z, b * 10.0 * ux * uz),
            (b * uy * (5.0 * uz * uz - 1.0), z, b * (5.0 * uz * uz - 1.0), b * 10.0 * uy * uz),
            (d * ux * uy * uz, d * uy * uz, d * ux * uz, d * ux * uy),
            (d * 0.5 * uz * (ux * ux - uy * uy), d * uz * ux, -d * uz * uy,
             d * 0.5 * (ux * ux - uy * uy)),
            (c * ux * (ux * ux - 3.0 * uy * uy), c * 3.0 * (ux * ux - uy * uy), -c * 6.0 * ux * uy, z),
            (c * uy * (3.0 * ux * ux - uy * uy), c * 6.0 * ux * uy, c * 3.0 * (ux * ux - uy * uy), z)]


def twnl_grad_table(sp, kpgx):
    """Gradient of twnl_table's rows with respect to the vector q = k+G(-A): (3, rows, ngw).
    Each row is v_l(|q|) Y(q/|q|), so d/dq_d = v_l' u_d Y + (v_l/|q|)(dY/du_d - u_d u.grad Y)
    with u = q/|q|; zero where q = 0 (twnl_table's direction factors vanish there).  The
    derivative a vector potential A brings in is -d/dq (q = k+G-A).  Pinned by central
    differences of twnl_table."""
    kx, ky, kz = kpgx[:, 0], kpgx[:, 1], kpgx[:, 2]
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    ki = np.where(k > 0.0, 1.0 / np.where(k > 0.0, k, 1.0), 0.0)
    ux, uy, uz = kx * ki, ky * ki, kz * ki
    u = (ux, uy, uz)
    out = []
    for l in range(sp.lmax + 1):
        if l == sp.llocal:
            continue
        v, dv = sp.dvnlg(l, k)
        for Y, gx, gy, gz in _harmonic_polys(l, ux, uy, uz):
            g = (gx, gy, gz)
            ug = ux * gx + uy * gy + uz * gz
            out.append([dv * u[d] * Y + v * ki * (g[d] - u[d] * ug) for d in range(3)])
    return np.array(out).transpose(1, 0, 2)


def structure_factor(tau, gx):
    """StructureFactor::update (StructureFactor.cc:85-173): sum_a exp(-i G.tau_a)."""
    arg = gx @ np.asarray(tau, dtype=np.float64).T
    return np.sum(np.cos(arg) - 1j * np.sin(arg), axis=1)


def local_tables(species, taus, gx, omega):
    """EnergyFunctional::cell_moved + atoms_moved (EnergyFunctional.cc:1889-1915,
    2001-2021): v_ion(G) = sum_s S_s(G) vlocg_s(|G|) / Omega and
    rho_ps(G) = sum_s S_s(G) rhopsg_s(|G|) / Omega on the density basis."""
    g = np.sqrt(np.sum(gx * gx, axis=1))
    vion = np.zeros(gx.shape[0], dtype=np.complex128)
    rhops = np.zeros(gx.shape[0], dtype=np.complex128)
    for sp, tau in zip(species, taus):
        sf = structure_factor(tau, gx)
        vion += sf * (sp.vlocg(g) / omega)
        rhops += sf * (sp.rhopsg(g) / omega)
    return vion, rhops


_WS = ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, -1, 0), (0, 1, 1), (0, 1, -1),
       (1, 0, 1), (-1, 0, 1), (1, 1, 1), (1, -1, -1), (1, 1, -1), (1, -1, 1))


def ws_planes(a):
    """UnitCell::set (UnitCell.cc:108-138): the 13 translations bounding the Wigner-Seitz
    cell and half their squared lengths."""
    a = np.asarray(a, dtype=np.float64)
    an = np.array([c[0] * a[0] + c[1] * a[1] + c[2] * a[2] for c in _WS])
    return an, 0.5 * np.sum(an * an, axis=1)


def min_wsdist(a):
    """UnitCell::min_wsdist (UnitCell.cc:241-248): the shortest of the 13 translations."""
    _, an2h = ws_planes(a)
    return float(np.min(np.sqrt(2.0 * an2h)))


def auto_rcps(zval, a, ecut):
    """AtomSet::set_rcps (AtomSet.cc:1102-1163): the Gaussian width for valence zval in
    cell a at the wavefunction cutoff ecut (Hartree)."""
    ellmin = min_wsdist(a)
    r = 0.3
    recipmax = ellmin / (1.0 * 2.0 * math.pi)
    while (zval / math.sqrt(math.pi)) * erfc(r * math.sqrt(ecut)) > 1.0e-7 * r / zval:
        r += 0.05
    if r > recipmax:
        r = recipmax
    while zval * erfc(ellmin / (math.sqrt(2.0) * r)) > 1.0e-8:
        r -= 0.05
    return r


def fold_in_ws(v, a):
    """UnitCell::fold_in_ws (UnitCell.cc:251-281)."""
    an, an2h = ws_planes(a)
    v = np.array(v, dtype=np.float64)
    eps = 1.0e-10
    done = False
    it = 0
    while not done and it < 10:
        done = True
        for i in range(13):
            sp = v @ an[i]
            if sp > an2h[i] + eps:
                done = False
                while True:
                    v = v - an[i]
                    if not (v @ an[i] > an2h[i] + eps):
                        break
            elif sp < -an2h[i] - eps:
                done = False
                while True:
                    v = v + an[i]
                    if not (v @ an[i] < -an2h[i] - eps):
                        break
            if not done:
                break
        it += 1
    return v


def esr(a, zv, rcps, taus):
    """The pseudocharge repulsion energy of EnergyFunctional::atoms_moved
    (EnergyFunctional.cc:1918-1988): pairs (l < m within a species, all pairs across),
    each folded into the Wigner-Seitz cell and summed over (2 newald + 1)^3 images with
    newald = int(0.5 * 5 * rckj / min_wsdist + 0.5)."""
    a = np.asarray(a, dtype=np.float64)
    wsd = min_wsdist(a)
    e = 0.0
    nsp = len(zv)
    for k in range(nsp):
        for j in range(k, nsp):
            rckj = math.sqrt(rcps[k] ** 2 + rcps[j] ** 2)
            newald = int(0.5 * 5 * rckj / wsd + 0.5)
            rng = np.arange(-newald, newald + 1)
            imgs = np.array([ux * a[0] + uy * a[1] + uz * a[2]
                             for ux in rng for uy in rng for uz in rng])
            tk, tj = np.asarray(taus[k]), np.asarray(taus[j])
            for l in range(len(tk) - (1 if k == j else 0)):
                m0 = l + 1 if k == j else 0
                for m in range(m0, len(tj)):
                    v = fold_in_ws(tk[l] - tj[m], a)
                    d = np.sqrt(np.sum((v[None, :] + imgs) ** 2, axis=1))
                    e += float(np.sum(zv[k] * zv[j] * erfc(d / rckj) / d))
    return e
