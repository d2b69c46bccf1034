// Wavefunction files.
//   raw: the per-task block of Wavefunction::write_fast / write_dump
//        (src/qball/Wavefunction.cc:2221-2404): the local c() block, mloc x nloc complex
//        doubles column-major, then eig[nst] and occ[nst] when the caller passes them (qball
//        writes them when there are empty states).  One process row: mloc = ld.
//   xml: SlaterDet::write (src/qball/SlaterDet.cc:2925-3343) for one process row, base64
//        encoding: a <slater_determinant> element with the diagonal density matrix (the
//        occupations) and, per state, the orbital on the basis' own grid as a base64
//        <grid_function> (real part at Γ, complex values otherwise), 72 characters a line.
// Reading an xml element transforms each grid function back to the plane-wave sphere.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/qbh.h"
#include "qbh_internal.h"

namespace qbh {
Basis* basis_of(qbh_basis b);
Ctx* ctx_of_basis(qbh_basis b);
}  // namespace qbh

using namespace qbh;

namespace {

const char kB64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

// RFC 4648 base64 with '=' padding
std::string b64_encode(const unsigned char* p, size_t n) {
  std::string out;
  out.reserve(4 * ((n + 2) / 3));
  size_t i = 0;
  for (; i + 3 <= n; i += 3) {
    const unsigned v = (p[i] << 16) | (p[i + 1] << 8) | p[i + 2];
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += kB64[(v >> 6) & 63];
    out += kB64[v & 63];
  }
  if (n - i == 2) {
    const unsigned v = (p[i] << 16) | (p[i + 1] << 8);
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += kB64[(v >> 6) & 63];
    out += '=';
  } else if (n - i == 1) {
    const unsigned v = p[i] << 16;
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += "==";
  }
  return out;
}

// decodes base64, skipping white space, into at most `cap` bytes; returns the bytes
// written, or cap + 1 when the text holds more than cap bytes (nothing past cap is written)
size_t b64_decode(const char* s, size_t n, unsigned char* out, size_t cap) {
  int val[256];
  for (int i = 0; i < 256; ++i) val[i] = -1;
  for (int i = 0; i < 64; ++i) val[(unsigned char)kB64[i]] = i;
  size_t o = 0;
  int q[4], nq = 0, pad = 0;
  for (size_t i = 0; i < n; ++i) {
    const unsigned char c = s[i];
    if (c <= ' ') continue;
    if (c == '=') {
      q[nq++] = 0;
      ++pad;
    } else if (val[c] >= 0) {
      q[nq++] = val[c];
    } else {
      break;
    }
    if (nq == 4) {
      const unsigned v = (q[0] << 18) | (q[1] << 12) | (q[2] << 6) | q[3];
      const size_t k = pad >= 2 ? 1 : pad == 1 ? 2 : 3;
      if (o + k > cap) return cap + 1;
      out[o++] = (v >> 16) & 255;
      if (pad < 2) out[o++] = (v >> 8) & 255;
      if (pad < 1) out[o++] = v & 255;
      nq = 0;
      if (pad) break;
    }
  }
  return o;
}

int iofail(Ctx* c, int code, const std::string& m) {
  if (c) c->last_error = m;
  return code;
}

std::string read_file(const char* path, long offset, bool& ok) {
  ok = false;
  FILE* f = std::fopen(path, "rb");
  if (!f) return {};
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  if (offset < 0 || offset > size) {
    std::fclose(f);
    return {};
  }
  std::string s((size_t)(size - offset), '\0');
  std::fseek(f, offset, SEEK_SET);
  const size_t got = std::fread(&s[0], 1, s.size(), f);
  std::fclose(f);
  ok = got == s.size();
  return s;
}

}  // namespace

extern "C" int qbh_wf_write_raw(const char* path, int append, const double* psi_dev, long ld,
                                int nst, const double* eig, const double* occ) {
  if (!path || !psi_dev || ld <= 0 || nst <= 0) return QBH_ERR_ARG;
  std::vector<double> h(2 * (size_t)ld * nst);
  if (hipMemcpy(h.data(), psi_dev, h.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return QBH_ERR_DEVICE;
  FILE* f = std::fopen(path, append ? "ab" : "wb");
  if (!f) return QBH_ERR_ARG;
  bool ok = std::fwrite(h.data(), sizeof(double), h.size(), f) == h.size();
  if (ok && eig && occ) {
    ok = std::fwrite(eig, sizeof(double), nst, f) == (size_t)nst &&
         std::fwrite(occ, sizeof(double), nst, f) == (size_t)nst;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? QBH_OK : QBH_ERR_ARG;
}

extern "C" int qbh_wf_read_raw(const char* path, long offset, double* psi_dev, long ld, int nst,
                               double* eig, double* occ, long* next_offset) {
  if (!path || !psi_dev || ld <= 0 || nst <= 0 || offset < 0) return QBH_ERR_ARG;
  FILE* f = std::fopen(path, "rb");
  if (!f) return QBH_ERR_ARG;
  std::vector<double> h(2 * (size_t)ld * nst);
  bool ok = std::fseek(f, offset, SEEK_SET) == 0 &&
            std::fread(h.data(), sizeof(double), h.size(), f) == h.size();
  long used = (long)(h.size() * sizeof(double));
  if (ok && eig && occ) {
    ok = std::fread(eig, sizeof(double), nst, f) == (size_t)nst &&
         std::fread(occ, sizeof(double), nst, f) == (size_t)nst;
    used += 2L * nst * (long)sizeof(double);
  }
  std::fclose(f);
  if (!ok) return QBH_ERR_ARG;
  if (hipMemcpy(psi_dev, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return QBH_ERR_DEVICE;
  if (next_offset) *next_offset = offset + used;
  return QBH_OK;
}

extern "C" int qbh_sd_write_xml(qbh_basis basis, const char* path, int append,
                                const double* psi_dev, long ld, int nst, const double* occ,
                                double weight, int ispin, int nspin) {
  if (!basis || !path || !psi_dev || !occ || nst <= 0) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  const GridDev& g = b->g;
  if (ld < g.ngw) return iofail(c, QBH_ERR_ARG, "ld < ngw");
  const long N = (long)g.np0 * g.np1 * g.np2;
  double* f = nullptr;
  if (hipMalloc(&f, 2 * sizeof(double) * N) != hipSuccess)
    return iofail(c, QBH_ERR_ALLOC, "grid buffer allocation failed");
  FILE* out = std::fopen(path, append ? "ab" : "wb");
  if (!out) {
    hipFree(f);
    return iofail(c, QBH_ERR_ARG, std::string("cannot open ") + path);
  }
  // header (SlaterDet.cc:2942-2970)
  std::ostringstream hdr;
  hdr << "<slater_determinant";
  if (nspin == 2) hdr << " spin=\"" << (ispin > 0 ? "down" : "up") << "\"";
  hdr << " kpoint=\"" << b->kpoint[0] << " " << b->kpoint[1] << " " << b->kpoint[2] << "\"\n"
      << "  weight=\"" << std::setprecision(12) << weight << "\""
      << " size=\"" << nst << "\">" << std::endl;
  hdr << "<density_matrix form=\"diagonal\" size=\"" << nst << "\">" << std::endl;
  hdr.setf(std::ios::fixed, std::ios::floatfield);
  hdr.setf(std::ios::right, std::ios::adjustfield);
  for (int i = 0; i < nst; ++i) {
    hdr << " " << std::setprecision(8) << occ[i];
    if (i % 10 == 9) hdr << std::endl;
  }
  if (nst % 10 != 0) hdr << std::endl;
  hdr << "</density_matrix>" << std::endl;
  bool ok = std::fputs(hdr.str().c_str(), out) >= 0;
  const bool real = g.real;
  std::vector<double> grid(2 * (size_t)N), vals(real ? N : 2 * N);
  int r = QBH_OK;
  for (int n = 0; n < nst && ok && r == QBH_OK; ++n) {
    r = qbh_fft_backward(basis, psi_dev + 2 * (size_t)ld * n, ld, 1, 0, f, 0);
    if (r) break;
    hipStreamSynchronize(c->stream);
    if (hipMemcpy(grid.data(), f, grid.size() * sizeof(double), hipMemcpyDeviceToHost) !=
        hipSuccess) {
      r = iofail(c, QBH_ERR_DEVICE, "grid copy failed");
      break;
    }
    if (real)
      for (long i = 0; i < N; ++i) vals[i] = grid[2 * i];
    else
      vals = grid;
    std::ostringstream gs;
    gs << "<grid_function type=\"" << (real ? "double" : "complex") << "\""
       << " nx=\"" << g.np0 << "\" ny=\"" << g.np1 << "\" nz=\"" << g.np2 << "\""
       << " encoding=\"base64\">" << std::endl;
    const std::string enc =
        b64_encode(reinterpret_cast<const unsigned char*>(vals.data()), vals.size() * sizeof(double));
    // Base64Transcoder::print: 72 characters a line, then the rest and a newline
    const size_t nl = enc.size() / 72;
    for (size_t i = 0; i < nl; ++i) gs << enc.substr(72 * i, 72) << '\n';
    gs << enc.substr(72 * nl) << '\n';
    gs << "</grid_function>\n";
    ok = std::fputs(gs.str().c_str(), out) >= 0;
  }
  if (ok && r == QBH_OK) ok = std::fputs("</slater_determinant>\n", out) >= 0;
  ok = (std::fclose(out) == 0) && ok;
  hipFree(f);
  if (r) return r;
  return ok ? QBH_OK : iofail(c, QBH_ERR_ARG, "write failed");
}

extern "C" int qbh_sd_read_xml(qbh_basis basis, const char* path, long offset, double* psi_dev,
                               long ld, int nst, double* occ, long* next_offset) {
  if (!basis || !path || !psi_dev || nst <= 0 || offset < 0) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  const GridDev& g = b->g;
  if (ld < g.ngw) return iofail(c, QBH_ERR_ARG, "ld < ngw");
  bool okf = false;
  const std::string s = read_file(path, offset, okf);
  if (!okf) return iofail(c, QBH_ERR_ARG, std::string("cannot read ") + path);
  const size_t sd0 = s.find("<slater_determinant");
  if (sd0 == std::string::npos) return iofail(c, QBH_ERR_ARG, "no <slater_determinant>");
  const size_t sd1 = s.find("</slater_determinant>", sd0);
  if (sd1 == std::string::npos) return iofail(c, QBH_ERR_ARG, "unterminated <slater_determinant>");
  if (occ) {
    const size_t d0 = s.find("<density_matrix", sd0);
    const size_t d1 = s.find("</density_matrix>", d0);
    if (d0 == std::string::npos || d1 == std::string::npos || d1 > sd1)
      return iofail(c, QBH_ERR_ARG, "no <density_matrix>");
    std::istringstream is(s.substr(s.find('>', d0) + 1, d1 - s.find('>', d0) - 1));
    for (int i = 0; i < nst; ++i)
      if (!(is >> occ[i])) return iofail(c, QBH_ERR_ARG, "short density matrix");
  }
  const long N = (long)g.np0 * g.np1 * g.np2;
  const bool real = g.real;
  std::vector<double> vals(real ? N : 2 * N), grid(2 * (size_t)N, 0.0);
  double* f = nullptr;
  if (hipMalloc(&f, 2 * sizeof(double) * N) != hipSuccess)
    return iofail(c, QBH_ERR_ALLOC, "grid buffer allocation failed");
  size_t pos = sd0;
  int r = QBH_OK;
  for (int n = 0; n < nst && r == QBH_OK; ++n) {
    const size_t g0 = s.find("<grid_function", pos);
    if (g0 == std::string::npos || g0 > sd1) {
      r = iofail(c, QBH_ERR_ARG, "missing <grid_function>");
      break;
    }
    const size_t gt = s.find('>', g0);
    const size_t g1 = gt == std::string::npos ? gt : s.find("</grid_function>", gt);
    if (g1 == std::string::npos || g1 > sd1) {
      r = iofail(c, QBH_ERR_ARG, "unterminated <grid_function>");
      break;
    }
    const std::string tag = s.substr(g0, gt - g0);
    std::ostringstream want;
    want << "nx=\"" << g.np0 << "\" ny=\"" << g.np1 << "\" nz=\"" << g.np2 << "\"";
    if (tag.find(want.str()) == std::string::npos || tag.find("encoding=\"base64\"") == std::string::npos ||
        tag.find(real ? "type=\"double\"" : "type=\"complex\"") == std::string::npos) {
      r = iofail(c, QBH_ERR_UNSUPPORTED, "grid_function does not match the basis grid / type");
      break;
    }
    const size_t cap = vals.size() * sizeof(double);
    const size_t nb = b64_decode(s.data() + gt + 1, g1 - gt - 1,
                                 reinterpret_cast<unsigned char*>(vals.data()), cap);
    if (nb != vals.size() * sizeof(double)) {
      r = iofail(c, QBH_ERR_ARG, "grid_function size differs from the basis grid");
      break;
    }
    double* dst = psi_dev + 2 * (size_t)ld * n;
    hipStreamSynchronize(c->stream);  // the previous band's transform still reads f
    if (real) {
      if (hipMemcpy(f, vals.data(), N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        r = iofail(c, QBH_ERR_DEVICE, "grid copy failed");
        break;
      }
      r = qbh_fft_forward_real(basis, f, 1.0, dst, 0);
    } else {
      if (hipMemcpy(f, vals.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        r = iofail(c, QBH_ERR_DEVICE, "grid copy failed");
        break;
      }
      r = qbh_fft_forward(basis, f, 1, 0, dst, ld, 0);
    }
    pos = g1;
  }
  hipStreamSynchronize(c->stream);
  hipFree(f);
  if (r) return r;
  if (next_offset) *next_offset = offset + (long)(sd1 + std::strlen("</slater_determinant>\n"));
  return QBH_OK;
}
