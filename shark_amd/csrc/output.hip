// output.hip — host-side writer of the reference's output directories (include/shark_output.h).
//
// The layout and the formatted lines of info.dat follow `output` (src/commons/write_output.f90:
// 2-300) and `title` (:302-330): one stream-unformatted file of FP64 values per field over the
// active cells, in cell order, scaled to physical units in the reference's operation order.
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <string>
#include <vector>

#include "../../include/shark_output.h"
#include "growth.h"
#include "output.h"

namespace shark {

// Fortran Ew.d (no Ee) as written by flang: [-]0.<d digits>E+dd, or [-]0.<d digits>+ddd when the
// decimal exponent needs three digits, right-justified in w characters.
std::string fortran_e(double v, int w, int d) {
  std::string s;
  if (!isfinite(v)) {
    s = isnan(v) ? "NaN" : (v < 0 ? "-Infinity" : "Infinity");
  } else {
    const bool neg = signbit(v);
    int ex = 0;
    std::string digits;
    if (v == 0.0) {
      digits.assign((size_t)d, '0');
    } else {
      char buf[64];
      snprintf(buf, sizeof(buf), "%.*e", d - 1, fabs(v));   // correctly rounded d significant digits
      digits.push_back(buf[0]);
      const char* p = buf + 2;
      while (*p && *p != 'e') digits.push_back(*p++);
      ex = atoi(p + 1) + 1;
    }
    s = neg ? "-0." : "0.";
    s += digits;
    char eb[16];
    const int ae = ex < 0 ? -ex : ex;
    if (ae <= 99) snprintf(eb, sizeof(eb), "E%c%02d", ex < 0 ? '-' : '+', ae);
    else snprintf(eb, sizeof(eb), "%c%03d", ex < 0 ? '-' : '+', ae);
    s += eb;
  }
  if ((int)s.size() < w) s = std::string((size_t)(w - (int)s.size()), ' ') + s;
  return s;
}

static int write_field(const std::string& dir, const char* name, const std::vector<double>& v,
                       std::string* err) {
  const std::string path = dir + "/" + name;
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) {
    *err = "cannot open " + path + ": " + strerror(errno);
    return 1;
  }
  const size_t n = fwrite(v.data(), sizeof(double), v.size(), f);
  const int rc = fclose(f);
  if (n != v.size() || rc != 0) {
    *err = "short write to " + path;
    return 1;
  }
  return 0;
}

int write_output_dir(const shark_output_info* info, int iout, const OutputGrid& g, const double* q,
                     const double* sdust, const double* position, const double* radii_c, std::string* err) {
  if (iout < 0 || iout > 99999) {
    *err = "iout must lie in 0..99999 (five-digit directory number)";
    return 1;
  }
  char nchar[8];
  snprintf(nchar, sizeof(nchar), "%05d", iout);
  std::string dir = info->base_dir && info->base_dir[0] ? std::string(info->base_dir) + "/" : std::string();
  dir += std::string("output_") + nchar;
  if (mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) {
    *err = "cannot create " + dir + ": " + strerror(errno);
    return 1;
  }
  // info.dat (write_output.f90:20-41)
  {
    const std::string path = dir + "/info.dat";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) {
      *err = "cannot open " + path + ": " + strerror(errno);
      return 1;
    }
    const int nvar = 5 + 4 * g.nd;
    fprintf(f, "nvar        =%11d\n", nvar);
    fprintf(f, "ndust        =%11d\n", g.nd);
    fprintf(f, "ndustpscal   =%11d\n", 0);
    fprintf(f, "NX        =%11d\n", g.nx);
    fprintf(f, "NY        =%11d\n", g.ny);
    fprintf(f, "MHD        =%11d\n", 0);
    fprintf(f, "GEOM        =%11d\n", g.geom);
    fprintf(f, "TURB        =%11d\n", 0);
    fprintf(f, "GRIDSPACE       =%11d\n", info->gridspace);
    fprintf(f, "time        =%s\n", fortran_e(info->time, 23, 15).c_str());
    fprintf(f, "unit_t        =%s\n", fortran_e(info->unit_t, 23, 15).c_str());
    fprintf(f, "unit_d        =%s\n", fortran_e(info->unit_d, 23, 15).c_str());
    fprintf(f, "unit_l        =%s\n", fortran_e(info->unit_l, 23, 15).c_str());
    fprintf(f, "unit_v        =%s\n", fortran_e(info->unit_v, 23, 15).c_str());
    fprintf(f, "unit_p        =%s\n", fortran_e(info->unit_p, 23, 15).c_str());
    fprintf(f, "charging       =%11d\n", 0);
    if (fclose(f) != 0) {
      *err = "short write to " + path;
      return 1;
    }
  }
  // active cells in cell order (iy outer, ix inner)
  std::vector<int64_t> cells;
  const int y0 = g.ndim == 2 ? 3 : 1, y1 = g.ndim == 2 ? g.ny + 2 : 1;
  for (int iy = y0; iy <= y1; ++iy)
    for (int ix = 3; ix <= g.nx + 2; ++ix) cells.push_back((int64_t)(ix - 1) + (int64_t)(iy - 1) * g.nxm);
  const int64_t nc = g.nc;
  const bool polar = g.geom == 2;
  const double ud = polar ? info->unit_dcol : info->unit_d;
  std::vector<double> v(cells.size());
  auto plane = [&](int var) { return q + (int64_t)var * nc; };
  auto put_scaled = [&](const double* src, double unit) {
    for (size_t i = 0; i < cells.size(); ++i) v[i] = src[cells[i]] * unit;
  };
  // x: position(:,1), or radii_c in polar geometry (:43-56)
  {
    const double* src = polar ? radii_c : position;
    for (size_t i = 0; i < cells.size(); ++i) v[i] = src[cells[i]];
    if (write_field(dir, "x", v, err)) return 1;
  }
  put_scaled(plane(0), ud);
  if (write_field(dir, "rho", v, err)) return 1;
  put_scaled(plane(1), info->unit_v);
  if (write_field(dir, "v", v, err)) return 1;
  if (g.ndim == 2) {
    for (size_t i = 0; i < cells.size(); ++i) v[i] = position[nc + cells[i]];
    if (write_field(dir, "y", v, err)) return 1;
  }
  put_scaled(plane(2), info->unit_v);
  if (write_field(dir, "vy", v, err)) return 1;
  put_scaled(plane(3), info->unit_v);
  if (write_field(dir, "vz", v, err)) return 1;
  // P: q*unit_P, times unit_l in polar geometry (:102-115)
  for (size_t i = 0; i < cells.size(); ++i) {
    double p = plane(4)[cells[i]] * info->unit_p;
    if (polar) p = p * info->unit_l;
    v[i] = p;
  }
  if (write_field(dir, "P", v, err)) return 1;
  if (g.nd > 0) {
    const int nd = g.nd;
    // rhod_tot: species summed in order, each scaled first (:157-176)
    for (size_t i = 0; i < cells.size(); ++i) {
      double t = 0.0;
      for (int k = 0; k < nd; ++k) t = t + plane(5 + k)[cells[i]] * ud;
      v[i] = t;
    }
    if (write_field(dir, "rhod_tot", v, err)) return 1;
    std::vector<double> w(cells.size() * (size_t)nd);
    auto species = [&](int base, const double* src_all, double unit, const char* name) {
      for (int k = 0; k < nd; ++k) {
        const double* src = src_all ? src_all + (int64_t)k * nc : plane(base + k);
        for (size_t i = 0; i < cells.size(); ++i) w[(size_t)k * cells.size() + i] = src[cells[i]] * unit;
      }
      return write_field(dir, name, w, err);
    };
    if (species(5, nullptr, ud, "rhod")) return 1;
    if (!sdust) {
      *err = "sdust is required to write the dust outputs";
      return 1;
    }
    if (species(0, sdust, info->unit_l, "sd")) return 1;
    if (species(5 + nd, nullptr, info->unit_v, "vd")) return 1;
    if (species(5 + 2 * nd, nullptr, info->unit_v, "vdy")) return 1;
    if (species(5 + 3 * nd, nullptr, info->unit_v, "vdz")) return 1;
  }
  return 0;
}

int output_info_ok(const shark_output_info* info, std::string* err) {
  if (!info) {
    *err = "null output info";
    return 0;
  }
  if (info->version != SHARK_OUTPUT_VERSION) {
    *err = "shark_output_info.version does not match SHARK_OUTPUT_VERSION";
    return 0;
  }
  if (info->mhd || info->turb || info->ndustpscal || info->charging) {
    *err = "MHD, turbulence, dust passive scalars and charging outputs are not built";
    return 0;
  }
  return 1;
}

}  // namespace shark

extern "C" int shark_output_write(const shark_output_info* info, int iout, const shark_cfg* cfg, const double* q,
                                  const double* sdust, const double* position, const double* radii_c) {
  std::string err;
  if (!shark::output_info_ok(info, &err)) return shark::fail_msg(SHARK_EINVAL, err.c_str());
  if (!cfg || !q || !radii_c || !position)
    return shark::fail_msg(SHARK_EINVAL, "shark_output_write: cfg, q, position and radii_c are required");
  shark::OutputGrid g;
  g.nx = cfg->nx;
  g.ny = cfg->ny > 1 ? cfg->ny : 1;
  g.ndim = cfg->ny > 1 ? 2 : 1;
  g.nd = cfg->ndust;
  g.nxm = cfg->nx + 4;
  g.geom = cfg->geom;
  g.nc = (int64_t)g.nxm * (g.ndim == 2 ? cfg->ny + 4 : 1);
  if (shark::write_output_dir(info, iout, g, q, sdust, position, radii_c, &err))
    return shark::fail_msg(SHARK_EIO, err.c_str());
  return SHARK_OK;
}
