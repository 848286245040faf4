// sph2d_output.hip -- the reference's on-disk formats, written from host arrays in
// the reference's Fortran layout (host code only; the context-level wrappers in
// sph2d_capi.hip download the device state and call these).
//
// Fortran edit descriptors are reproduced exactly: Fw.d (fixed), Ew.dEe (0.ddd
// mantissa, 3-digit exponent), list-directed integers (width 12, followed by one
// blank before the next character item, as in the shipped Free_Surface.dat and
// Final_Doamin_Results.dat), and list-directed lines begin with one blank.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sph2d_internal.h"

namespace {

int out_fail(int code, const std::string& msg) { return sph2d::set_error(code, msg); }

// Fortran Fw.d: right-justified fixed notation, asterisks when it does not fit,
// NaN / Infinity / -Infinity right-justified.
std::string fmt_f(double v, int w, int d) {
  char buf[512];
  if (std::isnan(v)) {
    std::snprintf(buf, sizeof buf, "%*s", w, "NaN");
  } else if (std::isinf(v)) {
    std::snprintf(buf, sizeof buf, "%*s", w, v > 0 ? "Infinity" : "-Infinity");
  } else {
    std::snprintf(buf, sizeof buf, "%#*.*f", w, d, v);   // # keeps the point when d = 0
  }
  std::string s(buf);
  if ((int)s.size() > w) s.assign(w, '*');
  return s;
}

// Fortran Ew.dEe with e = 3: [-]0.d1..dd E+eee, right-justified in w. The d
// significant digits are the correctly rounded ones (%.{d-1}e gives the same).
std::string fmt_e(double v, int w, int d) {
  char buf[512];
  if (std::isnan(v) || std::isinf(v)) return fmt_f(v, w, 0);
  std::string mant;
  int ex = 0;
  const bool neg = std::signbit(v) && v != 0.0;
  if (v == 0.0) {
    mant.assign(d, '0');
  } else {
    std::snprintf(buf, sizeof buf, "%.*e", d - 1, std::fabs(v));   // D.DDDDe+XX
    const char* p = buf;
    mant.push_back(*p++);
    if (*p == '.') ++p;
    while (*p && *p != 'e') mant.push_back(*p++);
    ex = std::atoi(p + 1) + 1;
  }
  std::snprintf(buf, sizeof buf, "%s0.%sE%c%03d", neg ? "-" : "", mant.c_str(), ex < 0 ? '-' : '+', std::abs(ex));
  std::string s(buf);
  if ((int)s.size() < w) s.insert(0, w - s.size(), ' ');
  if ((int)s.size() > w) s.assign(w, '*');
  return s;
}

// write(char_time,'(fW.D)') t followed by trim(adjustl(char_time))
std::string time_string(double t, int w, int d) {
  std::string s = fmt_f(t, w, d);
  const size_t a = s.find_first_not_of(' ');
  return a == std::string::npos ? std::string() : s.substr(a);
}

// A list-directed REAL(8) item as ifort writes it, 23 characters: the G23.15E3 choice
// of form on the value rounded to 15 significant digits -- F18.(15-k) and 5 blanks
// when 0.1 <= |x| < 1e15 (k digits before the point), else the E form with scale
// 1P (one digit before the point, 15 after, 3-digit exponent), right-justified.
// Zero takes the E form (0.000000000000000E+000).
std::string ld_real8(double v) {
  char buf[96];
  if (std::isnan(v) || std::isinf(v)) return fmt_f(v, 23, 0);
  const double a = std::fabs(v);
  if (a != 0.0) {
    std::snprintf(buf, sizeof buf, "%.14e", a);                 // the 15-digit rounding
    const int e10 = std::atoi(std::strchr(buf, 'e') + 1);
    if (e10 >= -1 && e10 < 15) {
      const int k = e10 + 1;
      std::snprintf(buf, sizeof buf, "%#18.*f     ", 15 - k, v);
      return buf;
    }
  }
  std::snprintf(buf, sizeof buf, "%.15e", a);                   // D.DDDDDDDDDDDDDDDe+XX
  char* e = std::strchr(buf, 'e');
  const int ex = std::atoi(e + 1);
  *e = 0;
  char item[64], out[64];
  std::snprintf(item, sizeof item, "%s%sE%c%03d", std::signbit(v) && v != 0.0 ? "-" : "", buf, ex < 0 ? '-' : '+',
                std::abs(ex));
  std::snprintf(out, sizeof out, "%23s", item);
  return out;
}

// List-directed integer item followed by the next character item.
std::string ld_int(int64_t v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%12lld ", (long long)v);
  return buf;
}

// A list-directed record sequence: items are appended to the current record and a
// character item that would pass column 80 starts a new record with a blank (the
// shipped Initial_Post_Proceeding_Grid.dat breaks its ZONE line this way).
struct LdRecords {
  std::string cur = " ";
  std::string out;
  void item(const std::string& s) {
    if (cur.size() > 1 && cur.size() + s.size() > 80) {
      out += cur + "\n";
      cur = " ";
    }
    cur += s;
  }
  std::string done() { return out + cur + "\n"; }
};

// The quadrilaterals of Distribute_Grid_For_Postproceeding.f90:47-58 (format 8I8).
void grid_elements(FILE* f, int nex, int ney) {
  const int nnx = nex + 1;
  char b[64];
  for (int i = 1; i <= ney; ++i)
    for (int j = 1; j <= nex; ++j) {
      const int k = j + (i - 1) * nex;
      const int L = k + (i - 1);
      std::snprintf(b, sizeof b, "%8d%8d%8d%8d\n", L, L + 1, L + nnx + 1, L + nnx);
      std::fputs(b, f);
    }
}

bool file_is_empty(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return true;
  const int c = std::fgetc(f);
  std::fclose(f);
  return c == EOF;
}

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() { if (f) std::fclose(f); }
  File(const File&) = delete;
  File& operator=(const File&) = delete;
};

int open_err(const char* path) {
  return out_fail(SPH2D_ERR_ARGUMENT, std::string("cannot open ") + (path ? path : "(null)") + ": " +
                                          std::strerror(errno));
}

bool ok_scales(const sph2d_scales* s) {
  return s && s->domain_size_y != 0.0 && s->velocity_0 != 0.0 && s->pressure_0 != 0.0;
}

// Write a string of the Tecplot binary header: each character as INT32, then 0
// (function_module.f90:85-101, Write_ASCII_to_BINARY).
void bin_string(FILE* f, const std::string& s) {
  for (unsigned char ch : s) {
    const int32_t v = ch;
    std::fwrite(&v, 4, 1, f);
  }
  const int32_t z = 0;
  std::fwrite(&z, 4, 1, f);
}
void bin_i32(FILE* f, int32_t v) { std::fwrite(&v, 4, 1, f); }
void bin_f32(FILE* f, float v) { std::fwrite(&v, 4, 1, f); }

}  // namespace

extern "C" {

int sph2d_format_fortran(double value, char kind, int32_t w, int32_t d, char* out, int32_t cap) {
  if (!out || cap <= w || w <= 0 || w > 400 || d < 0 || d >= w) return out_fail(SPH2D_ERR_ARGUMENT, "bad format");
  std::string s;
  if (kind == 'F') s = fmt_f(value, w, d);
  else if (kind == 'E') s = fmt_e(value, w, d);
  else return out_fail(SPH2D_ERR_ARGUMENT, "kind must be 'F' or 'E'");
  std::memcpy(out, s.c_str(), s.size() + 1);
  return SPH2D_OK;
}

// Parallel_Taylor_Green_Vortices_2D_Release.f90:185-199 (Final_Doamin_Results.dat).
int sph2d_write_final_results(const char* path, int32_t n, const double* pos, const double* vel,
                              const double* rho, const double* press, int32_t ld) {
  if (!path || n < 0 || !pos || !vel || !rho || !press || ld < n) return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  File F(path, "w");
  if (!F.f) return open_err(path);
  std::fprintf(F.f, " TITLE='DISTRIBUTION'\n");
  std::fprintf(F.f, " VARIABLES= 'X' 'Y' 'VX' 'VY' 'RHO' 'P' \n");
  std::fprintf(F.f, " ZONE I=%s F=POINT\n", ld_int(n).c_str());
  std::string line;
  for (int i = 0; i < n; ++i) {
    line.clear();
    for (double v : {pos[i], pos[ld + i], vel[i], vel[ld + i], rho[i], press[i]}) line += fmt_f(v, 20, 10);
    std::fprintf(F.f, "%s\n", line.c_str());
  }
  return SPH2D_OK;
}

// Output_for_tecplot.f90:70-137: one ZONE of video_result.dat (time_iteration.f90:147-151
// writes the header when the file is opened).
int sph2d_write_tecplot_zone(const char* path, double current_time, int32_t n, const double* pos,
                             const double* vel, const double* press, const double* vort_z, const double* ana,
                             int32_t ld, const sph2d_scales* sc) {
  if (!path || n < 0 || !pos || !vel || !press || !vort_z || !ana || ld < n || !ok_scales(sc))
    return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  const bool fresh = file_is_empty(path);
  File F(path, "a");
  if (!F.f) return open_err(path);
  if (fresh) {
    std::fprintf(F.f, " TITLE='DISTRIBUTION'\n");
    std::fprintf(F.f, " VARIABLES= 'X' 'Y' 'VX' 'VY' 'P' 'U' 'Vor' 'EVX' 'EVY' 'EP' 'EU' \n");
  }
  std::fprintf(F.f, " ZONE T=\" %s \", I=%s F=POINT\n", time_string(current_time, 12, 5).c_str(), ld_int(n).c_str());
  const double L = sc->domain_size_y, U = sc->velocity_0, P0 = sc->pressure_0, T0 = sc->time_0;
  std::string line;
  for (int i = 0; i < n; ++i) {
    const double u = vel[i], v = vel[ld + i];
    const double vm = std::sqrt(u * u + v * v);                 // DOT_PRODUCT form
    const double eu = std::log10(std::fabs((u - ana[i]) / U));
    const double ev = std::log10(std::fabs((v - ana[ld + i]) / U));
    const double ep = std::log10(std::fabs((press[i] - ana[2 * (size_t)ld + i]) / P0));
    const double em = std::log10(std::fabs((vm - ana[3 * (size_t)ld + i]) / U));
    line.clear();
    for (double x : {pos[i] / L, pos[ld + i] / L, u / U, v / U, press[i] / P0, vm / U, vort_z[i] * T0, eu, ev, ep, em})
      line += fmt_f(x, 20, 8);
    std::fprintf(F.f, "%s\n", line.c_str());
  }
  return SPH2D_OK;
}

// Output_for_tecplot.f90:141-158: one ZONE of Free_Surface.dat (header of
// time_iteration.f90:137-140 when the file is new); pos/normal of the surface
// particles only, in the order given.
int sph2d_write_surface_zone(const char* path, double current_time, int32_t n, const double* pos,
                             const double* normal, int32_t ld, const sph2d_scales* sc) {
  if (!path || n < 0 || (n > 0 && (!pos || !normal)) || ld < n || !ok_scales(sc))
    return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  const bool fresh = file_is_empty(path);
  File F(path, "a");
  if (!F.f) return open_err(path);
  if (fresh) {
    std::fprintf(F.f, " TITLE='DISTRIBUTION'\n");
    std::fprintf(F.f, " VARIABLES= 'X' 'Y' 'Nor_X' 'Nor_Y' \n");
  }
  std::fprintf(F.f, " ZONE T=\" %s \", I=%s F=POINT\n", time_string(current_time, 12, 5).c_str(), ld_int(n).c_str());
  const double L = sc->domain_size_y;
  std::string line;
  for (int i = 0; i < n; ++i) {
    line.clear();
    for (double x : {pos[i] / L, pos[ld + i] / L, normal[i], normal[ld + i]}) line += fmt_f(x, 20, 8);
    std::fprintf(F.f, "%s\n", line.c_str());
  }
  return SPH2D_OK;
}

// Output_for_tecplot.f90:162-291: the #!TDV101 binary file of one output step
// (point-packed, 6 double variables), native byte order as the reference writes it.
int sph2d_write_tecplot_binary(const char* path, double time, int32_t n, const double* pos, const double* vel,
                               const double* press, int32_t ld, const sph2d_scales* sc) {
  if (!path || n < 0 || !pos || !vel || !press || ld < n || !ok_scales(sc))
    return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  File F(path, "wb");
  if (!F.f) return open_err(path);
  FILE* f = F.f;
  const std::string t = time_string(time, 8, 4);
  std::fwrite("#!TDV101", 1, 8, f);
  bin_i32(f, 1);
  bin_string(f, "TITLE='" + t + "'");
  bin_i32(f, 6);
  for (const char* name : {"X", "Y", "VX", "VY", "P", "U"}) bin_string(f, name);
  bin_f32(f, 299.0f);                                  // zone marker
  bin_string(f, "ZONE " + t);
  bin_i32(f, -1);                                      // zone colour
  bin_i32(f, 0);                                       // ORDERED
  bin_i32(f, 1);                                       // point packing
  bin_i32(f, 0);                                       // variable location not specified
  bin_i32(f, 0);                                       // no face neighbour connections
  bin_i32(f, n); bin_i32(f, 1); bin_i32(f, 1);         // IMax, JMax, KMax
  bin_i32(f, 0);                                       // no auxiliary data
  bin_f32(f, 357.0f);                                  // EOHMARKER
  bin_f32(f, 299.0f);
  for (int k = 0; k < 6; ++k) bin_i32(f, 2);           // double
  bin_i32(f, 0);                                       // no variable sharing
  bin_i32(f, -1);                                      // no connectivity sharing
  const double L = sc->domain_size_y, U = sc->velocity_0, P0 = sc->pressure_0;
  std::vector<double> row(6);
  for (int i = 0; i < n; ++i) {
    const double u = vel[i], v = vel[ld + i];
    const double vm = std::sqrt(u * u + v * v);        // sqrt(u**2+v**2)
    row[0] = pos[i] / L; row[1] = pos[ld + i] / L; row[2] = u / U; row[3] = v / U;
    row[4] = press[i] / P0; row[5] = vm / U;
    std::fwrite(row.data(), 8, 6, f);
  }
  return std::ferror(f) ? out_fail(SPH2D_ERR_ARGUMENT, std::string("write failed: ") + path) : SPH2D_OK;
}

// Output_for_paraview.f90:188-263: VTK XML UnstructuredGrid, ASCII, points only.
int sph2d_write_paraview(const char* path, int32_t n, const double* pos, const double* vel, const double* press,
                         int32_t ld, const sph2d_scales* sc) {
  if (!path || n < 0 || !pos || !vel || !press || ld < n || !ok_scales(sc))
    return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  File F(path, "w");
  if (!F.f) return open_err(path);
  FILE* f = F.f;
  const double L = sc->domain_size_y, U = sc->velocity_0, P0 = sc->pressure_0;
  const std::string zero = fmt_e(0.0, 20, 9);
  std::fprintf(f, "<?xml version=\"1.0\"?>\n");
  std::fprintf(f, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"BigEndian\">\n");
  std::fprintf(f, "<UnstructuredGrid>\n");
  std::fprintf(f, " <Piece NumberOfPoints=\"%s\" NumberOfCells=\"%s\" >\n", ld_int(n).c_str(), ld_int(0).c_str());
  std::fprintf(f, "<Points>\n");
  std::fprintf(f, "<DataArray name=\"Position\" type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int i = 0; i < n; ++i)
    std::fprintf(f, "%s%s%s\n", fmt_e(pos[i] / L, 20, 9).c_str(), fmt_e(pos[ld + i] / L, 20, 9).c_str(), zero.c_str());
  std::fprintf(f, "</DataArray>\n</Points>\n");
  std::fprintf(f, "<PointData Scalars=\"Pressure\" Vectors=\"Velocity\" >\n");
  std::fprintf(f, "<DataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int i = 0; i < n; ++i)
    std::fprintf(f, "%s%s%s\n", fmt_e(vel[i] / U, 20, 9).c_str(), fmt_e(vel[ld + i] / U, 20, 9).c_str(), zero.c_str());
  std::fprintf(f, "</DataArray>\n");
  std::fprintf(f, "<DataArray type=\"Float64\" Name=\"Pressure\" format=\"ascii\">\n");
  for (int i = 0; i < n; ++i) std::fprintf(f, "%s\n", fmt_e(press[i] / P0, 20, 9).c_str());
  std::fprintf(f, "</DataArray>\n");
  std::fprintf(f, "</PointData>\n");
  std::fprintf(f, "<Cells>\n");
  std::fprintf(f, "<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n</DataArray>\n");
  std::fprintf(f, "<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n</DataArray>\n");
  std::fprintf(f, "<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n</DataArray>\n");
  std::fprintf(f, "</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  return SPH2D_OK;
}

// time_iteration.f90:143-144 (header) and :495-498 (one ZONE, format 9F20.8,3I):
// the restart record read back by Distribute_Latest_Particle.f90.
int sph2d_write_latest_results_zone(const char* path, int32_t i_time_step, int32_t n, const double* pos,
                                    const double* vel, const double* press, const double* rho,
                                    const double* mass, const double* energy, const double* c,
                                    const int32_t* initial_type, const int32_t* wavemaker_layer, int32_t ld) {
  if (!path || n < 0 || !pos || !vel || !press || !rho || !mass || !energy || !c || !initial_type || ld < n)
    return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  const bool fresh = file_is_empty(path);
  File F(path, "a");
  if (!F.f) return open_err(path);
  if (fresh) {
    std::fprintf(F.f, " TITLE='DISTRIBUTION'\n");
    std::fprintf(F.f, " VARIABLES= 'X' 'Y' 'VX' 'VY' 'P' 'Rho' 'Mass' 'Ene' 'C' 'Type' 'Wmak_L' 'Step' \n");
  }
  std::fprintf(F.f, " ZONE I=%s F=POINT\n", ld_int(n).c_str());
  std::string line;
  char ib[64];
  for (int i = 0; i < n; ++i) {
    line.clear();
    for (double v : {pos[i], pos[ld + i], vel[i], vel[ld + i], press[i], rho[i], mass[i], energy[i], c[i]})
      line += fmt_f(v, 20, 8);
    std::snprintf(ib, sizeof ib, "%12d%12d%12d", initial_type[i], wavemaker_layer ? wavemaker_layer[i] : 0,
                  i_time_step);
    line += ib;
    std::fprintf(F.f, "%s\n", line.c_str());
  }
  return SPH2D_OK;
}

// Distribute_Latest_Particle.f90:60-215: the zone to restart from (zone >= 1, or 0
// for the last one), stepping back one zone while the chosen one holds a NaN;
// returns its particles and the step it was saved at.
int sph2d_read_latest_results(const char* path, int32_t zone, int32_t capacity, int32_t* n_out, double* pos,
                              double* vel, double* press, double* rho, double* mass, double* energy, double* c,
                              int32_t* initial_type, int32_t* wavemaker_layer, int32_t* step_out, int32_t ld) {
  if (!path || !n_out || capacity < 0 || ld < capacity || zone < 0) return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  File F(path, "r");
  if (!F.f) return open_err(path);
  // Zones: offsets of their ZONE lines and their counts.
  std::vector<long> zpos;
  std::vector<long> zcnt;
  std::vector<char> buf(1 << 16);
  while (true) {
    const long at = std::ftell(F.f);
    if (!std::fgets(buf.data(), (int)buf.size(), F.f)) break;
    const char* p = buf.data();
    while (*p == ' ') ++p;
    if (std::strncmp(p, "ZONE", 4) == 0) {
      const char* q = std::strstr(p, "I=");
      if (!q) return out_fail(SPH2D_ERR_ARGUMENT, "ZONE line without I=");
      zpos.push_back(at);
      zcnt.push_back(std::strtol(q + 2, nullptr, 10));
    }
  }
  if (zpos.empty()) return out_fail(SPH2D_ERR_ARGUMENT, "no ZONE in the file");
  int z = zone == 0 ? (int)zpos.size() : zone;
  if (z > (int)zpos.size()) return out_fail(SPH2D_ERR_ARGUMENT, "zone beyond the last saved one");
  std::vector<double> row(9);
  for (; z >= 1; --z) {
    const long cnt = zcnt[z - 1];
    if (cnt > capacity) return out_fail(SPH2D_ERR_CAPACITY, "zone holds more particles than capacity");
    std::fseek(F.f, zpos[z - 1], SEEK_SET);
    if (!std::fgets(buf.data(), (int)buf.size(), F.f)) return out_fail(SPH2D_ERR_ARGUMENT, "truncated file");
    bool has_nan = false;
    int step = 0;
    for (long i = 0; i < cnt; ++i) {
      if (!std::fgets(buf.data(), (int)buf.size(), F.f)) return out_fail(SPH2D_ERR_ARGUMENT, "truncated zone");
      char* p = buf.data();
      for (int k = 0; k < 9; ++k) {
        char* e = nullptr;
        row[k] = std::strtod(p, &e);
        if (e == p) return out_fail(SPH2D_ERR_ARGUMENT, "bad number in zone");
        p = e;
        if (std::isnan(row[k])) has_nan = true;
      }
      const long t = std::strtol(p, &p, 10);
      const long wl = std::strtol(p, &p, 10);
      step = (int)std::strtol(p, &p, 10);
      if (pos) { pos[i] = row[0]; pos[ld + i] = row[1]; }
      if (vel) { vel[i] = row[2]; vel[ld + i] = row[3]; }
      if (press) press[i] = row[4];
      if (rho) rho[i] = row[5];
      if (mass) mass[i] = row[6];
      if (energy) energy[i] = row[7];
      if (c) c[i] = row[8];
      if (initial_type) initial_type[i] = (int32_t)t;
      if (wavemaker_layer) wavemaker_layer[i] = (int32_t)wl;
    }
    if (!has_nan) {
      *n_out = (int32_t)cnt;
      if (step_out) *step_out = step;
      return SPH2D_OK;
    }
  }
  return out_fail(SPH2D_ERR_NONFINITE, "every zone up to the requested one holds a NaN");
}

// Sampling_subroutine.f90:360-461 (format 310 = 20F20.8): one line appended to
// each of Kinetic_energy.dat, Maximum_velocity_magnitude.dat and the four Norm_*
// files in dir, from the 14 values of sph2d_sample.
int sph2d_write_sampling(const char* dir, double nondim_time, const double* out14) {
  if (!dir || !out14) return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  struct Item { const char* name; int first, count; };
  const Item items[] = {{"Kinetic_energy.dat", 0, 1},         {"Maximum_velocity_magnitude.dat", 1, 1},
                        {"Norm_Horizontal_velocity.dat", 2, 3}, {"Norm_Vertical_velocity.dat", 5, 3},
                        {"Norm_Pressure.dat", 8, 3},            {"Norm_Velocity_magnitude.dat", 11, 3}};
  for (const Item& it : items) {
    const std::string path = std::string(dir) + "/" + it.name;
    File F(path.c_str(), "a");
    if (!F.f) return open_err(path.c_str());
    std::string line = fmt_f(nondim_time, 20, 8);
    for (int k = 0; k < it.count; ++k) line += fmt_f(out14[it.first + k], 20, 8);
    std::fprintf(F.f, "%s\n", line.c_str());
  }
  return SPH2D_OK;
}

int sph2d_format_list_directed(double value, char* out, int32_t cap) {
  if (!out) return out_fail(SPH2D_ERR_ARGUMENT, "out is NULL");
  const std::string s = ld_real8(value);
  if ((int)s.size() >= cap) return out_fail(SPH2D_ERR_ARGUMENT, "cap too small");
  std::memcpy(out, s.c_str(), s.size() + 1);
  return SPH2D_OK;
}

// write(unit,*) of n REAL(8) values appended to path (Sampling_subroutine.f90:283,
// :287: the wave gauge files and Wave_front_results.dat): a leading blank, the items
// separated by one blank, a new record (again with a leading blank) when the next
// item would pass column 80.
int sph2d_write_list_directed(const char* path, int32_t n, const double* values) {
  if (!path || n < 0 || (n > 0 && !values)) return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  std::string rec = " ", all;
  for (int k = 0; k < n; ++k) {
    const std::string it = ld_real8(values[k]);
    if (rec.size() > 1 && rec.size() + 1 + it.size() > 80) {
      all += rec + "\n";
      rec = " ";
    }
    if (rec.size() > 1) rec += " ";
    rec += it;
  }
  all += rec + "\n";
  File F(path, "a");
  if (!F.f) return open_err(path);
  std::fputs(all.c_str(), F.f);
  return SPH2D_OK;
}

// write(unit,310) with 310 format(20F20.8) (Sampling_subroutine.f90:318-322, the
// pressure probe files): n values appended as records of at most 20.
int sph2d_write_f20_8(const char* path, int32_t n, const double* values) {
  if (!path || n < 0 || (n > 0 && !values)) return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  std::string all, rec;
  for (int k = 0; k < n; ++k) {
    rec += fmt_f(values[k], 20, 8);
    if ((k + 1) % 20 == 0 || k + 1 == n) {
      all += rec + "\n";
      rec.clear();
    }
  }
  if (n == 0) all = "\n";
  File F(path, "a");
  if (!F.f) return open_err(path);
  std::fputs(all.c_str(), F.f);
  return SPH2D_OK;
}

// Sampling_subroutine.f90:300-322: the scalar Kalman filter of each sampling point's
// MLS pressure (Q = 1.0e-6, R = 1.0e-5 as REAL(4) literals, x0 = 1). On the first
// sampling (current_time_number = i_time_step / sampling_time_step = 1) the output
// is the measurement and the state is (x0, measurement); later the gain
// K = P / (P + R) updates x and P. last_x, last_p hold the state between calls.
int sph2d_kalman_filter(int32_t n, int32_t current_time_number, const double* measured, double* last_x,
                        double* last_p, double* filtered) {
  if (n < 0 || (n > 0 && (!measured || !last_x || !last_p || !filtered)))
    return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  const double Q = (double)1.0e-6f, R = (double)1.0e-5f, x0 = 1.0;
  for (int m = 0; m < n; ++m) {
    if (current_time_number == 1) {
      filtered[m] = measured[m];
      last_x[m] = x0;
      last_p[m] = measured[m];
    } else {
      const double K = last_p[m] / (last_p[m] + R);
      filtered[m] = last_x[m] + K * (measured[m] - last_x[m]);
      last_x[m] = filtered[m];
      last_p[m] = last_p[m] - K * last_p[m] + Q;
    }
  }
  return SPH2D_OK;
}

// Initialize_SPH_Calculation.f90:385-398: n_points evenly spaced points from first
// to last (first + (i-1) (last - first) / (n_points - 1)), as columns x, y of out (ld,2).
int sph2d_sampling_line_points(int32_t n_points, const double* first, const double* last, double* out, int32_t ld) {
  if (n_points < 2 || !first || !last || !out || ld < n_points) return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  for (int j = 0; j < 2; ++j) {
    const double step = (last[j] - first[j]) / (n_points - 1);
    for (int i = 0; i < n_points; ++i) out[(size_t)j * ld + i] = first[j] + i * step;
  }
  return SPH2D_OK;
}

// Initialize_SPH_Calculation.f90:405-436: dir/<line>th_SamplingLine_Position.dat.
int sph2d_write_sampling_line_positions(const char* dir, int32_t line, int32_t n_points, const double* pos,
                                        int32_t ld) {
  if (!dir || !pos || n_points < 0 || ld < n_points) return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  const std::string path = std::string(dir) + "/" + std::to_string(line) + "th_SamplingLine_Position.dat";
  File F(path.c_str(), "w");
  if (!F.f) return open_err(path.c_str());
  std::fputs(" TITLE='DISTRIBUTION'\n VARIABLES= 'X' 'Y'\n", F.f);
  for (int i = 0; i < n_points; ++i)
    std::fprintf(F.f, "%s%s\n", fmt_f(pos[i], 12, 6).c_str(), fmt_f(pos[(size_t)ld + i], 12, 6).c_str());
  return SPH2D_OK;
}

// Sampling_subroutine.f90:490-512: dir/<line>th_<t, F7.4>_SamplingLine_Results.dat,
// format 20f12.6: x, y, analytic u/U, v/U, p/p0, MLS u/U, v/U, MLS p/p0, average p/p0.
int sph2d_write_sampling_line_results(const char* dir, int32_t line, double t_nondim, int32_t n_points,
                                      const double* pos, const double* ana_vel, const double* ana_press,
                                      const double* vel, const double* mls_press, const double* average_press,
                                      int32_t ld, const sph2d_scales* sc) {
  if (!dir || !pos || !ana_vel || !ana_press || !vel || !mls_press || !average_press || n_points < 0 ||
      ld < n_points || !ok_scales(sc))
    return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  const std::string path = std::string(dir) + "/" + std::to_string(line) + "th_" + time_string(t_nondim, 7, 4) +
                           "_SamplingLine_Results.dat";
  File F(path.c_str(), "w");
  if (!F.f) return open_err(path.c_str());
  const size_t L = ld;
  for (int i = 0; i < n_points; ++i) {
    const double v[9] = {pos[i],
                         pos[L + i],
                         ana_vel[i] / sc->velocity_0,
                         ana_vel[L + i] / sc->velocity_0,
                         ana_press[i] / sc->pressure_0,
                         vel[i] / sc->velocity_0,
                         vel[L + i] / sc->velocity_0,
                         mls_press[i] / sc->pressure_0,
                         average_press[i] / sc->pressure_0};
    std::string rec;
    for (double x : v) rec += fmt_f(x, 12, 6);
    std::fprintf(F.f, "%s\n", rec.c_str());
  }
  return SPH2D_OK;
}

// Distribute_Grid_For_Postproceeding.f90:60-95: Initial_Post_Proceeding_Grid.dat and
// Initial_Post_Proceeding_Node.dat of the (nex+1) x (ney+1) element nodes.
int sph2d_write_initial_grid(const char* grid_path, const char* node_path, int32_t nex, int32_t ney,
                             double element_dx) {
  if ((!grid_path && !node_path) || nex < 1 || ney < 1) return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  const int nnx = nex + 1, nny = ney + 1, nn = nnx * nny;
  auto nodes = [&](FILE* f) {
    std::string line;
    for (int i = 0; i < nny; ++i)
      for (int j = 0; j < nnx; ++j) {
        line = fmt_f(j * element_dx, 20, 10) + fmt_f(i * element_dx, 20, 10);
        std::fprintf(f, "%s\n", line.c_str());
      }
  };
  if (grid_path) {
    File F(grid_path, "w");
    if (!F.f) return open_err(grid_path);
    std::fprintf(F.f, " TITLE='mesh'\n VARIABLES= 'X' 'Y'\n");
    LdRecords z;
    z.item("ZONE T= '0' ,N=");
    z.item(ld_int(nn));
    z.item(",E=");
    z.item(ld_int((int64_t)nex * ney));
    z.item(",DATAPACKING=POINT,ZONETYPE=FEQUADRILATERAL ");
    std::fputs(z.done().c_str(), F.f);
    nodes(F.f);
    grid_elements(F.f, nex, ney);
  }
  if (node_path) {
    File F(node_path, "w");
    if (!F.f) return open_err(node_path);
    std::fprintf(F.f, " TITLE='Node'\n VARIABLES= 'X' 'Y'\n");
    std::fprintf(F.f, " ZONE I=%s F=POINT\n", ld_int(nn).c_str());
    nodes(F.f);
  }
  return SPH2D_OK;
}

// Output_for_Grid_PostProceding.f90:110-138: one ZONE of Post_Proceeding_Grid.dat
// (header of time_iteration.f90:153-157 when the file is new), 20F20.8 rows and
// the element connectivity.
int sph2d_write_grid_zone(const char* path, double current_time, int32_t nex, int32_t ney, double element_dx,
                          const double* vel, const double* press, const double* vort_z, const double* fai,
                          const double* ana, int32_t ld, const sph2d_scales* sc) {
  const int nnx = nex + 1, nny = ney + 1, nn = nnx * nny;
  if (!path || nex < 1 || ney < 1 || !vel || !press || !vort_z || !fai || !ana || ld < nn || !ok_scales(sc))
    return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  const bool fresh = file_is_empty(path);
  File F(path, "a");
  if (!F.f) return open_err(path);
  if (fresh) {
    std::fprintf(F.f, " TITLE='mesh'\n");
    std::fprintf(F.f, " VARIABLES= 'X' 'Y' 'VX' 'VY' 'P' 'U' 'Vor' 'EVX' 'EVY' 'EP' 'EU' 'Fai'\n");
  }
  LdRecords z;
  z.item("ZONE T=\" " + time_string(current_time, 12, 5) + " \",N=");
  z.item(ld_int(nn));
  z.item(",E=");
  z.item(ld_int((int64_t)nex * ney));
  z.item(",DATAPACKING=POINT,ZONETYPE=FEQUADRILATERAL ");
  std::fputs(z.done().c_str(), F.f);
  const double L = sc->domain_size_y, U = sc->velocity_0, P0 = sc->pressure_0, T0 = sc->time_0;
  std::string line;
  for (int k = 0; k < nn; ++k) {
    const double ex = (k % nnx) * element_dx, ey = (k / nnx) * element_dx;
    const double u = vel[k], v = vel[ld + k];
    const double vm = std::sqrt(u * u + v * v);
    const double eu = std::log10(std::fabs((u - ana[k]) / U));
    const double ev = std::log10(std::fabs((v - ana[ld + k]) / U));
    const double ep = std::log10(std::fabs((press[k] - ana[2 * (size_t)ld + k]) / P0));
    const double em = std::log10(std::fabs((vm - ana[3 * (size_t)ld + k]) / U));
    line.clear();
    for (double x : {ex / L, ey / L, u / U, v / U, press[k] / P0, vm / U, vort_z[k] * T0, eu, ev, ep, em, fai[k]})
      line += fmt_f(x, 20, 8);
    std::fprintf(F.f, "%s\n", line.c_str());
  }
  grid_elements(F.f, nex, ney);
  return SPH2D_OK;
}

// Output_for_Grid_PostProceding.f90:140-291: the #!TDV101 file of the grid
// (ordered zone nnx x nny, point-packed doubles X Y VX VY P U Fai).
int sph2d_write_grid_binary(const char* path, double time, int32_t nex, int32_t ney, double element_dx,
                            const double* vel, const double* press, const double* fai, int32_t ld,
                            const sph2d_scales* sc) {
  const int nnx = nex + 1, nny = ney + 1, nn = nnx * nny;
  if (!path || nex < 1 || ney < 1 || !vel || !press || !fai || ld < nn || !ok_scales(sc))
    return out_fail(SPH2D_ERR_ARGUMENT, "bad argument");
  File F(path, "wb");
  if (!F.f) return open_err(path);
  FILE* f = F.f;
  const std::string t = time_string(time, 8, 4);
  std::fwrite("#!TDV101", 1, 8, f);
  bin_i32(f, 1);
  bin_string(f, "TITLE='" + t + "'");
  bin_i32(f, 7);
  for (const char* name : {"X", "Y", "VX", "VY", "P", "U", "Fai"}) bin_string(f, name);
  bin_f32(f, 299.0f);
  bin_string(f, "ZONE " + t);
  bin_i32(f, -1); bin_i32(f, 0); bin_i32(f, 1); bin_i32(f, 0); bin_i32(f, 0);
  bin_i32(f, nnx); bin_i32(f, nny); bin_i32(f, 1);
  bin_i32(f, 0);
  bin_f32(f, 357.0f);
  bin_f32(f, 299.0f);
  for (int k = 0; k < 7; ++k) bin_i32(f, 2);
  bin_i32(f, 0);
  bin_i32(f, -1);
  const double L = sc->domain_size_y, U = sc->velocity_0, P0 = sc->pressure_0;
  double row[7];
  for (int k = 0; k < nn; ++k) {
    const double u = vel[k], v = vel[ld + k];
    row[0] = ((k % nnx) * element_dx) / L; row[1] = ((k / nnx) * element_dx) / L;
    row[2] = u / U; row[3] = v / U; row[4] = press[k] / P0; row[5] = std::sqrt(u * u + v * v) / U;
    row[6] = fai[k];
    std::fwrite(row, 8, 7, f);
  }
  return std::ferror(f) ? out_fail(SPH2D_ERR_ARGUMENT, std::string("write failed: ") + path) : SPH2D_OK;
}

}  // extern "C"
