// PLUTO's double-precision checkpoint format (.dbl output and restart) for
// the periodic box of a domain, written and read on the host:
//
//   data.NNNN.dbl  single_file mode, no header: every dumped variable in
//                  turn, native-endian FP64 (write_data.c:111-170).  Cell-
//                  centred variables hold the interior [k][j][i]; the
//                  staggered field bxNs holds one more face along N, the
//                  first one being the face below the domain's first cell
//                  (bin_io.c:216-259 with the -1 base of write_data.c:126-
//                  137), which a periodic box holds as the image of its
//                  last face.  Variables: rho vx1 vx2 vx3 Bx1 Bx2 Bx3 prs
//                  Bx1s Bx2s Bx3s for DIMENSIONS = COMPONENTS = 3, rho vx1
//                  vx2 Bx1 Bx2 prs Bx1s Bx2s for 2 (var_names.c:33-44,
//                  set_output.c:78-96, NVAR order of MHD/mod_defs.h).
//   dbl.out        one text line per file: "%d %12.6e %12.6e %ld ", then
//                  "single_file ", the endianness and the variable names,
//                  each followed by a blank (write_data.c:365-395).
//   restart.out    one 128-byte Restart record per dump (structs.h:286-292:
//                  int nstep, int nfile[16], double t, double dt, fill); the
//                  .dbl output is output type 0 (runtime_setup.c:224-229),
//                  the other 15 entries are -1 (never written); padding and
//                  fill bytes are written as zeros (PLUTO leaves them
//                  unset) (restart.c:260-299).
//
// Reading follows restart.c:17-257: the line of dbl.out gives the endianness
// (files written big-endian are byte-swapped on load), restart.out the
// time, dt and step, the data file the state; the image faces are dropped.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "domain.hpp"

namespace pgpu {
namespace {

constexpr int kOutputTypes = 16;   // MAX_OUTPUT_TYPES (pluto.h:99)

struct RestartRecord {             // layout of structs.h:286-292 on x86-64
  int32_t nstep;
  int32_t nfile[kOutputTypes];
  int32_t pad;
  double t;
  double dt;
  char fill[40];
};
static_assert(sizeof(RestartRecord) == 128, "Restart record is 128 bytes");

struct Box {
  int nx, ny, nz, dims;
  size_t n() const { return (size_t)nx * ny * nz; }
};

// primitive slots dumped, in PLUTO's variable order
std::vector<int> dumped_prims(int dims)
{
  if (dims == 3) return {0, 1, 2, 3, 4, 5, 6, 7};
  return {0, 1, 2, 4, 5, 7};   // COMPONENTS 2: no vx3, Bx3
}

std::string var_list(int dims)
{
  return dims == 3 ? "rho vx1 vx2 vx3 Bx1 Bx2 Bx3 prs Bx1s Bx2s Bx3s "
                   : "rho vx1 vx2 Bx1 Bx2 prs Bx1s Bx2s ";
}

std::string path_of(const char* dir, const char* name)
{
  std::string p = (dir && *dir) ? dir : ".";
  p += "/";
  p += name;
  return p;
}

// number of doubles in a data file of the box
size_t file_doubles(const Box& b)
{
  const size_t nc = b.n();
  const size_t ncell = dumped_prims(b.dims).size() * nc;
  size_t stag = (size_t)b.nz * b.ny * (b.nx + 1) + (size_t)b.nz * (b.ny + 1) * b.nx;
  if (b.dims == 3) stag += (size_t)(b.nz + 1) * b.ny * b.nx;
  return ncell + stag;
}

void swap8(double* p, size_t n)
{
  for (size_t q = 0; q < n; q++) {
    unsigned char* c = reinterpret_cast<unsigned char*>(p + q);
    for (int a = 0; a < 4; a++) {
      const unsigned char t = c[a];
      c[a] = c[7 - a];
      c[7 - a] = t;
    }
  }
}

bool little_endian()
{
  const uint16_t one = 1;
  unsigned char c;
  std::memcpy(&c, &one, 1);
  return c == 1;
}

int check_box(const Box& b)
{
  if (b.dims != 2 && b.dims != 3) return fail(PGPU_EINVAL, "dims must be 2 or 3 (got %d)", b.dims);
  if (b.nx < 1 || b.ny < 1 || b.nz < 1) return fail(PGPU_EINVAL, "empty box");
  if (b.dims == 2 && b.nz != 1) return fail(PGPU_EINVAL, "2D boxes have nz = 1");
  return PGPU_OK;
}

// the file image of a state
// low (optional): the face planes below the first cell, bx [nz][ny], by
// [nz][nx], bz [ny][nx]; without it they are the images of the last faces.
void pack(const Box& b, const double* prim, const double* bx, const double* by, const double* bz,
          std::vector<double>& out, const double* const* low = nullptr)
{
  const size_t nc = b.n();
  out.clear();
  out.reserve(file_doubles(b));
  for (int q : dumped_prims(b.dims)) out.insert(out.end(), prim + q * nc, prim + (q + 1) * nc);
  const int nx = b.nx, ny = b.ny, nz = b.nz;
  auto at = [&](const double* f, int i, int j, int k) { return f[((size_t)k * ny + j) * nx + i]; };
  for (int k = 0; k < nz; k++)           // Bx1s: faces i-1/2 .. nx-1/2
    for (int j = 0; j < ny; j++) {
      out.push_back(low ? low[0][(size_t)k * ny + j] : at(bx, nx - 1, j, k));
      for (int i = 0; i < nx; i++) out.push_back(at(bx, i, j, k));
    }
  for (int k = 0; k < nz; k++)           // Bx2s: faces j-1/2 .. ny-1/2
    for (int j = -1; j < ny; j++)
      for (int i = 0; i < nx; i++)
        out.push_back(j < 0 && low ? low[1][(size_t)k * nx + i] : at(by, i, j < 0 ? ny - 1 : j, k));
  if (b.dims == 3)
    for (int k = -1; k < nz; k++)        // Bx3s: faces k-1/2 .. nz-1/2
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          out.push_back(k < 0 && low ? low[2][(size_t)j * nx + i] : at(bz, i, j, k < 0 ? nz - 1 : k));
}

void unpack(const Box& b, const double* in, double* prim, double* bx, double* by, double* bz,
            double* const* low = nullptr)
{
  const size_t nc = b.n();
  for (int q = 0; q < 8; q++) std::fill(prim + q * nc, prim + (q + 1) * nc, 0.0);
  for (int q : dumped_prims(b.dims)) {
    std::memcpy(prim + q * nc, in, nc * sizeof(double));
    in += nc;
  }
  const int nx = b.nx, ny = b.ny, nz = b.nz;
  auto at = [&](double* f, int i, int j, int k) -> double& { return f[((size_t)k * ny + j) * nx + i]; };
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++) {
      if (low) low[0][(size_t)k * ny + j] = *in;
      in++;                               // the face below i = 0
      for (int i = 0; i < nx; i++) at(bx, i, j, k) = *in++;
    }
  for (int k = 0; k < nz; k++)
    for (int j = -1; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        if (j >= 0) at(by, i, j, k) = *in;
        else if (low) low[1][(size_t)k * nx + i] = *in;
        in++;
      }
  if (b.dims == 3) {
    for (int k = -1; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          if (k >= 0) at(bz, i, j, k) = *in;
          else if (low) low[2][(size_t)j * nx + i] = *in;
          in++;
        }
  } else {
    std::fill(bz, bz + nc, 0.0);
  }
}

int read_lines(const std::string& p, std::vector<std::string>& lines)
{
  lines.clear();
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) return -1;
  std::string cur;
  int c;
  while ((c = std::fgetc(f)) != EOF) {
    cur.push_back((char)c);
    if (c == '\n') {
      lines.push_back(cur);
      cur.clear();
    }
  }
  if (!cur.empty()) lines.push_back(cur);
  std::fclose(f);
  return 0;
}

}  // namespace
}  // namespace pgpu

using namespace pgpu;

static int write_dbl_impl(const char* dir, int nfile, int nx, int ny, int nz, int dims,
                          const double* prim, const double* bx, const double* by, const double* bz,
                          double time, double dt, long nstep, const double* const* low)
{
  const Box b{nx, ny, nz, dims};
  int rc = check_box(b);
  if (rc) return rc;
  if (nfile < 0 || nfile > 9999) return fail(PGPU_EINVAL, "nfile must be in [0, 9999] (got %d)", nfile);
  if (!prim || !bx || !by || (dims == 3 && !bz)) return fail(PGPU_EINVAL, "NULL state array");

  // data.NNNN.dbl
  std::vector<double> img;
  pack(b, prim, bx, by, bz, img, low);
  char name[64];
  std::snprintf(name, sizeof(name), "data.%04d.dbl", nfile);
  const std::string dp = path_of(dir, name);
  FILE* f = std::fopen(dp.c_str(), "wb");
  if (!f) return fail(PGPU_EINVAL, "cannot write %s", dp.c_str());
  const size_t w = std::fwrite(img.data(), sizeof(double), img.size(), f);
  std::fclose(f);
  if (w != img.size()) return fail(PGPU_EINVAL, "short write to %s", dp.c_str());

  // dbl.out: keep the lines of the files before this one, then this line
  const std::string op = path_of(dir, "dbl.out");
  std::vector<std::string> lines;
  if (nfile > 0 && read_lines(op, lines) == 0 && (int)lines.size() > nfile) lines.resize(nfile);
  if (nfile == 0) lines.clear();
  char head[128];
  std::snprintf(head, sizeof(head), "%d %12.6e %12.6e %ld ", nfile, time, dt, nstep);
  std::string line = head;
  line += "single_file ";
  line += little_endian() ? "little " : "big ";
  line += var_list(dims);
  line += "\n";
  lines.push_back(line);
  f = std::fopen(op.c_str(), "wb");
  if (!f) return fail(PGPU_EINVAL, "cannot write %s", op.c_str());
  for (const std::string& l : lines) std::fwrite(l.data(), 1, l.size(), f);
  std::fclose(f);

  // restart.out: record nfile
  RestartRecord r;
  std::memset(&r, 0, sizeof(r));
  r.nstep = (int32_t)nstep;
  for (int t = 0; t < kOutputTypes; t++) r.nfile[t] = -1;
  r.nfile[0] = nfile;
  r.t = time;
  r.dt = dt;
  const std::string rp = path_of(dir, "restart.out");
  f = std::fopen(rp.c_str(), nfile == 0 ? "wb" : "r+b");
  if (!f && nfile > 0) f = std::fopen(rp.c_str(), "wb");
  if (!f) return fail(PGPU_EINVAL, "cannot write %s", rp.c_str());
  std::fseek(f, (long)nfile * (long)sizeof(RestartRecord), SEEK_SET);
  std::fwrite(&r, sizeof(r), 1, f);
  std::fclose(f);
  return PGPU_OK;
}

extern "C" int pgpu_write_dbl_host(const char* dir, int nfile, int nx, int ny, int nz, int dims,
                                   const double* prim, const double* bx, const double* by,
                                   const double* bz, double time, double dt, long nstep)
{
  return write_dbl_impl(dir, nfile, nx, ny, nz, dims, prim, bx, by, bz, time, dt, nstep, nullptr);
}

static int read_dbl_impl(const char* dir, int nfile, int nx, int ny, int nz, int dims, double* prim,
                         double* bx, double* by, double* bz, double* time, double* dt, long* nstep,
                         double* const* low)
{
  const Box b{nx, ny, nz, dims};
  int rc = check_box(b);
  if (rc) return rc;
  if (nfile < 0 || nfile > 9999) return fail(PGPU_EINVAL, "nfile must be in [0, 9999] (got %d)", nfile);
  if (!prim || !bx || !by || !bz) return fail(PGPU_EINVAL, "NULL state array");

  // dbl.out line nfile: mode, endianness, variables
  std::vector<std::string> lines;
  const std::string op = path_of(dir, "dbl.out");
  if (read_lines(op, lines)) return fail(PGPU_EINVAL, "cannot find %s", op.c_str());
  if ((int)lines.size() <= nfile)
    return fail(PGPU_EINVAL, "output #%d does not exist in %s", nfile, op.c_str());
  int nf = -1;
  long ns = 0;
  double tt = 0.0, dd = 0.0;
  char mode[64] = {0}, endian[64] = {0};
  int used = 0;
  if (std::sscanf(lines[nfile].c_str(), "%d %lf %lf %ld %63s %63s %n", &nf, &tt, &dd, &ns, mode, endian,
                  &used) < 6 || nf != nfile)
    return fail(PGPU_EINVAL, "line %d of %s is not a dbl.out entry", nfile, op.c_str());
  if (std::strcmp(mode, "single_file"))
    return fail(PGPU_EINVAL, "only single_file .dbl output is supported (got %s)", mode);
  const bool swap = (!std::strcmp(endian, "big")) == little_endian();
  std::string vars = lines[nfile].substr(used);
  while (!vars.empty() && (vars.back() == '\n' || vars.back() == '\r')) vars.pop_back();
  if (vars != var_list(dims))
    return fail(PGPU_EINVAL, "variables '%s' are not the DIMENSIONS %d MHD + CT set", vars.c_str(), dims);

  // restart.out: the record of this file
  const std::string rp = path_of(dir, "restart.out");
  FILE* f = std::fopen(rp.c_str(), "rb");
  if (!f) return fail(PGPU_EINVAL, "cannot find %s", rp.c_str());
  RestartRecord r;
  bool found = false;
  while (std::fread(&r, sizeof(r), 1, f) == 1) {
    int32_t n0 = r.nfile[0];
    if (swap) {
      unsigned char* c = reinterpret_cast<unsigned char*>(&n0);
      std::swap(c[0], c[3]);
      std::swap(c[1], c[2]);
    }
    if (n0 == nfile) {
      found = true;
      break;
    }
  }
  std::fclose(f);
  if (!found) return fail(PGPU_EINVAL, "no record of output #%d in %s", nfile, rp.c_str());
  if (swap) {
    unsigned char* c = reinterpret_cast<unsigned char*>(&r.nstep);
    std::swap(c[0], c[3]);
    std::swap(c[1], c[2]);
    swap8(&r.t, 1);
    swap8(&r.dt, 1);
  }

  // data file
  char name[64];
  std::snprintf(name, sizeof(name), "data.%04d.dbl", nfile);
  const std::string dp = path_of(dir, name);
  f = std::fopen(dp.c_str(), "rb");
  if (!f) return fail(PGPU_EINVAL, "cannot find %s", dp.c_str());
  std::vector<double> img(file_doubles(b));
  const size_t got = std::fread(img.data(), sizeof(double), img.size(), f);
  const bool extra = std::fgetc(f) != EOF;
  std::fclose(f);
  if (got != img.size() || extra)
    return fail(PGPU_EINVAL, "%s does not hold a %d x %d x %d box", dp.c_str(), nx, ny, nz);
  if (swap) swap8(img.data(), img.size());
  unpack(b, img.data(), prim, bx, by, bz, low);
  if (time) *time = r.t;
  if (dt) *dt = r.dt;
  if (nstep) *nstep = r.nstep;
  return PGPU_OK;
}

extern "C" int pgpu_read_dbl_host(const char* dir, int nfile, int nx, int ny, int nz, int dims,
                                  double* prim, double* bx, double* by, double* bz, double* time,
                                  double* dt, long* nstep)
{
  return read_dbl_impl(dir, nfile, nx, ny, nz, dims, prim, bx, by, bz, time, dt, nstep, nullptr);
}

extern "C" int pgpu_write_dbl(pgpu_domain* d, const char* dir, int nfile)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (d->g.ghost) {
    // a bounded box: the interior and its own faces below the first cells
    const GhostBox& gb = d->gb;
    const long ni = (long)gb.n[0] * gb.n[1] * gb.n[2];
    std::vector<double> prim((size_t)8 * ni), f((size_t)3 * ni);
    std::vector<double> lo((size_t)gb.n[1] * gb.n[2] + (size_t)gb.n[0] * gb.n[2] +
                           (size_t)gb.n[0] * gb.n[1]);
    double* low[3] = {lo.data(), lo.data() + (size_t)gb.n[1] * gb.n[2],
                      lo.data() + (size_t)gb.n[1] * gb.n[2] + (size_t)gb.n[0] * gb.n[2]};
    int rc = pgpu_get_state(d, prim.data(), f.data(), f.data() + ni, f.data() + 2 * ni);
    if (!rc) rc = pgpu_ghosted_low_faces(d, low);
    double s[6];
    if (!rc) rc = pgpu_get_scalars(d, s);
    if (rc) return rc;
    const double* cl[3] = {low[0], low[1], low[2]};
    return write_dbl_impl(dir, nfile, gb.n[0], gb.n[1], gb.n[2], 3, prim.data(), f.data(),
                          f.data() + ni, f.data() + 2 * ni, s[3], s[0], (long)s[4], cl);
  }
  const Geo& g = d->g;
  std::vector<double> prim((size_t)8 * g.n), f((size_t)3 * g.n);
  int rc = pgpu_get_state(d, prim.data(), f.data(), f.data() + g.n, f.data() + 2 * g.n);
  if (rc) return rc;
  double s[6];
  rc = pgpu_get_scalars(d, s);
  if (rc) return rc;
  return pgpu_write_dbl_host(dir, nfile, g.nx, g.ny, g.nz, g.dims, prim.data(), f.data(),
                             f.data() + g.n, f.data() + 2 * g.n, s[3], s[0], (long)s[4]);
}

extern "C" int pgpu_read_dbl(pgpu_domain* d, const char* dir, int nfile)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  const Geo& g = d->g;
  double t = 0.0, dt = 0.0;
  long ns = 0;
  int rc;
  if (d->g.ghost) {
    const GhostBox& gb = d->gb;
    const long ni = (long)gb.n[0] * gb.n[1] * gb.n[2];
    std::vector<double> prim((size_t)8 * ni), f((size_t)3 * ni);
    std::vector<double> lo((size_t)gb.n[1] * gb.n[2] + (size_t)gb.n[0] * gb.n[2] +
                           (size_t)gb.n[0] * gb.n[1]);
    double* low[3] = {lo.data(), lo.data() + (size_t)gb.n[1] * gb.n[2],
                      lo.data() + (size_t)gb.n[1] * gb.n[2] + (size_t)gb.n[0] * gb.n[2]};
    rc = read_dbl_impl(dir, nfile, gb.n[0], gb.n[1], gb.n[2], 3, prim.data(), f.data(),
                       f.data() + ni, f.data() + 2 * ni, &t, &dt, &ns, low);
    if (rc) return rc;
    if (!(dt > 0.0)) return fail(PGPU_EINVAL, "restart record of output #%d has dt <= 0", nfile);
    const double* faces[3] = {f.data(), f.data() + ni, f.data() + 2 * ni};
    const double* cl[3] = {low[0], low[1], low[2]};
    rc = pgpu_ghosted_set_state(d, prim.data(), faces, cl);
  } else {
    std::vector<double> prim((size_t)8 * g.n), f((size_t)3 * g.n);
    rc = pgpu_read_dbl_host(dir, nfile, g.nx, g.ny, g.nz, g.dims, prim.data(), f.data(),
                            f.data() + g.n, f.data() + 2 * g.n, &t, &dt, &ns);
    if (rc) return rc;
    if (!(dt > 0.0)) return fail(PGPU_EINVAL, "restart record of output #%d has dt <= 0", nfile);
    rc = pgpu_set_state(d, prim.data(), f.data(), f.data() + g.n, f.data() + 2 * g.n);
  }
  if (rc) return rc;
  // the clock of the restarted run (RestartGet: g_time, g_dt, g_stepNumber);
  // the CFL settings of the last pgpu_set_dt stay
  DevScalars h;
  std::memset(&h, 0, sizeof(h));
  h.dt = dt;
  h.time = t;
  h.step = ns;
  if (hipMemcpyAsync(d->sc, &h, sizeof(h), hipMemcpyHostToDevice, d->stream) != hipSuccess ||
      hipStreamSynchronize(d->stream) != hipSuccess)
    return fail(PGPU_EDEVICE, "restoring the step clock failed");
  return PGPU_OK;
}
