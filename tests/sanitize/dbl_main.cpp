// AddressSanitizer / UndefinedBehaviorSanitizer run of the library's .dbl
// host code (pluto_amd/csrc/dbl_io.hip, compiled with the host pass
// sanitized): PLUTO .dbl files written and read back for 3D and 2D boxes of
// odd sizes, several outputs per directory (dbl.out, restart.out appended).
// tests/test_sanitize_cpu.py builds and runs it.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

extern "C" int pgpu_write_dbl_host(const char* dir, int nfile, int nx, int ny, int nz, int dims,
                                   const double* prim, const double* bx, const double* by,
                                   const double* bz, double time, double dt, long nstep);
extern "C" int pgpu_read_dbl_host(const char* dir, int nfile, int nx, int ny, int nz, int dims,
                                  double* prim, double* bx, double* by, double* bz, double* time,
                                  double* dt, long* nstep);

int main(int argc, char** argv)
{
  if (argc < 2) return 2;
  static const int shapes[][4] = {{5, 4, 3, 3}, {7, 6, 1, 2}, {3, 3, 3, 3}, {1, 9, 2, 3}};
  int q = 0;
  for (const auto& s : shapes) {
    char dir[4096];
    snprintf(dir, sizeof dir, "%s/box%d", argv[1], q++);
    char cmd[4200];
    snprintf(cmd, sizeof cmd, "mkdir -p %s", dir);
    if (system(cmd) != 0) return 2;
    const long n = (long)s[0] * s[1] * s[2];
    std::vector<double> p(8 * n), bx(n), by(n), bz(n), p2(8 * n), bx2(n), by2(n), bz2(n);
    for (long c = 0; c < 8 * n; c++) p[c] = 0.25 * (double)c - 3.0;
    for (long c = 0; c < n; c++) {
      bx[c] = (double)c;
      by[c] = -(double)c;
      bz[c] = s[3] == 3 ? 2.0 * (double)c : 0.0;
    }
    if (s[3] == 2)
      for (long c = 0; c < n; c++) p[3 * n + c] = p[6 * n + c] = 0.0;
    for (int f = 0; f < 3; f++)
      if (pgpu_write_dbl_host(dir, f, s[0], s[1], s[2], s[3], p.data(), bx.data(), by.data(),
                              bz.data(), 0.1 * f, 1e-3, f) != 0)
        return 3;
    double t = 0, dt = 0;
    long ns = 0;
    if (pgpu_read_dbl_host(dir, 2, s[0], s[1], s[2], s[3], p2.data(), bx2.data(), by2.data(),
                           bz2.data(), &t, &dt, &ns) != 0)
      return 4;
    if (memcmp(p.data(), p2.data(), 8 * n * sizeof(double)) != 0 ||
        memcmp(bx.data(), bx2.data(), n * sizeof(double)) != 0 ||
        memcmp(by.data(), by2.data(), n * sizeof(double)) != 0 || ns != 2 || t != 0.2)
      return 5;
    // a missing output is an error, not a read past the end
    if (pgpu_read_dbl_host(dir, 7, s[0], s[1], s[2], s[3], p2.data(), bx2.data(), by2.data(),
                           bz2.data(), &t, &dt, &ns) == 0)
      return 6;
  }
  printf("sanitized dbl run: ok\n");
  return 0;
}

// The device-domain entry points dbl_io.hip also serves (pgpu_write_dbl,
// pgpu_read_dbl); the host round trip above never calls them.
struct pgpu_domain;
namespace pgpu {
int fail(int code, const char*, ...) { return code; }
}
extern "C" int pgpu_get_state(pgpu_domain*, double*, double*, double*, double*) { abort(); }
extern "C" int pgpu_set_state(pgpu_domain*, const double*, const double*, const double*,
                              const double*) { abort(); }
extern "C" int pgpu_get_scalars(pgpu_domain*, double*) { abort(); }
int pgpu_ghosted_low_faces(pgpu_domain*, double* const*) { abort(); }
int pgpu_ghosted_set_state(pgpu_domain*, const double*, const double* const*,
                           const double* const*) { abort(); }
