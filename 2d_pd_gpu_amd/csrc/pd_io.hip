// pd_io.hip -- mesh input and result files in the reference's text formats
// (host code only).
//
//   readMeshFile                  src/SetupParticleSystem.cpp:11-60  -> pd_read_mesh
//   storeVelocity                 include/util.h:20-39               -> pd_write_velocity
//   outputLocalParticlesPositions src/SetupParticleSystem.cpp:646-658 -> pd_write_positions
//   outputGatheredPositions       src/SetupParticleSystem.cpp:552-627 -> pd_write_gathered
//
// Numbers go through a C++ stream exactly as the reference writes them:
// default stream precision (6 significant digits) unless the caller asks for
// more (17 reproduces every double).
#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "pd_context.h"

extern "C" long pd_read_mesh(const char* path, int reference_notch, double* X, double* V, long cap)
{
    if (!path) return fail(nullptr, PD_ERR_INVALID, "path is NULL"), -1;
    std::ifstream in(path);
    if (!in) return fail(nullptr, PD_ERR_INVALID, "cannot open mesh file %s", path), -1;
    std::string line;
    std::getline(in, line);  // header "Center_X Center_Y Area"
    long n = 0;
    while (std::getline(in, line)) {
        std::istringstream iss(line);
        long double x, y, area;
        if (!(iss >> x >> y >> area)) continue;
        // the reference cuts its two edge notches by dropping particles (:33)
        if (reference_notch && std::fabs(x) > 17.5L && std::fabs(y) < 2.0L) continue;
        if (X || V) {
            if (n >= cap) return fail(nullptr, PD_ERR_INVALID, "mesh %s: more than %ld particles", path, cap), -1;
            if (X) {
                X[2 * n] = (double)x;
                X[2 * n + 1] = (double)y;
            }
            if (V) V[n] = (double)area;
        }
        ++n;
    }
    return n;
}

static int open_out(pd_context* c, const char* path, int digits, std::ofstream& f)
{
    if (!c || !path) return fail(c, PD_ERR_INVALID, "NULL argument");
    f.open(path);
    if (!f) return fail(c, PD_ERR_INVALID, "cannot write %s", path);
    if (digits > 0) f << std::setprecision(digits);
    return PD_OK;
}

extern "C" int pd_write_velocity(pd_context* c, const char* path, int digits)
{
    std::ofstream f;
    if (int rc = open_out(c, path, digits, f)) return rc;
    std::vector<double> v(2 * (size_t)c->n);
    if (int rc = pd_get_state(c, nullptr, v.data(), nullptr, nullptr)) return rc;
    for (int i = 0; i < c->n; ++i) f << v[2 * i] << " " << v[2 * i + 1] << "\n";
    return f.good() ? PD_OK : fail(c, PD_ERR_INVALID, "write to %s failed", path);
}

extern "C" int pd_write_positions(pd_context* c, const char* path, const int* ids, int digits)
{
    std::ofstream f;
    if (int rc = open_out(c, path, digits, f)) return rc;
    std::vector<double> u(2 * (size_t)c->n);
    if (int rc = pd_get_state(c, u.data(), nullptr, nullptr, nullptr)) return rc;
    f << "ID, x, y, ux, uy \n";
    for (int i = 0; i < c->n; ++i)
        f << (ids ? ids[i] : i) << "," << c->X_caller[2 * i] << "," << c->X_caller[2 * i + 1] << ","
          << u[2 * i] << "," << u[2 * i + 1] << "\n";
    return f.good() ? PD_OK : fail(c, PD_ERR_INVALID, "write to %s failed", path);
}

extern "C" int pd_write_gathered(pd_context* c, const char* path, const int* ids, int digits)
{
    std::ofstream f;
    if (int rc = open_out(c, path, digits, f)) return rc;
    std::vector<double> u(2 * (size_t)c->n), st(c->n, 0.0);
    if (int rc = pd_get_state(c, u.data(), nullptr, nullptr, nullptr)) return rc;
    if (c->kappa)
        if (int rc = pd_get_damage(c, nullptr, st.data(), nullptr)) return rc;
    f << "ID, x, y, ux, uy, damage\n";
    for (int i = 0; i < c->n; ++i) {
        // the reference gathers the id as a long double (:560)
        const long double id = (long double)(ids ? ids[i] : i);
        f << id << "," << c->X_caller[2 * i] << "," << c->X_caller[2 * i + 1] << "," << u[2 * i] << ","
          << u[2 * i + 1] << "," << st[i] << "\n";
    }
    return f.good() ? PD_OK : fail(c, PD_ERR_INVALID, "write to %s failed", path);
}
